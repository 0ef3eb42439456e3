// The learned warm start's network on gfx950: 64 -> 128 -> 128 -> 2 with
// SiLU, bf16 weights and activations into v_mfma_f32_16x16x32_bf16 with f32
// accumulation; inference, the training rows (forward + backward through the
// layers), the weight-gradient GEMMs, the reductions and Adam.  No
// counterpart in the reference, whose separation is one Gauss-Seidel sweep
// per cell (shaders/physics/src/particles.rs:62-93); the network predicts
// the displacement that the plain separation passes produce, and those
// passes then start from its guess.
//
// MFMA layout (v_mfma_f32_16x16x32_bf16, wave64): D[i][j] = sum_k A[i][k]
// B[k][j] + C[i][j] over a 16 x 32 A, 32 x 16 B and 16 x 16 C.  Lane l
// supplies row l % 16 of A and column l % 16 of B, the 8 k entries of lane
// group g = l / 16 (k = 8g .. 8g + 7), and holds D[4g + r][l % 16],
// r = 0..3.
//
// The network runs transposed: a layer computes out^T = W^T . in^T, so the
// particles are the B/D columns.  Layer 1's B operand is then 16 bytes of a
// particle's feature row.  Layer 1's D registers hold, per lane, outputs
// 16t + 4g + r of one particle; two consecutive tiles give lane group g the
// 8 features {32s + 4g + 0..3, 32s + 16 + 4g + 0..3} -- exactly a B operand
// of layer 2's k-step s if the k order of that step is permuted the same way
// in A.  The weights are packed with that permutation (kmap below), so the
// layers chain in registers with no transpose.
#include <hip/hip_runtime.h>

#include "learn.h"

namespace wrach {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// f32 -> bf16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Feature of the permuted k order: slot e of lane group g in k-step s.
__host__ __device__ constexpr uint32_t kmap(uint32_t s, uint32_t g, uint32_t e) {
    return 32u * s + (e < 4u ? 4u * g + e : 16u + 4u * g + (e - 4u));
}

// SiLU and its derivative, in f32, from the hardware exp2 and reciprocal
// (v_exp_f32, v_rcp_f32: about 1 ulp each; the network is trained and run
// with the same function).
__device__ __forceinline__ float sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * 1.4426950408889634f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// Eight f32 to bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ bf16x8 pack8(const float (&v)[8]) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
    return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Forward pass of NC column tiles (16 particles each) by one wavefront.
// xb[c][s]: layer-1 B operands (feature rows).  frag: the packed
// fragments (global or LDS).  Leaves z1, z2 (pre-activations) and y (the
// output, summed over the lane groups: every lane holds its particle's y).
template <int NC, typename FragPtr>
__device__ __forceinline__ void mlp_forward(const bf16x8 (&xb)[NC][2], FragPtr frag,
                                            const float *b1, const float *b2, const float *w3,
                                            const float *b3, f32x4 (&z1)[8][NC], f32x4 (&z2)[8][NC],
                                            float (&y)[NC][2]) {
    const uint32_t l = lane_id(), g = l >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f32x4 bias;
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[r] = b1[16 * t + 4 * g + r];
#pragma unroll
        for (int c = 0; c < NC; ++c) z1[t][c] = bias;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 a = frag[(kFragW1 + t * 2 + s) * 64 + l];
#pragma unroll
            for (int c = 0; c < NC; ++c) z1[t][c] = mfma(a, xb[c][s], z1[t][c]);
        }
    }
    // Layer 2: B operands from layer 1's registers (permuted k).
    bf16x8 h1[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            float v[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = silu(z1[2 * s][c][r]);
                v[4 + r] = silu(z1[2 * s + 1][c][r]);
            }
            h1[c][s] = pack8(v);
        }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f32x4 bias;
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[r] = b2[16 * t + 4 * g + r];
#pragma unroll
        for (int c = 0; c < NC; ++c) z2[t][c] = bias;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 a = frag[(kFragW2F + t * 4 + s) * 64 + l];
#pragma unroll
            for (int c = 0; c < NC; ++c) z2[t][c] = mfma(a, h1[c][s], z2[t][c]);
        }
    }
    // Layer 3 (128 -> 2) on the vector ALUs: each lane group sums its 32
    // outputs, then the four groups are added.
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float y0 = 0.0f, y1 = 0.0f;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t o = 16 * t + 4 * g + r;
                const float a2 = silu(z2[t][c][r]);
                y0 = fmaf(a2, w3[2 * o], y0);
                y1 = fmaf(a2, w3[2 * o + 1], y1);
            }
        y0 += __shfl_xor(y0, 16, 64);
        y1 += __shfl_xor(y1, 16, 64);
        y0 += __shfl_xor(y0, 32, 64);
        y1 += __shfl_xor(y1, 32, 64);
        y[c][0] = y0 + b3[0];
        y[c][1] = y1 + b3[1];
    }
}

// Inference of NC column tiles with few live registers: layer 1 a pair of
// output tiles at a time (one k-step of layer 2's B operands), layer 2 one
// output tile at a time, folded straight into layer 3.  Same arithmetic as
// mlp_forward, in the same order.
template <int NC, typename FragPtr>
__device__ __forceinline__ void mlp_infer(const bf16x8 (&xb)[NC][2], FragPtr frag, const float *b1,
                                          const float *b2, const float *w3, const float *b3,
                                          float (&y)[NC][2]) {
    const uint32_t l = lane_id(), g = l >> 4;
    bf16x8 h1[NC][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        f32x4 z[2][NC];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int t = 2 * s + h;
            f32x4 bias;
#pragma unroll
            for (int r = 0; r < 4; ++r) bias[r] = b1[16 * t + 4 * g + r];
#pragma unroll
            for (int c = 0; c < NC; ++c) z[h][c] = bias;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const bf16x8 a = frag[(kFragW1 + t * 2 + k) * 64 + l];
#pragma unroll
                for (int c = 0; c < NC; ++c) z[h][c] = mfma(a, xb[c][k], z[h][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            float v[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = silu(z[0][c][r]);
                v[4 + r] = silu(z[1][c][r]);
            }
            h1[c][s] = pack8(v);
        }
    }
    float y0[NC], y1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) y0[c] = y1[c] = 0.0f;
    // One output tile at a time (not unrolled): its 4 fragments and 8 W3
    // entries are the only weights live, which keeps the kernel at 4
    // wavefronts per SIMD without spilling.
#pragma unroll 1
    for (int t = 0; t < 8; ++t) {
        f32x4 bias;
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[r] = b2[16 * t + 4 * g + r];
        f32x4 z[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) z[c] = bias;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bf16x8 a = frag[(kFragW2F + t * 4 + k) * 64 + l];
#pragma unroll
            for (int c = 0; c < NC; ++c) z[c] = mfma(a, h1[c][k], z[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t o = 16 * t + 4 * g + r;
                const float a2 = silu(z[c][r]);
                y0[c] = fmaf(a2, w3[2 * o], y0[c]);
                y1[c] = fmaf(a2, w3[2 * o + 1], y1[c]);
            }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float u0 = y0[c], u1 = y1[c];
        u0 += __shfl_xor(u0, 16, 64);
        u1 += __shfl_xor(u1, 16, 64);
        u0 += __shfl_xor(u0, 32, 64);
        u1 += __shfl_xor(u1, 32, 64);
        y[c][0] = u0 + b3[0];
        y[c][1] = u1 + b3[1];
    }
}

// Layer-1 B operands of particles row0 + 16c + l % 16 (rows >= n: zeros).
template <int NC>
__device__ __forceinline__ void load_rows(const uint16_t *feat, const uint32_t *sel, uint32_t row0,
                                          uint32_t n, bf16x8 (&xb)[NC][2]) {
    const uint32_t l = lane_id(), g = l >> 4;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t p = row0 + 16 * c + (l & 15);
        const uint32_t row = p < n ? (sel ? sel[p] : p) : 0u;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (p < n) {
                xb[c][s] = *reinterpret_cast<const bf16x8 *>(feat + (size_t)row * kFeat + 32 * s + 8 * g);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) xb[c][s][e] = __builtin_bit_cast(__bf16, (uint16_t)0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Inference: persistent blocks of 4 wavefronts, weights staged in LDS once
// per block, 32 particles per wavefront per step.

constexpr int kInferBlock = 512;
constexpr int kInferCols = 2;  // 16-particle column tiles per wavefront step

__global__ __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(4))) void learn_infer_kernel(MlpWeights w, InferArgs a) {
    __shared__ bf16x8 s_frag[(kW1Frags + kW2Frags) * 64];  // 48 KiB
    __shared__ __align__(16) float s_b1[kHid];
    __shared__ __align__(16) float s_b2[kHid];
    __shared__ __align__(16) float s_w3[kHid * kOutDim];
    __shared__ float s_b3[kOutDim];
    const bf16x8 *gfrag = reinterpret_cast<const bf16x8 *>(w.frag);
    for (uint32_t i = threadIdx.x; i < (kW1Frags + kW2Frags) * 64; i += kInferBlock) s_frag[i] = gfrag[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kHid; i += kInferBlock) {
        s_b1[i] = w.params[kOffB1 + i];
        s_b2[i] = w.params[kOffB2 + i];
    }
    for (uint32_t i = threadIdx.x; i < (uint32_t)(kHid * kOutDim); i += kInferBlock) s_w3[i] = w.w3[i];
    if (threadIdx.x < (uint32_t)kOutDim) s_b3[threadIdx.x] = w.params[kOffB3 + threadIdx.x];
    __syncthreads();
    const uint32_t l = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t per_step = 16 * kInferCols;
    for (uint32_t row0 = (blockIdx.x * (kInferBlock / 64) + wave) * per_step; row0 < a.n;
         row0 += gridDim.x * (kInferBlock / 64) * per_step) {
        bf16x8 xb[kInferCols][2];
        load_rows<kInferCols>(a.feat, nullptr, row0, a.n, xb);
        float y[kInferCols][2];
        // The weights stay in LDS: an address the optimiser cannot see
        // through keeps it from hoisting all 48 fragments into registers for
        // the whole loop (192 VGPRs, one wavefront per SIMD).
        const bf16x8 *fr = s_frag;
        const float *b1 = s_b1, *b2 = s_b2, *w3 = s_w3;
        asm volatile("" : "+v"(fr), "+v"(b1), "+v"(b2), "+v"(w3));
        mlp_infer<kInferCols>(xb, fr, b1, b2, w3, s_b3, y);
        if (l < 16) {
#pragma unroll
            for (int c = 0; c < kInferCols; ++c) {
                const uint32_t p = row0 + 16 * c + l;
                if (p < a.n) {
                    if (a.y_out) a.y_out[p] = make_float2(y[c][0], y[c][1]);
                    if (a.pout) {
                        const float2 q = a.pin[p];
                        const float dx = fminf(fmaxf(y[c][0], -kMaxWarm), kMaxWarm);
                        const float dy = fminf(fmaxf(y[c][1], -kMaxWarm), kMaxWarm);
                        a.pout[p] = make_float2(q.x + dx, q.y + dy);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Training rows: one wavefront per 16 batch rows.  Forward (kept in
// registers), the output error, the backward pass through the layers (dW2's
// transpose product on MFMA with the backward-packed W2), and the
// feature-major activations and gradients the weight GEMMs read.

__global__ __launch_bounds__(256) void learn_rows_kernel(MlpWeights w, TrainArgs a) {
    const uint32_t l = lane_id(), g = l >> 4, col = l & 15;
    const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t row0 = tile * 16;
    if (row0 >= a.bsz) return;  // whole wavefronts (bsz is a multiple of 64)
    const bf16x8 *frag = reinterpret_cast<const bf16x8 *>(w.frag);
    bf16x8 xb[1][2];
    load_rows<1>(a.feat, a.sel, row0, a.bsz, xb);
    f32x4 z1[8][1], z2[8][1];
    float y[1][2];
    mlp_forward<1>(xb, frag, w.params + kOffB1, w.params + kOffB2, w.w3, w.params + kOffB3, z1, z2, y);
    const uint32_t p = row0 + col;
    const float2 tg = a.target[a.sel[p]];
    const float e0 = y[0][0] - tg.x, e1 = y[0][1] - tg.y;
    const float dy0 = e0 * a.inv_bsz, dy1 = e1 * a.inv_bsz;
    // Squared error of the tile (lane group 0 holds each row once).
    float se = g == 0 ? e0 * e0 + e1 * e1 : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
    if (l == 0) a.sqerr[tile] = se;
    if (g == 0) {
        a.dyt[p] = dy0;
        a.dyt[a.bsz + p] = dy1;
    }
    // dL/dz2 = (W3 . dy) * SiLU'(z2); the B operands of the transpose
    // product, and the feature-major copies.
    bf16x8 d2[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float v[8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int t = 2 * s + h;
                const uint32_t o = 16 * t + 4 * g + r;
                const float z = z2[t][0][r];
                float da = w.w3[2 * o] * dy0;
                da = da + w.w3[2 * o + 1] * dy1;
                const float dz = da * silu_grad(z);
                v[4 * h + r] = dz;
                a.a2t[(size_t)o * a.bsz + p] = silu(z);
                a.dz2t[(size_t)o * a.bsz + p] = bf16_bits(dz);
            }
        d2[s] = pack8(v);
    }
    // dL/da1 = W2 . dz2 (rows = layer-2 inputs), then dL/dz1.
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma(frag[(kFragW2B + t * 4 + s) * 64 + l], d2[s], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t o = 16 * t + 4 * g + r;
            const float z = z1[t][0][r];
            a.a1t[(size_t)o * a.bsz + p] = bf16_bits(silu(z));
            a.dz1t[(size_t)o * a.bsz + p] = bf16_bits(acc[r] * silu_grad(z));
        }
    }
    // The inputs, feature-major, from the raw bits of the B operands (a
    // whole-vector cast: subscripting the bf16 vector type here read
    // element 0 for every element).
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 w4 = __builtin_bit_cast(u32x4, xb[0][s]);
#pragma unroll
        for (int e = 0; e < 8; ++e)
            a.xt[(size_t)(32 * s + 8 * g + e) * a.bsz + p] = (uint16_t)(w4[e >> 1] >> (16 * (e & 1)));
    }
}

// part[c][m][n] = sum over k in chunk c (k < ldk) of A[m][k] B[n][k]: a
// 32 x 32 output tile per block, one 16 x 16 MFMA tile per wavefront.
__global__ __launch_bounds__(256) void learn_gemm_nt_kernel(const uint16_t *A, const uint16_t *B,
                                                            uint32_t M, uint32_t N, uint32_t ldk,
                                                            uint32_t kc, float *part) {
    const uint32_t l = lane_id(), g = l >> 4, wave = threadIdx.x >> 6;
    const uint32_t m0 = blockIdx.x * 32 + (wave & 1) * 16, n0 = blockIdx.y * 32 + (wave >> 1) * 16;
    const uint32_t chunk = blockIdx.z;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t *ar = A + (size_t)(m0 + (l & 15)) * ldk + (size_t)chunk * kc + 8 * g;
    const uint16_t *br = B + (size_t)(n0 + (l & 15)) * ldk + (size_t)chunk * kc + 8 * g;
    const uint32_t kend = min(kc, ldk - chunk * kc);  // the last chunk may be shorter
    for (uint32_t k = 0; k < kend; k += 32)
        acc = mfma(*reinterpret_cast<const bf16x8 *>(ar + k), *reinterpret_cast<const bf16x8 *>(br + k), acc);
    float *out = part + (size_t)chunk * M * N;
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(size_t)(m0 + 4 * g + r) * N + n0 + (l & 15)] = acc[r];
}

// Sum of 256 per-thread partials in a fixed order (deterministic).
__device__ __forceinline__ float block_sum(float v, float *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    const float r = s[0];
    __syncthreads();
    return r;
}

// Blocks 0..kFeat*kHid/256 - 1 and the next kHid*kHid/256: dW1 and dW2
// from the GEMM partials (summed in chunk order), one element per thread.
// Then one block per bias entry / W3 row (column sums over the batch in a
// fixed order), and a last block for b3 and the loss.
constexpr uint32_t kGradBlocksW = (kFeat * kHid + kHid * kHid) / 256;

__global__ __launch_bounds__(256) void learn_reduce_kernel(TrainArgs a, const float *part1,
                                                           const float *part2, uint32_t chunks,
                                                           float *grad, float *loss_out) {
    __shared__ float s[256];
    const uint32_t blk = blockIdx.x;
    if (blk < kGradBlocksW) {
        const uint32_t e = blk * 256 + threadIdx.x;
        if (e < (uint32_t)(kFeat * kHid)) {
            float v = 0.0f;
            for (uint32_t c = 0; c < chunks; ++c) v += part1[(size_t)c * kFeat * kHid + e];
            grad[kOffW1 + e] = v;
        } else {
            const uint32_t f = e - kFeat * kHid;
            float v = 0.0f;
            for (uint32_t c = 0; c < chunks; ++c) v += part2[(size_t)c * kHid * kHid + f];
            grad[kOffW2 + f] = v;
        }
        return;
    }
    const uint32_t j = blk - kGradBlocksW;
    if (j < (uint32_t)kHid) {  // b1
        float v = 0.0f;
        for (uint32_t p = threadIdx.x; p < a.bsz; p += 256) v += bf16_float(a.dz1t[(size_t)j * a.bsz + p]);
        v = block_sum(v, s);
        if (threadIdx.x == 0) grad[kOffB1 + j] = v;
    } else if (j < 2u * kHid) {  // b2
        const uint32_t o = j - kHid;
        float v = 0.0f;
        for (uint32_t p = threadIdx.x; p < a.bsz; p += 256) v += bf16_float(a.dz2t[(size_t)o * a.bsz + p]);
        v = block_sum(v, s);
        if (threadIdx.x == 0) grad[kOffB2 + o] = v;
    } else if (j < 3u * kHid) {  // W3 row o
        const uint32_t o = j - 2 * kHid;
        float v0 = 0.0f, v1 = 0.0f;
        for (uint32_t p = threadIdx.x; p < a.bsz; p += 256) {
            const float h = a.a2t[(size_t)o * a.bsz + p];
            v0 = fmaf(h, a.dyt[p], v0);
            v1 = fmaf(h, a.dyt[a.bsz + p], v1);
        }
        v0 = block_sum(v0, s);
        v1 = block_sum(v1, s);
        if (threadIdx.x == 0) {
            grad[kOffW3 + 2 * o] = v0;
            grad[kOffW3 + 2 * o + 1] = v1;
        }
    } else {  // b3 and the loss
        float v0 = 0.0f, v1 = 0.0f, se = 0.0f;
        for (uint32_t p = threadIdx.x; p < a.bsz; p += 256) {
            v0 += a.dyt[p];
            v1 += a.dyt[a.bsz + p];
        }
        for (uint32_t t = threadIdx.x; t < a.bsz / 16; t += 256) se += a.sqerr[t];
        v0 = block_sum(v0, s);
        v1 = block_sum(v1, s);
        se = block_sum(se, s);
        if (threadIdx.x == 0) {
            grad[kOffB3] = v0;
            grad[kOffB3 + 1] = v1;
            if (loss_out) *loss_out = se * (0.5f * a.inv_bsz);
        }
    }
}

// Adam (Kingma & Ba, 2015) on the master vector, each step a rounded f32
// operation in this order (no contraction).
__global__ __launch_bounds__(256) void learn_adam_kernel(AdamArgs a) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kParams) return;
    const float gr = a.grad[i];
    float m = a.beta1 * a.m[i];
    m = m + (1.0f - a.beta1) * gr;
    float v = a.beta2 * a.v[i];
    v = v + (1.0f - a.beta2) * (gr * gr);
    a.m[i] = m;
    a.v[i] = v;
    const float mh = m * a.c1;
    const float vh = v * a.c2;
    const float step = a.lr * mh / (sqrtf(vh) + a.eps);
    a.params[i] = a.params[i] - step;
}

// Fragments and the rounded W3 from the master vector: one thread per
// fragment entry (kFrags * 512), the first 256 also round W3.
__global__ __launch_bounds__(256) void learn_pack_kernel(const float *params, uint16_t *frag, float *w3) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < (uint32_t)(kHid * kOutDim)) w3[i] = bf16_float(bf16_bits(params[kOffW3 + i]));
    if (i >= kFrags * kFragElems) return;
    const uint32_t f = i / kFragElems, l = (i / 8) % 64, e = i % 8, g = l / 16, row = l % 16;
    float v;
    if (f < kFragW2F) {  // W1^T: rows = outputs, k = inputs (natural order)
        const uint32_t t = (f - kFragW1) / 2, s = (f - kFragW1) % 2;
        v = params[kOffW1 + (32 * s + 8 * g + e) * kHid + 16 * t + row];
    } else if (f < kFragW2B) {  // W2^T: rows = outputs, k = inputs (permuted)
        const uint32_t t = (f - kFragW2F) / 4, s = (f - kFragW2F) % 4;
        v = params[kOffW2 + kmap(s, g, e) * kHid + 16 * t + row];
    } else {  // W2: rows = inputs, k = outputs (permuted)
        const uint32_t t = (f - kFragW2B) / 4, s = (f - kFragW2B) % 4;
        v = params[kOffW2 + (16 * t + row) * kHid + kmap(s, g, e)];
    }
    frag[i] = bf16_bits(v);
}

__global__ __launch_bounds__(256) void learn_target_kernel(const float2 *pk, const float2 *p0, float2 *target,
                                                           uint32_t row0, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 a = pk[i], b = p0[i];
    target[row0 + i] = make_float2(a.x - b.x, a.y - b.y);
}

}  // namespace

uint32_t resident_blocks_infer() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, learn_infer_kernel, kInferBlock, 0) != hipSuccess)
        return 0;
    return (uint32_t)(cus * (per > 0 ? per : 1));
}

hipError_t launch_learn_infer(const MlpWeights &w, const InferArgs &a, hipStream_t stream) {
    if (a.n == 0) return hipSuccess;
    static uint32_t resident = 0;
    if (!resident) resident = resident_blocks_infer();
    const uint32_t steps = (a.n + 16 * kInferCols - 1) / (16 * kInferCols);
    uint32_t grid = (steps + kInferBlock / 64 - 1) / (kInferBlock / 64);
    if (resident && grid > resident) grid = resident;
    hipLaunchKernelGGL(learn_infer_kernel, dim3(grid), dim3(kInferBlock), 0, stream, w, a);
    return hipGetLastError();
}

hipError_t launch_learn_rows(const MlpWeights &w, const TrainArgs &a, hipStream_t stream) {
    if (a.bsz == 0 || a.bsz % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(learn_rows_kernel, dim3(a.bsz / 64), dim3(256), 0, stream, w, a);
    return hipGetLastError();
}

hipError_t launch_learn_gemm_nt(const uint16_t *A, const uint16_t *B, uint32_t M, uint32_t N,
                                uint32_t ldk, uint32_t kc, uint32_t chunks, float *part,
                                hipStream_t stream) {
    if (M % 32 || N % 32 || kc % 32 || ldk % 32 || chunks == 0 || (size_t)kc * (chunks - 1) >= ldk ||
        (size_t)kc * chunks < ldk)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(learn_gemm_nt_kernel, dim3(M / 32, N / 32, chunks), dim3(256), 0, stream, A, B, M,
                       N, ldk, kc, part);
    return hipGetLastError();
}

hipError_t launch_learn_reduce(const TrainArgs &a, const float *part1, const float *part2,
                               uint32_t chunks, float *grad, float *loss_out, hipStream_t stream) {
    hipLaunchKernelGGL(learn_reduce_kernel, dim3(kGradBlocksW + 3 * kHid + 1), dim3(256), 0, stream, a,
                       part1, part2, chunks, grad, loss_out);
    return hipGetLastError();
}

hipError_t launch_learn_adam(const AdamArgs &a, hipStream_t stream) {
    hipLaunchKernelGGL(learn_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_learn_pack(const float *params, uint16_t *frag, float *w3, hipStream_t stream) {
    hipLaunchKernelGGL(learn_pack_kernel, dim3((kFrags * kFragElems + 255) / 256), dim3(256), 0, stream,
                       params, frag, w3);
    return hipGetLastError();
}

hipError_t launch_learn_target(const float2 *pk, const float2 *p0, float2 *target, uint32_t row0,
                               uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(learn_target_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, pk, p0, target,
                       row0, n);
    return hipGetLastError();
}

}  // namespace wrach
