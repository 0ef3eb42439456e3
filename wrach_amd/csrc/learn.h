// Internal interface of the learned warm start (learn.hip): the neighbour
// features, the 64 -> 128 -> 128 -> 2 SiLU network on bf16 MFMA, its
// training kernels and Adam.  Not installed; the public ABI is wrach.h.
#pragma once

#include "frame.h"

namespace wrach {

constexpr int kNbrs = 16;     // neighbours per particle
constexpr int kFeat = 64;     // kNbrs x (dx, dy, dvx, dvy)
constexpr int kHid = 128;     // hidden width of both layers
constexpr int kOutDim = 2;    // predicted displacement (dx, dy)
// Most particles in a particle's 5x5 cells for it to get features (else
// all-zero features): a bound on one lane's work, as kStencilBudget is for
// the 3x3 cells of a separation pass.
constexpr uint32_t kFeatBudget = 4096;

// The master parameters, f32, in one vector: W1[64][128] (input-major),
// b1[128], W2[128][128], b2[128], W3[128][2], b3[2].
constexpr uint32_t kOffW1 = 0;
constexpr uint32_t kOffB1 = kOffW1 + kFeat * kHid;
constexpr uint32_t kOffW2 = kOffB1 + kHid;
constexpr uint32_t kOffB2 = kOffW2 + kHid * kHid;
constexpr uint32_t kOffW3 = kOffB2 + kHid;
constexpr uint32_t kOffB3 = kOffW3 + kHid * kOutDim;
constexpr uint32_t kParams = kOffB3 + kOutDim;

// The dense layers' weights as v_mfma_f32_16x16x32_bf16 A operands: a
// fragment is 64 lanes x 8 bf16 (1 KiB), lane l holding row 16t + l % 16 of
// a 16-row tile t and the 8 entries of k-step s that its lane group l / 16
// supplies.  W1 (A = W1^T, rows = outputs): 8 x 2 fragments; W2 forward
// (A = W2^T): 8 x 4; W2 backward (A = W2, rows = inputs): 8 x 4.
constexpr uint32_t kFragElems = 512;
constexpr uint32_t kW1Frags = (kHid / 16) * (kFeat / 32);
constexpr uint32_t kW2Frags = (kHid / 16) * (kHid / 32);
constexpr uint32_t kFragW1 = 0;
constexpr uint32_t kFragW2F = kFragW1 + kW1Frags;
constexpr uint32_t kFragW2B = kFragW2F + kW2Frags;
constexpr uint32_t kFrags = kFragW2B + kW2Frags;

// Packed weights the kernels read: fragments (bf16 bits), W3 rounded to
// bf16 and held as f32, and the f32 master vector (biases).
struct MlpWeights {
    const uint16_t *frag;  // kFrags * kFragElems
    const float *w3;       // [128][2]
    const float *params;   // kParams
};

// Neighbour features of the packed state (pos, vel) for particles
// [0, p.n): row row0 + i of `feat` (64 bf16) for particle i.  Optional:
// each particle's cell (`cell`), and the overlap sums of these positions
// over the 3x3 cells as a separation pass measures them, added to `sums`
// (the layout of the stencil sums: slot block % kStencilSumSlots).
struct FeatureArgs {
    const float2 *pos;
    const float2 *vel;
    uint16_t *feat;
    uint32_t row0;
    uint32_t *cell;
    unsigned long long *sums;
};
hipError_t launch_learn_features(const FrameParams &p, const FrameBuffers &b, const FeatureArgs &a,
                                 hipStream_t stream);

// Inference over rows [0, n) of `feat`: y_out[i] = the network's output
// (if not null); pout[i] = pin[i] + clamp(y, +-kMaxWarm) (if not null).
constexpr float kMaxWarm = 1.0f;
struct InferArgs {
    const uint16_t *feat;
    uint32_t n;
    float2 *y_out;
    const float2 *pin;
    float2 *pout;
};
hipError_t launch_learn_infer(const MlpWeights &w, const InferArgs &a, hipStream_t stream);
// Blocks of the inference kernel the device holds at once.
uint32_t resident_blocks_infer();

// One training step's buffers (batch rows bsz, a multiple of 64).
struct TrainArgs {
    const uint16_t *feat;    // training rows
    const float2 *target;    // their targets
    const uint32_t *sel;     // batch row p is training row sel[p]
    uint32_t bsz;
    float inv_bsz;           // 1 / bsz
    // Feature-major activations and gradients ([feature][bsz]):
    uint16_t *xt;            // [64][bsz] the inputs
    uint16_t *a1t, *dz1t;    // [128][bsz] SiLU(z1), dL/dz1 (bf16)
    uint16_t *dz2t;          // [128][bsz] dL/dz2 (bf16)
    float *a2t;              // [128][bsz] SiLU(z2)
    float *dyt;              // [2][bsz] dL/dy
    float *sqerr;            // [bsz / 16] squared error per 16-row tile
};
hipError_t launch_learn_rows(const MlpWeights &w, const TrainArgs &a, hipStream_t stream);
// out[m][n] partial sums over k chunks: part[c][m][n] = sum over
// k in [c * kc, (c + 1) * kc) of A[m][k] * B[n][k] (bf16, row length ldk).
hipError_t launch_learn_gemm_nt(const uint16_t *A, const uint16_t *B, uint32_t M, uint32_t N,
                                uint32_t ldk, uint32_t kc, uint32_t chunks, float *part,
                                hipStream_t stream);
// grad[kOff..] from the GEMM partials (dW1, dW2), the column sums (b1,
// b2, W3, b3) and the loss (sum of sqerr / (2 bsz)) into loss_out.
hipError_t launch_learn_reduce(const TrainArgs &a, const float *part1, const float *part2,
                               uint32_t chunks, float *grad, float *loss_out, hipStream_t stream);
// Adam on the master vector (f32, in place), then the packed weights.
struct AdamArgs {
    float *params, *m, *v;
    const float *grad;
    float lr, beta1, beta2, eps;
    float c1, c2;  // 1 / (1 - beta1^t), 1 / (1 - beta2^t)
};
hipError_t launch_learn_adam(const AdamArgs &a, hipStream_t stream);
hipError_t launch_learn_pack(const float *params, uint16_t *frag, float *w3, hipStream_t stream);
// target[row0 + i] = pk[i] - p0[i] for i < n.
hipError_t launch_learn_target(const float2 *pk, const float2 *p0, float2 *target, uint32_t row0,
                               uint32_t n, hipStream_t stream);

}  // namespace wrach
