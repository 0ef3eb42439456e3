// Frame kernels of the particle step, written for gfx950 (wave64).
//
// One frame is four launches:
//   collide_kernel     one packed particle per lane, 64 consecutive ones per
//                      wavefront; `key` gives each particle's cell, so a
//                      cell is a segment of equal keys.  A cell that lies
//                      inside the wavefront is handled there: ranked by
//                      `src` (the canonical order), its first 9 particles
//                      separated (Gauss-Seidel, particles.rs:62-93), every
//                      particle finished -- integrate and limit
//                      (shaders/physics/src/particle.rs:40-82) -- and written
//                      to pos_out/vel_out at its canonical index, the new cell
//                      key computed (particles_per_cell.wgsl:14-27) and slots
//                      claimed with one atomic per neighbour direction.  A
//                      cell that runs past the wavefront's last particle is
//                      queued by size class.
//   cells_kernel       the queued cells: 65..kWideMax particles a wavefront
//                      each plus one Gauss-Seidel wavefront per three cells;
//                      2..64 particles in lane groups of 2..64 (one particle
//                      per lane); one block per larger cell.  Every path
//                      restores the canonical order (by
//                      `src`), separates the first 9 particles, finishes every
//                      particle (the overflow ones without collisions,
//                      cell.rs:79-95), writes pos_out/vel_out at the
//                      canonical index and claims slots with one atomic per
//                      neighbour direction.
//   scan_kernel        single-pass exclusive scan of the per-cell counts with
//                      decoupled look-back; writes [0, start_0.., N] and clears
//                      the counts for the next frame (prefix_sum.wgsl +
//                      cell.rs:116-131).
//   scatter_kernel     one lane per particle: dest = start[key] + slot
//                      (pack_new_particle_data.wgsl:39-44); records the source
//                      index so the next collide pass can rebuild the stable
//                      order that the reference's CPU packer defines
//                      (spatial_bin.rs:122-142), and the new cell in `key`.
//
// In-cell order.  Slots inside a new cell come from atomics, so the stored
// order is a permutation of the canonical one (ascending source index).  A
// source cell that sends several particles in one direction reserves them as
// one contiguous run with one atomic, in canonical order; the scatter flags
// the first particle of every run in bit 31 of `src`.  A cell's segment is
// therefore a handful of ascending runs whose source ranges do not
// interleave, and ordering the runs by their first source index restores the
// canonical order in O(n) (a full sort remains the fallback).
//
// Floating point: the library is compiled with -ffp-contract=off, so every
// a*b+c is a rounded multiply then a rounded add, and f32 division and
// sqrtf are the correctly rounded sequences -- the native Rust semantics of
// the reference's shader crate.
#include <hip/hip_ext.h>

#include "frame.h"
#include "learn.h"
#include "select_nets.h"

namespace wrach {
namespace {

constexpr uint32_t kRunFlag = 0x80000000u;  // in src and in kslot.y: first of a run
constexpr uint32_t kSrcMask = 0x7fffffffu;

// ---------------------------------------------------------------------------
// Per-particle arithmetic.

// The grid of one world: its first cell key, dimensions and walls.  One
// world or a batch of equal worlds: from FrameParams; worlds of different
// sizes: from the world's WorldGrid (grid_of, below).
struct Grid {
    uint32_t base, gx, gy;
    float x1, y1;
};

// Particle::integrate + enforce_boundaries + enforce_velocity
// (shaders/physics/src/particle.rs:40-82) with the gravity/dt extension
// placed just before the integration.  A zero gravity component adds -0.0
// (FrameParams), which leaves -0.0 velocities their sign, as skipping the add
// in the reference does.  The walls are the particle's world's.
__device__ __forceinline__ void finish_particle(const FrameParams &p, const Grid &g, float &px,
                                                float &py, float &vx, float &vy) {
    vx = vx + p.gdx;
    vy = vy + p.gdy;
    px = px + vx * p.dt;
    py = py + vy * p.dt;
    if (px > g.x1) { px = g.x1; vx = vx * -1.0f; }
    if (px < p.ax) { px = p.ax; vx = vx * -1.0f; }
    if (py > g.y1) { py = g.y1; vy = vy * -1.0f; }
    if (py < p.ay) { py = p.ay; vy = vy * -1.0f; }
    if (vx < -1.0f) vx = -1.0f;
    if (vx > 1.0f) vx = 1.0f;
    if (vy < -1.0f) vy = -1.0f;
    if (vy > 1.0f) vy = 1.0f;
}


// Cell coordinates of particles_per_cell.wgsl:14-27: floor((x - ax) / cs)
// per axis, for a finished particle.  The walls leave x in [ax, x1] with
// ax = 0 and x1 = W, and the division is monotonic, so the key lies in
// [0, gx - 1] (gx - 1 = floor(W / cs) on the host, the same f32 division);
// the max only maps a NaN position, which no upload lets in, to column 0.
__device__ __forceinline__ void cell_xy(const FrameParams &p, float x, float y, uint32_t &kx,
                                        uint32_t &ky) {
    float fx, fy;
    if (p.cs == 1.0f) {  // x / 1 is exact: skip the division sequence
        fx = floorf(x - p.ax);
        fy = floorf(y - p.ay);
    } else {
        fx = floorf((x - p.ax) / p.cs);
        fy = floorf((y - p.ay) / p.cs);
    }
    kx = (uint32_t)fmaxf(fx, 0.0f);
    ky = (uint32_t)fmaxf(fy, 0.0f);
}

// First cell key of the world that holds cell key c: 0 for one world; in a
// batch of equal worlds, the world's offset in the flattened key space (a
// particle never leaves its world: the walls keep it inside).
__device__ __forceinline__ uint32_t world_base(const FrameParams &p, uint32_t c) {
    if (p.world_cells == p.ncell) return 0u;
    return c / p.world_cells * p.world_cells;
}

// The grid of the world that holds cell key c (c < ncell).  Worlds of
// different sizes start on chunk boundaries, so the chunk names the world.
__device__ __forceinline__ Grid grid_of(const FrameParams &p, uint32_t c) {
    Grid g;
    if (p.grids) {
        const uint32_t last = (p.ncell - 1) >> kChunkShift;
        const uint32_t ch = (c >> kChunkShift) < last ? (c >> kChunkShift) : last;
        const WorldGrid w = p.grids[p.chunk_world[ch]];
        g.base = w.base;
        g.gx = w.gx;
        g.gy = w.gy;
        g.x1 = w.x1;
        g.y1 = w.y1;
    } else {
        g.base = world_base(p, c);
        g.gx = p.gx;
        g.gy = p.gy;
        g.x1 = p.x1;
        g.y1 = p.y1;
    }
    return g;
}

// Grid coordinates of cell key c inside its world g.
__device__ __forceinline__ void cell_coords(const Grid &g, uint32_t c, uint32_t &cx, uint32_t &cy) {
    const uint32_t l = c - g.base;
    cy = l / g.gx;
    cx = l - cy * g.gx;
}

__device__ __forceinline__ uint32_t cell_key(const FrameParams &p, const Grid &g, float x, float y) {
    uint32_t kx, ky;
    cell_xy(p, x, y, kx, ky);
    return g.base + ky * g.gx + kx;
}

// Claims `count` consecutive slots in cell k.  The frame's first claim in a
// chunk marks it, so that the scan visits only chunks that hold particles.
__device__ __forceinline__ uint32_t claim(const FrameBuffers &b, uint32_t k, uint32_t count) {
    const uint32_t old = atomicAdd(&b.cnt[k], count);
    if (old == 0) b.chunk_next[k >> kChunkShift] = 1;
    return old;
}

// First particle of cell k (k == ncell: the end).  Cells of a chunk the last
// scan found empty all start at the chunk's prefix (their idx entries are
// stale); a live chunk's are in idx.
__device__ __forceinline__ uint32_t cell_start(const FrameParams &p, const FrameBuffers &b,
                                               uint32_t k) {
    if (k >= p.ncell) return p.n;
    const uint32_t ch = k >> kChunkShift;
    const uint32_t s = b.chunk_live[ch] ? b.idx[k + 1] : b.chunk_prefix[ch];
    return s < p.n ? s : p.n;
}


// Neighbour direction of a new cell relative to the old one (cx, cy) of
// world g: 0..8 for the 3x3 neighbourhood (row-major from (-1,-1)), 9 for
// anything farther.
__device__ __forceinline__ uint32_t direction(const FrameParams &p, const Grid &g, float x, float y,
                                              uint32_t cx, uint32_t cy, uint32_t &key) {
    uint32_t kx, ky;
    cell_xy(p, x, y, kx, ky);
    key = g.base + ky * g.gx + kx;
    const int dx = (int)kx - (int)cx, dy = (int)ky - (int)cy;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1) return 9;
    return (uint32_t)((dy + 1) * 3 + (dx + 1));
}

// Lane index, opaque to the optimiser: values derived from it (lane masks,
// Gauss-Seidel partners) are then recomputed where they are used instead of
// being hoisted out of the work loops and held in registers for the whole
// kernel, which spills scalar registers.
__device__ __forceinline__ uint32_t lane_id() {
    uint32_t l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    return l;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << lane_id()) - 1;
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Orders LDS accesses of one wavefront: its LDS instructions execute in
// program order, so a compiler fence is all a wave-private region needs.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Batcher odd-even merge sort networks, generated at compile time so that
// every register-array index is a constant after unrolling.

template <int N>
struct Network {
    static constexpr int count() {
        int c = 0;
        for (int p = 1; p < N; p += p)
            for (int k = p; k > 0; k /= 2)
                for (int j = k % p; j + k < N; j += k + k)
                    for (int i = 0; i < k; ++i)
                        if (i + j + k < N && (i + j) / (p + p) == (i + j + k) / (p + p)) ++c;
        return c;
    }
    static constexpr int size = count();
    int lo[size];
    int hi[size];
    constexpr Network() : lo(), hi() {
        int c = 0;
        for (int p = 1; p < N; p += p)
            for (int k = p; k > 0; k /= 2)
                for (int j = k % p; j + k < N; j += k + k)
                    for (int i = 0; i < k; ++i)
                        if (i + j + k < N && (i + j) / (p + p) == (i + j + k) / (p + p)) {
                            lo[c] = i + j;
                            hi[c] = i + j + k;
                            ++c;
                        }
    }
};
static_assert(Network<2>::size == 1 && Network<4>::size == 5 && Network<8>::size == 19 &&
                  Network<16>::size == 63,
              "Batcher network sizes");

template <int N>
__device__ __forceinline__ void sort_by_src(uint32_t (&key)[N], float (&px)[N], float (&py)[N],
                                            float (&vx)[N], float (&vy)[N]) {
    constexpr Network<N> net{};
#pragma unroll
    for (int q = 0; q < Network<N>::size; ++q) {
        const int a = net.lo[q], c = net.hi[q];
        const bool sw = key[a] > key[c];
        const uint32_t k0 = key[a], k1 = key[c];
        const float x0 = px[a], x1 = px[c], y0 = py[a], y1 = py[c];
        const float u0 = vx[a], u1 = vx[c], w0 = vy[a], w1 = vy[c];
        key[a] = sw ? k1 : k0;  key[c] = sw ? k0 : k1;
        px[a] = sw ? x1 : x0;   px[c] = sw ? x0 : x1;
        py[a] = sw ? y1 : y0;   py[c] = sw ? y0 : y1;
        vx[a] = sw ? u1 : u0;   vx[c] = sw ? u0 : u1;
        vy[a] = sw ? w1 : w0;   vy[c] = sw ? w0 : w1;
    }
}

// ---------------------------------------------------------------------------
// Cells handled by one lane, in the readback's canonical pass: copy a cell
// of 1 or 2..NB particles to pos_out/vel_out in canonical order (ascending
// source index).

__device__ __forceinline__ void one_particle(const FrameBuffers &b, uint32_t s) {
    b.pos_out[s] = b.pos[s];
    b.vel_out[s] = b.vel[s];
}

template <int NB>
__device__ __forceinline__ void small_cell(const FrameBuffers &b, uint32_t s, uint32_t n) {
    uint32_t key[NB];
    float px[NB], py[NB], vx[NB], vy[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < (int)n) {
            const float2 q = b.pos[s + j];
            const float2 v = b.vel[s + j];
            px[j] = q.x; py[j] = q.y; vx[j] = v.x; vy[j] = v.y;
            key[j] = b.src[s + j] & kSrcMask;
        } else {
            px[j] = py[j] = vx[j] = vy[j] = 0.0f;
            key[j] = 0xffffffffu;
        }
    }
    sort_by_src<NB>(key, px, py, vx, vy);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j < (int)n) {
            b.pos_out[s + j] = make_float2(px[j], py[j]);
            b.vel_out[s + j] = make_float2(vx[j], vy[j]);
        }
    }
}

// ---------------------------------------------------------------------------
// Cells of more than kWideMax particles: the whole block cooperates.

// Ascending bitonic sort of P (a power of two, <= kLdsSort) 64-bit keys in
// LDS.  Callers synchronise before the call; the last stage ends with a
// barrier.
__device__ void bitonic_lds(unsigned long long *sk, uint32_t P) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < P; i += kCollideBlock) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long x = sk[i], y = sk[l];
                    const bool asc = (i & k) == 0;
                    if ((x > y) == asc) {
                        sk[i] = y;
                        sk[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Merge of two sorted runs of distinct keys into z, by the whole block (merge
// path: each thread finds its split with a binary search, then merges its
// share of the output sequentially).
__device__ void merge_block(const unsigned long long *x, uint32_t nx,
                            const unsigned long long *y, uint32_t ny,
                            unsigned long long *z) {
    const uint32_t tid = threadIdx.x;
    const uint64_t total = (uint64_t)nx + ny;
    const uint32_t d0 = (uint32_t)(total * tid / kCollideBlock);
    const uint32_t d1 = (uint32_t)(total * (tid + 1) / kCollideBlock);
    uint32_t lo = d0 > ny ? d0 - ny : 0;
    uint32_t hi = d0 < nx ? d0 : nx;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (x[mid] < y[d0 - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    uint32_t i = lo, j = d0 - lo;
    for (uint32_t d = d0; d < d1; ++d) {
        const bool take_x = j >= ny || (i < nx && x[i] < y[j]);
        z[d] = take_x ? x[i++] : y[j++];
    }
}

// Sorts (src << 32 | local index) for a cell of n > kLdsSort particles:
// LDS-sorted chunks, then block merges ping-ponging between the cell's own
// ranges of kslot and sortbuf (both unused by any other cell).  Returns the
// buffer holding the result.
__device__ __forceinline__ const unsigned long long *sort_large_cell(const FrameBuffers &b, uint32_t s, uint32_t n,
                                                     unsigned long long *sk) {
    const uint32_t tid = threadIdx.x;
    unsigned long long *a = reinterpret_cast<unsigned long long *>(b.kslot) + s;
    unsigned long long *t = b.sortbuf + s;
    for (uint32_t base = 0; base < n; base += kLdsSort) {
        const uint32_t len = min((uint32_t)kLdsSort, n - base);
        for (uint32_t i = tid; i < (uint32_t)kLdsSort; i += kCollideBlock)
            sk[i] = i < len ? ((unsigned long long)(b.src[s + base + i] & kSrcMask) << 32) | (base + i)
                            : ~0ull;
        __syncthreads();
        bitonic_lds(sk, kLdsSort);
        for (uint32_t i = tid; i < len; i += kCollideBlock) a[base + i] = sk[i];
        __syncthreads();
    }
    unsigned long long *from = a, *to = t;
    for (uint64_t w = kLdsSort; w < n; w *= 2) {
        for (uint64_t lo = 0; lo < n; lo += 2 * w) {
            const uint64_t mid = lo + w < n ? lo + w : n;
            const uint64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            merge_block(from + lo, (uint32_t)(mid - lo), from + mid, (uint32_t)(hi - mid), to + lo);
        }
        __syncthreads();
        unsigned long long *sw = from;
        from = to;
        to = sw;
    }
    return from;
}

template <int G>
__device__ __forceinline__ void collide_group(float &x, float &y, uint32_t r, uint32_t m,
                                              uint32_t gbase);

template <bool STEP>
__device__ __forceinline__ void big_cell(const FrameParams &p, const FrameBuffers &b, uint32_t s,
                                         uint32_t n, uint32_t c, unsigned long long *sk) {
    const uint32_t tid = threadIdx.x;
    const unsigned long long *perm;
    if (n <= (uint32_t)kLdsSort) {
        uint32_t P = 2;
        while (P < n) P <<= 1;
        for (uint32_t i = tid; i < P; i += kCollideBlock)
            sk[i] = i < n ? ((unsigned long long)(b.src[s + i] & kSrcMask) << 32) | i : ~0ull;
        __syncthreads();
        bitonic_lds(sk, P);
        perm = sk;
    } else {
        perm = sort_large_cell(b, s, n, sk);
    }
    if (STEP) {
        // Pass 1 over 256-rank chunks in canonical order: finish, write at
        // s + r, and rank each particle among the cell's particles going the
        // same way (ballots per wave, offsets across the block's waves, running
        // totals across chunks).  Ranks 0..8 are separated first by lanes
        // 0..8 of wave 0 (collide_group).  Per rank, `tmp` keeps (direction,
        // rank in it), or for a farther mover (key | bit 31, its slot); it is
        // whichever of kslot / sortbuf does not hold the order.
        __shared__ uint32_t s_wcnt[kWavesPerBlock][9], s_run[9], s_base[9];
        const uint32_t lane = tid & 63, wave = tid >> 6;
        const unsigned long long lt = lanemask_lt();
        const Grid g = grid_of(p, c);
        uint32_t cx, cy;
        cell_coords(g, c, cx, cy);
        uint2 *tmp = reinterpret_cast<const void *>(perm) == reinterpret_cast<const void *>(b.kslot + s)
                         ? reinterpret_cast<uint2 *>(b.sortbuf + s)
                         : b.kslot + s;
        if (tid < 9) s_run[tid] = 0;
        for (uint32_t base = 0; base < n; base += kCollideBlock) {
            const uint32_t r = base + tid;
            float2 q = make_float2(0.0f, 0.0f), v = make_float2(0.0f, 0.0f);
            if (r < n) {
                const uint32_t a = (uint32_t)perm[r];
                q = b.pos[s + a];
                v = b.vel[s + a];
            }
            if (base == 0 && wave == 0) collide_group<64>(q.x, q.y, lane, kMaxInCell, 0);
            uint32_t dir = 10, key = 0;
            if (r < n) {
                finish_particle(p, g, q.x, q.y, v.x, v.y);
                b.pos_out[s + r] = q;
                b.vel_out[s + r] = v;
                dir = direction(p, g, q.x, q.y, cx, cy, key);
            }
            uint32_t rin = 0;
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                const unsigned long long m = __ballot(dir == (uint32_t)d);
                if (dir == (uint32_t)d) rin = (uint32_t)__popcll(m & lt);
                if (lane == 0) s_wcnt[wave][d] = (uint32_t)__popcll(m);
            }
            __syncthreads();
            if (dir < 9) {
                uint32_t off = s_run[dir];
                for (uint32_t w = 0; w < wave; ++w) off += s_wcnt[w][dir];
                tmp[r] = make_uint2(dir, off + rin);
            } else if (dir == 9) {
                const uint32_t slot = claim(b, key, 1u);
                tmp[r] = make_uint2(key | kRunFlag, slot | kRunFlag);
            }
            __syncthreads();
            if (tid < 9) {
                uint32_t t = 0;
                for (uint32_t w = 0; w < (uint32_t)kWavesPerBlock; ++w) t += s_wcnt[w][tid];
                s_run[tid] += t;
            }
            __syncthreads();
        }
        // One atomic per occupied direction, then the slots.
        if (tid < 9) {
            const uint32_t tx = cx + tid % 3 - 1, ty = cy + tid / 3 - 1;
            s_base[tid] = s_run[tid] ? claim(b, g.base + ty * g.gx + tx, s_run[tid]) : 0u;
        }
        __syncthreads();
        for (uint32_t r = tid; r < n; r += kCollideBlock) {
            const uint2 t = tmp[r];
            if (t.x & kRunFlag) {
                b.kslot[s + r] = make_uint2(t.x & kSrcMask, t.y);
            } else {
                const uint32_t key = g.base + (cy + t.x / 3 - 1) * g.gx + (cx + t.x % 3 - 1);
                b.kslot[s + r] = make_uint2(key, (s_base[t.x] + t.y) | (t.y == 0 ? kRunFlag : 0u));
            }
        }
    } else {
        for (uint32_t r = tid; r < n; r += kCollideBlock) {
            const uint32_t a = (uint32_t)perm[r];
            b.pos_out[s + r] = b.pos[s + a];
            b.vel_out[s + r] = b.vel[s + a];
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The collide pass.

// Per-cell offsets with a guard for a corrupt offset array: never read past
// the particles.
__device__ __forceinline__ void clamp_range(const FrameParams &p, uint32_t s, uint32_t &n) {
    if (s > p.n) n = 0;
    else if (n > p.n - s) n = p.n - s;
}

// Size class of a queued cell of n particles (kClasses for none: 0 or 1
// particle).  For 2 <= n <= 64 the class is ceil(log2(n)) - 1.
__device__ __forceinline__ uint32_t size_class(uint32_t n) {
    if (n < 2) return (uint32_t)kClasses;
    if (n > kWideMax) return (uint32_t)kBlockClass;
    const uint32_t k = 31u - (uint32_t)__builtin_clz(n - 1);
    return k < (uint32_t)kWaveClass ? k : (uint32_t)kWaveClass;
}

// Wave-private LDS of a cell wave of the wide path (below).
constexpr int kWideRuns = 256;  // runs per cell ordered by runs (more: pairwise ranks)
struct WaveLds {
    uint32_t code[kWideMax];          // per rank >= 9: direction (bit 31 set); 0 for a far mover
    uint32_t head_key[kWideRuns];     // first source index of each run
    uint32_t head_pos[kWideRuns];     // stored position of each run's start
    uint32_t run_off[kWideRuns];      // canonical rank of each run's start minus its position
    uint2 head_kl[kWideRuns];         // (first source index, length) of each run
    uint2 dir_tbl[9];                 // per direction: (new cell key, first slot)
};

// Gauss-Seidel on the first m = min(n, 9) particles of a cell held by lanes
// gbase + r (r = canonical rank).  In the in-order loop (particles.rs:62-81)
// pair (i,j) can run at step i + j: the pairs of one step touch disjoint
// particles and every particle still sees its updates in the loop's order, so
// 2M - 3 steps (M = min(G, 9)) give the loop's result bit for bit.  Both lanes
// of a pair fetch the other's position and evaluate collide_pair's arithmetic
// on the same operands; the lower rank applies -dv, the higher +dv.  All
// lanes execute the shuffles.
// Step t of that schedule for the lane holding rank ri of a cell whose
// first mi ranks collide, ranks held by lanes gbase + r.  The pair rule is
// Particles::pairs (particles.rs:65-92): glam distance, skip beyond
// MIN_DISTANCE = 1, 0 -> 0.0001, force (0.5*(1-d))/d, equal and opposite
// displacement.
__device__ __forceinline__ void separate_step(float &x, float &y, int ri, int mi, int gbase, int t) {
    const int lane = (int)lane_id();
    const int partner = t - ri;
    const bool left = ri < partner && partner < mi;
    const bool right = partner >= 0 && partner < ri && ri < mi;
    const int from = (left || right) ? gbase + partner : lane;
    const float qx = __shfl(x, from, 64);
    const float qy = __shfl(y, from, 64);
    if (left || right) {
        const float xi = left ? x : qx, yi = left ? y : qy;
        const float xj = left ? qx : x, yj = left ? qy : y;
        const float dx = xi - xj;
        const float dy = yi - yj;
        float sq = dx * dx;
        sq = sq + dy * dy;
        float d = sqrtf(sq);
        if (!(d > 1.0f)) {
            if (d == 0.0f) d = 0.0001f;
            const float f = (0.5f * (1.0f - d)) / d;
            const float ex = (xj - xi) * f;
            const float ey = (yj - yi) * f;
            if (left) {
                x = x - ex;
                y = y - ey;
            } else {
                x = x + ex;
                y = y + ey;
            }
        }
    }
}

template <int G>
__device__ __forceinline__ void collide_group(float &x, float &y, uint32_t r, uint32_t m,
                                              uint32_t gbase) {
    constexpr int M = G < kMaxInCell ? G : kMaxInCell;
#pragma unroll
    for (int t = 1; t <= 2 * M - 3; ++t) separate_step(x, y, (int)r, (int)m, (int)gbase, t);
}

// ---------------------------------------------------------------------------
// collide_kernel: one packed particle per lane, the 64 consecutive particles
// i0 .. i0 + 63 per wavefront, whose cells come from `key` (written by the
// previous frame's scatter or by the upload): a cell is a segment of equal
// keys.
//  - A cell of at most kInWaveMax particles that lies inside the wavefront is
//    handled there, with the arithmetic of the cell kernels' lane groups:
//    ranked by source index, moved to lane (first lane + rank), separated,
//    finished and written at i0 + lane (its canonical index), slots claimed
//    with one atomic per direction.  The wavefront reads whole lines of pos,
//    vel, src and key and writes whole lines of pos_out, vel_out and kslot;
//    before the collapse under gravity almost every cell is such a segment.
//    Every lane runs the separation steps of the wavefront's largest such
//    cell, so the cap bounds them (2 * 3 - 3 = 3 steps); larger cells go to
//    the lane groups of cells_kernel, which are sorted by size.
//  - Any other cell whose first particle lies in the wavefront -- larger, or
//    running past the wavefront's last particle -- is queued by size class
//    for cells_kernel, with its range from the offsets.  The range must start
//    at that particle and agree with the keys, or the kernel raises
//    kErrLayout and queues nothing for it (keys and offsets that disagree
//    would make queued ranges overlap).
//  - The lanes of a cell that began before the wavefront leave it to the
//    wavefront that queued it.  The kernel has no barrier: a wavefront with
//    nothing to handle or queue returns after the keys, and queued cells go
//    to their lists with one atomic each (slice = wavefront % kSlices).
constexpr uint32_t kNoKey = 0xffffffffu;  // past the particles; cell keys stay below 2^31
constexpr uint32_t kInWaveMax = 3;

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = (uint32_t)__shfl_xor((int)v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Start and end of cell k (k < ncell) as cell_start gives them, the six
// loads in flight together.
__device__ __forceinline__ void cell_range(const FrameParams &p, const FrameBuffers &b, uint32_t k,
                                           uint32_t &st, uint32_t &en) {
    const uint32_t c0 = k >> kChunkShift;
    const uint32_t live0 = b.chunk_live[c0], at0 = b.idx[k + 1], pre0 = b.chunk_prefix[c0];
    uint32_t live1 = 1, at1 = p.n, pre1 = p.n;
    if (k + 1 < p.ncell) {
        const uint32_t c1 = (k + 1) >> kChunkShift;
        live1 = b.chunk_live[c1];
        at1 = b.idx[k + 2];
        pre1 = b.chunk_prefix[c1];
    }
    st = live0 ? at0 : pre0;
    en = live1 ? at1 : pre1;
    st = st < p.n ? st : p.n;
    en = en < p.n ? en : p.n;
}

// 64-particle windows per wavefront (the kernel takes any number: each
// window's loads and slot atomics are in flight together).  One is the
// fastest measured: with 2..4 the frame before the collapse is no faster and
// the collapsed frame is slower.
constexpr int kCollideWindows = 1;

__global__ __launch_bounds__(kCollideBlock) void collide_kernel(FrameParams p, FrameBuffers b) {
    constexpr int W = kCollideWindows;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * kWavesPerBlock + wave) * (64u * W);
    uint32_t k[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t i = base + 64u * j + lane;
        k[j] = i < p.n ? b.key[i] : kNoKey;
    }
    // The keys on either side of the wavefront's particles: lane 0 reads the
    // one before, lane 63 the one after.
    uint32_t before = kNoKey, after = kNoKey;
    if (lane == 0 && base > 0 && base < p.n) before = b.key[base - 1];
    if (lane == 63 && base + 64u * W < p.n) after = b.key[base + 64u * W];
    // A frame after one that raised an error (a scan that gave up its
    // look-back, keys that disagree with the offsets) reads no offsets: the
    // context is dead.  A vector load beside the keys (no barrier follows in
    // this kernel, so wavefronts need not agree).
    const uint32_t dead = b.error[lane >> 6];
    before = __builtin_amdgcn_readlane(before, 0);
    after = __builtin_amdgcn_readlane(after, 63);

    // Per window: my cell's first and last lane (tl == 64: past the window),
    // whether this window handles it (mine) or queues it (queue).
    int hl[W], tl[W];
    uint32_t m[W];
    bool mine[W], queue[W];
    bool work = false, bad = false;
    const unsigned long long lt = lanemask_lt(), le = lt | (1ull << lane);
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t i = base + 64u * j + lane;
        const bool valid = i < p.n;
        uint32_t kp = (uint32_t)__shfl_up((int)k[j], 1, 64);
        uint32_t kn = (uint32_t)__shfl_down((int)k[j], 1, 64);
        const uint32_t wb = j == 0 ? before : __builtin_amdgcn_readlane(k[j > 0 ? j - 1 : 0], 63);
        const uint32_t wa = j == W - 1 ? after : __builtin_amdgcn_readlane(k[j < W - 1 ? j + 1 : 0], 0);
        if (lane == 0) kp = wb;
        if (lane == 63) kn = wa;
        const bool head = valid && k[j] != kp;
        const bool tail = valid && k[j] != kn;
        const unsigned long long heads = __ballot(head), tails = __ballot(tail);
        const unsigned long long hb = heads & le, ta = tails & ~lt;
        hl[j] = hb ? 63 - (int)__builtin_clzll(hb) : -1;
        tl[j] = ta ? (int)__builtin_ctzll(ta) : 64;
        // A key past the grid is a layout no upload or frame writes.
        const bool kbad = valid && k[j] >= p.ncell;
        const bool inside = valid && !kbad && hl[j] >= 0 && tl[j] < 64;
        m[j] = inside ? (uint32_t)(tl[j] - hl[j] + 1) : 0u;
        mine[j] = inside && m[j] <= kInWaveMax;
        queue[j] = head && !kbad && !mine[j];
        work = work || mine[j] || queue[j];
        bad = bad || kbad;
    }
    if (__builtin_amdgcn_readfirstlane(dead)) return;
    if (bad) atomicOr(b.error, kErrLayout);
    if (__ballot(work) == 0) return;  // every lane votes; nothing to handle or queue

    // Loads: the queued cells' ranges and the handled particles, all windows.
    uint32_t qst[W], qen[W], sv[W];
    float2 q[W], w[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t i = base + 64u * j + lane;
        qst[j] = qen[j] = 0;
        if (queue[j]) cell_range(p, b, k[j], qst[j], qen[j]);
        sv[j] = 0;
        q[j] = w[j] = make_float2(0.0f, 0.0f);
        if (mine[j]) {
            sv[j] = b.src[i] & kSrcMask;
            q[j] = b.pos[i];
            w[j] = b.vel[i];
        }
    }

    // The first particle of a cell this window does not handle queues it.
    uint32_t cls[W], qs[W], qn[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t i = base + 64u * j + lane;
        const uint32_t i0 = base + 64u * j;
        cls[j] = (uint32_t)kClasses;
        qs[j] = qn[j] = 0;
        if (queue[j]) {
            const bool ok = tl[j] == 64 ? qen[j] >= i0 + 64 : qen[j] == i0 + (uint32_t)tl[j] + 1;
            if (qst[j] == i && ok && qen[j] <= p.n) {
                qs[j] = qst[j];
                qn[j] = qen[j] - qst[j];
                cls[j] = size_class(qn[j]);
            } else {
                atomicOr(b.error, kErrLayout);
            }
        }
    }

    // Handled cells: rank, separate, finish, claim (every window's atomics
    // in flight together), then write.
    uint32_t nkey[W], rin[W], leader[W], got[W];
    bool near[W];
    float px[W], py[W], vx[W], vy[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t mw = wave_max_u32(mine[j] ? m[j] : 0u);
        // Canonical rank: source indices inside a cell are distinct.
        uint32_t rank = 0;
        for (uint32_t t = 0; t < mw; ++t) {
            const bool in = mine[j] && t < m[j];
            const uint32_t o = (uint32_t)__shfl((int)sv[j], in ? hl[j] + (int)t : (int)lane, 64);
            rank += (in && o < sv[j]) ? 1u : 0u;
        }
        const int to = (mine[j] ? hl[j] + (int)rank : (int)lane) * 4;
        px[j] = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(q[j].x)));
        py[j] = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(q[j].y)));
        vx[j] = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(w[j].x)));
        vy[j] = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(w[j].y)));
        // Lane hl + r now holds rank r of its cell (m <= kInWaveMax < 9: every
        // rank collides).
        const int r = mine[j] ? (int)lane - hl[j] : 0;
        const int mi = mine[j] ? (int)m[j] : 0;
        for (int t = 1; t <= 2 * (int)mw - 3; ++t) separate_step(px[j], py[j], r, mi, mine[j] ? hl[j] : 0, t);

        uint32_t dir = 10;
        nkey[j] = 0;
        if (mine[j]) {
            const Grid g = grid_of(p, k[j]);
            uint32_t cx, cy;
            cell_coords(g, k[j], cx, cy);
            finish_particle(p, g, px[j], py[j], vx[j], vy[j]);
            dir = direction(p, g, px[j], py[j], cx, cy, nkey[j]);
        }
        // Lanes of my cell going the same way.
        unsigned long long same = 0;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const unsigned long long bm = __ballot(dir == (uint32_t)d);
            same = dir == (uint32_t)d ? bm : same;
        }
        const unsigned long long upto = tl[j] >= 63 ? ~0ull : (1ull << (tl[j] + 1)) - 1;
        const unsigned long long seg = mine[j] ? upto & ~((1ull << hl[j]) - 1) : 0ull;
        same &= seg;
        near[j] = dir < 9;
        rin[j] = (uint32_t)__popcll(same & lt);
        leader[j] = near[j] ? (uint32_t)__builtin_ctzll(same) : lane;
        got[j] = 0;
        if (mine[j] && (!near[j] || rin[j] == 0))
            got[j] = claim(b, nkey[j], near[j] ? (uint32_t)__popcll(same) : 1u);
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t i = base + 64u * j + lane;
        const uint32_t first = (uint32_t)__shfl((int)got[j], (int)leader[j], 64);
        if (mine[j]) {
            const uint32_t slot = near[j] ? (first + rin[j]) | (rin[j] == 0 ? kRunFlag : 0u) : (got[j] | kRunFlag);
            b.pos_out[i] = make_float2(px[j], py[j]);
            b.vel_out[i] = make_float2(vx[j], vy[j]);
            b.kslot[i] = make_uint2(nkey[j], slot);
        }
    }

    // Queued cells, one global atomic each into slice (wavefront % kSlices)
    // of their class; what does not fit goes to the class's overflow list.
    const uint32_t slice = (blockIdx.x * kWavesPerBlock + wave) % kSlices;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t c = cls[j];
        if (c < (uint32_t)kClasses) {
            const uint4 ent = make_uint4(qs[j], qn[j], k[j], 0);
            const uint32_t a = atomicAdd(&b.class_ctr[ctr_index((int)c, slice)], 1u);
            if (a < b.slice_cap[c]) {
                b.cells[b.class_off[c] + slice * b.slice_cap[c] + a] = ent;
            } else {
                const uint32_t o = atomicAdd(&b.class_ctr[ctr_index((int)c, kSlices)], 1u);
                if (o < b.class_cap[c]) b.cells[b.class_off[c] + kSlices * b.slice_cap[c] + o] = ent;
            }
        }
    }
}

// Cells of 2 <= n <= G particles (G = 2..64), 64 / G cells per wavefront,
// one particle per lane.  Each group of G lanes:
//  1. loads its cell's particles (lane j: stored position j);
//  2. ranks them by source index (the canonical order) and moves rank r to
//     lane gbase + r;
//  3. hands them to ranked_group (below).
template <int G>
__device__ __forceinline__ void ranked_group(const FrameParams &p, const FrameBuffers &b,
                                             float px, float py, float vx, float vy, uint32_t j,
                                             uint32_t n, uint32_t s, uint32_t c, uint32_t gbase);

// One lane's share of a group item: its cell (s, n, c) and, in lane j of
// the group, stored particle j.  group_load issues the loads; group_run
// waits for them, so a wavefront can load its next item while it computes.
struct GroupItem {
    uint32_t s, n, c, key;
    float2 q, w;
};

template <int G>
__device__ __forceinline__ uint4 group_entry(const uint4 *list, uint32_t count) {
    const uint32_t e = lane_id() / G;
    return e < count ? list[e] : make_uint4(0, 0, 0, 0);
}

template <int G>
__device__ __forceinline__ GroupItem group_load(const FrameBuffers &b, uint4 cell) {
    const uint32_t j = lane_id() % G;
    GroupItem it;
    it.s = cell.x;
    it.n = cell.y;
    it.c = cell.z;
    it.key = 0xffffffffu;
    it.q = it.w = make_float2(0.0f, 0.0f);
    if (j < it.n) {
        it.key = b.src[it.s + j] & kSrcMask;
        it.q = b.pos[it.s + j];
        it.w = b.vel[it.s + j];
    }
    return it;
}

template <int G>
__device__ __forceinline__ void group_run(const FrameParams &p, const FrameBuffers &b,
                                          const GroupItem &it) {
    const uint32_t lane = lane_id();
    const uint32_t j = lane % G;
    const uint32_t gbase = lane - j;
    const uint32_t s = it.s, n = it.n, c = it.c, key = it.key;
    const float2 q = it.q, w = it.w;
    // Source indices inside a cell are distinct; empty lanes keep their own
    // position after the n particles.
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < G; ++k) rank += __shfl(key, (int)gbase + k, 64) < key;
    if (j >= n) rank = j;
    const int to = (int)(gbase + rank) * 4;
    float px = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(q.x)));
    float py = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(q.y)));
    float vx = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(w.x)));
    float vy = __int_as_float(__builtin_amdgcn_ds_permute(to, __float_as_int(w.y)));

    // Lane gbase + r now holds rank r.
    ranked_group<G>(p, b, px, py, vx, vy, j, n, s, c, gbase);
}

// The particles of ranks 0..n-1 of a cell segment that starts at canonical
// index s, held by lanes gbase + r of a group of G lanes: separate the first
// min(n, 9) (collide_group), finish every particle, write it at s + r, and
// claim slots: one atomic per occupied neighbour direction (by the
// direction's first lane) reserves a run of consecutive slots in canonical
// order; a farther mover takes a slot of its own.
template <int G>
__device__ __forceinline__ void ranked_group(const FrameParams &p, const FrameBuffers &b,
                                             float px, float py, float vx, float vy, uint32_t j,
                                             uint32_t n, uint32_t s, uint32_t c, uint32_t gbase) {
    const uint32_t lane = lane_id();
    collide_group<G>(px, py, j, n < (uint32_t)kMaxInCell ? n : (uint32_t)kMaxInCell, gbase);

    const bool has = j < n;
    const Grid g = grid_of(p, c);
    uint32_t cx, cy;
    cell_coords(g, c, cx, cy);
    uint32_t dir = 10, nkey = 0;
    if (has) {
        finish_particle(p, g, px, py, vx, vy);
        dir = direction(p, g, px, py, cx, cy, nkey);
        b.pos_out[s + j] = make_float2(px, py);
        b.vel_out[s + j] = make_float2(vx, vy);
    }
    // Lanes of this group going the same way.
    unsigned long long same = 0;
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const unsigned long long m = __ballot(dir == (uint32_t)d);
        same = dir == (uint32_t)d ? m : same;
    }
    const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gbase);
    same &= gmask;
    const bool near = dir < 9;
    const uint32_t rin = (uint32_t)__popcll(same & lanemask_lt());
    const uint32_t leader = near ? (uint32_t)__builtin_ctzll(same) : lane;
    uint32_t got = 0;
    if (has && (!near || rin == 0))
        got = claim(b, nkey, near ? (uint32_t)__popcll(same) : 1u);
    const uint32_t base = (uint32_t)__shfl((int)got, (int)leader, 64);
    if (has) {
        const uint32_t slot = near ? (base + rin) | (rin == 0 ? kRunFlag : 0u) : (got | kRunFlag);
        b.kslot[s + j] = make_uint2(nkey, slot);
    }
}

// ---------------------------------------------------------------------------
// Cells of 65..kWideChunks*64 particles, three per block: waves 0..2 take one
// cell each, wave 3 separates the first nine particles of all three.
//
// Ranks 0..8 and ranks 9..n-1 of a cell claim their slots independently, each
// direction's share of either part as its own run.  Both runs are ascending
// in canonical index and every index of the first lies below every index of
// the second, so in the new cell they are two runs that do not interleave,
// which is all the order restoration of the next frame needs.  The cell
// waves therefore never wait for the Gauss-Seidel chain.

constexpr int kWideCells = kWavesPerBlock - 1;

struct WideStage {
    float4 q[kWideCells][kMaxInCell];  // ranks 0..8 of each cell: x, y, stored position (bits)
    float2 v[kWideCells][kMaxInCell];  // and their velocities
    uint2 sc[kWideCells];              // each cell's first particle and key
    uint32_t on[kWideCells];           // nonzero: wave 3 handles ranks 0..8
};

// Cell wave, before the block barrier: loads in stored order, canonical rank
// of every stored position (rk), ranks 0..8 (position, stored position and
// velocity) staged for wave 3.  The ranks
// come from the runs; a cell of more than kWideRuns runs is ranked pairwise
// (source indices in a cell are distinct).
__device__ __forceinline__ void wide_order(const FrameBuffers &b, uint32_t s, uint32_t n,
                                           WaveLds &L, float4 *stage, float2 *vstage,
                                           float2 (&pp)[kWideChunks],
                                           float2 (&vv)[kWideChunks],
                                           uint32_t (&rk)[kWideChunks]) {
    const uint32_t lane = lane_id();
    const unsigned long long le = lanemask_lt() | (1ull << lane);
    uint32_t sv[kWideChunks];
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q) {
        const uint32_t i = 64 * q + lane;
        sv[q] = 0;
        pp[q] = vv[q] = make_float2(0.0f, 0.0f);
        if (i < n) {
            sv[q] = b.src[s + i];
            pp[q] = b.pos[s + i];
            vv[q] = b.vel[s + i];
        }
    }
    // Runs: flagged particles and position 0 start one.
    uint32_t runs = 0, rid[kWideChunks];
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q) {
        const uint32_t i = 64 * q + lane;
        const bool start = i < n && (i == 0 || (sv[q] & kRunFlag));
        const unsigned long long m = __ballot(start);
        rid[q] = runs + (uint32_t)__popcll(m & le) - 1;
        if (start && rid[q] < (uint32_t)kWideRuns) {
            L.head_key[rid[q]] = sv[q] & kSrcMask;
            L.head_pos[rid[q]] = i;
        }
        runs += (uint32_t)__popcll(m);
    }
    if (runs <= (uint32_t)kWideRuns) {
        wave_sync();
        // Lane l owns runs l, l + 64, ...
        constexpr int kOwn = kWideRuns / 64;
#pragma unroll
        for (int k = 0; k < kOwn; ++k) {
            const uint32_t r = lane + 64 * k;
            if (r < runs) {
                const uint32_t end = r + 1 < runs ? L.head_pos[r + 1] : n;
                L.head_kl[r] = make_uint2(L.head_key[r], end - L.head_pos[r]);
            }
        }
        wave_sync();
        // Canonical rank of a run's start: particles in runs with a smaller
        // first source index.  Stored as rank - position so that
        // rank = run_off + i.
        uint32_t hk[kOwn], off[kOwn];
#pragma unroll
        for (int k = 0; k < kOwn; ++k) {
            const uint32_t r = lane + 64 * k;
            hk[k] = r < runs ? L.head_key[r] : 0u;
            off[k] = 0;
        }
        if (runs <= 64) {  // the common case: one run per lane at most
#pragma unroll 4
            for (uint32_t j = 0; j < runs; ++j) {
                const uint2 e = L.head_kl[j];
                off[0] += e.x < hk[0] ? e.y : 0u;
            }
        } else {
#pragma unroll 2
            for (uint32_t j = 0; j < runs; ++j) {
                const uint2 e = L.head_kl[j];
#pragma unroll
                for (int k = 0; k < kOwn; ++k) off[k] += e.x < hk[k] ? e.y : 0u;
            }
        }
#pragma unroll
        for (int k = 0; k < kOwn; ++k) {
            const uint32_t r = lane + 64 * k;
            if (r < runs) L.run_off[r] = off[k] - L.head_pos[r];
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < kWideChunks; ++q) {
            const uint32_t i = 64 * q + lane;
            rk[q] = i < n ? L.run_off[rid[q]] + i : 0xffffffffu;
        }
    } else {
        // Pairwise: rank = number of smaller source indices (code[] is free
        // until after the block barrier).
        wave_sync();
#pragma unroll
        for (int q = 0; q < kWideChunks; ++q) {
            const uint32_t i = 64 * q + lane;
            if (i < n) L.code[i] = sv[q] & kSrcMask;
            rk[q] = 0;
        }
        wave_sync();
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t o = L.code[j];
#pragma unroll
            for (int q = 0; q < kWideChunks; ++q) rk[q] += o < (sv[q] & kSrcMask);
        }
#pragma unroll
        for (int q = 0; q < kWideChunks; ++q)
            if (64 * q + lane >= n) rk[q] = 0xffffffffu;
    }
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q) {
        const uint32_t i = 64 * q + lane;
        if (i < n && rk[q] < (uint32_t)kMaxInCell) {
            stage[rk[q]] = make_float4(pp[q].x, pp[q].y, __uint_as_float(i), 0.0f);
            vstage[rk[q]] = vv[q];
        }
    }
}

// Cell wave, after the barrier: ranks 9..n-1.  Finish in stored order, write
// at s + rank, keep each rank's direction (or a far mover's key) in LDS; then
// in canonical order count per direction, one atomic per occupied direction,
// and the slots.
__device__ __forceinline__ void wide_rest(const FrameParams &p, const FrameBuffers &b, uint32_t s,
                                          uint32_t n, uint32_t c, WaveLds &L,
                                          const float2 (&pp)[kWideChunks],
                                          const float2 (&vv)[kWideChunks],
                                          const uint32_t (&rk)[kWideChunks]) {
    const uint32_t lane = lane_id();
    const unsigned long long lt = lanemask_lt();
    const Grid g = grid_of(p, c);
    uint32_t cx, cy;
    cell_coords(g, c, cx, cy);
    uint32_t *code = L.code;
    uint32_t fkey[kWideChunks];  // a farther mover's new key, or ~0
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q) {
        fkey[q] = 0xffffffffu;
        // rk is ~0 past the particles: one compare covers both bounds.
        if (rk[q] - (uint32_t)kMaxInCell < n - (uint32_t)kMaxInCell) {
            float px = pp[q].x, py = pp[q].y, vx = vv[q].x, vy = vv[q].y;
            finish_particle(p, g, px, py, vx, vy);
            b.pos_out[s + rk[q]] = make_float2(px, py);
            b.vel_out[s + rk[q]] = make_float2(vx, vy);
            uint32_t key;
            const uint32_t dir = direction(p, g, px, py, cx, cy, key);
            code[rk[q]] = dir < 9 ? (kRunFlag | (dir << 27)) : 0u;
            if (dir == 9) fkey[q] = key;
        }
    }
    wave_sync();
    // Canonical order, chunk by chunk: each near mover's rank in its
    // direction, written back into code[] (a rolled loop keeps the ballot
    // masks of one chunk live at a time).
    uint32_t count[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) count[d] = 0;
#pragma unroll 1
    for (uint32_t r0 = kMaxInCell; r0 < n; r0 += 64) {
        const uint32_t r = r0 + lane;
        const uint32_t w = r < n ? code[r] : 0u;
        const uint32_t dir = (w & kRunFlag) ? (w >> 27) & 15u : 10u;
        uint32_t rin = 0;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            const unsigned long long m = __ballot(dir == (uint32_t)d);
            if (dir == (uint32_t)d) rin = count[d] + (uint32_t)__popcll(m & lt);
            count[d] += (uint32_t)__popcll(m);
        }
        if (dir < 9) code[r] = kRunFlag | (dir << 27) | rin;
    }
    // Every atomic of the cell in one round: one per occupied direction
    // (lanes 0..8), one per farther mover.
    uint32_t mine = 0;
#pragma unroll
    for (int d = 0; d < 9; ++d) mine = lane == (uint32_t)d ? count[d] : mine;
    uint32_t got = 0, tkey = 0;
    if (lane < 9) {
        const uint32_t tx = cx + lane % 3 - 1, ty = cy + lane / 3 - 1;
        tkey = g.base + ty * g.gx + tx;
        if (mine > 0) got = claim(b, tkey, mine);
    }
    uint32_t fslot[kWideChunks];
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q)
        fslot[q] = fkey[q] != 0xffffffffu ? claim(b, fkey[q], 1u) : 0u;
    // Per direction: (new key, first slot), then the slots.
    if (lane < 9) L.dir_tbl[lane] = make_uint2(tkey, got);
    wave_sync();
#pragma unroll 1
    for (uint32_t r0 = kMaxInCell; r0 < n; r0 += 64) {
        const uint32_t r = r0 + lane;
        const uint32_t w = r < n ? code[r] : 0u;
        if (w & kRunFlag) {
            const uint32_t rin = w & 0x07ffffffu;
            const uint2 t = L.dir_tbl[(w >> 27) & 15u];
            b.kslot[s + r] = make_uint2(t.x, (t.y + rin) | (rin == 0 ? kRunFlag : 0u));
        }
    }
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q)
        if (fkey[q] != 0xffffffffu) b.kslot[s + rk[q]] = make_uint2(fkey[q], fslot[q] | kRunFlag);
}

// One block item: cells list[0..2], of which `count` exist.
__device__ __forceinline__ void wide_item(const FrameParams &p, const FrameBuffers &b,
                                          const uint4 *list, uint32_t count, WaveLds *lds,
                                          WideStage &st) {
    // readfirstlane: the compiler then knows the wave index is uniform and
    // branches on it instead of running both sides under lane masks.
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id();
    // Defined on every path, so that they are not carried around the item
    // loop (and kept live through wave 3's branch).
    float2 pp[kWideChunks], vv[kWideChunks];
    uint32_t rk[kWideChunks];
#pragma unroll
    for (int q = 0; q < kWideChunks; ++q) {
        pp[q] = vv[q] = make_float2(0.0f, 0.0f);
        rk[q] = 0;
    }
    uint4 cell = make_uint4(0, 0, 0, 0);
    bool mine = false, fast = false;
    __syncthreads();  // the previous item's stage has been read
    if (wave < (uint32_t)kWideCells) {
        const uint32_t e = wave;
        mine = e < count;
        if (mine) {
            cell = list[e];
            if (cell.y <= kWideMax) {  // always, by the class; guards a corrupt list
                wide_order(b, cell.x, cell.y, lds[wave], st.q[wave], st.v[wave], pp, vv, rk);
                fast = true;
            }
        }
        if (lane == 0) {
            st.on[wave] = fast ? 1u : 0u;
            st.sc[wave] = make_uint2(cell.x, cell.z);
        }
    }
    __syncthreads();
    if (wave == (uint32_t)kWideCells) {
        // Lanes 16g + j, j < 9: rank j of cell g.
        const uint32_t g = lane / 16, j = lane % 16;
        const uint32_t e = g;
        uint32_t s = 0, c = 0, m = 0;
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float2 v = make_float2(0.0f, 0.0f);
        if (g < (uint32_t)kWideCells && e < count && st.on[g]) {
            // The cell waves staged all of it: no global load before the
            // Gauss-Seidel chain.
            const uint2 sc = st.sc[g];
            s = sc.x;
            c = sc.y;
            m = kMaxInCell;
            if (j < (uint32_t)kMaxInCell) {
                q = st.q[g][j];
                v = st.v[g][j];
            }
        }
        ranked_group<16>(p, b, q.x, q.y, v.x, v.y, j, m, s, c, 16 * g);
    } else if (fast) {
        wide_rest(p, b, cell.x, cell.y, cell.z, lds[wave], pp, vv, rk);
    }
}

// Work items of the queued classes.  Class k's items take `per` cells each
// (64 / G for a group class, three for class 6, one for class 7) and are
// numbered slice by slice, then the overflow list.  ItemMap holds, per
// class, the inclusive prefix of items over the slices (in LDS), the slice
// fills and the totals.
struct ItemMap {
    uint32_t incl[kClasses][kSlices];
    uint32_t fill[kClasses][kSlices + 1];  // [kSlices]: the overflow list
    uint32_t in_slices[kClasses];          // items in the slices
    uint32_t items[kClasses];              // all items of the class
    uint32_t off[kClasses];                // class_off and slice_cap, kept here
    uint32_t scap[kClasses];               // rather than in scalar registers
};

__host__ __device__ constexpr uint32_t cells_per_item(int k) {
    return k < kGroupClasses ? (32u >> k) : k == kWaveClass ? (uint32_t)kWideCells : 1u;
}

// Fills the map: wave w takes classes w and w + 4, lane l slice l.
__device__ __forceinline__ void build_item_map(const FrameBuffers &b, ItemMap &m) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k = (int)wave + 4 * h;
        // constant indices into the kernel-argument arrays (a variable index
        // would copy them to scratch memory)
        uint32_t scap = 0, ccap = 0, off = 0;
#pragma unroll
        for (int kk = 0; kk < kClasses; ++kk) {
            scap = k == kk ? b.slice_cap[kk] : scap;
            ccap = k == kk ? b.class_cap[kk] : ccap;
            off = k == kk ? b.class_off[kk] : off;
        }
        const uint32_t per = cells_per_item(k);
        const uint32_t f = min(b.class_ctr[ctr_index(k, (int)lane)], scap);
        const uint32_t incl = wave_inclusive_sum((f + per - 1) / per);
        m.incl[k][lane] = incl;
        m.fill[k][lane] = f;
        if (lane == 63) {
            const uint32_t fo = min(b.class_ctr[ctr_index(k, kSlices)], ccap);
            m.fill[k][kSlices] = fo;
            m.in_slices[k] = incl;
            m.items[k] = incl + (fo + per - 1) / per;
            m.off[k] = off;
            m.scap[k] = scap;
        }
    }
}

// Item v (below m.items[k], the same in every lane) of class k: its first
// cell and how many cells of its list segment start there.
__device__ __forceinline__ const uint4 *locate(const FrameBuffers &b, const ItemMap &m, int k,
                                               uint32_t v, uint32_t &avail) {
    const uint32_t per = cells_per_item(k);
    uint32_t seg, local;
    if (v < m.in_slices[k]) {
        seg = (uint32_t)__popcll(__ballot(m.incl[k][lane_id()] <= v));
        local = v - (seg ? m.incl[k][seg - 1] : 0u);
    } else {
        seg = kSlices;
        local = v - m.in_slices[k];
    }
    avail = m.fill[k][seg] - local * per;
    return b.cells + m.off[k] + seg * m.scap[k] + local * per;
}

// Work items [start, stop) of the global numbering that belong to group
// class k (G = 2^(k+1) lanes per cell) and fall to wavefront gwave.
// Items of 2..8-lane groups are short chains of dependent loads and a few
// steps of work, and a wavefront takes several per frame before the
// collapse: their loads are pipelined, the next item's cells and particles
// in flight while the current one computes.
template <int G>
__device__ __forceinline__ void group_class(const FrameParams &p, const FrameBuffers &b,
                                            const ItemMap &m, int k, uint32_t start,
                                            uint32_t stop, uint32_t gwave, uint32_t nwaves) {
    uint32_t v = gwave;
    if (v < start) v += (start - v + nwaves - 1) / nwaves * nwaves;
    if (v >= stop) return;
    uint32_t avail;
    const uint4 *list = locate(b, m, k, v - start, avail);
    if (G > 8) {
        for (;;) {
            group_run<G>(p, b, group_load<G>(b, group_entry<G>(list, avail)));
            v += nwaves;
            if (v >= stop) break;
            list = locate(b, m, k, v - start, avail);
        }
        return;
    }
    // Entries run two items ahead, particles one.
    GroupItem cur = group_load<G>(b, group_entry<G>(list, avail));
    uint4 ent = make_uint4(0, 0, 0, 0);
    if (v + nwaves < stop) {
        list = locate(b, m, k, v + nwaves - start, avail);
        ent = group_entry<G>(list, avail);
    }
    for (;;) {
        const uint32_t vn = v + nwaves;
        const bool more = vn < stop;
        GroupItem nxt;
        if (more) nxt = group_load<G>(b, ent);
        if (vn + nwaves < stop) {
            list = locate(b, m, k, vn + nwaves - start, avail);
            ent = group_entry<G>(list, avail);
        }
        group_run<G>(p, b, cur);
        if (!more) break;
        cur = nxt;
        v = vn;
    }
}

// Every queued cell:
//  1. class 6 (65..kWideMax particles), three cells per block item, on the
//     first blocks;
//  2. the group classes from 64 lanes per cell down to 2, wavefront work
//     items spread over every wavefront, numbered from the first block
//     without a class-6 item so that those take group work first;
//  3. class 7 (over kWideMax particles), one block each (block-wide LDS sort,
//     or merge sort past kLdsSort).
__global__ __launch_bounds__(kCollideBlock) void cells_kernel(FrameParams p, FrameBuffers b) {
    __shared__ union {
        struct {
            WaveLds w[kWavesPerBlock];
            WideStage stage;
        } wide;
        unsigned long long sort[kLdsSort];
    } sh;
    __shared__ ItemMap map;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t gwave = blockIdx.x * kWavesPerBlock + wave;
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    build_item_map(b, map);
    __syncthreads();

    const uint32_t items6 = map.items[kWaveClass];
    for (uint32_t it = blockIdx.x; it < items6; it += gridDim.x) {
        uint32_t avail;
        const uint4 *list = locate(b, map, kWaveClass, it, avail);
        wide_item(p, b, list, avail, sh.wide.w, sh.wide.stage);
    }

    // end[k]: one past the last work item of group class k, largest first.
    uint32_t end[kGroupClasses];
    {
        uint32_t at = 0;
#pragma unroll
        for (int k = kGroupClasses - 1; k >= 0; --k) {
            at += map.items[k];
            end[k] = at;
        }
    }
    const uint32_t busy = (items6 < gridDim.x ? items6 : gridDim.x) * kWavesPerBlock;
    const uint32_t rw = gwave >= busy ? gwave - busy : gwave + nwaves - busy;
    // One loop per group class keeps each variant's registers to itself.
    group_class<64>(p, b, map, 5, 0, end[5], rw, nwaves);
    group_class<32>(p, b, map, 4, end[5], end[4], rw, nwaves);
    group_class<16>(p, b, map, 3, end[4], end[3], rw, nwaves);
    group_class<8>(p, b, map, 2, end[3], end[2], rw, nwaves);
    group_class<4>(p, b, map, 1, end[2], end[1], rw, nwaves);
    group_class<2>(p, b, map, 0, end[1], end[0], rw, nwaves);

    const uint32_t nb = map.items[kBlockClass];
    if (nb == 0) return;
    __syncthreads();
    for (uint32_t e = blockIdx.x; e < nb; e += gridDim.x) {
        uint32_t avail;
        const uint4 cell = *locate(b, map, kBlockClass, e, avail);
        big_cell<true>(p, b, cell.x, cell.y, cell.z, sh.sort);
    }
}

// Canonical copy of the state into pos_out/vel_out (readback): the same
// ordering code as the step, no arithmetic.
// The readback's canonical order: cells of up to kRegCell particles here,
// larger ones queued (b.canon_ctr, entries in b.cells, which the frame's
// cell lists leave free between frames) for canonical_big_kernel, one block
// per cell over the whole GPU: at the collapsed lattice about 2,000 floor
// cells of ~460 particles are neighbours, and a block that sorted every big
// cell among its own 256 would hold the readback for milliseconds.
__global__ __launch_bounds__(kCollideBlock) void canonical_kernel(FrameParams p, FrameBuffers b) {
    const uint32_t c = blockIdx.x * kCollideBlock + threadIdx.x;
    if (c >= p.ncell || *b.error) return;  // a dead context's offsets are not read
    const uint32_t s = b.idx[c + 1];
    uint32_t n = b.idx[c + 2] - s;
    clamp_range(p, s, n);
    if (n == 1) {
        one_particle(b, s);
    } else if (n == 2) {
        small_cell<2>(b, s, n);
    } else if (n > 2 && n <= 4) {
        small_cell<4>(b, s, n);
    } else if (n > 4 && n <= 8) {
        small_cell<8>(b, s, n);
    } else if (n > 8 && n <= (uint32_t)kRegCell) {
        small_cell<kRegCell>(b, s, n);
    } else if (n > (uint32_t)kRegCell) {
        const uint32_t q = atomicAdd(b.canon_ctr, 1u);
        b.cells[q] = make_uint4(s, n, c, 0u);
    }
}

__global__ __launch_bounds__(kCollideBlock) void canonical_big_kernel(FrameParams p, FrameBuffers b) {
    __shared__ unsigned long long sk[kLdsSort];
    const uint32_t total = *b.error ? 0u : *b.canon_ctr;
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        const uint4 e = b.cells[q];
        big_cell<false>(p, b, e.x, e.y, 0, sk);
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of the counts: one tile of THREADS x ITEMS cells per block; status
// words are tagged with the frame's epoch so they never need clearing.

constexpr unsigned long long kFlagAggregate = 1ull;
constexpr unsigned long long kFlagInclusive = 2ull;
constexpr uint32_t kEpochMask = 0x3fffffffu;
constexpr int kLookWindows = 4;  // 64-tile windows read per look-back round

__device__ __forceinline__ unsigned long long status_word(unsigned long long flag, uint32_t epoch,
                                                          uint32_t value) {
    return (flag << 62) | ((unsigned long long)(epoch & kEpochMask) << 32) | value;
}

// Exclusive prefix of `tile` from its predecessors' status words, by one
// wavefront.  Lane l of window w reads tile j - 64w - l.  A tile waits only
// on tiles dispatched before it, so the wait ends; should it not within
// p.wait_limit rounds, the tile raises kErrLookBack and takes 0 (the frame
// kernels clamp every offset range to the particles, so what follows stays
// in bounds, and the host reports the error).
__device__ uint32_t look_back(const FrameParams &p, const FrameBuffers &b, uint32_t tile,
                              uint32_t epoch) {
    const uint32_t lane = lane_id();
    uint32_t excl = 0;
    int64_t j = (int64_t)tile - 1;
    epoch = (epoch + p.epoch_skew) & kEpochMask;
    for (uint32_t round = 0;; ++round) {
        if (round >= p.wait_limit) {
            if (lane == 0) atomicOr(b.error, kErrLookBack);
            return 0u;
        }
        unsigned long long w[kLookWindows];
#pragma unroll
        for (int q = 0; q < kLookWindows; ++q) {
            const int64_t jj = j - 64 * q - (int64_t)lane;
            w[q] = jj >= 0 ? __hip_atomic_load(&b.tile_state[jj], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : status_word(kFlagInclusive, epoch, 0);
        }
        bool finished = false, stalled = false;
#pragma unroll
        for (int q = 0; q < kLookWindows; ++q) {
            if (!finished && !stalled) {
                const unsigned long long flag = w[q] >> 62;
                const bool ok = flag != 0 && (uint32_t)((w[q] >> 32) & kEpochMask) == epoch;
                const unsigned long long okm = __ballot(ok);
                const unsigned long long incm = __ballot(ok && flag == kFlagInclusive);
                const uint32_t val = (uint32_t)w[q];
                if (incm != 0) {
                    const int first = __builtin_ctzll(incm);
                    const unsigned long long need =
                        first == 63 ? ~0ull : ((1ull << (first + 1)) - 1);
                    if ((okm & need) == need) {
                        excl += wave_sum((int)lane <= first ? val : 0u);
                        finished = true;
                    } else {
                        stalled = true;
                    }
                } else if (okm == ~0ull) {
                    excl += wave_sum(val);
                    j -= 64;
                } else {
                    stalled = true;
                }
            }
        }
        if (finished) return excl;
        if (stalled) __builtin_amdgcn_s_sleep(1);
    }
}

// THREADS x ITEMS counts per tile (kScanShapes); ITEMS is a multiple of 4.
template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void scan_kernel(FrameParams p, FrameBuffers b) {
    static_assert(ITEMS % 4 == 0 && THREADS % 64 == 0, "128-bit count loads, whole wavefronts");
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_wave[THREADS / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63, wave = tid >> 6;
    // Tiles follow block order: a block waits only on lower-numbered blocks,
    // which were dispatched before it.
    const uint32_t tile = blockIdx.x;
    const uint32_t epoch = *b.epoch & kEpochMask;
    const uint32_t base = tile * (THREADS * ITEMS) + tid * ITEMS;
    // A chunk in which no slot was claimed this frame holds only zero counts:
    // its counts are not read and its offsets not written (the readback fills
    // them in from chunk_prefix).
    const uint32_t chunk = base >> kChunkShift;
    const bool live = base < p.ncell && b.chunk_next[chunk] != 0;

    // Load ITEMS counts and clear them for the next frame.
    uint32_t v[ITEMS];
    if (!live) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) v[i] = 0;
    } else if (base + ITEMS <= p.ncell) {
        uint4 *q = reinterpret_cast<uint4 *>(b.cnt + base);
#pragma unroll
        for (int i = 0; i < ITEMS / 4; ++i) {
            const uint4 w = q[i];
            v[4 * i] = w.x; v[4 * i + 1] = w.y; v[4 * i + 2] = w.z; v[4 * i + 3] = w.w;
        }
        // Clear only what holds counts: after the first frames most of the
        // grid is empty, and these stores are a third of the kernel's bytes.
#pragma unroll
        for (int i = 0; i < ITEMS / 4; ++i)
            if (v[4 * i] | v[4 * i + 1] | v[4 * i + 2] | v[4 * i + 3]) q[i] = make_uint4(0, 0, 0, 0);
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (base + i < p.ncell) {
                v[i] = b.cnt[base + i];
                b.cnt[base + i] = 0;
            } else {
                v[i] = 0;
            }
        }
    }
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) mine += v[i];
    const uint32_t incl = wave_inclusive_sum(mine);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) {
        const uint32_t ws = s_wave[w];
        if ((uint32_t)w < wave) wave_off += ws;
        agg += ws;
    }

    if (wave == 0) {
        uint32_t excl = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&b.tile_state[0], status_word(kFlagInclusive, epoch, agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&b.tile_state[tile], status_word(kFlagAggregate, epoch, agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            excl = look_back(p, b, tile, epoch);
            if (lane == 0)
                __hip_atomic_store(&b.tile_state[tile],
                                   status_word(kFlagInclusive, epoch, excl + agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_prefix = excl;
    }
    __syncthreads();

    // idx[1 + k] = start_k; idx[0] = 0; idx[C + 1] = N.
    uint32_t run = s_prefix + wave_off + incl - mine;
    uint32_t out[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        out[i] = run;
        run += v[i];
    }
    uint32_t *dst = b.idx + 1 + base;
    if (!live) {
    } else if (base + ITEMS <= p.ncell) {
        uint4 *q = reinterpret_cast<uint4 *>(dst);
#pragma unroll
        for (int i = 0; i < ITEMS / 4; ++i)
            q[i] = make_uint4(out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (base + i < p.ncell) dst[i] = out[i];
    }
    if ((base & (kChunkCells - 1)) == 0 && base < p.ncell) {  // the chunk's first thread
        b.chunk_prefix[chunk] = out[0];
        b.chunk_live[chunk] = live ? 1 : 0;
        b.chunk_next[chunk] = 0;  // every thread of the chunk read it before the barriers above
    }
    if (tile == 0 && tid == 0) {
        b.idx[0] = 0;
    }
    if (tile == 0)  // the cell kernel of this frame has finished with the lists
        for (uint32_t i = tid; i < (uint32_t)(kClasses * (kSlices + 1)); i += THREADS)
            b.class_ctr[i * kCtrStride] = 0;
    if (tile == b.scan_tiles - 1 && tid == 0) b.idx[p.ncell + 1] = s_prefix + agg;
}

// ---------------------------------------------------------------------------

// PER particles per lane (kScatterPerLane), strided across the wavefront: every
// load of the dependent chain (kslot -> start offset -> destination) is in
// flight for all of them at once.
template <int PER>
__global__ __launch_bounds__(kScatterBlock) void scatter_kernel(FrameParams p, FrameBuffers b) {
    // Blocks go to the 8 XCDs in turn (block b to XCD b % 8): block b takes
    // tile (b % 8) * (grid / 8) + b / 8, so that each XCD scatters one
    // contiguous eighth of the particles.  Their destinations then share
    // lines mostly with destinations of the same XCD, whose L2 merges the
    // partial lines before writing them back.  The grid is a multiple of 8.
    const uint32_t tile = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t w0 = (tile * kScatterBlock + (threadIdx.x & ~63u)) * PER;
    const uint32_t ln = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.epoch = *b.epoch + 1;  // the frame's scan has finished
    uint2 ks[PER];
    float2 q[PER], v[PER];
    uint32_t start[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t i = w0 + 64 * j + ln;
        if (i < p.n) {
            ks[j] = b.kslot[i];
            q[j] = b.pos_out[i];
            v[j] = b.vel_out[i];
        }
    }
    // A scan that gave up its look-back (b.error) left offsets that are not
    // valid: the scatter then writes nothing (a vector load beside the
    // others, checked before the stores).
    const uint32_t dead = b.error[lane_id() >> 6];
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (w0 + 64 * j + ln < p.n) start[j] = b.idx[ks[j].x + 1];
    if (__builtin_amdgcn_readfirstlane(dead)) return;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const uint32_t i = w0 + 64 * j + ln;
        if (i < p.n) {
            const uint32_t dest = start[j] + (ks[j].y & kSrcMask);
            if (dest < p.n) {  // always, with a consistent frame
                b.pos[dest] = q[j];
                b.vel[dest] = v[j];
                b.src[dest] = i | (ks[j].y & kRunFlag);
                b.key[dest] = ks[j].x;
            }
        }
    }
}

// Offsets of the chunks the last scan found empty: every entry is the offset
// at the chunk's first cell.
__global__ __launch_bounds__(256) void materialize_kernel(FrameParams p, FrameBuffers b) {
    const uint32_t chunk = blockIdx.x;
    if (b.chunk_live[chunk]) return;
    const uint32_t v = b.chunk_prefix[chunk];
    for (uint32_t i = threadIdx.x; i < kChunkCells; i += 256) {
        const uint32_t c = chunk * kChunkCells + i;
        if (c < p.ncell) b.idx[1 + c] = v;
    }
}

__global__ void fill32_kernel(uint32_t *dst, uint32_t v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

__global__ void copy32_kernel(uint32_t *dst, const uint32_t *src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void iota_kernel(uint32_t *dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

// key[i]: the last cell c whose start idx[c + 1] is at most i (offsets given
// whole, non-decreasing from idx[1] = 0, so that cell's range holds i).
__global__ void keys_from_offsets_kernel(const uint32_t *idx, uint32_t ncell, uint32_t n,
                                         uint32_t *key) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = ncell;  // idx[lo + 1] <= i < idx[hi + 1] (idx[ncell + 1] = n)
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (idx[mid + 1] <= i) lo = mid;
        else hi = mid;
    }
    key[i] = lo;
}

// ---------------------------------------------------------------------------
// Upload binning (SpatialBin::create_packed_data, spatial_bin.rs:103-149, on
// the device).  The CPU packer is a stable counting sort of the store by
// cell; here each particle takes a slot with one atomic, so a cell's
// particles land in arbitrary order, each marked as a run of its own.  The
// order inside a cell is then the one every frame already restores (runs
// ordered by source index, here the store index), and the readback's
// canonical pass applies it as well: the packed triple is the CPU packer's.

// Rust's `f as i32`: saturating, NaN -> 0.
__device__ __forceinline__ int32_t saturating_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 0x7fffffff;
    if (f <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)f;
}

// f32::div_euclid for b > 0 (spatial_bin.rs:48-64): the truncated IEEE
// quotient, one less when the remainder is negative.
__device__ __forceinline__ float div_euclid_pos(float a, float b) {
    const float q = truncf(a / b);
    return fmodf(a, b) < 0.0f ? q - 1.0f : q;
}

// The packer's cell of a store particle (spatial_bin.rs:48-64) in a grid of
// gx x gy cells; false outside it.
__device__ __forceinline__ bool store_cell(const UploadParams &u, float2 q, uint32_t gx, uint32_t gy,
                                           uint32_t &key) {
    // The cell relative to the grid's bottom-left cell (the view anchor's).
    const int64_t cx = (int64_t)saturating_i32(div_euclid_pos(q.x, u.cs)) - u.blx;
    const int64_t cy = (int64_t)saturating_i32(div_euclid_pos(q.y, u.cs)) - u.bly;
    const bool in = cx >= 0 && cy >= 0 && cx < (int64_t)gx && cy < (int64_t)gy;
    key = in ? (uint32_t)cy * gx + (uint32_t)cx : 0xffffffffu;
    return in;
}

__global__ __launch_bounds__(256) void upload_bin_kernel(UploadParams u, FrameBuffers b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (i < u.n) {
        uint2 ks = make_uint2(0xffffffffu, 0u);
        // Its world's grid in a batch: worlds of one size one after another,
        // or each world's own grid.
        uint32_t gx = u.gx, gy = u.gy, base = 0;
        if (u.world) {
            if (u.grids) {
                const WorldGrid g = u.grids[u.world[i]];
                gx = g.gx;
                gy = g.gy;
                base = g.base;
            } else {
                base = u.world[i] * u.world_cells;
            }
        }
        in = store_cell(u, u.pos[i], gx, gy, ks.x);
        if (in) {
            ks.x += base;
            ks.y = claim(b, ks.x, 1u);
        }
        u.ks[i] = ks;
    }
    const unsigned long long m = __ballot(in);
    if (lane_id() == 0 && m != 0) atomicAdd(u.valid, (uint32_t)__popcll(m));
}

// The store's particles outside the grid, appended at dpos/dvel[base + ...]
// (wrach_store_write_back) in their store order: a moved view can bring
// them into cells later, where that order is their in-cell order.  Two
// launches: each block counts its own (counts[block]); the host turns the
// counts into exclusive offsets; each block then writes its particles from
// its offset on, ranked within the block by ballots.
__global__ __launch_bounds__(256) void store_outside_count_kernel(UploadParams u, uint32_t *counts) {
    __shared__ uint32_t s_n[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false;
    if (i < u.n) {
        uint32_t key;
        out = !store_cell(u, u.pos[i], u.gx, u.gy, key);
    }
    const unsigned long long m = __ballot(out);
    if (lane_id() == 0) s_n[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
}

__global__ __launch_bounds__(256) void store_outside_kernel(UploadParams u, float2 *dpos,
                                                            float2 *dvel, uint32_t base,
                                                            const uint32_t *offsets) {
    __shared__ uint32_t s_n[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false;
    float2 q = make_float2(0.0f, 0.0f);
    if (i < u.n) {
        q = u.pos[i];
        uint32_t key;
        out = !store_cell(u, q, u.gx, u.gy, key);
    }
    const unsigned long long m = __ballot(out);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) s_n[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += s_n[w];
    if (out) {
        const uint32_t j = base + offsets[blockIdx.x] + before + (uint32_t)__popcll(m & lanemask_lt());
        dpos[j] = q;
        dvel[j] = u.vel[i];
    }
}

__global__ __launch_bounds__(256) void upload_scatter_kernel(UploadParams u, FrameBuffers b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.epoch = *b.epoch + 1;  // the scan has finished
    if (i >= u.n) return;
    const uint2 ks = u.ks[i];
    if (ks.x == 0xffffffffu) return;
    const uint32_t dest = b.idx[ks.x + 1] + ks.y;
    if (dest >= u.cap) return;  // never, with a consistent scan
    b.pos[dest] = u.pos[i];
    b.vel[dest] = u.vel[i];
    b.src[dest] = i | kRunFlag;
    b.key[dest] = ks.x;
}

// ---------------------------------------------------------------------------
// Mode S: separation passes over the 3x3 cells around each particle, before
// the reference step (an extension: the reference collides within one cell,
// shaders/physics/src/cell.rs:60-69).  A pass is Jacobi -- every particle
// reads the positions the pass started from -- and a particle's pair terms
// are rounded to multiples of 2^-30 and summed exactly, so its
// displacement does not depend on the order of the particles inside a cell
// (which the device layout keeps canonical only through `src`).

// A term is below 0.62 in magnitude (|dx| / d is at most 1 + 2^-22, or
// 1.22 where dx * dx is subnormal), so term * 2^30 is below 2^31 and
// rintf (round half to even, as llrintf) makes it an integer-valued float.
// The terms are summed as doubles: integers below 2^53 add exactly, and at
// most kStencilBudget of them stay far below that, so the double holds the
// same integer as an int64 sum, in two instructions a term in place of
// five (conversion, sign extension, 64-bit add).
constexpr float kStencilScale = 1073741824.0f;            // 2^30
constexpr float kStencilUnit = 9.313225746154785e-10f;    // 2^-30

// sqrtf(x) for x in [2^-96, 1 + 2^-23]: the hardware reciprocal root y,
// s = x * y, and one Newton step on the residual, fma(x - s * s, y / 2, s).
// That this equals the correctly rounded root for every x of the range is
// a property of the hardware reciprocal root, not a theorem;
// check_pair_arithmetic_kernel tests it for all 805,306,370 values of x on
// the device in use (wrach_check_pair_arithmetic).  hipcc's sqrtf takes
// v_sqrt_f32 (one unit in the last place either way), corrects it with
// two residual tests, and scales inputs below 2^-96.
__device__ __forceinline__ float sqrt_rn_above_2m96(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float s = x * y;
    const float e = __builtin_fmaf(-s, s, x);
    return __builtin_fmaf(e, 0.5f * y, s);
}

// (1 - d) * 2^29 / d correctly rounded, for d in [2^-48, 1] (so 1 - d is
// the overlap and the numerator a = (1 - d) * 2^29 is 0 or in
// [2^5, 2^29]): the hardware reciprocal y of d, the quotient a * y and one
// residual correction fma(fma(-d, q, a), y, q).  hipcc's f32 division adds
// a Newton step on y, a second correction, and v_div_scale / v_div_fixup
// (identities here: the exponent difference stays below 96, nothing is
// subnormal, and the fix-up only rewrites special operands).  That the
// short sequence equals the correctly rounded quotient for every d of the
// range is not a theorem but a property of the hardware reciprocal;
// check_pair_arithmetic_kernel tests it for all 402,653,185 values of d
// on the device in use (wrach_check_pair_arithmetic).
__device__ __forceinline__ float overlap_force_rn(float d) {
    const float a = (1.0f - d) * 0x1p29f;
    const float y = __builtin_amdgcn_rcpf(d);
    const float q = a * y;
    const float r = __builtin_fmaf(-d, q, a);
    return __builtin_fmaf(r, y, q);
}


// The pair arithmetic's short sequences against hipcc's correctly rounded
// ones over their whole input ranges: sqrt_rn_above_2m96(x) against
// sqrtf(x) for every float x in [2^-96, 1 + 2^-23], overlap_force_rn(d)
// against ((1 - d) * 2^29) / d for every float d in [2^-48, 1].  Adds the
// mismatches to out[0] and out[1] and the inputs tried to out[2] and
// out[3] (one atomic per wavefront and counter).
__global__ __launch_bounds__(256) void check_pair_arithmetic_kernel(unsigned long long *out) {
    const uint32_t s_lo = __float_as_uint(0x1p-96f), s_hi = __float_as_uint(1.00000011920928955078125f);
    const uint32_t d_lo = __float_as_uint(0x1p-48f), d_hi = __float_as_uint(1.0f);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bad_s = 0, bad_d = 0, n_s = 0, n_d = 0;
    for (uint64_t u = s_lo + t0; u <= s_hi; u += stride) {
        const float x = __uint_as_float((uint32_t)u);
        bad_s += __float_as_uint(sqrt_rn_above_2m96(x)) != __float_as_uint(sqrtf(x));
        ++n_s;
    }
    for (uint64_t u = d_lo + t0; u <= d_hi; u += stride) {
        const float d = __uint_as_float((uint32_t)u);
        const float full = ((1.0f - d) * 0x1p29f) / d;
        bad_d += __float_as_uint(overlap_force_rn(d)) != __float_as_uint(full);
        ++n_d;
    }
    const uint32_t v[4] = {wave_sum(bad_s), wave_sum(bad_d), wave_sum(n_s), wave_sum(n_d)};
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k)
            if (v[k]) atomicAdd(&out[k], (unsigned long long)v[k]);
}

// The particles of the 3x3 cells around (cx, cy) as three ranges, one per
// row cy - 1 .. cy + 1: the cells of one row are contiguous in the packed
// layout, and [start(x0), start(x1 + 1)) covers columns x0..x1 (past the
// row's last column, start(x1 + 1) is the next row's first cell, which
// begins where the row ends).
__device__ __forceinline__ void stencil_rows(const FrameParams &p, const FrameBuffers &b,
                                             const Grid &g, int cx, int cy, uint32_t rs[3],
                                             uint32_t re[3]) {
    const int gx = (int)g.gx, gy = (int)g.gy;
    const uint32_t base = g.base;
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < gx ? cx + 1 : gx - 1;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int y = cy + t - 1;
        rs[t] = re[t] = 0;
        if (y < 0 || y >= gy) continue;  // the world's own rows only
        rs[t] = cell_start(p, b, base + (uint32_t)(y * gx + x0));
        re[t] = cell_start(p, b, base + (uint32_t)(y * gx + x1 + 1));
        if (re[t] < rs[t]) re[t] = rs[t];
    }
}

// The grid of the world that holds packed particle i: the last world whose
// first cell starts at or before i (an empty world starts where the next
// one does, so it is never the last such).  One world: its own grid.
__device__ __forceinline__ Grid particle_grid(const FrameParams &p, const FrameBuffers &b, uint32_t i) {
    if (p.grids) {
        uint32_t lo = 0, hi = p.n_grids;  // world lo starts at or before i
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (cell_start(p, b, p.grids[mid].base) <= i) lo = mid;
            else hi = mid;
        }
        return grid_of(p, p.grids[lo].base);
    }
    if (p.world_cells == p.ncell) return grid_of(p, 0u);
    uint32_t lo = 0, hi = p.ncell / p.world_cells;
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (cell_start(p, b, mid * p.world_cells) <= i) lo = mid;
        else hi = mid;
    }
    return grid_of(p, lo * p.world_cells);
}

// collide_pair's rule for particle i alone, given dx = x_i - x_j and
// dy = y_i - y_j: whether j is within reach, the overlap 1 - d (a multiple
// of 2^-24 in [0, 1]) and F = f * 2^30 for f = (0.5 * (1 - d)) / d.  The
// reference subtracts (p_j - p_i) * f from p_i (particles.rs:65-92), which
// is exactly (dx, dy) * f added: p_j - p_i = -(p_i - p_j) in IEEE
// arithmetic, up to the sign of a zero, which the rounding of the term to
// 2^-30 units drops.  That term, rint((dx * f) * 2^30), is rint(dx * F):
// f is 0 or at least 2^-25 (normal), so F is exactly f * 2^30, and
// dx * f * 2^30 rounds as dx * f does, scaled, unless dx * f is subnormal,
// where both round to 0.
// `other`: j is not i (tested with the distance, so that the caller's load
// of p_j need not wait for it).
__device__ __forceinline__ bool stencil_reach(float dx, float dy, bool other, float &ov, float &F) {
    float sq = dx * dx;
    sq = sq + dy * dy;
    // sqrtf(sq) > 1 exactly when sq > 1 + 2^-23: sqrtf(1 + 2^-23) rounds to
    // 1 and the root of the next float above it does not, so d <= 1 below.
    // Most neighbours of a crowded cell are out of reach; they skip the root
    // and the division with the same outcome.
    if (!other | (sq > 1.00000011920928955078125f)) return false;
    if (sq < 0x1p-96f) {  // d below 2^-48, or 0: the full sequences
        float d = sqrtf(sq);
        if (d == 0.0f) d = 0.0001f;
        ov = 1.0f - d;
        F = ((0.5f * ov) / d) * kStencilScale;
        return true;
    }
    const float d = sqrt_rn_above_2m96(sq);  // in [2^-48, 1]: never 0
    ov = 1.0f - d;
    // d is in [2^-48, 1]: overlap_force_rn's range, and the quotient is
    // (0.5 * ov / d) * 2^30 exactly, as f is normal.
    F = overlap_force_rn(d);
    return true;
}

// One neighbour of particle i (`other`: not i itself) for stencil_particle.
// ov is a multiple of 2^-24 in [0, 1], so a lane's sum of at most
// kStencilBudget of them is a multiple of 2^-24 below 2^12: exact in a
// double, and the sum times 2^24 an integer below 2^36.
template <bool MOVE, bool SUMS>
__device__ __forceinline__ void stencil_visit(bool other, float xi, float yi, float2 q,
                                              uint32_t &worst, uint32_t &pairs,
                                              double &osum, double &ax, double &ay) {
    const float dx = xi - q.x, dy = yi - q.y;
    float ov, F;
    if (!stencil_reach(dx, dy, other, ov, F)) return;
    worst = max(worst, __float_as_uint(ov));
    pairs += 1;
    if (SUMS) osum += (double)ov;
    if (MOVE) {
        ax += (double)rintf(dx * F);
        ay += (double)rintf(dy * F);
    }
}

// Particle i (in cell (cx, cy) at the frame's start) against the 3x3 cells
// around it: the largest overlap, the number of overlapping pairs, with
// MOVE the sum of its pair terms in 2^-30 units, and with SUMS the sum of
// the overlaps in 2^-24 units.  Over kStencilBudget particles in those
// cells: none of these (a bound on the work of one lane).
template <bool MOVE, bool SUMS, bool HALF>
__device__ __forceinline__ float stencil_particle(const FrameParams &p, const FrameBuffers &b,
                                                  const float2 *pin, uint32_t i, uint32_t k,
                                                  float xi, float yi, double &ax, double &ay,
                                                  uint32_t &pairs, double &osum, bool &over) {
    uint32_t rs[3], re[3];
    const Grid g = grid_of(p, k);
    uint32_t cx, cy;
    cell_coords(g, k, cx, cy);
    stencil_rows(p, b, g, (int)cx, (int)cy, rs, re);
    const uint64_t total = (uint64_t)(re[0] - rs[0]) + (re[1] - rs[1]) + (re[2] - rs[2]);
    if (total > kStencilBudget) {
        over = true;
        return 0.0f;
    }
    uint32_t worst = 0;  // the largest overlap's bits (overlaps lie in [0, 1])
    // HALF: each pair once.  A particle takes its pairs with the later
    // particles of its own cell, with the cell to its right and with the
    // three cells above; every other pair of its 3x3 cells is one of these
    // seen from the other particle.
    if (HALF) rs[1] = min(max(rs[1], cell_start(p, b, k)), re[1]);
    for (int t = HALF ? 1 : 0; t < 3; ++t) {
        uint32_t j = rs[t];
        const uint32_t e = re[t];
        // Four neighbours a step, all loaded before the first is used.
        for (; j + 4 <= e; j += 4) {
            // One address, four loads at constant offsets from it.
            const float2 *qp = pin + j;
            float2 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = qp[u];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                stencil_visit<MOVE, SUMS>(HALF && t == 1 ? j + u > i : j + u != i, xi, yi, q[u], worst, pairs,
                                          osum, ax, ay);
        }
        for (; j < e; ++j)
            stencil_visit<MOVE, SUMS>(HALF && t == 1 ? j > i : j != i, xi, yi, pin[j], worst, pairs, osum, ax, ay);
    }
    return __uint_as_float(worst);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Raises *w to the block's largest v -- one atomic per block, and none when
// the word already holds as much (every block hitting one word serialises)
// -- and, with sums, adds the block's overlap sum and pair count to its
// slot (blockIdx % kStencilSumSlots) of them.
__device__ __forceinline__ void block_report(uint32_t *w, float v, unsigned long long *sums,
                                             unsigned long long osum, unsigned long long ocnt) {
    __shared__ float s_max[4];
    __shared__ unsigned long long s_sum[4], s_cnt[4];
    v = wave_max(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        osum += __shfl_xor(osum, o, 64);
        ocnt += __shfl_xor(ocnt, o, 64);
    }
    if (lane_id() == 0) {
        s_max[threadIdx.x >> 6] = v;
        s_sum[threadIdx.x >> 6] = osum;
        s_cnt[threadIdx.x >> 6] = ocnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        if (m > 0.0f && __float_as_uint(m) > __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(w, __float_as_uint(m));
        const unsigned long long cnt = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        if (sums && cnt) {
            unsigned long long *slot = sums + (blockIdx.x % kStencilSumSlots) * kStencilSumStride;
            atomicAdd(slot, s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3]);
            atomicAdd(slot + 1, cnt);
        }
    }
}

// The largest overlap at positions a.pin, raised into a.words[a.in_slot]
// (non-negative floats order as their bits).  a.first: the frame-start
// positions, whose cells every pass of the frame uses.
// Particle i's cell at the frame's start: derived from its position by the
// frame's first stencil kernel (a.first), which records it; read after.
__device__ __forceinline__ uint32_t stencil_cell(const FrameParams &p, const FrameBuffers &b,
                                                 const StencilArgs &a, uint32_t i, float2 q) {
    if (!a.first) return a.cell[i];
    const uint32_t k = cell_key(p, particle_grid(p, b, i), q.x, q.y);
    a.cell[i] = k;
    return k;
}

__device__ __forceinline__ unsigned long long *stencil_sums(const StencilArgs &a) {
    return a.stat ? a.sums + (a.stat - 1) * kStencilSumSlots * kStencilSumStride : nullptr;
}

// The learned frame's fallback: whether the warm start's mean overlap
// (sums set 2) exceeds the frame start's (set 0), compared exactly as
// sum_w * n_0 > sum_0 * n_w in 128 bits.  Every block of the move computes
// it from the same words; wavefront 0 reads the 64 slots.
__device__ __forceinline__ bool warm_rejected(const StencilArgs &a) {
    __shared__ uint32_t s_rej;
    if (threadIdx.x < 64) {
        const unsigned long long *s0 = a.sums + (size_t)threadIdx.x * kStencilSumStride;
        const unsigned long long *s2 = a.sums + (size_t)(2 * kStencilSumSlots + threadIdx.x) * kStencilSumStride;
        unsigned long long v[4] = {s0[0], s0[1], s2[0], s2[1]};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        if (threadIdx.x == 0) {
            const unsigned __int128 lhs = (unsigned __int128)v[2] * v[1];
            const unsigned __int128 rhs = (unsigned __int128)v[0] * v[3];
            s_rej = lhs > rhs ? 1u : 0u;
        }
    }
    __syncthreads();
    return s_rej != 0;
}

// The particle of lane t (see StencilArgs::perm).
__device__ __forceinline__ uint32_t lane_particle(const FrameParams &p, const StencilArgs &a, uint32_t t) {
    return a.perm && t < p.n ? a.perm[t] : t;
}

__global__ __launch_bounds__(256) void stencil_measure_kernel(FrameParams p, FrameBuffers b,
                                                              StencilArgs a) {
    const uint32_t i = lane_particle(p, a, blockIdx.x * 256 + threadIdx.x);
    float worst = 0.0f;
    uint32_t pairs = 0;
    double osum = 0.0;  // overlaps: multiples of 2^-24, at most kStencilBudget of them
    bool over = false;
    const bool half = a.half && a.words[kOverBudgetWord] == 0;  // the whole grid alike
    if (i < p.n) {
        const float2 q = a.pin[i];
        const uint32_t k = stencil_cell(p, b, a, i, q);
        double ax = 0.0, ay = 0.0;
        if (k < p.ncell) {
            if (half)
                worst = a.stat ? stencil_particle<false, true, true>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over)
                               : stencil_particle<false, false, true>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over);
            else
                worst = a.stat ? stencil_particle<false, true, false>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over)
                               : stencil_particle<false, false, false>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over);
        }
    }
    if (over) atomicOr(&a.words[kOverBudgetWord], 1u);
    const unsigned long long sides = half ? 2ull : 1ull;  // each pair seen once or from both sides
    block_report(&a.words[a.in_slot], worst, stencil_sums(a),
                 sides * (unsigned long long)(osum * (double)kOverlapScale), sides * pairs);
}

// One pass: a.pin -> a.pout.  With a.check it runs only while the
// overlap at its start (a.words[a.in_slot], measured before, the same for
// the whole grid) exceeds the tolerance, and is a copy otherwise.  Either
// way it raises a.words[a.in_slot] to the largest overlap of its input, so
// that a frame without a check (tolerance 0, where a pass over no overlap
// moves nothing) needs no measure between passes.
__global__ __launch_bounds__(256) void stencil_move_kernel(FrameParams p, FrameBuffers b,
                                                           StencilArgs a) {
    const uint32_t i = lane_particle(p, a, blockIdx.x * 256 + threadIdx.x);
    if (a.check && !(__uint_as_float(a.words[a.in_slot]) > a.tol)) {
        if (i < p.n) a.pout[i] = a.pin[i];
        return;  // the whole grid: no barrier follows
    }
    float worst = 0.0f;
    double ax = 0.0, ay = 0.0;  // integers: the sums of the terms in 2^-30 units
    uint32_t pairs = 0;
    double osum = 0.0;  // overlaps: multiples of 2^-24, at most kStencilBudget of them
    bool over = false;
    float2 q = make_float2(0.0f, 0.0f);
    if (a.pin_alt) {
        const bool rej = warm_rejected(a);
        if (blockIdx.x == 0 && threadIdx.x == 0) a.words[kFallbackWord] = rej ? 1u : 0u;
        if (rej) a.pin = a.pin_alt;  // a copy of the arguments: the whole grid decides alike
    }
    if (i < p.n) {
        q = a.pin[i];
        const uint32_t k = stencil_cell(p, b, a, i, q);
        if (k < p.ncell)
            worst = a.stat ? stencil_particle<true, true, false>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over)
                           : stencil_particle<true, false, false>(p, b, a.pin, i, k, q.x, q.y, ax, ay, pairs, osum, over);
    }
    if (over) atomicOr(&a.words[kOverBudgetWord], 1u);
    block_report(&a.words[a.in_slot], worst, stencil_sums(a),
                 (unsigned long long)(osum * (double)kOverlapScale), pairs);
    if (i >= p.n) return;
    // The mean of the pair corrections (constraint averaging): a particle
    // moves at most 0.5 per pass.
    float dx = (float)ax * kStencilUnit;  // the integer rounded once to f32
    float dy = (float)ay * kStencilUnit;
    if (pairs > 1) {
        dx = dx / (float)pairs;
        dy = dy / (float)pairs;
    }
    a.pout[i] = make_float2(q.x + dx, q.y + dy);
}

__global__ __launch_bounds__(256) void stencil_clear_kernel(uint32_t *w, unsigned long long *sums) {
    for (uint32_t i = threadIdx.x; i < kStencilWords; i += 256) w[i] = 0;
    for (uint32_t i = threadIdx.x; i < kStencilSumSets * kStencilSumSlots; i += 256) {
        sums[i * kStencilSumStride] = 0;
        sums[i * kStencilSumStride + 1] = 0;
    }
}

__global__ __launch_bounds__(256) void copy_pos_kernel(float2 *dst, const float2 *src, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// The lane order of the separation kernels.  A pass's sums are exact and
// each lane writes its own particle, so any order gives the same result;
// this one puts each crowded cell's particles in order of x (columns of
// particles are the common crowd), so that the 64 lanes of a wavefront lie
// close together and more of their neighbours are out of every lane's
// reach at once.  Cells of kCrowdMin..kCrowdMax particles, one thread per
// cell, one atomic per wavefront; the grid also sets perm to the identity.
__global__ __launch_bounds__(256) void crowd_list_kernel(FrameParams p, FrameBuffers b, StencilArgs a) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    for (uint64_t t = c; t < p.n; t += (uint64_t)gridDim.x * 256) a.perm[t] = (uint32_t)t;
    bool crowded = false;
    if (c < p.ncell) {
        const uint32_t s0 = cell_start(p, b, c), s1 = cell_start(p, b, c + 1);
        crowded = s1 >= s0 && s1 - s0 >= kCrowdMin && s1 - s0 <= kCrowdMax;
    }
    const unsigned long long m = __ballot(crowded);
    if (m == 0) return;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(&a.words[kCrowdCountWord], (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, 64);
    if (crowded) {
        const uint32_t slot = base + (uint32_t)__popcll(m & lanemask_lt());
        if (slot < a.crowd_cap) a.crowd[slot] = c;
    }
}

constexpr uint32_t kOrderSlices = 64;    // x slices of a crowded cell: one per lane
constexpr uint32_t kCrowdOrderGrid = 1024;  // blocks of crowd_order_kernel (4 wavefronts each)

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

// One wavefront per listed cell (the wavefronts stride over the list): the
// cell's particles counted into kOrderSlices equal slices of their x range
// (LDS atomics give each its rank in its slice), the slices' starts by a
// wavefront scan, then perm over the cell's range in slice order; inside a
// slice the order is the atomics'.  Slices of 1/64 of the range are far
// finer than the 64 lanes of a wavefront need.  perm holds the identity
// before.
__global__ __launch_bounds__(256) void crowd_order_kernel(FrameParams p, FrameBuffers b, StencilArgs a,
                                                          const float2 *pos) {
    __shared__ uint32_t slice_count[4][kOrderSlices];
    __shared__ uint32_t tag[4][kCrowdMax];  // a particle's slice << 16 | its rank in the slice
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id();
    const uint32_t listed = min(a.words[kCrowdCountWord], a.crowd_cap);
    for (uint32_t e = blockIdx.x * 4 + wave; e < listed; e += gridDim.x * 4) {
        const uint32_t c = a.crowd[e];
        const uint32_t s0 = cell_start(p, b, c), s1 = cell_start(p, b, c + 1);
        const uint32_t n = s1 > s0 ? s1 - s0 : 0u;
        if (n > kCrowdMax) continue;  // never: the list holds cells of at most kCrowdMax
        float lo = INFINITY, hi = -INFINITY;
        for (uint32_t r = lane; r < n; r += 64) {
            const float x = pos[s0 + r].x;
            lo = fminf(lo, x);
            hi = fmaxf(hi, x);
        }
        lo = wave_min(lo);
        hi = wave_max(hi);
        // Not finite (an infinite range): every particle in slice 0.
        const float scale = hi > lo ? (float)kOrderSlices / (hi - lo) : 0.0f;
        slice_count[wave][lane] = 0;
        wave_sync();
        for (uint32_t r = lane; r < n; r += 64) {
            // fmaxf takes 0 for a NaN; the slice is at most kOrderSlices - 1.
            const float f = fminf(fmaxf((pos[s0 + r].x - lo) * scale, 0.0f), (float)(kOrderSlices - 1));
            const uint32_t sl = (uint32_t)f;
            const uint32_t rank = atomicAdd(&slice_count[wave][sl], 1u);
            tag[wave][r] = sl << 16 | rank;
        }
        wave_sync();
        const uint32_t cnt = slice_count[wave][lane];
        wave_sync();
        slice_count[wave][lane] = wave_inclusive_sum(cnt) - cnt;
        wave_sync();
        for (uint32_t r = lane; r < n; r += 64) {
            const uint32_t t = tag[wave][r];
            const uint32_t slot = slice_count[wave][t >> 16] + (t & 0xffffu);
            if (slot < n) a.perm[s0 + slot] = s0 + r;  // always: the counts sum to n
        }
        wave_sync();  // this cell's LDS reads are done before the next cell's writes
    }
}

// ---------------------------------------------------------------------------
// Neighbour features of the learned warm start (an extension with no
// counterpart in the reference).
//
// Particle i's candidates are the other particles of the 5x5 cells around
// its cell (two cells each way, the world's own cells only), five row ranges
// of the packed layout.  It keeps the 16 nearest by (squared distance, packed
// index) -- the packed index is canonical, so the choice is unique -- and its
// feature row holds, for the r-th nearest, (x_j - x_i, y_j - y_i,
// vx_j - vx_i, vy_j - vy_i) rounded to bf16 (nearest even) at entries
// 4r..4r+3; fewer than 16 candidates leave zeros.  More than kFeatBudget
// candidates: an all-zero row (a bound on one lane's work).
//
// With a.sums it also adds the overlaps of i's pairs over the 3x3 cells
// (the pairs a separation pass visits, under the pass's budget) to the
// stencil sums, as stencil_particle does: every pair closer than 1 lies in
// the 3x3 cells, and the 3x3 ranges are tracked exactly inside the 5x5 ones.

// f32 -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Candidate order: key = (squared distance bits << 32) | packed index; equal
// distances are ordered by the bits of the candidates' (x, y, vx, vy) as
// unsigned integers, not by their index, so that the choice does not depend
// on the order of the particles inside a cell (two candidates equal in all
// four are interchangeable).  ~0 (no candidate) is last.
__device__ __forceinline__ bool cand_less(const float2 *pos, const float2 *vel, unsigned long long x,
                                          unsigned long long y) {
    const uint32_t hx = (uint32_t)(x >> 32), hy = (uint32_t)(y >> 32);
    if (hx != hy) return hx < hy;
    if (x == y) return false;
    const uint32_t jx = (uint32_t)x, jy = (uint32_t)y;
    const float2 px = pos[jx], py = pos[jy], vx = vel[jx], vy = vel[jy];
    const uint32_t a[4] = {__float_as_uint(px.x), __float_as_uint(px.y), __float_as_uint(vx.x),
                           __float_as_uint(vx.y)};
    const uint32_t c[4] = {__float_as_uint(py.x), __float_as_uint(py.y), __float_as_uint(vy.x),
                           __float_as_uint(vy.y)};
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (a[t] != c[t]) return a[t] < c[t];
    return false;
}

constexpr int kSortLog2 = 5;
constexpr int kSortCands = 1 << kSortLog2;  // candidates sorted by the network

// Inserts candidate c into the ascending list k[0..15] (dropping the last).
// Squared distances are finite, so a real key's high word never equals the
// empty key's (0xffffffff), and cand_less reads particles only for real keys.
__device__ __forceinline__ void insert16(const float2 *pos, const float2 *vel, unsigned long long (&k)[kNbrs],
                                         unsigned long long c) {
    if (!cand_less(pos, vel, c, k[kNbrs - 1])) return;
    k[kNbrs - 1] = c;
#pragma unroll
    for (int t = kNbrs - 1; t > 0; --t) {
        const unsigned long long lo = k[t - 1], hi = k[t];
        const bool sw = cand_less(pos, vel, hi, lo);
        k[t - 1] = sw ? hi : lo;
        k[t] = sw ? lo : hi;
    }
}

__global__ __launch_bounds__(256) void learn_features_kernel(FrameParams p, FrameBuffers b,
                                                             FeatureArgs a) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long osum = 0, ocnt = 0;
    if (i < p.n) {
        const float2 q = a.pos[i];
        const float2 w = a.vel[i];
        const Grid g = particle_grid(p, b, i);
        const uint32_t base = g.base;
        const uint32_t k = cell_key(p, g, q.x, q.y);
        if (a.cell) a.cell[i] = k;
        uint32_t ucx, ucy;
        cell_coords(g, k, ucx, ucy);
        const int cx = (int)ucx, cy = (int)ucy;
        const int gx = (int)g.gx, gy = (int)g.gy;
        const int x0 = cx > 1 ? cx - 2 : 0, x1 = cx + 2 < gx ? cx + 2 : gx - 1;
        const int i0 = cx > 0 ? cx - 1 : 0, i1 = cx + 1 < gx ? cx + 1 : gx - 1;
        uint32_t rs[5], re[5], is[5], ie[5];
        uint64_t tot5 = 0, tot3 = 0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            const int y = cy + t - 2;
            rs[t] = re[t] = is[t] = ie[t] = 0;
            if (y < 0 || y >= gy) continue;
            const uint32_t row = base + (uint32_t)(y * gx);
            rs[t] = cell_start(p, b, row + (uint32_t)x0);
            re[t] = cell_start(p, b, row + (uint32_t)x1 + 1);
            if (re[t] < rs[t]) re[t] = rs[t];
            tot5 += re[t] - rs[t];
            if (t >= 1 && t <= 3) {
                is[t] = cell_start(p, b, row + (uint32_t)i0);
                ie[t] = cell_start(p, b, row + (uint32_t)i1 + 1);
                if (ie[t] < is[t]) ie[t] = is[t];
                tot3 += ie[t] - is[t];
            }
        }
        const bool feats = tot5 <= kFeatBudget;
        const bool sums = a.sums != nullptr && tot3 <= kStencilBudget;
        // The overlap sums: the 3x3 cells, as a separation pass visits them.
        if (sums) {
#pragma unroll
            for (int t = 1; t <= 3; ++t) {
                for (uint32_t j = is[t]; j < ie[t]; ++j) {
                    if (j == i) continue;
                    const float2 o = a.pos[j];
                    float ov, f;
                    if (stencil_reach(q.x - o.x, q.y - o.y, true, ov, f)) {
                        osum += (uint32_t)rintf(ov * kOverlapScale);
                        ocnt += 1;
                    }
                }
            }
        }
        // The 16 nearest.  The first kSortCands candidates are ordered in
        // registers by Batcher's odd-even merge network for 32 keys, pruned
        // to the 174 comparators that decide the first 17 places
        // (select_nets.h), comparing the 64-bit key (squared distance bits,
        // packed index) as one integer.  That is the exact order unless two
        // of the first 17 share a distance: a wavefront where any lane finds
        // such a tie orders its keys again with the exact comparator (a
        // bitonic network of 240 compare-exchanges over cand_less).  Any
        // further candidates are inserted into the best 16.  (Networks for
        // 24 and 28 keys beside it, chosen per wavefront, took 137-165
        // VGPRs: one wavefront less per SIMD.)
        unsigned long long key[kNbrs];
#pragma unroll
        for (int r = 0; r < kNbrs; ++r) key[r] = ~0ull;
        uint32_t cum[5], m5 = 0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            cum[t] = m5;
            m5 += re[t] - rs[t];
        }
        unsigned long long ck[kSortCands];
        bool tie = false;
        if (feats) {
#pragma unroll
            for (int u = 0; u < kSortCands; ++u) {
                ck[u] = ~0ull;
                if ((uint32_t)u < m5) {
                    uint32_t j = 0;
#pragma unroll
                    for (int t = 0; t < 5; ++t)
                        if ((uint32_t)u >= cum[t] && (uint32_t)u < cum[t] + (re[t] - rs[t]))
                            j = rs[t] + ((uint32_t)u - cum[t]);
                    if (j != i) {
                        const float2 o = a.pos[j];
                        const float dx = o.x - q.x;
                        const float dy = o.y - q.y;
                        float sq = dx * dx;
                        sq = sq + dy * dy;
                        ck[u] = ((unsigned long long)__float_as_uint(sq) << 32) | j;
                    }
                }
            }
#define WRACH_SEL_CE(x, y)                                 \
    {                                                      \
        const unsigned long long lo_ = ck[x], hi_ = ck[y]; \
        const bool sw_ = hi_ < lo_;                        \
        ck[x] = sw_ ? hi_ : lo_;                           \
        ck[y] = sw_ ? lo_ : hi_;                           \
    }
            SEL_NET_32(WRACH_SEL_CE)
#undef WRACH_SEL_CE
            // Real keys never have the empty key's high word (distances are
            // finite), so only real neighbours can tie.
#pragma unroll
            for (int r = 0; r < kNbrs; ++r)
                tie = tie || (ck[r + 1] != ~0ull && (uint32_t)(ck[r] >> 32) == (uint32_t)(ck[r + 1] >> 32));
        }
        const bool exact = __ballot(tie) != 0;  // every lane votes
        if (feats) {
            if (exact) {
#pragma unroll
                for (int kb = 1; kb <= kSortLog2; ++kb) {
#pragma unroll
                    for (int jb = kb - 1; jb >= 0; --jb) {
                        const int kk = 1 << kb, jj = 1 << jb;
#pragma unroll
                        for (int u = 0; u < kSortCands; ++u) {
                            const int l = u ^ jj;
                            if (l > u) {
                                const bool up = (u & kk) == 0;
                                const unsigned long long x = ck[u], y = ck[l];
                                const bool sw = up ? cand_less(a.pos, a.vel, y, x) : cand_less(a.pos, a.vel, x, y);
                                ck[u] = sw ? y : x;
                                ck[l] = sw ? x : y;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kNbrs; ++r) key[r] = ck[r];
            for (uint32_t u = kSortCands; u < m5; ++u) {
                uint32_t j = 0;
#pragma unroll
                for (int t = 0; t < 5; ++t)
                    if (u >= cum[t] && u < cum[t] + (re[t] - rs[t])) j = rs[t] + (u - cum[t]);
                if (j == i) continue;
                const float2 o = a.pos[j];
                const float dx = o.x - q.x;
                const float dy = o.y - q.y;
                float sq = dx * dx;
                sq = sq + dy * dy;
                insert16(a.pos, a.vel, key, ((unsigned long long)__float_as_uint(sq) << 32) | j);
            }
        }
        uint4 *row = reinterpret_cast<uint4 *>(a.feat + (size_t)(a.row0 + i) * kFeat);
#pragma unroll
        for (int r = 0; r < kNbrs; r += 2) {
            uint32_t h[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (key[r + u] != ~0ull) {
                    const uint32_t j = (uint32_t)key[r + u];
                    const float2 o = a.pos[j];
                    const float2 ov = a.vel[j];
                    h[2 * u] = (uint32_t)bf16_bits(o.x - q.x) | ((uint32_t)bf16_bits(o.y - q.y) << 16);
                    h[2 * u + 1] = (uint32_t)bf16_bits(ov.x - w.x) | ((uint32_t)bf16_bits(ov.y - w.y) << 16);
                }
            }
            row[r / 2] = make_uint4(h[0], h[1], h[2], h[3]);
        }
    }
    if (a.sums) {
        __shared__ unsigned long long s_sum[4], s_cnt[4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            osum += __shfl_xor(osum, o, 64);
            ocnt += __shfl_xor(ocnt, o, 64);
        }
        if (lane_id() == 0) {
            s_sum[threadIdx.x >> 6] = osum;
            s_cnt[threadIdx.x >> 6] = ocnt;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long cnt = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            if (cnt) {
                unsigned long long *slot = a.sums + (blockIdx.x % kStencilSumSlots) * kStencilSumStride;
                atomicAdd(slot, s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3]);
                atomicAdd(slot + 1, cnt);
            }
        }
    }
}

}  // namespace

// The frame launchers take optional start/stop events, which hipExtLaunchKernel
// stamps with the kernel's own start and end (wrach_step_profiled).
template <typename K>
static hipError_t launch(K kernel, uint32_t grid, uint32_t block, hipStream_t stream,
                         hipEvent_t start, hipEvent_t stop, const FrameParams &p,
                         const FrameBuffers &b) {
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, start, stop, 0, p, b);
    else
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, p, b);
    return hipGetLastError();
}

hipError_t launch_collide(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                          hipEvent_t start, hipEvent_t stop) {
    // One block of 4 wavefronts per 256 * kCollideWindows packed particles
    // (one block for none: the events of a profiled frame still get a kernel).
    constexpr uint32_t per_block = kCollideBlock * kCollideWindows;
    const uint32_t blocks = p.n ? (p.n + per_block - 1) / per_block : 1u;
    return launch(collide_kernel, blocks, kCollideBlock, stream, start, stop, p, b);
}

hipError_t launch_cells(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                        hipEvent_t start, hipEvent_t stop) {
    return launch(cells_kernel, b.cells_grid, kCollideBlock, stream, start, stop, p, b);
}

static uint32_t resident_blocks(const void *kernel) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kCollideBlock, 0) != hipSuccess)
        return 0;
    return (uint32_t)(per_cu > 0 ? per_cu : 1) * (uint32_t)(cus > 0 ? cus : 1);
}

uint32_t resident_blocks_cells() { return resident_blocks(reinterpret_cast<const void *>(cells_kernel)); }

hipError_t launch_canonical(const FrameParams &p, const FrameBuffers &b, hipStream_t stream) {
    const uint32_t grid = (p.ncell + kCollideBlock - 1) / kCollideBlock;
    hipLaunchKernelGGL(canonical_kernel, dim3(grid), dim3(kCollideBlock), 0, stream, p, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    static uint32_t big_grid = 0;
    if (!big_grid) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, canonical_big_kernel, kCollideBlock, 0) ==
                hipSuccess)
            big_grid = (uint32_t)(cus * (per > 0 ? per : 1));
        if (!big_grid) big_grid = 1024;
    }
    hipLaunchKernelGGL(canonical_big_kernel, dim3(big_grid), dim3(kCollideBlock), 0, stream, p, b);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_fill32(b.canon_ctr, 0u, 1, stream);  // for the next readback
}

template <int S>
static hipError_t launch_scan_as(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                                 hipEvent_t start, hipEvent_t stop) {
    constexpr ScanShape sh = kScanShapes[S];
    return launch(scan_kernel<sh.threads, sh.items>, b.scan_tiles, sh.threads, stream, start, stop,
                  p, b);
}

hipError_t launch_scan(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                       hipEvent_t start, hipEvent_t stop) {
    static_assert(kNumScanShapes == 9, "one case per scan shape");
    switch (b.scan_shape) {
        case 8: return launch_scan_as<8>(p, b, stream, start, stop);
        case 7: return launch_scan_as<7>(p, b, stream, start, stop);
        case 6: return launch_scan_as<6>(p, b, stream, start, stop);
        case 1: return launch_scan_as<1>(p, b, stream, start, stop);
        case 2: return launch_scan_as<2>(p, b, stream, start, stop);
        case 3: return launch_scan_as<3>(p, b, stream, start, stop);
        case 4: return launch_scan_as<4>(p, b, stream, start, stop);
        case 5: return launch_scan_as<5>(p, b, stream, start, stop);
        default: return launch_scan_as<0>(p, b, stream, start, stop);
    }
}

template <int S>
static hipError_t launch_scatter_as(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                                    hipEvent_t start, hipEvent_t stop) {
    constexpr int per = kScatterPerLane[S];
    constexpr uint32_t per_block = kScatterBlock * per;
    const uint32_t tiles = (p.n + per_block - 1) / per_block;
    return launch(scatter_kernel<per>, (tiles + 7) / 8 * 8, kScatterBlock, stream, start, stop, p, b);
}

hipError_t launch_scatter(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                          hipEvent_t start, hipEvent_t stop) {
    static_assert(kNumScatterShapes == 3, "one case per scatter shape");
    switch (b.scatter_shape) {
        case 1: return launch_scatter_as<1>(p, b, stream, start, stop);
        case 2: return launch_scatter_as<2>(p, b, stream, start, stop);
        default: return launch_scatter_as<0>(p, b, stream, start, stop);
    }
}

hipError_t launch_materialize(const FrameParams &p, const FrameBuffers &b, hipStream_t stream) {
    const uint32_t chunks = (p.ncell + kChunkCells - 1) / kChunkCells;
    if (chunks == 0) return hipSuccess;
    hipLaunchKernelGGL(materialize_kernel, dim3(chunks), dim3(256), 0, stream, p, b);
    return hipGetLastError();
}

hipError_t launch_fill32(uint32_t *dst, uint32_t v, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_copy32(uint32_t *dst, const uint32_t *src, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t *dst, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dst, n);
    return hipGetLastError();
}

hipError_t launch_keys_from_offsets(const uint32_t *idx, uint32_t ncell, uint32_t n, uint32_t *key,
                                    hipStream_t stream) {
    if (n == 0 || ncell == 0) return hipSuccess;
    hipLaunchKernelGGL(keys_from_offsets_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, idx, ncell,
                       n, key);
    return hipGetLastError();
}

hipError_t launch_upload_bin(const UploadParams &u, const FrameBuffers &b, hipStream_t stream) {
    if (u.n == 0) return hipSuccess;
    hipLaunchKernelGGL(upload_bin_kernel, dim3((u.n + 255) / 256), dim3(256), 0, stream, u, b);
    return hipGetLastError();
}

hipError_t launch_store_outside_count(const UploadParams &u, uint32_t *counts, hipStream_t stream) {
    if (u.n == 0) return hipSuccess;
    hipLaunchKernelGGL(store_outside_count_kernel, dim3((u.n + 255) / 256), dim3(256), 0, stream, u,
                       counts);
    return hipGetLastError();
}

hipError_t launch_store_outside(const UploadParams &u, float2 *dpos, float2 *dvel, uint32_t base,
                                const uint32_t *offsets, hipStream_t stream) {
    if (u.n == 0) return hipSuccess;
    hipLaunchKernelGGL(store_outside_kernel, dim3((u.n + 255) / 256), dim3(256), 0, stream, u, dpos,
                       dvel, base, offsets);
    return hipGetLastError();
}

hipError_t launch_upload_scatter(const UploadParams &u, const FrameBuffers &b, hipStream_t stream) {
    const uint32_t blocks = u.n ? (u.n + 255) / 256 : 1;  // block 0 advances the epoch
    hipLaunchKernelGGL(upload_scatter_kernel, dim3(blocks), dim3(256), 0, stream, u, b);
    return hipGetLastError();
}

hipError_t launch_stencil_measure(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                                  hipStream_t stream) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(stencil_measure_kernel, dim3((p.n + 255) / 256), dim3(256), 0, stream, p, b, a);
    return hipGetLastError();
}

hipError_t launch_stencil_move(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                               hipStream_t stream) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(stencil_move_kernel, dim3((p.n + 255) / 256), dim3(256), 0, stream, p, b, a);
    return hipGetLastError();
}

hipError_t launch_learn_features(const FrameParams &p, const FrameBuffers &b, const FeatureArgs &a,
                                 hipStream_t stream) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(learn_features_kernel, dim3((p.n + 255) / 256), dim3(256), 0, stream, p, b, a);
    return hipGetLastError();
}

hipError_t launch_stencil_clear(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                                const float2 *pos, hipStream_t stream) {
    hipLaunchKernelGGL(stencil_clear_kernel, dim3(1), dim3(256), 0, stream, a.words, a.sums);
    if (a.perm && p.n && p.ncell) {
        hipLaunchKernelGGL(crowd_list_kernel, dim3((p.ncell + 255) / 256), dim3(256), 0, stream, p, b, a);
        hipLaunchKernelGGL(crowd_order_kernel, dim3(kCrowdOrderGrid), dim3(256), 0, stream, p, b, a, pos);
    }
    return hipGetLastError();
}

hipError_t launch_check_pair_arithmetic(unsigned long long *out, hipStream_t stream) {
    hipLaunchKernelGGL(check_pair_arithmetic_kernel, dim3(4096), dim3(256), 0, stream, out);
    return hipGetLastError();
}

hipError_t launch_copy_pos(float2 *dst, const float2 *src, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_pos_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, dst, src, n);
    return hipGetLastError();
}

}  // namespace wrach
