// Host side of libwrach_hip.so: the C ABI declared in include/wrach.h.
//
// The context owns the host particle store (the reference's ParticleStore,
// runners/bevy/src/particle_store.rs:14-27), the device buffers of the
// compute worker (runners/bevy/src/compute/builder.rs:70-84), one HIP stream
// and the captured frame graphs.
#include "frame.h"
#include "learn.h"
#include "wrach.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define WRACH_HIP(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(WRACH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Rust's `f as i32` (saturating, NaN -> 0).
int32_t sat_i32(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
    if (f <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
    return (int32_t)f;
}

// f32::div_euclid for a positive divisor.
float div_euclid(float a, float b) {
    const float q = std::trunc(a / b);
    return std::fmod(a, b) < 0.0f ? q - 1.0f : q;
}

// SpatialBin::get_cell_coord (runners/bevy/src/spatial_bin.rs:48-64).
inline void cell_coord(float x, float y, float cs, int32_t &cx, int32_t &cy) {
    cx = sat_i32(div_euclid(x, cs));
    cy = sat_i32(div_euclid(y, cs));
}

// ParticleStore::max_particles_per_frame (particle_store.rs:116-133);
// 0 if it does not fit u32 (where the reference's `expect` would panic).
uint32_t max_particles_per_frame(uint64_t cells, uint32_t cs) {
    const uint64_t normal = cells * cs * cs;
    const uint64_t total = normal + 10 * ((normal + 99) / 100);
    return total > 0xffffffffull ? 0 : (uint32_t)total;
}

template <typename T>
hipError_t dmalloc(T **p, size_t count) {
    return hipMalloc((void **)p, count * sizeof(T) + 256);
}

// Grid, in cells, of the viewport (0, 0, W, H): SpatialBin::get_active_cells
// (spatial_bin.rs:68-95) with bottom-left (0,0), so gx = floor(W/cs) + 1.
constexpr uint64_t kMaxCells = (1ull << 31) - 8;

}  // namespace

struct wrach_ctx {
    wrach_config cfg{};
    wrach_world_settings settings{};
    wrach::FrameParams fp{};
    wrach::FrameBuffers buf{};
    uint32_t capacity = 0;
    uint64_t ncell = 0;
    void *idx_alloc = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph_one = nullptr;
    hipGraphExec_t graph_many = nullptr;
    bool use_graph = true;
    bool canonical_valid = false;
    // The store: every particle ever added, in insertion order.  Cells of
    // the hashmap store keep insertion order, so a stable sort by cell of
    // this list is the packed layout (particle_store.rs:54-59).  On the host
    // with WRACH_FLAG_HOST_PACK, else on the device (dstore_*), where dks
    // holds each store particle's (cell, slot) during a re-pack.
    bool host_pack = false;
    std::vector<float> store_pos, store_vel;
    float2 *dstore_pos = nullptr, *dstore_vel = nullptr;
    uint2 *dks = nullptr;
    uint32_t *dvalid = nullptr;
    size_t dstore_n = 0, dstore_cap = 0;
    uint32_t *dstore_world = nullptr;  // world of each store particle (a batch)
    uint32_t worlds = 1;               // worlds of a batch (one grid each, keys one after another)
    uint64_t world_cells = 0;          // cells of one world (worlds of one size)
    // Worlds of different sizes: each world's grid and configuration on the
    // host, the grids and the world of every chunk on the device.
    std::vector<wrach::WorldGrid> grids;
    std::vector<wrach_config> world_cfg;
    wrach::WorldGrid *dgrids = nullptr;
    uint32_t *dchunk_world = nullptr;
    int32_t blx = 0, bly = 0;  // the grid's bottom-left cell: view_anchor / cell_size
    uint32_t cells_grid0 = 1;  // the cells kernel's grid chosen at creation
    uint64_t chunks = 0;       // chunks of kChunkCells cells
    // Solver: WRACH_SOLVER_R, or WRACH_SOLVER_S with up to `passes`
    // separation passes before the reference step (stencil_* kernels).
    uint32_t solver = WRACH_SOLVER_R;
    bool quiet = false;  // WRACH_SOLVER_QUIET: no measure after the passes
    uint32_t passes = 0;
    float overlap_tol = 0.0f;
    uint32_t *scell = nullptr;   // frame-start cell of each particle (capacity entries)
    uint32_t *swords = nullptr;  // wrach::kStencilWords words
    unsigned long long *ssums = nullptr;  // wrach::kStencilSumEntries entries
    uint32_t *sperm = nullptr;   // the separation kernels' lane order (capacity entries)
    uint32_t *scrowd = nullptr;  // the frame's crowded cells (crowd_cap entries)
    uint32_t crowd_cap = 0;
    uint32_t *err_host = nullptr;  // pinned copy of buf.error, read at each synchronisation
    // The learned warm start (wrach_learn_*, WRACH_SOLVER_L): the network's
    // master parameters, gradient and Adam moments (f32, wrach::kParams
    // each), its packed weights, the frame's feature rows and a third
    // position buffer for the passes after the warm start.
    float *lparams = nullptr, *lgrad = nullptr, *lm = nullptr, *lv = nullptr, *lw3 = nullptr;
    uint16_t *lfrag = nullptr;
    uint32_t adam_t = 0;
    uint16_t *lfeat = nullptr;
    float2 *lbuf = nullptr;
    // Training set (rows of features and target displacements) and the
    // workspace of one batch.
    uint16_t *tfeat = nullptr;
    float2 *ttarget = nullptr;
    size_t trows = 0, tcap = 0;
    uint32_t *tsel = nullptr;
    size_t tsel_cap = 0;
    uint32_t tbsz = 0;
    uint16_t *txt = nullptr, *ta1t = nullptr, *tdz1t = nullptr, *tdz2t = nullptr;
    float *ta2t = nullptr, *tdyt = nullptr, *tsq = nullptr, *tpart1 = nullptr, *tpart2 = nullptr;
    float *tloss = nullptr;  // one loss per step of a training call
    size_t tloss_cap = 0;
};

namespace {

constexpr uint32_t kFramesPerGraph = 16;

void free_graphs(wrach_ctx *c) {
    if (c->graph_one) (void)hipGraphExecDestroy(c->graph_one);
    if (c->graph_many) (void)hipGraphExecDestroy(c->graph_many);
    c->graph_one = c->graph_many = nullptr;
}

// Synchronises the stream and reports a wait a kernel gave up
// (FrameBuffers::error, raised by the scan's bounded look-back).  The word
// is never cleared: the offsets of that frame, and so every later frame of
// the context, are not valid.
int settle(wrach_ctx *c) {
    WRACH_HIP(hipMemcpyAsync(c->err_host, c->buf.error, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    WRACH_HIP(hipStreamSynchronize(c->stream));
    if (*c->err_host & wrach::kErrLookBack)
        return fail(WRACH_E_DEVICE, "the exclusive scan's look-back waited past its bound: the offsets "
                                    "of a frame are not valid and the context must be destroyed");
    if (*c->err_host & wrach::kErrLayout)
        return fail(WRACH_E_DEVICE, "the collide kernel found particle cells that disagree with the "
                                    "offsets: the layout is not valid and the context must be destroyed");
    return WRACH_OK;
}

#define WRACH_SETTLE(c)                 \
    do {                                \
        const int rc_ = settle(c);      \
        if (rc_ != WRACH_OK) return rc_; \
    } while (0)

// The separation kernels' buffers; the lane order when both of its
// buffers are there.
wrach::StencilArgs stencil_args(const wrach_ctx *c) {
    wrach::StencilArgs a{};
    a.cell = c->scell;
    a.words = c->swords;
    a.sums = c->ssums;
    if (c->sperm && c->scrowd) {
        a.perm = c->sperm;
        a.crowd = c->scrowd;
        a.crowd_cap = c->crowd_cap;
    }
    return a;
}

// Mode S's passes: P0 = pos; pass k reads pos or pos_out (k even or odd)
// and writes the other (pos_out is free until the collide kernel writes
// it); an odd count leaves the result in pos_out and a copy brings it back.
int enqueue_stencil(wrach_ctx *c) {
    if (c->solver != WRACH_SOLVER_S || c->passes == 0 || c->fp.n == 0) return WRACH_OK;
    wrach::StencilArgs a = stencil_args(c);
    a.tol = c->overlap_tol;
    WRACH_HIP(wrach::launch_stencil_clear(c->fp, c->buf, a, c->buf.pos, c->stream));
    const uint32_t passes = c->passes;
    // With a tolerance, the start of every pass is measured before it (the
    // pass runs only above the tolerance).  With tolerance 0 a pass over no
    // overlap moves nothing, so each pass measures its own start and only
    // the end has a measure of its own.
    const bool check = c->overlap_tol > 0.0f;
    a.first = 1;  // the frame's first kernel derives the cells
    if (check) {
        a.pin = c->buf.pos;
        a.in_slot = 0;
        a.stat = 1;  // the overlap sums before the passes
        WRACH_HIP(wrach::launch_stencil_measure(c->fp, c->buf, a, c->stream));
        a.first = 0;
    }
    for (uint32_t k = 0; k < passes; ++k) {
        float2 *in = k % 2 == 0 ? c->buf.pos : c->buf.pos_out;
        float2 *out = k % 2 == 0 ? c->buf.pos_out : c->buf.pos;
        a.pin = in;
        a.pout = out;
        a.in_slot = k;
        a.check = check ? 1 : 0;
        a.stat = !check && k == 0 ? 1 : 0;
        WRACH_HIP(wrach::launch_stencil_move(c->fp, c->buf, a, c->stream));
        a.first = 0;
        if (check ? (k + 1 < passes || !c->quiet) : (k + 1 == passes && !c->quiet)) {
            a.pin = out;
            a.in_slot = k + 1;
            a.check = 0;
            a.stat = k + 1 == passes ? 2 : 0;  // the last measure: the sums after the passes
            a.half = 1;  // the frame's first kernel has run: each pair from one side
            WRACH_HIP(wrach::launch_stencil_measure(c->fp, c->buf, a, c->stream));
            a.half = 0;
        }
    }
    if (passes % 2) WRACH_HIP(wrach::launch_copy_pos(c->buf.pos, c->buf.pos_out, c->fp.n, c->stream));
    return WRACH_OK;
}

wrach::MlpWeights mlp_weights(const wrach_ctx *c) {
    wrach::MlpWeights w{};
    w.frag = c->lfrag;
    w.w3 = c->lw3;
    w.params = c->lparams;
    return w;
}

// The learned frame (WRACH_SOLVER_L), before the reference step.  With X =
// pos (the frame start), Y = pos_out and Z = lbuf:
//   features of X (and each particle's cell, and X's overlap sums, set 0);
//   the network's warm start X -> Y;
//   pass 1 Y -> Z (its input's sums: set 2);
//   pass 2 Z -> Y, reading X instead of Z when the warm start left more mean
//     overlap than X had (the fallback to the plain passes);
//   the other passes end in X (a copy when they cannot).
// One pass: Y -> X, no fallback.  No pass: the warm start, copied to X.
int enqueue_learned(wrach_ctx *c) {
    if (c->solver != WRACH_SOLVER_L || c->fp.n == 0) return WRACH_OK;
    const uint32_t k = c->passes;
    float2 *X = c->buf.pos, *Y = c->buf.pos_out, *Z = c->lbuf;
    wrach::StencilArgs a = stencil_args(c);
    WRACH_HIP(wrach::launch_stencil_clear(c->fp, c->buf, a, X, c->stream));
    wrach::FeatureArgs f{};
    f.pos = X;
    f.vel = c->buf.vel;
    f.feat = c->lfeat;
    f.row0 = 0;
    f.cell = c->scell;
    f.sums = c->ssums;  // set 0: the frame start
    WRACH_HIP(wrach::launch_learn_features(c->fp, c->buf, f, c->stream));
    wrach::InferArgs in{};
    in.feat = c->lfeat;
    in.n = c->fp.n;
    in.pin = X;
    in.pout = Y;
    WRACH_HIP(wrach::launch_learn_infer(mlp_weights(c), in, c->stream));
    std::vector<std::pair<float2 *, float2 *>> plan;
    float2 *cur = Y;
    if (k == 1) {
        plan.push_back({Y, X});
        cur = X;
    } else if (k >= 2) {
        plan.push_back({Y, Z});
        plan.push_back({Z, Y});
        uint32_t r = k - 2;
        if (r > 0 && r % 2 == 0) {
            plan.push_back({Y, Z});
            plan.push_back({Z, X});
            cur = X;
            r -= 2;
        }
        for (; r > 0; --r) {
            float2 *nx = cur == X ? Y : X;
            plan.push_back({cur, nx});
            cur = nx;
        }
    }
    a.tol = 0.0f;
    a.first = 0;  // the cells came from the features
    for (uint32_t j = 0; j < plan.size(); ++j) {
        a.pin = plan[j].first;
        a.pout = plan[j].second;
        a.in_slot = j;
        a.check = 0;
        a.stat = j == 0 ? 3 : 0;  // set 2: the warm start
        a.pin_alt = j == 1 ? X : nullptr;
        WRACH_HIP(wrach::launch_stencil_move(c->fp, c->buf, a, c->stream));
    }
    if (cur != X) WRACH_HIP(wrach::launch_copy_pos(X, cur, c->fp.n, c->stream));
    if (!c->quiet) {
        a.pin = X;
        a.pout = nullptr;
        a.in_slot = k;
        a.stat = 2;  // set 1: after the passes
        a.pin_alt = nullptr;
        a.half = plan.empty() ? 0 : 1;  // a pass has checked the budget
        WRACH_HIP(wrach::launch_stencil_measure(c->fp, c->buf, a, c->stream));
    }
    return WRACH_OK;
}

int enqueue_frame(wrach_ctx *c) {
    const int learned = enqueue_learned(c);
    if (learned) return learned;
    const int rc = enqueue_stencil(c);
    if (rc) return rc;
    WRACH_HIP(wrach::launch_collide(c->fp, c->buf, c->stream));
    WRACH_HIP(wrach::launch_cells(c->fp, c->buf, c->stream));
    WRACH_HIP(wrach::launch_scan(c->fp, c->buf, c->stream));
    WRACH_HIP(wrach::launch_scatter(c->fp, c->buf, c->stream));
    return WRACH_OK;
}

int capture(wrach_ctx *c, uint32_t frames, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    WRACH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = WRACH_OK;
    for (uint32_t f = 0; f < frames && rc == WRACH_OK; ++f) rc = enqueue_frame(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != WRACH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(WRACH_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(WRACH_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return WRACH_OK;
}

void release(wrach_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graphs(c);
    (void)hipFree(c->idx_alloc);
    (void)hipFree(c->buf.pos);
    (void)hipFree(c->buf.vel);
    (void)hipFree(c->buf.src);
    (void)hipFree(c->buf.key);
    (void)hipFree(c->buf.pos_out);
    (void)hipFree(c->buf.vel_out);
    (void)hipFree(c->buf.kslot);
    (void)hipFree(c->buf.sortbuf);
    (void)hipFree(c->buf.cnt);
    (void)hipFree(c->buf.tile_state);
    (void)hipFree(c->buf.chunk_next);
    (void)hipFree(c->buf.chunk_live);
    (void)hipFree(c->buf.chunk_prefix);
    (void)hipFree(c->buf.epoch);
    (void)hipFree(c->buf.error);
    (void)hipFree(c->buf.canon_ctr);
    if (c->err_host) (void)hipHostFree(c->err_host);
    (void)hipFree(c->dstore_pos);
    (void)hipFree(c->dstore_vel);
    (void)hipFree(c->dks);
    (void)hipFree(c->dstore_world);
    (void)hipFree(c->dgrids);
    (void)hipFree(c->dchunk_world);
    (void)hipFree(c->dvalid);
    (void)hipFree(c->buf.cells);
    (void)hipFree(c->buf.class_ctr);
    (void)hipFree(c->scell);
    (void)hipFree(c->swords);
    (void)hipFree(c->ssums);
    (void)hipFree(c->sperm);
    (void)hipFree(c->scrowd);
    for (void *q : {(void *)c->lparams, (void *)c->lgrad, (void *)c->lm, (void *)c->lv, (void *)c->lw3,
                    (void *)c->lfrag, (void *)c->lfeat, (void *)c->lbuf, (void *)c->tfeat,
                    (void *)c->ttarget, (void *)c->tsel, (void *)c->txt, (void *)c->ta1t,
                    (void *)c->tdz1t, (void *)c->tdz2t, (void *)c->ta2t, (void *)c->tdyt, (void *)c->tsq,
                    (void *)c->tpart1, (void *)c->tpart2, (void *)c->tloss})
        (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// The packer: ParticleStore::create_packed_data (particle_store.rs:90-103)
// -> SpatialBin::create_packed_data (spatial_bin.rs:103-149).  Stable
// counting sort of the store by row-major cell of the active grid; particles
// outside the grid are skipped.  indices = [0, 0, c0, c0+c1, ..., N].
uint32_t pack_store(const wrach_ctx *c, std::vector<uint32_t> &indices,
                    std::vector<float> &pos, std::vector<float> &vel) {
    const size_t total = c->store_pos.size() / 2;
    const uint32_t gx = c->settings.grid_dimensions[0], gy = c->settings.grid_dimensions[1];
    const float cs = (float)c->cfg.cell_size;
    std::vector<int64_t> cell(total);
    indices.assign(c->ncell + 2, 0);
    for (size_t i = 0; i < total; ++i) {
        int32_t cx, cy;
        cell_coord(c->store_pos[2 * i], c->store_pos[2 * i + 1], cs, cx, cy);
        if (cx < 0 || cy < 0 || (uint32_t)cx >= gx || (uint32_t)cy >= gy) {
            cell[i] = -1;
            continue;
        }
        cell[i] = (int64_t)cy * gx + cx;
        indices[cell[i] + 2] += 1;
    }
    uint64_t run = 0;
    for (uint64_t k = 2; k < c->ncell + 2; ++k) {
        run += indices[k];
        indices[k] = (uint32_t)run;  // end of cell k-2 == start of cell k-1
    }
    if (run > 0xffffffffull) return 0xffffffffu;
    const uint32_t n = (uint32_t)run;
    pos.resize(2 * (size_t)n);
    vel.resize(2 * (size_t)n);
    std::vector<uint32_t> cursor(indices.begin() + 1, indices.end() - 1);  // start_k
    for (size_t i = 0; i < total; ++i) {
        if (cell[i] < 0) continue;
        const uint32_t d = cursor[cell[i]]++;
        pos[2 * (size_t)d] = c->store_pos[2 * i];
        pos[2 * (size_t)d + 1] = c->store_pos[2 * i + 1];
        vel[2 * (size_t)d] = c->store_vel[2 * i];
        vel[2 * (size_t)d + 1] = c->store_vel[2 * i + 1];
    }
    return n;
}


// wrach_add_particles with WRACH_FLAG_HOST_PACK: the store on the host,
// packed there (pack_store) and the packed arrays uploaded.
int add_host_packed(wrach_ctx *c, const float *pos_xy, const float *vel_xy, uint32_t n) {
    const size_t old = c->store_pos.size();
    c->store_pos.insert(c->store_pos.end(), pos_xy, pos_xy + 2 * (size_t)n);
    c->store_vel.insert(c->store_vel.end(), vel_xy, vel_xy + 2 * (size_t)n);
    std::vector<uint32_t> indices;
    std::vector<float> pos, vel;
    const uint32_t packed = pack_store(c, indices, pos, vel);
    if (packed > c->capacity) {
        c->store_pos.resize(old);
        c->store_vel.resize(old);
        return fail(WRACH_E_CAPACITY, "particles in frame (" + std::to_string(packed) +
                                          ") exceed the buffer capacity (" +
                                          std::to_string(c->capacity) + ")");
    }
    WRACH_HIP(hipMemcpyAsync(c->buf.idx, indices.data(), indices.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c->stream));
    if (packed) {
        WRACH_HIP(hipMemcpyAsync(c->buf.pos, pos.data(), pos.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
        WRACH_HIP(hipMemcpyAsync(c->buf.vel, vel.data(), vel.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    }
    WRACH_HIP(wrach::launch_iota(c->buf.src, packed, c->stream));
    WRACH_HIP(wrach::launch_keys_from_offsets(c->buf.idx, (uint32_t)c->ncell, packed, c->buf.key, c->stream));
    // every offset came from the host: no chunk's offsets are stale
    WRACH_HIP(hipMemsetAsync(c->buf.chunk_live, 1, c->chunks, c->stream));
    WRACH_SETTLE(c);
    if (c->fp.n != packed) free_graphs(c);
    c->fp.n = packed;
    c->settings.particles_in_frame_count = packed;
    return WRACH_OK;
}

// Device store arrays: room for `cap` particles, grown by doubling.
size_t store_cap_for(size_t n) {
    size_t cap = 1024;
    while (cap < n) cap *= 2;
    return cap;
}

int alloc_store(size_t cap, float2 **p, float2 **v, uint2 **k) {
    *p = *v = nullptr;
    *k = nullptr;
    hipError_t e = hipMalloc(p, cap * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(v, cap * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(k, cap * sizeof(uint2));
    if (e != hipSuccess) {
        (void)hipFree(*p);
        (void)hipFree(*v);
        (void)hipFree(*k);
        *p = *v = nullptr;
        *k = nullptr;
        return fail(WRACH_E_DEVICE, std::string("device store: ") + hipGetErrorString(e));
    }
    return WRACH_OK;
}

// Replaces the store arrays (the old ones are no longer in use).
void swap_store(wrach_ctx *c, float2 *p, float2 *v, uint2 *k, size_t cap) {
    (void)hipFree(c->dstore_pos);
    (void)hipFree(c->dstore_vel);
    (void)hipFree(c->dks);
    c->dstore_pos = p;
    c->dstore_vel = v;
    c->dks = k;
    c->dstore_cap = cap;
}

// Room for `total` store particles on the device, keeping the first
// dstore_n.
int reserve_device_store(wrach_ctx *c, size_t total) {
    if (total <= c->dstore_cap) return WRACH_OK;
    const size_t cap = store_cap_for(total);
    float2 *p, *v;
    uint2 *k;
    {
        const int rc = alloc_store(cap, &p, &v, &k);
        if (rc) return rc;
    }
    hipError_t e = hipSuccess;
    if (c->dstore_n) {
        e = hipMemcpyAsync(p, c->dstore_pos, c->dstore_n * sizeof(float2), hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(v, c->dstore_vel, c->dstore_n * sizeof(float2), hipMemcpyDeviceToDevice,
                               c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        (void)hipFree(p);
        (void)hipFree(v);
        (void)hipFree(k);
        return fail(WRACH_E_DEVICE, std::string("device store: ") + hipGetErrorString(e));
    }
    swap_store(c, p, v, k, cap);
    if (c->worlds > 1) {  // the world of each store particle, kept alongside
        uint32_t *w = nullptr;
        e = hipMalloc(&w, cap * sizeof(uint32_t));
        if (e == hipSuccess && c->dstore_n)
            e = hipMemcpyAsync(w, c->dstore_world, c->dstore_n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipFree(w);
            return fail(WRACH_E_DEVICE, std::string("device store: ") + hipGetErrorString(e));
        }
        (void)hipFree(c->dstore_world);
        c->dstore_world = w;
    }
    return WRACH_OK;
}

// The binning parameters over the first n store particles.
wrach::UploadParams store_params(const wrach_ctx *c, size_t n) {
    wrach::UploadParams u{};
    u.pos = c->dstore_pos;
    u.vel = c->dstore_vel;
    u.ks = c->dks;
    u.valid = c->dvalid;
    u.n = (uint32_t)n;
    u.cs = (float)c->cfg.cell_size;
    u.gx = c->settings.grid_dimensions[0];
    u.gy = c->settings.grid_dimensions[1];
    u.cap = c->capacity;
    u.world = c->worlds > 1 ? c->dstore_world : nullptr;
    u.world_cells = (uint32_t)c->world_cells;
    u.grids = c->dgrids;
    u.blx = c->blx;
    u.bly = c->bly;
    return u;
}

// wrach_add_particles without WRACH_FLAG_HOST_PACK: the new particles are
// appended to the device store, which is binned on the device (cell and slot
// per particle, the frame's scan, a scatter).  Nothing of the state changes
// before the capacity check has passed.
int add_device_packed(wrach_ctx *c, const float *pos_xy, const float *vel_xy, uint32_t n,
                      uint32_t world) {
    const size_t total = c->dstore_n + n;
    // a store index and the run flag share a 32-bit source word
    if (total > (size_t)0x7fffffffu) return fail(WRACH_E_CAPACITY, "store exceeds 2^31 - 1 particles");
    {
        const int rc = reserve_device_store(c, total);
        if (rc) return rc;
    }
    if (n) {
        WRACH_HIP(hipMemcpyAsync(c->dstore_pos + c->dstore_n, pos_xy, (size_t)n * sizeof(float2),
                                 hipMemcpyHostToDevice, c->stream));
        WRACH_HIP(hipMemcpyAsync(c->dstore_vel + c->dstore_n, vel_xy, (size_t)n * sizeof(float2),
                                 hipMemcpyHostToDevice, c->stream));
        if (c->worlds > 1)
            WRACH_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(c->dstore_world + c->dstore_n),
                                        (int)world, n, c->stream));
    }
    const wrach::UploadParams u = store_params(c, total);
    WRACH_HIP(hipMemsetAsync(c->dvalid, 0, sizeof(uint32_t), c->stream));
    WRACH_HIP(wrach::launch_upload_bin(u, c->buf, c->stream));
    uint32_t packed = 0;
    WRACH_HIP(hipMemcpyAsync(&packed, c->dvalid, sizeof packed, hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    if (packed > c->capacity) {
        // the counts are the only state touched: clear them for the next frame
        WRACH_HIP(hipMemsetAsync(c->buf.cnt, 0, (c->ncell + 16) * sizeof(uint32_t), c->stream));
        WRACH_SETTLE(c);
        return fail(WRACH_E_CAPACITY, "particles in frame (" + std::to_string(packed) +
                                          ") exceed the buffer capacity (" +
                                          std::to_string(c->capacity) + ")");
    }
    WRACH_HIP(wrach::launch_scan(c->fp, c->buf, c->stream));
    WRACH_HIP(wrach::launch_upload_scatter(u, c->buf, c->stream));
    WRACH_SETTLE(c);
    c->dstore_n = total;
    if (c->fp.n != packed) free_graphs(c);
    c->fp.n = packed;
    c->settings.particles_in_frame_count = packed;
    return WRACH_OK;
}

}  // namespace

extern "C" {

const char *wrach_last_error(void) { return g_err.c_str(); }

void wrach_config_default(wrach_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->width = 480;
    cfg->height = 352;
    cfg->cell_size = 3;
    cfg->boundaries_as_dimensions = 0;
    cfg->gravity[0] = cfg->gravity[1] = 0.0f;
    cfg->dt = 1.0f;
    cfg->device = 0;
}

namespace {

// world_cfgs (n_worlds entries) or null: worlds of different sizes, each
// world's dimensions from its entry, every other field from cfg (entry 0).
int create_ctx(const wrach_config *cfg, uint32_t n_worlds, const wrach_config *world_cfgs,
               wrach_ctx **out) {
    if (!cfg || !out) return fail(WRACH_E_INVALID, "wrach_create: null argument");
    *out = nullptr;
    if (cfg->cell_size == 0) return fail(WRACH_E_INVALID, "cell_size must be >= 1");
    if (cfg->flags & ~WRACH_FLAG_HOST_PACK) return fail(WRACH_E_INVALID, "unknown flags");
    if (n_worlds == 0) return fail(WRACH_E_INVALID, "a batch holds at least one world");
    if (n_worlds > 1 && (cfg->flags & WRACH_FLAG_HOST_PACK))
        return fail(WRACH_E_INVALID, "a batch of worlds bins its store on the device");
    if (!std::isfinite(cfg->gravity[0]) || !std::isfinite(cfg->gravity[1]) || !std::isfinite(cfg->dt))
        return fail(WRACH_E_INVALID, "gravity and dt must be finite");
    int ndev = 0;
    WRACH_HIP(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev) return fail(WRACH_E_INVALID, "no such HIP device");
    WRACH_HIP(hipSetDevice(cfg->device));

    const float cs = (float)cfg->cell_size;
    int32_t tx, ty;
    cell_coord((float)cfg->width, (float)cfg->height, cs, tx, ty);
    const uint64_t gx = (uint64_t)tx + 1, gy = (uint64_t)ty + 1;
    const uint64_t world_cells = gx * gy;
    uint64_t ncell = world_cells * n_worlds;  // a batch: the worlds' grids one after another
    uint64_t cap64 = 0;
    std::vector<wrach::WorldGrid> grids;
    if (world_cfgs) {
        // Each world's grid starts on a chunk boundary, so that the chunk of
        // a cell names its world (the padding cells between grids stay empty).
        uint64_t at = 0;
        for (uint32_t w = 0; w < n_worlds; ++w) {
            int32_t wx, wy;
            cell_coord((float)world_cfgs[w].width, (float)world_cfgs[w].height, cs, wx, wy);
            const uint64_t wgx = (uint64_t)wx + 1, wgy = (uint64_t)wy + 1, wcells = wgx * wgy;
            at = (at + wrach::kChunkCells - 1) / wrach::kChunkCells * wrach::kChunkCells;
            if (at + wcells > kMaxCells)
                return fail(WRACH_E_TOO_MANY_CELLS, "grid has more cells than the 32-bit offsets allow");
            const uint32_t wcap = max_particles_per_frame(wcells, cfg->cell_size);
            if (wcap == 0) return fail(WRACH_E_TOO_MANY_CELLS, "particle capacity does not fit in u32");
            wrach::WorldGrid g{};
            g.base = (uint32_t)at;
            g.gx = (uint32_t)wgx;
            g.gy = (uint32_t)wgy;
            g.x1 = 0.0f + (float)world_cfgs[w].width;
            g.y1 = 0.0f + (float)world_cfgs[w].height;
            grids.push_back(g);
            cap64 += wcap;
            at += wcells;
        }
        ncell = at;
    } else {
        if (ncell > kMaxCells)
            return fail(WRACH_E_TOO_MANY_CELLS, "grid has more cells than the 32-bit offsets allow");
        const uint32_t world_cap = max_particles_per_frame(world_cells, cfg->cell_size);
        if (world_cap == 0) return fail(WRACH_E_TOO_MANY_CELLS, "particle capacity does not fit in u32");
        cap64 = (uint64_t)world_cap * n_worlds;
    }
    // Bit 31 of a particle's source index marks the start of a run.
    if (cap64 > 0x7fffffffu)
        return fail(WRACH_E_TOO_MANY_CELLS, "particle capacity exceeds 2^31 - 1");
    const uint32_t cap = (uint32_t)cap64;

    wrach_ctx *c = new wrach_ctx();
    c->cfg = *cfg;
    c->host_pack = (cfg->flags & WRACH_FLAG_HOST_PACK) != 0;
    c->worlds = n_worlds;
    c->world_cells = world_cells;
    c->ncell = ncell;
    c->capacity = cap;
    c->settings.view_dimensions[0] = (float)cfg->width;
    c->settings.view_dimensions[1] = (float)cfg->height;
    c->settings.view_anchor[0] = 0.0f;
    c->settings.view_anchor[1] = 0.0f;
    c->settings.grid_dimensions[0] = (uint32_t)gx;
    c->settings.grid_dimensions[1] = (uint32_t)gy;
    c->settings.cell_size = cfg->cell_size;
    c->settings.particles_in_frame_count = 0;

    wrach::FrameParams &fp = c->fp;
    fp.ax = 0.0f;
    fp.ay = 0.0f;
    fp.x1 = fp.ax + c->settings.view_dimensions[0];
    fp.y1 = fp.ay + c->settings.view_dimensions[1];
    fp.cs = cs;
    // gravity * dt rounded to f32 once here, as the device would per particle
    fp.gdx = cfg->gravity[0] != 0.0f ? cfg->gravity[0] * cfg->dt : -0.0f;
    fp.gdy = cfg->gravity[1] != 0.0f ? cfg->gravity[1] * cfg->dt : -0.0f;
    fp.dt = cfg->dt;
    fp.gx = (uint32_t)gx;
    fp.gy = (uint32_t)gy;
    fp.ncell = (uint32_t)ncell;
    fp.n = 0;
    fp.world_cells = (uint32_t)world_cells;
    fp.wait_limit = wrach::kDefaultWaitLimit;
    fp.epoch_skew = 0;

    wrach::FrameBuffers &b = c->buf;
    b.scan_shape = wrach::default_scan_shape(ncell);
    b.scatter_shape = 0;
    b.scan_tiles = wrach::scan_tiles_for((uint32_t)ncell, b.scan_shape);
    // Queued-cell regions: a class of cells holding at least m particles has
    // at most capacity / m cells.  Each class: kSlices slices sharing that
    // bound, then an overflow list that can hold all of it.
    size_t ncells_total = 0;
    for (int k = 0; k < wrach::kClasses; ++k) {
        b.class_cap[k] = cap / wrach::kClassMinN[k] + 64;
        b.slice_cap[k] = (b.class_cap[k] + wrach::kSlices - 1) / wrach::kSlices + 16;
        b.class_off[k] = (uint32_t)ncells_total;
        ncells_total += (size_t)wrach::kSlices * b.slice_cap[k] + b.class_cap[k];
    }
    if (ncells_total > 0xffffffffu) {
        delete c;
        return fail(WRACH_E_TOO_MANY_CELLS, "queued-cell lists exceed 2^32 entries");
    }
    auto bail = [&](int rc) { release(c); return rc; };
    // Grids: the blocks the device holds at once (the kernels loop over
    // their work), or fewer for a small world.
    {
        const uint32_t want = (b.class_cap[0] * 2 / 64 + wrach::kWavesPerBlock) / wrach::kWavesPerBlock;
        const uint32_t resident = wrach::resident_blocks_cells();
        if (resident == 0) return bail(fail(WRACH_E_DEVICE, "occupancy query failed"));
        b.cells_grid = want < resident ? want : resident;
        if (b.cells_grid == 0) b.cells_grid = 1;
        c->cells_grid0 = b.cells_grid;
    }
#define WRACH_HIP_C(expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return bail(fail(WRACH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    WRACH_HIP_C(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const size_t nidx = ncell + 2;
    WRACH_HIP_C(hipMalloc(&c->idx_alloc, (nidx + 64) * sizeof(uint32_t)));
    b.idx = reinterpret_cast<uint32_t *>(c->idx_alloc) + 3;  // &idx[1] on a 16-byte boundary
    WRACH_HIP_C(dmalloc(&b.pos, cap));
    WRACH_HIP_C(dmalloc(&b.vel, cap));
    WRACH_HIP_C(dmalloc(&b.src, cap));
    WRACH_HIP_C(dmalloc(&b.key, cap));
    WRACH_HIP_C(dmalloc(&b.pos_out, cap));
    WRACH_HIP_C(dmalloc(&b.vel_out, cap));
    WRACH_HIP_C(dmalloc(&b.kslot, cap));
    WRACH_HIP_C(dmalloc(&b.sortbuf, cap));
    WRACH_HIP_C(dmalloc(&b.cnt, ncell + 16));
    // status words for the smallest scan tile of kScanShapes
    const size_t max_tiles = (ncell + wrach::kMinScanTile - 1) / wrach::kMinScanTile;
    WRACH_HIP_C(dmalloc(&b.tile_state, max_tiles));
    c->chunks = (ncell + wrach::kChunkCells - 1) / wrach::kChunkCells;
    WRACH_HIP_C(dmalloc(&b.chunk_next, c->chunks));
    WRACH_HIP_C(dmalloc(&b.chunk_live, c->chunks));
    WRACH_HIP_C(dmalloc(&b.chunk_prefix, c->chunks));
    WRACH_HIP_C(dmalloc(&b.epoch, 1));
    WRACH_HIP_C(dmalloc(&b.error, 1));
    WRACH_HIP_C(dmalloc(&b.canon_ctr, 1));
    WRACH_HIP_C(hipHostMalloc((void **)&c->err_host, sizeof(uint32_t), hipHostMallocDefault));
    *c->err_host = 0;
    WRACH_HIP_C(dmalloc(&b.cells, ncells_total));
    WRACH_HIP_C(dmalloc(&b.class_ctr, wrach::kCtrWords));
    WRACH_HIP_C(dmalloc(&c->dvalid, 1));
    WRACH_HIP_C(hipMemsetAsync(c->idx_alloc, 0, (nidx + 64) * sizeof(uint32_t), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.pos, 0, cap * sizeof(float2), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.vel, 0, cap * sizeof(float2), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.pos_out, 0, cap * sizeof(float2), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.vel_out, 0, cap * sizeof(float2), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.cnt, 0, (ncell + 16) * sizeof(uint32_t), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.tile_state, 0, max_tiles * sizeof(unsigned long long), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.chunk_next, 0, c->chunks, c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.chunk_live, 0, c->chunks, c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.chunk_prefix, 0, c->chunks * sizeof(uint32_t), c->stream));
    {
        static const uint32_t first_epoch = 1;  // status words start at 0, never valid
        WRACH_HIP_C(hipMemcpyAsync(b.epoch, &first_epoch, sizeof first_epoch, hipMemcpyHostToDevice,
                                   c->stream));
    }
    WRACH_HIP_C(hipMemsetAsync(b.class_ctr, 0, wrach::kCtrWords * sizeof(uint32_t), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.error, 0, sizeof(uint32_t), c->stream));
    WRACH_HIP_C(hipMemsetAsync(b.canon_ctr, 0, sizeof(uint32_t), c->stream));
    if (world_cfgs) {
        std::vector<uint32_t> chunk_world(c->chunks);
        for (uint32_t w = 0; w < n_worlds; ++w) {
            const uint64_t first = grids[w].base / wrach::kChunkCells;
            const uint64_t last = w + 1 < n_worlds ? grids[w + 1].base / wrach::kChunkCells : c->chunks;
            for (uint64_t ch = first; ch < last; ++ch) chunk_world[ch] = w;
        }
        WRACH_HIP_C(dmalloc(&c->dgrids, n_worlds));
        WRACH_HIP_C(dmalloc(&c->dchunk_world, c->chunks));
        WRACH_HIP_C(hipMemcpyAsync(c->dgrids, grids.data(), n_worlds * sizeof(wrach::WorldGrid),
                                   hipMemcpyHostToDevice, c->stream));
        WRACH_HIP_C(hipMemcpyAsync(c->dchunk_world, chunk_world.data(), c->chunks * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, c->stream));
        c->grids = grids;
        c->world_cfg.assign(world_cfgs, world_cfgs + n_worlds);
        fp.grids = c->dgrids;
        fp.chunk_world = c->dchunk_world;
        fp.n_grids = n_worlds;
    }
    WRACH_HIP_C(hipStreamSynchronize(c->stream));
#undef WRACH_HIP_C
    *out = c;
    g_err.clear();
    return WRACH_OK;
}

}  // namespace

int wrach_create(const wrach_config *cfg, wrach_ctx **out) { return create_ctx(cfg, 1, nullptr, out); }

int wrach_create_batch(const wrach_config *cfg, uint32_t n_worlds, wrach_ctx **out) {
    return create_ctx(cfg, n_worlds, nullptr, out);
}

int wrach_create_worlds(const wrach_config *cfgs, uint32_t n_worlds, wrach_ctx **out) {
    if (!cfgs || !out) return fail(WRACH_E_INVALID, "wrach_create_worlds: null argument");
    if (n_worlds == 0) return fail(WRACH_E_INVALID, "a batch holds at least one world");
    bool same = true;
    for (uint32_t w = 1; w < n_worlds; ++w) {
        const wrach_config &a = cfgs[0], &b = cfgs[w];
        if (b.cell_size != a.cell_size || b.boundaries_as_dimensions != a.boundaries_as_dimensions ||
            b.gravity[0] != a.gravity[0] || b.gravity[1] != a.gravity[1] || b.dt != a.dt ||
            b.flags != a.flags || b.device != a.device)
            return fail(WRACH_E_INVALID, "world " + std::to_string(w) +
                                             ": the worlds of a batch differ only in their dimensions");
        same = same && b.width == a.width && b.height == a.height;
    }
    // Worlds of one size: the equal-world layout (no tables).
    return create_ctx(&cfgs[0], n_worlds, same ? nullptr : cfgs, out);
}

void wrach_destroy(wrach_ctx *ctx) { release(ctx); }

namespace {

int add_to_world(wrach_ctx *c, uint32_t world, const float *pos_xy, const float *vel_xy, uint32_t n) {
    if (n > 0 && (!pos_xy || !vel_xy)) return fail(WRACH_E_INVALID, "null particle arrays");
    for (uint32_t i = 0; i < 2 * n; ++i)
        if (std::isnan(pos_xy[i]) || std::isnan(vel_xy[i]))
            return fail(WRACH_E_NONFINITE, "NaN in particle " + std::to_string(i / 2));
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    const int rc = c->host_pack ? add_host_packed(c, pos_xy, vel_xy, n)
                                : add_device_packed(c, pos_xy, vel_xy, n, world);
    if (rc) return rc;
    c->canonical_valid = false;
    g_err.clear();
    return WRACH_OK;
}

}  // namespace

int wrach_add_particles(wrach_ctx *c, const float *pos_xy, const float *vel_xy, uint32_t n) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (c->worlds > 1) return fail(WRACH_E_INVALID, "a batch of worlds: use wrach_add_particles_to_world");
    return add_to_world(c, 0, pos_xy, vel_xy, n);
}

int wrach_add_particles_to_world(wrach_ctx *c, uint32_t world, const float *pos_xy, const float *vel_xy,
                                 uint32_t n) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (world >= c->worlds)
        return fail(WRACH_E_INVALID, "no world " + std::to_string(world) + " in a batch of " +
                                         std::to_string(c->worlds));
    return add_to_world(c, world, pos_xy, vel_xy, n);
}

namespace {

// Captures the graphs wrach_step(nframes) launches (kFramesPerGraph frames,
// and one frame for the remainder) that are not captured yet.
int ensure_graphs(wrach_ctx *c, uint32_t nframes) {
    if (nframes >= kFramesPerGraph && !c->graph_many) {
        const int rc = capture(c, kFramesPerGraph, &c->graph_many);
        if (rc) return rc;
    }
    if (nframes % kFramesPerGraph && !c->graph_one) {
        const int rc = capture(c, 1, &c->graph_one);
        if (rc) return rc;
    }
    return WRACH_OK;
}

}  // namespace

int wrach_prepare(wrach_ctx *c, uint32_t nframes) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (nframes == 0 || c->fp.n == 0 || !c->use_graph) return WRACH_OK;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    return ensure_graphs(c, nframes);
}

int wrach_test_scan_wait(wrach_ctx *c, uint32_t wait_limit, uint32_t epoch_skew) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (wait_limit == 0) return fail(WRACH_E_INVALID, "the wait limit must be at least one round");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    free_graphs(c);  // the graphs hold the frame parameters by value
    c->fp.wait_limit = wait_limit;
    c->fp.epoch_skew = epoch_skew;
    return WRACH_OK;
}

int wrach_step(wrach_ctx *c, uint32_t nframes) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (nframes == 0 || c->fp.n == 0) return WRACH_OK;  // an empty world does not change
    WRACH_HIP(hipSetDevice(c->cfg.device));
    c->canonical_valid = false;
    if (!c->use_graph) {
        for (uint32_t f = 0; f < nframes; ++f) {
            const int rc = enqueue_frame(c);
            if (rc) return rc;
        }
        return WRACH_OK;
    }
    {
        const int rc = ensure_graphs(c, nframes);
        if (rc) return rc;
    }
    for (uint32_t f = 0; f < nframes / kFramesPerGraph; ++f)
        WRACH_HIP(hipGraphLaunch(c->graph_many, c->stream));
    for (uint32_t f = 0; f < nframes % kFramesPerGraph; ++f)
        WRACH_HIP(hipGraphLaunch(c->graph_one, c->stream));
    return WRACH_OK;
}

int wrach_step_profiled(wrach_ctx *c, uint32_t nframes, double *kernel_ms) {
    if (!c || !kernel_ms) return fail(WRACH_E_INVALID, "null argument");
    if (nframes == 0 || c->fp.n == 0) return WRACH_OK;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    c->canonical_valid = false;
    constexpr int K = WRACH_NUM_FRAME_KERNELS;
    std::vector<hipEvent_t> ev(2 * K * (size_t)nframes, nullptr);
    for (auto &e : ev) WRACH_HIP(hipEventCreate(&e));
    int rc = WRACH_OK;
    for (uint32_t f = 0; f < nframes && rc == WRACH_OK; ++f) {
        hipEvent_t *fe = &ev[2 * K * (size_t)f];
        hipError_t (*const launch[K])(const wrach::FrameParams &, const wrach::FrameBuffers &,
                                      hipStream_t, hipEvent_t, hipEvent_t) = {
            wrach::launch_collide, wrach::launch_cells, wrach::launch_scan, wrach::launch_scatter};
        rc = enqueue_learned(c);  // the learned warm start and its passes, untimed
        if (rc) break;
        rc = enqueue_stencil(c);  // mode S's passes, untimed
        if (rc) break;
        hipError_t e = hipSuccess;
        // The launch stamps both events with the kernel's own start and end.
        for (int k = 0; k < K && e == hipSuccess; ++k)
            e = launch[k](c->fp, c->buf, c->stream, fe[2 * k], fe[2 * k + 1]);
        if (e != hipSuccess) rc = fail(WRACH_E_DEVICE, std::string("profiled frame: ") + hipGetErrorString(e));
    }
    hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == WRACH_OK && se != hipSuccess)
        rc = fail(WRACH_E_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
    if (rc == WRACH_OK) rc = settle(c);
    if (rc == WRACH_OK) {
        for (uint32_t f = 0; f < nframes; ++f) {
            for (int k = 0; k < K; ++k) {
                float ms = 0.0f;
                hipError_t e = hipEventElapsedTime(&ms, ev[2 * K * (size_t)f + 2 * k],
                                                   ev[2 * K * (size_t)f + 2 * k + 1]);
                if (e != hipSuccess) {
                    rc = fail(WRACH_E_DEVICE, std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
                    break;
                }
                kernel_ms[k] += ms;
            }
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

int wrach_sync(wrach_ctx *c) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    WRACH_SETTLE(c);
    return WRACH_OK;
}

static_assert(WRACH_MAX_PASSES == wrach::kMaxStencilPasses, "one pass limit");
static_assert(WRACH_STENCIL_BUDGET == wrach::kStencilBudget, "one budget");
static_assert(WRACH_LEARN_PARAMS == wrach::kParams, "one parameter count");
static_assert(WRACH_LEARN_FEATURES == (unsigned)wrach::kFeat, "one feature width");
static_assert(WRACH_LEARN_BUDGET == wrach::kFeatBudget, "one feature budget");

namespace {

// The separation passes' buffers (mode S and L, and the learned samples):
// each on its own, so that a failed allocation leaves the context as it
// was and the next call allocates what is still missing.
int ensure_stencil(wrach_ctx *c) {
    if (!c->scell) {
        uint32_t *cell = nullptr;
        WRACH_HIP(dmalloc(&cell, c->capacity ? c->capacity : 1));
        c->scell = cell;
    }
    if (!c->swords) {
        uint32_t *words = nullptr;
        WRACH_HIP(dmalloc(&words, wrach::kStencilWords));
        WRACH_HIP(hipMemset(words, 0, wrach::kStencilWords * sizeof(uint32_t)));
        c->swords = words;
    }
    if (!c->ssums) {
        unsigned long long *sums = nullptr;
        WRACH_HIP(dmalloc(&sums, wrach::kStencilSumEntries));
        WRACH_HIP(hipMemset(sums, 0, wrach::kStencilSumEntries * sizeof(unsigned long long)));
        c->ssums = sums;
    }
    if (!c->sperm) {
        uint32_t *perm = nullptr;
        WRACH_HIP(dmalloc(&perm, c->capacity ? c->capacity : 1));
        c->sperm = perm;
    }
    if (!c->scrowd) {
        // A crowded cell holds at least kCrowdMin particles.
        const uint32_t cap = c->capacity / wrach::kCrowdMin + 1;
        uint32_t *crowd = nullptr;
        WRACH_HIP(dmalloc(&crowd, cap));
        c->scrowd = crowd;
        c->crowd_cap = cap;
    }
    return WRACH_OK;
}

// The learned frame's buffers: feature rows and a third position buffer.
int ensure_learned_frame(wrach_ctx *c) {
    const size_t rows = c->capacity ? c->capacity : 1;
    if (!c->lfeat) {
        uint16_t *f = nullptr;
        WRACH_HIP(dmalloc(&f, rows * wrach::kFeat));
        c->lfeat = f;
    }
    if (!c->lbuf) {
        float2 *b = nullptr;
        WRACH_HIP(dmalloc(&b, rows));
        c->lbuf = b;
    }
    return WRACH_OK;
}

}  // namespace

int wrach_set_solver(wrach_ctx *c, uint32_t mode, uint32_t max_passes, float overlap_tol) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    const uint32_t base = mode & ~WRACH_SOLVER_QUIET;
    const bool quiet = (mode & WRACH_SOLVER_QUIET) != 0 && (base == WRACH_SOLVER_S || base == WRACH_SOLVER_L);
    if (quiet) mode = base;
    if (mode != WRACH_SOLVER_R && mode != WRACH_SOLVER_S && mode != WRACH_SOLVER_L)
        return fail(WRACH_E_INVALID, "no such solver");
    if (mode != WRACH_SOLVER_R) {
        if (max_passes > wrach::kMaxStencilPasses)
            return fail(WRACH_E_INVALID, "more than " + std::to_string(wrach::kMaxStencilPasses) + " passes");
        if (!(overlap_tol >= 0.0f) || std::isinf(overlap_tol))
            return fail(WRACH_E_INVALID, "overlap tolerance must be finite and >= 0");
    }
    if (mode == WRACH_SOLVER_L) {
        if (overlap_tol != 0.0f)
            return fail(WRACH_E_INVALID, "the learned solver runs a fixed number of passes (tolerance 0)");
        if (!c->lparams) return fail(WRACH_E_STATE, "the learned solver needs parameters (wrach_learn_set_params)");
    }
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    if (mode != WRACH_SOLVER_R) {
        const int rc = ensure_stencil(c);
        if (rc) return rc;
    }
    if (mode == WRACH_SOLVER_L) {
        const int rc = ensure_learned_frame(c);
        if (rc) return rc;
    }
    if (c->swords) WRACH_HIP(hipMemset(c->swords, 0, wrach::kStencilWords * sizeof(uint32_t)));
    if (c->ssums) WRACH_HIP(hipMemset(c->ssums, 0, wrach::kStencilSumEntries * sizeof(unsigned long long)));
    free_graphs(c);  // captured with the previous solver
    c->solver = mode;
    c->quiet = quiet;
    c->passes = mode != WRACH_SOLVER_R ? max_passes : 0;
    c->overlap_tol = mode == WRACH_SOLVER_S ? overlap_tol : 0.0f;
    g_err.clear();
    return WRACH_OK;
}

int wrach_solver_stats(wrach_ctx *c, wrach_solver_report *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    uint32_t w[wrach::kStencilWords] = {};
    std::vector<unsigned long long> sums(wrach::kStencilSumEntries, 0ull);
    if (c->swords && c->ssums) {
        WRACH_HIP(hipSetDevice(c->cfg.device));
        WRACH_SETTLE(c);
        WRACH_HIP(hipMemcpy(w, c->swords, sizeof w, hipMemcpyDeviceToHost));
        WRACH_HIP(hipMemcpy(sums.data(), c->ssums, sums.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost));
    }
    wrach_solver_report r{};
    std::memcpy(&r.max_overlap_before, &w[0], sizeof r.max_overlap_before);
    std::memcpy(&r.max_overlap_after, &w[c->passes], sizeof r.max_overlap_after);
    // The passes that ran: those whose start exceeded the tolerance (a pass
    // over no more overlap than that moves nothing, or is skipped).
    for (uint32_t k = 0; k < c->passes; ++k) {
        float start;
        std::memcpy(&start, &w[k], sizeof start);
        if (start > c->overlap_tol) r.passes += 1;
    }
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint32_t which = 0; which < 2; ++which)
        for (uint32_t slot = 0; slot < wrach::kStencilSumSlots; ++slot) {
            const size_t e = ((size_t)which * wrach::kStencilSumSlots + slot) * wrach::kStencilSumStride;
            acc[2 * which] += sums[e];
            acc[2 * which + 1] += sums[e + 1];
        }
    r.overlap_sum_before = acc[0];
    r.pairs_before = acc[1];
    r.overlap_sum_after = acc[2];
    r.pairs_after = acc[3];
    r.after_measured = c->solver != WRACH_SOLVER_R && (c->passes > 0 || c->solver == WRACH_SOLVER_L) &&
                               !c->quiet ? 1u : 0u;
    *out = r;
    return WRACH_OK;
}

int wrach_set_graph(wrach_ctx *c, int enable) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    c->use_graph = enable != 0;
    return WRACH_OK;
}

int wrach_capacity(wrach_ctx *c, uint32_t *max_particles, uint32_t *n_indices) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (max_particles) *max_particles = c->capacity;
    if (n_indices) *n_indices = (uint32_t)(c->ncell + 2);
    return WRACH_OK;
}

int wrach_settings(wrach_ctx *c, wrach_world_settings *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    *out = c->settings;
    return WRACH_OK;
}

int wrach_device_state(wrach_ctx *c, wrach_device_view *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_HIP(wrach::launch_materialize(c->fp, c->buf, c->stream));  // every offset, on the stream
    out->indices = c->buf.idx;
    out->positions = reinterpret_cast<const float *>(c->buf.pos);
    out->velocities = reinterpret_cast<const float *>(c->buf.vel);
    out->n_particles = c->fp.n;
    out->n_indices = (uint32_t)(c->ncell + 2);
    out->stream = c->stream;
    return WRACH_OK;
}

int wrach_canonical_device(wrach_ctx *c, wrach_device_view *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    if (!c->canonical_valid) {
        // the offsets of chunks the last scan found empty, then the order
        WRACH_HIP(wrach::launch_materialize(c->fp, c->buf, c->stream));
        if (c->fp.n > 0) WRACH_HIP(wrach::launch_canonical(c->fp, c->buf, c->stream));
        c->canonical_valid = true;
    }
    out->indices = c->buf.idx;
    out->positions = reinterpret_cast<const float *>(c->fp.n ? c->buf.pos_out : c->buf.pos);
    out->velocities = reinterpret_cast<const float *>(c->fp.n ? c->buf.vel_out : c->buf.vel);
    out->n_particles = c->fp.n;
    out->n_indices = (uint32_t)(c->ncell + 2);
    out->stream = c->stream;
    return WRACH_OK;
}

int wrach_store_write_back(wrach_ctx *c) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (c->worlds > 1) return fail(WRACH_E_INVALID, "store write-back is for a single world");
    wrach_device_view view{};
    {
        const int rc = wrach_canonical_device(c, &view);
        if (rc) return rc;
    }
    const uint32_t n = c->fp.n;
    if (c->host_pack) {
        // the frame's particles, then the old store's particles outside the grid
        std::vector<float> sp(2 * (size_t)n), sv(2 * (size_t)n);
        if (n) {
            WRACH_HIP(hipMemcpyAsync(sp.data(), view.positions, sp.size() * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream));
            WRACH_HIP(hipMemcpyAsync(sv.data(), view.velocities, sv.size() * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream));
        }
        WRACH_SETTLE(c);
        const uint32_t gx = c->settings.grid_dimensions[0], gy = c->settings.grid_dimensions[1];
        const float cs = (float)c->cfg.cell_size;
        for (size_t i = 0; i < c->store_pos.size() / 2; ++i) {
            int32_t cx, cy;
            cell_coord(c->store_pos[2 * i], c->store_pos[2 * i + 1], cs, cx, cy);
            if (cx >= 0 && cy >= 0 && (uint32_t)cx < gx && (uint32_t)cy < gy) continue;
            sp.insert(sp.end(), {c->store_pos[2 * i], c->store_pos[2 * i + 1]});
            sv.insert(sv.end(), {c->store_vel[2 * i], c->store_vel[2 * i + 1]});
        }
        c->store_pos.swap(sp);
        c->store_vel.swap(sv);
        g_err.clear();
        return WRACH_OK;
    }
    // Device store: new arrays holding the frame's particles, then the old
    // store's particles outside the grid.
    const size_t cap = store_cap_for((size_t)n + c->dstore_n);
    float2 *p, *v;
    uint2 *k;
    {
        const int rc = alloc_store(cap, &p, &v, &k);
        if (rc) return rc;
    }
    uint32_t outside = 0;
    hipError_t e = hipSuccess;
    if (n) {
        e = hipMemcpyAsync(p, view.positions, (size_t)n * sizeof(float2), hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(v, view.velocities, (size_t)n * sizeof(float2), hipMemcpyDeviceToDevice,
                               c->stream);
    }
    // The old store's particles outside the grid, in store order: per-block
    // counts, their exclusive offsets on the host, then the copy.
    const size_t blocks = (c->dstore_n + 255) / 256;
    uint32_t *dcounts = nullptr;
    std::vector<uint32_t> counts(blocks);
    if (blocks) {
        if (e == hipSuccess) e = hipMalloc(&dcounts, blocks * sizeof(uint32_t));
        if (e == hipSuccess) e = wrach::launch_store_outside_count(store_params(c, c->dstore_n), dcounts, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(counts.data(), dcounts, blocks * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) {
            for (size_t b = 0; b < blocks; ++b) {
                const uint32_t m = counts[b];
                counts[b] = outside;
                outside += m;
            }
            e = hipMemcpyAsync(dcounts, counts.data(), blocks * sizeof(uint32_t), hipMemcpyHostToDevice,
                               c->stream);
        }
        if (e == hipSuccess) e = wrach::launch_store_outside(store_params(c, c->dstore_n), p, v, n, dcounts, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dcounts);
    if (e != hipSuccess) {
        (void)hipFree(p);
        (void)hipFree(v);
        (void)hipFree(k);
        return fail(WRACH_E_DEVICE, std::string("store write-back: ") + hipGetErrorString(e));
    }
    swap_store(c, p, v, k, cap);
    c->dstore_n = (size_t)n + outside;
    g_err.clear();
    return WRACH_OK;
}

namespace {

void place_view(wrach_ctx *c, float x, float y) {
    const float cs = (float)c->cfg.cell_size;
    c->settings.view_anchor[0] = x;
    c->settings.view_anchor[1] = y;
    c->fp.ax = x;
    c->fp.ay = y;
    c->fp.x1 = x + c->settings.view_dimensions[0];
    c->fp.y1 = y + c->settings.view_dimensions[1];
    c->blx = (int32_t)(x / cs);
    c->bly = (int32_t)(y / cs);
}

}  // namespace

int wrach_set_view_anchor(wrach_ctx *c, float x, float y) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (c->worlds > 1) return fail(WRACH_E_INVALID, "the worlds of a batch keep their views at the origin");
    if (c->host_pack) return fail(WRACH_E_INVALID, "moving the view needs the device store (no WRACH_FLAG_HOST_PACK)");
    const float cs = (float)c->cfg.cell_size;
    const float w = c->settings.view_dimensions[0], h = c->settings.view_dimensions[1];
    // Integers below 2^24 keep x + W and every x - anchor exact in f32.
    if (!(x >= 0.0f) || !(y >= 0.0f) || !(x + w <= 16777216.0f) || !(y + h <= 16777216.0f))
        return fail(WRACH_E_INVALID, "the view must lie inside [0, 2^24] on both axes");
    if (std::fmod(x, cs) != 0.0f || std::fmod(y, cs) != 0.0f)
        return fail(WRACH_E_INVALID, "the view anchor must be a multiple of the cell size");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    const float ox = c->settings.view_anchor[0], oy = c->settings.view_anchor[1];
    if (x == ox && y == oy) return WRACH_OK;
    // The store takes the simulated particles and keeps those outside the
    // view in their order; the new view's particles are packed from it.
    int rc = wrach_store_write_back(c);
    if (rc) return rc;
    place_view(c, x, y);
    free_graphs(c);  // the captured frames hold the old walls
    c->canonical_valid = false;
    rc = add_device_packed(c, nullptr, nullptr, 0, 0);
    if (rc) {
        // Too many particles for the new view: the packed state is still the
        // old view's, which the written-back store matches.
        const std::string err = wrach_last_error();
        place_view(c, ox, oy);
        return fail(rc, err);
    }
    g_err.clear();
    return WRACH_OK;
}

int wrach_load_packed(wrach_ctx *c, const uint32_t *indices, size_t n_indices, const float *pos_xy,
                      const float *vel_xy, uint32_t n) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (c->worlds > 1) return fail(WRACH_E_INVALID, "a checkpoint restores a single world");
    if (!indices || (n > 0 && (!pos_xy || !vel_xy))) return fail(WRACH_E_INVALID, "null arrays");
    if (n_indices != c->ncell + 2)
        return fail(WRACH_E_INVALID, "indices must hold the grid cells + 2 offsets (" +
                                         std::to_string(c->ncell + 2) + ")");
    if (indices[0] != 0 || indices[1] != 0) return fail(WRACH_E_INVALID, "indices must start [0, 0, ...]");
    for (size_t i = 1; i + 1 < n_indices; ++i)
        if (indices[i + 1] < indices[i])
            return fail(WRACH_E_INVALID, "indices decrease at " + std::to_string(i + 1));
    if (indices[n_indices - 1] != n)
        return fail(WRACH_E_INVALID, "the last offset is not the particle count");
    if (n > c->capacity)
        return fail(WRACH_E_CAPACITY, "particles (" + std::to_string(n) + ") exceed the buffer capacity (" +
                                          std::to_string(c->capacity) + ")");
    for (size_t i = 0; i < 2 * (size_t)n; ++i)
        if (std::isnan(pos_xy[i]) || std::isnan(vel_xy[i]))
            return fail(WRACH_E_NONFINITE, "NaN in particle " + std::to_string(i / 2));
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    float2 *p = nullptr, *v = nullptr;
    uint2 *k = nullptr;
    size_t cap = c->dstore_cap;
    if (!c->host_pack && n > c->dstore_cap) {
        cap = store_cap_for(n);
        const int rc = alloc_store(cap, &p, &v, &k);
        if (rc) return rc;
    }
    WRACH_HIP(hipMemcpyAsync(c->buf.idx, indices, n_indices * sizeof(uint32_t), hipMemcpyHostToDevice,
                             c->stream));
    if (n) {
        WRACH_HIP(hipMemcpyAsync(c->buf.pos, pos_xy, (size_t)n * sizeof(float2), hipMemcpyHostToDevice,
                                 c->stream));
        WRACH_HIP(hipMemcpyAsync(c->buf.vel, vel_xy, (size_t)n * sizeof(float2), hipMemcpyHostToDevice,
                                 c->stream));
    }
    WRACH_HIP(wrach::launch_iota(c->buf.src, n, c->stream));
    WRACH_HIP(wrach::launch_keys_from_offsets(c->buf.idx, (uint32_t)c->ncell, n, c->buf.key, c->stream));
    WRACH_HIP(hipMemsetAsync(c->buf.chunk_live, 1, c->chunks, c->stream));  // offsets from the host
    if (c->host_pack) {
        c->store_pos.assign(pos_xy, pos_xy + 2 * (size_t)n);
        c->store_vel.assign(vel_xy, vel_xy + 2 * (size_t)n);
    } else {
        if (p) swap_store(c, p, v, k, cap);
        if (n) {
            WRACH_HIP(hipMemcpyAsync(c->dstore_pos, c->buf.pos, (size_t)n * sizeof(float2),
                                     hipMemcpyDeviceToDevice, c->stream));
            WRACH_HIP(hipMemcpyAsync(c->dstore_vel, c->buf.vel, (size_t)n * sizeof(float2),
                                     hipMemcpyDeviceToDevice, c->stream));
        }
        c->dstore_n = n;
    }
    WRACH_SETTLE(c);
    if (c->fp.n != n) free_graphs(c);
    c->fp.n = n;
    c->settings.particles_in_frame_count = n;
    c->canonical_valid = false;
    g_err.clear();
    return WRACH_OK;
}

namespace {

// Candidate grids of the cells kernel around the creation default g0.
void cells_grid_candidates(uint32_t g0, std::vector<uint32_t> &out) {
    out.clear();
    for (const uint32_t g : {g0, std::max(1u, g0 * 3 / 4), std::max(1u, g0 / 2), g0 + g0 / 2})
        if (std::find(out.begin(), out.end(), g) == out.end()) out.push_back(g);
}

wrach_tuning current_tuning(const wrach_ctx *c) {
    wrach_tuning t{};
    t.scan_threads = (uint32_t)wrach::kScanShapes[c->buf.scan_shape].threads;
    t.scan_items = (uint32_t)wrach::kScanShapes[c->buf.scan_shape].items;
    t.scatter_per_lane = (uint32_t)wrach::kScatterPerLane[c->buf.scatter_shape];
    t.cells_grid = c->buf.cells_grid;
    t.frame_us = 0.0f;
    return t;
}

void apply_shapes(wrach_ctx *c, uint32_t scan_shape, uint32_t scatter_shape, uint32_t grid) {
    if (c->buf.scan_shape != scan_shape || c->buf.scatter_shape != scatter_shape ||
        c->buf.cells_grid != grid)
        free_graphs(c);  // captured with the old shapes
    c->buf.scan_shape = scan_shape;
    c->buf.scatter_shape = scatter_shape;
    c->buf.cells_grid = grid;
    c->buf.scan_tiles = wrach::scan_tiles_for((uint32_t)c->ncell, scan_shape);
}

}  // namespace

int wrach_get_tuning(wrach_ctx *c, wrach_tuning *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    *out = current_tuning(c);
    return WRACH_OK;
}

int wrach_set_tuning(wrach_ctx *c, const wrach_tuning *t) {
    if (!c || !t) return fail(WRACH_E_INVALID, "null argument");
    int scan = -1, scatter = -1;
    for (int i = 0; i < wrach::kNumScanShapes; ++i)
        if ((uint32_t)wrach::kScanShapes[i].threads == t->scan_threads &&
            (uint32_t)wrach::kScanShapes[i].items == t->scan_items)
            scan = i;
    for (int i = 0; i < wrach::kNumScatterShapes; ++i)
        if ((uint32_t)wrach::kScatterPerLane[i] == t->scatter_per_lane) scatter = i;
    if (scan < 0) return fail(WRACH_E_INVALID, "no such scan tile shape");
    if (scatter < 0) return fail(WRACH_E_INVALID, "no such scatter width");
    if (t->cells_grid == 0 || t->cells_grid > 4 * c->cells_grid0)
        return fail(WRACH_E_INVALID, "cells grid outside 1.." + std::to_string(4 * c->cells_grid0));
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    apply_shapes(c, (uint32_t)scan, (uint32_t)scatter, t->cells_grid);
    return WRACH_OK;
}

int wrach_autotune(wrach_ctx *c, uint32_t frames, wrach_tuning *out) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (frames == 0) frames = 1;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    const uint32_t n = c->fp.n;
    if (n == 0) {  // nothing to time: an empty world does not change
        if (out) *out = current_tuning(c);
        return WRACH_OK;
    }
    // The state every trial starts from: offsets, positions, velocities,
    // source words and cell keys (counts, list counters and scan words are
    // clean between frames).
    const size_t nidx = c->ncell + 2;
    void *s_idx = nullptr, *s_pos = nullptr, *s_vel = nullptr, *s_src = nullptr, *s_key = nullptr,
         *s_live = nullptr, *s_pre = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(s_idx);
        (void)hipFree(s_pos);
        (void)hipFree(s_vel);
        (void)hipFree(s_src);
        (void)hipFree(s_key);
        (void)hipFree(s_live);
        (void)hipFree(s_pre);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
    };
    hipError_t e = hipMalloc(&s_idx, nidx * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s_pos, (size_t)n * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&s_vel, (size_t)n * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&s_src, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s_key, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s_live, c->chunks);
    if (e == hipSuccess) e = hipMalloc(&s_pre, c->chunks * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = copy(s_idx, c->buf.idx, nidx * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(s_pos, c->buf.pos, (size_t)n * sizeof(float2));
    if (e == hipSuccess) e = copy(s_vel, c->buf.vel, (size_t)n * sizeof(float2));
    if (e == hipSuccess) e = copy(s_src, c->buf.src, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(s_key, c->buf.key, (size_t)n * sizeof(uint32_t));
    if (e == hipSuccess) e = copy(s_live, c->buf.chunk_live, c->chunks);
    if (e == hipSuccess) e = copy(s_pre, c->buf.chunk_prefix, c->chunks * sizeof(uint32_t));
    if (e != hipSuccess) {
        cleanup();
        return fail(WRACH_E_DEVICE, std::string("autotune snapshot: ") + hipGetErrorString(e));
    }
    auto restore = [&]() {
        hipError_t r = copy(c->buf.idx, s_idx, nidx * sizeof(uint32_t));
        if (r == hipSuccess) r = copy(c->buf.pos, s_pos, (size_t)n * sizeof(float2));
        if (r == hipSuccess) r = copy(c->buf.vel, s_vel, (size_t)n * sizeof(float2));
        if (r == hipSuccess) r = copy(c->buf.src, s_src, (size_t)n * sizeof(uint32_t));
        if (r == hipSuccess) r = copy(c->buf.key, s_key, (size_t)n * sizeof(uint32_t));
        if (r == hipSuccess) r = copy(c->buf.chunk_live, s_live, c->chunks);
        if (r == hipSuccess) r = copy(c->buf.chunk_prefix, s_pre, c->chunks * sizeof(uint32_t));
        if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
        return r;
    };
    // Microseconds per frame of the current shapes, state restored after.
    auto trial = [&](float &us) -> int {
        int rc = enqueue_frame(c);
        if (rc) return rc;
        WRACH_HIP(hipEventRecord(e0, c->stream));
        for (uint32_t f = 0; f < frames && rc == WRACH_OK; ++f) rc = enqueue_frame(c);
        if (rc) return rc;
        WRACH_HIP(hipEventRecord(e1, c->stream));
        WRACH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        WRACH_HIP(hipEventElapsedTime(&ms, e0, e1));
        us = 1000.0f * ms / (float)frames;
        WRACH_HIP(restore());
        return WRACH_OK;
    };
    uint32_t scan = c->buf.scan_shape, scatter = c->buf.scatter_shape, grid = c->buf.cells_grid;
    float best_us = 0.0f;
    int rc = WRACH_OK;
    // One knob at a time, the others at their current best.
    for (int knob = 0; knob < 3 && rc == WRACH_OK; ++knob) {
        std::vector<uint32_t> cand;
        if (knob == 0)
            for (int i = 0; i < wrach::kNumScanShapes; ++i) cand.push_back((uint32_t)i);
        else if (knob == 1)
            for (int i = 0; i < wrach::kNumScatterShapes; ++i) cand.push_back((uint32_t)i);
        else
            cells_grid_candidates(c->cells_grid0, cand);
        uint32_t best = knob == 0 ? scan : knob == 1 ? scatter : grid;
        float best_here = -1.0f;
        for (const uint32_t v : cand) {
            c->buf.scan_shape = knob == 0 ? v : scan;
            c->buf.scatter_shape = knob == 1 ? v : scatter;
            c->buf.cells_grid = knob == 2 ? v : grid;
            c->buf.scan_tiles = wrach::scan_tiles_for((uint32_t)c->ncell, c->buf.scan_shape);
            float us = 0.0f;
            rc = trial(us);
            if (rc) break;
            if (best_here < 0.0f || us < best_here) {
                best_here = us;
                best = v;
            }
        }
        if (knob == 0) scan = best;
        else if (knob == 1) scatter = best;
        else grid = best;
        best_us = best_here;
    }
    if (rc) (void)restore();  // a failed trial: put the state back as far as the device allows
    cleanup();
    // The trials changed the shapes in place: the captured graphs are stale.
    free_graphs(c);
    c->buf.scan_shape = scan;
    c->buf.scatter_shape = scatter;
    c->buf.cells_grid = grid;
    c->buf.scan_tiles = wrach::scan_tiles_for((uint32_t)c->ncell, scan);
    c->canonical_valid = false;
    if (rc) return rc;
    if (out) {
        *out = current_tuning(c);
        out->frame_us = best_us;
    }
    g_err.clear();
    return WRACH_OK;
}

int wrach_read(wrach_ctx *c, float *pos_xy, float *vel_xy, size_t particle_cap,
               uint32_t *indices, size_t indices_cap, uint32_t *n_particles) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    const uint32_t n = c->fp.n;
    if ((pos_xy || vel_xy) && particle_cap < n)
        return fail(WRACH_E_BUFFER, "particle arrays hold " + std::to_string(particle_cap) +
                                        " particles, the frame has " + std::to_string(n));
    if (indices && indices_cap < c->ncell + 2)
        return fail(WRACH_E_BUFFER, "indices array too small");
    wrach_device_view v{};
    int rc = wrach_canonical_device(c, &v);
    if (rc) return rc;
    if (indices)
        WRACH_HIP(hipMemcpyAsync(indices, v.indices, (c->ncell + 2) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
    if (pos_xy && n)
        WRACH_HIP(hipMemcpyAsync(pos_xy, v.positions, (size_t)n * 2 * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
    if (vel_xy && n)
        WRACH_HIP(hipMemcpyAsync(vel_xy, v.velocities, (size_t)n * 2 * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    if (n_particles) *n_particles = n;
    return WRACH_OK;
}

namespace {

// Cells of world w's grid, and its first cell key in the batch's key space.
size_t world_cell_count(const wrach_ctx *c, uint32_t w) {
    if (!c->grids.empty()) return (size_t)c->grids[w].gx * c->grids[w].gy;
    return (size_t)c->world_cells;
}

size_t world_first_cell(const wrach_ctx *c, uint32_t w) {
    if (!c->grids.empty()) return c->grids[w].base;
    return (size_t)w * c->world_cells;
}

}  // namespace

int wrach_settings_world(wrach_ctx *c, uint32_t world, wrach_world_settings *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    if (world >= c->worlds)
        return fail(WRACH_E_INVALID, "no world " + std::to_string(world) + " in a batch of " +
                                         std::to_string(c->worlds));
    wrach_world_settings s = c->settings;
    if (!c->grids.empty()) {
        s.view_dimensions[0] = (float)c->world_cfg[world].width;
        s.view_dimensions[1] = (float)c->world_cfg[world].height;
        s.grid_dimensions[0] = c->grids[world].gx;
        s.grid_dimensions[1] = c->grids[world].gy;
    }
    // The world's particles: the offsets at its first cell and past its last.
    const size_t first = world_first_cell(c, world), wc = world_cell_count(c, world);
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_HIP(wrach::launch_materialize(c->fp, c->buf, c->stream));
    uint32_t ends[2] = {0, 0};
    WRACH_HIP(hipMemcpyAsync(&ends[0], c->buf.idx + 1 + first, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    WRACH_HIP(hipMemcpyAsync(&ends[1], c->buf.idx + 1 + first + wc, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    WRACH_SETTLE(c);
    s.particles_in_frame_count = ends[1] >= ends[0] ? ends[1] - ends[0] : 0u;
    *out = s;
    return WRACH_OK;
}

int wrach_read_world(wrach_ctx *c, uint32_t world, float *pos_xy, float *vel_xy, size_t particle_cap,
                     uint32_t *indices, size_t indices_cap, uint32_t *n_particles) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (world >= c->worlds)
        return fail(WRACH_E_INVALID, "no world " + std::to_string(world) + " in a batch of " +
                                         std::to_string(c->worlds));
    const size_t wc = world_cell_count(c, world), first = world_first_cell(c, world);
    if (indices && indices_cap < wc + 2) return fail(WRACH_E_BUFFER, "indices array too small");
    wrach_device_view v{};
    {
        const int rc = wrach_canonical_device(c, &v);
        if (rc) return rc;
    }
    // The world's offsets: entries 1 + first .. 1 + first + wc of the
    // batch's, less the first (the world's first particle).
    std::vector<uint32_t> g(wc + 1);
    WRACH_HIP(hipMemcpyAsync(g.data(), v.indices + 1 + first, (wc + 1) * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    if (g[0] > g[wc] || g[wc] > c->fp.n)
        return fail(WRACH_E_STATE, "world " + std::to_string(world) + ": offsets outside the particles");
    const uint32_t off = g[0], n = g[wc] - g[0];
    if ((pos_xy || vel_xy) && particle_cap < n)
        return fail(WRACH_E_BUFFER, "particle arrays hold " + std::to_string(particle_cap) +
                                        " particles, the world has " + std::to_string(n));
    if (indices) {
        indices[0] = 0;
        for (size_t k = 0; k <= wc; ++k) indices[1 + k] = g[k] - off;
    }
    if (pos_xy && n)
        WRACH_HIP(hipMemcpyAsync(pos_xy, v.positions + 2 * (size_t)off, (size_t)n * 2 * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
    if (vel_xy && n)
        WRACH_HIP(hipMemcpyAsync(vel_xy, v.velocities + 2 * (size_t)off, (size_t)n * 2 * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    if (n_particles) *n_particles = n;
    return WRACH_OK;
}

// ---------------------------------------------------------------------------
// The learned warm start (no counterpart in the reference).

extern "C++" {
namespace {

constexpr float kAdamBeta1 = 0.9f, kAdamBeta2 = 0.999f, kAdamEps = 1e-8f;

int ensure_learn(wrach_ctx *c) {
    if (!c->lparams) {
        float *q = nullptr;
        WRACH_HIP(dmalloc(&q, wrach::kParams));
        WRACH_HIP(hipMemset(q, 0, wrach::kParams * sizeof(float)));
        c->lparams = q;
    }
    for (float **q : {&c->lgrad, &c->lm, &c->lv}) {
        if (!*q) {
            float *t = nullptr;
            WRACH_HIP(dmalloc(&t, wrach::kParams));
            WRACH_HIP(hipMemset(t, 0, wrach::kParams * sizeof(float)));
            *q = t;
        }
    }
    if (!c->lw3) {
        float *t = nullptr;
        WRACH_HIP(dmalloc(&t, (size_t)wrach::kHid * wrach::kOutDim));
        c->lw3 = t;
    }
    if (!c->lfrag) {
        uint16_t *t = nullptr;
        WRACH_HIP(dmalloc(&t, (size_t)wrach::kFrags * wrach::kFragElems));
        c->lfrag = t;
    }
    return WRACH_OK;
}

// The workspace of a training batch of bsz rows (a multiple of 64).
int ensure_batch(wrach_ctx *c, uint32_t bsz) {
    if (c->tbsz >= bsz) return WRACH_OK;
    for (void *q : {(void *)c->txt, (void *)c->ta1t, (void *)c->tdz1t, (void *)c->tdz2t, (void *)c->ta2t,
                    (void *)c->tdyt, (void *)c->tsq, (void *)c->tpart1, (void *)c->tpart2})
        (void)hipFree(q);
    c->txt = c->ta1t = c->tdz1t = c->tdz2t = nullptr;
    c->ta2t = c->tdyt = c->tsq = c->tpart1 = c->tpart2 = nullptr;
    c->tbsz = 0;
    const size_t b = bsz;
    const size_t chunks = b / 32;  // the most k chunks a GEMM uses
    WRACH_HIP(dmalloc(&c->txt, (size_t)wrach::kFeat * b));
    WRACH_HIP(dmalloc(&c->ta1t, (size_t)wrach::kHid * b));
    WRACH_HIP(dmalloc(&c->tdz1t, (size_t)wrach::kHid * b));
    WRACH_HIP(dmalloc(&c->tdz2t, (size_t)wrach::kHid * b));
    WRACH_HIP(dmalloc(&c->ta2t, (size_t)wrach::kHid * b));
    WRACH_HIP(dmalloc(&c->tdyt, 2 * b));
    WRACH_HIP(dmalloc(&c->tsq, b / 16));
    WRACH_HIP(dmalloc(&c->tpart1, (size_t)wrach::kFeat * wrach::kHid * std::min<size_t>(chunks, 64)));
    WRACH_HIP(dmalloc(&c->tpart2, (size_t)wrach::kHid * wrach::kHid * std::min<size_t>(chunks, 64)));
    c->tbsz = bsz;
    return WRACH_OK;
}

// k chunk of the weight GEMMs over a batch of bsz rows (a multiple of 64):
// at most 64 chunks, each a multiple of 32, the last one possibly shorter.
uint32_t gemm_chunk(uint32_t bsz) {
    const uint32_t chunks = std::min<uint32_t>(64u, bsz / 32u);
    return (bsz / chunks + 31u) / 32u * 32u;
}

// One training step on the batch rows sel[0..bsz) of (feat, target):
// gradient into lgrad, the loss into *loss_dev (if not null).
int enqueue_gradient(wrach_ctx *c, const uint16_t *feat, const float2 *target, const uint32_t *sel,
                     uint32_t bsz, float *loss_dev) {
    wrach::TrainArgs t{};
    t.feat = feat;
    t.target = target;
    t.sel = sel;
    t.bsz = bsz;
    t.inv_bsz = 1.0f / (float)bsz;
    t.xt = c->txt;
    t.a1t = c->ta1t;
    t.dz1t = c->tdz1t;
    t.dz2t = c->tdz2t;
    t.a2t = c->ta2t;
    t.dyt = c->tdyt;
    t.sqerr = c->tsq;
    WRACH_HIP(wrach::launch_learn_rows(mlp_weights(c), t, c->stream));
    const uint32_t kc = gemm_chunk(bsz), chunks = (bsz + kc - 1) / kc;
    WRACH_HIP(wrach::launch_learn_gemm_nt(c->txt, c->tdz1t, wrach::kFeat, wrach::kHid, bsz, kc, chunks,
                                          c->tpart1, c->stream));
    WRACH_HIP(wrach::launch_learn_gemm_nt(c->ta1t, c->tdz2t, wrach::kHid, wrach::kHid, bsz, kc, chunks,
                                          c->tpart2, c->stream));
    WRACH_HIP(wrach::launch_learn_reduce(t, c->tpart1, c->tpart2, chunks, c->lgrad, loss_dev, c->stream));
    return WRACH_OK;
}

int enqueue_adam(wrach_ctx *c, float lr) {
    c->adam_t += 1;
    wrach::AdamArgs a{};
    a.params = c->lparams;
    a.m = c->lm;
    a.v = c->lv;
    a.grad = c->lgrad;
    a.lr = lr;
    a.beta1 = kAdamBeta1;
    a.beta2 = kAdamBeta2;
    a.eps = kAdamEps;
    a.c1 = (float)(1.0 / (1.0 - std::pow((double)kAdamBeta1, (double)c->adam_t)));
    a.c2 = (float)(1.0 / (1.0 - std::pow((double)kAdamBeta2, (double)c->adam_t)));
    WRACH_HIP(wrach::launch_learn_adam(a, c->stream));
    WRACH_HIP(wrach::launch_learn_pack(c->lparams, c->lfrag, c->lw3, c->stream));
    return WRACH_OK;
}

template <typename T>
struct DevTemp {  // a device buffer for the length of one call
    T *p = nullptr;
    ~DevTemp() { (void)hipFree(p); }
};

bool valid_batch(uint32_t n) { return n >= 64 && n % 64 == 0 && n <= (1u << 20); }

}  // namespace
}  // extern "C++"

int wrach_learn_set_params(wrach_ctx *c, const float *params, size_t n) {
    if (!c || !params) return fail(WRACH_E_INVALID, "null argument");
    if (n != wrach::kParams)
        return fail(WRACH_E_INVALID, "the network has " + std::to_string(wrach::kParams) + " parameters");
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return fail(WRACH_E_NONFINITE, "non-finite parameter");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    {
        const int rc = ensure_learn(c);
        if (rc) return rc;
    }
    WRACH_HIP(hipMemcpy(c->lparams, params, n * sizeof(float), hipMemcpyHostToDevice));
    WRACH_HIP(hipMemset(c->lm, 0, n * sizeof(float)));
    WRACH_HIP(hipMemset(c->lv, 0, n * sizeof(float)));
    c->adam_t = 0;
    WRACH_HIP(wrach::launch_learn_pack(c->lparams, c->lfrag, c->lw3, c->stream));
    WRACH_SETTLE(c);
    return WRACH_OK;
}

int wrach_learn_get_params(wrach_ctx *c, float *params, size_t n) {
    if (!c || !params) return fail(WRACH_E_INVALID, "null argument");
    if (n < wrach::kParams) return fail(WRACH_E_BUFFER, "parameter array too small");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    WRACH_HIP(hipMemcpy(params, c->lparams, wrach::kParams * sizeof(float), hipMemcpyDeviceToHost));
    return WRACH_OK;
}

int wrach_learn_features(wrach_ctx *c, uint16_t *rows, size_t rows_cap, uint64_t overlap[2]) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    const uint32_t n = c->fp.n;
    if (rows && rows_cap < n)
        return fail(WRACH_E_BUFFER, "feature array holds " + std::to_string(rows_cap) + " rows, the frame has " +
                                        std::to_string(n));
    WRACH_HIP(hipSetDevice(c->cfg.device));
    {
        const int rc = ensure_learned_frame(c);
        if (rc) return rc;
    }
    DevTemp<unsigned long long> sums;
    WRACH_HIP(dmalloc(&sums.p, (size_t)wrach::kStencilSumSlots * wrach::kStencilSumStride));
    WRACH_HIP(hipMemsetAsync(sums.p, 0,
                             (size_t)wrach::kStencilSumSlots * wrach::kStencilSumStride * sizeof(unsigned long long),
                             c->stream));
    // On the canonical copy, so that row i is canonical particle i (the
    // features themselves do not depend on the order inside a cell).
    wrach_device_view v{};
    {
        const int rc = wrach_canonical_device(c, &v);
        if (rc) return rc;
    }
    wrach::FeatureArgs f{};
    f.pos = reinterpret_cast<const float2 *>(v.positions);
    f.vel = reinterpret_cast<const float2 *>(v.velocities);
    f.feat = c->lfeat;
    f.sums = sums.p;
    WRACH_HIP(wrach::launch_learn_features(c->fp, c->buf, f, c->stream));
    if (rows && n)
        WRACH_HIP(hipMemcpyAsync(rows, c->lfeat, (size_t)n * wrach::kFeat * sizeof(uint16_t), hipMemcpyDeviceToHost,
                                 c->stream));
    std::vector<unsigned long long> hs((size_t)wrach::kStencilSumSlots * wrach::kStencilSumStride);
    WRACH_HIP(hipMemcpyAsync(hs.data(), sums.p, hs.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    WRACH_SETTLE(c);
    if (overlap) {
        overlap[0] = overlap[1] = 0;
        for (uint32_t slot = 0; slot < wrach::kStencilSumSlots; ++slot) {
            overlap[0] += hs[(size_t)slot * wrach::kStencilSumStride];
            overlap[1] += hs[(size_t)slot * wrach::kStencilSumStride + 1];
        }
    }
    return WRACH_OK;
}

int wrach_learn_forward(wrach_ctx *c, const uint16_t *rows, uint32_t n, float *y) {
    if (!c || (n && (!rows || !y))) return fail(WRACH_E_INVALID, "null argument");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    if (n == 0) return WRACH_OK;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    DevTemp<uint16_t> f;
    DevTemp<float2> out;
    WRACH_HIP(dmalloc(&f.p, (size_t)n * wrach::kFeat));
    WRACH_HIP(dmalloc(&out.p, (size_t)n));
    WRACH_HIP(hipMemcpyAsync(f.p, rows, (size_t)n * wrach::kFeat * sizeof(uint16_t), hipMemcpyHostToDevice,
                             c->stream));
    wrach::InferArgs in{};
    in.feat = f.p;
    in.n = n;
    in.y_out = out.p;
    WRACH_HIP(wrach::launch_learn_infer(mlp_weights(c), in, c->stream));
    WRACH_HIP(hipMemcpyAsync(y, out.p, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    return WRACH_OK;
}

int wrach_learn_gradient(wrach_ctx *c, const uint16_t *rows, const float *targets, uint32_t n, float *grad,
                         float *loss) {
    if (!c || !rows || !targets || !grad) return fail(WRACH_E_INVALID, "null argument");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    if (!valid_batch(n)) return fail(WRACH_E_INVALID, "a batch is a positive multiple of 64 rows, at most 2^20");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    {
        const int rc = ensure_batch(c, n);
        if (rc) return rc;
    }
    DevTemp<uint16_t> f;
    DevTemp<float2> t;
    DevTemp<uint32_t> sel;
    DevTemp<float> l;
    WRACH_HIP(dmalloc(&f.p, (size_t)n * wrach::kFeat));
    WRACH_HIP(dmalloc(&t.p, (size_t)n));
    WRACH_HIP(dmalloc(&sel.p, (size_t)n));
    WRACH_HIP(dmalloc(&l.p, 1));
    std::vector<uint32_t> iota(n);
    for (uint32_t i = 0; i < n; ++i) iota[i] = i;
    WRACH_HIP(hipMemcpyAsync(f.p, rows, (size_t)n * wrach::kFeat * sizeof(uint16_t), hipMemcpyHostToDevice,
                             c->stream));
    WRACH_HIP(hipMemcpyAsync(t.p, targets, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    WRACH_HIP(hipMemcpyAsync(sel.p, iota.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    {
        const int rc = enqueue_gradient(c, f.p, t.p, sel.p, n, l.p);
        if (rc) return rc;
    }
    WRACH_HIP(hipMemcpyAsync(grad, c->lgrad, wrach::kParams * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    float hl = 0.0f;
    WRACH_HIP(hipMemcpyAsync(&hl, l.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    if (loss) *loss = hl;
    return WRACH_OK;
}

int wrach_learn_adam(wrach_ctx *c, const float *grad, float lr) {
    if (!c || !grad) return fail(WRACH_E_INVALID, "null argument");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    if (!(lr > 0.0f) || std::isinf(lr)) return fail(WRACH_E_INVALID, "learning rate must be finite and > 0");
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_HIP(hipMemcpyAsync(c->lgrad, grad, wrach::kParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
    {
        const int rc = enqueue_adam(c, lr);
        if (rc) return rc;
    }
    WRACH_SETTLE(c);
    return WRACH_OK;
}

int wrach_learn_sample(wrach_ctx *c, uint32_t target_passes, uint64_t *rows_total) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    if (target_passes == 0 || target_passes > wrach::kMaxStencilPasses)
        return fail(WRACH_E_INVALID, "target passes must be 1.." + std::to_string(wrach::kMaxStencilPasses));
    const uint32_t n = c->fp.n;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    {
        int rc = ensure_stencil(c);
        if (!rc) rc = ensure_learned_frame(c);
        if (rc) return rc;
    }
    if (c->trows + n > c->tcap) {  // grow the training set, keeping its rows
        size_t cap = c->tcap ? c->tcap : 1u << 20;
        while (cap < c->trows + n) cap *= 2;
        uint16_t *f = nullptr;
        float2 *t = nullptr;
        hipError_t e = hipMalloc((void **)&f, cap * wrach::kFeat * sizeof(uint16_t) + 256);
        if (e == hipSuccess) e = hipMalloc((void **)&t, cap * sizeof(float2) + 256);
        if (e == hipSuccess && c->trows)
            e = hipMemcpy(f, c->tfeat, c->trows * wrach::kFeat * sizeof(uint16_t), hipMemcpyDeviceToDevice);
        if (e == hipSuccess && c->trows)
            e = hipMemcpy(t, c->ttarget, c->trows * sizeof(float2), hipMemcpyDeviceToDevice);
        if (e != hipSuccess) {
            (void)hipFree(f);
            (void)hipFree(t);
            return fail(WRACH_E_DEVICE, std::string("training set: ") + hipGetErrorString(e));
        }
        (void)hipFree(c->tfeat);
        (void)hipFree(c->ttarget);
        c->tfeat = f;
        c->ttarget = t;
        c->tcap = cap;
    }
    if (n) {
        c->canonical_valid = false;  // pos_out is used below
        wrach::FeatureArgs f{};
        f.pos = c->buf.pos;
        f.vel = c->buf.vel;
        f.feat = c->tfeat;
        f.row0 = (uint32_t)c->trows;
        f.cell = c->scell;
        WRACH_HIP(wrach::launch_learn_features(c->fp, c->buf, f, c->stream));
        // The plain passes from the state, which stays as it is: pos ->
        // pos_out -> lbuf -> pos_out ...
        wrach::StencilArgs a = stencil_args(c);
        WRACH_HIP(wrach::launch_stencil_clear(c->fp, c->buf, a, c->buf.pos, c->stream));
        const float2 *in = c->buf.pos;
        for (uint32_t k = 0; k < target_passes; ++k) {
            float2 *out = k % 2 == 0 ? c->buf.pos_out : c->lbuf;
            a.pin = in;
            a.pout = out;
            a.in_slot = k;
            WRACH_HIP(wrach::launch_stencil_move(c->fp, c->buf, a, c->stream));
            in = out;
        }
        WRACH_HIP(wrach::launch_learn_target(in, c->buf.pos, c->ttarget, (uint32_t)c->trows, n, c->stream));
    }
    WRACH_SETTLE(c);
    c->trows += n;
    if (rows_total) *rows_total = c->trows;
    return WRACH_OK;
}

int wrach_learn_clear_samples(wrach_ctx *c) {
    if (!c) return fail(WRACH_E_INVALID, "null context");
    c->trows = 0;
    return WRACH_OK;
}

int wrach_learn_train(wrach_ctx *c, const uint32_t *order, size_t n_order, uint32_t batch, float lr,
                      float *losses) {
    if (!c || !order) return fail(WRACH_E_INVALID, "null argument");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    if (!valid_batch(batch)) return fail(WRACH_E_INVALID, "a batch is a positive multiple of 64 rows, at most 2^20");
    if (!(lr > 0.0f) || std::isinf(lr)) return fail(WRACH_E_INVALID, "learning rate must be finite and > 0");
    for (size_t i = 0; i < n_order; ++i)
        if (order[i] >= c->trows) return fail(WRACH_E_INVALID, "a batch row names no training sample");
    const size_t steps = n_order / batch;
    if (steps == 0) return WRACH_OK;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    WRACH_SETTLE(c);
    {
        const int rc = ensure_batch(c, batch);
        if (rc) return rc;
    }
    if (c->tsel_cap < steps * batch) {
        (void)hipFree(c->tsel);
        c->tsel = nullptr;
        c->tsel_cap = 0;
        WRACH_HIP(dmalloc(&c->tsel, steps * batch));
        c->tsel_cap = steps * batch;
    }
    if (c->tloss_cap < steps) {
        (void)hipFree(c->tloss);
        c->tloss = nullptr;
        c->tloss_cap = 0;
        WRACH_HIP(dmalloc(&c->tloss, steps));
        c->tloss_cap = steps;
    }
    WRACH_HIP(hipMemcpyAsync(c->tsel, order, steps * batch * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    for (size_t s = 0; s < steps; ++s) {
        int rc = enqueue_gradient(c, c->tfeat, c->ttarget, c->tsel + s * batch, batch, c->tloss + s);
        if (!rc) rc = enqueue_adam(c, lr);
        if (rc) return rc;
    }
    if (losses)
        WRACH_HIP(hipMemcpyAsync(losses, c->tloss, steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    WRACH_SETTLE(c);
    return WRACH_OK;
}

int wrach_learn_profile(wrach_ctx *c, uint32_t reps, double *ms) {
    if (!c || !ms) return fail(WRACH_E_INVALID, "null argument");
    if (!c->lparams) return fail(WRACH_E_STATE, "no parameters set");
    ms[0] = ms[1] = 0.0;
    if (reps == 0 || c->fp.n == 0) return WRACH_OK;
    WRACH_HIP(hipSetDevice(c->cfg.device));
    {
        const int rc = ensure_learned_frame(c);
        if (rc) return rc;
    }
    WRACH_SETTLE(c);
    std::vector<hipEvent_t> ev(3 * (size_t)reps, nullptr);
    for (auto &e : ev) WRACH_HIP(hipEventCreate(&e));
    int rc = WRACH_OK;
    wrach::FeatureArgs f{};
    f.pos = c->buf.pos;
    f.vel = c->buf.vel;
    f.feat = c->lfeat;
    wrach::InferArgs in{};
    in.feat = c->lfeat;
    in.n = c->fp.n;
    in.pin = c->buf.pos;
    in.pout = c->lbuf;  // the state does not change
    for (uint32_t r = 0; r < reps && rc == WRACH_OK; ++r) {
        hipError_t e = hipEventRecord(ev[3 * r], c->stream);
        if (e == hipSuccess) e = wrach::launch_learn_features(c->fp, c->buf, f, c->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[3 * r + 1], c->stream);
        if (e == hipSuccess) e = wrach::launch_learn_infer(mlp_weights(c), in, c->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[3 * r + 2], c->stream);
        if (e != hipSuccess) rc = fail(WRACH_E_DEVICE, std::string("learn profile: ") + hipGetErrorString(e));
    }
    if (rc == WRACH_OK) rc = settle(c);
    for (uint32_t r = 0; r < reps && rc == WRACH_OK; ++r) {
        float a = 0.0f, b = 0.0f;
        hipError_t e = hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]);
        if (e != hipSuccess) rc = fail(WRACH_E_DEVICE, std::string("hipEventElapsedTime: ") + hipGetErrorString(e));
        ms[0] += a;
        ms[1] += b;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    c->canonical_valid = false;
    return rc;
}

int wrach_learned_stats(wrach_ctx *c, wrach_learned_report *out) {
    if (!c || !out) return fail(WRACH_E_INVALID, "null argument");
    wrach_learned_report r{};
    if (c->swords && c->ssums) {
        WRACH_HIP(hipSetDevice(c->cfg.device));
        WRACH_SETTLE(c);
        std::vector<unsigned long long> sums(wrach::kStencilSumEntries, 0ull);
        uint32_t fb = 0;
        WRACH_HIP(hipMemcpy(sums.data(), c->ssums, sums.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        WRACH_HIP(hipMemcpy(&fb, c->swords + wrach::kFallbackWord, sizeof fb, hipMemcpyDeviceToHost));
        for (uint32_t slot = 0; slot < wrach::kStencilSumSlots; ++slot) {
            const size_t e0 = (size_t)slot * wrach::kStencilSumStride;
            const size_t e2 = ((size_t)2 * wrach::kStencilSumSlots + slot) * wrach::kStencilSumStride;
            r.start_overlap_sum += sums[e0];
            r.start_pairs += sums[e0 + 1];
            r.warm_overlap_sum += sums[e2];
            r.warm_pairs += sums[e2 + 1];
        }
        r.fell_back = fb;
    }
    *out = r;
    return WRACH_OK;
}

// ---------------------------------------------------------------------------
// One frame with no allocation (no counterpart in the reference): for
// callers that cannot allocate device memory.  Every buffer beyond the
// caller's packed triple lives in `scratch`, a file-scope device array that
// each GPU holds until the process ends; the host reaches it only through
// the address hipGetSymbolAddress returns for the current device.

}  // extern "C"

namespace {

constexpr size_t kNoAllocBytes = WRACH_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocDevices] = {};

// The scratch array's address on the current device (looked up once per
// device, with that device current).
int scratch_address(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kNoAllocDevices) return fail(WRACH_E_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(WRACH_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return WRACH_OK;
}

// Carves the frame's buffers for n particles and ncell cells out of `base`
// (256-byte aligned pieces); returns the bytes used (base may be null to
// size only).
size_t carve_noalloc(unsigned char *base, uint32_t n, uint64_t ncell, wrach::FrameBuffers &b,
                     size_t &ncells_total) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t np = n ? n : 1;
    b.idx = reinterpret_cast<uint32_t *>(take((ncell + 2 + 64) * sizeof(uint32_t)));
    if (b.idx) b.idx += 3;  // &idx[1] on a 16-byte boundary
    b.pos = reinterpret_cast<float2 *>(take(np * sizeof(float2)));
    b.vel = reinterpret_cast<float2 *>(take(np * sizeof(float2)));
    b.src = reinterpret_cast<uint32_t *>(take(np * sizeof(uint32_t)));
    b.key = reinterpret_cast<uint32_t *>(take(np * sizeof(uint32_t)));
    b.pos_out = reinterpret_cast<float2 *>(take(np * sizeof(float2)));
    b.vel_out = reinterpret_cast<float2 *>(take(np * sizeof(float2)));
    b.kslot = reinterpret_cast<uint2 *>(take(np * sizeof(uint2)));
    b.sortbuf = reinterpret_cast<unsigned long long *>(take(np * sizeof(unsigned long long)));
    b.cnt = reinterpret_cast<uint32_t *>(take((ncell + 16) * sizeof(uint32_t)));
    const size_t tiles = (ncell + wrach::kMinScanTile - 1) / wrach::kMinScanTile;
    b.tile_state = reinterpret_cast<unsigned long long *>(take(tiles * sizeof(unsigned long long)));
    const size_t chunks = (ncell + wrach::kChunkCells - 1) / wrach::kChunkCells;
    b.chunk_next = take(chunks);
    b.chunk_live = take(chunks);
    b.chunk_prefix = reinterpret_cast<uint32_t *>(take(chunks * sizeof(uint32_t)));
    b.epoch = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t)));
    b.error = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t)));
    b.canon_ctr = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t)));
    b.class_ctr = reinterpret_cast<uint32_t *>(take(wrach::kCtrWords * sizeof(uint32_t)));
    ncells_total = 0;
    for (int k = 0; k < wrach::kClasses; ++k) {
        b.class_cap[k] = (uint32_t)(np / wrach::kClassMinN[k] + 64);
        b.slice_cap[k] = (b.class_cap[k] + wrach::kSlices - 1) / wrach::kSlices + 16;
        b.class_off[k] = (uint32_t)ncells_total;
        ncells_total += (size_t)wrach::kSlices * b.slice_cap[k] + b.class_cap[k];
    }
    b.cells = reinterpret_cast<uint4 *>(take(ncells_total * sizeof(uint4)));
    return off;
}

}  // namespace

extern "C" {

int wrach_frame_noalloc(const wrach_world_settings *ws, const float gravity[2], float dt,
                        uint32_t *indices, float *pos_xy, float *vel_xy, uint32_t n, void *stream) {
    if (!ws || !gravity || !indices || (n && (!pos_xy || !vel_xy)))
        return fail(WRACH_E_INVALID, "null argument");
    if (n != ws->particles_in_frame_count)
        return fail(WRACH_E_INVALID, "n differs from the settings' particles_in_frame_count");
    if (ws->cell_size == 0 || ws->grid_dimensions[0] == 0 || ws->grid_dimensions[1] == 0)
        return fail(WRACH_E_INVALID, "cell size and grid dimensions must be >= 1");
    if (!std::isfinite(gravity[0]) || !std::isfinite(gravity[1]) || !std::isfinite(dt) ||
        !std::isfinite(ws->view_anchor[0]) || !std::isfinite(ws->view_anchor[1]) ||
        !std::isfinite(ws->view_dimensions[0]) || !std::isfinite(ws->view_dimensions[1]))
        return fail(WRACH_E_INVALID, "gravity, dt and the view must be finite");
    if (n > 0x7fffffffu) return fail(WRACH_E_CAPACITY, "more than 2^31 - 1 particles");
    const uint64_t gx = ws->grid_dimensions[0], gy = ws->grid_dimensions[1];
    const uint64_t ncell = gx * gy;
    if (ncell > kMaxCells) return fail(WRACH_E_TOO_MANY_CELLS, "grid has more cells than the 32-bit offsets allow");
    wrach::FrameBuffers b{};
    size_t ncells_total = 0;
    const size_t need = carve_noalloc(nullptr, n, ncell, b, ncells_total);
    if (need > kNoAllocBytes)
        return fail(WRACH_E_CAPACITY, "the frame needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                          std::to_string(kNoAllocBytes) + " the library holds");
    int dev = 0;
    WRACH_HIP(hipGetDevice(&dev));
    unsigned char *base = nullptr;
    {
        const int rc = scratch_address(dev, &base);
        if (rc) return rc;
    }
    carve_noalloc(base, n, ncell, b, ncells_total);
    b.scan_shape = wrach::default_scan_shape(ncell);
    b.scatter_shape = 0;
    b.scan_tiles = wrach::scan_tiles_for((uint32_t)ncell, b.scan_shape);
    {
        const uint32_t want = (b.class_cap[0] * 2 / 64 + wrach::kWavesPerBlock) / wrach::kWavesPerBlock;
        const uint32_t resident = wrach::resident_blocks_cells();
        if (resident == 0) return fail(WRACH_E_DEVICE, "occupancy query failed");
        b.cells_grid = want < resident ? want : resident;
        if (b.cells_grid == 0) b.cells_grid = 1;
    }
    wrach::FrameParams fp{};
    fp.ax = ws->view_anchor[0];
    fp.ay = ws->view_anchor[1];
    fp.x1 = fp.ax + ws->view_dimensions[0];
    fp.y1 = fp.ay + ws->view_dimensions[1];
    fp.cs = (float)ws->cell_size;
    fp.gdx = gravity[0] != 0.0f ? gravity[0] * dt : -0.0f;
    fp.gdy = gravity[1] != 0.0f ? gravity[1] * dt : -0.0f;
    fp.dt = dt;
    fp.gx = (uint32_t)gx;
    fp.gy = (uint32_t)gy;
    fp.ncell = (uint32_t)ncell;
    fp.n = n;
    fp.world_cells = (uint32_t)ncell;
    fp.wait_limit = wrach::kDefaultWaitLimit;
    fp.epoch_skew = 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t chunks = (ncell + wrach::kChunkCells - 1) / wrach::kChunkCells;
    const size_t tiles = (ncell + wrach::kMinScanTile - 1) / wrach::kMinScanTile;
    // The array keeps the last call's contents: clear what starts at zero.
    WRACH_HIP(wrach::launch_fill32(b.cnt, 0u, ncell + 16, st));
    WRACH_HIP(wrach::launch_fill32(reinterpret_cast<uint32_t *>(b.tile_state), 0u, 2 * tiles, st));
    WRACH_HIP(wrach::launch_fill32(reinterpret_cast<uint32_t *>(b.chunk_next), 0u, (chunks + 3) / 4, st));
    WRACH_HIP(wrach::launch_fill32(reinterpret_cast<uint32_t *>(b.chunk_live), 0x01010101u, (chunks + 3) / 4, st));
    WRACH_HIP(wrach::launch_fill32(b.epoch, 1u, 1, st));  // status words start at 0, never valid
    WRACH_HIP(wrach::launch_fill32(b.error, 0u, 1, st));
    WRACH_HIP(wrach::launch_fill32(b.canon_ctr, 0u, 1, st));
    WRACH_HIP(wrach::launch_fill32(b.class_ctr, 0u, wrach::kCtrWords, st));
    // The caller's packed triple becomes the state (each cell one run).
    WRACH_HIP(wrach::launch_copy32(b.idx, indices, ncell + 2, st));
    if (n) {
        WRACH_HIP(wrach::launch_copy32(reinterpret_cast<uint32_t *>(b.pos), reinterpret_cast<const uint32_t *>(pos_xy),
                                       2 * (size_t)n, st));
        WRACH_HIP(wrach::launch_copy32(reinterpret_cast<uint32_t *>(b.vel), reinterpret_cast<const uint32_t *>(vel_xy),
                                       2 * (size_t)n, st));
        WRACH_HIP(wrach::launch_iota(b.src, n, st));
        WRACH_HIP(wrach::launch_keys_from_offsets(b.idx, (uint32_t)ncell, n, b.key, st));
        WRACH_HIP(wrach::launch_collide(fp, b, st));
        WRACH_HIP(wrach::launch_cells(fp, b, st));
        WRACH_HIP(wrach::launch_scan(fp, b, st));
        WRACH_HIP(wrach::launch_scatter(fp, b, st));
        WRACH_HIP(wrach::launch_materialize(fp, b, st));
        WRACH_HIP(wrach::launch_canonical(fp, b, st));
        WRACH_HIP(wrach::launch_copy32(reinterpret_cast<uint32_t *>(pos_xy), reinterpret_cast<const uint32_t *>(b.pos_out),
                                       2 * (size_t)n, st));
        WRACH_HIP(wrach::launch_copy32(reinterpret_cast<uint32_t *>(vel_xy), reinterpret_cast<const uint32_t *>(b.vel_out),
                                       2 * (size_t)n, st));
        WRACH_HIP(wrach::launch_copy32(indices, b.idx, ncell + 2, st));
    }
    return WRACH_OK;
}

int wrach_check_pair_arithmetic(int device, uint64_t out[4]) {
    if (!out) return fail(WRACH_E_INVALID, "null argument");
    int ndev = 0;
    WRACH_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(WRACH_E_INVALID, "no such device");
    WRACH_HIP(hipSetDevice(device));
    unsigned long long *d = nullptr;
    WRACH_HIP(dmalloc(&d, 4));
    unsigned long long h[4] = {};
    hipError_t e = hipMemset(d, 0, sizeof h);
    if (e == hipSuccess) e = wrach::launch_check_pair_arithmetic(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);  // waits for the kernel
    (void)hipFree(d);
    if (e != hipSuccess) return fail(WRACH_E_DEVICE, std::string("pair arithmetic check: ") + hipGetErrorString(e));
    for (int k = 0; k < 4; ++k) out[k] = h[k];
    g_err.clear();
    return WRACH_OK;
}

int wrach_worlds(wrach_ctx *c, uint32_t *n_worlds) {
    if (!c || !n_worlds) return fail(WRACH_E_INVALID, "null argument");
    *n_worlds = c->worlds;
    return WRACH_OK;
}

}  // extern "C"
