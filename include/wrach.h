/*
 * wrach.h -- C ABI of the MI355X (gfx950) particle step.
 *
 * Drop-in boundary for wrach's `runners/api` surface: the Rust host's
 * `WrachAPI` (runners/api/src/lib.rs:17-88) and the wgpu compute worker it
 * drives (runners/bevy/src/compute/builder.rs:24-92, plugin/build.rs:88-158)
 * are replaced by these entry points.  Plain pointers and sizes only; the
 * library copies host arrays in and out and retains no caller pointer.
 *
 * Errors: every int-returning call returns WRACH_OK (0) or a negative
 * wrach_status; wrach_last_error() then describes the failure.  Nothing
 * unwinds across the ABI.  Where the reference panics (an assert or an
 * `expect`), the matching call returns an error code instead.
 *
 * Threading: a context is single-owner, like `&mut self` on WrachAPI
 * (runners/api/src/lib.rs:49,78).  Calls on one context must be serialised
 * by the caller; different contexts are independent.  wrach_step is
 * asynchronous on the context's HIP stream; wrach_read and wrach_sync
 * synchronise it.
 */
#ifndef WRACH_H
#define WRACH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WRACH_ABI_VERSION 1

typedef struct wrach_ctx wrach_ctx;

typedef enum wrach_status {
    WRACH_OK = 0,
    WRACH_E_INVALID = -1,        /* null pointer, zero cell size, bad flag */
    WRACH_E_TOO_MANY_CELLS = -2, /* grid too large (builder.rs:39-42 asserts) */
    WRACH_E_CAPACITY = -3,       /* more particles than the buffers hold */
    WRACH_E_NONFINITE = -4,      /* NaN position or velocity in an upload */
    WRACH_E_DEVICE = -5,         /* HIP runtime error, or a device wait that gave up */
    WRACH_E_BUFFER = -6,         /* caller's output array too small */
    WRACH_E_STATE = -7           /* device state failed a consistency check */
} wrach_status;

/* Config flags.  WRACH_FLAG_HOST_PACK: wrach_add_particles packs the store on
 * the host, as ParticleStore::create_packed_data does (particle_store.rs:
 * 90-103); without it the store is binned on the device (same packed
 * triple, state.rs:90-101 semantics either way). */
#define WRACH_FLAG_HOST_PACK 0x1u

/* WrachConfig (runners/bevy/src/config_app.rs:10-35) plus extensions.
 * gravity = (0,0) and dt = 1 give the reference's step bit for bit. */
typedef struct wrach_config {
    uint16_t width;                   /* dimensions.0 */
    uint16_t height;                  /* dimensions.1 */
    uint16_t cell_size;               /* spatial-bin cell side, >= 1 */
    uint8_t boundaries_as_dimensions; /* accepted and unused, as in the reference */
    uint8_t reserved0;
    float gravity[2];                 /* extension: v += g*dt before integrate */
    float dt;                         /* extension: p += v*dt */
    uint32_t flags;                   /* WRACH_FLAG_* bits; 0 for the defaults */
    int32_t device;                   /* HIP device ordinal */
} wrach_config;

/* ShaderWorldSettings, the 32-byte uniform (runners/bevy/src/
 * config_shader.rs:15-29, shaders/shared/src/lib.rs:19-31). */
typedef struct wrach_world_settings {
    float view_dimensions[2];
    float view_anchor[2];
    uint32_t grid_dimensions[2];
    uint32_t cell_size;
    uint32_t particles_in_frame_count;
} wrach_world_settings;

/* Device-resident state, for zero-copy consumers (a renderer, a caller that
 * keeps its own stream).  Inside each cell the particles are a permutation
 * of the canonical order until wrach_read (or wrach_canonical_device) runs. */
typedef struct wrach_device_view {
    const uint32_t *indices;   /* grid cells + 2 entries */
    const float *positions;    /* 2 floats per particle */
    const float *velocities;   /* 2 floats per particle */
    uint32_t n_particles;
    uint32_t n_indices;
    void *stream;              /* hipStream_t the frames run on */
} wrach_device_view;

/* Number of kernels timed by wrach_step_profiled, in launch order:
 * 0 collide/integrate/key (one-particle cells; fuller cells are queued),
 * 1 the queued cells (two or more particles), 2 exclusive scan, 3 stable
 * scatter. */
#define WRACH_NUM_FRAME_KERNELS 4

/* WrachConfig::default (config_app.rs:26-35): 480x352, cell size 3
 * (SPATIAL_BIN_CELL_SIZE, shaders/shared/src/lib.rs:34). */
void wrach_config_default(wrach_config *cfg);

/* WrachAPI::new (runners/api/src/lib.rs:30-45) + PhysicsComputeWorker::build
 * (runners/bevy/src/compute/builder.rs:24-92): grid from the viewport
 * (0,0,W,H), buffer capacity max_particles_per_frame
 * (runners/bevy/src/particle_store.rs:116-133), zeroed device buffers. */
int wrach_create(const wrach_config *cfg, wrach_ctx **out);

/* Drop of WrachAPI: frees device buffers, stream and graphs.  NULL is a no-op. */
void wrach_destroy(wrach_ctx *ctx);

/* WrachAPI::add_particles (runners/api/src/lib.rs:78-81) ->
 * WrachState::add_particles (runners/bevy/src/state.rs:90-101): appends n
 * particles (pos_xy, vel_xy: 2n floats each) to the store, re-packs the
 * whole store (particle_store.rs:54-103, spatial_bin.rs:103-149; particles
 * outside the grid stay in the store but are not simulated), replacing the
 * device state.  The store lives on the device and only the n new particles
 * cross PCIe; with WRACH_FLAG_HOST_PACK it lives on the host and the packed
 * arrays are uploaded.  On error the store is left unchanged. */
int wrach_add_particles(wrach_ctx *ctx, const float *pos_xy,
                        const float *vel_xy, uint32_t n);

/* The compute part of WrachAPI::tick / app.update() (runners/api/src/
 * lib.rs:49-52; passes chained at builder.rs:86-89): nframes frames of
 * collide + integrate + stable re-bin.  Asynchronous. */
int wrach_step(wrach_ctx *ctx, uint32_t nframes);

/* The readback of the `tick` system (plugin/build.rs:135-158) and
 * WrachAPI::read_data (runners/api/src/lib.rs:58-74): synchronises, then
 * copies the packed triple in canonical order.  Any output pointer may be
 * NULL.  pos_xy/vel_xy receive *n_particles particles (particle_cap is in
 * particles); indices receives the grid cells + 2 offsets
 * [0, start_0, ..., start_{C-1}, N].  Returns WRACH_E_BUFFER if a non-NULL
 * array is too small. */
int wrach_read(wrach_ctx *ctx, float *pos_xy, float *vel_xy,
               size_t particle_cap, uint32_t *indices, size_t indices_cap,
               uint32_t *n_particles);

/* Buffer sizes of PhysicsComputeWorker::build (builder.rs:26-55):
 * max_particles = max_particles_per_frame(), n_indices = cells + 2. */
int wrach_capacity(wrach_ctx *ctx, uint32_t *max_particles,
                   uint32_t *n_indices);

/* Current ShaderWorldSettings (WrachState.shader_settings, state.rs:22). */
int wrach_settings(wrach_ctx *ctx, wrach_world_settings *out);

/* Blocks until all submitted frames have finished. */
int wrach_sync(wrach_ctx *ctx);

/* Device pointers of the live state (no copy, no synchronisation).  The
 * frames leave the offsets of empty 1024-cell chunks as they were; this call
 * enqueues their fill on the context's stream, so work ordered after it on
 * that stream sees every offset. */
int wrach_device_state(wrach_ctx *ctx, wrach_device_view *out);

/* Writes the canonical (in-cell source order) copy of the state into
 * device staging buffers and returns them in *out; asynchronous.  The
 * pointers stay valid until the next wrach_step / wrach_read. */
int wrach_canonical_device(wrach_ctx *ctx, wrach_device_view *out);

/* ParticleStore::update_from_gpu, a TODO in the reference
 * (particle_store.rs:74-85, with cells_to_read_from_gpu, :22-26): the store
 * becomes the simulated state -- the frame's particles in canonical packed
 * order, then the store's particles outside the grid in their store order
 * (they are not simulated until a moved view covers them) -- so that a
 * later wrach_add_particles continues from the
 * simulated state instead of restarting from the uploads.  The device state
 * is unchanged.  Synchronises. */
int wrach_store_write_back(wrach_ctx *ctx);

/* Restores a packed triple saved with wrach_read (PackedData,
 * spatial_bin.rs:22-33; the TODO "Save GPU data to CPU memory and disk",
 * particle_store.rs:74-76).  n_indices must be the grid cells + 2 and
 * indices [0, 0, ...] non-decreasing up to n_particles <= capacity; NaN is
 * rejected as in wrach_add_particles.  The state and the store become the
 * triple's particles (the frames take each index range as that cell's
 * particles, as the reference's pass does).  On error nothing changes. */
int wrach_load_packed(wrach_ctx *ctx, const uint32_t *indices, size_t n_indices,
                      const float *pos_xy, const float *vel_xy, uint32_t n_particles);

/* A batch of worlds (no counterpart in the reference, which runs one world
 * per Bevy App, runners/api/src/lib.rs:30-45): n_worlds independent worlds
 * of one configuration, stepped together by wrach_step -- one launch per
 * frame kernel for all of them over a flattened (world, cell) key space,
 * world w owning cells [w*C, (w+1)*C) of the C = grid cells per world.
 * Each world behaves exactly as a context of its own would.  The store is
 * binned on the device (WRACH_FLAG_HOST_PACK is refused) and the capacity is
 * n_worlds times one world's, shared.  wrach_read, wrach_capacity and
 * wrach_canonical_device see the whole batch; wrach_add_particles,
 * wrach_store_write_back and wrach_load_packed are for one world only
 * (WRACH_E_INVALID on a batch).  wrach_create is wrach_create_batch with
 * n_worlds = 1. */
int wrach_create_batch(const wrach_config *cfg, uint32_t n_worlds, wrach_ctx **out);

/* A batch of worlds of different sizes (no counterpart in the reference):
 * world w has the dimensions of cfgs[w]; every other field (cell_size,
 * boundaries_as_dimensions, gravity, dt, flags, device) must be the same in
 * all n_worlds entries (WRACH_E_INVALID otherwise).  One launch per frame
 * kernel for all worlds, as for wrach_create_batch: every world's grid
 * starts on a multiple of 1,024 cells of the flattened key space (the cells
 * between grids stay empty), and the kernels take each cell's world from
 * its chunk of 1,024.  The capacity is the sum of the worlds'.  Entries of
 * one size give wrach_create_batch's layout. */
int wrach_create_worlds(const wrach_config *cfgs, uint32_t n_worlds, wrach_ctx **out);

/* wrach_add_particles for world `world` of a batch (0 for a single world):
 * appends to the store (tagged with the world) and re-packs the batch. */
int wrach_add_particles_to_world(wrach_ctx *ctx, uint32_t world, const float *pos_xy,
                                 const float *vel_xy, uint32_t n);

/* wrach_read of one world of a batch: its particles and its own offsets
 * [0, start_0, ..., N_world] over its C + 2 entries (C: that world's grid
 * cells). */
int wrach_read_world(wrach_ctx *ctx, uint32_t world, float *pos_xy, float *vel_xy,
                     size_t particle_cap, uint32_t *indices, size_t indices_cap,
                     uint32_t *n_particles);

/* Worlds in the context (1 unless made by wrach_create_batch or
 * wrach_create_worlds). */
int wrach_worlds(wrach_ctx *ctx, uint32_t *n_worlds);

/* wrach_settings of one world of a batch: its view and grid dimensions, the
 * cell size and its particles (synchronises). */
int wrach_settings_world(wrach_ctx *ctx, uint32_t world, wrach_world_settings *out);

/* Launch shapes of the frame (no counterpart in the reference, whose passes
 * have fixed workgroup sizes).  Each field takes one of the listed values. */
typedef struct wrach_tuning {
    uint32_t scan_threads;     /* threads per scan tile: 1024 or 512 */
    uint32_t scan_items;       /* counts per thread: (1024, 8), (512, 16), (1024, 4), (512, 8),
                                  (1024, 16), (1024, 32), (1024, 64), (512, 64), (256, 128); a
                                  context starts at (1024, 8), at (1024, 64) above 2^23
                                  cells and at (256, 128) above 2^26 */
    uint32_t scatter_per_lane; /* particles per lane of the scatter: 2, 1 or 4 */
    uint32_t cells_grid;       /* blocks of the queued-cell kernel, 1..4x the default */
    float frame_us;            /* wrach_autotune: the chosen shapes' frame time; else 0 */
} wrach_tuning;

/* The current launch shapes (frame_us = 0). */
int wrach_get_tuning(wrach_ctx *ctx, wrach_tuning *out);

/* Sets the launch shapes; WRACH_E_INVALID for a value outside the lists
 * above.  Results do not depend on the shapes. */
int wrach_set_tuning(wrach_ctx *ctx, const wrach_tuning *tuning);

/* Autotuner: runs `frames` frames (after one untimed frame) with each
 * candidate shape of one knob at a time -- scan tile, scatter width, cells
 * grid -- on the current state, keeps the fastest, and restores the state
 * after every trial, so the simulation continues as if this call had not
 * been made.  The choice reflects the current state (occupancy changes as a
 * world evolves).  *out (may be NULL) receives it.  Synchronises. */
int wrach_autotune(wrach_ctx *ctx, uint32_t frames, wrach_tuning *out);

/* Moves the view (the simulated window) to anchor (x, y), its size kept:
 * the active window over the full store (runners/bevy/src/
 * particle_store.rs:19-26, SpatialBin::get_active_cells, spatial_bin.rs:
 * 68-95; the reference keeps the anchor at (0, 0), builder.rs:61).  The
 * simulated particles are written back to the store (wrach_store_write_back),
 * which keeps the particles outside the view in their store order, and the
 * store's particles inside the new view are packed as wrach_add_particles
 * packs them; walls and cell keys then follow the new view.  The anchor
 * must be a multiple of the cell size, non-negative, with x + width and
 * y + height at most 2^24 (WRACH_E_INVALID otherwise); not for a batch or a
 * host-packed store.  WRACH_E_CAPACITY when the new view holds more
 * particles than the capacity: the view then stays where it was. */
int wrach_set_view_anchor(wrach_ctx *ctx, float x, float y);

/* Solvers.  WRACH_SOLVER_R (the default) is the reference's step:
 * collisions within one cell (shaders/physics/src/cell.rs:52-75).
 * WRACH_SOLVER_S is an extension with no counterpart in the reference: each
 * frame first runs up to max_passes separation passes over the 3x3 cells
 * around every particle, then the reference step.  A pass is Jacobi: every
 * particle i moves by the mean, over the other particles j in the 3x3 cells
 * around its frame-start cell with d <= 1, of the reference's pair
 * displacement -(p_j - p_i) * (0.5 * (1 - d)) / d (particles.rs:65-92;
 * d == 0 counts as 0.0001).  The terms are rounded to multiples of 2^-30 and
 * summed exactly, so the result does not depend on the order of particles
 * in a cell.  A pass runs only while the largest overlap 1 - d at its start
 * exceeds overlap_tol.  A particle whose 3x3 cells hold more than
 * WRACH_STENCIL_BUDGET particles is not displaced by the passes. */
#define WRACH_SOLVER_R 0u
#define WRACH_SOLVER_S 1u
/* WRACH_SOLVER_L: the learned warm start before the passes of WRACH_SOLVER_S
 * (see "The learned warm start" below). */
#define WRACH_SOLVER_L 2u
/* OR-ed into WRACH_SOLVER_S: no measure after the passes (they are the
 * same), so the report's "after" figures are not taken (after_measured 0). */
#define WRACH_SOLVER_QUIET 0x100u
#define WRACH_MAX_PASSES 64u
#define WRACH_STENCIL_BUDGET 4096u

/* Selects the solver for the following frames.  WRACH_E_INVALID for an
 * unknown mode, max_passes > WRACH_MAX_PASSES, or a negative, infinite or
 * NaN tolerance.  max_passes = 0 makes WRACH_SOLVER_S the reference's step.
 * In a batch every world's particles meet only their own world's cells; the
 * tolerance and the statistics cover the whole batch. */
int wrach_set_solver(wrach_ctx *ctx, uint32_t mode, uint32_t max_passes, float overlap_tol);

/* The last WRACH_SOLVER_S frame's overlaps (1 - d over the pairs its passes
 * visit) before and after its passes.  The sums are exact: each overlap is
 * rounded to a multiple of 2^-24 and they are added as integers, every pair
 * counted from both of its particles, so the mean overlap is
 * overlap_sum / 2^24 / pairs.  Zeros before the first such frame. */
typedef struct wrach_solver_report {
    float max_overlap_before;     /* largest 1 - d */
    float max_overlap_after;
    uint32_t passes;              /* passes that ran */
    uint32_t after_measured;      /* 1: the "after" figures were taken (not WRACH_SOLVER_QUIET) */
    uint64_t overlap_sum_before;  /* sum of 1 - d, in units of 2^-24 */
    uint64_t pairs_before;        /* overlapping pairs, each counted twice */
    uint64_t overlap_sum_after;
    uint64_t pairs_after;
} wrach_solver_report;

/* Fills *out for the last WRACH_SOLVER_S frame.  Synchronises the stream. */
int wrach_solver_stats(wrach_ctx *ctx, wrach_solver_report *out);

/* 1 (default): frames are replayed from a captured hipGraph; 0: plain
 * stream launches. */
int wrach_set_graph(wrach_ctx *ctx, int enable);

/* Runs nframes frames with plain launches and a hipEvent pair around each
 * kernel, synchronises, and adds each kernel's summed elapsed time in
 * milliseconds to kernel_ms[0..WRACH_NUM_FRAME_KERNELS). */
int wrach_step_profiled(wrach_ctx *ctx, uint32_t nframes, double *kernel_ms);

/* Captures the frame graphs that wrach_step(ctx, nframes) launches, without
 * running a frame, so that the first such call starts at once (no
 * counterpart in the reference).  A no-op with graphs off or no particles.
 * The graphs are captured again after wrach_set_solver, wrach_set_tuning,
 * wrach_autotune, a moved view or a new upload size. */
int wrach_prepare(wrach_ctx *ctx, uint32_t nframes);

/* Testing hook for the scan's bounded look-back (no counterpart in the
 * reference).  A scan tile gives up its look-back after wait_limit rounds
 * (default 2^20, about a second); the next synchronising call then returns
 * WRACH_E_DEVICE and the context must be destroyed.  epoch_skew != 0 makes
 * every look-back wait for an epoch no tile writes, so that a test can force
 * that path from the host; all offsets stay clamped to the particles, so the
 * frames that follow stay in bounds.  Production code never calls this. */
int wrach_test_scan_wait(wrach_ctx *ctx, uint32_t wait_limit, uint32_t epoch_skew);

/* Self-check of the separation passes' pair arithmetic on one device (no
 * counterpart in the reference).  Modes S and L take the root of a squared
 * distance x in [2^-96, 1 + 2^-23] and the force (1 - d) * 2^29 / d for d
 * in [2^-48, 1] by sequences shorter than hipcc's sqrtf and division
 * (one Newton step from the hardware reciprocal root, one residual
 * correction from the hardware reciprocal), which are correctly rounded by
 * a property of those hardware approximations, not by a theorem.  This
 * runs both over every float of their
 * ranges against the correctly rounded sequences: out[0] and out[1] are
 * the mismatches of the root and the force (0 and 0 where modes S and L
 * are exact), out[2] and out[3] the inputs tried (805,306,370 and
 * 402,653,185).  One kernel over 1.2 * 10^9 inputs; the call waits for it. */
int wrach_check_pair_arithmetic(int device, uint64_t out[4]);

/* The learned warm start (an extension with no counterpart in the
 * reference, whose separation is one Gauss-Seidel sweep per cell,
 * shaders/physics/src/particles.rs:62-93).
 *
 * Features.  For each particle, the 16 nearest other particles of the 5x5
 * cells around its cell, ordered by squared distance and then by the bits
 * of their (x, y, vx, vy), so that the choice does not depend on the order
 * of particles inside a cell.  Row entries 4r..4r+3 hold the r-th nearest's
 * (x_j - x_i, y_j - y_i, vx_j - vx_i, vy_j - vy_i) rounded to bf16 (nearest
 * even); missing neighbours, and particles whose 5x5 cells hold more than
 * WRACH_LEARN_BUDGET particles, give zeros.
 *
 * Network.  64 -> 128 -> 128 -> 2, SiLU after the two hidden layers:
 * z1 = x W1 + b1, a1 = SiLU(z1), z2 = bf16(a1) W2 + b2, a2 = SiLU(z2),
 * y = a2 W3 + b3, with W1, W2, W3 rounded to bf16 and f32 accumulation (the
 * hidden layers on bf16 MFMA).  Parameters: one f32 vector of
 * WRACH_LEARN_PARAMS entries, W1[64][128] (input-major), b1[128],
 * W2[128][128], b2[128], W3[128][2], b3[2].  The output predicts the
 * displacement the plain separation passes would give a particle.
 *
 * WRACH_SOLVER_L with k passes (wrach_set_solver, tolerance 0): features
 * of the frame start, then position + y (each component clamped to +-1),
 * then k passes of WRACH_SOLVER_S from there, then the reference step.  The
 * plain passes are the fallback: when the warm start leaves more mean
 * overlap than the frame start had, the second pass restarts from the
 * frame start (with k >= 2).  wrach_solver_stats reports the passes (its
 * "before" maximum is the warm start's); wrach_learned_stats the means. */
#define WRACH_LEARN_FEATURES 64u
#define WRACH_LEARN_PARAMS 25090u
#define WRACH_LEARN_BUDGET 4096u

/* Sets the network's parameters (WRACH_LEARN_PARAMS floats, the layout
 * above) and resets Adam's moments and step count.  WRACH_E_NONFINITE for
 * a NaN or infinity. */
int wrach_learn_set_params(wrach_ctx *ctx, const float *params, size_t n);

/* Copies the current parameters out (n >= WRACH_LEARN_PARAMS). */
int wrach_learn_get_params(wrach_ctx *ctx, float *params, size_t n);

/* Feature rows (64 bf16 bit patterns each) of the current state in
 * canonical packed order, as wrach_read orders the particles; rows may be
 * NULL.  overlap (may be NULL) receives the overlap sum (1 - d in 2^-24
 * units) and the number of overlapping pairs at these positions, as
 * wrach_solver_stats counts them.  Synchronises. */
int wrach_learn_features(wrach_ctx *ctx, uint16_t *rows, size_t rows_cap, uint64_t overlap[2]);

/* The network on n given feature rows: y[2i], y[2i+1] for row i. */
int wrach_learn_forward(wrach_ctx *ctx, const uint16_t *rows, uint32_t n, float *y);

/* The gradient of the loss sum_i |y_i - t_i|^2 / (2n) over n rows (n a
 * positive multiple of 64) with respect to the parameters, and the loss.
 * The parameters do not change. */
int wrach_learn_gradient(wrach_ctx *ctx, const uint16_t *rows, const float *targets, uint32_t n,
                         float *grad, float *loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with a given gradient:
 * m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2; w -= lr (m c1) /
 * (sqrt(v c2) + eps), c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t), each a
 * rounded f32 operation. */
int wrach_learn_adam(wrach_ctx *ctx, const float *grad, float lr);

/* Appends the current state's particles to the training set: their
 * feature rows and, as targets, the displacement that target_passes
 * passes of WRACH_SOLVER_S give them (the state does not change; the
 * solver's report does).  *rows_total (may be NULL): the set's size. */
int wrach_learn_sample(wrach_ctx *ctx, uint32_t target_passes, uint64_t *rows_total);

/* Empties the training set. */
int wrach_learn_clear_samples(wrach_ctx *ctx);

/* floor(n_order / batch) Adam steps on the device, step s on training rows
 * order[s*batch .. (s+1)*batch) (batch a positive multiple of 64); losses
 * (may be NULL) receives each step's loss before its update. */
int wrach_learn_train(wrach_ctx *ctx, const uint32_t *order, size_t n_order, uint32_t batch,
                      float lr, float *losses);

/* The last learned frame: overlap sums (2^-24 units) and pairs at the frame
 * start and after the warm start, and whether it fell back. */
typedef struct wrach_learned_report {
    uint64_t start_overlap_sum;
    uint64_t start_pairs;
    uint64_t warm_overlap_sum;
    uint64_t warm_pairs;
    uint32_t fell_back;
    uint32_t reserved;
} wrach_learned_report;

int wrach_learned_stats(wrach_ctx *ctx, wrach_learned_report *out);

/* Runs the learned frame's two kernels -- features of the current state,
 * then the network's warm start into a scratch buffer -- reps times with
 * events around each, synchronises, and adds their summed elapsed
 * milliseconds to ms[0] (features) and ms[1] (network).  The state does
 * not change. */
int wrach_learn_profile(wrach_ctx *ctx, uint32_t reps, double *ms);

/* One frame with no allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It is wrach_step
 * of one frame followed by the canonical order of wrach_read, on a packed
 * triple the caller holds in device memory and that is replaced in place:
 * indices (grid cells + 2 offsets, the layout of wrach_read), pos_xy and
 * vel_xy (2n floats each).  ws gives the view, grid, cell size and
 * particles_in_frame_count (= n); gravity and dt as in wrach_config.  The
 * frames' other buffers live in one device array of
 * WRACH_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the
 * process ends; a frame that needs more (about 8 bytes per cell plus 96 per
 * particle) returns WRACH_E_CAPACITY before any work.  Asynchronous on
 * `stream` (a hipStream_t; NULL: the null stream); calls no hipMalloc,
 * hipFree or allocating routine.  Calls that overlap on one GPU share that
 * array and race: serialise them (one stream, or wait between). */
#define WRACH_NOALLOC_SCRATCH_BYTES (32u << 20)
int wrach_frame_noalloc(const wrach_world_settings *ws, const float gravity[2], float dt,
                        uint32_t *indices, float *pos_xy, float *vel_xy, uint32_t n, void *stream);

/* Message of the last failed call on this thread ("" if none). */
const char *wrach_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* WRACH_H */
