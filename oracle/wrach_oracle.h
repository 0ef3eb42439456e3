/*
 * wrach_oracle.h -- CPU restatement of wrach's per-frame particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library (wrach_amd/csrc).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function restates a specific piece of the reference (tombh/wrach);
 * the citation is on each declaration.  Arithmetic is scalar IEEE-754 f32
 * with correctly rounded division and square root and no FMA contraction
 * (the file is compiled with -ffp-contract=off), which is what the
 * reference's shader crate computes when built natively (Rust never
 * contracts a*b+c).
 *
 * Parity pins (what checks each function independently of its own code):
 *   wo_cell_coord        golden vectors of runners/bevy/src/spatial_bin.rs:156-183
 *   wo_active_grid       golden vectors of runners/bevy/src/spatial_bin.rs:185-248
 *   wo_max_particles     runners/bevy/src/particle_store.rs:231-234 (74),
 *                        runners/api/src/lib.rs:122-125 (164)
 *   wo_pack              runners/bevy/src/particle_store.rs:143-229 and
 *                        runners/bevy/src/compute/03_prefix_sum.rs:196-199;
 *                        brute force (Python stable sort) on random inputs
 *   wo_collide_cell      shaders/physics/src/particles.rs:128-146 (golden),
 *                        closed forms for a separated pair, coincident pair
 *   wo_finish_particle   closed forms: integrate, wall clamp+reflect order,
 *                        velocity clamp, exactly-on-wall, NaN-free
 *   wo_physics           reference GPU tests after 4 ticks
 *                        (03_prefix_sum.rs:150-260, 04_pack_particle_data.rs:
 *                        72-140) plus a pure-Python brute-force step on tiny
 *                        worlds
 *   wo_rebin             brute force (Python stable sort by cell), invariants
 *                        (permutation, sortedness, counts)
 *   wo_step_mt           equality with wo_step (serial) on every tested case
 *   wo_stencil_pass      brute force (numpy all-pairs Jacobi, float64 sums) on
 *                        random worlds; closed form for one overlapping pair
 *                        (separation 1 after a pass); invariance under a
 *                        permutation of the particles inside their cells;
 *                        overlap equal to the all-pairs maximum
 *   wo_step_s            zero passes equal wo_step; a tolerance above the
 *                        first overlap runs no pass
 *   wo_learn_features    brute force (numpy: every particle's distances to all
 *                        others, Chebyshev cell distance <= 2, a lexsort); the
 *                        sums equal wo_stencil_pass's; order inside a cell
 *                        does not matter (permuted cells give permuted rows)
 */
#ifndef WRACH_ORACLE_H
#define WRACH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-cell collision cap: SPATIAL_BIN_CELL_SIZE^2 * CELL_LEEWAY = 3*3*1.0
 * (shaders/physics/src/cell.rs:21,29-30).  It does not follow the runtime
 * cell size. */
#define WO_MAX_PARTICLES_IN_CELL 9

/* Step parameters.  The first eight fields are the reference's 32-byte
 * WorldSettings uniform (shaders/shared/src/lib.rs:19-31,
 * runners/bevy/src/config_shader.rs:15-29, assets/shaders/types.wgsl:3-15).
 * gravity and dt are extensions: gravity == (0,0) and dt == 1 reproduce the
 * reference bit for bit. */
typedef struct wo_params {
    float view_dimensions[2];
    float view_anchor[2];
    uint32_t grid_dimensions[2];
    uint32_t cell_size;
    uint32_t particles_in_frame_count;
    float gravity[2];
    float dt;
} wo_params;

/* SpatialBin::get_cell_coord (runners/bevy/src/spatial_bin.rs:48-64):
 * f32::div_euclid by the cell size, then a saturating `as i32`. */
void wo_cell_coord(float x, float y, float cell_size, int32_t out[2]);

/* SpatialBin::get_active_cells (spatial_bin.rs:68-95): bottom-left cell
 * coordinate and grid dimensions of viewport (x0, y0, x1, y1). */
void wo_active_grid(const float viewport[4], float cell_size,
                    int32_t bottom_left[2], uint32_t grid[2]);

/* ParticleStore::max_particles_per_frame (particle_store.rs:116-133):
 * n_cells*cs^2 plus 10 * ceil(n_cells*cs^2 / 100).  Returns 0 on u32
 * overflow. */
uint32_t wo_max_particles(uint64_t n_cells, uint32_t cell_size);

/* ParticleStore::add_particle + SpatialBin::create_packed_data
 * (particle_store.rs:54-59, spatial_bin.rs:103-149): particles in insertion
 * order; cells of the active grid in row-major order; insertion order inside
 * a cell; particles outside the grid are not packed.  indices gets
 * [0, 0, c0, c0+c1, ..., N] (grid cells + 2 entries).  Returns the number of
 * packed particles. */
uint32_t wo_pack(const float *pos_xy, const float *vel_xy, uint32_t n,
                 const float viewport[4], float cell_size,
                 uint32_t *indices, float *pos_out, float *vel_out);

/* Particles::new + pairs + push_close_particles_apart
 * (shaders/physics/src/particles.rs:31-93): in-place Gauss-Seidel separation
 * of m <= 9 particles, pairs (i<j) in index order.  pos_xy is updated in
 * place (2*m floats). */
void wo_collide_cell(float *pos_xy, uint32_t m);

/* Particle::integrate + enforce_limits (shaders/physics/src/particle.rs:
 * 40-82), with the gravity/dt extension placed just before integrate. */
void wo_finish_particle(const wo_params *p, float pos[2], float vel[2]);

/* The physics pass over every cell (shaders/physics/src/lib.rs:41-64,
 * cell.rs:52-112): collide the first min(n,9) particles of each cell,
 * integrate/limit all of them, write at the same packed index.  Reads the
 * full `indices` before anything is cleared (the reference's in-kernel
 * clear races, cell.rs:116-131). */
void wo_physics(const wo_params *p, const uint32_t *indices,
                const float *pos_in, const float *vel_in,
                float *pos_out, float *vel_out);

/* Cell key of a position written by the physics pass
 * (assets/shaders/particles_per_cell.wgsl:14-27): floor((x-ax)/cs) etc.
 * Returns -1 if the position falls outside the grid. */
int64_t wo_cell_key(const wo_params *p, float x, float y);

/* Re-bin (count + exclusive scan + scatter: particles_per_cell.wgsl,
 * prefix_sum.wgsl, pack_new_particle_data.wgsl) as a STABLE counting sort by
 * cell key: inside a cell, particles keep their pos_out index order.
 * indices gets [0, start_0, ..., start_{C-1}, N].  Returns 0, or -1 if a key
 * falls outside the grid. */
int wo_rebin(const wo_params *p, const float *pos_out, const float *vel_out,
             uint32_t *indices, float *pos_in, float *vel_in);

/* One frame: wo_physics then wo_rebin.  Scratch: 4*N floats.  Returns 0 or
 * -1 (key outside the grid). */
int wo_step(const wo_params *p, uint32_t *indices, float *pos, float *vel,
            float *scratch);

/* The same frame on `nthreads` host threads: cells split into contiguous
 * ranges for the physics pass, then a parallel stable counting sort
 * (per-thread windowed histograms, scan, per-thread stable scatter).
 * Output identical to wo_step.  Scratch: 4*N floats. */
int wo_step_mt(const wo_params *p, uint32_t *indices, float *pos, float *vel,
               float *scratch, int nthreads);

/* Mode S (an extension; the reference collides within one cell only,
 * shaders/physics/src/cell.rs:60-69): before the reference step, separation
 * passes over the 3x3 cells around each particle.
 *
 * A pass is Jacobi: every particle i takes, from every other particle j in
 * the 3x3 cells around its frame-start cell (wo_cell_key of its position
 * when the frame began), the pair rule of wo_collide_cell
 * (particles.rs:65-92): skip d > 1, d == 0 -> 0.0001, force
 * (0.5 * (1 - d)) / d, p_i -= (p_j - p_i) * force.  The terms are rounded to
 * multiples of 1/WO_STENCIL_SCALE (llrintf, nearest-even) and summed in
 * int64, exactly, so the displacement does not depend on the order of the
 * particles in a cell.  The displacement is their mean (constraint averaging
 * of position-based dynamics, which bounds a particle's move to 0.5 per
 * pass): with m > 1 overlapping neighbours, D = (float)sum / WO_STENCIL_SCALE
 * then D / (float)m, each a rounded f32 operation; p_i' = p_i + D.  A
 * particle whose 3x3 cells hold more than WO_STENCIL_BUDGET particles
 * (itself included) takes no displacement and measures no overlap: a bound
 * on the work per particle.  With cell sides >= 1 every pair closer than 1
 * lies in the 3x3 cells, so one pass equals the all-pairs Jacobi pass. */
#define WO_STENCIL_SCALE 1073741824.0f  /* 2^30: a term (|term| <= 0.5) fits 32 bits */
#define WO_STENCIL_BUDGET 4096
/* Overlaps are summed as multiples of 1/WO_OVERLAP_SCALE (llrintf). */
#define WO_OVERLAP_SCALE 16777216.0f

/* One pass over the packed state: positions pin, frame-start positions
 * pcell (the cells), indices idx of the frame.  move: write pout.  Returns
 * the largest overlap 1 - d over the pairs it visits (0 if none); sums (if
 * not NULL): the sum of those overlaps in 1/WO_OVERLAP_SCALE units and the
 * number of pairs, each pair counted from both of its particles. */
float wo_stencil_pass(const wo_params *p, const uint32_t *indices, const float *pcell,
                      const float *pin, float *pout, int move, uint64_t sums[2]);

/* A mode S frame: up to max_passes passes, each run only while the overlap
 * at its start exceeds overlap_tol, then wo_step.  overlap[0]: before the
 * passes, overlap[1]: after them; *passes: the passes that ran; sums: the
 * overlap sums and pair counts of wo_stencil_pass before and after. */
int wo_step_s(const wo_params *p, uint32_t *indices, float *pos, float *vel, float *scratch,
              uint32_t max_passes, float overlap_tol, float overlap[2], uint32_t *passes,
              uint64_t sums[4]);

/* The learned warm start's neighbour features (an extension): for each
 * particle i of the packed state, the other particles of the 5x5 cells
 * around its cell (wo_cell_key of its position), the 16 nearest by squared
 * distance, then by the bits of (x, y, vx, vy) as unsigned integers; row i
 * of feat holds, for the r-th, (x_j - x_i, y_j - y_i, vx_j - vx_i,
 * vy_j - vy_i) rounded to bf16 (nearest even) at 4r..4r+3, zeros past the
 * last.  More than WO_LEARN_BUDGET particles in the 5x5 cells: a zero row.
 * sums (if not NULL): the overlap sum and pair count of wo_stencil_pass at
 * these positions (its 3x3 cells and budget). */
#define WO_LEARN_NEIGHBOURS 16
#define WO_LEARN_FEATURES 64
#define WO_LEARN_BUDGET 4096
void wo_learn_features(const wo_params *p, const uint32_t *indices, const float *pos,
                       const float *vel, uint16_t *feat, uint64_t sums[2]);

#ifdef __cplusplus
}
#endif

#endif /* WRACH_ORACLE_H */
