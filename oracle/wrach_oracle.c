/*
 * wrach_oracle.c -- CPU restatement of wrach's per-frame particle step.
 *
 * TEST INFRASTRUCTURE ONLY (see wrach_oracle.h).  Compiled with
 * -ffp-contract=off: every a*b+c below is a rounded multiply then a rounded
 * add, as in the reference's Rust.
 */
#define _GNU_SOURCE
#include "wrach_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Rust's `f as i32`: saturating, NaN -> 0. */
static int32_t sat_i32(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

/* f32::div_euclid: q = trunc(a / b); if a % b < 0 then q - 1 (b > 0). */
static float div_euclid(float a, float b)
{
    float q = truncf(a / b);
    if (fmodf(a, b) < 0.0f) return b > 0.0f ? q - 1.0f : q + 1.0f;
    return q;
}

/* spatial_bin.rs:48-64 */
void wo_cell_coord(float x, float y, float cell_size, int32_t out[2])
{
    out[0] = sat_i32(div_euclid(x, cell_size));
    out[1] = sat_i32(div_euclid(y, cell_size));
}

/* spatial_bin.rs:68-95: rows bottom_left.y..=top_right.y, columns
 * bottom_left.x..=top_right.x; an inverted viewport yields an empty grid. */
void wo_active_grid(const float viewport[4], float cell_size,
                    int32_t bottom_left[2], uint32_t grid[2])
{
    int32_t tr[2];
    wo_cell_coord(viewport[0], viewport[1], cell_size, bottom_left);
    wo_cell_coord(viewport[2], viewport[3], cell_size, tr);
    int64_t gx = (int64_t)tr[0] - bottom_left[0] + 1;
    int64_t gy = (int64_t)tr[1] - bottom_left[1] + 1;
    if (gx <= 0 || gy <= 0) { grid[0] = 0; grid[1] = 0; return; }
    grid[0] = (uint32_t)gx;
    grid[1] = (uint32_t)gy;
}

/* particle_store.rs:116-133 */
uint32_t wo_max_particles(uint64_t n_cells, uint32_t cell_size)
{
    uint64_t normal = n_cells * (uint64_t)cell_size * cell_size;
    uint64_t one_percent = (normal + 99) / 100;
    uint64_t total = normal + 10 * one_percent;
    if (normal > UINT32_MAX || total > UINT32_MAX) return 0;
    return (uint32_t)total;
}

/* particle_store.rs:54-59 + spatial_bin.rs:103-149 as a stable counting sort
 * over the active grid (the hashmap keeps insertion order per cell and the
 * packer walks cells row-major). */
uint32_t wo_pack(const float *pos_xy, const float *vel_xy, uint32_t n,
                 const float viewport[4], float cell_size,
                 uint32_t *indices, float *pos_out, float *vel_out)
{
    int32_t bl[2];
    uint32_t grid[2];
    wo_active_grid(viewport, cell_size, bl, grid);
    uint64_t ncell = (uint64_t)grid[0] * grid[1];
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    for (uint64_t c = 0; c < ncell + 2; ++c) indices[c] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t cc[2];
        wo_cell_coord(pos_xy[2 * i], pos_xy[2 * i + 1], cell_size, cc);
        int64_t cx = (int64_t)cc[0] - bl[0], cy = (int64_t)cc[1] - bl[1];
        if (cx < 0 || cy < 0 || cx >= grid[0] || cy >= grid[1]) {
            cell[i] = -1;
            continue;
        }
        cell[i] = cy * (int64_t)grid[0] + cx;
        indices[cell[i] + 2] += 1;
    }
    for (uint64_t c = 2; c < ncell + 2; ++c) indices[c] += indices[c - 1];
    uint32_t packed = ncell ? indices[ncell + 1] : 0;
    /* cursor for cell c is indices[c + 1] (its start); walk in insertion
     * order, then restore the starts by shifting back. */
    for (uint32_t i = 0; i < n; ++i) {
        if (cell[i] < 0) continue;
        uint32_t d = indices[cell[i] + 1]++;
        pos_out[2 * d] = pos_xy[2 * i];
        pos_out[2 * d + 1] = pos_xy[2 * i + 1];
        vel_out[2 * d] = vel_xy[2 * i];
        vel_out[2 * d + 1] = vel_xy[2 * i + 1];
    }
    /* after the walk indices[c+1] holds end_c == start_{c+1}: rebuild. */
    for (uint64_t c = ncell + 1; c >= 2; --c) indices[c] = indices[c - 1];
    indices[1] = 0;
    indices[0] = 0;
    free(cell);
    return packed;
}

/* particles.rs:62-93.  glam's Vec2::distance is sqrt(dx*dx + dy*dy) of the
 * difference; force = (0.5 * (1.0 - d)) / d; dv = (p_j - p_i) * force;
 * p_i -= dv; p_j += dv. */
void wo_collide_cell(float *pos, uint32_t m)
{
    for (uint32_t i = 0; i < m; ++i) {
        for (uint32_t j = i + 1; j < m; ++j) {
            float dx = pos[2 * i] - pos[2 * j];
            float dy = pos[2 * i + 1] - pos[2 * j + 1];
            float sq = dx * dx;
            sq = sq + dy * dy;
            float d = sqrtf(sq);
            if (d > 1.0f) continue;            /* MIN_DISTANCE particles.rs:18 */
            if (d == 0.0f) d = 0.0001f;         /* particles.rs:74-76 */
            float force = (0.5f * (1.0f - d)) / d;
            float vx = pos[2 * j] - pos[2 * i];
            float vy = pos[2 * j + 1] - pos[2 * i + 1];
            vx = vx * force;
            vy = vy * force;
            pos[2 * i] = pos[2 * i] - vx;
            pos[2 * i + 1] = pos[2 * i + 1] - vy;
            pos[2 * j] = pos[2 * j] + vx;
            pos[2 * j + 1] = pos[2 * j + 1] + vy;
        }
    }
}

/* Rust f32::clamp(min, max): `if x < min {min}`, then `if x > max {max}`. */
static float clamp_rs(float x, float lo, float hi)
{
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* particle.rs:40-82 (+ extension): [v += g*dt]; p += v*dt; walls in the
 * order x>max, x<min, y>max, y<min with v *= -1; v clamped to [-1, 1]. */
void wo_finish_particle(const wo_params *p, float pos[2], float vel[2])
{
    if (p->gravity[0] != 0.0f) vel[0] = vel[0] + p->gravity[0] * p->dt;
    if (p->gravity[1] != 0.0f) vel[1] = vel[1] + p->gravity[1] * p->dt;
    if (p->dt == 1.0f) {
        pos[0] = pos[0] + vel[0];
        pos[1] = pos[1] + vel[1];
    } else {
        pos[0] = pos[0] + vel[0] * p->dt;
        pos[1] = pos[1] + vel[1] * p->dt;
    }
    float x0 = p->view_anchor[0], y0 = p->view_anchor[1];
    float x1 = p->view_anchor[0] + p->view_dimensions[0];
    float y1 = p->view_anchor[1] + p->view_dimensions[1];
    if (pos[0] > x1) { pos[0] = x1; vel[0] = vel[0] * -1.0f; }
    if (pos[0] < x0) { pos[0] = x0; vel[0] = vel[0] * -1.0f; }
    if (pos[1] > y1) { pos[1] = y1; vel[1] = vel[1] * -1.0f; }
    if (pos[1] < y0) { pos[1] = y0; vel[1] = vel[1] * -1.0f; }
    vel[0] = clamp_rs(vel[0], -1.0f, 1.0f);
    vel[1] = clamp_rs(vel[1], -1.0f, 1.0f);
}

/* cell.rs:52-112 for cells [c0, c1).  Cell c (0-based) is the reference's
 * thread k = c + 1: start = indices[c+1], count = indices[c+2] - start. */
static void physics_cells(const wo_params *p, const uint32_t *idx,
                          const float *pin, const float *vin,
                          float *pout, float *vout, uint64_t c0, uint64_t c1)
{
    float lp[2 * WO_MAX_PARTICLES_IN_CELL];
    for (uint64_t c = c0; c < c1; ++c) {
        uint32_t s = idx[c + 1];
        uint32_t n = idx[c + 2] - s;
        uint32_t m = n > WO_MAX_PARTICLES_IN_CELL ? WO_MAX_PARTICLES_IN_CELL : n;
        for (uint32_t i = 0; i < m; ++i) {
            lp[2 * i] = pin[2 * (uint64_t)(s + i)];
            lp[2 * i + 1] = pin[2 * (uint64_t)(s + i) + 1];
        }
        wo_collide_cell(lp, m);
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t g = (uint64_t)s + i;
            float q[2], v[2];
            if (i < m) { q[0] = lp[2 * i]; q[1] = lp[2 * i + 1]; }
            else { q[0] = pin[2 * g]; q[1] = pin[2 * g + 1]; }  /* cell.rs:79-95 */
            v[0] = vin[2 * g];
            v[1] = vin[2 * g + 1];
            wo_finish_particle(p, q, v);
            pout[2 * g] = q[0];
            pout[2 * g + 1] = q[1];
            vout[2 * g] = v[0];
            vout[2 * g + 1] = v[1];
        }
    }
}

void wo_physics(const wo_params *p, const uint32_t *idx,
                const float *pin, const float *vin, float *pout, float *vout)
{
    uint64_t ncell = (uint64_t)p->grid_dimensions[0] * p->grid_dimensions[1];
    physics_cells(p, idx, pin, vin, pout, vout, 0, ncell);
}

/* particles_per_cell.wgsl:14-27 / pack_new_particle_data.wgsl:19-37. */
int64_t wo_cell_key(const wo_params *p, float x, float y)
{
    float cs = (float)p->cell_size;
    float fx = floorf((x - p->view_anchor[0]) / cs);
    float fy = floorf((y - p->view_anchor[1]) / cs);
    if (!(fx >= 0.0f) || !(fy >= 0.0f)) return -1;
    if (fx >= (float)p->grid_dimensions[0] || fy >= (float)p->grid_dimensions[1])
        return -1;
    return (int64_t)fy * p->grid_dimensions[0] + (int64_t)fx;
}

int wo_rebin(const wo_params *p, const float *pout, const float *vout,
             uint32_t *idx, float *pin, float *vin)
{
    uint64_t ncell = (uint64_t)p->grid_dimensions[0] * p->grid_dimensions[1];
    uint32_t n = p->particles_in_frame_count;
    uint32_t *key = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    memset(idx, 0, sizeof(uint32_t) * (ncell + 2));
    for (uint32_t i = 0; i < n; ++i) {
        int64_t k = wo_cell_key(p, pout[2 * (uint64_t)i], pout[2 * (uint64_t)i + 1]);
        if (k < 0) { free(key); return -1; }
        key[i] = (uint32_t)k;
        idx[k + 2] += 1;
    }
    for (uint64_t c = 2; c < ncell + 2; ++c) idx[c] += idx[c - 1];
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t d = idx[key[i] + 1]++;
        pin[2 * (uint64_t)d] = pout[2 * (uint64_t)i];
        pin[2 * (uint64_t)d + 1] = pout[2 * (uint64_t)i + 1];
        vin[2 * (uint64_t)d] = vout[2 * (uint64_t)i];
        vin[2 * (uint64_t)d + 1] = vout[2 * (uint64_t)i + 1];
    }
    for (uint64_t c = ncell + 1; c >= 2; --c) idx[c] = idx[c - 1];
    idx[1] = 0;
    idx[0] = 0;
    free(key);
    return 0;
}

int wo_step(const wo_params *p, uint32_t *idx, float *pos, float *vel,
            float *scratch)
{
    uint64_t n = p->particles_in_frame_count;
    float *po = scratch, *vo = scratch + 2 * n;
    wo_physics(p, idx, pos, vel, po, vo);
    return wo_rebin(p, po, vo, idx, pos, vel);
}

/* ---------------------------------------------------------------------- */
/* Multi-threaded frame (CPU baseline).                                    */

typedef struct mt_shared {
    const wo_params *p;
    uint32_t *idx;
    float *pos, *vel, *po, *vo;
    uint32_t *key;
    uint64_t ncell;
    uint32_t n;
    int nthreads;
    uint32_t **hist;      /* per thread: counts over [kmin, kmax] */
    int64_t *kmin, *kmax;
    uint64_t *range_sum;  /* per thread: total over its cell range */
    int bad;
    pthread_barrier_t bar;
} mt_shared;

typedef struct mt_arg { mt_shared *sh; int t; } mt_arg;

static void *mt_worker(void *vp)
{
    mt_arg *a = (mt_arg *)vp;
    mt_shared *sh = a->sh;
    int t = a->t, T = sh->nthreads;
    uint64_t c0 = sh->ncell * t / T, c1 = sh->ncell * (t + 1) / T;
    uint64_t i0 = (uint64_t)sh->n * t / T, i1 = (uint64_t)sh->n * (t + 1) / T;

    /* 1. physics over a contiguous cell range: writes [idx[c0+1], idx[c1+1]). */
    physics_cells(sh->p, sh->idx, sh->pos, sh->vel, sh->po, sh->vo, c0, c1);
    pthread_barrier_wait(&sh->bar);

    /* 2. keys of a contiguous particle range and a windowed histogram. */
    int64_t lo = INT64_MAX, hi = -1;
    for (uint64_t i = i0; i < i1; ++i) {
        int64_t k = wo_cell_key(sh->p, sh->po[2 * i], sh->po[2 * i + 1]);
        if (k < 0) { sh->bad = 1; k = 0; }
        sh->key[i] = (uint32_t)k;
        if (k < lo) lo = k;
        if (k > hi) hi = k;
    }
    if (hi < 0) { lo = 0; hi = -1; }
    sh->kmin[t] = lo;
    sh->kmax[t] = hi;
    uint32_t *h = (uint32_t *)calloc((size_t)(hi - lo + 1 > 0 ? hi - lo + 1 : 1), 4);
    for (uint64_t i = i0; i < i1; ++i) h[sh->key[i] - lo] += 1;
    sh->hist[t] = h;
    pthread_barrier_wait(&sh->bar);

    /* 3. per cell range: totals, then (after a serial prefix over ranges)
     * exclusive starts and per-thread cursors in thread order. */
    uint64_t sum = 0;
    for (uint64_t c = c0; c < c1; ++c) {
        uint32_t cnt = 0;
        for (int u = 0; u < T; ++u)
            if ((int64_t)c >= sh->kmin[u] && (int64_t)c <= sh->kmax[u])
                cnt += sh->hist[u][c - sh->kmin[u]];
        sum += cnt;
    }
    sh->range_sum[t] = sum;
    pthread_barrier_wait(&sh->bar);
    uint64_t base = 0;
    for (int u = 0; u < t; ++u) base += sh->range_sum[u];
    for (uint64_t c = c0; c < c1; ++c) {
        sh->idx[c + 1] = (uint32_t)base;
        for (int u = 0; u < T; ++u) {
            if ((int64_t)c >= sh->kmin[u] && (int64_t)c <= sh->kmax[u]) {
                uint32_t *hc = &sh->hist[u][c - sh->kmin[u]];
                uint32_t cnt = *hc;
                *hc = (uint32_t)base;   /* becomes thread u's cursor */
                base += cnt;
            }
        }
    }
    if (t == T - 1) { sh->idx[0] = 0; sh->idx[sh->ncell + 1] = (uint32_t)base; }
    pthread_barrier_wait(&sh->bar);

    /* 4. stable scatter of the thread's particle range. */
    for (uint64_t i = i0; i < i1; ++i) {
        uint32_t d = h[sh->key[i] - lo]++;
        sh->pos[2 * (uint64_t)d] = sh->po[2 * i];
        sh->pos[2 * (uint64_t)d + 1] = sh->po[2 * i + 1];
        sh->vel[2 * (uint64_t)d] = sh->vo[2 * i];
        sh->vel[2 * (uint64_t)d + 1] = sh->vo[2 * i + 1];
    }
    return NULL;
}

int wo_step_mt(const wo_params *p, uint32_t *idx, float *pos, float *vel,
               float *scratch, int nthreads)
{
    if (nthreads <= 1) return wo_step(p, idx, pos, vel, scratch);
    mt_shared sh;
    memset(&sh, 0, sizeof sh);
    sh.p = p;
    sh.idx = idx;
    sh.pos = pos;
    sh.vel = vel;
    sh.n = p->particles_in_frame_count;
    sh.po = scratch;
    sh.vo = scratch + 2 * (uint64_t)sh.n;
    sh.ncell = (uint64_t)p->grid_dimensions[0] * p->grid_dimensions[1];
    sh.nthreads = nthreads;
    sh.key = (uint32_t *)malloc(sizeof(uint32_t) * (sh.n ? sh.n : 1));
    sh.hist = (uint32_t **)calloc(nthreads, sizeof(uint32_t *));
    sh.kmin = (int64_t *)calloc(nthreads, sizeof(int64_t));
    sh.kmax = (int64_t *)calloc(nthreads, sizeof(int64_t));
    sh.range_sum = (uint64_t *)calloc(nthreads, sizeof(uint64_t));
    pthread_barrier_init(&sh.bar, NULL, nthreads);
    pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
    mt_arg *args = (mt_arg *)calloc(nthreads, sizeof(mt_arg));
    for (int t = 0; t < nthreads; ++t) {
        args[t].sh = &sh;
        args[t].t = t;
        pthread_create(&th[t], NULL, mt_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&sh.bar);
    for (int t = 0; t < nthreads; ++t) free(sh.hist[t]);
    free(sh.hist); free(sh.kmin); free(sh.kmax); free(sh.range_sum);
    free(sh.key); free(th); free(args);
    return sh.bad ? -1 : 0;
}

/* ---------------------------------------------------------------------- */
/* Mode S: separation passes over the 3x3 cells around each particle.     */
/* An extension with no counterpart in the reference, whose collisions     */
/* are same-cell only (shaders/physics/src/cell.rs:60-69).                 */

/* Cell (cx, cy) of a particle from its frame-start position, as the
 * re-bin keys it (wo_cell_key); -1 outside the grid. */
static int64_t stencil_cell(const wo_params *p, const float *pcell, uint64_t i)
{
    return wo_cell_key(p, pcell[2 * i], pcell[2 * i + 1]);
}

/* Particles in the 3x3 cells around cell k (clipped at the grid). */
static uint64_t stencil_count(const wo_params *p, const uint32_t *idx, int64_t k)
{
    int64_t gx = p->grid_dimensions[0], gy = p->grid_dimensions[1];
    int64_t cx = k % gx, cy = k / gx;
    uint64_t total = 0;
    for (int64_t y = cy - 1; y <= cy + 1; ++y) {
        if (y < 0 || y >= gy) continue;
        for (int64_t x = cx - 1; x <= cx + 1; ++x) {
            if (x < 0 || x >= gx) continue;
            int64_t m = y * gx + x;
            total += idx[m + 2] - idx[m + 1];
        }
    }
    return total;
}

/* One pair of the pass: particles.rs:65-92's distance, cutoff, 0 -> 0.0001
 * and force, for the displacement of i (returned through ex, ey: the
 * reference subtracts (p_j - p_i) * force from p_i) and the overlap 1 - d. */
static int stencil_pair(float xi, float yi, float xj, float yj, float *ex, float *ey,
                        float *overlap)
{
    float dx = xi - xj;
    float dy = yi - yj;
    float sq = dx * dx;
    sq = sq + dy * dy;
    float d = sqrtf(sq);
    if (d > 1.0f) return 0;
    if (d == 0.0f) d = 0.0001f;
    *overlap = 1.0f - d;
    float force = (0.5f * (1.0f - d)) / d;
    *ex = (xj - xi) * force;
    *ey = (yj - yi) * force;
    return 1;
}

float wo_stencil_pass(const wo_params *p, const uint32_t *idx, const float *pcell,
                      const float *pin, float *pout, int move, uint64_t sums[2])
{
    uint64_t n = p->particles_in_frame_count;
    int64_t gx = p->grid_dimensions[0], gy = p->grid_dimensions[1];
    float worst = 0.0f;
    if (sums) sums[0] = sums[1] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        float xi = pin[2 * i], yi = pin[2 * i + 1];
        int64_t k = stencil_cell(p, pcell, i);
        int64_t ax = 0, ay = 0;  /* the displacement in WO_STENCIL_SCALE units */
        uint32_t pairs = 0;      /* overlapping neighbours */
        if (k >= 0 && stencil_count(p, idx, k) <= WO_STENCIL_BUDGET) {
            int64_t cx = k % gx, cy = k / gx;
            for (int64_t y = cy - 1; y <= cy + 1; ++y) {
                if (y < 0 || y >= gy) continue;
                for (int64_t x = cx - 1; x <= cx + 1; ++x) {
                    if (x < 0 || x >= gx) continue;
                    int64_t m = y * gx + x;
                    for (uint64_t j = idx[m + 1]; j < idx[m + 2]; ++j) {
                        if (j == i) continue;
                        float ex, ey, ov;
                        if (!stencil_pair(xi, yi, pin[2 * j], pin[2 * j + 1], &ex, &ey, &ov))
                            continue;
                        if (ov > worst) worst = ov;
                        if (sums) {
                            sums[0] += (uint64_t)llrintf(ov * WO_OVERLAP_SCALE);
                            sums[1] += 1;
                        }
                        pairs += 1;
                        ax -= llrintf(ex * WO_STENCIL_SCALE);
                        ay -= llrintf(ey * WO_STENCIL_SCALE);
                    }
                }
            }
        }
        if (move) {
            float dx = (float)ax * (1.0f / WO_STENCIL_SCALE);
            float dy = (float)ay * (1.0f / WO_STENCIL_SCALE);
            if (pairs > 1) {  /* the mean of the pair corrections */
                dx = dx / (float)pairs;
                dy = dy / (float)pairs;
            }
            pout[2 * i] = xi + dx;
            pout[2 * i + 1] = yi + dy;
        }
    }
    return worst;
}

int wo_step_s(const wo_params *p, uint32_t *idx, float *pos, float *vel, float *scratch,
              uint32_t max_passes, float overlap_tol, float overlap[2], uint32_t *passes,
              uint64_t sums[4])
{
    uint64_t n = p->particles_in_frame_count;
    float *a = (float *)malloc(sizeof(float) * 2 * (n ? n : 1));
    float *b = (float *)malloc(sizeof(float) * 2 * (n ? n : 1));
    memcpy(a, pos, sizeof(float) * 2 * n);
    uint32_t ran = 0;
    float ov = wo_stencil_pass(p, idx, pos, a, NULL, 0, sums);
    overlap[0] = ov;
    sums[2] = sums[0];
    sums[3] = sums[1];
    for (uint32_t k = 0; k < max_passes; ++k) {
        if (!(ov > overlap_tol)) break;     /* the pass would start within tolerance */
        wo_stencil_pass(p, idx, pos, a, b, 1, NULL);
        float *t = a; a = b; b = t;
        ov = wo_stencil_pass(p, idx, pos, a, NULL, 0, sums + 2);
        ran = k + 1;
    }
    overlap[1] = ov;
    *passes = ran;
    memcpy(pos, a, sizeof(float) * 2 * n);
    free(a);
    free(b);
    return wo_step(p, idx, pos, vel, scratch);
}

/* ---------------------------------------------------------------------- */
/* The learned warm start's neighbour features (an extension with no       */
/* counterpart in the reference).                                          */

static uint16_t bf16_rne(float f)
{
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

static uint32_t f32_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

typedef struct feat_key {
    uint32_t k[5];  /* squared distance, x, y, vx, vy (bits, compared unsigned) */
    uint64_t j;
} feat_key;

static int feat_less(const feat_key *a, const feat_key *b)
{
    for (int t = 0; t < 5; ++t)
        if (a->k[t] != b->k[t]) return a->k[t] < b->k[t];
    return 0;
}

void wo_learn_features(const wo_params *p, const uint32_t *idx, const float *pos, const float *vel,
                       uint16_t *feat, uint64_t sums[2])
{
    uint64_t n = p->particles_in_frame_count;
    int64_t gx = p->grid_dimensions[0], gy = p->grid_dimensions[1];
    if (sums) sums[0] = sums[1] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        float xi = pos[2 * i], yi = pos[2 * i + 1];
        int64_t k = wo_cell_key(p, xi, yi);
        uint16_t *row = feat + i * WO_LEARN_FEATURES;
        memset(row, 0, WO_LEARN_FEATURES * sizeof(uint16_t));
        if (k < 0) continue;
        int64_t cx = k % gx, cy = k / gx;
        uint64_t tot5 = 0, tot3 = 0;
        for (int64_t y = cy - 2; y <= cy + 2; ++y) {
            if (y < 0 || y >= gy) continue;
            for (int64_t x = cx - 2; x <= cx + 2; ++x) {
                if (x < 0 || x >= gx) continue;
                int64_t m = y * gx + x;
                uint64_t c = idx[m + 2] - idx[m + 1];
                tot5 += c;
                if (y >= cy - 1 && y <= cy + 1 && x >= cx - 1 && x <= cx + 1) tot3 += c;
            }
        }
        feat_key best[WO_LEARN_NEIGHBOURS];
        int nb = 0;
        for (int64_t y = cy - 2; y <= cy + 2; ++y) {
            if (y < 0 || y >= gy) continue;
            for (int64_t x = cx - 2; x <= cx + 2; ++x) {
                if (x < 0 || x >= gx) continue;
                int inner = y >= cy - 1 && y <= cy + 1 && x >= cx - 1 && x <= cx + 1;
                int64_t m = y * gx + x;
                for (uint64_t j = idx[m + 1]; j < idx[m + 2]; ++j) {
                    if (j == i) continue;
                    float xj = pos[2 * j], yj = pos[2 * j + 1];
                    if (sums && inner && tot3 <= WO_STENCIL_BUDGET) {
                        float ex, ey, ov;
                        if (stencil_pair(xi, yi, xj, yj, &ex, &ey, &ov)) {
                            sums[0] += (uint64_t)llrintf(ov * WO_OVERLAP_SCALE);
                            sums[1] += 1;
                        }
                    }
                    if (tot5 > WO_LEARN_BUDGET) continue;
                    float dx = xj - xi;
                    float dy = yj - yi;
                    float sq = dx * dx;
                    sq = sq + dy * dy;
                    feat_key c;
                    c.k[0] = f32_bits(sq);
                    c.k[1] = f32_bits(xj);
                    c.k[2] = f32_bits(yj);
                    c.k[3] = f32_bits(vel[2 * j]);
                    c.k[4] = f32_bits(vel[2 * j + 1]);
                    c.j = j;
                    /* insertion into the ascending list of the nb nearest */
                    if (nb == WO_LEARN_NEIGHBOURS && !feat_less(&c, &best[nb - 1])) continue;
                    int at = nb < WO_LEARN_NEIGHBOURS ? nb++ : WO_LEARN_NEIGHBOURS - 1;
                    while (at > 0 && feat_less(&c, &best[at - 1])) {
                        best[at] = best[at - 1];
                        --at;
                    }
                    best[at] = c;
                }
            }
        }
        for (int r = 0; r < nb; ++r) {
            uint64_t j = best[r].j;
            row[4 * r + 0] = bf16_rne(pos[2 * j] - xi);
            row[4 * r + 1] = bf16_rne(pos[2 * j + 1] - yi);
            row[4 * r + 2] = bf16_rne(vel[2 * j] - vel[2 * i]);
            row[4 * r + 3] = bf16_rne(vel[2 * j + 1] - vel[2 * i + 1]);
        }
    }
}
