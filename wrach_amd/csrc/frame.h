// Internal interface between the host side of the library (abi.hip) and the
// frame kernels (kernels.hip).  Not installed; the public ABI is wrach.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wrach {

// One world of a batch of worlds of different sizes: its grid starts at
// cell key `base` (a multiple of kChunkCells), its view at the origin.
struct WorldGrid {
    uint32_t base;    // first cell key
    uint32_t gx, gy;  // grid dimensions
    float x1, y1;     // the walls: the world's width and height
};

// Per-frame constants, passed by value to every kernel.  x1/y1 are
// view_anchor + view_dimensions evaluated in f32, the viewport corner of
// Particle::enforce_boundaries (shaders/physics/src/particle.rs:46-52).
struct FrameParams {
    float ax, ay;        // view_anchor
    float x1, y1;        // view_anchor + view_dimensions
    float cs;            // cell size as f32 (particles_per_cell.wgsl:18)
    // Extension: v += gravity * dt, then p += v * dt.  gdx/gdy hold
    // gravity * dt, or -0.0 for a zero component: adding -0.0 leaves every
    // value unchanged, -0.0 included, which is what skipping the add does.
    // p + v * 1 is p + v exactly, so dt == 1 needs no special case.
    float gdx, gdy;
    float dt;
    uint32_t gx, gy;     // grid_dimensions
    uint32_t ncell;      // gx * gy, times the worlds of a batch
    uint32_t n;          // particles_in_frame_count (all worlds of a batch)
    // A batch of equal worlds: world w owns cell keys [w * world_cells,
    // (w + 1) * world_cells); one world: world_cells == ncell.
    uint32_t world_cells;
    // A batch of worlds of different sizes: one WorldGrid per world and the
    // world of every chunk of kChunkCells cells (null for one world or a
    // batch of equal worlds; gx, gy, x1, y1 and world_cells then apply).
    const WorldGrid *grids;
    const uint32_t *chunk_world;
    uint32_t n_grids;
    // The scan's look-back gives up after wait_limit rounds without finding
    // its prefix and raises kErrLookBack in FrameBuffers::error (the host
    // turns that into WRACH_E_DEVICE at its next synchronisation).
    // epoch_skew is 0 except in a test that forces that path from the host:
    // the look-back then expects an epoch no status word carries.
    uint32_t wait_limit;
    uint32_t epoch_skew;
};

// Look-back rounds before the scan gives up.  A round is a few L2 round
// trips (about a microsecond); a legitimate wait is bounded by the time the
// earlier tiles, dispatched first, take to finish (under a millisecond for
// the largest grids here), so this only ends a wait that would never end.
constexpr uint32_t kDefaultWaitLimit = 1u << 20;
// Bits of FrameBuffers::error.
constexpr uint32_t kErrLookBack = 1u;
// The collide kernel found a cell whose keys and offsets disagree (a layout
// no upload or frame produces); it queues nothing for that cell.
constexpr uint32_t kErrLayout = 2u;

// Device buffers of one world.  `idx` has ncell + 2 entries and is placed 3
// words past a 256-byte boundary so that &idx[1] is 16-byte aligned (the
// scan writes idx[1..] with 128-bit stores).
struct FrameBuffers {
    uint32_t *idx;        // [0, start_0, ..., start_{C-1}, N]
    float2 *pos, *vel;    // cell-sorted state; in-cell order is by `src`
    uint32_t *src;        // canonical index of each particle in the previous frame
    uint32_t *key;        // cell of each packed particle (the segment of idx that holds it)
    float2 *pos_out, *vel_out;  // post-collide state in canonical order
    uint2 *kslot;         // (new cell key, slot inside the new cell)
    unsigned long long *sortbuf;  // merge-sort ping-pong for very full cells
    uint32_t *cnt;        // per-cell counters, zero between frames
    unsigned long long *tile_state;  // scan look-back words
    // Per chunk of kChunkCells cells: chunk_next is set by the first slot
    // claimed in the chunk during a frame (the scan visits only those and
    // clears it); chunk_live is what the last scan found; chunk_prefix the
    // offset at the chunk's first cell.  idx entries of a chunk that is not
    // live are left as they were until a readback fills them in.
    uint8_t *chunk_next, *chunk_live;
    uint32_t *chunk_prefix;
    uint32_t *epoch;      // frame counter tagging the scan's status words; the scatter advances it
    uint32_t *error;      // kErr* bits raised by a kernel that gave up a wait; zero otherwise
    uint32_t *canon_ctr;  // the readback's queue of large cells (zero outside launch_canonical)
    uint4 *cells;         // queued cells (s, n, cell, 0), one region per size class
    uint32_t *class_ctr;  // fill of every slice and overflow list (ctr_index); reset by the scan
    uint32_t class_off[8];  // first entry of each class's region in `cells`
    uint32_t slice_cap[8];  // entries per slice of a class
    uint32_t class_cap[8];  // most cells a class can hold (its overflow list's size)
    uint32_t scan_tiles;
    uint32_t cells_grid;  // blocks of the queued-cell kernel
    uint32_t scan_shape;     // index into kScanShapes
    uint32_t scatter_shape;  // index into kScatterPerLane
};

constexpr int kCollideBlock = 256;
constexpr uint32_t kChunkShift = 10;  // a chunk: the cells of one collide block
constexpr uint32_t kChunkCells = 1u << kChunkShift;
// Launch shapes the autotuner (wrach_autotune) chooses among; entry 0 of
// each table is the default (the scan's: see default_scan_shape).  Scan
// tiles: threads x counts per thread (1024 x 8 is the fastest in the
// 1M-particle frame).  Scatter: particles per lane of a 256-thread block.
struct ScanShape {
    int threads, items;
};
// The shapes of 16k to 64k cells per tile are for large grids that are
// mostly empty (one large world, a batch of worlds), where the number of
// tiles and the look-back between them set the time.  A 1024 x 32 or
// 1024 x 64 tile needs 74 or 105 registers, so one tile runs on a CU at a
// time; 512 x 64 and 256 x 128 tiles (103 and 200 registers) run two to a
// CU, one loading while the other looks back.
constexpr ScanShape kScanShapes[] = {{1024, 8}, {512, 16}, {1024, 4}, {512, 8}, {1024, 16}, {1024, 32},
                                     {1024, 64}, {512, 64}, {256, 128}};
constexpr int kNumScanShapes = 9;
// The scan shape a context starts with: entry 0 up to kLargeScanGrid cells,
// 1024 x 64 up to kHugeScanGrid, 256 x 128 above.  Measured (median of
// interleaved rounds): the 1M lattice's 2.1M cells scan in 8.1 us at
// 1024 x 8, 18.3 at 1024 x 32, 31.7 at 1024 x 64 and 17.4 at 256 x 128;
// the 4M blobs' 16.8M cells in 67.5, 57.8, 50.3 and 53.4 us; the
// 512-world batch's 227M cells at 1024 x 32, 1024 x 64, 512 x 64 and
// 256 x 128 in 401, 344, 338 and 322 us.
constexpr uint64_t kLargeScanGrid = 1ull << 23;
constexpr uint64_t kHugeScanGrid = 1ull << 26;
constexpr uint32_t kLargeScanShape = 6;
constexpr uint32_t kHugeScanShape = 8;
static_assert(kScanShapes[kLargeScanShape].threads == 1024 && kScanShapes[kLargeScanShape].items == 64,
              "the large-grid scan shape");
static_assert(kScanShapes[kHugeScanShape].threads == 256 && kScanShapes[kHugeScanShape].items == 128,
              "the huge-grid scan shape");
constexpr uint32_t default_scan_shape(uint64_t ncell) {
    return ncell > kHugeScanGrid ? kHugeScanShape : ncell > kLargeScanGrid ? kLargeScanShape : 0u;
}
constexpr uint32_t kMinScanTile = 4096;  // smallest threads x items above (sizes tile_state)
constexpr int kScatterBlock = 256;
constexpr int kScatterPerLane[] = {2, 1, 4};
constexpr int kNumScatterShapes = 3;
constexpr int kMaxInCell = 9;            // cell.rs:29-30
constexpr int kRegCell = 16;             // cells up to this size sort in registers
constexpr int kLdsSort = 2048;           // keys per LDS bitonic sort (block path)
constexpr int kWavesPerBlock = kCollideBlock / 64;

// Size classes of queued cells.  Classes 0..5 hold cells of 2, 3-4, 5-8,
// 9-16, 17-32 and 33-64 particles, processed by groups of 2..64 lanes with
// one particle per lane; class 6 (65..kWideMax) takes a wavefront per cell
// with its particles in registers; class 7 (more) a block per cell.
constexpr int kWideChunks = 11;           // 64-particle chunks of a wide-path wavefront
constexpr uint32_t kWideMax = 64 * kWideChunks;
constexpr int kClasses = 8;
constexpr int kGroupClasses = 6;
constexpr int kWaveClass = 6;
constexpr int kBlockClass = 7;
// Fewest particles a cell of each class holds (sizes the list regions).
constexpr uint32_t kClassMinN[kClasses] = {2, 3, 5, 9, 17, 33, 65, kWideMax + 1};
static_assert(sizeof(FrameBuffers::class_off) == kClasses * sizeof(uint32_t), "one region per class");

// A class's region is kSlices slices of slice_cap entries, then an overflow
// list of class_cap entries.  Collide block b appends to slice b % kSlices
// and spills to the overflow list when the slice is full: counters that many
// blocks hit serialise their atomics, one instruction at a time per cache
// line, so every counter has a 128-byte line of its own.
constexpr int kSlices = 64;
constexpr int kCtrStride = 32;  // words between counters
__host__ __device__ constexpr uint32_t ctr_index(int k, int s) {  // s == kSlices: overflow
    return (uint32_t)((k * (kSlices + 1) + s) * kCtrStride);
}
constexpr uint32_t kCtrWords = ctr_index(kClasses, 0);

inline uint32_t scan_tiles_for(uint32_t ncell, uint32_t shape) {
    const uint32_t tile = (uint32_t)(kScanShapes[shape].threads * kScanShapes[shape].items);
    return (ncell + tile - 1) / tile;
}

// Launchers (kernels.hip).  Each enqueues on `stream` and returns the launch
// status; the frame kernels' optional events are stamped with the kernel's
// start and end.
hipError_t launch_collide(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                          hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_cells(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                        hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// Blocks of the queued-cell kernel that the device holds at once (all CUs).
uint32_t resident_blocks_cells();
hipError_t launch_scan(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                       hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_scatter(const FrameParams &p, const FrameBuffers &b, hipStream_t stream,
                          hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_canonical(const FrameParams &p, const FrameBuffers &b, hipStream_t stream);
// Fills the idx entries of chunks the last scan found empty (before a readback).
hipError_t launch_materialize(const FrameParams &p, const FrameBuffers &b, hipStream_t stream);
hipError_t launch_iota(uint32_t *dst, uint32_t n, hipStream_t stream);
// key[i] = the cell whose range [idx[c + 1], idx[c + 2]) holds packed
// particle i, for offsets given whole (an upload from the host, a
// checkpoint): a binary search per particle.
hipError_t launch_keys_from_offsets(const uint32_t *idx, uint32_t ncell, uint32_t n, uint32_t *key,
                                    hipStream_t stream);
// n words of v / of src (kernels, so that a captured stream records no memset).
hipError_t launch_fill32(uint32_t *dst, uint32_t v, size_t n, hipStream_t stream);
hipError_t launch_copy32(uint32_t *dst, const uint32_t *src, size_t n, hipStream_t stream);

// Upload binning on the device: the store (every particle added, in
// insertion order) and where its particles go.
struct UploadParams {
    const float2 *pos;  // store positions
    const float2 *vel;  // store velocities
    uint2 *ks;          // per store particle: (cell key or ~0, slot in the cell)
    uint32_t *valid;    // particles inside the grid (one counter)
    uint32_t n;         // store size, < 2^31
    float cs;           // cell side
    uint32_t gx, gy;    // grid dimensions
    uint32_t cap;       // packed-array capacity
    const uint32_t *world;  // a batch: world of each store particle (else null)
    uint32_t world_cells;   // cells of one world (a world's keys start at world * world_cells)
    const WorldGrid *grids;  // worlds of different sizes: each world's grid (else null)
    int32_t blx, bly;       // the grid's bottom-left cell: the view anchor / cell size
};
// Cell and slot of every store particle, counts into b.cnt, inside-grid
// total into u.valid (zeroed by the caller).
hipError_t launch_upload_bin(const UploadParams &u, const FrameBuffers &b, hipStream_t stream);
// The store's particles outside the grid in store order, copied to
// dpos/dvel from index base on: counts[b] receives block b's number (blocks
// of 256 store particles), and after the caller has turned them into
// exclusive offsets, launch_store_outside writes them.
hipError_t launch_store_outside_count(const UploadParams &u, uint32_t *counts, hipStream_t stream);
hipError_t launch_store_outside(const UploadParams &u, float2 *dpos, float2 *dvel, uint32_t base,
                                const uint32_t *offsets, hipStream_t stream);
// After launch_scan: copy each particle to idx[key + 1] + slot, marked as a
// run of its own (src = store index | run flag).
hipError_t launch_upload_scatter(const UploadParams &u, const FrameBuffers &b, hipStream_t stream);

// Mode S: separation passes over the 3x3 cells around each particle before
// the reference step (the kernels' comment has the rule).  `words` holds the
// largest overlap (float bits) after 0..max_passes passes for the last
// frame; the passes that ran are those whose start exceeded the tolerance.
constexpr uint32_t kMaxStencilPasses = 64;
constexpr uint32_t kStencilWords = kMaxStencilPasses + 4;
constexpr uint32_t kStencilBudget = 4096;  // most particles one lane visits in its 3x3 cells
// Overlap sums (1 - d in 2^-24 units) and pair counts of the frame's first
// and last measure (sets 0 and 1), and of the learned warm start's output
// (set 2), striped over kStencilSumSlots counters 128 bytes apart (block b
// adds to slot b % kStencilSumSlots): entry
// ((which * kStencilSumSlots + slot) * kStencilSumStride) holds the sum,
// the next one the count.
constexpr uint32_t kStencilSumSlots = 64;
constexpr uint32_t kStencilSumStride = 16;  // u64 entries between slots
constexpr uint32_t kStencilSumSets = 3;
constexpr uint32_t kStencilSumEntries = kStencilSumSets * kStencilSumSlots * kStencilSumStride;
// words[kFallbackWord]: 1 when the last learned frame fell back to the
// frame-start positions (the warm start left more mean overlap than it
// found), else 0.
constexpr uint32_t kFallbackWord = kMaxStencilPasses + 1;
// words[kCrowdCountWord]: the crowded cells listed for the frame.
constexpr uint32_t kCrowdCountWord = kMaxStencilPasses + 2;
// words[kOverBudgetWord]: nonzero once a stencil kernel of the frame has met
// a particle whose 3x3 cells hold more than kStencilBudget particles (the
// cells are the frame start's, so the frame's first pass or measure finds
// every such particle).
constexpr uint32_t kOverBudgetWord = kMaxStencilPasses + 3;
constexpr float kOverlapScale = 16777216.0f;  // 2^24
// Crowded cells (kCrowdMin..kCrowdMax particles): the separation kernels'
// lanes take their particles in order of x (the lane order; the results
// do not depend on it), so that a wavefront's particles lie close together
// and skip more of their neighbours together.
constexpr uint32_t kCrowdMin = 64;
constexpr uint32_t kCrowdMax = 1024;
struct StencilArgs {
    const float2 *pin;  // positions entering the kernel
    float2 *pout;       // a move's output
    uint32_t *cell;     // each particle's cell at the frame's start
    uint32_t *words;    // kStencilWords words
    unsigned long long *sums;  // kStencilSumEntries entries
    uint32_t stat;      // a measure: 0, or 1 / 2 to add its overlaps to the first / last sums
    uint32_t in_slot;   // a measure: the word it raises; a move: the word of its start
    uint32_t first;     // the frame's first kernel: derives and writes `cell`
    uint32_t check;     // a move: skip (copy) when the measured start is within tol
    float tol;          // a move runs while the overlap at its start exceeds this
    // A move after a learned warm start (else null): it reads pin_alt (the
    // frame-start positions) instead of pin when the mean overlap of the
    // warm start (sums set 2) exceeds that of the frame start (set 0), and
    // block 0 records the choice in words[kFallbackWord].
    const float2 *pin_alt;
    // The lane order (null: lane t takes particle t): lane t takes particle
    // perm[t], a permutation that orders each crowded cell by x; the
    // crowded cells of the frame (words[kCrowdCountWord] of them, at most
    // crowd_cap).
    uint32_t *perm;
    uint32_t *crowd;
    uint32_t crowd_cap;
    // A measure after a pass or measure of the frame: each pair from one
    // side only (the particle's own cell past it, the cell to its right and
    // the three cells above), the pair count and overlap sum doubled, unless
    // words[kOverBudgetWord] is set (a particle over the budget takes no
    // pair from its own side, so the two sides differ).
    uint32_t half;
};
hipError_t launch_stencil_measure(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                                  hipStream_t stream);
hipError_t launch_stencil_move(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                               hipStream_t stream);
// The frame's first stencil launch: clears the overlap words and sums and,
// with a.perm, orders the lanes of the crowded cells by the x of `pos`.
hipError_t launch_stencil_clear(const FrameParams &p, const FrameBuffers &b, const StencilArgs &a,
                                const float2 *pos, hipStream_t stream);
hipError_t launch_copy_pos(float2 *dst, const float2 *src, uint32_t n, hipStream_t stream);
// Adds to out[0..3] (zeroed by the caller) the mismatches of the pair
// arithmetic's short root and force sequences against the correctly
// rounded ones over their whole input ranges, and the inputs tried.
hipError_t launch_check_pair_arithmetic(unsigned long long *out, hipStream_t stream);

}  // namespace wrach
