//! `extern "C"` binding of `libwrach_hip.so`, one to one with
//! `include/wrach.h`, and [`api::WrachAPI`], a wrapper with the surface of
//! wrach's `WrachAPI` (`runners/api/src/lib.rs:17-88`).
//!
//! The structs are `#[repr(C)]` with the header's field order; the sizes and
//! offsets asserted below are the ones the header gives (a test of the
//! repository compiles the header with a C compiler and checks them against
//! this file).
#![allow(non_camel_case_types)]

pub mod api;

use core::ffi::{c_char, c_int, c_void};

/// Opaque context (`wrach_ctx`).
#[repr(C)]
pub struct wrach_ctx {
    _p: [u8; 0],
}

pub const WRACH_OK: c_int = 0;
pub const WRACH_E_INVALID: c_int = -1;
pub const WRACH_E_TOO_MANY_CELLS: c_int = -2;
pub const WRACH_E_CAPACITY: c_int = -3;
pub const WRACH_E_NONFINITE: c_int = -4;
pub const WRACH_E_DEVICE: c_int = -5;
pub const WRACH_E_BUFFER: c_int = -6;
pub const WRACH_E_STATE: c_int = -7;

/// `wrach_config.flags`: pack uploads on the host (default: bin on the device).
pub const WRACH_FLAG_HOST_PACK: u32 = 0x1;
pub const WRACH_NUM_FRAME_KERNELS: usize = 4;
pub const WRACH_SOLVER_R: u32 = 0;
pub const WRACH_SOLVER_S: u32 = 1;
pub const WRACH_SOLVER_L: u32 = 2;
pub const WRACH_SOLVER_QUIET: u32 = 0x100;
pub const WRACH_MAX_PASSES: u32 = 64;
pub const WRACH_STENCIL_BUDGET: u32 = 4096;
pub const WRACH_LEARN_FEATURES: u32 = 64;
pub const WRACH_LEARN_PARAMS: u32 = 25090;
pub const WRACH_LEARN_BUDGET: u32 = 4096;
pub const WRACH_NOALLOC_SCRATCH_BYTES: u32 = 32 << 20;

/// `WrachConfig` (`runners/bevy/src/config_app.rs:10-35`) plus extensions.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct wrach_config {
    pub width: u16,
    pub height: u16,
    pub cell_size: u16,
    pub boundaries_as_dimensions: u8,
    pub reserved0: u8,
    pub gravity: [f32; 2],
    pub dt: f32,
    pub flags: u32,
    pub device: i32,
}

/// `ShaderWorldSettings`, the 32-byte uniform (`runners/bevy/src/config_shader.rs:15-29`).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct wrach_world_settings {
    pub view_dimensions: [f32; 2],
    pub view_anchor: [f32; 2],
    pub grid_dimensions: [u32; 2],
    pub cell_size: u32,
    pub particles_in_frame_count: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct wrach_device_view {
    pub indices: *const u32,
    pub positions: *const f32,
    pub velocities: *const f32,
    pub n_particles: u32,
    pub n_indices: u32,
    pub stream: *mut c_void,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct wrach_tuning {
    pub scan_threads: u32,
    pub scan_items: u32,
    pub scatter_per_lane: u32,
    pub cells_grid: u32,
    pub frame_us: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct wrach_solver_report {
    pub max_overlap_before: f32,
    pub max_overlap_after: f32,
    pub passes: u32,
    pub after_measured: u32,
    pub overlap_sum_before: u64,
    pub pairs_before: u64,
    pub overlap_sum_after: u64,
    pub pairs_after: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct wrach_learned_report {
    pub start_overlap_sum: u64,
    pub start_pairs: u64,
    pub warm_overlap_sum: u64,
    pub warm_pairs: u64,
    pub fell_back: u32,
    pub reserved: u32,
}

// Layouts of include/wrach.h on x86-64 and aarch64 (checked at compile time).
const _: () = assert!(core::mem::size_of::<wrach_config>() == 28);
const _: () = assert!(core::mem::offset_of!(wrach_config, gravity) == 8);
const _: () = assert!(core::mem::offset_of!(wrach_config, device) == 24);
const _: () = assert!(core::mem::size_of::<wrach_world_settings>() == 32);
const _: () = assert!(core::mem::size_of::<wrach_device_view>() == 40);
const _: () = assert!(core::mem::offset_of!(wrach_device_view, stream) == 32);
const _: () = assert!(core::mem::size_of::<wrach_tuning>() == 20);
const _: () = assert!(core::mem::size_of::<wrach_solver_report>() == 48);
const _: () = assert!(core::mem::offset_of!(wrach_solver_report, overlap_sum_before) == 16);
const _: () = assert!(core::mem::size_of::<wrach_learned_report>() == 40);

#[link(name = "wrach_hip")]
extern "C" {
    pub fn wrach_config_default(cfg: *mut wrach_config);
    pub fn wrach_create(cfg: *const wrach_config, out: *mut *mut wrach_ctx) -> c_int;
    pub fn wrach_destroy(ctx: *mut wrach_ctx);
    pub fn wrach_add_particles(ctx: *mut wrach_ctx, pos_xy: *const f32, vel_xy: *const f32, n: u32) -> c_int;
    pub fn wrach_step(ctx: *mut wrach_ctx, nframes: u32) -> c_int;
    pub fn wrach_read(ctx: *mut wrach_ctx, pos_xy: *mut f32, vel_xy: *mut f32, particle_cap: usize,
                      indices: *mut u32, indices_cap: usize, n_particles: *mut u32) -> c_int;
    pub fn wrach_capacity(ctx: *mut wrach_ctx, max_particles: *mut u32, n_indices: *mut u32) -> c_int;
    pub fn wrach_settings(ctx: *mut wrach_ctx, out: *mut wrach_world_settings) -> c_int;
    pub fn wrach_sync(ctx: *mut wrach_ctx) -> c_int;
    pub fn wrach_device_state(ctx: *mut wrach_ctx, out: *mut wrach_device_view) -> c_int;
    pub fn wrach_canonical_device(ctx: *mut wrach_ctx, out: *mut wrach_device_view) -> c_int;
    pub fn wrach_store_write_back(ctx: *mut wrach_ctx) -> c_int;
    pub fn wrach_load_packed(ctx: *mut wrach_ctx, indices: *const u32, n_indices: usize,
                             pos_xy: *const f32, vel_xy: *const f32, n_particles: u32) -> c_int;
    pub fn wrach_create_batch(cfg: *const wrach_config, n_worlds: u32, out: *mut *mut wrach_ctx) -> c_int;
    pub fn wrach_add_particles_to_world(ctx: *mut wrach_ctx, world: u32, pos_xy: *const f32,
                                        vel_xy: *const f32, n: u32) -> c_int;
    pub fn wrach_read_world(ctx: *mut wrach_ctx, world: u32, pos_xy: *mut f32, vel_xy: *mut f32,
                            particle_cap: usize, indices: *mut u32, indices_cap: usize,
                            n_particles: *mut u32) -> c_int;
    pub fn wrach_worlds(ctx: *mut wrach_ctx, n_worlds: *mut u32) -> c_int;
    pub fn wrach_create_worlds(cfgs: *const wrach_config, n_worlds: u32, out: *mut *mut wrach_ctx) -> c_int;
    pub fn wrach_settings_world(ctx: *mut wrach_ctx, world: u32, out: *mut wrach_world_settings) -> c_int;
    pub fn wrach_get_tuning(ctx: *mut wrach_ctx, out: *mut wrach_tuning) -> c_int;
    pub fn wrach_set_tuning(ctx: *mut wrach_ctx, tuning: *const wrach_tuning) -> c_int;
    pub fn wrach_autotune(ctx: *mut wrach_ctx, frames: u32, out: *mut wrach_tuning) -> c_int;
    pub fn wrach_set_view_anchor(ctx: *mut wrach_ctx, x: f32, y: f32) -> c_int;
    pub fn wrach_set_solver(ctx: *mut wrach_ctx, mode: u32, max_passes: u32, overlap_tol: f32) -> c_int;
    pub fn wrach_solver_stats(ctx: *mut wrach_ctx, out: *mut wrach_solver_report) -> c_int;
    pub fn wrach_set_graph(ctx: *mut wrach_ctx, enable: c_int) -> c_int;
    pub fn wrach_step_profiled(ctx: *mut wrach_ctx, nframes: u32, kernel_ms: *mut f64) -> c_int;
    pub fn wrach_prepare(ctx: *mut wrach_ctx, nframes: u32) -> c_int;
    pub fn wrach_test_scan_wait(ctx: *mut wrach_ctx, wait_limit: u32, epoch_skew: u32) -> c_int;
    pub fn wrach_check_pair_arithmetic(device: c_int, out: *mut u64) -> c_int;
    pub fn wrach_learn_set_params(ctx: *mut wrach_ctx, params: *const f32, n: usize) -> c_int;
    pub fn wrach_learn_get_params(ctx: *mut wrach_ctx, params: *mut f32, n: usize) -> c_int;
    pub fn wrach_learn_features(ctx: *mut wrach_ctx, rows: *mut u16, rows_cap: usize, overlap: *mut u64) -> c_int;
    pub fn wrach_learn_forward(ctx: *mut wrach_ctx, rows: *const u16, n: u32, y: *mut f32) -> c_int;
    pub fn wrach_learn_gradient(ctx: *mut wrach_ctx, rows: *const u16, targets: *const f32, n: u32,
                                grad: *mut f32, loss: *mut f32) -> c_int;
    pub fn wrach_learn_adam(ctx: *mut wrach_ctx, grad: *const f32, lr: f32) -> c_int;
    pub fn wrach_learn_sample(ctx: *mut wrach_ctx, target_passes: u32, rows_total: *mut u64) -> c_int;
    pub fn wrach_learn_clear_samples(ctx: *mut wrach_ctx) -> c_int;
    pub fn wrach_learn_train(ctx: *mut wrach_ctx, order: *const u32, n_order: usize, batch: u32, lr: f32,
                             losses: *mut f32) -> c_int;
    pub fn wrach_learned_stats(ctx: *mut wrach_ctx, out: *mut wrach_learned_report) -> c_int;
    pub fn wrach_learn_profile(ctx: *mut wrach_ctx, reps: u32, ms: *mut f64) -> c_int;
    pub fn wrach_frame_noalloc(ws: *const wrach_world_settings, gravity: *const f32, dt: f32,
                               indices: *mut u32, pos_xy: *mut f32, vel_xy: *mut f32, n: u32,
                               stream: *mut c_void) -> c_int;
    pub fn wrach_last_error() -> *const c_char;
}
