//! `WrachAPI` over the HIP library: the surface of wrach's
//! `runners/api/src/lib.rs:17-88` (`new`, `tick`, `read_data`,
//! `add_particles`, `get_simulation_state`, the `positions` / `velocities`
//! fields) with the Bevy app and its wgpu worker replaced by one
//! `wrach_ctx`.  Errors, which the reference turns into panics, are
//! returned as [`WrachError`] by the `try_*` methods; the reference-shaped
//! methods panic with the library's message, as the reference does.
use core::ffi::CStr;

use crate::*;

/// `bevy::math::Vec2` for this crate's purposes.
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct Vec2 {
    pub x: f32,
    pub y: f32,
}

impl Vec2 {
    #[must_use]
    pub const fn new(x: f32, y: f32) -> Self {
        Self { x, y }
    }
}

/// `wrach_bevy::Particle` (`runners/bevy/src/particle_store.rs`).
#[derive(Clone, Copy, Debug, Default)]
pub struct Particle {
    pub position: Vec2,
    pub velocity: Vec2,
}

/// `wrach_bevy::WrachConfig` (`runners/bevy/src/config_app.rs:10-35`) plus
/// the library's extensions (gravity, dt, device).
#[derive(Clone, Copy, Debug)]
pub struct WrachConfig {
    pub dimensions: (u16, u16),
    pub boundaries_as_dimensions: bool,
    pub cell_size: u16,
    pub gravity: (f32, f32),
    pub dt: f32,
    pub device: i32,
}

impl Default for WrachConfig {
    fn default() -> Self {
        let mut c = wrach_config::default();
        // SAFETY: the library fills a plain struct.
        unsafe { wrach_config_default(&mut c) };
        Self {
            dimensions: (c.width, c.height),
            boundaries_as_dimensions: c.boundaries_as_dimensions != 0,
            cell_size: c.cell_size,
            gravity: (c.gravity[0], c.gravity[1]),
            dt: c.dt,
            device: c.device,
        }
    }
}

/// A failed call: the status code and `wrach_last_error`'s message.
#[derive(Clone, Debug)]
pub struct WrachError {
    pub code: i32,
    pub message: String,
}

fn check(code: core::ffi::c_int) -> Result<(), WrachError> {
    if code == WRACH_OK {
        return Ok(());
    }
    // SAFETY: the library returns a NUL-terminated thread-local string.
    let message = unsafe { CStr::from_ptr(wrach_last_error()) }.to_string_lossy().into_owned();
    Err(WrachError { code, message })
}

/// `wrach_bevy::PackedData` (`runners/bevy/src/spatial_bin.rs:22-33`).
#[derive(Clone, Debug, Default)]
pub struct PackedData {
    pub indices: Vec<u32>,
    pub positions: Vec<Vec2>,
    pub velocities: Vec<Vec2>,
}

/// The part of `WrachState` (`runners/bevy/src/state.rs:17-33`) a caller
/// reads: the settings uniform and the last packed data.
#[derive(Clone, Debug, Default)]
pub struct WrachState {
    pub shader_settings: wrach_world_settings,
    pub packed_data: PackedData,
}

/// Main struct for Wrach physics simulations, on the HIP library.
pub struct WrachAPI {
    ctx: *mut wrach_ctx,
    capacity: usize,
    state: WrachState,
    /// All the positions of the particles (capacity long, zero-padded, as
    /// the reference's staging buffer is).
    pub positions: Vec<(f32, f32)>,
    /// All the velocities of the particles.
    pub velocities: Vec<(f32, f32)>,
}

impl WrachAPI {
    /// `WrachAPI::new` (`runners/api/src/lib.rs:30-45`).
    #[must_use]
    pub fn new(config: WrachConfig) -> Self {
        Self::try_new(config).unwrap_or_else(|e| panic!("wrach_create: {}", e.message))
    }

    pub fn try_new(config: WrachConfig) -> Result<Self, WrachError> {
        let c = wrach_config {
            width: config.dimensions.0,
            height: config.dimensions.1,
            cell_size: config.cell_size,
            boundaries_as_dimensions: u8::from(config.boundaries_as_dimensions),
            reserved0: 0,
            gravity: [config.gravity.0, config.gravity.1],
            dt: config.dt,
            flags: 0,
            device: config.device,
        };
        let mut ctx = core::ptr::null_mut();
        // SAFETY: plain pointers to live values.
        check(unsafe { wrach_create(&c, &mut ctx) })?;
        let (mut cap, mut nidx) = (0u32, 0u32);
        // SAFETY: ctx was just created.
        check(unsafe { wrach_capacity(ctx, &mut cap, &mut nidx) })?;
        Ok(Self {
            ctx,
            capacity: cap as usize,
            state: WrachState::default(),
            positions: Vec::new(),
            velocities: Vec::new(),
        })
    }

    /// `WrachAPI::tick` (`lib.rs:49-52`): one frame, then `read_data`.
    pub fn tick(&mut self) {
        self.try_tick().unwrap_or_else(|e| panic!("wrach tick: {}", e.message));
    }

    pub fn try_tick(&mut self) -> Result<(), WrachError> {
        // SAFETY: ctx is live for &mut self.
        check(unsafe { wrach_step(self.ctx, 1) })?;
        self.readback()?;
        self.read_data();
        Ok(())
    }

    /// The readback of the `tick` system (`runners/bevy/src/plugin/build.rs:135-158`).
    fn readback(&mut self) -> Result<(), WrachError> {
        let mut nidx = 0u32;
        // SAFETY: ctx is live.
        check(unsafe { wrach_capacity(self.ctx, core::ptr::null_mut(), &mut nidx) })?;
        let mut pos = vec![0.0f32; 2 * self.capacity];
        let mut vel = vec![0.0f32; 2 * self.capacity];
        let mut idx = vec![0u32; nidx as usize];
        let mut n = 0u32;
        // SAFETY: the arrays hold capacity particles and nidx offsets.
        check(unsafe {
            wrach_read(self.ctx, pos.as_mut_ptr(), vel.as_mut_ptr(), self.capacity, idx.as_mut_ptr(),
                       idx.len(), &mut n)
        })?;
        // SAFETY: ctx is live; settings is a plain struct.
        check(unsafe { wrach_settings(self.ctx, &mut self.state.shader_settings) })?;
        let pairs = |a: &[f32]| a.chunks_exact(2).map(|p| Vec2::new(p[0], p[1])).collect::<Vec<_>>();
        self.state.packed_data = PackedData { indices: idx, positions: pairs(&pos), velocities: pairs(&vel) };
        Ok(())
    }

    /// `WrachAPI::read_data` (`lib.rs:58-74`).
    pub fn read_data(&mut self) {
        self.positions = self.state.packed_data.positions.iter().map(|p| (p.x, p.y)).collect();
        self.velocities = self.state.packed_data.velocities.iter().map(|p| (p.x, p.y)).collect();
    }

    /// `WrachAPI::add_particles` (`lib.rs:78-81`).
    pub fn add_particles(&mut self, particles: Vec<Particle>) {
        self.try_add_particles(&particles).unwrap_or_else(|e| panic!("wrach add_particles: {}", e.message));
    }

    pub fn try_add_particles(&mut self, particles: &[Particle]) -> Result<(), WrachError> {
        let pos: Vec<f32> = particles.iter().flat_map(|p| [p.position.x, p.position.y]).collect();
        let vel: Vec<f32> = particles.iter().flat_map(|p| [p.velocity.x, p.velocity.y]).collect();
        let n = u32::try_from(particles.len())
            .map_err(|_| WrachError { code: WRACH_E_CAPACITY, message: "more than 2^32 particles".into() })?;
        // SAFETY: 2n floats each.
        check(unsafe { wrach_add_particles(self.ctx, pos.as_ptr(), vel.as_ptr(), n) })
    }

    /// `WrachAPI::get_simulation_state` (`lib.rs:85-88`).
    #[must_use]
    pub fn get_simulation_state(&self) -> &WrachState {
        &self.state
    }

    /// The raw context, for the other entry points of `include/wrach.h`.
    #[must_use]
    pub fn raw(&self) -> *mut wrach_ctx {
        self.ctx
    }
}

impl Drop for WrachAPI {
    fn drop(&mut self) {
        // SAFETY: ctx came from wrach_create and is destroyed once.
        unsafe { wrach_destroy(self.ctx) };
    }
}

#[cfg(test)]
mod test {
    use super::*;

    /// The reference's `test_api_returns_data` (`runners/api/src/lib.rs:101-126`).
    #[test]
    fn test_api_returns_data() {
        let mut wrach = WrachAPI::new(WrachConfig { dimensions: (10, 10), cell_size: 3, ..Default::default() });
        let particles = vec![Particle { position: Vec2::new(5.0, 5.0), velocity: Vec2::new(0.5, 0.5) }; 3];
        wrach.add_particles(particles);
        for _ in 0..5 {
            wrach.tick();
        }
        assert_eq!(wrach.positions.len(), 164);
        assert_ne!(wrach.positions[0], (0.0, 0.0));
        assert_eq!(wrach.velocities.len(), 164);
        assert_ne!(wrach.velocities[0], (0.0, 0.0));
    }
}
