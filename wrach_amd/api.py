"""Python mirror of wrach's `runners/api` surface over the C ABI.

`WrachAPI`, `WrachConfig`, `Particle` and `WrachState` keep the names,
argument meaning and error behaviour of runners/api/src/lib.rs:17-88,
runners/bevy/src/config_app.rs:10-35 and runners/bevy/src/state.rs:17-59, so
tests read like the reference's own.  `World` is the thin numpy-level handle
used by the bench and the parity tests.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import WrachError, check

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


@dataclass
class WrachConfig:
    """WrachConfig (config_app.rs:10-35); defaults of config_app.rs:26-35.

    gravity and dt are extensions: (0, 0) and 1.0 reproduce the reference.
    host_pack packs uploads on the host (the reference's CPU packer) instead
    of binning the store on the device; the packed triple is the same."""

    dimensions: tuple = (480, 352)
    boundaries_as_dimensions: bool = False
    cell_size: int = 3
    gravity: tuple = (0.0, 0.0)
    dt: float = 1.0
    device: int = 0
    host_pack: bool = False

    def to_c(self) -> _lib.WrachConfigC:
        c = _lib.WrachConfigC()
        _lib.load().wrach_config_default(ctypes.byref(c))
        c.width, c.height = int(self.dimensions[0]), int(self.dimensions[1])
        c.cell_size = int(self.cell_size)
        c.boundaries_as_dimensions = 1 if self.boundaries_as_dimensions else 0
        c.gravity[0], c.gravity[1] = float(self.gravity[0]), float(self.gravity[1])
        c.dt = float(self.dt)
        c.device = int(self.device)
        c.flags = _lib.WRACH_FLAG_HOST_PACK if self.host_pack else 0
        return c


@dataclass
class Particle:
    """Particle (state.rs:41-51)."""

    position: tuple
    velocity: tuple


@dataclass
class PackedData:
    """PackedData (spatial_bin.rs:22-33); positions/velocities are capacity
    long and zero padded, as the staging buffers the reference reads back."""

    indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    velocities: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))


@dataclass
class ShaderWorldSettings:
    """ShaderWorldSettings (config_shader.rs:15-29)."""

    view_dimensions: tuple
    view_anchor: tuple
    grid_dimensions: tuple
    cell_size: int
    particles_in_frame_count: int


@dataclass
class WrachState:
    """The parts of WrachState (state.rs:17-33) a caller reads."""

    config: WrachConfig
    shader_settings: ShaderWorldSettings
    packed_data: PackedData


class World:
    """One simulated world on one GPU (a wrach_ctx)."""

    def __init__(self, config, worlds: int = 1):
        """worlds > 1: a batch of equal worlds stepped together
        (wrach_create_batch); a list of configurations: a batch of worlds of
        different sizes (wrach_create_worlds).  add_particles(..., world=w)
        and read_world(w) address one of them."""
        L = _lib.load()
        self._L = L
        h = ctypes.c_void_p()
        if isinstance(config, (list, tuple)):
            self.config = config[0]
            self.worlds = len(config)
            cfgs = (_lib.WrachConfigC * self.worlds)(*[c.to_c() for c in config])
            check(L.wrach_create_worlds(cfgs, self.worlds, ctypes.byref(h)))
        else:
            self.config = config
            self.worlds = int(worlds)
            cfg = config.to_c()
            if self.worlds == 1:
                check(L.wrach_create(ctypes.byref(cfg), ctypes.byref(h)))
            else:
                check(L.wrach_create_batch(ctypes.byref(cfg), self.worlds, ctypes.byref(h)))
        self._h = h
        cap, nidx = ctypes.c_uint32(), ctypes.c_uint32()
        check(L.wrach_capacity(h, ctypes.byref(cap), ctypes.byref(nidx)))
        self.capacity = cap.value
        self.n_indices = nidx.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.wrach_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_particles(self, pos, vel, world: int | None = None) -> None:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 2)
        if pos.shape != vel.shape:
            raise ValueError("positions and velocities differ in length")
        if world is None:
            check(self._L.wrach_add_particles(self._h, pos.ctypes.data_as(_f32p),
                                              vel.ctypes.data_as(_f32p), len(pos)))
        else:
            check(self._L.wrach_add_particles_to_world(self._h, int(world), pos.ctypes.data_as(_f32p),
                                                       vel.ctypes.data_as(_f32p), len(pos)))

    def world_settings(self, world: int):
        """One world's settings (wrach_settings_world): view and grid
        dimensions, cell size, particles."""
        s = _lib.WorldSettingsC()
        check(self._L.wrach_settings_world(self._h, int(world), ctypes.byref(s)))
        return s

    def read_world(self, world: int):
        """One world of a batch: (indices, positions[N,2], velocities[N,2])."""
        ws = self.world_settings(world)
        cells = ws.grid_dimensions[0] * ws.grid_dimensions[1]
        n = ws.particles_in_frame_count
        pos = np.zeros((max(n, 1), 2), np.float32)
        vel = np.zeros((max(n, 1), 2), np.float32)
        idx = np.zeros(cells + 2, np.uint32)
        got = ctypes.c_uint32()
        check(self._L.wrach_read_world(self._h, int(world), pos.ctypes.data_as(_f32p),
                                       vel.ctypes.data_as(_f32p), pos.shape[0],
                                       idx.ctypes.data_as(_u32p), len(idx), ctypes.byref(got)))
        return idx, pos[:got.value], vel[:got.value]

    def step(self, nframes: int = 1) -> None:
        check(self._L.wrach_step(self._h, int(nframes)))

    def prepare(self, nframes: int) -> None:
        """Captures the graphs step(nframes) launches, running no frame
        (wrach_prepare)."""
        check(self._L.wrach_prepare(self._h, int(nframes)))

    def test_scan_wait(self, wait_limit: int, epoch_skew: int = 0) -> None:
        """Testing hook (wrach_test_scan_wait): the scan's look-back bound and
        an epoch skew that makes every look-back give up."""
        check(self._L.wrach_test_scan_wait(self._h, int(wait_limit), int(epoch_skew)))

    _TUNING = ("scan_threads", "scan_items", "scatter_per_lane", "cells_grid", "frame_us")

    def tuning(self) -> dict:
        """Current launch shapes (wrach_get_tuning)."""
        t = _lib.TuningC()
        check(self._L.wrach_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k in self._TUNING}

    def set_tuning(self, **shapes) -> None:
        """Sets launch shapes (wrach_set_tuning); unnamed ones stay."""
        t = _lib.TuningC()
        check(self._L.wrach_get_tuning(self._h, ctypes.byref(t)))
        for k, v in shapes.items():
            if k not in self._TUNING[:4]:
                raise ValueError(f"unknown launch shape {k}")
            setattr(t, k, int(v))
        check(self._L.wrach_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, frames: int = 20) -> dict:
        """Times the candidate launch shapes on the current state and keeps
        the fastest (wrach_autotune); the state is unchanged."""
        t = _lib.TuningC()
        check(self._L.wrach_autotune(self._h, int(frames), ctypes.byref(t)))
        return {k: getattr(t, k) for k in self._TUNING}

    def store_write_back(self) -> None:
        """The store becomes the simulated state (wrach_store_write_back)."""
        check(self._L.wrach_store_write_back(self._h))

    def load_packed(self, indices, pos, vel) -> None:
        """Restores a packed triple saved with read() (wrach_load_packed)."""
        indices = np.ascontiguousarray(indices, np.uint32)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 2)
        if pos.shape != vel.shape:
            raise ValueError("positions and velocities differ in length")
        check(self._L.wrach_load_packed(self._h, indices.ctypes.data_as(_u32p), len(indices),
                                        pos.ctypes.data_as(_f32p), vel.ctypes.data_as(_f32p),
                                        len(pos)))

    def sync(self) -> None:
        check(self._L.wrach_sync(self._h))

    def set_view_anchor(self, x: float, y: float) -> None:
        """wrach_set_view_anchor: move the simulated window over the store."""
        check(self._L.wrach_set_view_anchor(self._h, float(x), float(y)))

    def set_solver(self, mode: str = "R", max_passes: int = 0, overlap_tol: float = 0.0,
                   quiet: bool = False) -> None:
        """wrach_set_solver: "R", the reference's step; "S", up to
        max_passes separation passes over the 3x3 cells around each particle
        (each while the largest overlap exceeds overlap_tol), then that step;
        "L", the learned warm start, then max_passes such passes (tolerance
        0), then that step.  quiet (S, L): no measure after the passes
        (WRACH_SOLVER_QUIET)."""
        modes = {"R": _lib.WRACH_SOLVER_R, "S": _lib.WRACH_SOLVER_S, "L": _lib.WRACH_SOLVER_L}
        if mode not in modes:
            raise ValueError(f"unknown solver {mode!r}")
        m = modes[mode] | (_lib.WRACH_SOLVER_QUIET if quiet else 0)
        check(self._L.wrach_set_solver(self._h, m, int(max_passes), float(overlap_tol)))

    def after_measured(self) -> bool:
        """Whether the last report holds the figures after the passes."""
        r = _lib.SolverReportC()
        check(self._L.wrach_solver_stats(self._h, ctypes.byref(r)))
        return bool(r.after_measured)

    def solver_stats(self):
        """The last mode S frame (wrach_solver_stats): (largest overlap
        before the passes, after them, passes run, (overlap sum in 2^-24
        units, pairs) before + the same after).  The mean overlap is
        sum / 2**24 / pairs."""
        r = _lib.SolverReportC()
        check(self._L.wrach_solver_stats(self._h, ctypes.byref(r)))
        return (r.max_overlap_before, r.max_overlap_after, r.passes,
                (r.overlap_sum_before, r.pairs_before, r.overlap_sum_after, r.pairs_after))

    # -------------------------------------------------- learned warm start

    def learn_set_params(self, params) -> None:
        """wrach_learn_set_params: the network's f32 parameter vector
        (resets Adam)."""
        params = np.ascontiguousarray(params, np.float32).ravel()
        check(self._L.wrach_learn_set_params(self._h, params.ctypes.data_as(_f32p), len(params)))

    def learn_params(self) -> np.ndarray:
        out = np.zeros(_lib.WRACH_LEARN_PARAMS, np.float32)
        check(self._L.wrach_learn_get_params(self._h, out.ctypes.data_as(_f32p), len(out)))
        return out

    def learn_features(self):
        """(feature rows as uint16 bf16 bits [N, 64] in canonical order,
        (overlap sum in 2^-24 units, pairs)) of the current state."""
        n = self.n_particles
        rows = np.zeros((max(n, 1), _lib.WRACH_LEARN_FEATURES), np.uint16)
        ov = np.zeros(2, np.uint64)
        check(self._L.wrach_learn_features(self._h, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                           rows.shape[0], ov.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return rows[:n], (int(ov[0]), int(ov[1]))

    def learn_forward(self, rows) -> np.ndarray:
        rows = np.ascontiguousarray(rows, np.uint16).reshape(-1, _lib.WRACH_LEARN_FEATURES)
        y = np.zeros((len(rows), 2), np.float32)
        check(self._L.wrach_learn_forward(self._h, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                          len(rows), y.ctypes.data_as(_f32p)))
        return y

    def learn_gradient(self, rows, targets):
        rows = np.ascontiguousarray(rows, np.uint16).reshape(-1, _lib.WRACH_LEARN_FEATURES)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 2)
        grad = np.zeros(_lib.WRACH_LEARN_PARAMS, np.float32)
        loss = ctypes.c_float()
        check(self._L.wrach_learn_gradient(self._h, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                           targets.ctypes.data_as(_f32p), len(rows),
                                           grad.ctypes.data_as(_f32p), ctypes.byref(loss)))
        return grad, loss.value

    def learn_adam(self, grad, lr: float) -> None:
        grad = np.ascontiguousarray(grad, np.float32).ravel()
        check(self._L.wrach_learn_adam(self._h, grad.ctypes.data_as(_f32p), float(lr)))

    def learn_sample(self, target_passes: int) -> int:
        """Appends the state's features and target_passes-pass displacements
        to the training set; returns its size."""
        total = ctypes.c_uint64()
        check(self._L.wrach_learn_sample(self._h, int(target_passes), ctypes.byref(total)))
        return total.value

    def learn_clear_samples(self) -> None:
        check(self._L.wrach_learn_clear_samples(self._h))

    def learn_train(self, order, batch: int, lr: float) -> np.ndarray:
        """Adam steps over consecutive batches of `order` (training rows);
        returns each step's loss."""
        order = np.ascontiguousarray(order, np.uint32)
        steps = len(order) // int(batch)
        losses = np.zeros(max(steps, 1), np.float32)
        check(self._L.wrach_learn_train(self._h, order.ctypes.data_as(_u32p), len(order), int(batch),
                                        float(lr), losses.ctypes.data_as(_f32p)))
        return losses[:steps]

    def learn_profile(self, reps: int) -> dict:
        """Summed milliseconds of the features and network kernels over
        reps runs on the current state (wrach_learn_profile)."""
        ms = (ctypes.c_double * 2)()
        check(self._L.wrach_learn_profile(self._h, int(reps), ms))
        return {"features": ms[0], "network": ms[1]}

    def learned_stats(self) -> dict:
        r = _lib.LearnedReportC()
        check(self._L.wrach_learned_stats(self._h, ctypes.byref(r)))
        return {"start_overlap_sum": r.start_overlap_sum, "start_pairs": r.start_pairs,
                "warm_overlap_sum": r.warm_overlap_sum, "warm_pairs": r.warm_pairs,
                "fell_back": bool(r.fell_back)}

    def set_graph(self, enable: bool) -> None:
        check(self._L.wrach_set_graph(self._h, 1 if enable else 0))

    def settings(self) -> ShaderWorldSettings:
        s = _lib.WorldSettingsC()
        check(self._L.wrach_settings(self._h, ctypes.byref(s)))
        return ShaderWorldSettings(tuple(s.view_dimensions), tuple(s.view_anchor),
                                   tuple(s.grid_dimensions), s.cell_size,
                                   s.particles_in_frame_count)

    @property
    def n_particles(self) -> int:
        return self.settings().particles_in_frame_count

    def read(self):
        """Canonical packed triple (indices, positions[N,2], velocities[N,2])."""
        n = self.n_particles
        pos = np.zeros((max(n, 1), 2), np.float32)
        vel = np.zeros((max(n, 1), 2), np.float32)
        idx = np.zeros(self.n_indices, np.uint32)
        got = ctypes.c_uint32()
        check(self._L.wrach_read(self._h, pos.ctypes.data_as(_f32p), vel.ctypes.data_as(_f32p),
                                 pos.shape[0], idx.ctypes.data_as(_u32p), len(idx),
                                 ctypes.byref(got)))
        return idx, pos[:got.value], vel[:got.value]

    def step_profiled(self, nframes: int):
        """Per-kernel summed milliseconds over nframes (HIP events)."""
        ms = (ctypes.c_double * _lib.NUM_FRAME_KERNELS)()
        check(self._L.wrach_step_profiled(self._h, int(nframes), ms))
        return dict(zip(_lib.FRAME_KERNEL_NAMES, list(ms)))

    def device_state(self, canonical: bool = False) -> _lib.DeviceViewC:
        v = _lib.DeviceViewC()
        fn = self._L.wrach_canonical_device if canonical else self._L.wrach_device_state
        check(fn(self._h, ctypes.byref(v)))
        return v


class WrachAPI:
    """WrachAPI (runners/api/src/lib.rs:17-88)."""

    def __init__(self, config: WrachConfig | None = None):
        self.config = config if config is not None else WrachConfig()
        self.world = World(self.config)
        self.positions: list = []
        self.velocities: list = []
        self._packed = PackedData()

    def tick(self) -> None:
        """One frame, then read_data (lib.rs:49-52)."""
        self.world.step(1)
        self._readback()
        self.read_data()

    def _readback(self) -> None:
        # plugin/build.rs:135-158: the three staging buffers, capacity long.
        idx, pos, vel = self.world.read()
        cap = self.world.capacity
        p = np.zeros((cap, 2), np.float32)
        v = np.zeros((cap, 2), np.float32)
        p[:len(pos)] = pos
        v[:len(vel)] = vel
        self._packed = PackedData(idx, p, v)

    def read_data(self) -> None:
        """lib.rs:58-74: positions/velocities as (x, y) tuples."""
        self.positions = [(float(x), float(y)) for x, y in self._packed.positions]
        self.velocities = [(float(x), float(y)) for x, y in self._packed.velocities]

    def add_particles(self, particles) -> None:
        """lib.rs:78-81 -> state.rs:90-101."""
        if not particles:
            pos = np.zeros((0, 2), np.float32)
            vel = np.zeros((0, 2), np.float32)
        else:
            pos = np.array([p.position for p in particles], np.float32)
            vel = np.array([p.velocity for p in particles], np.float32)
        self.world.add_particles(pos, vel)

    def get_simulation_state(self) -> WrachState:
        return WrachState(self.config, self.world.settings(), self._packed)

    def close(self) -> None:
        self.world.close()


__all__ = ["WrachAPI", "WrachConfig", "Particle", "PackedData", "WrachState",
           "ShaderWorldSettings", "World", "WrachError"]
