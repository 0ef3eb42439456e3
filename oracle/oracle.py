"""ctypes front end of the CPU oracle (wrach_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (wrach_amd) never
imports this module.

Each wrapper names the reference code it restates; the C file carries the
arithmetic.  Arrays are numpy: positions/velocities as (N, 2) float32,
indices as uint32.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
MAX_PARTICLES_IN_CELL = 9  # shaders/physics/src/cell.rs:29-30

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


class Params(ctypes.Structure):
    """WorldSettings (shaders/shared/src/lib.rs:19-31) + gravity/dt extension."""

    _fields_ = [
        ("view_dimensions", ctypes.c_float * 2),
        ("view_anchor", ctypes.c_float * 2),
        ("grid_dimensions", ctypes.c_uint32 * 2),
        ("cell_size", ctypes.c_uint32),
        ("particles_in_frame_count", ctypes.c_uint32),
        ("gravity", ctypes.c_float * 2),
        ("dt", ctypes.c_float),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.wo_cell_coord.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _i32p]
        L.wo_active_grid.argtypes = [_f32p, ctypes.c_float, _i32p, _u32p]
        L.wo_max_particles.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        L.wo_max_particles.restype = ctypes.c_uint32
        L.wo_pack.argtypes = [_f32p, _f32p, ctypes.c_uint32, _f32p, ctypes.c_float,
                              _u32p, _f32p, _f32p]
        L.wo_pack.restype = ctypes.c_uint32
        L.wo_collide_cell.argtypes = [_f32p, ctypes.c_uint32]
        L.wo_finish_particle.argtypes = [ctypes.POINTER(Params), _f32p, _f32p]
        L.wo_physics.argtypes = [ctypes.POINTER(Params), _u32p, _f32p, _f32p, _f32p, _f32p]
        L.wo_cell_key.argtypes = [ctypes.POINTER(Params), ctypes.c_float, ctypes.c_float]
        L.wo_cell_key.restype = ctypes.c_int64
        L.wo_rebin.argtypes = [ctypes.POINTER(Params), _f32p, _f32p, _u32p, _f32p, _f32p]
        L.wo_rebin.restype = ctypes.c_int
        L.wo_step.argtypes = [ctypes.POINTER(Params), _u32p, _f32p, _f32p, _f32p]
        L.wo_step.restype = ctypes.c_int
        L.wo_step_mt.argtypes = [ctypes.POINTER(Params), _u32p, _f32p, _f32p, _f32p, ctypes.c_int]
        L.wo_step_mt.restype = ctypes.c_int
        L.wo_stencil_pass.argtypes = [ctypes.POINTER(Params), _u32p, _f32p, _f32p, _f32p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint64)]
        L.wo_stencil_pass.restype = ctypes.c_float
        L.wo_step_s.argtypes = [ctypes.POINTER(Params), _u32p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                ctypes.c_float, _f32p, _u32p, ctypes.POINTER(ctypes.c_uint64)]
        L.wo_step_s.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def cell_coord(x: float, y: float, cell_size: float):
    """SpatialBin::get_cell_coord (runners/bevy/src/spatial_bin.rs:48-64)."""
    out = np.zeros(2, np.int32)
    lib().wo_cell_coord(x, y, cell_size, _p(out, _i32p))
    return int(out[0]), int(out[1])


def active_grid(viewport, cell_size: float):
    """SpatialBin::get_active_cells (spatial_bin.rs:68-95) -> (bottom_left, grid)."""
    vp = np.asarray(viewport, np.float32)
    bl = np.zeros(2, np.int32)
    g = np.zeros(2, np.uint32)
    lib().wo_active_grid(_p(vp, _f32p), cell_size, _p(bl, _i32p), _p(g, _u32p))
    return (int(bl[0]), int(bl[1])), (int(g[0]), int(g[1]))


def max_particles(n_cells: int, cell_size: int) -> int:
    """ParticleStore::max_particles_per_frame (particle_store.rs:116-133)."""
    return int(lib().wo_max_particles(n_cells, cell_size))


def pack(pos, vel, viewport, cell_size: float):
    """ParticleStore::create_packed_data (particle_store.rs:90-103,
    spatial_bin.rs:103-149) -> (indices, positions, velocities)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 2)
    _, g = active_grid(viewport, cell_size)
    vp = np.asarray(viewport, np.float32)
    idx = np.zeros(g[0] * g[1] + 2, np.uint32)
    po = np.zeros_like(pos)
    vo = np.zeros_like(vel)
    n = lib().wo_pack(_p(pos, _f32p), _p(vel, _f32p), len(pos), _p(vp, _f32p), cell_size,
                      _p(idx, _u32p), _p(po, _f32p), _p(vo, _f32p))
    return idx, po[:n].copy(), vo[:n].copy()


def collide_cell(pos):
    """Particles::pairs (shaders/physics/src/particles.rs:62-93) on <= 9 particles."""
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 2).copy()
    lib().wo_collide_cell(_p(p, _f32p), len(p))
    return p


def params(width, height, cell_size, n, gravity=(0.0, 0.0), dt=1.0, anchor=(0.0, 0.0)) -> Params:
    """WorldSettings as built by PhysicsComputeWorker::build (builder.rs:53-65)."""
    vp = (anchor[0], anchor[1], anchor[0] + width, anchor[1] + height)
    _, g = active_grid(vp, float(cell_size))
    P = Params()
    P.view_dimensions[:] = (float(width), float(height))
    P.view_anchor[:] = anchor
    P.grid_dimensions[:] = g
    P.cell_size = int(cell_size)
    P.particles_in_frame_count = int(n)
    P.gravity[:] = gravity
    P.dt = dt
    return P


def finish_particle(P: Params, pos, vel):
    """Particle::integrate + enforce_limits (shaders/physics/src/particle.rs:40-82)."""
    p = np.array(pos, np.float32)
    v = np.array(vel, np.float32)
    lib().wo_finish_particle(ctypes.byref(P), _p(p, _f32p), _p(v, _f32p))
    return p, v


def physics(P: Params, idx, pos, vel):
    """The physics pass (shaders/physics/src/lib.rs:41-64, cell.rs:52-112)."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    po = np.zeros_like(pos)
    vo = np.zeros_like(vel)
    lib().wo_physics(ctypes.byref(P), _p(idx, _u32p), _p(pos, _f32p), _p(vel, _f32p),
                     _p(po, _f32p), _p(vo, _f32p))
    return po, vo


def cell_key(P: Params, x: float, y: float) -> int:
    """particles_per_cell.wgsl:14-27 (-1 if outside the grid)."""
    return int(lib().wo_cell_key(ctypes.byref(P), x, y))


def rebin(P: Params, pos_out, vel_out):
    """Count + scan + stable scatter (particles_per_cell.wgsl, prefix_sum.wgsl,
    pack_new_particle_data.wgsl with in-cell order by source index)."""
    pos_out = np.ascontiguousarray(pos_out, np.float32)
    vel_out = np.ascontiguousarray(vel_out, np.float32)
    C = P.grid_dimensions[0] * P.grid_dimensions[1]
    idx = np.zeros(C + 2, np.uint32)
    pi = np.zeros_like(pos_out)
    vi = np.zeros_like(vel_out)
    rc = lib().wo_rebin(ctypes.byref(P), _p(pos_out, _f32p), _p(vel_out, _f32p),
                        _p(idx, _u32p), _p(pi, _f32p), _p(vi, _f32p))
    if rc != 0:
        raise ValueError("a particle left the grid")
    return idx, pi, vi


def step(P: Params, idx, pos, vel, nsteps: int = 1, nthreads: int = 1):
    """nsteps frames of physics + stable re-bin; returns new (idx, pos, vel)."""
    idx = np.array(idx, np.uint32)
    pos = np.array(pos, np.float32)
    vel = np.array(vel, np.float32)
    n = P.particles_in_frame_count
    scratch = np.zeros(4 * max(n, 1), np.float32)
    L = lib()
    for _ in range(nsteps):
        if nthreads > 1:
            rc = L.wo_step_mt(ctypes.byref(P), _p(idx, _u32p), _p(pos, _f32p), _p(vel, _f32p),
                              _p(scratch, _f32p), nthreads)
        else:
            rc = L.wo_step(ctypes.byref(P), _p(idx, _u32p), _p(pos, _f32p), _p(vel, _f32p),
                           _p(scratch, _f32p))
        if rc != 0:
            raise ValueError("a particle left the grid")
    return idx, pos, vel


def stencil_pass(P: Params, idx, pcell, pin, move: bool = True):
    """One mode S separation pass (an extension; see wrach_oracle.h):
    returns (new positions or None, largest overlap 1 - d visited)."""
    idx = np.ascontiguousarray(idx, np.uint32)
    pcell = np.ascontiguousarray(pcell, np.float32)
    pin = np.ascontiguousarray(pin, np.float32)
    pout = np.zeros_like(pin)
    ov = lib().wo_stencil_pass(ctypes.byref(P), _p(idx, _u32p), _p(pcell, _f32p), _p(pin, _f32p),
                               _p(pout, _f32p), 1 if move else 0, None)
    return (pout if move else None), float(ov)


def stencil_overlap_sums(P: Params, idx, pcell, pin):
    """(sum of the overlaps 1 - d in 2^-24 units, overlapping pairs) over
    the pairs one pass visits, each pair counted from both particles."""
    idx = np.ascontiguousarray(idx, np.uint32)
    pcell = np.ascontiguousarray(pcell, np.float32)
    pin = np.ascontiguousarray(pin, np.float32)
    sums = np.zeros(2, np.uint64)
    lib().wo_stencil_pass(ctypes.byref(P), _p(idx, _u32p), _p(pcell, _f32p), _p(pin, _f32p),
                          None, 0, _p(sums, ctypes.POINTER(ctypes.c_uint64)))
    return int(sums[0]), int(sums[1])


def step_s(P: Params, idx, pos, vel, max_passes: int, overlap_tol: float = 0.0, nsteps: int = 1):
    """nsteps mode S frames: separation passes, then the reference step.
    Returns (idx, pos, vel, stats of the last frame: largest overlap before
    and after the passes, the passes run, and the overlap sums in 2^-24
    units and pair counts before and after)."""
    idx = np.array(idx, np.uint32)
    pos = np.array(pos, np.float32)
    vel = np.array(vel, np.float32)
    n = P.particles_in_frame_count
    scratch = np.zeros(4 * max(n, 1), np.float32)
    ov = np.zeros(2, np.float32)
    ran = np.zeros(1, np.uint32)
    sums = np.zeros(4, np.uint64)
    stats = None
    for _ in range(nsteps):
        rc = lib().wo_step_s(ctypes.byref(P), _p(idx, _u32p), _p(pos, _f32p), _p(vel, _f32p),
                             _p(scratch, _f32p), int(max_passes), float(overlap_tol),
                             _p(ov, _f32p), _p(ran, _u32p), _p(sums, ctypes.POINTER(ctypes.c_uint64)))
        if rc != 0:
            raise ValueError("a particle left the grid")
        stats = (float(ov[0]), float(ov[1]), int(ran[0]), tuple(int(x) for x in sums))
    return idx, pos, vel, stats
