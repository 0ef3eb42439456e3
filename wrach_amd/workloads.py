"""Seeded synthetic inputs for the particle-step workloads.

All generators use numpy's PCG64 bit generator with an explicit seed, draw
in float64 and round once to float32.  Every workload uses particle radius
0.5 (MIN_DISTANCE = 1, shaders/physics/src/particles.rs:18), cell size 1,
gravity (0, -0.1) px/step^2, dt = 1 and walls at [0, W] x [0, H].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import WrachConfig

GRAVITY = (0.0, -0.1)


@dataclass
class Workload:
    name: str
    config: WrachConfig
    pos: np.ndarray      # (N, 2) float32
    vel: np.ndarray      # (N, 2) float32
    steps: int           # frames the workload is defined over

    @property
    def n(self) -> int:
        return len(self.pos)

    @property
    def cells(self) -> int:
        w, h = self.config.dimensions
        cs = self.config.cell_size
        return (w // cs + 1) * (h // cs + 1)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uniform_lower_half(n: int, width: int, height: int, seed: int, gravity=GRAVITY,
                       name: str = "uniform") -> Workload:
    """N particles uniform in [0, W) x [0, H/2), velocities uniform in [-1, 1]^2."""
    r = _rng(seed)
    pos = np.empty((n, 2), np.float64)
    pos[:, 0] = r.uniform(0.0, width, n)
    pos[:, 1] = r.uniform(0.0, height / 2.0, n)
    vel = r.uniform(-1.0, 1.0, (n, 2))
    cfg = WrachConfig(dimensions=(width, height), cell_size=1, gravity=gravity)
    return Workload(name, cfg, pos.astype(np.float32), vel.astype(np.float32), 100)


def small_uniform(seed: int = 0, gravity=GRAVITY) -> Workload:
    """16,384 particles in a 512x512 world, lower half; 100 frames."""
    return uniform_lower_half(16_384, 512, 512, seed, gravity, "uniform-16k-512x512")


def lattice_1m(seed: int = 1, n: int = 1_000_000, width: int = 2048, height: int = 1024,
               gravity=GRAVITY) -> Workload:
    """1,000,000 particles at the pixel centres of the lower half of a
    2048x1024 world (row-major, first N), jitter U(-0.1, 0.1), velocities
    N(0, 0.5); 1,000 frames."""
    r = _rng(seed)
    i = np.arange(n)
    x = (i % width) + 0.5 + r.uniform(-0.1, 0.1, n)
    y = (i // width) + 0.5 + r.uniform(-0.1, 0.1, n)
    vel = r.normal(0.0, 0.5, (n, 2))
    cfg = WrachConfig(dimensions=(width, height), cell_size=1, gravity=gravity)
    name = "lattice-1m-2048x1024" if (n, width, height) == (1_000_000, 2048, 1024) else \
        f"lattice-{n}-{width}x{height}"
    return Workload(name, cfg, np.stack([x, y], 1).astype(np.float32), vel.astype(np.float32), 1000)


def gaussian_blobs(seed: int = 2, n: int = 4_000_000, side: int = 4096, blobs: int = 64,
                   gravity=GRAVITY) -> Workload:
    """4,000,000 particles from 64 Gaussian blobs (centres uniform in the
    world, sigma uniform in [8, 128] px), clamped to the world; velocities
    N(0, 1); 500 frames."""
    r = _rng(seed)
    centres = r.uniform(0.0, side, (blobs, 2))
    sigma = r.uniform(8.0, 128.0, blobs)
    which = r.integers(0, blobs, n)
    pos = centres[which] + r.normal(0.0, 1.0, (n, 2)) * sigma[which, None]
    pos = np.clip(pos, 0.0, float(side))
    vel = r.normal(0.0, 1.0, (n, 2))
    cfg = WrachConfig(dimensions=(side, side), cell_size=1, gravity=gravity)
    name = "blobs-4m-4096x4096" if (n, side, blobs) == (4_000_000, 4096, 64) else \
        f"blobs-{n}-{side}x{side}"
    return Workload(name, cfg, pos.astype(np.float32), vel.astype(np.float32), 500)


def many_worlds(seed: int = 3, worlds: int = 512, gravity=GRAVITY) -> list[Workload]:
    """512 independent worlds: side uniform from {256, 512, 1024}, particle
    count log-uniform in [1,000, 65,536], particles as in small_uniform."""
    r = _rng(seed)
    sides = r.choice([256, 512, 1024], worlds)
    counts = np.exp(r.uniform(np.log(1000.0), np.log(65536.0), worlds)).astype(np.int64)
    return [uniform_lower_half(int(c), int(s), int(s), seed * 100_003 + w, gravity,
                               f"world{w}-{int(c)}-{int(s)}")
            for w, (s, c) in enumerate(zip(sides, counts))]


BY_NAME = {
    "uniform-16k-512x512": small_uniform,
    "lattice-1m-2048x1024": lattice_1m,
    "blobs-4m-4096x4096": gaussian_blobs,
}
