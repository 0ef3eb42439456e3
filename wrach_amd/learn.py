"""Host side of the learned warm start (WRACH_SOLVER_L): the network's
initial parameters, training pairs from the plain solver's own rollouts,
and training on the device.  The kernels (features, the bf16-MFMA network,
its backward pass and Adam) are in csrc/learn.hip and csrc/kernels.hip.

No counterpart in the reference: its separation is one Gauss-Seidel sweep
per cell per frame (shaders/physics/src/particles.rs:62-93,
shaders/physics/src/cell.rs:69).  Here the plain separation passes of
WRACH_SOLVER_S are the teacher (the targets are their displacements), the
fallback (the learned frame returns to them when the warm start leaves more
overlap than it found) and the judge (their exact overlap sums measure
both).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .api import World

FEAT, HID, OUT = 64, 128, 2
N_PARAMS = _lib.WRACH_LEARN_PARAMS
SHAPES = (("W1", (FEAT, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, OUT)), ("b3", (OUT,)))


def init_params(seed: int) -> np.ndarray:
    """Weights uniform in +-1/sqrt(fan_in), biases zero (seeded PCG64)."""
    r = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            bound = 1.0 / np.sqrt(shape[0])
            parts.append(r.uniform(-bound, bound, shape).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(shape, np.float32).ravel())
    v = np.concatenate(parts)
    assert len(v) == N_PARAMS
    return v


def collect(world: World, sample_frames, state_passes: int, target_passes: int) -> int:
    """Steps `world` under WRACH_SOLVER_S with state_passes passes per frame
    and, at each frame count in sample_frames (ascending), appends every
    particle's features and its target_passes-pass displacement to the
    training set.  Returns the set's size."""
    world.set_solver("S", state_passes, quiet=True)
    done, total = 0, 0
    for f in sample_frames:
        if f > done:
            world.step(f - done)
            done = f
        total = world.learn_sample(target_passes)
    return total


def train(world: World, rows: int, steps: int, batch: int = 8192, lr: float = 1e-3,
          seed: int = 0) -> np.ndarray:
    """`steps` Adam steps of `batch` rows, the rows drawn epoch by epoch
    from seeded permutations of the training set.  Returns the losses."""
    r = np.random.Generator(np.random.PCG64(seed))
    per_epoch = rows // batch
    losses = []
    left = steps
    while left > 0:
        k = min(left, per_epoch)
        order = r.permutation(rows)[:k * batch].astype(np.uint32)
        losses.append(world.learn_train(order, batch, lr))
        left -= k
    return np.concatenate(losses) if losses else np.zeros(0, np.float32)
