"""The learned warm start on the GPU against its CPU oracle (oracle/learn.py,
wo_learn_features), through the C ABI.

Bars:
- features and their overlap sums: bit-exact (integer/selection work and
  f32 differences rounded to bf16);
- the network's output and gradient: the device sums products on MFMA in
  another order than numpy, and a hidden activation or dL/dz that lands
  next to a bf16 rounding boundary can round the other way, so the outputs
  are compared to 1e-3 of their scale (max abs error <= 1e-3 * (1 + max|y|))
  and each gradient block to 1e-2 of its norm (L2), with cosine > 0.9999;
- Adam: bit-exact (elementwise f32 in the same operation order);
- the learned frame: bit-exact against the oracle's passes wherever the
  warm start is exact (a zero network, a constant one) or rejected (the
  fallback), and a trained network must need fewer passes than an untrained
  one to reach the plain solver's mean overlap.
"""
import sys

import numpy as np
import pytest

from oracle import learn as ol
from wrach_amd import World, WrachConfig, WrachError, workloads
from wrach_amd import _lib
from wrach_amd import learn as wl

from .test_gpu_parity import assert_same_state, oracle_world

pytestmark = pytest.mark.gpu

BLOCKS = (("W1", 0, 8192), ("b1", 8192, 8320), ("W2", 8320, 24704), ("b2", 24704, 24832),
          ("W3", 24832, 25088), ("b3", 25088, 25090))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def lattice(W=256, H=128, seed=1):
    return workloads.lattice_1m(seed=seed, n=W * H // 2, width=W, height=H)


def test_features_match_the_oracle(oracle):
    """A lattice after 6 mode S frames and a crowded uniform world, cell
    sizes 1 and 3: rows and sums bit for bit."""
    cases = []
    lw = lattice()
    cases.append(("lattice", lw.config, lw.pos, lw.vel, 6))
    r = np.random.Generator(np.random.PCG64(21))
    n = 8000  # about 19 per 3x3 cell in the lower 40 rows; capacity 10,791
    pos = np.stack([r.uniform(0, 96, n), r.uniform(0, 40, n)], 1).astype(np.float32)
    vel = r.uniform(-1, 1, (n, 2)).astype(np.float32)
    cases.append(("crowded cs 3", WrachConfig(dimensions=(96, 96), cell_size=3, gravity=(0.0, -0.1)),
                  pos, vel, 2))
    # Exact ties: coincident copies, half with their original's velocity.
    n = 3000
    pos = np.stack([r.uniform(0, 64, n), r.uniform(0, 32, n)], 1).astype(np.float32)
    vel = r.uniform(-1, 1, (n, 2)).astype(np.float32)
    pos[-400:] = pos[:400]
    vel[-200:] = vel[:200]
    cases.append(("coincident copies", WrachConfig(dimensions=(64, 64), cell_size=1), pos, vel, 0))
    for name, cfg, pos, vel, frames in cases:
        say(f"[learn] features: {name}")
        with World(cfg) as w:
            w.add_particles(pos, vel)
            w.set_solver("S", 2)
            w.step(frames)
            rows, sums = w.learn_features()
            idx, p, v = w.read()
        W, H = cfg.dimensions
        P = oracle.params(W, H, cfg.cell_size, len(p), gravity=cfg.gravity)
        ref_rows, ref_sums = ol.features(P, idx, p, v)
        bad = np.flatnonzero((rows != ref_rows).any(axis=1))
        assert bad.size == 0, f"{name}: {bad.size} rows differ, first {bad[:5]}"
        assert sums == ref_sums, name
        assert (rows != 0).any()


def _rand_rows(seed, n):
    r = np.random.Generator(np.random.PCG64(seed))
    return ol.bf16_bits(r.normal(0, 0.6, (n, 64)).astype(np.float32))


def _params(seed):
    r = np.random.Generator(np.random.PCG64(seed + 100))
    p = wl.init_params(seed)
    p[8192:8320] = r.normal(0, 0.1, 128)
    p[24704:24832] = r.normal(0, 0.1, 128)
    p[25088:] = r.normal(0, 0.1, 2)
    return p


@pytest.mark.parametrize("n", [37, 4096, 100_003])
def test_forward_matches_the_oracle(n):
    params = _params(1)
    rows = _rand_rows(2, n)
    say(f"[learn] forward, {n} rows")
    with World(WrachConfig(dimensions=(16, 16), cell_size=1)) as w:
        w.learn_set_params(params)
        np.testing.assert_array_equal(w.learn_params(), params)
        y = w.learn_forward(rows)
    ref, _ = ol.forward(params, rows)
    err = float(np.max(np.abs(y - ref)))
    scale = 1.0 + float(np.max(np.abs(ref)))
    say(f"[learn] forward max abs err {err:.3e} (scale {scale:.3f})")
    assert np.isfinite(y).all()
    assert err <= 1e-3 * scale


@pytest.mark.parametrize("n", [64, 8192])
def test_gradient_matches_the_oracle(n):
    params = _params(3)
    rows = _rand_rows(4, n)
    t = np.random.Generator(np.random.PCG64(5)).normal(0, 0.3, (n, 2)).astype(np.float32)
    say(f"[learn] gradient, batch {n}")
    with World(WrachConfig(dimensions=(16, 16), cell_size=1)) as w:
        w.learn_set_params(params)
        g, loss = w.learn_gradient(rows, t)
        np.testing.assert_array_equal(w.learn_params(), params)  # unchanged
    ref, ref_loss = ol.gradient(params, rows, t)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    bad = []
    for name, a, b in BLOCKS:
        d = np.linalg.norm(g[a:b] - ref[a:b])
        nrm = np.linalg.norm(ref[a:b])
        cos = float(g[a:b] @ ref[a:b] / (np.linalg.norm(g[a:b]) * nrm))
        say(f"[learn] {name}: rel L2 err {d / nrm:.2e}, cos {cos:.6f}")
        if not (d <= 1e-2 * nrm and cos > 0.9999):
            bad.append(name)
    assert not bad, bad


def test_adam_is_bit_exact():
    params = _params(6)
    r = np.random.Generator(np.random.PCG64(7))
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    with World(WrachConfig(dimensions=(16, 16), cell_size=1)) as w:
        w.learn_set_params(params)
        for step in range(1, 5):
            g = (r.normal(0, 1, len(params)) * 10.0 ** (-step)).astype(np.float32)
            w.learn_adam(g, 1e-3)
            params, m, v = ol.adam(params, m, v, g, step, 1e-3)
            np.testing.assert_array_equal(w.learn_params().view(np.uint32), params.view(np.uint32))


def _oracle_learned_frame(oracle, P, ref, warm, passes):
    """The learned frame's passes from given warm-start positions, then the
    reference step; returns (state, fell_back)."""
    idx, p, v = ref
    start = oracle.stencil_overlap_sums(P, idx, p, p)
    wsum = oracle.stencil_overlap_sums(P, idx, p, warm)
    cur, _ = oracle.stencil_pass(P, idx, p, warm)
    fell = False
    if passes >= 2:
        fell = wsum[0] * start[1] > start[0] * wsum[1]
        cur, _ = oracle.stencil_pass(P, idx, p, p if fell else cur)
        for _ in range(passes - 2):
            cur, _ = oracle.stencil_pass(P, idx, p, cur)
    return oracle.step(P, idx, cur, v), fell


@pytest.mark.parametrize("passes", [1, 2, 3, 4, 5])
def test_learned_frame_with_zero_network_is_mode_s(oracle, passes):
    """A network of zeros warm-starts nowhere: the learned frame is mode S
    with the same passes, bit for bit, and never falls back."""
    lw = lattice(128, 64)
    say(f"[learn] zero network, {passes} passes")
    P, ref = oracle_world(oracle, lw.config, lw.pos, lw.vel)
    with World(lw.config) as w:
        w.add_particles(lw.pos, lw.vel)
        w.learn_set_params(np.zeros(_lib.WRACH_LEARN_PARAMS, np.float32))
        w.set_solver("L", passes)
        for f in range(3):
            w.step(1)
            ref = oracle.step_s(P, *ref, passes)[:3]
            assert_same_state(w.read(), ref, f"zero network, frame {f + 1}")
            assert not w.learned_stats()["fell_back"]


@pytest.mark.parametrize("passes", [1, 3])
def test_learned_frame_with_constant_network(oracle, passes):
    """W3 = 0 and b3 = (0.25, -0.125): the warm start is the frame start
    shifted by b3 exactly; the passes and the fallback choice match the
    oracle bit for bit."""
    lw = lattice(128, 64, seed=4)
    params = wl.init_params(9)
    params[24832:25088] = 0.0
    params[25088:] = [0.25, -0.125]
    say(f"[learn] constant network, {passes} passes")
    P, ref = oracle_world(oracle, lw.config, lw.pos, lw.vel)
    with World(lw.config) as w:
        w.add_particles(lw.pos, lw.vel)
        w.learn_set_params(params)
        w.set_solver("L", passes)
        for f in range(3):
            w.step(1)
            warm = ref[1] + np.array([0.25, -0.125], np.float32)
            ref, fell = _oracle_learned_frame(oracle, P, ref, warm, passes)
            assert_same_state(w.read(), ref, f"constant network, frame {f + 1}")
            assert w.learned_stats()["fell_back"] == fell


def _onto_nearest_network():
    """A network that moves each particle onto its nearest neighbour: hidden
    units 0 and 1 carry the nearest's (dx, dy) shifted by 20 (where SiLU is
    the identity to 1e-8), and the output removes the shift; the bf16
    rounding of the hidden values (1/16 at 20) leaves the particle within
    1/16 of that neighbour."""
    P = {k: np.zeros(shape, np.float32) for k, shape in wl.SHAPES}
    P["W1"][0, 0] = P["W1"][1, 1] = 1.0
    P["b1"][:2] = 20.0
    P["W2"][0, 0] = P["W2"][1, 1] = 1.0
    P["W3"][0, 0] = P["W3"][1, 1] = 1.0
    P["b3"][:] = -20.0
    return np.concatenate([P[k].ravel() for k, _ in wl.SHAPES])


def test_learned_frame_falls_back_on_a_bad_warm_start(oracle):
    """A network that moves every particle onto its nearest neighbour leaves
    far more mean overlap than the frame start had: the frame falls back,
    and its positions are those of passes 2..k from the frame start (k - 1
    plain passes)."""
    lw = lattice(128, 64, seed=5)
    params = _onto_nearest_network()
    k = 4
    say("[learn] fallback")
    P, ref = oracle_world(oracle, lw.config, lw.pos, lw.vel)
    with World(lw.config) as w:
        w.add_particles(lw.pos, lw.vel)
        w.step(2)
        ref = oracle.step(P, *ref, nsteps=2)
        w.learn_set_params(params)
        w.set_solver("L", k)
        w.step(1)
        st = w.learned_stats()
        assert st["fell_back"]
        assert st["warm_overlap_sum"] * st["start_pairs"] > st["start_overlap_sum"] * st["warm_pairs"]
        ref = oracle.step_s(P, *ref, k - 1)[:3]
        assert_same_state(w.read(), ref, "after the fallback")


def test_learned_solver_argument_errors():
    with World(WrachConfig(dimensions=(32, 32), cell_size=1)) as w:
        with pytest.raises(WrachError) as e:
            w.set_solver("L", 2)  # no parameters yet
        assert e.value.code == _lib.WRACH_E_STATE
        w.learn_set_params(np.zeros(_lib.WRACH_LEARN_PARAMS, np.float32))
        with pytest.raises(WrachError) as e:
            w.set_solver("L", 2, overlap_tol=0.1)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError) as e:
            w.learn_set_params(np.full(_lib.WRACH_LEARN_PARAMS, np.nan, np.float32))
        assert e.value.code == _lib.WRACH_E_NONFINITE
        with pytest.raises(WrachError):
            w.learn_gradient(np.zeros((100, 64), np.uint16), np.zeros((100, 2), np.float32))
        with pytest.raises(WrachError):
            w.learn_train(np.array([0], np.uint32), 64, 1e-3)  # no samples


def _mean_overlap(world, passes, mode, params=None):
    """Mean overlap after `passes` passes of `mode` on one frame from the
    world's current state (the state itself is restored afterwards)."""
    idx, p, v = world.read()
    if params is not None:
        world.learn_set_params(params)
    world.set_solver(mode, passes)
    world.step(1)
    st = world.solver_stats()
    sa, na = st[3][2], st[3][3]
    world.load_packed(idx, p, v)
    return sa / 2**24 / na


def test_training_beats_the_untrained_network(oracle):
    """On a 256x128 lattice: training pairs from mode S rollouts (2 passes
    per frame, targets after 8 passes), 400 Adam steps of 4096 rows on the
    device.  The loss falls, and the trained warm start reaches the mean
    overlap of 8 plain passes in fewer passes than the untrained one."""
    lw = lattice(256, 128, seed=6)
    say("[learn] train on a 256x128 lattice")
    with World(lw.config) as w:
        w.add_particles(lw.pos, lw.vel)
        rows = wl.collect(w, [3, 5, 7, 9, 11, 13], state_passes=2, target_passes=8)
        assert rows == 6 * lw.n
        untrained = wl.init_params(11)
        w.learn_set_params(untrained)
        losses = wl.train(w, rows, steps=400, batch=4096, lr=1e-3, seed=1)
        say(f"[learn] loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        assert losses[-1] < 0.7 * losses[0]
        trained = w.learn_params()
        w.set_solver("S", 2, quiet=True)
        w.step(2)
        target = _mean_overlap(w, 8, "S")
        need = {}
        for name, params in (("trained", trained), ("untrained", untrained)):
            need[name] = None
            for k in range(1, 9):
                if _mean_overlap(w, k, "L", params) <= target:
                    need[name] = k
                    break
        say(f"[learn] passes to reach 8 plain passes' mean overlap {target:.4f}: {need}")
        assert need["trained"] is not None
        assert need["untrained"] is None or need["trained"] < need["untrained"]
