"""Pins the CPU oracle (oracle/wrach_oracle.c) before it is trusted.

Each function is checked against (a) the reference's own test vectors
(tests/golden/reference_vectors.json), (b) closed forms and special cases
the mathematics fixes, and (c) independent brute force (tests/pystep.py:
numpy float32 scalars that round every operation, Python's stable sort).
"""
import json
import math
import os

import numpy as np
import pytest

from tests import pystep

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
F = np.float32


def f32(a):
    return np.asarray(a, np.float32)


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


# ----------------------------------------------------------------- collision

def test_collision_golden(oracle):
    g = GOLD["collision"][0]
    out = oracle.collide_cell(g["positions"])
    assert np.array_equal(bits(out), bits(g["expected_positions"]))
    d = float(np.hypot(*(out[0] - out[1])))
    lo, hi = g["new_distance_open_interval"]
    assert lo < d < hi


def test_collision_matches_strict_f32_evaluation(oracle):
    """A plausible mistake the golden vector cannot see (dx == dy there):
    FMA contraction.  numpy float32 rounds every operation."""
    r = np.random.default_rng(5)
    for _ in range(2000):
        a = r.uniform(0, 64, 2)
        b = a + r.uniform(-1.2, 1.2, 2)
        pts = f32([a, b])
        assert np.array_equal(bits(oracle.collide_cell(pts)), bits(pystep.collide(pts)))


def test_collision_separated_pair_closed_form(oracle):
    """Two particles at distance d <= 1 end 1 apart, with their midpoint kept."""
    r = np.random.default_rng(6)
    for _ in range(500):
        a = r.uniform(10, 50, 2)
        ang = r.uniform(0, 2 * math.pi)
        d = r.uniform(0.05, 1.0)
        b = a + d * np.array([math.cos(ang), math.sin(ang)])
        pts = f32([a, b])
        out = oracle.collide_cell(pts).astype(np.float64)
        d_old = float(np.hypot(*(pts[1].astype(np.float64) - pts[0])))
        if d_old > 1.0:
            continue
        assert abs(np.hypot(*(out[1] - out[0])) - 1.0) < 2e-5
        assert np.allclose(out.mean(0), pts.astype(np.float64).mean(0), atol=2e-5)
        # pushed apart along the joining line
        assert np.dot(out[1] - out[0], pts[1].astype(np.float64) - pts[0]) > 0


def test_collision_special_distances(oracle):
    # d > 1: untouched
    p = f32([[3.0, 3.0], [4.5, 3.0]])
    assert np.array_equal(bits(oracle.collide_cell(p)), bits(p))
    # d == 1 exactly: force 0, untouched
    p = f32([[3.0, 3.0], [4.0, 3.0]])
    assert np.array_equal(bits(oracle.collide_cell(p)), bits(p))
    # coincident: d := 0.0001 but the difference vector is 0 -> untouched
    p = f32([[2.5, 2.5], [2.5, 2.5]])
    assert np.array_equal(bits(oracle.collide_cell(p)), bits(p))


def test_collision_is_in_place_gauss_seidel_in_index_order(oracle):
    """m particles = the 2-particle primitive applied to pairs (i<j) in
    lexicographic order, each on the already-updated positions.  A Jacobi
    update or another pair order fails this."""
    r = np.random.default_rng(7)
    for m in range(2, 10):
        for _ in range(50):
            pts = f32(r.uniform(5, 5.9, (m, 2)))
            want = pts.copy()
            for i in range(m):
                for j in range(i + 1, m):
                    want[[i, j]] = oracle.collide_cell(want[[i, j]])
            assert np.array_equal(bits(oracle.collide_cell(pts)), bits(want))
    # the order matters for these inputs (so the check above has teeth)
    pts = f32([[5.0, 5.0], [5.3, 5.0], [5.6, 5.1]])
    rev = pts.copy()
    for i, j in [(1, 2), (0, 2), (0, 1)]:
        rev[[i, j]] = oracle.collide_cell(rev[[i, j]])
    assert not np.array_equal(bits(oracle.collide_cell(pts)), bits(rev))


# ------------------------------------------------------------ finish/limits

@pytest.mark.parametrize("pos,vel,want_pos,want_vel", [
    ((5.0, 5.0), (0.5, 0.5), (5.5, 5.5), (0.5, 0.5)),          # plain integrate
    ((9.5, 5.0), (1.0, 0.0), (10.0, 5.0), (-1.0, 0.0)),        # right wall: clamp + reflect
    ((0.2, 0.3), (-0.5, -0.5), (0.0, 0.0), (0.5, 0.5)),         # left/bottom walls
    ((10.0, 10.0), (0.0, 0.0), (10.0, 10.0), (0.0, 0.0)),       # exactly on the wall: untouched
    ((5.0, 5.0), (2.0, -3.0), (7.0, 2.0), (1.0, -1.0)),         # integrate with v, then clamp v
    ((9.0, 9.0), (3.0, 3.0), (10.0, 10.0), (-1.0, -1.0)),       # reflect (-3) then clamp
    ((10.5, 5.0), (0.0, 0.0), (10.0, 5.0), (-0.0, 0.0)),        # outside, v=0 -> v = -0.0
])
def test_finish_closed_forms(oracle, pos, vel, want_pos, want_vel):
    P = oracle.params(10, 10, 1, 0)
    p, v = oracle.finish_particle(P, pos, vel)
    assert np.array_equal(bits(p), bits(want_pos))
    assert np.array_equal(bits(v), bits(want_vel))


def test_finish_gravity_extension(oracle):
    # g = 0 leaves a -0.0 velocity alone (no +0.0 add)
    P = oracle.params(10, 10, 1, 0, gravity=(0.0, -0.1))
    p, v = oracle.finish_particle(P, (5.0, 5.0), (-0.0, 0.5))
    assert bits(v)[0] == 0x80000000
    assert v[1] == F(F(0.5) + F(-0.1))
    assert p[1] == F(F(5.0) + v[1])
    # matches strict float32 evaluation for random inputs
    r = np.random.default_rng(8)
    for _ in range(500):
        pos = r.uniform(-2, 12, 2)
        vel = r.uniform(-2, 2, 2)
        g = (0.0, float(r.uniform(-0.3, 0.3)))
        P = oracle.params(10, 10, 1, 0, gravity=g)
        p, v = oracle.finish_particle(P, pos, vel)
        wp, wv = pystep.finish(f32(pos), f32(vel), 10, 10, g)
        assert np.array_equal(bits(p), bits(wp)) and np.array_equal(bits(v), bits(wv))


# ---------------------------------------------------------- binning/packing

@pytest.mark.parametrize("case", GOLD["cell_coord"], ids=lambda c: c["source"].split(":")[-1])
def test_cell_coord_golden(oracle, case):
    assert oracle.cell_coord(*case["position"], case["cell_size"]) == tuple(case["expected"])


@pytest.mark.parametrize("case", GOLD["active_cells"], ids=lambda c: c["source"].split(":")[-1])
def test_active_cells_golden(oracle, case):
    (bx, by), (gx, gy) = oracle.active_grid(case["viewport"], case["cell_size"])
    cells = [[bx + x, by + y] for y in range(gy) for x in range(gx)]
    assert cells == case["expected_cells"]
    if "expected_grid" in case:
        assert [gx, gy] == case["expected_grid"]


@pytest.mark.parametrize("case", GOLD["pack"], ids=lambda c: c["source"].split(" ")[-1])
def test_pack_golden(oracle, case):
    idx, pos, vel = oracle.pack(case["positions"], case["velocities"], case["viewport"], case["cell_size"])
    assert idx.tolist() == case["expected_indices"]
    assert np.array_equal(bits(pos), bits(case["expected_positions"]))
    assert np.array_equal(bits(vel), bits(case["expected_velocities"]))


@pytest.mark.parametrize("case", GOLD["capacity"], ids=lambda c: str(c["expected"]))
def test_capacity_golden(oracle, case):
    _, (gx, gy) = oracle.active_grid(case["viewport"], case["cell_size"])
    assert oracle.max_particles(gx * gy, case["cell_size"]) == case["expected"]
    # closed form: C*cs^2 + 10*ceil(C*cs^2/100)
    normal = gx * gy * case["cell_size"] ** 2
    assert case["expected"] == normal + 10 * -(-normal // 100)


def test_cell_coord_div_euclid_cases(oracle):
    # floor semantics for negatives, exact multiples, tiny negatives
    assert oracle.cell_coord(-0.0001, 5.0, 3) == (-1, 1)
    assert oracle.cell_coord(-0.0, 6.0, 3) == (0, 2)
    assert oracle.cell_coord(2.9999998, 3.0, 3) == (0, 1)
    assert oracle.cell_coord(float("inf"), float("-inf"), 3) == (2**31 - 1, -2**31)


def test_pack_brute_force(oracle):
    r = np.random.default_rng(9)
    for cs in (1, 2, 3, 5):
        for _ in range(20):
            n = int(r.integers(0, 60))
            pos = f32(r.uniform(-3, 14, (n, 2)))
            vel = f32(r.uniform(-1, 1, (n, 2)))
            idx, p, v = oracle.pack(pos, vel, (0, 0, 12, 9), cs)
            widx, wp, wv = pystep.pack(pos, vel, 12, 9, cs)
            assert idx.tolist() == widx.tolist()
            assert np.array_equal(bits(p), bits(wp)) and np.array_equal(bits(v), bits(wv))


def test_rebin_brute_force_and_invariants(oracle):
    r = np.random.default_rng(10)
    for cs in (1, 3):
        P = oracle.params(16, 12, cs, 0)
        gx, gy = P.grid_dimensions
        for _ in range(20):
            n = int(r.integers(1, 80))
            po = f32(r.uniform(0, [16, 12], (n, 2)))
            po[: n // 4] = np.round(po[: n // 4])  # exact cell boundaries
            vo = f32(r.uniform(-1, 1, (n, 2)))
            P.particles_in_frame_count = n
            idx, pi, vi = oracle.rebin(P, po, vo)
            keys = [oracle.cell_key(P, x, y) for x, y in po]
            order = sorted(range(n), key=lambda i: keys[i])  # stable
            assert np.array_equal(bits(pi), bits(po[order]))
            assert np.array_equal(bits(vi), bits(vo[order]))
            assert idx[0] == 0 and idx[1] == 0 and idx[-1] == n and len(idx) == gx * gy + 2
            assert np.all(np.diff(idx.astype(np.int64)) >= 0)
            for c in range(gx * gy):
                for i in range(idx[c + 1], idx[c + 2]):
                    assert oracle.cell_key(P, *pi[i]) == c


def _rand_packed(oracle, r, w, h, cs, n, spread=1.0):
    pos = f32(r.uniform(0, [w, h], (n, 2)))
    if spread < 1.0:  # crowd particles to make full cells
        pos = f32(np.minimum(pos * spread, [w, h]))
    vel = f32(r.uniform(-1.5, 1.5, (n, 2)))
    return oracle.pack(pos, vel, (0, 0, w, h), cs)


@pytest.mark.parametrize("cs", [1, 2, 3])
def test_step_brute_force_tiny_worlds(oracle, cs):
    """Whole frame: per-cell ranges, cap of 9, overflow path, walls, stable
    re-bin -- against the pure-Python step."""
    r = np.random.default_rng(11 + cs)
    for trial in range(6):
        w, h = 9, 7
        n = int(r.integers(1, 70))
        spread = 0.3 if trial % 2 else 1.0
        idx, pos, vel = _rand_packed(oracle, r, w, h, cs, n, spread)
        g = (0.0, -0.1) if trial % 3 == 0 else (0.0, 0.0)
        P = oracle.params(w, h, cs, len(pos), gravity=g)
        for _ in range(3):
            widx, wpos, wvel = pystep.step(idx, pos, vel, w, h, cs, gravity=g)
            idx, pos, vel = oracle.step(P, idx, pos, vel)
            assert idx.tolist() == widx.tolist()
            assert np.array_equal(bits(pos), bits(wpos))
            assert np.array_equal(bits(vel), bits(wvel))


def test_overfull_cell_only_first_nine_collide(oracle):
    """cell.rs:29-30,79-95: 12 coincident-free particles in one cell; ranks
    9..11 are integrated without collisions."""
    pts = f32([[6.1 + 0.07 * i, 6.2 + 0.05 * i] for i in range(12)])
    vel = f32(np.zeros((12, 2)))  # p + 0 == p and no wall is reached
    idx, pos, v = oracle.pack(pts, vel, (0, 0, 15, 15), 3)
    P = oracle.params(15, 15, 3, 12)
    po, vo = oracle.physics(P, idx, pos, v)
    assert np.array_equal(bits(po[9:]), bits(pos[9:]))
    assert np.array_equal(bits(po[:9]), bits(oracle.collide_cell(pos[:9])))


@pytest.mark.parametrize("case", GOLD["gpu_frames"], ids=lambda c: c["source"].split(" ")[-1])
def test_reference_frame_tests_on_oracle(oracle, case):
    """The reference's GPU tests, run through the oracle: pack, then ticks."""
    w, h = case["dimensions"]
    cs = case["cell_size"]
    idx, pos, vel = oracle.pack(case["positions"], case["velocities"], (0, 0, w, h), cs)
    cpu_idx = idx.copy()
    P = oracle.params(w, h, cs, len(pos))
    idx, pos, vel = oracle.step(P, idx, pos, vel, nsteps=case["ticks"])
    if "expected_indices" in case:
        assert idx.tolist() == case["expected_indices"]
    if case.get("indices_equal_cpu_pack"):
        assert idx.tolist() == cpu_idx.tolist()
    if "expected_positions" in case:
        assert np.array_equal(bits(pos), bits(case["expected_positions"]))
    if "expected_readback_length" in case:
        _, (gx, gy) = oracle.active_grid((0, 0, w, h), cs)
        assert oracle.max_particles(gx * gy, cs) == case["expected_readback_length"]
        assert tuple(pos[0]) != (0.0, 0.0) and tuple(vel[0]) != (0.0, 0.0)


@pytest.mark.parametrize("nthreads", [2, 3, 8])
def test_multithreaded_step_equals_serial(oracle, nthreads):
    r = np.random.default_rng(12)
    for cs, spread in ((1, 1.0), (1, 0.2), (3, 0.5)):
        idx, pos, vel = _rand_packed(oracle, r, 64, 48, cs, 1500, spread)
        P = oracle.params(64, 48, cs, len(pos), gravity=(0.0, -0.1))
        a = oracle.step(P, idx, pos, vel, nsteps=4)
        b = oracle.step(P, idx, pos, vel, nsteps=4, nthreads=nthreads)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


def test_regression_fixture(oracle):
    """Committed rollout of a small world (tests/golden/make_fixtures.py)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "rollout_small.npz")
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    P = oracle.params(meta["width"], meta["height"], meta["cell_size"], int(z["pos0"].shape[0]),
                      gravity=tuple(meta["gravity"]))
    idx, pos, vel = z["idx0"], z["pos0"], z["vel0"]
    for k in range(meta["steps"]):
        idx, pos, vel = oracle.step(P, idx, pos, vel)
        assert np.array_equal(idx, z["idx"][k])
        assert np.array_equal(bits(pos), bits(z["pos"][k]))
        assert np.array_equal(bits(vel), bits(z["vel"][k]))


# -------------------------------------------------- mode S (an extension)

def _packed_world(oracle, n, side, cs, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    pos = f32(rng.uniform(0, side * spread, (n, 2)))
    vel = f32(rng.uniform(-1, 1, (n, 2)))
    idx, p, v = oracle.pack(pos, vel, (0, 0, side, side), float(cs))
    P = oracle.params(side, side, cs, len(p))
    return P, idx, p, v


def _brute_force_pass(pin):
    """All-pairs Jacobi pass in numpy: the pair rule in float32 operation by
    operation, the mean of a particle's terms in float64."""
    x, y = pin[:, 0], pin[:, 1]
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    sq = (dx * dx).astype(F) + (dy * dy).astype(F)
    d = np.sqrt(sq.astype(F)).astype(F)
    near = d <= F(1)
    np.fill_diagonal(near, False)
    d = np.where(d == 0, F(0.0001), d).astype(F)
    force = ((F(0.5) * (F(1) - d)).astype(F) / d).astype(F)
    ex = ((x[None, :] - x[:, None]).astype(F) * force).astype(F)
    ey = ((y[None, :] - y[:, None]).astype(F) * force).astype(F)
    disp = np.stack([-(np.where(near, ex, 0).astype(np.float64)).sum(1),
                     -(np.where(near, ey, 0).astype(np.float64)).sum(1)], 1)
    disp /= np.maximum(near.sum(1), 1)[:, None]  # the mean over the overlapping neighbours
    ov = np.where(near, F(1) - d, 0).max() if near.any() else 0.0
    return pin.astype(np.float64) + disp, float(ov)


def _brute_force_overlap_sums(pin):
    """Sum of 1 - d in 2^-24 units (float32 product, round half to even)
    and count, over ordered pairs i != j with d <= 1 (d == 0 as 0.0001)."""
    x, y = pin[:, 0], pin[:, 1]
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    d = np.sqrt((dx * dx).astype(F) + (dy * dy).astype(F)).astype(F)
    near = d <= F(1)
    np.fill_diagonal(near, False)
    d = np.where(d == 0, F(0.0001), d).astype(F)
    units = np.rint(((F(1) - d).astype(F) * F(16777216.0)).astype(F)).astype(np.int64)
    return int(units[near].sum()), int(near.sum())


def test_stencil_pair_separates_to_one(oracle):
    """Two particles 0.6 apart in neighbouring cells: each moves 0.2 away
    from the other, so the pair ends exactly 1 apart, midpoint kept."""
    idx, p, _ = oracle.pack(f32([[2.5, 2.5], [3.1, 2.5]]), f32([[0, 0], [0, 0]]), (0, 0, 8, 8), 1.0)
    P = oracle.params(8, 8, 1, 2)
    out, ov = oracle.stencil_pass(P, idx, p, p)
    assert out[1, 0] - out[0, 0] == F(1.0)
    assert abs((out[0, 0] + out[1, 0]) / 2 - 2.8) < 1e-6 and np.all(out[:, 1] == F(2.5))
    dx = F(2.5) - F(3.1)
    assert ov == float(F(1) - np.sqrt(dx * dx + F(0) * F(0)))


@pytest.mark.parametrize("cs", [1, 3])
@pytest.mark.parametrize("seed", [0, 1])
def test_stencil_pass_equals_all_pairs_brute_force(oracle, cs, seed):
    """With cell sides >= 1 every pair closer than 1 shares the 3x3 cells,
    so the pass must equal the all-pairs Jacobi pass: positions to the
    fixed-point rounding (2^-33 per term), the overlap exactly."""
    P, idx, p, _ = _packed_world(oracle, 400, 24, cs, seed)
    out, ov = oracle.stencil_pass(P, idx, p, p)
    ref, ref_ov = _brute_force_pass(p)
    assert np.abs(out - ref).max() < 2e-6
    assert ov == ref_ov
    assert oracle.stencil_overlap_sums(P, idx, p, p) == _brute_force_overlap_sums(p)
    assert np.abs(out - p).max() > 0.01  # the world has overlaps to resolve


def test_stencil_pass_ignores_order_inside_cells(oracle):
    """The fixed-point sum makes the pass independent of the order of the
    particles within their cells: a permuted packed layout gives every
    particle bit-identical output."""
    P, idx, p, _ = _packed_world(oracle, 600, 16, 1, 3)
    rng = np.random.default_rng(7)
    perm = np.arange(len(p))
    for c in range(len(idx) - 2):
        s, e = int(idx[c + 1]), int(idx[c + 2])
        perm[s:e] = s + rng.permutation(e - s)
    q = p[perm]
    out, ov = oracle.stencil_pass(P, idx, p, p)
    out2, ov2 = oracle.stencil_pass(P, idx, q, q)
    assert np.array_equal(bits(out2), bits(out[perm])) and ov == ov2


def test_stencil_budget_bounds_the_work(oracle):
    """A particle whose 3x3 cells hold more than the budget (4096) takes no
    displacement; a sparse particle beside the crowd still does."""
    crowd = np.full((4200, 2), 5.5, np.float32) + f32(np.random.default_rng(0).uniform(-0.4, 0.4, (4200, 2)))
    pos = np.concatenate([crowd, f32([[20.5, 20.5], [21.2, 20.5]])])
    idx, p, _ = oracle.pack(pos, np.zeros_like(pos), (0, 0, 32, 32), 1.0)
    P = oracle.params(32, 32, 1, len(p))
    out, _ = oracle.stencil_pass(P, idx, p, p)
    in_crowd = np.abs(p[:, 0] - 5.5) < 1
    assert np.array_equal(out[in_crowd], p[in_crowd])
    assert out[~in_crowd][1, 0] - out[~in_crowd][0, 0] == F(1.0)


def test_step_s_zero_passes_is_the_reference_step(oracle):
    P, idx, p, v = _packed_world(oracle, 500, 20, 1, 5)
    a = oracle.step(P, idx, p, v, nsteps=3)
    b = oracle.step_s(P, idx, p, v, max_passes=0, nsteps=3)
    for x, y in zip(a, b[:3]):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    assert b[3][2] == 0


def test_step_s_passes_and_tolerance(oracle):
    """Passes reduce the largest overlap; a tolerance at or above the first
    overlap runs none; the frame after the passes is the reference step on
    the separated positions."""
    P, idx, p, v = _packed_world(oracle, 500, 20, 1, 6)
    _, _, _, (before, after, ran, sums) = oracle.step_s(P, idx, p, v, max_passes=4)
    assert ran == 4 and after < before
    assert sums[0] / sums[1] > sums[2] / sums[3]  # the mean overlap falls too
    _, _, _, st = oracle.step_s(P, idx, p, v, max_passes=4, overlap_tol=before)
    assert st[:3] == (before, before, 0) and st[3][:2] == st[3][2:]
    moved, _ = oracle.stencil_pass(P, idx, p, p)
    i2, p2, v2, st1 = oracle.step_s(P, idx, p, v, max_passes=1)
    r = oracle.step(P, idx, moved, v)
    assert st1[2] == 1 and np.array_equal(bits(p2), bits(r[1])) and np.array_equal(i2, r[0])
