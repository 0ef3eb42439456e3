"""Mode S (separation passes over the 3x3 cells around each particle, then
the reference step) on the HIP path vs the oracle's wo_step_s, through the
C ABI (needs a GPU).

Mode S is an extension with no counterpart in the reference.  Its passes sum
each particle's pair terms in 2^-30 fixed point, so the device's in-cell
order does not matter and the bar is the same as mode R's: indices exact,
positions and velocities bit for bit, and the overlap statistics equal.
"""
import numpy as np
import pytest

from wrach_amd import World, WrachConfig, WrachError, workloads
from wrach_amd import _lib
from tests.test_gpu_parity import _batch_worlds, assert_same_state, f32, oracle_world, say

pytestmark = pytest.mark.gpu


def run_stencil(oracle, cfg, pos, vel, frames, passes, tol=0.0, every=2, label="", graph=True):
    say(f"[stencil] {label}: N={len(pos)} dims={cfg.dimensions} cs={cfg.cell_size} passes={passes} "
        f"tol={tol} frames={frames} graph={graph}")
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        w.set_graph(graph)
        w.set_solver("S", passes, tol)
        done = 0
        while done < frames:
            k = min(every, frames - done)
            w.step(k)
            *ref, st = oracle.step_s(P, *ref, max_passes=passes, overlap_tol=tol, nsteps=k)
            done += k
            assert_same_state(w.read(), tuple(ref), f"{label} frame {done}")
            got = w.solver_stats()
            assert got == st, f"{label} frame {done}: stats {got} against {st}"
        return st


def uniform(n, w, h, seed):
    rng = np.random.default_rng(seed)
    pos = f32(np.stack([rng.uniform(0, w, n), rng.uniform(0, h / 2, n)], 1))
    vel = f32(rng.uniform(-1, 1, (n, 2)))
    return pos, vel


@pytest.mark.parametrize("passes", [1, 2, 3])
def test_stencil_uniform_16k(oracle, passes):
    pos, vel = uniform(16384, 512, 512, 11)
    cfg = WrachConfig(dimensions=(512, 512), cell_size=1, gravity=(0.0, -0.1))
    st = run_stencil(oracle, cfg, pos, vel, 8, passes, label="uniform-16k")
    assert st[2] == passes and st[1] < st[0]


def test_stencil_cell_size_3(oracle):
    pos, vel = uniform(6000, 240, 180, 12)
    cfg = WrachConfig(dimensions=(240, 180), cell_size=3)
    run_stencil(oracle, cfg, pos, vel, 6, 2, label="cs3")


def test_stencil_dense_clusters(oracle):
    """Clusters of 10-60 particles per cell (mode R's group and wide classes
    beside the passes), and every lane visiting hundreds of neighbours."""
    rng = np.random.default_rng(13)
    centres = rng.uniform(5, 59, (40, 2))
    pos = np.concatenate([c + rng.normal(0, 0.6, (rng.integers(10, 60), 2)) for c in centres])
    pos = f32(np.clip(pos, 0, 63.99))
    vel = f32(rng.normal(0, 0.3, pos.shape))
    cfg = WrachConfig(dimensions=(64, 64), cell_size=1, gravity=(0.0, -0.05))
    run_stencil(oracle, cfg, pos, vel, 6, 2, label="clusters")


def test_stencil_tolerance_stops_the_passes(oracle):
    """A tolerance between the first and the best overlap: the passes stop
    early, after the count the oracle stops at."""
    pos, vel = uniform(4000, 128, 128, 14)
    cfg = WrachConfig(dimensions=(128, 128), cell_size=1)
    P, ref = oracle_world(oracle, cfg, pos, vel)
    *_, st8 = oracle.step_s(P, *ref, max_passes=8)
    tol = (st8[0] + st8[1]) / 2
    st = run_stencil(oracle, cfg, pos, vel, 4, 8, tol=tol, every=1, label="tolerance")
    assert 0 < st[2] < 8 or st[0] <= tol


def test_stencil_budget_guard(oracle):
    """A cell of 4,200 particles: its lanes and their neighbours' exceed the
    4,096-particle budget and take no displacement; the rest of the world
    is separated as usual."""
    rng = np.random.default_rng(15)
    crowd = f32(np.full((4200, 2), 20.5) + rng.uniform(-0.45, 0.45, (4200, 2)))
    rest, _ = uniform(800, 128, 128, 16)
    pos = np.concatenate([crowd, rest])
    vel = f32(rng.uniform(-0.2, 0.2, pos.shape))
    cfg = WrachConfig(dimensions=(128, 128), cell_size=1)  # capacity 18,305 (cells x 1.1)
    run_stencil(oracle, cfg, pos, vel, 3, 1, every=1, label="budget")


def test_stencil_plain_launches_match_graphs(oracle):
    pos, vel = uniform(8000, 256, 256, 17)
    cfg = WrachConfig(dimensions=(256, 256), cell_size=1, gravity=(0.0, -0.1))
    run_stencil(oracle, cfg, pos, vel, 20, 3, every=10, label="plain launches", graph=False)
    run_stencil(oracle, cfg, pos, vel, 20, 3, every=10, label="graph of 16 frames", graph=True)


def test_stencil_zero_passes_and_back_to_mode_r(oracle):
    """Mode S with no passes is the reference's step; switching back to R
    mid-run continues the reference's step."""
    pos, vel = uniform(8000, 256, 256, 18)
    cfg = WrachConfig(dimensions=(256, 256), cell_size=1, gravity=(0.0, -0.1))
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        w.set_solver("S", 0)
        w.step(5)
        ref = oracle.step(P, *ref, nsteps=5)
        assert_same_state(w.read(), ref, "zero passes")
        w.set_solver("S", 2)
        w.step(3)
        *ref, _ = oracle.step_s(P, *ref, max_passes=2, nsteps=3)
        assert_same_state(w.read(), tuple(ref), "two passes")
        w.set_solver("R")
        w.step(4)
        ref = oracle.step(P, *ref, nsteps=4)
        assert_same_state(w.read(), ref, "back to R")


def test_stencil_argument_errors(oracle):
    say("[stencil] errors")
    cfg = WrachConfig(dimensions=(32, 32), cell_size=1)
    with World(cfg) as w:
        for args in [("S", _lib.WRACH_MAX_PASSES + 1, 0.0), ("S", 1, float("nan")), ("S", 1, -1.0),
                     ("S", 1, float("inf"))]:
            with pytest.raises(WrachError) as e:
                w.set_solver(*args)
            assert e.value.code == _lib.WRACH_E_INVALID, args
        with pytest.raises(WrachError) as e:
            _lib.check(w._L.wrach_set_solver(w._h, 7, 1, 0.0))
        assert e.value.code == _lib.WRACH_E_INVALID
        w.add_particles(f32([[3.5, 3.5], [4.1, 3.5]]), f32([[0, 0], [0, 0]]))
        w.set_solver("S", 1)
        w.step(1)
        assert w.solver_stats()[2] == 1


def test_stencil_lattice_1m_full_size(oracle):
    """The 1,000,000-particle lattice, 2048x1024, with two passes per frame:
    exact against the oracle for 3 frames."""
    wl = workloads.lattice_1m()
    run_stencil(oracle, wl.config, wl.pos, wl.vel, 3, 2, every=3, label=wl.name)


def test_stencil_with_autotune_and_profiled_frames(oracle):
    """The autotuner's trial frames and the event-timed frames run mode S's
    passes too: the world continues bit for bit after autotuning (which
    restores the state) and after profiled frames (which step it)."""
    pos, vel = uniform(12000, 384, 256, 19)
    cfg = WrachConfig(dimensions=(384, 256), cell_size=1, gravity=(0.0, -0.1))
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        w.set_solver("S", 2)
        w.step(4)
        *ref, _ = oracle.step_s(P, *ref, max_passes=2, nsteps=4)
        w.autotune(2)
        assert_same_state(w.read(), tuple(ref), "after autotune")
        ms = w.step_profiled(3)
        assert all(v > 0 for v in ms.values())
        *ref, st = oracle.step_s(P, *ref, max_passes=2, nsteps=3)
        assert_same_state(w.read(), tuple(ref), "after profiled frames")
        assert w.solver_stats() == st
        w.step(5)
        *ref, _ = oracle.step_s(P, *ref, max_passes=2, nsteps=5)
        assert_same_state(w.read(), tuple(ref), "graph frames after")


@pytest.mark.parametrize("cs", [1, 3])
def test_stencil_batch_matches_each_world_alone(oracle, cs):
    """Mode S on a batch: every world's particles meet only their own
    world's cells, so each world matches the single-world oracle bit for bit
    (with no tolerance, a world without overlaps is unchanged by the passes
    that the others run); the batch's statistics are the worlds' combined."""
    W, H = 48, 32
    say(f"[stencil] batch of 5 worlds {W}x{H} cs={cs}, 2 passes")
    cfg = WrachConfig(dimensions=(W, H), cell_size=cs, gravity=(0.0, -0.1))
    worlds = _batch_worlds(cs, W, H, 70 + cs)
    with World(cfg, worlds=len(worlds)) as b:
        for w, (p, v) in enumerate(worlds):
            b.add_particles(p, v, world=w)
        b.set_solver("S", 2)
        refs = [oracle_world(oracle, cfg, p, v) for p, v in worlds]
        for step in range(3):
            b.step(3)
            stats = []
            for w, (P, ref) in enumerate(refs):
                *ref, st = oracle.step_s(P, *ref, max_passes=2, nsteps=3)
                refs[w] = (P, tuple(ref))
                stats.append(st)
                assert_same_state(b.read_world(w), tuple(ref), f"world {w}, frame {3 * (step + 1)}")
            before, after, ran, sums = b.solver_stats()
            assert before == max(st[0] for st in stats) and after == max(st[1] for st in stats)
            assert sums == tuple(sum(st[3][j] for st in stats) for j in range(4))
            assert ran <= 2 and (ran > 0) == (before > 0)


@pytest.mark.parametrize("tol", [0.0, 0.3])
def test_stencil_quiet_skips_only_the_closing_measure(oracle, tol):
    """WRACH_SOLVER_QUIET: the same frames as mode S (bit for bit against the
    oracle), the same overlap before the passes and passes run; the figures
    after the passes are not taken."""
    pos, vel = uniform(6000, 160, 128, 21)
    cfg = WrachConfig(dimensions=(160, 128), cell_size=1, gravity=(0.0, -0.1))
    say(f"[stencil] quiet, tol={tol}")
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        w.set_solver("S", 3, tol, quiet=True)
        for _ in range(3):
            w.step(2)
            *ref, st = oracle.step_s(P, *ref, max_passes=3, overlap_tol=tol, nsteps=2)
            assert_same_state(w.read(), tuple(ref), "quiet")
            before, after, ran, sums = w.solver_stats()
            assert (before, ran, sums[:2]) == (st[0], st[2], st[3][:2])
            assert after == 0.0 and sums[2:] == (0, 0) and not w.after_measured()
        w.set_solver("S", 3, tol)
        w.step(1)
        *ref, st = oracle.step_s(P, *ref, max_passes=3, overlap_tol=tol, nsteps=1)
        assert w.solver_stats() == st and w.after_measured()
    with World(cfg) as w:
        with pytest.raises(WrachError) as e:
            _lib.check(w._L.wrach_set_solver(w._h, _lib.WRACH_SOLVER_R | _lib.WRACH_SOLVER_QUIET, 0, 0.0))
        assert e.value.code == _lib.WRACH_E_INVALID


# Crowded cells: 64..1,024 particles with at most 4,096 in their 3x3 cells
# (a lane's work stays within the budget), beside cells of every other kind.
CROWD_MIN, CROWD_MAX = 64, 1024


def crowded_cells_world(seed, side=112, origin=(0, 0)):
    """Cells of exactly 63, 64, 1,024 and 1,025 particles; a 3x3 block of
    crowded cells (crowded neighbours in all eight directions); crowded
    cells beside a cell over the budget, at a corner and on an edge;
    diagonal crowded neighbours; over a sparse background.  Cell size 1,
    cells given relative to `origin`.  (Worlds hold about as many particles
    as cells, so the world is larger than the cells it uses.)"""
    rng = np.random.default_rng(seed)
    designed = {(10 + i, 10 + j): 100 for i in range(3) for j in range(3)}
    designed.update({(20, 10): 64, (21, 10): 63, (20, 11): 64, (21, 11): 64,
                     (30, 10): 1024, (31, 10): 1025, (30, 11): 70,
                     (40, 30): 200, (41, 30): 4000, (39, 30): 100, (39, 31): 80,
                     (0, 0): 300, (1, 0): 65, (0, 1): 90,
                     (side - 1, 20): 500, (side - 1, 21): 200, (side - 2, 19): 64,
                     (25, 40): 128, (26, 41): 128, (24, 39): 128, (26, 39): 128, (24, 41): 128})
    bg = rng.uniform(0, side, (2000, 2))
    keep = np.array([tuple(c) not in designed for c in np.floor(bg).astype(int)])
    parts = [bg[keep]]
    for (cx, cy), k in designed.items():
        parts.append(np.stack([cx + rng.uniform(0.001, 0.999, k), cy + rng.uniform(0.001, 0.999, k)], 1))
    pos = f32(np.concatenate(parts) + np.asarray(origin, np.float64))
    vel = f32(rng.normal(0, 0.05, pos.shape))
    return pos, vel


def crowded_cells(P, idx):
    """The crowded cells of a packed layout (a check on the test's inputs,
    computed from the oracle's offsets), and each cell's count."""
    gx, gy = P.grid_dimensions[0], P.grid_dimensions[1]
    n = P.particles_in_frame_count
    st = np.minimum(np.asarray(idx[1:gx * gy + 2], np.int64), n)
    cnt = (st[1:] - st[:-1]).reshape(gy, gx)
    pad = np.pad(cnt, 1)
    tot = sum(pad[1 + dy:1 + dy + gy, 1 + dx:1 + dx + gx] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    return (cnt >= CROWD_MIN) & (cnt <= CROWD_MAX) & (tot <= 4096), cnt


def test_stencil_crowded_cells(oracle):
    """Crowded cells beside cells of every other kind, 6 frames of 2 passes
    (graphs) and 4 frames of 3 passes with a tolerance (plain launches),
    bit for bit with the statistics."""
    pos, vel = crowded_cells_world(31)
    cfg = WrachConfig(dimensions=(112, 112), cell_size=1)
    P, (idx, _, _) = oracle_world(oracle, cfg, pos, vel)
    crowd, cnt = crowded_cells(P, idx)
    assert crowd[10:13, 10:13].all() and crowd[10, 20] and not crowd[10, 21] and cnt[10, 21] == 63
    assert crowd[10, 30] and cnt[10, 30] == CROWD_MAX and not crowd[10, 31] and cnt[10, 31] == CROWD_MAX + 1
    assert not crowd[30, 40] and crowd[30, 39] and crowd[0, 0] and crowd[20, 111]
    run_stencil(oracle, cfg, pos, vel, 6, 2, every=1, label="crowded cells")
    run_stencil(oracle, cfg, pos, vel, 4, 3, tol=0.2, every=2, label="crowded cells, tolerance", graph=False)


def test_stencil_crowded_cells_in_a_batch(oracle):
    """Two worlds of different sizes with crowded cells in one batch: each
    cell's 3x3 cells clipped to its own world, each world bit for bit
    against the oracle alone."""
    ws = [(WrachConfig(dimensions=(112, 112), cell_size=1), *crowded_cells_world(32)),
          (WrachConfig(dimensions=(128, 120), cell_size=1), *crowded_cells_world(33, side=120, origin=(4, 0)))]
    say("[stencil] crowded cells in a batch of 2 worlds, 2 passes")
    with World([c for c, _, _ in ws]) as b:
        for w, (c, p, v) in enumerate(ws):
            b.add_particles(p, v, world=w)
        b.set_solver("S", 2)
        refs = [oracle_world(oracle, c, p, v) for c, p, v in ws]
        for frames in (1, 3):
            b.step(frames)
            for w, (P, ref) in enumerate(refs):
                ref = oracle.step_s(P, *ref, 2, nsteps=frames)[:3]
                refs[w] = (P, ref)
                assert_same_state(b.read_world(w), ref, f"world {w}")


def test_stencil_collapsed_lattice(oracle):
    """A lattice 256 cells wide and 488 rows high collapsed under gravity
    (1,000 frames of the reference step, about 380 particles in each floor
    cell), then mode S frames from that state: the oracle's passes from the
    same state, bit for bit."""
    wl = workloads.lattice_1m(n=256 * 488, width=256, height=1024)
    cfg = wl.config
    say(f"[stencil] collapsed lattice {wl.name}")
    with World(cfg) as w:
        w.add_particles(wl.pos, wl.vel)
        w.step(1000)
        state = w.read()
        W, H = cfg.dimensions
        P = oracle.params(W, H, cfg.cell_size, len(state[1]), gravity=cfg.gravity, dt=cfg.dt)
        crowd, cnt = crowded_cells(P, state[0])
        say(f"[stencil] collapsed: {int(crowd.sum())} crowded cells holding {int(cnt[crowd].sum())} particles")
        assert crowd.sum() >= 200
        w.set_solver("S", 2)
        ref = state
        for frame in range(2):
            w.step(1)
            *ref, st = oracle.step_s(P, *ref, max_passes=2, nsteps=1)
            assert_same_state(w.read(), tuple(ref), f"collapsed, mode S frame {frame + 1}")
            assert w.solver_stats() == st


def test_stencil_pair_arithmetic_at_its_edges(oracle):
    """The pass's root and division take short sequences for squared
    distances from 2^-96 to 1 + 2^-23 and hipcc's full ones below: pairs on
    both sides of 2^-96, subnormal squares, coincident pairs, d exactly 1,
    the reach bound 1 + 2^-23 and just past it, near the world's origin
    where differences can be that small; bit for bit over 3 frames."""
    base = []
    for e in (20, 40, 47, 48, 49, 60, 70, 74, 75, 100, 140, 149):
        d = np.float32(2.0 ** -e)
        base += [[0.0, 0.0], [float(d), 0.0], [0.0, float(d)], [float(d), float(d)]]
    base += [[0.25, 0.25], [0.25, 0.25], [0.25, 0.25]]                      # coincident
    base += [[0.5, 0.3], [1.5, 0.3]]                                          # d == 1
    one = float(np.sqrt(np.float32(1.0 + 2.0 ** -23)))
    base += [[0.1, 0.7], [0.1 + one, 0.7]]                                    # d rounds to 1
    base += [[0.2, 0.9], [float(np.nextafter(np.float32(1.2), np.float32(2))), 0.9]]  # just past reach
    r = np.random.default_rng(41)
    tiny = r.uniform(0, 2.0 ** -40, (200, 2))                                 # a crowd at the origin
    pos = f32(np.concatenate([f32(base), tiny, r.uniform(0, 24, (300, 2))]))
    vel = f32(np.zeros_like(pos))
    cfg = WrachConfig(dimensions=(24, 24), cell_size=1)
    run_stencil(oracle, cfg, pos, vel, 3, 2, every=1, label="pair arithmetic edges")


def test_pair_arithmetic_short_sequences_are_exact():
    """The root and force sequences that modes S and L use for squared
    distances in [2^-96, 1 + 2^-23] and distances in [2^-48, 1] give hipcc's
    correctly rounded sqrtf and division on every float of those ranges.  The
    root's Newton step from the hardware reciprocal root and the force's
    one-correction sequence from the hardware reciprocal are exact only by
    a property of those approximations, so this runs on the device itself,
    over all 805,306,370 and 402,653,185 inputs."""
    import ctypes

    say("[stencil] pair arithmetic: exhaustive root and force check on device 0")
    L = _lib.load()
    out = (ctypes.c_uint64 * 4)()
    _lib.check(L.wrach_check_pair_arithmetic(0, out))
    assert list(out) == [0, 0, 805306370, 402653185]
