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
