"""HIP path vs the oracle, through the C ABI (needs a GPU).

Bar: cell offsets, per-cell counts and in-cell order are compared exactly;
positions and velocities are compared bit for bit, which implies the
stated per-step tolerance (max abs error <= 1e-5 world units, TOL below).
"""
import json
import os
import sys

import numpy as np
import pytest

from wrach_amd import Particle, World, WrachAPI, WrachConfig, WrachError, workloads
from wrach_amd import _lib

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def f32(a):
    return np.asarray(a, np.float32)


def assert_same_state(gpu, ref, what=""):
    gi, gp, gv = gpu
    ri, rp, rv = ref
    assert np.array_equal(gi, ri), f"{what}: indices differ at {np.flatnonzero(gi != ri)[:10]}"
    assert gp.shape == rp.shape and gv.shape == rv.shape, what
    # The tolerance is checked where both states are finite; a NaN or
    # infinity on one side only fails here (the reference's clamps let an
    # infinite velocity through, so both may hold one); the bit comparison
    # below is stricter still.
    for g, r, name in ((gp, rp, "positions"), (gv, rv, "velocities")):
        assert np.array_equal(np.isfinite(g), np.isfinite(r)), f"{what}: non-finite {name} differ"
        assert np.array_equal(np.isnan(g), np.isnan(r)), f"{what}: NaN {name} differ"
    fin_p, fin_v = np.isfinite(rp), np.isfinite(rv)
    diff = np.concatenate([np.abs(gp - rp)[fin_p], np.abs(gv - rv)[fin_v], np.zeros(1, np.float32)])
    err = float(np.max(diff))
    assert err <= TOL, f"{what}: max abs error {err}"
    bad = np.flatnonzero(np.any(gp.view(np.uint32) != rp.view(np.uint32), axis=1) |
                         np.any(gv.view(np.uint32) != rv.view(np.uint32), axis=1))
    assert bad.size == 0, f"{what}: {bad.size} particles differ in bits, first {bad[:8]}"


def oracle_world(oracle, cfg: WrachConfig, pos, vel):
    w, h = cfg.dimensions
    idx, p, v = oracle.pack(pos, vel, (0, 0, w, h), cfg.cell_size)
    P = oracle.params(w, h, cfg.cell_size, len(p), gravity=cfg.gravity, dt=cfg.dt)
    return P, (idx, p, v)


def run_parity(oracle, cfg, pos, vel, steps, every=1, label="", threads=1):
    say(f"[parity] {label}: N={len(pos)} dims={cfg.dimensions} cs={cfg.cell_size} g={cfg.gravity} steps={steps}")
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        assert_same_state(w.read(), ref, f"{label} upload")
        done = 0
        while done < steps:
            k = min(every, steps - done)
            w.step(k)
            ref = oracle.step(P, *ref, nsteps=k, nthreads=threads)
            done += k
            assert_same_state(w.read(), ref, f"{label} frame {done}")
    return ref


# ------------------------------------------------------ the reference's tests

@pytest.mark.parametrize("case", GOLD["gpu_frames"], ids=lambda c: c["source"].split(" ")[-1])
def test_reference_gpu_tests_through_wrachapi(oracle, case):
    say(f"[ref] {case['source']}")
    w, h = case["dimensions"]
    cfg = WrachConfig(dimensions=(w, h), cell_size=case["cell_size"])
    api = WrachAPI(cfg)
    parts = [Particle(tuple(p), tuple(v)) for p, v in zip(case["positions"], case["velocities"])]
    api.add_particles(parts)
    for _ in range(case["ticks"]):
        api.tick()
    state = api.get_simulation_state()
    cpu_idx, _, _ = oracle.pack(case["positions"], case["velocities"], (0, 0, w, h), cfg.cell_size)
    if "expected_indices" in case:
        assert state.packed_data.indices.tolist() == case["expected_indices"]
    if case.get("indices_equal_cpu_pack"):
        assert state.packed_data.indices.tolist() == cpu_idx.tolist()
    if "expected_positions" in case:
        n = len(case["expected_positions"])
        got = state.packed_data.positions[:n]
        # canonical order is exact here; the reference sorts the first cell
        assert np.array_equal(got.view(np.uint32), f32(case["expected_positions"]).view(np.uint32))
    if "expected_readback_length" in case:
        assert len(api.positions) == case["expected_readback_length"]
        assert len(api.velocities) == case["expected_readback_length"]
        assert api.positions[0] != (0.0, 0.0)
        assert api.velocities[0] != (0.0, 0.0)
    api.close()


# ------------------------------------------------------ step parity, small

@pytest.mark.parametrize("gravity", [(0.0, 0.0), (0.0, -0.1)], ids=["reference", "gravity"])
def test_uniform_16k_every_frame(oracle, gravity):
    """16,384 particles in 512x512, compared after every one of 100 frames."""
    wl = workloads.small_uniform(seed=0, gravity=gravity)
    run_parity(oracle, wl.config, wl.pos, wl.vel, 100, every=1, label=wl.name)


@pytest.mark.parametrize("cs", [2, 3, 5])
def test_cell_sizes(oracle, cs):
    r = np.random.default_rng(30 + cs)
    n = 6000
    pos = f32(r.uniform(0, [200, 120], (n, 2)))
    vel = f32(r.uniform(-1.3, 1.3, (n, 2)))
    cfg = WrachConfig(dimensions=(200, 120), cell_size=cs, gravity=(0.0, -0.1))
    run_parity(oracle, cfg, pos, vel, 30, every=3, label=f"cs{cs}")


# ------------------------------------------------------ edge cases

def test_empty_world(oracle):
    say("[edge] empty world")
    cfg = WrachConfig(dimensions=(30, 20), cell_size=1)
    with World(cfg) as w:
        w.step(5)
        idx, p, v = w.read()
        assert idx.tolist() == [0] * (31 * 21 + 2) and len(p) == 0
        w.add_particles(np.zeros((0, 2)), np.zeros((0, 2)))
        w.step(2)
        assert w.read()[0].tolist() == [0] * (31 * 21 + 2)


def test_single_and_coincident_particles(oracle):
    pos = f32([[5.5, 5.5], [7.25, 3.5], [7.25, 3.5], [7.25, 3.5], [0.0, 0.0], [20.0, 10.0]])
    vel = f32([[0.3, -0.2], [0.0, 0.0], [0.0, 0.0], [0.1, 0.1], [0.0, 0.0], [0.5, 0.5]])
    cfg = WrachConfig(dimensions=(20, 10), cell_size=1)
    run_parity(oracle, cfg, pos, vel, 20, label="coincident+walls")


@pytest.mark.parametrize("per_cell", [2, 3, 4, 5, 8, 9, 10, 16, 17, 33, 40, 64, 65, 300, 577, 704,
                                      705, 1024, 2048, 2049, 5000])
def test_full_cells(oracle, per_cell):
    """Every size class: lane groups (2..64), the wide path (65..704, its
    particles in registers) and the block path past it (LDS sort up to 2048,
    merge sort past it), all with the 9-particle collision cap."""
    r = np.random.default_rng(per_cell)
    cells = [(3, 4), (10, 2), (11, 2), (20, 7)]
    pos, vel = [], []
    for cx, cy in cells:
        pos.append(r.uniform([cx, cy], [cx + 1, cy + 1], (per_cell, 2)))
        vel.append(r.uniform(-1.2, 1.2, (per_cell, 2)))
    # cell size 1 gives ~1.1 particles of capacity per cell
    # (particle_store.rs:116-133), so the world is sized for 4 x 5000 + 500.
    extra = r.uniform(0, [160, 120], (500, 2))
    pos = f32(np.concatenate(pos + [extra]))
    vel = f32(np.concatenate(vel + [r.uniform(-1, 1, (500, 2))]))
    perm = r.permutation(len(pos))  # insertion order differs from cell order
    cfg = WrachConfig(dimensions=(160, 120), cell_size=1, gravity=(0.0, -0.1))
    steps = 6 if per_cell >= 2048 else 12
    run_parity(oracle, cfg, pos[perm], vel[perm], steps, label=f"full cells x{per_cell}")


def test_particles_outside_grid_are_dropped(oracle):
    pos = f32([[1.0, 1.0], [-0.5, 2.0], [9.0, 9.0], [10.0, 3.0], [3.0, 10.5], [1e30, 1.0]])
    vel = f32(np.full((6, 2), 0.25))
    cfg = WrachConfig(dimensions=(10, 10), cell_size=1)
    ref = run_parity(oracle, cfg, pos, vel, 4, label="off-grid")
    assert len(ref[1]) == 4  # x = 10 and y = 10.5 still bin into the 11x11 grid


def test_velocity_and_wall_extremes(oracle):
    pos = f32([[0.0, 0.0], [64.0, 32.0], [63.9, 0.1], [0.1, 31.9], [32.0, 16.0], [10.0, 10.0]])
    vel = f32([[-5.0, -5.0], [7.0, 3.0], [1e9, -1e9], [-0.0, 0.0], [np.inf, -np.inf], [1.0, -1.0]])
    cfg = WrachConfig(dimensions=(64, 32), cell_size=1, gravity=(0.0, -0.1))
    run_parity(oracle, cfg, pos, vel, 10, label="extremes")


@pytest.mark.parametrize("per_cell", [5, 20])
def test_far_movers_across_rows(oracle, per_cell):
    """Particles leaving a cell at column 0 for the last column one row down
    (key - cell == -1, the same difference as the left neighbour), mixed
    with ordinary neighbours, from a register-path cell (5 particles) and a
    wave-path cell (20)."""
    r = np.random.default_rng(77 + per_cell)
    W, H = 40, 20
    pos = f32(r.uniform([0.0, 5.0], [1.0, 6.0], (per_cell, 2)))
    vel = f32(r.uniform(-1, 1, (per_cell, 2)))
    vel[::3] = (W + 5.0, -1.0)
    cfg = WrachConfig(dimensions=(W, H), cell_size=1)
    run_parity(oracle, cfg, pos, vel, 6, label=f"far movers x{per_cell}")


@pytest.mark.parametrize("runs", [100, 300])
def test_many_runs_in_one_cell(oracle, runs):
    """A cell fed by `runs` far movers from distinct source cells (a run of
    one each) plus a near run: frame 2 orders it by runs (100, within the
    wide path's 256) or through the block-wide sort (300)."""
    r = np.random.default_rng(runs)
    W, H = 420, 40
    tx, ty = 10, 20
    d = np.arange(2, runs + 2, dtype=np.float64)
    pos = np.stack([tx + d + r.uniform(0.3, 0.7, runs), np.full(runs, ty) + r.uniform(0.3, 0.7, runs)], 1)
    vel = np.stack([-d, np.zeros(runs)], 1)
    near = np.stack([np.full(80, tx + 0.5) + r.uniform(-0.4, 0.4, 80), np.full(80, ty + 0.5)], 1)
    pos = f32(np.concatenate([pos, near, r.uniform(0, [W, H], (200, 2))]))
    vel = f32(np.concatenate([vel, np.zeros((80, 2)), r.uniform(-1, 1, (200, 2))]))
    cfg = WrachConfig(dimensions=(W, H), cell_size=1)
    run_parity(oracle, cfg, pos, vel, 4, label=f"many runs x{runs}")


def test_collapsed_columns(oracle):
    """64x512 world, a jittered lattice of 64 x 256 particles under gravity:
    the columns land on the floor as 64 cells of ~256 particles, Gauss-Seidel
    pushes send particles several cells sideways (far movers), corner cells
    collect many runs.  Exact every 25 frames for 400 frames."""
    r = np.random.default_rng(11)
    xs, ys = np.meshgrid(np.arange(64) + 0.5, np.arange(256) + 0.5)
    pos = f32(np.stack([xs.ravel(), ys.ravel()], 1) + r.uniform(-0.1, 0.1, (64 * 256, 2)))
    vel = f32(r.normal(0, 0.5, (64 * 256, 2)))
    cfg = WrachConfig(dimensions=(64, 512), cell_size=1, gravity=(0.0, -0.1))
    run_parity(oracle, cfg, pos, vel, 400, every=25, label="collapsed columns")


def test_errors_leave_state_usable(oracle):
    say("[edge] error paths")
    cfg = WrachConfig(dimensions=(6, 6), cell_size=2)  # capacity 74
    with World(cfg) as w:
        assert w.capacity == 74
        w.add_particles(f32([[1, 1]] * 70), f32([[0, 0]] * 70))
        with pytest.raises(WrachError) as e:
            w.add_particles(f32([[2, 2]] * 5), f32([[0, 0]] * 5))
        assert e.value.code == _lib.WRACH_E_CAPACITY
        assert w.n_particles == 70
        with pytest.raises(WrachError) as e:
            w.add_particles(f32([[np.nan, 1]]), f32([[0, 0]]))
        assert e.value.code == _lib.WRACH_E_NONFINITE
        w.step(3)
        assert w.read()[0][-1] == 70


def test_add_particles_repacks_the_store(oracle):
    """state.rs:90-101: add_particles re-packs the whole host store (which the
    frames never write back, particle_store.rs:77-85) and replaces the device
    state."""
    say("[edge] add after ticks")
    r = np.random.default_rng(40)
    a_p, a_v = f32(r.uniform(0, 30, (200, 2))), f32(r.uniform(-1, 1, (200, 2)))
    b_p, b_v = f32(r.uniform(0, 30, (50, 2))), f32(r.uniform(-1, 1, (50, 2)))
    cfg = WrachConfig(dimensions=(30, 30), cell_size=1)
    with World(cfg) as w:
        w.add_particles(a_p, a_v)
        w.step(7)
        w.add_particles(b_p, b_v)
        _, ref = oracle_world(oracle, cfg, np.concatenate([a_p, b_p]), np.concatenate([a_v, b_v]))
        assert_same_state(w.read(), ref, "repacked")


def _upload_batches(cs, W, H, seed):
    """Appends for the upload tests: uniform particles reaching outside the
    world, particles on cell borders and one float below them, and special
    coordinates (signed zeros, the far edges, infinities, huge and subnormal
    negatives), each batch with finite velocities."""
    r = np.random.default_rng(seed)
    out = []
    for n in (300, 1, 0, 250):
        p = r.uniform(-5, max(W, H) + 8, (n, 2)).astype(np.float32)
        q = n // 3
        border = (r.integers(-1, max(W, H) // cs + 2, (q, 2)) * cs).astype(np.float32)
        p[:q] = border
        p[q:2 * q] = np.nextafter(border, np.float32(-np.inf))
        out.append((p, r.uniform(-1, 1, (n, 2)).astype(np.float32)))
    special = f32([[0.0, 0.0], [-0.0, -0.0], [W, H], [W, 0], [0, H], [np.inf, 1], [-np.inf, 1],
                   [1, np.inf], [3e38, 2], [-3e38, 2], [-1e-30, 5], [5, -1e-30], [-1e-40, 5],
                   [5, -1e-40], [1e-40, 1e-40], [3 * cs, 2 * cs], [-cs, cs]])
    out.append((special, f32(r.uniform(-1, 1, special.shape))))
    return out


@pytest.mark.parametrize("cs", [1, 3, 7])
def test_device_upload_matches_host_packer(oracle, cs):
    """The device binning of the store (the default) against the host packer
    (WRACH_FLAG_HOST_PACK) and the oracle's pack (spatial_bin.rs:48-149),
    after every append; then frames from both states agree bit for bit."""
    W, H = 61, 47
    say(f"[upload] device vs host packer, cs={cs} dims=({W},{H})")
    cfg_d = WrachConfig(dimensions=(W, H), cell_size=cs, gravity=(0.0, -0.1))
    cfg_h = WrachConfig(dimensions=(W, H), cell_size=cs, gravity=(0.0, -0.1), host_pack=True)
    allp, allv = [], []
    with World(cfg_d) as wd, World(cfg_h) as wh:
        for k, (p, v) in enumerate(_upload_batches(cs, W, H, 200 + cs)):
            wd.add_particles(p, v)
            wh.add_particles(p, v)
            allp.append(p)
            allv.append(v)
            _, ref = oracle_world(oracle, cfg_d, np.concatenate(allp), np.concatenate(allv))
            assert_same_state(wd.read(), ref, f"device pack after append {k}")
            assert_same_state(wh.read(), ref, f"host pack after append {k}")
        P, _ = oracle_world(oracle, cfg_d, np.concatenate(allp), np.concatenate(allv))
        wd.step(6)
        wh.step(6)
        ref = oracle.step(P, *ref, nsteps=6)
        assert_same_state(wd.read(), ref, "6 frames after the device pack")
        assert_same_state(wh.read(), ref, "6 frames after the host pack")


def test_device_upload_full_cells_and_capacity(oracle):
    """Device binning of crowded cells (one particle per store index, so the
    first frame orders up to hundreds of one-particle runs per cell) and a
    capacity error that leaves store and state as they were."""
    say("[upload] crowded cells and capacity")
    r = np.random.default_rng(9)
    cfg = WrachConfig(dimensions=(40, 30), cell_size=1, gravity=(0.0, -0.1))
    with World(cfg) as w:
        cap = w.capacity
        cells = r.integers(0, 8, (cap - 200,))
        p = f32(np.stack([cells * 4.0 + 0.25, np.full(len(cells), 3.5)], 1))
        p += f32(r.uniform(0, 0.5, p.shape))
        v = f32(r.uniform(-1, 1, p.shape))
        w.add_particles(p, v)
        _, ref = oracle_world(oracle, cfg, p, v)
        assert_same_state(w.read(), ref, "crowded upload")
        with pytest.raises(WrachError) as e:
            w.add_particles(f32(r.uniform(0, 40, (300, 2))), f32(np.zeros((300, 2))))
        assert e.value.code == _lib.WRACH_E_CAPACITY
        assert_same_state(w.read(), ref, "after the refused append")
        P, _ = oracle_world(oracle, cfg, p, v)
        w.step(4)
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=4), "4 frames")
        extra_p, extra_v = f32(r.uniform(0, 40, (150, 2))), f32(r.uniform(-1, 1, (150, 2)))
        w.add_particles(extra_p, extra_v)
        _, ref = oracle_world(oracle, cfg, np.concatenate([p, extra_p]), np.concatenate([v, extra_v]))
        assert_same_state(w.read(), ref, "append after the refusal")


def test_checkpoint_round_trip(oracle):
    """A packed triple saved with read() and restored into a fresh context
    (wrach_load_packed) continues exactly as the original does."""
    wl = workloads.small_uniform(seed=5)
    say("[checkpoint] save, restore, continue")
    w_, h_ = wl.config.dimensions
    with World(wl.config) as a, World(wl.config) as b:
        a.add_particles(wl.pos, wl.vel)
        a.step(12)
        idx, pos, vel = a.read()
        b.load_packed(idx, pos, vel)
        assert_same_state(b.read(), (idx, pos, vel), "restored")
        a.step(9)
        b.step(9)
        P = oracle.params(w_, h_, wl.config.cell_size, len(pos), gravity=wl.config.gravity,
                          dt=wl.config.dt)
        ref = oracle.step(P, idx, pos, vel, nsteps=9)
        assert_same_state(a.read(), ref, "original after 9 more frames")
        assert_same_state(b.read(), ref, "restored after 9 frames")
        # the restored context's store is the checkpoint: a re-pack gives it back
        b.load_packed(idx, pos, vel)
        b.add_particles(f32(np.zeros((0, 2))), f32(np.zeros((0, 2))))
        assert_same_state(b.read(), (idx, pos, vel), "re-pack of the restored store")


def _window_layout(seed: int, sizes):
    """Positions inside a 64x64 world (cell size 1, cells row-major) whose
    packed layout holds cells of the given particle counts in order, one
    per cell from cell 0 on; velocities small and random."""
    r = np.random.default_rng(seed)
    gx = 65
    pos, vel = [], []
    for c, m in enumerate(sizes):
        cx, cy = c % gx, c // gx
        if m == 0 or cx >= 64 or cy >= 64:
            continue
        pos.append(np.stack([cx + r.uniform(0.05, 0.95, m), cy + r.uniform(0.05, 0.95, m)], 1))
        vel.append(r.uniform(-0.6, 0.6, (m, 2)))
    return f32(np.concatenate(pos)), f32(np.concatenate(vel))


@pytest.mark.parametrize("case", ["straddle", "mixed", "wide"])
def test_cells_across_wavefront_windows(oracle, case):
    """collide_kernel handles cells of up to 3 particles inside a window of
    64 packed particles and queues the others (larger, or running past the
    window).  Cells of 1..5 particles placed across the window edges at 64,
    128, ..., a cell spanning several windows and one filling a window
    exactly, loaded as an exact packed triple and stepped against the oracle."""
    if case == "straddle":
        # 62 singles, then a 3-cell over 62..64 (across the edge at 64), a 2-cell, ...
        sizes = [1] * 62 + [3, 2] + [1] * 59 + [4] + [2] * 3 + [5] + [1] * 50 + [2, 3, 4, 5] * 6
    elif case == "mixed":
        r = np.random.default_rng(7)
        sizes = list(r.choice([0, 1, 1, 1, 2, 2, 3, 4, 5, 7], 700))
    else:
        # a cell of 150 over three windows, one of 64 filling the window
        # 192..255 exactly, then cells of 700 and 900
        sizes = [1] * 40 + [150] + [1] * 2 + [64] + [3] * 10 + [700] + [2] * 30 + [900] + [1] * 20
    pos, vel = _window_layout(11, sizes)
    cfg = WrachConfig(dimensions=(64, 64), cell_size=1, gravity=(0.0, -0.05))
    P, ref = oracle_world(oracle, cfg, pos, vel)
    say(f"[windows] {case}: {len(pos)} particles")
    with World(cfg) as w:
        w.load_packed(*ref)
        for frames in (1, 1, 3, 10):
            w.step(frames)
            ref = oracle.step(P, *ref, nsteps=frames)
            assert_same_state(w.read(), ref, f"{case} after {frames} more")


def test_load_packed_rejects_bad_triples(oracle):
    say("[checkpoint] rejected triples leave the state")
    cfg = WrachConfig(dimensions=(20, 12), cell_size=1)
    r = np.random.default_rng(31)
    with World(cfg) as w:
        w.add_particles(f32(r.uniform(0, 20, (120, 2))), f32(r.uniform(-1, 1, (120, 2))))
        w.step(3)
        idx, pos, vel = w.read()
        before = (idx.copy(), pos.copy(), vel.copy())
        k = int(np.flatnonzero(np.diff(idx.astype(np.int64)) > 0)[1]) + 1
        bad_order = idx.copy()
        bad_order[k] = bad_order[k + 1] + 1
        cases = [
            ("short indices", idx[:-1], pos, vel, _lib.WRACH_E_INVALID),
            ("first offset", np.concatenate([[1], idx[1:]]), pos, vel, _lib.WRACH_E_INVALID),
            ("decreasing", bad_order, pos, vel, _lib.WRACH_E_INVALID),
            ("last offset", idx, pos[:-1], vel[:-1], _lib.WRACH_E_INVALID),
            ("NaN", idx, np.where(np.arange(len(pos))[:, None] == 5, np.nan, pos), vel,
             _lib.WRACH_E_NONFINITE),
        ]
        cap = w.capacity
        big = np.zeros_like(idx)
        big[2:] = cap + 1
        cases.append(("over capacity", big, np.ones((cap + 1, 2)), np.zeros((cap + 1, 2)),
                      _lib.WRACH_E_CAPACITY))
        for name, i_, p_, v_, code in cases:
            with pytest.raises(WrachError) as e:
                w.load_packed(i_, p_, v_)
            assert e.value.code == code, name
            assert_same_state(w.read(), before, f"after the rejected {name}")


@pytest.mark.parametrize("host", [False, True], ids=["device-store", "host-store"])
def test_store_write_back_then_append(oracle, host):
    """wrach_store_write_back (update_from_gpu, particle_store.rs:74-85): the
    store becomes the simulated particles in canonical order, so a re-pack
    reproduces the state and an append continues from it."""
    say(f"[write-back] host_pack={host}")
    cfg = WrachConfig(dimensions=(40, 30), cell_size=1, gravity=(0.0, -0.1), host_pack=host)
    r = np.random.default_rng(52)
    a_p = f32(r.uniform(-10, 50, (400, 2)))
    a_v = f32(r.uniform(-1, 1, (400, 2)))
    b_p = f32(r.uniform(0, 40, (80, 2)))
    b_v = f32(r.uniform(-1, 1, (80, 2)))
    with World(cfg) as w:
        w.add_particles(a_p, a_v)
        w.step(8)
        idx8, p8, v8 = w.read()
        w.store_write_back()
        assert_same_state(w.read(), (idx8, p8, v8), "write-back leaves the state")
        w.add_particles(f32(np.zeros((0, 2))), f32(np.zeros((0, 2))))
        assert_same_state(w.read(), (idx8, p8, v8), "re-pack of the written-back store")
        w.add_particles(b_p, b_v)
        P, ref = oracle_world(oracle, cfg, np.concatenate([p8, b_p]), np.concatenate([v8, b_v]))
        assert_same_state(w.read(), ref, "append after the write-back")
        w.step(5)
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=5), "5 frames after the append")


SCAN_SHAPES = [(1024, 8), (512, 16), (1024, 4), (512, 8), (1024, 16), (1024, 32), (1024, 64), (512, 64),
               (256, 128)]


@pytest.mark.parametrize("scan", SCAN_SHAPES, ids=lambda s: f"scan{s[0]}x{s[1]}")
@pytest.mark.parametrize("per_lane", [2, 1, 4], ids=lambda k: f"scatter{k}")
def test_every_launch_shape_matches_the_oracle(oracle, scan, per_lane):
    """Every scan tile shape and scatter width the autotuner can pick, on a
    world with one- to several-hundred-particle cells and a scan of many
    tiles, over frames with the captured graph."""
    say(f"[tuning] scan {scan} scatter {per_lane}")
    r = np.random.default_rng(11)
    cfg = WrachConfig(dimensions=(300, 200), cell_size=1, gravity=(0.0, -0.1))
    pos = np.concatenate([r.uniform(0, 300, (9000, 2)),
                          np.repeat(f32([[17.5, 3.5], [150.25, 40.75], [299.5, 0.5]]), [90, 300, 40], 0)
                          + r.uniform(-0.2, 0.2, (430, 2))]).astype(np.float32)
    vel = f32(r.uniform(-1, 1, pos.shape))
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        g0 = w.tuning()["cells_grid"]
        w.set_tuning(scan_threads=scan[0], scan_items=scan[1], scatter_per_lane=per_lane,
                     cells_grid=max(1, g0 // 2) if per_lane == 1 else g0)
        t = w.tuning()
        assert (t["scan_threads"], t["scan_items"], t["scatter_per_lane"]) == (*scan, per_lane)
        assert_same_state(w.read(), ref, "upload (scan of the store)")
        w.step(17)
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=17), "17 frames")


@pytest.mark.parametrize("scan", [(1024, 64), (512, 64), (256, 128)], ids=lambda s: f"scan{s[0]}x{s[1]}")
def test_largest_scan_tiles_over_many_tiles(oracle, scan):
    """Scan tiles of 32k and 64k counts over a 1201 x 801 grid (15 or 30
    tiles, the last one partial), with chunks that hold no particle between
    live ones."""
    say(f"[tuning] scan {scan} over many tiles")
    r = np.random.default_rng(12)
    cfg = WrachConfig(dimensions=(1200, 800), cell_size=1, gravity=(0.0, -0.1))
    pos = np.concatenate([r.uniform((0, 0), (1200, 40), (6000, 2)), r.uniform((0, 300), (1200, 310), (3000, 2)),
                          np.repeat(f32([[1199.5, 799.5], [600.25, 400.75]]), [70, 200], 0)
                          + r.uniform(-0.2, 0.2, (270, 2))]).astype(np.float32)
    vel = f32(r.uniform(-1, 1, pos.shape))
    P, ref = oracle_world(oracle, cfg, pos, vel)
    with World(cfg) as w:
        w.set_tuning(scan_threads=scan[0], scan_items=scan[1])
        w.add_particles(pos, vel)
        assert_same_state(w.read(), ref, "upload (scan of the store)")
        w.step(9)
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=9), "9 frames")


def test_scan_shape_follows_the_grid_size():
    """A context starts with 1024 x 8 scan tiles up to 2^23 cells, 1024 x 64
    up to 2^26 and 256 x 128 above, where the look-back over thousands of
    tiles costs more than the larger tiles: worlds of exactly 2^23 and 2^26
    cells and just above."""
    say("[tuning] default scan shape by grid size")
    for dims, shape in (((2048, 1024), (1024, 8)),
                        ((4095, 2047), (1024, 8)),      # 4096 x 2048 cells
                        ((4096, 2048), (1024, 64)),     # 4097 x 2049 cells
                        ((8191, 8191), (1024, 64)),     # 8192 x 8192 cells
                        ((8192, 8191), (256, 128))):    # 8193 x 8192 cells
        with World(WrachConfig(dimensions=dims, cell_size=1)) as w:
            t = w.tuning()
            assert (t["scan_threads"], t["scan_items"]) == shape, dims


def test_autotune_keeps_the_state(oracle):
    """wrach_autotune times every candidate on the live state and restores
    it after each trial: the world continues as if it had not run."""
    wl = workloads.small_uniform(seed=6)
    say("[tuning] autotune")
    P, ref = oracle_world(oracle, wl.config, wl.pos, wl.vel)
    with World(wl.config) as w:
        w.add_particles(wl.pos, wl.vel)
        w.step(4)
        got = w.autotune(3)
        assert (got["scan_threads"], got["scan_items"]) in SCAN_SHAPES
        assert got["scatter_per_lane"] in (1, 2, 4) and got["cells_grid"] >= 1 and got["frame_us"] > 0
        assert {k: v for k, v in w.tuning().items() if k != "frame_us"} == \
            {k: v for k, v in got.items() if k != "frame_us"}
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=4), "after autotune")
        w.step(6)
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=10), "6 frames after autotune")
        with pytest.raises(WrachError) as e:
            w.set_tuning(scan_threads=256)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError):
            w.set_tuning(scatter_per_lane=3)
        with pytest.raises(WrachError):
            w.set_tuning(cells_grid=0)


def _batch_worlds(cs, W, H, seed):
    """Worlds of uneven particle counts: uniform, empty, one particle,
    crowded cells (a few hundred particles in three cells), and a small one."""
    r = np.random.default_rng(seed)
    out = []
    for n in (300, 0, 1, 0, 57):
        out.append((f32(r.uniform(0, [W, H], (n, 2))), f32(r.uniform(-1, 1, (n, 2)))))
    crowd = np.repeat(f32([[10.5, 5.5], [30.2, 20.7], [47.9, 0.3]]), [260, 400, 90], 0)
    crowd += f32(r.uniform(-0.4, 0.4, crowd.shape))
    out[3] = (f32(np.clip(crowd, 0, [W, H])), f32(r.uniform(-1, 1, crowd.shape)))
    return out


@pytest.mark.parametrize("cs", [1, 3])
def test_batch_of_worlds_matches_each_world_alone(oracle, cs):
    """wrach_create_batch: worlds of uneven particle counts stepped together
    over one flattened key space; each world bit for bit against the
    single-world oracle, after the upload, after frames, after an append to
    one world."""
    W, H = 48, 32
    say(f"[batch] 5 worlds {W}x{H} cs={cs}")
    cfg = WrachConfig(dimensions=(W, H), cell_size=cs, gravity=(0.0, -0.1))
    worlds = _batch_worlds(cs, W, H, 60 + cs)
    with World(cfg, worlds=len(worlds)) as b:
        for w, (p, v) in enumerate(worlds):
            b.add_particles(p, v, world=w)
        refs = [oracle_world(oracle, cfg, p, v) for p, v in worlds]
        for w, (P, ref) in enumerate(refs):
            assert_same_state(b.read_world(w), ref, f"world {w} upload")
        b.step(21)
        for w, (P, ref) in enumerate(refs):
            assert_same_state(b.read_world(w), oracle.step(P, *ref, nsteps=21), f"world {w}, 21 frames")
        # an append re-packs the batch from its store (state.rs:90-101, per world)
        extra = f32([[5.5, 5.5], [6.5, 5.5], [40.25, 30.75]])
        b.add_particles(extra, f32(np.zeros_like(extra)), world=1)
        worlds[1] = (np.concatenate([worlds[1][0], extra]), np.concatenate([worlds[1][1], np.zeros_like(extra)]))
        refs = [oracle_world(oracle, cfg, p, v) for p, v in worlds]
        b.step(5)
        for w, (P, ref) in enumerate(refs):
            assert_same_state(b.read_world(w), oracle.step(P, *ref, nsteps=5), f"world {w} after the append")
        assert b.read()[0][-1] == sum(int(ref[0][-1]) for _, ref in refs)  # the batch's offsets


def test_many_worlds_batch_full_size(oracle):
    """The 512-world workload's 169 worlds of side 256 (11.2M cells, 2.65M
    particles) as one batch: the first, a middle and the last world bit for
    bit against the single-world oracle after 10 frames; the batch's
    particle total against the sum of the worlds'."""
    ws = [w for w in workloads.many_worlds() if w.config.dimensions[0] == 256]
    say(f"[batch] many-worlds side 256: {len(ws)} worlds")
    with World(ws[0].config, worlds=len(ws)) as b:
        for k, w in enumerate(ws):
            b.add_particles(w.pos, w.vel, world=k)
        assert b.n_particles == sum(w.n for w in ws)
        b.step(10)
        for k in (0, len(ws) // 2, len(ws) - 1):
            P, ref = oracle_world(oracle, ws[k].config, ws[k].pos, ws[k].vel)
            assert_same_state(b.read_world(k), oracle.step(P, *ref, nsteps=10, nthreads=THREADS),
                              f"world {k} of {len(ws)}")


@pytest.mark.parametrize("side", [512, 1024])
def test_many_worlds_batches_of_sides_512_and_1024(oracle, side):
    """The other two batches of the 512-world workload at default launch
    shapes (no autotune): 184 worlds of side 512 (48.4M cells) and 159 of
    side 1024 (167M cells).  The first, a middle and the last world of each
    bit for bit against the single-world oracle after 10 frames, and the
    batch's particle total."""
    ws = [w for w in workloads.many_worlds() if w.config.dimensions[0] == side]
    say(f"[batch] many-worlds side {side}: {len(ws)} worlds, {sum(w.n for w in ws)} particles")
    with World(ws[0].config, worlds=len(ws)) as b:
        for k, w in enumerate(ws):
            b.add_particles(w.pos, w.vel, world=k)
        assert b.n_particles == sum(w.n for w in ws)
        b.step(10)
        b.sync()
        say(f"[batch] side {side}: 10 frames done")
        for k in (0, len(ws) // 2, len(ws) - 1):
            P, ref = oracle_world(oracle, ws[k].config, ws[k].pos, ws[k].vel)
            assert_same_state(b.read_world(k), oracle.step(P, *ref, nsteps=10, nthreads=THREADS),
                              f"world {k} of {len(ws)}, side {side}")
        assert b.read()[0][-1] == sum(w.n for w in ws)


def _mixed_worlds(cs, seed):
    """Worlds of different sizes (sides 48x32, 100x100, 17x65, 256x40 and a
    1x1), of uneven particle counts, crowded cells among them."""
    r = np.random.default_rng(seed)
    dims = [(48, 32), (100, 100), (17, 65), (256, 40), (1, 1)]
    counts = [300, 2500, 0, 900, 3]
    out = []
    for (W, H), n in zip(dims, counts):
        p = r.uniform(0, [W, H], (n, 2))
        if n >= 900:  # a crowded cell or two
            p[: n // 4] = np.array([W / 3, H / 2]) + r.uniform(-0.4, 0.4, (n // 4, 2))
        out.append((WrachConfig(dimensions=(W, H), cell_size=cs, gravity=(0.0, -0.1)),
                    f32(np.clip(p, 0, [W, H])), f32(r.uniform(-1, 1, (n, 2)))))
    return out


@pytest.mark.parametrize("cs", [1, 3])
def test_worlds_of_different_sizes_match_each_world_alone(oracle, cs):
    """wrach_create_worlds: worlds of different sizes in one key space (each
    grid on a 1,024-cell boundary), stepped together; each world bit for bit
    against the single-world oracle after the upload, after frames and after
    an append, and its settings (wrach_settings_world)."""
    ws = _mixed_worlds(cs, 90 + cs)
    say(f"[mixed] {len(ws)} worlds of different sizes, cs={cs}")
    with World([c for c, _, _ in ws]) as b:
        assert b.worlds == len(ws)
        for w, (c, p, v) in enumerate(ws):
            b.add_particles(p, v, world=w)
        refs = [oracle_world(oracle, c, p, v) for c, p, v in ws]
        for w, ((c, p, _), (P, ref)) in enumerate(zip(ws, refs)):
            s_ = b.world_settings(w)
            gx, gy = c.dimensions[0] // cs + 1, c.dimensions[1] // cs + 1
            assert (s_.grid_dimensions[0], s_.grid_dimensions[1]) == (gx, gy)
            assert tuple(s_.view_dimensions) == tuple(float(d) for d in c.dimensions)
            assert s_.particles_in_frame_count == len(p)
            assert_same_state(b.read_world(w), ref, f"world {w} upload")
        for frames in (1, 16, 7):
            b.step(frames)
            refs = [(P, oracle.step(P, *ref, nsteps=frames)) for P, ref in refs]
            for w, (P, ref) in enumerate(refs):
                assert_same_state(b.read_world(w), ref, f"world {w} after +{frames}")
        extra = f32([[5.5, 5.5], [6.5, 5.5], [12.25, 30.75]])
        c1, p1, v1 = ws[1]
        b.add_particles(extra, f32(np.zeros_like(extra)), world=1)
        ws[1] = (c1, np.concatenate([p1, extra]), np.concatenate([v1, np.zeros_like(extra)]))
        refs = [oracle_world(oracle, c, p, v) for c, p, v in ws]
        b.step(5)
        for w, (P, ref) in enumerate(refs):
            assert_same_state(b.read_world(w), oracle.step(P, *ref, nsteps=5), f"world {w} after the append")
        assert b.read()[0][-1] == sum(len(p) for _, p, _ in ws)


def test_worlds_of_different_sizes_mode_s(oracle):
    """Mode S passes over a batch of worlds of different sizes: every
    world's 3x3 cells clipped to its own grid, each world bit for bit against
    the oracle's passes alone."""
    ws = _mixed_worlds(1, 95)
    say("[mixed] mode S, 2 passes")
    with World([c for c, _, _ in ws]) as b:
        for w, (c, p, v) in enumerate(ws):
            b.add_particles(p, v, world=w)
        b.set_solver("S", 2)
        refs = [oracle_world(oracle, c, p, v) for c, p, v in ws]
        b.step(4)
        for w, (P, ref) in enumerate(refs):
            assert_same_state(b.read_world(w), oracle.step_s(P, *ref, 2, nsteps=4)[:3], f"world {w}, mode S")


def test_worlds_of_different_sizes_argument_errors(oracle):
    say("[mixed] errors")
    a = WrachConfig(dimensions=(16, 16), cell_size=1)
    with pytest.raises(WrachError) as e:
        World([a, WrachConfig(dimensions=(32, 8), cell_size=2)])
    assert e.value.code == _lib.WRACH_E_INVALID
    with pytest.raises(WrachError) as e:
        World([a, WrachConfig(dimensions=(32, 8), cell_size=1, gravity=(0.0, -0.1))])
    assert e.value.code == _lib.WRACH_E_INVALID
    with World([a, WrachConfig(dimensions=(32, 8), cell_size=1)]) as b:
        with pytest.raises(WrachError) as e:
            b.world_settings(2)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError) as e:
            b.add_particles(f32([[1.5, 1.5]]), f32([[0.0, 0.0]]), world=2)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError):
            b.add_particles(f32([[1.5, 1.5]]), f32([[0.0, 0.0]]))
    # entries of one size: the equal-world batch
    with World([a, a, a]) as b:
        assert b.worlds == 3 and b.n_indices == 3 * 17 * 17 + 2


def test_many_worlds_in_one_batch(oracle):
    """The 512-world workload (sides 256, 512 and 1024) as one batch of
    worlds of different sizes: one launch per frame kernel.  Two worlds of
    each side bit for bit against the single-world oracle after 10 and 100
    frames, and the batch's particle total."""
    ws = workloads.many_worlds()
    say(f"[mixed] many-worlds: {len(ws)} worlds in one batch, {sum(w.n for w in ws)} particles")
    with World([w.config for w in ws]) as b:
        for k, w in enumerate(ws):
            b.add_particles(w.pos, w.vel, world=k)
        assert b.n_particles == sum(w.n for w in ws)
        picks = []
        for side in (256, 512, 1024):
            ks = [k for k, w in enumerate(ws) if w.config.dimensions[0] == side]
            picks += [ks[0], ks[-1]]
        refs = {k: oracle_world(oracle, ws[k].config, ws[k].pos, ws[k].vel) for k in picks}
        for frames in (10, 90):
            b.step(frames)
            b.sync()
            say(f"[mixed] many-worlds: +{frames} frames done")
            for k in picks:
                P, ref = refs[k]
                refs[k] = (P, oracle.step(P, *ref, nsteps=frames, nthreads=THREADS))
                assert_same_state(b.read_world(k), refs[k][1], f"world {k}, side {ws[k].config.dimensions[0]}")
        assert b.read()[0][-1] == sum(w.n for w in ws)


def test_batch_argument_errors(oracle):
    say("[batch] errors")
    cfg = WrachConfig(dimensions=(16, 16), cell_size=1)
    with World(cfg, worlds=3) as b:
        p = f32([[1.5, 1.5]])
        with pytest.raises(WrachError) as e:
            b.add_particles(p, p)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError) as e:
            b.add_particles(p, p, world=3)
        assert e.value.code == _lib.WRACH_E_INVALID
        with pytest.raises(WrachError):
            b.store_write_back()
        with pytest.raises(WrachError):
            b.load_packed(np.zeros(3 * 17 * 17 + 2, np.uint32), np.zeros((0, 2)), np.zeros((0, 2)))
        b.add_particles(p, p, world=2)
        assert b.read_world(2)[1].shape == (1, 2) and b.read_world(0)[1].shape == (0, 2)
    with pytest.raises(WrachError) as e:
        World(WrachConfig(dimensions=(16, 16), cell_size=1, host_pack=True), worlds=2)
    assert e.value.code == _lib.WRACH_E_INVALID


def test_graph_replay_matches_plain_launches(oracle):
    """Replays of the captured frame graph, with eager work and readbacks in
    between, against plain stream launches."""
    wl = workloads.small_uniform(seed=3)
    say("[graph] 16 + 1 frame graphs vs plain launches")
    with World(wl.config) as a, World(wl.config) as b:
        a.add_particles(wl.pos, wl.vel)
        b.add_particles(wl.pos, wl.vel)
        b.set_graph(False)
        for k in (1, 16, 3, 16, 16, 2, 33):
            a.step(k)
            b.step(k)
            sa, sb = a.read(), b.read()
            assert_same_state(sa, sb, f"graph after +{k}")


def test_profiled_step_matches_and_times(oracle):
    wl = workloads.small_uniform(seed=4)
    say("[profile] step_profiled")
    P, ref = oracle_world(oracle, wl.config, wl.pos, wl.vel)
    with World(wl.config) as w:
        w.add_particles(wl.pos, wl.vel)
        ms = w.step_profiled(5)
        assert set(ms) == {"collide", "cells", "scan", "scatter"}
        assert all(v > 0 for v in ms.values())
        assert_same_state(w.read(), oracle.step(P, *ref, nsteps=5), "profiled")


def test_device_views(oracle):
    wl = workloads.small_uniform(seed=5)
    with World(wl.config) as w:
        w.add_particles(wl.pos, wl.vel)
        w.step(3)
        v = w.device_state()
        c = w.device_state(canonical=True)
        assert v.n_particles == c.n_particles == wl.n
        assert v.n_indices == c.n_indices == wl.cells + 2
        assert v.positions and c.positions and v.stream


# ------------------------------------------------------ full-size workloads

def test_lattice_1m_full_size(oracle):
    """1,000,000 particles, 2048x1024, cs 1, gravity: exact for 12 frames."""
    wl = workloads.lattice_1m()
    run_parity(oracle, wl.config, wl.pos, wl.vel, 12, every=4, label=wl.name, threads=THREADS)


def test_lattice_1m_rollout_and_invariants(oracle):
    """The workload's 1,000 frames, past the collapse onto the floor (about
    2,050 cells of ~460 particles from frame ~500 on): still exact against
    the oracle at frames 300, 800 and 1,000, plus the size-independent
    invariants at the end (particle count, sorted offsets, every particle in
    its cell, inside the walls, |v| <= 1)."""
    wl = workloads.lattice_1m()
    say(f"[rollout] {wl.name} 1000 frames")
    P, ref = oracle_world(oracle, wl.config, wl.pos, wl.vel)
    with World(wl.config) as w:
        w.add_particles(wl.pos, wl.vel)
        w.step(300)
        ref = oracle.step(P, *ref, nsteps=300, nthreads=THREADS)
        assert_same_state(w.read(), ref, "lattice 300")
        w.step(500)
        ref = oracle.step(P, *ref, nsteps=500, nthreads=THREADS)
        assert_same_state(w.read(), ref, "lattice 800")
        w.step(200)
        idx, pos, vel = w.read()
    ref = oracle.step(P, *ref, nsteps=200, nthreads=THREADS)
    assert_same_state((idx, pos, vel), ref, "lattice 1000")
    W, H = wl.config.dimensions
    gx = W + 1
    assert idx[0] == 0 and idx[1] == 0 and idx[-1] == wl.n
    assert np.all(np.diff(idx.astype(np.int64)) >= 0)
    keys = np.floor(pos[:, 1]).astype(np.int64) * gx + np.floor(pos[:, 0]).astype(np.int64)
    cell_of = np.repeat(np.arange(len(idx) - 2), np.diff(idx[1:].astype(np.int64)))
    assert np.array_equal(keys, cell_of)
    assert pos.min() >= 0 and pos[:, 0].max() <= W and pos[:, 1].max() <= H
    assert np.abs(vel).max() <= 1.0


def test_blobs_4m_full_size(oracle):
    """4,000,000 particles in 64 Gaussian blobs, 4096x4096 (16.8M cells, more
    than the reference's 4,194,304-cell scan limit): exact for 3 frames."""
    wl = workloads.gaussian_blobs()
    run_parity(oracle, wl.config, wl.pos, wl.vel, 3, every=1, label=wl.name, threads=THREADS)


def test_blobs_4m_rollout_500(oracle):
    """The 4,000,000-particle blobs workload for 500 frames: binning, offsets,
    in-cell order and state exact against the oracle at frames 250 and 500,
    and the invariants at the end (every particle in its cell, offsets
    sorted, inside the walls, |v| <= 1)."""
    wl = workloads.gaussian_blobs()
    say(f"[rollout] {wl.name} 500 frames")
    P, ref = oracle_world(oracle, wl.config, wl.pos, wl.vel)
    with World(wl.config) as w:
        w.add_particles(wl.pos, wl.vel)
        for upto in (250, 500):
            w.step(250)
            ref = oracle.step(P, *ref, nsteps=250, nthreads=THREADS)
            say(f"[rollout] {wl.name} frame {upto}")
            assert_same_state(w.read(), ref, f"blobs frame {upto}")
        idx, pos, vel = w.read()
    W, H = wl.config.dimensions
    gx = W // wl.config.cell_size + 1
    assert idx[0] == 0 and idx[1] == 0 and idx[-1] == len(pos)
    assert np.all(np.diff(idx.astype(np.int64)) >= 0)
    cs = wl.config.cell_size
    keys = (np.floor(pos[:, 1] / cs).astype(np.int64) * gx + np.floor(pos[:, 0] / cs).astype(np.int64))
    cell_of = np.repeat(np.arange(len(idx) - 2), np.diff(idx[1:].astype(np.int64)))
    assert np.array_equal(keys, cell_of)
    assert pos.min() >= 0 and pos[:, 0].max() <= W and pos[:, 1].max() <= H
    assert np.abs(vel).max() <= 1.0
