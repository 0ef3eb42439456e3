"""The movable view (wrach_set_view_anchor) on the HIP path against the
oracle's model of the reference's store (needs a GPU).

The model follows ParticleStore (runners/bevy/src/particle_store.rs:14-27,
54-103) with the write-back this library adds: the store is a list of
particles in insertion order; a view packs the store's particles that lie in
its cells, in store order (oracle.pack with the view as the viewport,
spatial_bin.rs:68-149); frames step the packed state (oracle.step with the
view's anchor: walls and cell keys follow it).  Moving the view writes the
simulated particles back -- the store becomes them, in packed order,
followed by the store's particles outside the old view, in their order --
and packs the new view.
"""
import numpy as np
import pytest

from wrach_amd import World, WrachConfig, WrachError
from wrach_amd import _lib
from tests.test_gpu_parity import assert_same_state, f32, say

pytestmark = pytest.mark.gpu


class StoreModel:
    def __init__(self, oracle, cfg):
        self.o = oracle
        self.cfg = cfg
        self.anchor = (0.0, 0.0)
        self.pos = np.zeros((0, 2), np.float32)
        self.vel = np.zeros((0, 2), np.float32)

    def viewport(self):
        (ax, ay), (w, h) = self.anchor, self.cfg.dimensions
        return (ax, ay, ax + w, ay + h)

    def inside(self):
        """The packer's test: the particle's cell (div_euclid per axis) lies
        in the view's active grid."""
        cs = float(self.cfg.cell_size)
        bl, grid = self.o.active_grid(self.viewport(), cs)
        out = np.zeros(len(self.pos), bool)
        for i, (x, y) in enumerate(self.pos):
            cx, cy = self.o.cell_coord(float(x), float(y), cs)
            out[i] = bl[0] <= cx < bl[0] + grid[0] and bl[1] <= cy < bl[1] + grid[1]
        return out

    def add(self, pos, vel):
        self.pos = np.concatenate([self.pos, f32(pos)])
        self.vel = np.concatenate([self.vel, f32(vel)])

    def pack(self):
        w, h = self.cfg.dimensions
        idx, p, v = self.o.pack(self.pos, self.vel, self.viewport(), float(self.cfg.cell_size))
        P = self.o.params(w, h, self.cfg.cell_size, len(p), gravity=self.cfg.gravity, dt=self.cfg.dt,
                          anchor=self.anchor)
        return P, (idx, p, v)

    def move(self, state, anchor):
        keep = ~self.inside()
        self.pos = np.concatenate([state[1], self.pos[keep]])
        self.vel = np.concatenate([state[2], self.vel[keep]])
        self.anchor = (float(anchor[0]), float(anchor[1]))


def spread(n, xmax, ymax, seed):
    rng = np.random.default_rng(seed)
    pos = f32(np.stack([rng.uniform(0, xmax, n), rng.uniform(0, ymax, n)], 1))
    vel = f32(rng.uniform(-1, 1, (n, 2)))
    return pos, vel


@pytest.mark.parametrize("cs", [1, 2])
def test_view_moves_over_the_store(oracle, cs):
    """A 64x48 view over particles spread across 256x60: frames, moves to
    four anchors (including back to the origin and one off the floor), an
    upload after a move, and frames after each."""
    cfg = WrachConfig(dimensions=(64, 48), cell_size=cs, gravity=(0.0, -0.1))
    pos, vel = spread(2500, 256, 60, 30 + cs)
    say(f"[view] cs={cs}: {len(pos)} particles over 256x60, view 64x48")
    model = StoreModel(oracle, cfg)
    model.add(pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        P, ref = model.pack()
        assert_same_state(w.read(), ref, "first view")
        w.step(5)
        ref = oracle.step(P, *ref, nsteps=5)
        assert_same_state(w.read(), ref, "first view, 5 frames")
        for k, anchor in enumerate([(64, 0), (130, 4), (0, 0), (32, 12)]):
            model.move(ref, anchor)
            P, ref = model.pack()
            w.set_view_anchor(*anchor)
            assert tuple(w.settings().view_anchor) == (float(anchor[0]), float(anchor[1]))
            assert_same_state(w.read(), ref, f"view at {anchor}")
            if k == 1:
                extra, evel = spread(300, 256, 60, 40 + cs)
                model.add(extra, evel)
                P, ref = model.pack()
                w.add_particles(extra, evel)
                assert_same_state(w.read(), ref, f"upload at {anchor}")
            w.step(4)
            ref = oracle.step(P, *ref, nsteps=4)
            assert_same_state(w.read(), ref, f"view at {anchor}, 4 frames")


def test_view_errors_and_capacity(oracle):
    cfg = WrachConfig(dimensions=(64, 48), cell_size=2)
    say("[view] errors")
    with World(cfg) as w:
        for xy in [(3.0, 0.0), (0.0, 1.0), (-2.0, 0.0), (float("nan"), 0.0), (16777216.0, 0.0)]:
            with pytest.raises(WrachError) as e:
                w.set_view_anchor(*xy)
            assert e.value.code == _lib.WRACH_E_INVALID, xy
        # A crowd beyond the view, more than the view's capacity (33 x 25
        # cells x 4 x 1.1 = 3,630): moving there fails and changes nothing.
        rng = np.random.default_rng(50)
        crowd = f32(np.stack([rng.uniform(130, 190, 4000), rng.uniform(0, 48, 4000)], 1))
        near, nvel = spread(500, 60, 48, 51)
        w.add_particles(np.concatenate([near, crowd]), np.concatenate([nvel, np.zeros_like(crowd)]))
        w.step(3)
        before = w.read()
        with pytest.raises(WrachError) as e:
            w.set_view_anchor(128.0, 0.0)
        assert e.value.code == _lib.WRACH_E_CAPACITY
        assert tuple(w.settings().view_anchor) == (0.0, 0.0)
        after = w.read()
        assert all(np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                  b.view(np.uint32) if b.dtype == np.float32 else b)
                   for a, b in zip(before, after))
        w.step(2)
        w.set_view_anchor(0.0, 0.0)  # where it is: nothing to do
    with World(cfg, worlds=2) as b:
        with pytest.raises(WrachError) as e:
            b.set_view_anchor(64.0, 0.0)
        assert e.value.code == _lib.WRACH_E_INVALID
    with World(WrachConfig(dimensions=(64, 48), cell_size=2, host_pack=True)) as h:
        with pytest.raises(WrachError) as e:
            h.set_view_anchor(64.0, 0.0)
        assert e.value.code == _lib.WRACH_E_INVALID


def test_view_with_mode_s(oracle):
    """Mode S on a moved view: the stencil's cells are the view's (keys
    relative to the anchor); frames bit for bit against the store model
    stepped by the oracle's mode S."""
    cfg = WrachConfig(dimensions=(64, 48), cell_size=1, gravity=(0.0, -0.1))
    pos, vel = spread(3000, 200, 50, 60)
    say("[view] mode S on a moved view")
    model = StoreModel(oracle, cfg)
    model.add(pos, vel)
    with World(cfg) as w:
        w.add_particles(pos, vel)
        w.set_solver("S", 2)
        P, ref = model.pack()
        w.step(3)
        *ref, _ = oracle.step_s(P, *ref, max_passes=2, nsteps=3)
        assert_same_state(w.read(), tuple(ref), "origin, mode S")
        for anchor in [(70, 0), (120, 2)]:
            model.move(ref, anchor)
            P, ref = model.pack()
            w.set_view_anchor(*anchor)
            w.step(4)
            *ref, st = oracle.step_s(P, *ref, max_passes=2, nsteps=4)
            assert_same_state(w.read(), tuple(ref), f"view at {anchor}, mode S")
            assert w.solver_stats() == st
