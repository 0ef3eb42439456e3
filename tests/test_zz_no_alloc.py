"""wrach_frame_noalloc (the allocation-free frame) against wrach_step +
wrach_read of a context loaded with the same packed triple, through the C
ABI, bit for bit: two calls in a row on one GPU, the second with a smaller
world and fewer particles, so that the second starts from what the first
left in the library's scratch array.  Also: a frame too large for the
scratch array is refused before any work.

The test allocates the caller's device buffers itself (hipMalloc through
ctypes); the entry point allocates nothing.
"""
import ctypes
import sys

import numpy as np
import pytest

from wrach_amd import World, WrachConfig, WrachError, _lib
from wrach_amd._lib import check

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        self.h = h

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        if out.nbytes:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def packed(oracle, n, W, H, cs, seed):
    r = np.random.Generator(np.random.PCG64(seed))
    pos = np.stack([r.uniform(0, W, n), r.uniform(0, H / 2, n)], 1).astype(np.float32)
    pos[: n // 10] = (np.array([W / 3, H / 4]) + r.normal(0, 1.5, (n // 10, 2))).astype(np.float32)
    vel = r.uniform(-1, 1, (n, 2)).astype(np.float32)
    return oracle.pack(pos, vel, (0, 0, W, H), cs)


def noalloc_frame(hip, cfg, idx, pos, vel):
    L = _lib.load()
    W, H = cfg.dimensions
    cs = cfg.cell_size
    ws = _lib.WorldSettingsC()
    ws.view_dimensions[:] = [float(W), float(H)]
    ws.view_anchor[:] = [0.0, 0.0]
    ws.grid_dimensions[:] = [W // cs + 1, H // cs + 1]
    ws.cell_size = cs
    ws.particles_in_frame_count = len(pos)
    g = (ctypes.c_float * 2)(*cfg.gravity)
    di, dp, dv = hip.upload(idx), hip.upload(pos), hip.upload(vel)
    try:
        check(L.wrach_frame_noalloc(ctypes.byref(ws), g, ctypes.c_float(cfg.dt), di, dp, dv, len(pos), None))
        assert hip.h.hipDeviceSynchronize() == 0
        return hip.download(di, idx), hip.download(dp, pos), hip.download(dv, vel)
    finally:
        hip.free(di, dp, dv)


def reference_frame(cfg, idx, pos, vel):
    with World(cfg) as w:
        w.load_packed(idx, pos, vel)
        w.step(1)
        return w.read()


def same(a, b, what):
    for x, y, name in zip(a, b, ("indices", "positions", "velocities")):
        assert x.shape == y.shape, f"{what}: {name} shape"
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{what}: {name} differ"


def test_noalloc_frame_matches_step_and_read(oracle):
    hip = Hip()
    cases = [
        ("large", WrachConfig(dimensions=(512, 512), cell_size=1, gravity=(0.0, -0.1)), 60_000, 1),
        ("smaller", WrachConfig(dimensions=(200, 96), cell_size=3, gravity=(0.0, -0.1)), 5_000, 2),
    ]
    for name, cfg, n, seed in cases:
        W, H = cfg.dimensions
        idx, pos, vel = packed(oracle, n, W, H, cfg.cell_size, seed)
        say(f"[noalloc] {name}: {len(pos)} particles, {W}x{H}, cs {cfg.cell_size}")
        got = noalloc_frame(hip, cfg, idx, pos, vel)
        want = reference_frame(cfg, idx, pos, vel)
        same(got, want, name)
        P = oracle.params(W, H, cfg.cell_size, len(pos), gravity=cfg.gravity)
        same(got, oracle.step(P, idx, pos, vel), f"{name} vs the oracle")


def test_noalloc_refuses_a_frame_beyond_its_scratch(oracle):
    hip = Hip()
    cfg = WrachConfig(dimensions=(4096, 4096), cell_size=1)  # 16.8M cells: > 32 MiB of offsets
    idx = np.zeros(4097 * 4097 + 2, np.uint32)
    with pytest.raises(WrachError) as e:
        noalloc_frame(hip, cfg, idx, np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32))
    assert e.value.code == _lib.WRACH_E_CAPACITY
