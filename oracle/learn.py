"""CPU restatement of the learned warm start's arithmetic (an extension with
no counterpart in the reference; its separation is one Gauss-Seidel sweep
per cell, shaders/physics/src/particles.rs:62-93).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
network, its gradient and Adam; the product (wrach_amd) never imports it.

Each function states its arithmetic in numpy float32 with the roundings the
device applies (bf16 weights, bf16 hidden activations into the second layer,
bf16 dL/dz into the weight products, f32 everywhere else).  The device sums
in another order (MFMA), so comparisons use the tolerances written in the
tests.  With exact=True the same formulas run in float64 without any bf16
rounding: that version is pinned against torch's autograd (tests), which
pins the formulas; the float32 version differs from it only by the stated
roundings.

Parameter vector (f32): W1[64][128] (input-major), b1[128], W2[128][128],
b2[128], W3[128][2], b3[2].
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle as _o

FEAT, HID, OUT = 64, 128, 2
N_PARAMS = FEAT * HID + HID + HID * HID + HID + HID * OUT + OUT
ADAM_BETA1, ADAM_BETA2, ADAM_EPS = np.float32(0.9), np.float32(0.999), np.float32(1e-8)


# ------------------------------------------------------------------ bf16

def bf16_bits(x) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even (NaN kept quiet)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return np.where(nan, ((u >> 16) | 0x40).astype(np.uint16), r)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x) -> np.ndarray:
    return bf16_value(bf16_bits(x))


# ------------------------------------------------------------- features

def features(P, idx, pos, vel):
    """wo_learn_features: (rows [N, 64] bf16 bits, (overlap sum in 2^-24
    units, pairs)) of the packed state."""
    L = _o.lib()
    if not hasattr(L, "_learn_ready"):
        L.wo_learn_features.argtypes = [ctypes.POINTER(_o.Params), _o._u32p, _o._f32p, _o._f32p,
                                        ctypes.POINTER(ctypes.c_uint16), ctypes.POINTER(ctypes.c_uint64)]
        L.wo_learn_features.restype = None
        L._learn_ready = True
    idx = np.ascontiguousarray(idx, np.uint32)
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n = len(pos)
    rows = np.zeros((max(n, 1), FEAT), np.uint16)
    sums = np.zeros(2, np.uint64)
    L.wo_learn_features(ctypes.byref(P), _o._p(idx, _o._u32p), _o._p(pos, _o._f32p), _o._p(vel, _o._f32p),
                        rows.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                        sums.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return rows[:n], (int(sums[0]), int(sums[1]))


# --------------------------------------------------------------- network

def unpack(vec, dtype=np.float32) -> dict:
    v = np.asarray(vec, dtype)
    o = 0
    out = {}
    for name, shape in (("W1", (FEAT, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
                        ("W3", (HID, OUT)), ("b3", (OUT,))):
        k = int(np.prod(shape))
        out[name] = v[o:o + k].reshape(shape)
        o += k
    assert o == N_PARAMS
    return out


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[k]).ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def _sigmoid(z):
    one = z.dtype.type(1)
    return one / (one + np.exp(-z))


def forward(params, rows, exact: bool = False):
    """Outputs y [N, 2] and the intermediates (x, z1, h1, z2, a2)."""
    t = np.float64 if exact else np.float32
    p = unpack(params, t)
    rnd = (lambda a: a) if exact else (lambda a: bf16_round(a).astype(t))
    x = bf16_value(rows).astype(t)
    W1, W2, W3 = rnd(p["W1"]), rnd(p["W2"]), rnd(p["W3"])
    z1 = x @ W1 + p["b1"]
    h1 = rnd(z1 * _sigmoid(z1))          # the second layer's bf16 input
    z2 = h1 @ W2 + p["b2"]
    a2 = z2 * _sigmoid(z2)
    y = a2 @ W3 + p["b3"]
    return y.astype(t), (x, z1, h1, z2, a2, W1, W2, W3)


def gradient(params, rows, targets, exact: bool = False):
    """(gradient vector, loss) of sum_i |y_i - t_i|^2 / (2n)."""
    t = np.float64 if exact else np.float32
    rnd = (lambda a: a) if exact else (lambda a: bf16_round(a).astype(t))
    y, (x, z1, h1, z2, a2, W1, W2, W3) = forward(params, rows, exact)
    tg = np.asarray(targets, t).reshape(-1, 2)
    n = len(y)
    e = y - tg
    loss = (e * e).sum() / t(2 * n)
    dy = e * (t(1) / t(n))
    s2 = _sigmoid(z2)
    dz2 = (dy @ W3.T) * (s2 * (t(1) + z2 * (t(1) - s2)))
    dz2b = rnd(dz2)
    s1 = _sigmoid(z1)
    dz1 = (dz2b @ W2.T) * (s1 * (t(1) + z1 * (t(1) - s1)))
    dz1b = rnd(dz1)
    g = {"W1": x.T @ dz1b, "b1": dz1b.sum(0), "W2": h1.T @ dz2b, "b2": dz2b.sum(0),
         "W3": a2.T @ dy, "b3": dy.sum(0)}
    return pack(g).astype(t), float(loss)


def adam(params, m, v, grad, step: int, lr: float):
    """One Adam step in float32, each operation rounded in the order the
    device runs them; returns (params, m, v)."""
    f = np.float32
    b1, b2, eps = ADAM_BETA1, ADAM_BETA2, ADAM_EPS
    c1 = f(1.0 / (1.0 - float(b1) ** step))
    c2 = f(1.0 / (1.0 - float(b2) ** step))
    g = np.asarray(grad, f)
    m = b1 * np.asarray(m, f)
    m = m + (f(1) - b1) * g
    v = b2 * np.asarray(v, f)
    v = v + (f(1) - b2) * (g * g)
    mh = m * c1
    vh = v * c2
    stepv = f(lr) * mh / (np.sqrt(vh) + eps)
    return np.asarray(params, f) - stepv, m, v
