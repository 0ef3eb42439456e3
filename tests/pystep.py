"""Independent brute-force restatements used to pin the C oracle.

Test infrastructure.  Arithmetic is numpy float32 scalars, which round after
every operation and never fuse a*b+c; orderings use Python's stable sorted().
Only for small worlds.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def div_euclid(a: float, b: float) -> float:
    """Rust f32::div_euclid for b > 0 (floor division with a positive remainder)."""
    q = math.trunc(float(F(a) / F(b)))
    if math.fmod(float(F(a)), float(F(b))) < 0.0:
        q -= 1
    return q


def cell_coord(x, y, cs):
    return int(div_euclid(x, cs)), int(div_euclid(y, cs))


def pack(pos, vel, width, height, cs):
    """Store + packer: insertion order per cell, row-major cells, off-grid dropped."""
    gx, gy = int(div_euclid(width, cs)) + 1, int(div_euclid(height, cs)) + 1
    keyed = []
    for i, (p, v) in enumerate(zip(pos, vel)):
        cx, cy = cell_coord(p[0], p[1], cs)
        if 0 <= cx < gx and 0 <= cy < gy:
            keyed.append((cy * gx + cx, i))
    keyed = sorted(keyed, key=lambda t: t[0])  # stable: keeps insertion order
    counts = [0] * (gx * gy)
    for k, _ in keyed:
        counts[k] += 1
    idx = [0, 0]
    run = 0
    for c in counts:
        run += c
        idx.append(run)
    P = np.array([pos[i] for _, i in keyed], np.float32).reshape(-1, 2)
    V = np.array([vel[i] for _, i in keyed], np.float32).reshape(-1, 2)
    return np.array(idx, np.uint32), P, V


def collide_pair(p, i, j):
    dx = F(p[i][0]) - F(p[j][0])
    dy = F(p[i][1]) - F(p[j][1])
    d = F(np.sqrt(F(F(dx * dx) + F(dy * dy))))
    if d > F(1.0):
        return
    if d == F(0.0):
        d = F(0.0001)
    f = F(F(F(0.5) * F(F(1.0) - d)) / d)
    ex = F(F(p[j][0]) - F(p[i][0])) * f
    ey = F(F(p[j][1]) - F(p[i][1])) * f
    p[i][0] = F(p[i][0]) - F(ex)
    p[i][1] = F(p[i][1]) - F(ey)
    p[j][0] = F(p[j][0]) + F(ex)
    p[j][1] = F(p[j][1]) + F(ey)


def collide(points):
    p = [[F(a), F(b)] for a, b in points]
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            collide_pair(p, i, j)
    return np.array(p, np.float32).reshape(-1, 2)


def finish(pos, vel, width, height, gravity=(0.0, 0.0), dt=1.0):
    px, py = F(pos[0]), F(pos[1])
    vx, vy = F(vel[0]), F(vel[1])
    g0, g1, dt = F(gravity[0]), F(gravity[1]), F(dt)
    if g0 != 0:
        vx = F(vx + F(g0 * dt))
    if g1 != 0:
        vy = F(vy + F(g1 * dt))
    if dt == 1:
        px, py = F(px + vx), F(py + vy)
    else:
        px, py = F(px + F(vx * dt)), F(py + F(vy * dt))
    x1, y1 = F(width), F(height)
    if px > x1:
        px, vx = x1, F(vx * F(-1.0))
    if px < 0:
        px, vx = F(0.0), F(vx * F(-1.0))
    if py > y1:
        py, vy = y1, F(vy * F(-1.0))
    if py < 0:
        py, vy = F(0.0), F(vy * F(-1.0))
    vx = F(-1.0) if vx < -1 else vx
    vx = F(1.0) if vx > 1 else vx
    vy = F(-1.0) if vy < -1 else vy
    vy = F(1.0) if vy > 1 else vy
    return (px, py), (vx, vy)


def step(idx, pos, vel, width, height, cs, gravity=(0.0, 0.0), dt=1.0, cap=9):
    """One frame on a packed triple: per cell collide the first `cap`, finish
    all, then re-bin with a stable sort by cell key."""
    gx, gy = int(div_euclid(width, cs)) + 1, int(div_euclid(height, cs)) + 1
    n = len(pos)
    po = np.zeros((n, 2), np.float32)
    vo = np.zeros((n, 2), np.float32)
    for c in range(gx * gy):
        s, e = int(idx[c + 1]), int(idx[c + 2])
        m = min(e - s, cap)
        local = collide(pos[s:s + m]) if m else np.zeros((0, 2), np.float32)
        for i in range(s, e):
            q = local[i - s] if i - s < m else pos[i]
            (a, b), (u, w) = finish(q, vel[i], width, height, gravity, dt)
            po[i] = (a, b)
            vo[i] = (u, w)
    keys = []
    for i in range(n):
        fx = math.floor(float(F(po[i][0]) / F(cs)))
        fy = math.floor(float(F(po[i][1]) / F(cs)))
        assert 0 <= fx < gx and 0 <= fy < gy
        keys.append((fy * gx + fx, i))
    keys = sorted(keys, key=lambda t: t[0])
    counts = [0] * (gx * gy)
    for k, _ in keys:
        counts[k] += 1
    new_idx = [0, 0]
    run = 0
    for c in counts[:-1]:
        run += c
        new_idx.append(run)
    new_idx.append(n)
    order = [i for _, i in keys]
    return np.array(new_idx, np.uint32), po[order].copy(), vo[order].copy()
