"""The learned warm start's oracle (oracle/learn.py, wo_learn_features),
pinned on the CPU by checks independent of its own code:

- the features against a numpy brute force over all pairs (Chebyshev cell
  distance <= 2, a lexsort by the stated key), on worlds with crowded cells,
  coincident particles and exact distance ties, cell sizes 1 and 2;
- the features' overlap sums against wo_stencil_pass's;
- invariance of the features under a permutation inside each cell;
- the network's gradient (float64, no bf16) against torch's autograd and
  against central differences;
- Adam against torch.optim.Adam;
- bf16 rounding against torch's float32 -> bfloat16 conversion.
"""
import numpy as np
import pytest
import torch

from oracle import learn as ol
from oracle import oracle as o


def packed_world(n, W, H, cs, seed, crowd=0.0, dup=0):
    r = np.random.Generator(np.random.PCG64(seed))
    pos = np.stack([r.uniform(0, W, n), r.uniform(0, H, n)], 1).astype(np.float32)
    if crowd:
        k = int(crowd * n)
        pos[:k] = (np.array([W / 2, H / 2]) + r.normal(0, 0.7, (k, 2))).astype(np.float32)
    if dup:  # exact duplicates of positions (ties in distance), some with equal velocities
        pos[-dup:] = pos[:dup]
    vel = r.uniform(-1, 1, (n, 2)).astype(np.float32)
    if dup:
        vel[-dup // 2:] = vel[:dup // 2]
    idx, p, v = o.pack(pos, vel, (0, 0, W, H), cs)
    P = o.params(W, H, cs, len(p))
    return P, idx, p, v


def brute_features(P, idx, pos, vel, cs):
    n = len(pos)
    gx = P.grid_dimensions[0]
    cell = np.repeat(np.arange(len(idx) - 2), np.diff(idx[1:].astype(np.int64)))
    cx, cy = cell % gx, cell // gx
    rows = np.zeros((n, 64), np.uint16)
    for i in range(n):
        near = (np.abs(cx - cx[i]) <= 2) & (np.abs(cy - cy[i]) <= 2)
        near[i] = False
        if near.sum() > ol_budget():
            continue
        j = np.flatnonzero(near)
        dx = (pos[j, 0] - pos[i, 0]).astype(np.float32)
        dy = (pos[j, 1] - pos[i, 1]).astype(np.float32)
        sq = (dx * dx).astype(np.float32)
        sq = (sq + (dy * dy).astype(np.float32)).astype(np.float32)
        keys = [vel[j, 1].view(np.uint32), vel[j, 0].view(np.uint32), pos[j, 1].view(np.uint32),
                pos[j, 0].view(np.uint32), sq.view(np.uint32)]
        order = np.lexsort(keys)[:16]
        jj = j[order]
        f = np.stack([pos[jj, 0] - pos[i, 0], pos[jj, 1] - pos[i, 1],
                      vel[jj, 0] - vel[i, 0], vel[jj, 1] - vel[i, 1]], 1).astype(np.float32)
        rows[i, :4 * len(jj)] = ol.bf16_bits(f.ravel())
    return rows


def ol_budget():
    return 4096


@pytest.mark.parametrize("cs,crowd,dup", [(1, 0.0, 0), (1, 0.3, 40), (2, 0.1, 20)])
def test_features_match_brute_force(cs, crowd, dup):
    P, idx, p, v = packed_world(1500, 48, 32, cs, seed=cs * 7 + dup, crowd=crowd, dup=dup)
    rows, _ = ol.features(P, idx, p, v)
    np.testing.assert_array_equal(rows, brute_features(P, idx, p, v, cs))


def test_features_sums_equal_the_pass_sums():
    P, idx, p, v = packed_world(3000, 64, 32, 1, seed=5, crowd=0.2, dup=10)
    _, sums = ol.features(P, idx, p, v)
    assert sums == o.stencil_overlap_sums(P, idx, p, p)
    assert sums[1] > 0


def test_features_do_not_depend_on_order_inside_cells():
    P, idx, p, v = packed_world(2000, 40, 40, 1, seed=9, crowd=0.4, dup=30)
    rows, sums = ol.features(P, idx, p, v)
    r = np.random.Generator(np.random.PCG64(1))
    perm = np.arange(len(p))
    for c in range(len(idx) - 2):
        s, e = idx[c + 1], idx[c + 2]
        perm[s:e] = s + r.permutation(e - s)
    rows2, sums2 = ol.features(P, idx, p[perm], v[perm])
    np.testing.assert_array_equal(rows2, rows[perm])
    assert sums2 == sums


def test_features_budget_gives_zero_rows():
    # 5000 particles inside one cell's 5x5 neighbourhood: over the budget.
    r = np.random.Generator(np.random.PCG64(3))
    pos = (np.array([10.5, 10.5]) + r.uniform(-2.4, 2.4, (5000, 2))).astype(np.float32)
    vel = np.zeros_like(pos)
    idx, p, v = o.pack(pos, vel, (0, 0, 32, 32), 1)
    P = o.params(32, 32, 1, len(p))
    rows, _ = ol.features(P, idx, p, v)
    gx = P.grid_dimensions[0]
    cell = np.repeat(np.arange(len(idx) - 2), np.diff(idx[1:].astype(np.int64)))
    centre = cell == 10 * gx + 10  # its 5x5 cells hold all 5000 particles
    assert centre.sum() > 0 and (rows[centre] == 0).all()
    corner = cell == 8 * gx + 8    # its 5x5 cells hold about 9/25 of them
    assert corner.sum() > 0 and (rows[corner] != 0).any(axis=1).all()


def _torch_loss(vec, x, t):
    W1 = vec[:8192].reshape(64, 128)
    b1 = vec[8192:8320]
    W2 = vec[8320:24704].reshape(128, 128)
    b2 = vec[24704:24832]
    W3 = vec[24832:25088].reshape(128, 2)
    b3 = vec[25088:25090]
    h = torch.nn.functional.silu(x @ W1 + b1)
    h = torch.nn.functional.silu(h @ W2 + b2)
    y = h @ W3 + b3
    return ((y - t) ** 2).sum() / (2 * len(x))


def _rand_problem(seed, n=128):
    r = np.random.Generator(np.random.PCG64(seed))
    from wrach_amd.learn import init_params
    params = init_params(seed)
    params[8192:8320] = r.normal(0, 0.1, 128)  # non-zero biases exercise their gradient
    params[24704:24832] = r.normal(0, 0.1, 128)
    params[25088:] = r.normal(0, 0.1, 2)
    rows = ol.bf16_bits(r.normal(0, 0.6, (n, 64)).astype(np.float32))
    t = r.normal(0, 0.3, (n, 2))
    return params, rows, t


def test_gradient_formulas_match_autograd():
    params, rows, t = _rand_problem(1)
    g, loss = ol.gradient(params, rows, t, exact=True)
    vec = torch.tensor(params.astype(np.float64), requires_grad=True)
    x = torch.tensor(ol.bf16_value(rows).astype(np.float64))
    tl = _torch_loss(vec, x, torch.tensor(t))
    tl.backward()
    np.testing.assert_allclose(loss, tl.item(), rtol=1e-12)
    np.testing.assert_allclose(g, vec.grad.numpy(), rtol=1e-9, atol=1e-13)


def test_gradient_matches_central_differences():
    params, rows, t = _rand_problem(2, n=64)
    g, _ = ol.gradient(params, rows, t, exact=True)
    r = np.random.Generator(np.random.PCG64(4))
    for k in r.choice(len(params), 24, replace=False):
        e = np.zeros(len(params))
        e[k] = 1e-6
        lp = ol.gradient(params.astype(np.float64) + e, rows, t, exact=True)[1]
        lm = ol.gradient(params.astype(np.float64) - e, rows, t, exact=True)[1]
        assert abs((lp - lm) / 2e-6 - g[k]) <= 1e-6 + 1e-5 * abs(g[k]), k


def test_float32_gradient_is_the_rounded_version():
    """The device's roundings move the gradient by bf16-sized amounts only."""
    params, rows, t = _rand_problem(3)
    g32, l32 = ol.gradient(params, rows, t)
    g64, l64 = ol.gradient(params, rows, t, exact=True)
    cos = float(g32 @ g64 / np.linalg.norm(g32) / np.linalg.norm(g64))
    assert cos > 0.999
    assert abs(l32 - l64) <= 2e-2 * l64


def test_adam_matches_torch():
    r = np.random.Generator(np.random.PCG64(6))
    w = r.normal(0, 1, 1000).astype(np.float32)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    tw = torch.tensor(w.astype(np.float64), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        g = r.normal(0, 1, 1000).astype(np.float32) * (10.0 ** -step)
        w, m, v = ol.adam(w, m, v, g, step, 1e-3)
        tw.grad = torch.tensor(g.astype(np.float64))
        opt.step()
        np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-6, atol=1e-8)


def test_bf16_rounding_matches_torch():
    r = np.random.Generator(np.random.PCG64(8))
    x = np.concatenate([r.normal(0, 1, 10000), r.normal(0, 1e-30, 1000), [0.0, -0.0, 1.0, 65504.0,
                        np.float32(1.00390625), np.float32(1.01171875), 3.4e38]]).astype(np.float32)
    mine = ol.bf16_bits(x)
    ref = torch.tensor(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(mine, ref)
