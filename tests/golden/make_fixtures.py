"""Writes tests/golden/rollout_small.npz: a 20-frame rollout of a small world
computed by the oracle, kept as a regression fixture (it pins the oracle's
output over time, and GPU tests compare against it too).

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as o  # noqa: E402


def main():
    r = np.random.Generator(np.random.PCG64(2024))
    w, h, cs, n, steps = 48, 32, 1, 600, 20
    pos = r.uniform(0, [w, h / 2], (n, 2)).astype(np.float32)
    pos[:40] = pos[40:80]  # some coincident pairs
    vel = r.uniform(-1, 1, (n, 2)).astype(np.float32)
    idx, p, v = o.pack(pos, vel, (0, 0, w, h), cs)
    P = o.params(w, h, cs, len(p), gravity=(0.0, -0.1))
    I, PP, VV = [], [], []
    ci, cp, cv = idx, p, v
    for _ in range(steps):
        ci, cp, cv = o.step(P, ci, cp, cv)
        I.append(ci)
        PP.append(cp)
        VV.append(cv)
    meta = dict(width=w, height=h, cell_size=cs, gravity=[0.0, -0.1], steps=steps, seed=2024)
    np.savez_compressed(os.path.join(HERE, "rollout_small.npz"), idx0=idx, pos0=p, vel0=v,
                        idx=np.stack(I), pos=np.stack(PP), vel=np.stack(VV),
                        meta=np.array(json.dumps(meta)))


if __name__ == "__main__":
    main()
