"""Writes tests/golden/tgv64_100.json: the oracle's Taylor-Green vortex at 64^3
(README.md:33-44 quick-start state, periodic [0,2pi)^3, one species gamma=1.4,
rho=1, p=71.43, C6 derivative, C8 filter every step, Cook artificial
properties on, cfl 0.5) after 100 run! steps: the dt history, kinetic energy,
enstrophy, and the conserved state at a fixed set of probe points.

Run from the repository root:  python tests/golden/make_tgv64.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
import cases  # noqa: E402

N, STEPS = 64, 100


def probes(n, count=64, seed=4):
    rng = np.random.default_rng(seed)
    return [tuple(int(x) for x in rng.integers(0, n, 3)) for _ in range(count)]


def main(out=os.path.join(ROOT, "tests", "golden", "tgv64_100.json")):
    n, L = (N, N, N), (2 * np.pi,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    s = cases.oracle_solver(n, L, q, art=True, cfl=0.5)
    ke0, en0 = s.diagnostics()
    t0 = time.time()
    dts = s.run(1e9, STEPS, record=STEPS)
    wall = time.time() - t0
    Qi = [s.interior(s.Q[..., c]) for c in range(5)]
    ke, en = s.diagnostics()
    pts = probes(N)
    rec = {
        "case": "tgv64", "steps": STEPS, "n": N, "cfl": 0.5, "art": True, "p0": 71.43,
        "t": s.time[0], "dts": [float(x) for x in dts],
        "ke0": ke0, "enstrophy0": en0, "ke": ke, "enstrophy": en,
        "probes": pts, "Q_probes": [[float(Qi[c][p]) for c in range(5)] for p in pts],
        "Q_sum": [float(Qi[c].sum()) for c in range(5)],
        "Q_abs_sum": [float(np.abs(Qi[c]).sum()) for c in range(5)],
        "oracle_wall_s": wall,
    }
    with open(out, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(f"wrote {out}: ke={ke:.15e} enstrophy={en:.15e} wall={wall:.1f}s")


if __name__ == "__main__":
    main()
