"""Writes tests/golden/shock2d_1024x512_200.json: the oracle's two-dimensional
He/CO2 shock / interface case at its full size, 1024 x 512 (Mach-1.5 shock,
single-mode interface of amplitude 0.05 lambda smeared over 3 cells, two
species, NSCBC inflow at low x and outflow at high x, periodic y, Cook
artificial properties on, cfl 0.4, filter every step; the setup of
tests/cases.py shock_interface) after 200 run! steps: the dt history,
kinetic energy, enstrophy, and the conserved state at a fixed set of probe
points.

Run from the repository root:  python tests/golden/make_shock2d.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
import cases  # noqa: E402

N, STEPS = (1024, 512, 1), 200


def probes(n, count=64, seed=5):
    rng = np.random.default_rng(seed)
    pts = [(int(rng.integers(0, n[0])), int(rng.integers(0, n[1])), 0) for _ in range(count)]
    # and points through the interface and next to both NSCBC faces
    pts += [(0, 7, 0), (1, 300, 0), (n[0] - 1, 11, 0), (n[0] - 2, 400, 0), (256, 0, 0), (256, 255, 0)]
    return pts


def main(out=os.path.join(ROOT, "tests", "golden", "shock2d_1024x512_200.json")):
    q, L, faces = cases.shock_interface(N)
    bcs = ((1, 1), (0, 0), (0, 0))
    codes = cases.nscbc_codes(bcs, faces)
    s = cases.oracle_solver(N, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=codes, nscbc=faces)
    ke0, en0 = s.diagnostics()
    t0 = time.time()
    dts = s.run(1e9, STEPS, record=STEPS)
    wall = time.time() - t0
    nc = s.n_cons
    Qi = [s.interior(s.Q[..., c]) for c in range(nc)]
    ke, en = s.diagnostics()
    pts = probes(N)
    rec = {
        "case": "shock2d", "steps": STEPS, "n": list(N), "cfl": 0.4, "art": True,
        "t": s.time[0], "dts": [float(x) for x in dts],
        "ke0": ke0, "enstrophy0": en0, "ke": ke, "enstrophy": en,
        "probes": pts, "Q_probes": [[float(Qi[c][p]) for c in range(nc)] for p in pts],
        "Q_sum": [float(Qi[c].sum()) for c in range(nc)],
        "Q_abs_sum": [float(np.abs(Qi[c]).sum()) for c in range(nc)],
        "oracle_wall_s": wall,
    }
    with open(out, "w") as fh:
        json.dump(rec, fh, indent=1)
    print(f"wrote {out}: ke={ke:.15e} enstrophy={en:.15e} wall={wall:.1f}s")


if __name__ == "__main__":
    main()
