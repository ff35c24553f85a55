"""Taylor-Green vortex 64^3, 100 run! steps on the GPU (FP64) against the
oracle's committed fixture tests/golden/tgv64_100.json (written by
tests/golden/make_tgv64.py).  Bar: 1e-10 relative on kinetic energy and
enstrophy, 1e-12 on every dt, 1e-10 normwise on the probed state."""
import json
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_tgv64_100_steps_fp64():
    import compactles_hip as ch
    ref = json.load(open(os.path.join(HERE, "golden", "tgv64_100.json")))
    n, L = (64, 64, 64), (2 * np.pi,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    ctx = ch.Context(0)
    gs = cases.gpu_solver(ctx, n, L, q, art=True, cfl=0.5)
    ke0, en0 = gs.diagnostics()
    assert abs(ke0 - ref["ke0"]) <= 1e-13 * ref["ke0"]
    assert abs(en0 - ref["enstrophy0"]) <= 1e-12 * ref["enstrophy0"]
    dts = gs.run(1e9, ref["steps"], record=ref["steps"])
    assert np.abs(dts - np.array(ref["dts"])).max() <= 1e-12 * max(ref["dts"])
    assert abs(gs.t - ref["t"]) <= 1e-12 * ref["t"]
    Qh = gs.allocate_state()
    gs.download(Qh)
    Qi = [Qh[4:-4, 4:-4, 4:-4, c] for c in range(5)]
    scale = [max(abs(v[c]) for v in ref["Q_probes"]) for c in range(5)]
    # momentum components share one scale (rho*w of this flow is round-off noise)
    mom = max(scale[1:4])
    scale = [scale[0], mom, mom, mom, scale[4]]
    for p, vals in zip(ref["probes"], ref["Q_probes"]):
        for c in range(5):
            assert abs(Qi[c][tuple(p)] - vals[c]) <= 1e-10 * scale[c]
    ke, en = gs.diagnostics()
    assert abs(ke - ref["ke"]) <= 1e-10 * ref["ke"]
    assert abs(en - ref["enstrophy"]) <= 1e-10 * ref["enstrophy"]
