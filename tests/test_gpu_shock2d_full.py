"""The two-dimensional He/CO2 shock / interface case at its full size,
1024 x 512 (two species, NSCBC inflow and outflow, art on, cfl 0.4):

  * FP64, 200 run! steps against the oracle's committed fixture
    tests/golden/shock2d_1024x512_200.json (tests/golden/make_shock2d.py):
    every dt to 1e-12, kinetic energy and enstrophy to 1e-10, the probed state
    to 1e-10 normwise per component;
  * the FP32-mixed form against the FP64 GPU run after the case's full 2000
    steps: kinetic-energy relative difference (the bar is 1e-5), printed with
    the enstrophy and state differences."""
import json
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _solver(ctx, precision):
    n = (1024, 512, 1)
    q, L, faces = cases.shock_interface(n)
    return cases.gpu_solver(ctx, n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=((1, 1), (0, 0), (0, 0)),
                            nscbc=faces, precision=precision)


def test_shock2d_full_size_200_steps_fp64(ctx):
    ref = json.load(open(os.path.join(HERE, "golden", "shock2d_1024x512_200.json")))
    print("case shock2d 1024x512 FP64, 200 steps", flush=True)
    gs = _solver(ctx, "fp64")
    ke0, en0 = gs.diagnostics()
    # the same initial state, summed in another order over 524288 points
    e_ke0 = abs(ke0 / ref["ke0"] - 1)
    dts = gs.run(1e9, ref["steps"], record=ref["steps"])
    err_dt = np.abs(dts - np.array(ref["dts"])).max() / max(ref["dts"])
    ke, en = gs.diagnostics()
    nc = len(ref["Q_probes"][0])
    Qi = [gs.field(f"Q[{c + 1}]") for c in range(nc)]
    scale = [max(abs(v[c]) for v in ref["Q_probes"]) for c in range(nc)]
    mom = max(scale[nc - 4:nc - 1])
    scale[nc - 4:nc - 1] = [mom] * 3
    err_q = max(abs(Qi[c][tuple(p)] - vals[c]) / scale[c] for p, vals in zip(ref["probes"], ref["Q_probes"])
                for c in range(nc))
    print(f"  initial KE {e_ke0:.2e}; after {ref['steps']} steps: dt {err_dt:.2e}, t {abs(gs.t / ref['t'] - 1):.2e}, "
          f"KE {abs(ke / ref['ke'] - 1):.2e}, enstrophy {abs(en / ref['enstrophy'] - 1):.2e}, probed state {err_q:.2e}")
    assert e_ke0 <= 1e-11
    assert err_dt <= 1e-12
    assert abs(gs.t - ref["t"]) <= 1e-12 * ref["t"]
    assert abs(ke - ref["ke"]) <= 1e-10 * ref["ke"]
    assert abs(en - ref["enstrophy"]) <= 1e-10 * ref["enstrophy"]
    assert err_q <= 1e-10
    gs.close()


def test_shock2d_full_run_fp32_mixed_error(ctx):
    print("case shock2d 1024x512, 2000 steps, FP64 and FP32-mixed", flush=True)
    out = {}
    for prec in ("fp64", "fp32"):
        gs = _solver(ctx, prec)
        dts = gs.run(1e9, 2000, record=2000)
        ke, en = gs.diagnostics()
        out[prec] = (dts, ke, en, gs.t, [gs.field(f"Q[{c + 1}]") for c in range(6)])
        gs.close()
    d64, ke64, en64, t64, q64 = out["fp64"]
    d32, ke32, en32, t32, q32 = out["fp32"]
    eke = abs(ke32 / ke64 - 1)
    een = abs(en32 / en64 - 1)
    # per component, the three momentum components on one scale (w is zero in 2-D)
    scale = [np.abs(b).max() for b in q64]
    scale[2:5] = [max(scale[2:5])] * 3
    eq = max(np.abs(a - b).max() / s for a, b, s in zip(q32, q64, scale))
    edt = np.abs(d32 - d64).max() / d64.max()
    print(f"  after 2000 steps (t = {t64:.6e} / {t32:.6e}): KE {eke:.2e}, enstrophy {een:.2e}, dt {edt:.2e}, "
          f"state {eq:.2e}")
    assert np.all(np.isfinite(q32[0])) and eke <= 1e-5
