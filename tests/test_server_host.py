"""Host logic of the serving loop (no GPU is touched): admission refuses a
request that can never fit the HBM budget, and a request whose run fails
(here: no such device) is reported as failed with its message instead of
hanging the server."""
import numpy as np
import pytest

import cases


def _tgv_spec(ch, n):
    prob = ch.Problem(domain=((0.0, 2 * np.pi),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3,
                      ic=lambda X, Y, Z: ch.Prim(u=(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0 * X),
                                                 p=71.43, rho=1.0))
    return prob, ch.Numerics(n_global=(n, n, n))


def test_footprint_grows_with_grid_and_precision():
    import compactles_hip as ch
    prob, num = _tgv_spec(ch, 32)
    a = ch.SolverSpec(prob, num).footprint()
    prob, num = _tgv_spec(ch, 64)
    b = ch.SolverSpec(prob, num).footprint()
    assert 7.5 < b / a < 8.1
    num32 = ch.Numerics(n_global=(64, 64, 64), precision="fp32")
    c = ch.SolverSpec(prob, num32).footprint()
    assert c < b
    # the TGV case's allocation table: Q, du, dQ (15), primitives and Y (9), per dimension three velocity
    # gradients, a temperature gradient and five fluxes (27), mu*, beta*, kappa*, the mu*/beta* and
    # kappa* sensors, strain, e (7): 58 FP64 fields of 64^3, plus a few per cent of plan tables at most
    assert 58 * 8 * 64 ** 3 <= b <= 1.05 * 58 * 8 * 64 ** 3


def test_server_rejects_request_beyond_budget():
    import compactles_hip as ch
    sv = ch.Server(device=999, hbm_budget=1 << 20, runners=1)
    prob, num = _tgv_spec(ch, 32)
    with pytest.raises(ch.ClesError) as e:
        sv.submit(prob, num, tfinal=1.0)
    assert "budget" in str(e.value)
    sv.close()


def test_server_reports_failed_run():
    import compactles_hip as ch
    sv = ch.Server(device=999, hbm_budget=64 << 30, runners=2)
    prob, num = _tgv_spec(ch, 16)
    ids = [sv.submit(prob, num, tfinal=1.0, nmax=3) for _ in range(3)]
    for rid in ids:
        assert sv.wait(rid, timeout=60.0) == ch.server.FAILED
        Q, status, t, steps = sv.result(rid)
        assert status != 0 and steps == 0
    used, running, queued = sv.stats()
    assert used == 0 and running == 0 and queued == 0
    sv.close()
