"""run! issued as replays of a captured per-step graph, with the dt decision
(predicted_dt, check_step floors, tfinal clip) in device memory, against the
same solver launching every kernel from the host: the dt sequence and the
state must be bitwise identical, including when a run is split into several
calls with host work between the replays, when it ends on tfinal, with NSCBC
faces, in the FP32-mixed form, and when a floor stops it."""
import time

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _tgv(ctx, n=(32, 32, 32), precision="fp64", **kw):
    L = (TWO_PI,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    return cases.gpu_solver(ctx, n, L, q, art=True, precision=precision, **kw)


def _state(s):
    Q = s.allocate_state()
    s.download(Q)
    return Q


@pytest.mark.parametrize("mode", ["step_graph", "device_dt"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_modes_equal_host(ctx, precision, mode):
    a, b = _tgv(ctx, precision=precision), _tgv(ctx, precision=precision)
    a.set_run_mode(mode)
    b.set_run_mode("host_dt")
    da = a.run(1e9, 20, record=20)
    db = b.run(1e9, 20, record=20)
    assert np.array_equal(da, db)
    assert a.t == b.t and a.step_count == b.step_count == 20
    assert np.array_equal(_state(a), _state(b))


def test_graph_split_runs_with_host_work_between(ctx):
    a, b = _tgv(ctx), _tgv(ctx)
    b.set_run_mode("host_dt")
    d1 = a.run(1e9, 7, record=7)
    Q1 = _state(a)                      # host work between replays
    d2 = a.run(1e9, 15, record=8)
    ke_mid = a.diagnostics()
    d3 = a.run(1e9, 20, record=5)
    db = b.run(1e9, 20, record=20)
    assert np.array_equal(np.concatenate([d1, d2, d3]), db)
    assert np.array_equal(_state(a), _state(b))
    assert np.isfinite(Q1).all() and np.isfinite(ke_mid).all()


def test_graph_ends_on_tfinal(ctx):
    a, b = _tgv(ctx), _tgv(ctx)
    b.set_run_mode("host_dt")
    a.run(0.0123, 10 ** 6)
    b.run(0.0123, 10 ** 6)
    assert a.t == b.t == 0.0123 or (a.t == b.t and abs(a.t - 0.0123) < 1e-15)
    assert a.step_count == b.step_count
    assert np.array_equal(_state(a), _state(b))


def test_graph_nscbc(ctx):
    n = (96, 32, 1)
    q, L, faces = cases.shock_interface(n)
    bcs = ((1, 1), (0, 0), (0, 0))
    sol = [cases.gpu_solver(ctx, n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=bcs, nscbc=faces)
           for _ in range(2)]
    sol[1].set_run_mode("host_dt")
    d = [s.run(1e9, 25, record=25) for s in sol]
    assert np.array_equal(d[0], d[1])
    assert np.array_equal(_state(sol[0]), _state(sol[1]))


def test_graph_floor_failure_matches_eager(ctx):
    """A dt_min floor above every possible dt stops run! at its first step with
    the same reason, step count and state in both modes."""
    import compactles_hip as ch
    res = []
    for mode in ("step_graph", "device_dt", "host_dt"):
        prob = ch.Problem(domain=((0.0, TWO_PI),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3,
                          ic=lambda X, Y, Z: ch.Prim(u=(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0 * X),
                                                     p=71.43, rho=1.0))
        num = ch.Numerics(n_global=(32, 32, 32), control=ch.StepControl(dt_min=1.0))
        s, _ = ch.setup(prob, num, ctx)
        s.set_run_mode(mode)
        with pytest.raises(ch.SolverFailure) as e:
            s.run(1e9, 10)
        res.append((e.value.reason, s.step_count, s.t, _state(s)))
    for r in res:
        assert r[0] == "dt_min" and r[1] == 0 and r[2] == 0.0
        assert np.array_equal(r[3], res[0][3])


def test_run_modes_timing_small_grids(ctx):
    """Time per step of the three run modes at 32^3 and 64^3 (printed with
    pytest -s)."""
    for n in (32, 64):
        out = {}
        for mode in ("step_graph", "device_dt", "host_dt"):
            s = _tgv(ctx, n=(n, n, n))
            s.set_run_mode(mode)
            s.run(1e9, 3)
            t0 = time.perf_counter()
            s.run(1e9, 43)
            out[mode] = (time.perf_counter() - t0) / 40 * 1e3
        print(f"\nTGV {n}^3 ms/step: " + ", ".join(f"{k} {v:.3f}" for k, v in out.items()))
        assert out["step_graph"] < 1.5 * out["host_dt"]

