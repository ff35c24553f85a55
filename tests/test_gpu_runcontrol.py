"""run!'s outer-loop controls on the GPU (cles_run_ctl): StepControl recovery
and the callbacks' soft landing / stop / rewind, against the oracle running
the same triggers, and against the GPU's own composed run.  Bar: dt history
<= 1e-12 relative, state <= 1e-10; bitwise where the same kernels run the same
steps."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi
N16, L3 = (16, 16, 16), (TWO_PI,) * 3


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _q():
    X, Y, Z = cases.grid(N16, L3)
    return cases.conserved(cases.tgv_prim(X, Y, Z))


def _gpu(ctx, cfl, **ctl):
    import compactles_hip as ch
    return cases.gpu_solver(ctx, N16, L3, _q(), art=True, cfl=cfl, control=ch.StepControl(**ctl))


def _orc(cfl, **ctl):
    return cases.oracle_solver(N16, L3, _q(), art=True, cfl=cfl, **ctl)


def _state(s):
    Q = s.allocate_state()
    s.download(Q)
    return Q


def _interior_err(Qg, os_):
    """Normwise per component; the three momenta share one scale (rho w of
    the TGV state is zero up to rounding)."""
    refs = [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)]
    mom = max(np.abs(refs[c]).max() for c in (1, 2, 3))
    errs = []
    for c, ref in enumerate(refs):
        got = Qg[4:-4, 4:-4, 4:-4, c]
        scale = mom if c in (1, 2, 3) else np.abs(ref).max()
        errs.append(np.abs(got - ref).max() / scale)
    return max(errs)


def test_rollback_on_gpu(ctx):
    import compactles_hip as ch
    ctl = dict(retries=2, cfl_backoff=0.5, savepoint_interval=10)
    a = _gpu(ctx, 6.0, **ctl)
    da = a.run(1e9, 60, record=70)
    assert a.retries_used == 1 and a.cfl == 3.0 and a.step_count == 60
    # the oracle takes the same decisions.  The blow-up amplifies rounding
    # exponentially (the two dt histories part by 1e-7 at step 15), so the
    # failing step itself may differ by one and only the decisions compare.
    o = _orc(6.0, **ctl)
    o.run(1e9, 60)
    assert o.retries_used == 1 and o.cfl == 3.0 and o.time[1] == 60
    # the GPU's own composed run: on to the failure in runs of ten steps,
    # keeping the state at every savepoint step (the multiples of 10); the
    # failure leaves the same lagged artificial coefficients; restore the last
    # savepoint by hand, lower the CFL.  The step the blow-up lands on depends
    # on the last bits of every step before it, so it is found, not assumed.
    b = _gpu(ctx, 6.0)
    saved, db1 = {}, []
    with pytest.raises(ch.SolverFailure) as e:
        while b.step_count < 60:
            saved[b.step_count] = (_state(b), (b.t, b.step_count))
            db1.append(b.run(1e9, b.step_count + 10, record=10))
    fail = b.step_count
    sp = fail // 10 * 10
    assert e.value.reason == "negative_density" and fail > 10
    Qs, ts = saved[sp]
    b.upload(Qs)
    b.set_time(*ts)
    b.cfl = 3.0
    db2 = b.run(1e9, 60, record=60 - sp)
    assert np.array_equal(da[:sp], np.concatenate(db1[:sp // 10]))
    assert np.array_equal(da[fail:fail + 60 - sp], db2)
    assert np.array_equal(_state(a), _state(b))


def test_retries_exhausted_on_gpu(ctx):
    import compactles_hip as ch
    s = _gpu(ctx, 12.0, retries=1, cfl_backoff=0.9, savepoint_interval=10)
    with pytest.raises(ch.SolverFailure) as e:
        s.run(1e9, 60)
    assert e.value.reason == "negative_density" and s.retries_used == 1 and s.cfl == 12.0 * 0.9


def _triggers(log):
    import compactles_hip as ch
    return (ch.Callback(ch.AtTime([0.3, 0.71]), lambda s_, Q: log.append(("at", s_.clock.step, s_.clock.t))),
            ch.Callback(ch.EveryTime(0.25), lambda s_, Q: log.append(("every", s_.clock.step, s_.clock.t))),
            ch.Callback(ch.EveryStep(7), lambda s_, Q: log.append(("step", s_.clock.step))),
            ch.Callback(ch.EveryStep(1), lambda s_, Q: s_.clock.t >= 1.0))


def test_callbacks_match_oracle(ctx):
    from compactles_hip.callbacks import HookAdapter
    lg, lo = [], []
    g = _gpu(ctx, 0.5)
    dg = g.run(5.0, 10 ** 6, record=500, callback=_triggers(lg))
    o = _orc(0.5)
    ad = HookAdapter(_triggers(lo), o, o.Q, 0.0, 0)
    do = o.run(5.0, 10 ** 6, record=500, hooks=C.byref(ad.hooks))
    ad.check()
    assert g.step_count == o.time[1] and len(dg) == len(do)
    assert np.abs(dg - do).max() <= 1e-12 * do.max()
    assert [e[:2] for e in lg] == [e[:2] for e in lo]
    for e in lg:
        if e[0] in ("at", "every"):
            m = min((0.3, 0.71) if e[0] == "at" else (0.25 * round(e[2] / 0.25),), key=lambda x: abs(x - e[2]))
            assert abs(e[2] - m) <= 8 * np.spacing(max(m, 1.0)), e
    assert 1.0 <= g.t < 1.0 + dg.max()
    assert _interior_err(_state(g), o) <= 1e-10


def test_callbacks_take_the_host_loop_in_device_modes(ctx):
    la, lb = [], []
    a = _gpu(ctx, 0.5)
    da = a.run(5.0, 10 ** 6, record=500, callback=_triggers(la))
    b = _gpu(ctx, 0.5)
    b.set_run_mode("step_graph")
    db = b.run(5.0, 10 ** 6, record=500, callback=_triggers(lb))
    assert np.array_equal(da, db) and la == lb
    assert np.array_equal(_state(a), _state(b))


def test_effect_reads_the_device_state(ctx):
    import compactles_hip as ch
    seen = []
    s = _gpu(ctx, 0.5)

    def effect(solver, Q):
        q = np.asarray(Q)
        seen.append((solver.clock.step, float(q[4:-4, 4:-4, 4:-4, 0].sum())))
    s.run(1e9, 6, callback=ch.Callback(ch.EveryStep(3), effect))
    ref = _gpu(ctx, 0.5)
    ref.run(1e9, 3)
    Q3 = _state(ref)
    assert [k for k, _ in seen] == [3, 6]
    assert seen[0][1] == float(Q3[4:-4, 4:-4, 4:-4, 0].sum())
