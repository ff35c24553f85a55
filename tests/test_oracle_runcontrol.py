"""run!'s outer-loop controls in the oracle (CPU): StepControl recovery
(timestep.jl:286-351) and the callbacks' soft landing, stop requests and
rewind (timestep.jl:352-399, callbacks.jl), driven through the frontend's
trigger classes.  What pins them:
 - a rolled-back run ends bitwise equal to the run composed by hand from its
   pieces (run to the savepoint at the first CFL, lower the CFL, run on);
 - landed instants are hit to 8 eps and no landing step is below half the
   CFL step (the soft landing's property, timestep.jl:371-379);
 - trigger semantics follow callbacks.jl (each AtTime instant once per
   trajectory, EveryStep counts, a True effect stops the run, an instant
   between the savepoint and a failure fires again after the rollback)."""
import math

import numpy as np
import pytest

import cases

TWO_PI = 2 * np.pi
N16, L3 = (16, 16, 16), (TWO_PI,) * 3


def _tgv(**kw):
    X, Y, Z = cases.grid(N16, L3)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    return cases.oracle_solver(N16, L3, q, art=True, **kw)


def _hooks(cb, s):
    import ctypes as C

    from compactles_hip.callbacks import HookAdapter
    t, st = s.time
    ad = HookAdapter(cb, s, s.Q, t, st)
    return ad, C.byref(ad.hooks)


def test_rollback_equals_composed_run():
    # cfl 6 loses positivity entering step 25; the savepoint is step 20.  The
    # savepoint holds Q, t and step only (timestep.jl:286, stepcontrol.jl
    # Savepoint): the lagged artificial coefficients that max_rate reads stay
    # those of the failed trajectory, so the composed run reaches the failure
    # too before restoring the state by hand.
    a = _tgv(cfl=6.0, retries=2, cfl_backoff=0.5, savepoint_interval=10)
    da = a.run(1e9, 60, record=70)
    assert a.retries_used == 1 and a.cfl == 3.0 and a.time[1] == 60
    b = _tgv(cfl=6.0)
    db1 = b.run(1e9, 20, record=20)
    Q20, t20 = b.Q.copy(), b.time
    with pytest.raises(RuntimeError, match="code 2"):
        b.run(1e9, 60, record=5)
    assert b.time[1] == 25
    b.Q[...] = Q20
    b.set_time(*t20)
    b.cfl = 3.0
    db2 = b.run(1e9, 60, record=40)
    # A took steps 0..24 at cfl 6, then steps 20..59 again at cfl 3
    assert np.array_equal(da[:20], db1) and np.array_equal(da[25:65], db2)
    assert np.array_equal(a.Q, b.Q) and a.time == b.time


def test_no_retries_left_fails():
    s = _tgv(cfl=12.0, retries=1, cfl_backoff=0.9, savepoint_interval=10)
    with pytest.raises(RuntimeError, match="code 2"):
        s.run(1e9, 60)
    assert s.retries_used == 1 and s.cfl == 12.0 * 0.9
    s = _tgv(cfl=6.0)
    with pytest.raises(RuntimeError, match="code 2"):
        s.run(1e9, 60)
    assert s.time[1] == 25


def test_at_time_soft_landing():
    import compactles_hip as ch
    s0 = _tgv()
    d0 = s0.run(1e9, 12, record=12)
    marks = [0.6 * float(np.cumsum(d0)[3]), float(np.cumsum(d0)[8]) + 0.35 * d0[9]]
    hits = []
    cb = ch.Callback(ch.AtTime(marks), lambda solver, Q: hits.append(solver.clock.t))
    s = _tgv()
    ad, hooks = _hooks(cb, s)
    d = s.run(marks[1] + 0.5, 1000, record=1000, hooks=hooks)
    ad.check()
    assert len(hits) == 2
    for h, m in zip(hits, marks):
        assert abs(h - m) <= 8 * np.spacing(max(abs(m), 1.0))
    # no landing step is below half of the unclipped CFL step it replaced
    t = np.concatenate([[0.0], np.cumsum(d)])
    for m in marks:
        k = int(np.argmin(np.abs(t - m)))   # the step that landed ends at t[k]
        assert d[k - 1] >= 0.4 * np.median(d0)


def test_every_time_and_every_step_and_stop():
    import compactles_hip as ch
    at, steps = [], []
    cbs = (ch.Callback(ch.EveryTime(0.05), lambda s_, Q: at.append(s_.clock.t)),
           ch.Callback(ch.EveryStep(3), lambda s_, Q: steps.append(s_.clock.step)),
           ch.Callback(ch.EveryStep(10), lambda s_, Q: s_.clock.step >= 20))
    s = _tgv()
    ad, hooks = _hooks(cbs, s)
    s.run(1e9, 1000, hooks=hooks)
    ad.check()
    assert s.time[1] == 20                       # stopped by the third callback at step 20
    assert steps == list(range(3, 21, 3))
    assert len(at) >= 2
    for k, t in enumerate(at):
        assert abs(t - 0.05 * (k + 1)) <= 8 * np.spacing(max(t, 1.0))


def test_rewind_refires_instant_after_rollback():
    import compactles_hip as ch
    probe = _tgv(cfl=6.0)
    d = probe.run(1e9, 24, record=24)
    t = np.concatenate([[0.0], np.cumsum(d)])
    mark = 0.5 * (t[20] + t[21])                 # between the savepoint (20) and the failure
    hits = []
    cb = ch.Callback(ch.AtTime(mark), lambda s_, Q: hits.append((s_.clock.t, s_.clock.step)))
    s = _tgv(cfl=6.0, retries=2, cfl_backoff=0.5, savepoint_interval=10)
    ad, hooks = _hooks(cb, s)
    s.run(1e9, 40, hooks=hooks)
    ad.check()
    assert s.retries_used == 1 and len(ad.rewinds) == 1 and ad.rewinds[0][1] == 20
    assert len(hits) == 2 and hits[0][1] != hits[1][1]
    for h, _ in hits:
        assert abs(h - mark) <= 8 * np.spacing(max(mark, 1.0))


def test_when_state_and_bare_callable():
    import compactles_hip as ch
    seen = []
    s = _tgv()
    rho0 = float(np.asarray(s.Q)[..., 0].max())
    cond = ch.WhenState(lambda solver, Q: solver.clock.step >= 4 and np.asarray(Q)[..., 0].max() != rho0)
    cbs = [ch.Callback(cond, lambda s_, Q: seen.append(s_.clock.step)), lambda s_, Q: True]
    ad, hooks = _hooks(cbs, s)
    s.run(1e9, 8, hooks=hooks)
    ad.check()
    assert seen == [4] and s.time[1] == 8        # once=True; the bare callable's True does not stop


def test_effect_exception_is_raised_after_the_run():
    import compactles_hip as ch

    def boom(solver, Q):
        raise ValueError("effect failed")
    s = _tgv()
    ad, hooks = _hooks(ch.Callback(ch.EveryStep(2), boom), s)
    s.run(1e9, 10, hooks=hooks)
    assert s.time[1] == 2
    with pytest.raises(ValueError, match="effect failed"):
        ad.check()


def test_landing_steps_validation():
    import compactles_hip as ch
    with pytest.raises(ValueError, match="landing_steps"):
        ch.StepControl(landing_steps=0)
    assert math.isinf(ch.AtTime([]).next_time(None))
