"""Stacks of same-shape two-dimensional problems (cles_stack_*): each instance
of a stack against the same problem run alone by a plain solver whose run!
decides dt on the device: dt history, time, step count and state bitwise
equal, in FP64 and FP32-mixed, on the periodic 2-D Taylor-Green case and on the
two-gas shock / interface case with NSCBC faces; instances with their own end
times and step caps; a stack of one."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _tgv2d(ch, n, precision, amp):
    L = (2 * np.pi, 2 * np.pi, 1.0)
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)),
                      bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3, ic=lambda X, Y, Z: None,
                      eos=ch.single_species(gamma=1.4), transport=ch.Transport(mu0=2e-3))
    num = ch.Numerics(n_global=n, cfl=0.5, precision=precision)
    spec = ch.SolverSpec(prob, num)
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z, amp=amp))
    Q = spec.allocate_state()
    sl = tuple(slice(spec.halo[k], spec.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    return prob, num, Q


def _shock(ch, n, precision, scale):
    from compactles_hip import workloads as W
    prob, num, Q = W.shock_problem(ch, n, precision)
    Q = Q.copy()
    Q[..., 2:5] *= scale   # the momentum components, so that every instance has its own dt sequence
    Q[..., 5] += 0.5 * (scale * scale - 1) * (Q[..., 2] ** 2 + Q[..., 3] ** 2) / (scale * scale) / \
        np.maximum(Q[..., 0] + Q[..., 1], 1e-300)
    return prob, num, Q


def _alone(ctx, prob, num, Q, tf, nmax):
    import compactles_hip as ch
    s = ch.Solver(ctx, prob, num)
    s.upload(Q)
    s.set_run_mode("device_dt")
    s.run(tf, nmax)
    out = s.allocate_state()
    s.download(out)
    t, steps = s.t, s.step_count
    s.close()
    return out, t, steps


@pytest.mark.parametrize("case", ["tgv2d", "shock"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stack_instances_equal_alone(ctx, case, precision):
    import compactles_hip as ch
    if case == "tgv2d":
        n, vals = (48, 40, 1), (1.0, 0.7, 1.3, 0.9)
        make = lambda v: _tgv2d(ch, n, precision, v)   # noqa: E731
        tf, nmax = [0.3, 0.2, 0.35, 1e9], [2 ** 62, 2 ** 62, 2 ** 62, 23]
    else:
        n, vals = (96, 32, 1), (1.0, 0.95, 1.05)
        make = lambda v: _shock(ch, n, precision, v)   # noqa: E731
        tf, nmax = [0.03, 0.02, 1e9], [2 ** 62, 2 ** 62, 31]
    probs = [make(v) for v in vals]
    st = ch.Stack(ctx, probs[0][0], probs[0][1], len(vals))
    for i, (_, _, Q) in enumerate(probs):
        st.upload(i, Q)
    taken = st.run(tf, nmax)
    t, steps, status = st.state()
    print(f"\n{case} {precision}: stack of {len(vals)}, {taken} iterations, steps {list(steps)}")
    assert (status == 0).all()
    for i, (prob, num, Q) in enumerate(probs):
        ref, t1, s1 = _alone(ctx, prob, num, Q, tf[i], nmax[i])
        assert steps[i] == s1 and t[i] == t1, (i, steps[i], s1, t[i], t1)
        assert np.array_equal(st.download(i), ref), i
    # a stack of one
    one = ch.Stack(ctx, probs[1][0], probs[1][1], 1)
    one.upload(0, probs[1][2])
    one.run([tf[1]], [nmax[1]])
    ref, t1, s1 = _alone(ctx, *probs[1], tf[1], nmax[1])
    assert np.array_equal(one.download(0), ref)
    one.close()
    st.close()


def test_stack_refuses_three_dimensional_problems(ctx):
    import compactles_hip as ch
    prob, num, Q = _tgv2d(ch, (16, 16, 1), "fp64", 1.0)
    num3 = ch.Numerics(n_global=(16, 16, 8), cfl=0.5)
    with pytest.raises(ValueError):
        ch.Stack(ctx, prob, num3, 2)
