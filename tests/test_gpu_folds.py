"""Paired coordinate-singularity folds on the GPU (folds.jl: the antipodal
even / odd butterfly) against the oracle on the same inputs: the reference's
known-answer tests on the device (a resolved theta through the axis, odd
< 1e-5 and even < 1e-4, runtests.jl:250-268; the spherical origin and poles,
dT/dr < 1e-4 and (1/r) dT/dtheta < 1e-8, runtests.jl:269-284) and equal to the
oracle's to 1e-12; freestream (runtests.jl:305-340); Cartesian-smooth flows
through the axis, the origin and the poles (RHS, grad_u, step and filter
1e-12; 20 run! steps from primed artificial coefficients: dt 1e-12, state,
kinetic energy and enstrophy 1e-10); the FP32-mixed form; refusals."""
import numpy as np
import pytest

import cases
import helpers as Hh
import oracle as O
from test_gpu_batch import _state_err
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
AX = ((O.BC_AXIS, O.BC_SLIPWALL), (0, 0), (0, 0))
SPH = ((O.BC_ORIGIN, O.BC_SLIPWALL), (O.BC_POLE, O.BC_POLE), (0, 0))
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _gpu(ctx, n, L, q, kind, precision="fp64", art=True, mu0=0.0):
    import compactles_hip as ch
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    if kind == "axis":
        bcs, metric = ((ch.AxisBC(), ch.SlipWallBC()), per, per), ch.CylindricalMetric()
    else:
        bcs, metric = ((ch.OriginBC(), ch.SlipWallBC()), (ch.PoleBC(), ch.PoleBC()), per), ch.SphericalMetric()
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=bcs, metric=metric,
                      ic=lambda X, Y, Z: None, transport=ch.Transport(mu0=mu0))
    num = ch.Numerics(n_global=n, art=ch.ArtParams(enabled=art), precision=precision)
    s = ch.Solver(ctx, prob, num)
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def _nodes(n, L, kind):
    r = (np.arange(n[0]) + 0.5) * L[0] / (n[0] - 0.5)
    if kind == "axis":
        t = np.arange(n[1]) * L[1] / n[1] if n[1] > 1 else np.zeros(1)
    else:
        t = (np.arange(n[1]) + 0.5) * L[1] / n[1]
    z = np.arange(n[2]) * L[2] / n[2] if n[2] > 1 else np.zeros(1)
    return np.meshgrid(r, t, z, indexing="ij")


def _pair(ctx, n, L, kind, prim_fn, precision="fp64", art=True, mu0=0.0):
    X, Y, Z = _nodes(n, L, kind)
    q = cases.conserved(prim_fn(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, bcs=AX if kind == "axis" else SPH,
                              metric="cylindrical" if kind == "axis" else "spherical", art=art, mu0=mu0)
    gs = _gpu(ctx, n, L, q, kind, precision=precision, art=art, mu0=mu0)
    return os_, gs, (X, Y, Z)


def _cyl_flow(L):
    """A Cartesian-smooth field through the axis, in (u_r, u_theta, u_z)."""
    def prim(X, Y, Z):
        g = np.exp(-4 * X * X)
        k = 2 * np.pi / L[2]
        Ux, Uy, Uz = 0.1 * g * np.cos(k * Z), 0.05 * g, 0.1 * g
        ur = Ux * np.cos(Y) + Uy * np.sin(Y)
        ut = -Ux * np.sin(Y) + Uy * np.cos(Y) + 0.2 * X * g
        return ur, ut, Uz, 1 + 0.1 * g, 1 + 0.05 * g * np.cos(k * Z)
    return prim


def _sph_flow(collapsed_phi=False):
    """A Cartesian-smooth field through the origin and the poles, in
    (u_r, u_theta, u_phi), with a rotation about the polar axis."""
    def prim(X, Y, Z):
        g = np.exp(-4 * X * X)
        U = (0 * g, 0 * g, 0.08 * g) if collapsed_phi else (0.1 * g, 0.05 * g, 0.08 * g)
        er = (np.sin(Y) * np.cos(Z), np.sin(Y) * np.sin(Z), np.cos(Y))
        et = (np.cos(Y) * np.cos(Z), np.cos(Y) * np.sin(Z), -np.sin(Y))
        ep = (-np.sin(Z), np.cos(Z), 0 * Z)
        ur = sum(a * b for a, b in zip(U, er))
        ut = sum(a * b for a, b in zip(U, et))
        up = sum(a * b for a, b in zip(U, ep)) + 0.2 * X * np.sin(Y) * g
        return ur, ut, up, 1 + 0.1 * g, 1 + 0.05 * g
    return prim


CASES = {
    "axis_resolved": ((32, 16, 12), (1.0, TWO_PI, 0.5), "axis"),
    "sphere": ((24, 12, 12), (1.0, np.pi, TWO_PI), "sphere"),
    "sphere_phi_collapsed": ((24, 16, 1), (1.0, np.pi, 1.0), "sphere"),
}


def _flow(case):
    n, L, kind = CASES[case]
    return _cyl_flow(L) if kind == "axis" else _sph_flow(collapsed_phi=n[2] == 1)


def _q_err(gs, os_, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    return _state_err([Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)],
                      [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)])


def test_resolved_theta_axis_known_answers(ctx):
    n, L = (48, 16, 1), (1.0, TWO_PI, 1.0)

    def prim(X, Y, Z):
        g = np.exp(-4 * X * X)
        return X * np.cos(Y) * g, 0 * X, g * (1 + 0.5 * np.cos(2 * Y)), 1.0 + 0 * X, 1.0 + 0 * X
    os_, gs, (X, Y, Z) = _pair(ctx, n, L, "axis", prim, art=False)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    g = np.exp(-4 * X * X)
    assert np.abs(gs.field("grad_u[11]") - np.cos(Y) * (1 - 8 * X * X) * g).max() < 1e-5
    assert np.abs(gs.field("grad_u[13]") - (-8 * X * g * (1 + 0.5 * np.cos(2 * Y)))).max() < 1e-4
    for nm in ("grad_u[11]", "grad_u[13]", "grad_u[21]", "grad_u[22]"):
        assert Hh.relerr(gs.field(nm), os_.interior(os_.field(nm))) <= 1e-12, nm


def test_spherical_origin_and_poles_known_answers(ctx):
    n, L = (40, 16, 12), (1.0, np.pi, TWO_PI)
    os_, gs, (X, Y, Z) = _pair(ctx, n, L, "sphere",
                               lambda X, Y, Z: (0 * X, 0 * X, 0 * X, np.exp(-4 * X * X), 1.0 + 0 * X), art=False)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    assert np.abs(gs.field("grad_T[1]") - (-8 * X * np.exp(-4 * X * X))).max() < 1e-4
    assert np.abs(gs.field("grad_T[2]")).max() < 1e-8
    ref1 = os_.interior(os_.field("grad_T[1]"))
    assert Hh.relerr(gs.field("grad_T[1]"), ref1) <= 1e-12
    # grad_T[2] is zero up to rounding: compared on the scale of the gradient
    assert np.abs(gs.field("grad_T[2]") - os_.interior(os_.field("grad_T[2]"))).max() <= 1e-12 * np.abs(ref1).max()


@pytest.mark.parametrize("case", sorted(CASES))
def test_freestream_paired_folds(ctx, case):
    n, L, kind = CASES[case]
    _, gs, _ = _pair(ctx, n, L, kind, lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X), art=False)
    gs.apply_bcs()
    gs.compute_rhs(False)
    for c in range(5):
        assert np.abs(gs.field(f"dQ[{c + 1}]")).max() < 1e-12, (case, c)


@pytest.mark.parametrize("case", sorted(CASES))
def test_rhs_step_filter_match_oracle(ctx, case):
    n, L, kind = CASES[case]
    os_, gs, _ = _pair(ctx, n, L, kind, _flow(case), mu0=2e-3)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    # the gradient tensor on one scale: near the axis and the poles some
    # components are small sums of O(u / r) curvature terms that cancel
    # (|result| ~ 1e-3 of the terms), so their own relative error is rounding
    # of the terms
    refs = {(a, b): os_.interior(os_.field(f"grad_u[{a}{b}]")) for a in (1, 2, 3) for b in (1, 2, 3)}
    scale = max(np.abs(r).max() for r in refs.values())
    for (a, b), ref in refs.items():
        assert np.abs(gs.field(f"grad_u[{a}{b}]") - ref).max() <= 1e-12 * scale, (case, a, b)
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    os_.filter_state()
    gs.filter_state()
    assert _q_err(gs, os_, n) <= 1e-12


@pytest.mark.parametrize("case", sorted(CASES))
def test_run_matches_oracle(ctx, case):
    """From primed artificial coefficients: the run's first dt reads the
    lagged coefficients, which a fresh solver holds at zero; near the axis
    the first RHS's beta* then exceeds that step's stability limit.  The
    collapsed-phi sphere runs without artificial terms: with them, the oracle
    alone turns a 1e-13 change of the pressure into 8e-11 of the dt history in
    20 steps (the sensors' switches near the poles; 5e-15 without)."""
    n, L, kind = CASES[case]
    os_, gs, _ = _pair(ctx, n, L, kind, _flow(case), mu0=2e-3, art=case != "sphere_phi_collapsed")
    for s in (os_, gs):
        s.apply_bcs()
        s.compute_rhs(False)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-10 * ke1
    assert abs(en1 - en2) <= 1e-10 * en1


def test_fp32_mixed_paired_folds(ctx):
    errs = {}
    for case in ("axis_resolved", "sphere"):
        n, L, kind = CASES[case]
        os_, gs, _ = _pair(ctx, n, L, kind, _flow(case), precision="fp32", mu0=2e-3)
        for s in (os_, gs):
            s.apply_bcs()
            s.compute_rhs(False)
        os_.run(1e9, 10)
        gs.run(1e9, 10)
        errs[case] = _q_err(gs, os_, n)
    print("\nFP32-mixed paired folds, 10 steps vs FP64 oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-5


def test_fold_refusals(ctx):
    import compactles_hip as ch
    per = (ch.PeriodicBC(), ch.PeriodicBC())

    def mk(bcs, metric, L=(1.0, np.pi, TWO_PI), n=(24, 12, 12), origin=(0.0, 0.0, 0.0)):
        prob = ch.Problem(domain=tuple((origin[k], origin[k] + L[k]) for k in range(3)), bcs=bcs, metric=metric,
                          ic=lambda X, Y, Z: None)
        return ch.Solver(ctx, prob, ch.Numerics(n_global=n))
    with pytest.raises(ch.ClesError, match="both ends"):
        mk(((ch.OriginBC(), ch.SlipWallBC()), (ch.PoleBC(), ch.SlipWallBC()), per), ch.SphericalMetric())
    with pytest.raises(ch.ClesError, match="SphericalMetric"):
        mk(((ch.OriginBC(), ch.SlipWallBC()), (ch.PoleBC(), ch.PoleBC()), per), ch.CylindricalMetric())
    with pytest.raises(ch.ClesError, match="symmetric about pi/2"):
        mk(((ch.OriginBC(), ch.SlipWallBC()), (ch.SlipWallBC(), ch.SlipWallBC()), per), ch.SphericalMetric(),
           L=(1.0, 1.0, TWO_PI), origin=(0.0, 0.3, 0.0))
    with pytest.raises(ch.ClesError, match="even phi"):
        mk(((ch.OriginBC(), ch.SlipWallBC()), (ch.PoleBC(), ch.PoleBC()), per), ch.SphericalMetric(), n=(24, 12, 11))
    with pytest.raises(ch.ClesError, match="even theta"):
        mk(((ch.AxisBC(), ch.SlipWallBC()), per, per), ch.CylindricalMetric(), L=(1.0, TWO_PI, 0.5), n=(24, 15, 12))
