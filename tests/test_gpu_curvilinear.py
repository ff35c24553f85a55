"""Curvilinear metrics on the GPU (cles_solver_create_geom, metric.jl:80-269)
against the oracle on the same inputs: one RHS (dQ and the curvature-corrected
grad_u 1e-12; sensor-fed coefficients 1e-9), a step and a filter (state
1e-12), 20 run! steps (dt history 1e-12, state, kinetic energy and enstrophy
1e-10), for cylindrical annuli (theta resolved and collapsed with swirl),
spherical shells (phi resolved and collapsed) and a stretched radius;
freestream on the device; the FP32-mixed form against the FP64 oracle;
NSCBC on radial faces; refusals (NSCBC on an angular face, batches, r <= 0,
theta outside (0, pi))."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_batch import _state_err
from test_gpu_solver import _compare_rhs

pytestmark = pytest.mark.gpu
SLIP, PER = (1, 1), (0, 0)
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _cluster(n, lo, hi, xc=0.5, a=0.4):
    xi = np.arange(n) / (n - 1)
    L = hi - lo
    x = lo + L * (xi - (a / (2 * np.pi)) * (np.sin(2 * np.pi * (xi - xc)) + np.sin(2 * np.pi * xc)))
    return x, L * (1 - a * np.cos(2 * np.pi * (xi - xc)))


def _flow(origin, L, metric="cylindrical"):
    """A smooth swirling state that satisfies the walls: sine modes in r (and,
    spherical, in theta) vanishing where the normal velocity must, whole
    periods along periodic angles and z.  Spherical: rotation about the
    polar axis (u_phi = 0.3 r sin theta)."""
    if metric == "spherical":
        def sph(X, Y, Z):
            sr = np.sin(np.pi * (X - origin[0]) / L[0])
            st = np.sin(np.pi * (Y - origin[1]) / L[1])
            k3 = 2 * np.pi / L[2]
            ur = 0.08 * sr * np.cos(k3 * Z) * (0.5 + st)
            ut = 0.05 * sr * st * np.sin(k3 * Z)
            up = 0.3 * X * np.sin(Y) + 0.04 * sr * np.cos(k3 * Z)
            p = 1.0 + 0.05 * np.cos(np.pi * (Y - origin[1]) / L[1]) * np.sin(k3 * Z) + 0.02 * sr
            rho = 1.0 + 0.04 * np.sin(k3 * Z + Y) * sr
            return ur, ut, up, p, rho
        return sph

    def prim(X, Y, Z):
        s = np.sin(np.pi * (X - origin[0]) / L[0])
        k2, k3 = 2 * np.pi / L[1], 2 * np.pi / L[2]
        ur = 0.08 * s * np.cos(k2 * (Y - origin[1])) * np.cos(k3 * Z)
        ut = 0.3 * X + 0.05 * s * np.sin(k3 * Z)
        uz = 0.06 * np.cos(k2 * (Y - origin[1])) * s
        p = 1.0 + 0.05 * np.cos(k2 * (Y - origin[1])) * np.sin(k3 * Z) + 0.02 * s
        rho = 1.0 + 0.04 * np.sin(k2 * (Y - origin[1]) + k3 * Z) * s
        return ur, ut, uz, p, rho
    return prim


CASES = {
    "cyl_annulus": ((24, 16, 12), (1.0, TWO_PI, 0.5), (0.3, 0.0, 0.0), "cylindrical", (SLIP, PER, PER)),
    "cyl_axisymmetric_swirl": ((32, 1, 16), (1.0, 1.0, 0.5), (0.3, 0.0, 0.0), "cylindrical", (SLIP, PER, PER)),
    "sph_shell": ((24, 12, 16), (1.0, np.pi / 2, TWO_PI), (0.5, np.pi / 4, 0.0), "spherical", (SLIP, SLIP, PER)),
    "sph_phi_collapsed": ((24, 16, 1), (1.0, np.pi / 2, 1.0), (0.5, np.pi / 4, 0.0), "spherical", (SLIP, SLIP, PER)),
}


def _pair(ctx, case, precision="fp64", stretch_r=False, prim=None, **kw):
    import compactles_hip as ch
    n, L, origin, metric, bcs = CASES[case]
    periodic = tuple(b == PER for b in bcs)
    X, Y, Z = cases.grid(n, L, periodic=periodic, origin=origin)
    x = jac = st = None
    if stretch_r:
        xr, mr = _cluster(n[0], origin[0], origin[0] + L[0])
        X = np.broadcast_to(xr[:, None, None], X.shape)
        x, jac = (xr, None, None), (mr, None, None)
        st = (ch.sine_cluster(origin[0], origin[0] + L[0], 0.5, 0.4), None, None)
    q = cases.conserved((prim or _flow(origin, L, metric))(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, jac=jac, x=x, bcs=bcs, origin=origin, metric=metric, **kw)
    gs = cases.gpu_solver(ctx, n, L, q, bcs=bcs, origin=origin, metric=metric, precision=precision, stretch=st, **kw)
    return os_, gs, n


def _q_err(gs, os_, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    got = [Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)]
    ref = [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)]
    return _state_err(got, ref)


@pytest.mark.parametrize("case", sorted(CASES))
def test_rhs_step_filter_match_oracle(ctx, case):
    os_, gs, n = _pair(ctx, case, mu0=2e-3, art=True)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    for a in (1, 2, 3):
        for b in (1, 2, 3):
            ref = os_.interior(os_.field(f"grad_u[{a}{b}]"))
            got = gs.field(f"grad_u[{a}{b}]")
            if np.abs(ref).max() == 0:
                assert np.abs(got).max() == 0, (a, b)
            else:
                assert Hh.relerr(got, ref) <= 1e-12, (case, a, b, Hh.relerr(got, ref))
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    dt = 0.5 / r1
    os_.step(dt, True)
    gs.step(dt, True)
    os_.filter_state()
    gs.filter_state()
    assert _q_err(gs, os_, n) <= 1e-12


@pytest.mark.parametrize("case", ["cyl_annulus", "cyl_axisymmetric_swirl", "sph_shell", "sph_phi_collapsed"])
def test_run_matches_oracle(ctx, case):
    os_, gs, n = _pair(ctx, case, mu0=2e-3, art=True)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-10 * ke1
    assert abs(en1 - en2) <= 1e-10 * en1


@pytest.mark.parametrize("case", ["cyl_annulus", "sph_shell"])
def test_stretched_radius_matches_oracle(ctx, case):
    os_, gs, n = _pair(ctx, case, stretch_r=True, mu0=2e-3, art=True)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    d1 = os_.run(1e9, 15, record=15)
    d2 = gs.run(1e9, 15, record=15)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10


@pytest.mark.parametrize("case", sorted(CASES))
def test_freestream_on_gpu(ctx, case):
    _, gs, n = _pair(ctx, case, art=False,
                     prim=lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    gs.apply_bcs()
    gs.compute_rhs(False)
    for c in range(5):
        assert np.abs(gs.field(f"dQ[{c + 1}]")).max() < 1e-12, (case, c)


def test_fp32_mixed_curvilinear(ctx):
    errs = {}
    for case in ("cyl_annulus", "sph_shell"):
        os_, gs, n = _pair(ctx, case, precision="fp32", mu0=2e-3, art=True)
        os_.run(1e9, 10)
        gs.run(1e9, 10)
        errs[case] = _q_err(gs, os_, n)
    print("\nFP32-mixed curvilinear, 10 steps vs FP64 oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-5


def test_refusals(ctx):
    import compactles_hip as ch
    n, L, origin, metric, bcs = CASES["cyl_annulus"]
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True), origin=origin)
    q = cases.conserved(_flow(origin, L)(X, Y, Z))
    with pytest.raises(ch.ClesError, match="r > 0"):
        cases.gpu_solver(ctx, n, L, q, bcs=bcs, origin=(0.0, 0.0, 0.0), metric="cylindrical")
    with pytest.raises(ch.ClesError, match="theta"):
        cases.gpu_solver(ctx, n, (1.0, 3.2, TWO_PI), q, bcs=(SLIP, SLIP, PER), origin=(0.5, 0.1, 0.0),
                         metric="spherical")
    faces = {(1, 1): {"pinf": 1.0, "sigma": 0.25, "Lref": 1.0, "beta_t": 0.0}}
    with pytest.raises(ch.ClesError, match="normal metric scale factor is one"):
        cases.gpu_solver(ctx, n, (1.0, 1.0, 0.5), q, bcs=(SLIP, SLIP, PER), origin=origin, metric="cylindrical",
                         nscbc=faces)
    dummy = ch.DirichletBC(lambda x, y, z, t: None)
    prob = ch.Problem(domain=((0.5, 1.5),) * 3, bcs=((dummy, dummy), (ch.PeriodicBC(), ch.PeriodicBC()),
                                                     (ch.PeriodicBC(), ch.PeriodicBC())),
                      ic=lambda x, y, z: None, metric=ch.CylindricalMetric())
    inst = [ch.Instance(n=64, domain=(0.5, 1.5), ic=None, lo=ch.Prim(rho=1.0, p=1.0), hi=ch.Prim(rho=1.0, p=1.0))]
    with pytest.raises(ch.ClesError, match="Cartesian"):
        ch.Batch(ctx, prob, ch.Numerics(n_global=(1, 1, 1)), inst)


@pytest.mark.parametrize("metric", ["cylindrical", "spherical"])
def test_nscbc_radial_faces_match_oracle(ctx, metric):
    """NSCBC on r faces (normal scale factor one): a radial outflow through an
    annulus / shell with inflow at the inner radius (nscbc.jl:59-306 with the
    curvature-corrected grad_u and inv_h).  Artificial terms off: with them on,
    the oracle alone turns a 1e-13 change of the state into 1e-9 (spherical)
    of the dt history within 15 steps (the sensors' |.| and switches at the
    faces); the sensors under these metrics are compared above."""
    if metric == "cylindrical":
        n, L, origin, bcs = (28, 16, 12), (1.0, TWO_PI, 0.5), (0.5, 0.0, 0.0), (SLIP, PER, PER)
    else:
        n, L, origin, bcs = (28, 12, 16), (1.0, np.pi / 2, TWO_PI), (0.5, np.pi / 4, 0.0), (SLIP, SLIP, PER)
    X, Y, Z = cases.grid(n, L, periodic=tuple(b == PER for b in bcs), origin=origin)
    k3 = 2 * np.pi / L[2]
    st = np.sin(np.pi * (Y - origin[1]) / L[1]) if metric == "spherical" else np.cos(Y)
    ur = 0.25 * origin[0] / X + 0.03 * np.cos(k3 * Z) * st
    ut = 0.04 * np.sin(np.pi * (X - origin[0])) * st
    uz = 0.05 * np.cos(k3 * Z)
    p = 1.0 + 0.03 * np.sin(np.pi * (X - origin[0])) * np.cos(k3 * Z)
    rho = 1.0 + 0.02 * np.cos(k3 * Z)
    q = cases.conserved((ur, ut, uz, p, rho))
    faces = {(0, 0): {"u": (0.26, 0.0, 0.0), "T": 1.0, "Y": (1.0,), "eta_u": 0.28, "eta_T": 0.35, "eta_t": 0.22,
                      "eta_Y": 0.3},
             (0, 1): {"pinf": 0.98, "sigma": 0.3, "beta_t": -1.0}}
    os_ = cases.oracle_solver(n, L, q, bcs=cases.nscbc_codes(bcs, faces), nscbc=faces, origin=origin, metric=metric,
                              mu0=1e-3, art=False)
    gs = cases.gpu_solver(ctx, n, L, q, bcs=bcs, nscbc=faces, origin=origin, metric=metric, mu0=1e-3, art=False)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    _compare_rhs(gs, os_, n)
    d1 = os_.run(1e9, 15, record=15)
    d2 = gs.run(1e9, 15, record=15)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
