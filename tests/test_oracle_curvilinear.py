"""The oracle's curvilinear metrics (metric.jl:80-269), pinned on the CPU by
the reference's own tests and by closed forms:
 - freestream preservation in cylindrical and spherical shells, resolved and
   collapsed angular dimensions, with and without a radial stretch
   (runtests.jl:305-340; the reference's bar 1e-8, met here to 1e-12: the
   spherical case needs the discrete-GCL cot(theta)/r); with a stretched
   radius the residual is the scheme's truncation error and converges;
 - rigid rotation u_theta = Omega r has zero strain (runtests.jl:285-303):
   a wrong sign or a missing curvature correction leaves |S| = Omega;
 - curvature_rate for collapsed angular dimensions (runtests.jl:641-665), and
   its exact value |u_theta| / r_min (cylindrical), 2 |u| / r_min (spherical
   at theta = pi/2);
 - a radial pressure p = a + b r gives d(rho u_r)/dt = -b exactly (the
   scheme and its closures differentiate r p and r^2 p exactly): pins A_1,
   inv_J and the Pi_thetatheta (+ Pi_phiphi) sources;
 - p = a + b theta in a cylindrical wedge gives -b / r exactly (pins inv_J
   and the theta scale factor); in a spherical wedge the same limit is
   reached at the scheme's order;
 - the volume integral of 1 over an annulus is pi (R^2 - r0^2) L_z exactly
   (the trapezoid rule integrates J = r exactly)."""
import math

import numpy as np
import pytest

import cases
import oracle as O

SLIP, PER = (1, 1), (0, 0)


def _solver(n, L, origin, metric, prim_fn, bcs, jac=None, x=None, art=False, mu0=0.0, **kw):
    periodic = tuple(b == PER for b in bcs)
    X, Y, Z = cases.grid(n, L, periodic=periodic, origin=origin)
    if x is not None and x[0] is not None:
        X = np.broadcast_to(x[0][:, None, None], X.shape)
    q = cases.conserved(prim_fn(X, Y, Z))
    s = cases.oracle_solver(n, L, q, jac=jac, x=x, bcs=bcs, origin=origin, metric=metric, art=art, mu0=mu0, **kw)
    return s, (X, Y, Z)


def _uniform(X, Y, Z):
    z = 0 * X
    return z, z, z, 1.0 + z, 1.0 + z


def _cluster(n, lo, hi, xc=0.5, a=0.4):
    xi = np.arange(n) / (n - 1)
    L = hi - lo
    x = lo + L * (xi - (a / (2 * np.pi)) * (np.sin(2 * np.pi * (xi - xc)) + np.sin(2 * np.pi * xc)))
    return x, L * (1 - a * np.cos(2 * np.pi * (xi - xc)))


FREESTREAM = {
    "cyl_annulus": ((24, 12, 12), (1.0, 2 * np.pi, 0.5), (0.2, 0.0, 0.0), "cylindrical", (SLIP, PER, PER)),
    "cyl_axisymmetric": ((32, 1, 12), (1.0, 1.0, 0.5), (0.2, 0.0, 0.0), "cylindrical", (SLIP, PER, PER)),
    "cyl_2d_r_theta": ((24, 16, 1), (1.0, 2 * np.pi, 1.0), (0.3, 0.0, 0.0), "cylindrical", (SLIP, PER, PER)),
    "sph_shell": ((24, 12, 12), (1.0, np.pi / 2, 2 * np.pi), (0.5, np.pi / 4, 0.0), "spherical",
                  (SLIP, SLIP, PER)),
    "sph_phi_collapsed": ((24, 16, 1), (1.0, np.pi / 2, 1.0), (0.5, np.pi / 4, 0.0), "spherical",
                          (SLIP, SLIP, PER)),
    "sph_theta_collapsed": ((24, 1, 12), (1.0, 1.0, 2 * np.pi), (0.5, np.pi / 2, 0.0), "spherical", (SLIP, PER, PER)),
}


@pytest.mark.parametrize("case", sorted(FREESTREAM))
def test_freestream(case):
    n, L, origin, metric, bcs = FREESTREAM[case]
    s, _ = _solver(n, L, origin, metric, _uniform, bcs)
    s.apply_bcs()
    s.compute_rhs(False)
    m = np.abs(s.interior(s.dQ)).max()
    assert m < 1e-12, (case, m)


def test_freestream_stretched_radius_converges():
    """With a stretched radius the radial balance -inv_J D_xi(r) p + p / r
    (+ p / r spherical, with r^2) closes only as far as the compact
    derivative of the node radii matches the mapping Jacobian: the residual
    sits in the radial momentum alone and falls at the closures' third order."""
    for metric, L, origin, bcs in (("cylindrical", (1.0, 2 * np.pi, 0.5), (0.2, 0.0, 0.0), (SLIP, PER, PER)),
                                   ("spherical", (1.0, np.pi / 2, 2 * np.pi), (0.2, np.pi / 4, 0.0),
                                    (SLIP, SLIP, PER))):
        errs = []
        for nr in (24, 48):
            n = (nr, 12, 12)
            x, m = _cluster(nr, 0.2, 1.2)
            s, _ = _solver(n, L, origin, metric, _uniform, bcs, jac=(m, None, None), x=(x, None, None))
            s.apply_bcs()
            s.compute_rhs(False)
            dQ = s.interior(s.dQ)
            for c in (0, 2, 3, 4):
                assert np.abs(dQ[..., c]).max() < 1e-12, (metric, c)
            errs.append(np.abs(dQ[..., 1]).max())
        assert errs[1] < errs[0] / 6, (metric, errs)


def test_stretched_curvilinear_needs_coordinates():
    n = (24, 12, 12)
    _, m = _cluster(n[0], 0.2, 1.2)
    with pytest.raises(ValueError, match="node coordinates"):
        _solver(n, (1.0, 2 * np.pi, 0.5), (0.2, 0.0, 0.0), "cylindrical", _uniform, (SLIP, PER, PER),
                jac=(m, None, None))


def test_rigid_rotation_zero_strain():
    n = (32, 16, 12)
    s, _ = _solver(n, (1.0, 2 * np.pi, 0.5), (0.2, 0.0, 0.0), "cylindrical",
                   lambda X, Y, Z: (0 * X, 0.3 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X), (SLIP, PER, PER))
    s.compute_rhs(False)
    g = {(a, b): s.interior(s.field(f"grad_u[{a}{b}]")) for a in (1, 2, 3) for b in (1, 2, 3)}
    smax = max(np.abs(0.5 * (g[a, b] + g[b, a])).max() for a in (1, 2, 3) for b in (1, 2, 3))
    assert smax < 1e-8
    # the antisymmetric part is the rotation: (1/r) d(r u_theta)/dr = 2 Omega
    assert np.abs(g[1, 2] - g[2, 1] - 0.6).max() < 1e-10


@pytest.mark.parametrize("metric", ["cylindrical", "spherical"])
def test_curvature_rate(metric):
    def mk(uang):
        s, _ = _solver((64, 1, 1), (1.0, 1.0, 1.0), (0.5, np.pi / 2, 0.0), metric,
                       lambda X, Y, Z: (0 * X, uang + 0 * X, uang + 0 * X, 1.0 + 0 * X, 1.0 + 0 * X), (SLIP, PER, PER))
        return s.max_rate()[0]
    r0, r1 = mk(0.0), mk(2.0)
    assert r1 > r0
    # the largest rate sits at r = 0.5: |u_theta| / r, spherical also |u_phi| (1/r + |cot(pi/2)| / r)
    extra = 2.0 / 0.5 if metric == "cylindrical" else 2.0 / 0.5 + 2.0 * (1 / 0.5 + abs(math.cos(np.pi / 2)) / 0.5)
    assert abs((r1 - r0) - extra) <= 1e-12 * r1


@pytest.mark.parametrize("metric", ["cylindrical", "spherical"])
def test_radial_pressure_gradient_is_exact(metric):
    a, b = 1.0, 0.3
    n = (24, 12, 12)
    if metric == "cylindrical":
        L, origin, bcs = (1.0, 2 * np.pi, 0.5), (0.5, 0.0, 0.0), (SLIP, PER, PER)
    else:
        L, origin, bcs = (1.0, np.pi / 2, 2 * np.pi), (0.5, np.pi / 4, 0.0), (SLIP, SLIP, PER)
    s, _ = _solver(n, L, origin, metric, lambda X, Y, Z: (0 * X, 0 * X, 0 * X, a + b * X, 1.0 + 0 * X), bcs)
    s.compute_rhs(False)
    dQ = s.interior(s.dQ)
    assert np.abs(dQ[..., 1] + b).max() < 1e-11
    for c in (0, 2, 3, 4):
        assert np.abs(dQ[..., c]).max() < 1e-11, c


def test_theta_pressure_gradient_cylindrical_wedge():
    a, b = 1.0, 0.3
    n = (20, 24, 10)
    s, (X, Y, Z) = _solver(n, (1.0, 1.0, 0.5), (0.5, 0.0, 0.0), "cylindrical",
                           lambda X, Y, Z: (0 * X, 0 * X, 0 * X, a + b * Y, 1.0 + 0 * X), (SLIP, SLIP, PER))
    s.compute_rhs(False)
    dQ = s.interior(s.dQ)
    assert np.abs(dQ[..., 2] + b / X).max() < 1e-11
    assert np.abs(dQ[..., 1]).max() < 1e-11


def test_theta_pressure_gradient_spherical_converges():
    a, b = 1.0, 0.3
    errs = []
    for nt in (16, 32):
        n = (12, nt, 10)
        s, (X, Y, Z) = _solver(n, (1.0, np.pi / 2, 2 * np.pi), (0.5, np.pi / 4, 0.0), "spherical",
                               lambda X, Y, Z: (0 * X, 0 * X, 0 * X, a + b * Y, 1.0 + 0 * X), (SLIP, SLIP, PER))
        s.compute_rhs(False)
        errs.append(np.abs(s.interior(s.dQ)[..., 2] + b / X).max())
    assert errs[1] < errs[0] / 6, errs      # third order (the closures) at least
    assert errs[1] < 1e-4


def test_volume_integral_annulus():
    n = (32, 16, 12)
    r0, R, Lz = 0.2, 1.2, 0.5
    s, _ = _solver(n, (R - r0, 2 * np.pi, Lz), (r0, 0.0, 0.0), "cylindrical", _uniform, (SLIP, PER, PER))
    ke, _ = s.diagnostics()
    assert ke == 0.0
    one = np.ones(s.Q.shape[:3], order="F")
    vol = s.volume_integral(one)
    assert abs(vol - np.pi * (R * R - r0 * r0) * Lz) <= 1e-13 * vol


def test_viscous_couette_is_steady_in_momentum():
    """Circular Couette flow u_theta = A r + B / r with dp/dr = rho u_theta^2 / r
    is steady: tau_rtheta = -2 mu B / r^2 has zero divergence once the
    -Pi_thetar / r source and the grad_u corrections are in.  The momentum
    residual converges (second order); a wrong sign or a dropped curvature
    term leaves an O(mu B / r^3) residual."""
    A, B, p0, mu = 0.4, 0.15, 2.0, 0.05

    def prim(X, Y, Z):
        ut = A * X + B / X
        p = p0 + A * A * X * X / 2 + 2 * A * B * np.log(X) - B * B / (2 * X * X)
        return 0 * X, ut, 0 * X, p, 1.0 + 0 * X

    errs = []
    for nr in (24, 48):
        s, _ = _solver((nr, 1, 1), (1.0, 1.0, 1.0), (0.5, 0.0, 0.0), "cylindrical", prim, (SLIP, PER, PER),
                       mu0=mu)
        s.compute_rhs(False)
        dQ = s.interior(s.dQ)
        errs.append(max(np.abs(dQ[..., 1]).max(), np.abs(dQ[..., 2]).max()))
    # second order: the viscous term differentiates twice through third-order closures
    assert errs[1] < errs[0] / 3, errs
    assert errs[1] < 0.05 * (2 * mu * B / 0.5 ** 3), errs   # a dropped source leaves 2 mu B / r^3


def test_nscbc_faces_need_unit_scale_factor():
    """runtests.jl:514-532: r is a length under CylindricalMetric, theta is
    not; an NSCBC face on theta is a setup error (inflow and outflow)."""
    n, L, origin = (12, 12, 12), (1.0, 1.0, 1.0), (1.0, 0.0, 0.0)
    out = {"pinf": 1.0, "sigma": 0.25, "beta_t": -1.0}
    inflow = {"u": (0.0, 0.1, 0.0), "T": 1.0, "Y": (1.0,)}
    _solver(n, L, origin, "cylindrical", _uniform, (SLIP, SLIP, PER))
    for faces in ({(1, 0): out}, {(1, 1): inflow}):
        with pytest.raises(ValueError, match="normal metric scale factor is one"):
            _solver(n, L, origin, "cylindrical", _uniform, cases.nscbc_codes((SLIP, SLIP, PER), faces), nscbc=faces)
    for faces in ({(0, 1): out}, {(2, 0): out}):   # r and z faces are fine (z made closed here)
        bcs = (SLIP, SLIP, SLIP) if (2, 0) in faces else (SLIP, SLIP, PER)
        _solver(n, L, origin, "cylindrical", _uniform, cases.nscbc_codes(bcs, faces), nscbc=faces)
    with pytest.raises(ValueError, match="normal metric scale factor is one"):
        _solver(n, (1.0, 1.0, 1.0), (1.0, 0.5, 0.0), "spherical", _uniform,
                cases.nscbc_codes((SLIP, SLIP, SLIP), {(2, 1): out}), nscbc={(2, 1): out})
