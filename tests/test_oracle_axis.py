"""The oracle's cylindrical axis (AxisBC, boundary.jl:55-66, rhs.jl:155-205,
folds.jl), pinned on the CPU by the reference's own tests:
 - d/dr through the fold of an odd r exp(-4 r^2) (< 5e-6) and of an even
   exp(-4 r^2) (< 2e-5) at the half-offset nodes, with C6 and with C10
   (runtests.jl:171-189, 231-249); a wrong parity leaves O(1) errors there;
 - freestream with the axis (runtests.jl:305-340; 1e-12 here);
 - the volume of the unit cylinder per radian, 1/2, within h^2 (the outer
   wall's edge term only: the folded edge carries a full weight,
   runtests.jl:979-986);
 - three steps of the axisymmetric pressure pulse stay finite
   (runtests.jl:1880-1902);
 - the paired folds (folds.jl): a resolved theta through the axis
   (runtests.jl:250-268: odd < 1e-5, even < 1e-4), the spherical origin and
   poles (runtests.jl:269-284: dT/dr < 1e-4, (1/r) dT/dtheta < 1e-8), and
   their freestream (1e-12);
 - setup refusals (wrong metric, origin, face, stretched r, odd counts,
   one-sided poles, an asymmetric theta range)."""
import numpy as np
import pytest

import cases
import oracle as O

AX = ((O.BC_AXIS, O.BC_SLIPWALL), (0, 0), (0, 0))


def _axis_solver(n, L, prim_fn, deriv="lele_d1_6", art=False, mu0=0.0, **kw):
    h = L[0] / (n[0] - 0.5)
    r = (np.arange(n[0]) + 0.5) * h
    X, Y, Z = np.meshgrid(r, np.zeros(n[1]) if n[1] == 1 else np.arange(n[1]) * L[1] / n[1],
                          np.arange(n[2]) * L[2] / n[2] if n[2] > 1 else np.zeros(1), indexing="ij")
    q = cases.conserved(prim_fn(X, Y, Z))
    s = cases.oracle_solver(n, L, q, bcs=AX, metric="cylindrical", deriv=deriv, art=art, mu0=mu0, **kw)
    return s, (X, Y, Z)


@pytest.mark.parametrize("deriv", ["lele_d1_6", "lele_d1_10"])
def test_derivatives_through_the_fold(deriv):
    n, L = (64, 1, 12), (1.0, 1.0, 0.5)

    def prim(X, Y, Z):
        return X * np.exp(-4 * X * X), 0 * X, np.exp(-4 * X * X), 1.0 + 0 * X, 1.0 + 0 * X
    s, (X, Y, Z) = _axis_solver(n, L, prim, deriv=deriv)
    s.compute_rhs(False)
    odd = s.interior(s.field("grad_u[11]"))       # d u_r / dr, u_r odd
    even = s.interior(s.field("grad_u[13]"))      # d u_z / dr, u_z even
    assert np.abs(odd - (1 - 8 * X * X) * np.exp(-4 * X * X)).max() < 5e-6
    assert np.abs(even - (-8 * X * np.exp(-4 * X * X))).max() < 2e-5


def test_freestream_with_axis():
    s, _ = _axis_solver((32, 1, 12), (1.0, 1.0, 0.5), lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    s.apply_bcs()
    s.compute_rhs(False)
    assert np.abs(s.interior(s.dQ)).max() < 1e-12


def test_volume_of_the_unit_cylinder():
    s, _ = _axis_solver((64, 1, 12), (1.0, 1.0, 1.0), lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    h = 1.0 / (64 - 0.5)
    v = s.volume_integral(np.ones(s.Q.shape[:3], order="F"))
    assert abs(v - 0.5) <= 1e-3 * 0.5
    assert v - 0.5 < h * h


def test_axisymmetric_pressure_pulse_smoke():
    s, _ = _axis_solver((48, 1, 1), (1.0, 1.0, 1.0),
                        lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1 + np.exp(-40 * (X - 0.4) ** 2), 1.0 + 0 * X), art=True)
    s.run(1e9, 3)
    assert np.isfinite(s.interior(s.Q)).all()


def test_axis_refusals():
    n, L = (32, 1, 12), (1.0, 1.0, 0.5)
    z = np.zeros(n)
    q = cases.conserved((z, z, z, 1 + z, 1 + z))
    with pytest.raises(ValueError, match="CylindricalMetric"):
        cases.oracle_solver(n, L, q, bcs=AX, metric="cartesian")
    with pytest.raises(ValueError, match="origin"):
        cases.oracle_solver(n, L, q, bcs=AX, metric="cylindrical", origin=(0.1, 0.0, 0.0))
    with pytest.raises(ValueError, match="low end"):
        cases.oracle_solver(n, L, q, bcs=((O.BC_SLIPWALL, O.BC_AXIS), (0, 0), (0, 0)), metric="cylindrical")
    with pytest.raises(ValueError, match="stretched"):
        cases.oracle_solver(n, L, q, bcs=AX, metric="cylindrical", jac=(np.ones(n[0]), None, None),
                            x=(np.linspace(0, 1, n[0]), None, None))
    n2 = (32, 15, 12)
    z2 = np.zeros(n2)
    with pytest.raises(ValueError, match="even theta point count"):
        cases.oracle_solver(n2, (1.0, 2 * np.pi, 0.5), cases.conserved((z2, z2, z2, 1 + z2, 1 + z2)), bcs=AX,
                            metric="cylindrical")


SPH = ((O.BC_ORIGIN, O.BC_SLIPWALL), (O.BC_POLE, O.BC_POLE), (0, 0))


def _sph_solver(n, prim_fn, art=False, mu0=0.0, bcs=SPH, L=(1.0, np.pi, 2 * np.pi), origin=(0.0, 0.0, 0.0)):
    hr = L[0] / (n[0] - 0.5)
    r = (np.arange(n[0]) + 0.5) * hr
    ht = L[1] / n[1]
    th = origin[1] + (np.arange(n[1]) + 0.5) * ht if bcs[1][0] == O.BC_POLE else \
        origin[1] + np.arange(n[1]) * L[1] / max(n[1] - 1, 1)
    ph = np.arange(n[2]) * L[2] / n[2] if n[2] > 1 else np.zeros(1)
    X, Y, Z = np.meshgrid(r, th, ph, indexing="ij")
    q = cases.conserved(prim_fn(X, Y, Z))
    s = cases.oracle_solver(n, L, q, bcs=bcs, metric="spherical", art=art, mu0=mu0, origin=origin)
    return s, (X, Y, Z)


def test_resolved_theta_axis_pairing():
    """runtests.jl:250-268: an odd r cos(theta) exp(-4 r^2) (< 1e-5) and an
    even exp(-4 r^2) (1 + cos(2 theta) / 2) (< 1e-4) through the paired axis."""
    n, L = (48, 16, 1), (1.0, 2 * np.pi, 1.0)
    h = L[0] / (n[0] - 0.5)
    r = (np.arange(n[0]) + 0.5) * h
    th = np.arange(n[1]) * L[1] / n[1]
    X, Y, Z = np.meshgrid(r, th, np.zeros(1), indexing="ij")
    g = np.exp(-4 * X * X)
    q = cases.conserved((X * np.cos(Y) * g, 0 * X, g * (1 + 0.5 * np.cos(2 * Y)), 1.0 + 0 * X, 1.0 + 0 * X))
    s = cases.oracle_solver(n, L, q, bcs=AX, metric="cylindrical", art=False)
    s.compute_rhs(False)
    odd = s.interior(s.field("grad_u[11]"))
    even = s.interior(s.field("grad_u[13]"))
    assert np.abs(odd - np.cos(Y) * (1 - 8 * X * X) * g).max() < 1e-5
    assert np.abs(even - (-8 * X * g * (1 + 0.5 * np.cos(2 * Y)))).max() < 1e-4


def test_spherical_origin_and_poles_gaussian():
    """runtests.jl:269-284: T = exp(-4 r^2) is smooth at the origin and the
    poles: dT/dr < 1e-4, (1/r) dT/dtheta < 1e-8."""
    s, (X, Y, Z) = _sph_solver((40, 16, 12), lambda X, Y, Z: (0 * X, 0 * X, 0 * X, np.exp(-4 * X * X), 1.0 + 0 * X))
    s.compute_rhs(False)
    assert np.abs(s.interior(s.field("grad_T[1]")) - (-8 * X * np.exp(-4 * X * X))).max() < 1e-4
    assert np.abs(s.interior(s.field("grad_T[2]"))).max() < 1e-8


@pytest.mark.parametrize("case", ["axis_resolved", "sphere_full", "sphere_phi_collapsed"])
def test_freestream_paired_folds(case):
    """runtests.jl:305-340 freestream with the paired folds (1e-12 here)."""
    uni = lambda X, Y, Z: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X)
    if case == "axis_resolved":
        n, L = (24, 16, 1), (1.0, 2 * np.pi, 1.0)
        z = np.zeros(n)
        s = cases.oracle_solver(n, L, cases.conserved((z, z, z, 1 + z, 1 + z)), bcs=AX, metric="cylindrical",
                                art=False)
    elif case == "sphere_full":
        s, _ = _sph_solver((24, 12, 12), uni)
    else:
        s, _ = _sph_solver((24, 12, 1), uni, L=(1.0, np.pi, 1.0))
    s.apply_bcs()
    s.compute_rhs(False)
    assert np.abs(s.interior(s.dQ)).max() < 1e-12, case


def test_sphere_refusals():
    z = np.zeros((24, 12, 12))
    q = cases.conserved((z, z, z, 1 + z, 1 + z))
    with pytest.raises(ValueError, match="both ends"):
        cases.oracle_solver((24, 12, 12), (1.0, np.pi, 2 * np.pi), q,
                            bcs=((O.BC_ORIGIN, O.BC_SLIPWALL), (O.BC_POLE, O.BC_SLIPWALL), (0, 0)), metric="spherical")
    with pytest.raises(ValueError, match="SphericalMetric"):
        cases.oracle_solver((24, 12, 12), (1.0, np.pi, 2 * np.pi), q, bcs=SPH, metric="cylindrical")
    with pytest.raises(ValueError, match="symmetric about pi/2"):
        cases.oracle_solver((24, 12, 12), (1.0, 1.0, 2 * np.pi), q,
                            bcs=((O.BC_ORIGIN, O.BC_SLIPWALL), (O.BC_SLIPWALL, O.BC_SLIPWALL), (0, 0)),
                            metric="spherical", origin=(0.0, 0.3, 0.0))
    z2 = np.zeros((24, 12, 11))
    with pytest.raises(ValueError, match="even phi"):
        cases.oracle_solver((24, 12, 11), (1.0, np.pi, 2 * np.pi), cases.conserved((z2, z2, z2, 1 + z2, 1 + z2)),
                            bcs=SPH, metric="spherical")
