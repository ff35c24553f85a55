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
 - setup refusals (wrong metric, origin, face, stretched r, resolved theta)."""
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
    n2 = (32, 16, 12)
    z2 = np.zeros(n2)
    with pytest.raises(ValueError, match="resolved theta"):
        cases.oracle_solver(n2, (1.0, 2 * np.pi, 0.5), cases.conserved((z2, z2, z2, 1 + z2, 1 + z2)), bcs=AX,
                            metric="cylindrical")
