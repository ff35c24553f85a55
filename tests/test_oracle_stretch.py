"""Stretched Cartesian grids (Numerics.stretch, metric.jl:45-75) in the oracle,
pinned by what the mathematics fixes:
 - freestream preservation with sine_cluster(0, 1, 0.5, 0.4) in x between slip
   walls (the reference's own case, runtests.jl:305-340);
 - d/dx of sin(2 pi x) on the clustered grid (a missing or wrong 1/h scaling
   is an O(1) error) and its convergence under refinement;
 - the general divergence dQ = -inv_J D_d(A_d F_d): a pressure linear in xi
   gives -0.1 / m(xi) along the stretched x, and one linear in the uniform y
   gives -0.1 exactly, which needs A_y = J / h_y (compact schemes and their
   closures are exact on linear data);
 - max_rate against its closed form on a uniform state, and the kinetic energy
   volume integral against its exact value (the trapezoid rule is exact on the
   sine map's Jacobian)."""
import math

import numpy as np
import pytest

import cases

A, XC = 0.4, 0.5


def _map(n):
    xi = np.arange(n) / (n - 1)
    x = xi - (A / (2 * np.pi)) * (np.sin(2 * np.pi * (xi - XC)) + math.sin(2 * np.pi * XC))
    m = 1.0 * (1 - A * np.cos(2 * np.pi * (xi - XC)))
    return xi, x, m


def _solver(n, prim_fn, walls_y=False, **kw):
    xi, x, m = _map(n[0])
    L = (1.0, 1.0, 1.0)
    hy = L[1] / (n[1] - 1) if walls_y else L[1] / n[1]
    y = np.arange(n[1]) * hy
    z = np.arange(n[2]) * (L[2] / n[2])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    XI = np.broadcast_to(xi[:, None, None], n)
    q = cases.conserved(prim_fn(X, Y, Z, XI))
    bcs = ((1, 1), (1, 1) if walls_y else (0, 0), (0, 0))
    kw.setdefault("art", False)
    return cases.oracle_solver(n, L, q, jac=[m, None, None], bcs=bcs, **kw), X, Y, XI, m


def test_freestream_is_preserved():
    n = (24, 12, 12)
    s, *_ = _solver(n, lambda X, Y, Z, XI: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    s.apply_bcs()
    s.compute_rhs(False)
    assert np.abs(s.dQ).max() < 1e-12


def test_gradient_on_the_clustered_grid():
    errs = []
    for n0 in (24, 48):
        n = (n0, 10, 10)
        s, X, *_ = _solver(n, lambda X, Y, Z, XI: (0.01 * np.sin(2 * np.pi * X), 0 * X, 0 * X, 1.0 + 0 * X,
                                                     1.0 + 0 * X))
        s.compute_rhs(False)
        g = s.interior(s.field("grad_u[11]"))
        ref = 0.01 * 2 * np.pi * np.cos(2 * np.pi * X)
        errs.append(np.abs(g - ref).max() / np.abs(ref).max())
    # measured 9.1e-3 and 5.7e-4 (closure-limited, about fourth order); a
    # missing or wrong 1/h_x would leave an O(1) error at every n
    assert errs[0] < 1.2e-2 and errs[1] < 8e-4 and errs[0] / errs[1] > 10, errs


def test_divergence_general_path_on_linear_pressure():
    n = (20, 14, 10)
    # p linear in xi: d(rho u)/dt = -dp/dx = -0.1 / m(xi); p linear in y: d(rho v)/dt = -0.1
    s, X, Y, XI, m = _solver(n, lambda X, Y, Z, XI: (0 * X, 0 * X, 0 * X, 1.0 + 0.1 * XI + 0.1 * Y, 1.0 + 0 * X),
                             walls_y=True)
    s.compute_rhs(False)
    dmx = s.interior(s.dQ[..., 1])
    dmy = s.interior(s.dQ[..., 2])
    want_x = -0.1 / m[:, None, None]
    assert np.abs(dmx - want_x).max() < 1e-12
    assert np.abs(dmy + 0.1).max() < 1e-12


def test_max_rate_closed_form():
    n = (20, 10, 10)
    s, X, Y, XI, m = _solver(n, lambda X, Y, Z, XI: (0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    rate, rho_min = s.max_rate()
    c = math.sqrt(1.4)
    hx, hy, hz = 1.0 / (n[0] - 1), 1.0 / n[1], 1.0 / n[2]
    want = max(c * ((1 / mi) / hx + 1.0 / hy + 1.0 / hz) for mi in m)
    assert abs(rate - want) <= 1e-13 * want and rho_min == 1.0


def test_kinetic_energy_volume():
    n = (20, 10, 10)
    u0 = 0.3
    s, *_ = _solver(n, lambda X, Y, Z, XI: (u0 + 0 * X, 0 * X, 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    ke, _ = s.diagnostics()
    assert abs(ke - 0.5 * u0 * u0 * 1.0) <= 1e-13


def test_frontend_validation():
    import compactles_hip as ch
    walls = (ch.SlipWallBC(), ch.SlipWallBC())
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    prob = ch.Problem(domain=((0.0, 1.0),) * 3, bcs=(walls, per, per), ic=lambda X, Y, Z: None)
    ok = ch.SolverSpec(prob, ch.Numerics(n_global=(16, 10, 10), stretch=(ch.sine_cluster(0.0, 1.0, 0.5, 0.4), None,
                                                                        None)))
    xi, x, m = _map(16)
    assert np.allclose(ok.xcoord(0), x, rtol=0, atol=1e-15) and np.allclose(ok.jac[0], m, rtol=0, atol=1e-15)
    assert ok.h[0] == 1.0 / 15
    with pytest.raises(ValueError, match="does not span the domain"):
        ch.SolverSpec(prob, ch.Numerics(n_global=(16, 8, 8), stretch=(ch.sine_cluster(0.0, 2.0, 0.5, 0.4), None,
                                                                       None)))
    with pytest.raises(ValueError, match="non-periodic"):
        ch.SolverSpec(prob, ch.Numerics(n_global=(16, 8, 8), stretch=(None, ch.sine_cluster(0.0, 1.0, 0.5, 0.4),
                                                                       None)))
    with pytest.raises(ValueError, match="0 ≤ a < 1"):
        ch.sine_cluster(0.0, 1.0, 0.5, 1.0)
