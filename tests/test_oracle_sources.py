"""The oracle's ConstantBodyForce (sources.jl), pinned by the reference's test
(runtests.jl:374-387): on a uniform periodic state the RHS is the source
alone, rho g on momentum (2, 4, -6) and (rho u) . g = 2.6 on energy; two
forces add in order."""
import numpy as np

import cases


def _uniform(forces):
    n, L = (12, 12, 12), (2 * np.pi,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved((0.5 + 0 * X, 0.25 + 0 * X, -0.1 + 0 * X, 1.0 + 0 * X, 2.0 + 0 * X))
    return cases.oracle_solver(n, L, q, art=False, body_forces=forces)


def test_body_force_known_answer():
    s = _uniform([(1.0, 2.0, -3.0)])
    s.compute_rhs(False)
    dQ = s.interior(s.dQ)
    for c, v in ((1, 2.0), (2, 4.0), (3, -6.0), (4, 2.6)):
        assert np.abs(dQ[..., c] - v).max() < 1e-10, c
    assert np.abs(dQ[..., 0]).max() < 1e-12


def test_two_forces_add():
    s = _uniform([(1.0, 2.0, -3.0), (0.5, 0.0, 1.0)])
    s.compute_rhs(False)
    dQ = s.interior(s.dQ)
    for c, v in ((1, 3.0), (2, 4.0), (3, -4.0), (4, 2.6 + 2 * (0.25 - 0.1))):
        assert np.abs(dQ[..., c] - v).max() < 1e-10, c
