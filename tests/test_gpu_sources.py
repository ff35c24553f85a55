"""ConstantBodyForce on the GPU against the oracle: the reference's known
answer (runtests.jl:374-387) on the device, and a gravity-driven run in a
walled box (RHS 1e-12; 20 run! steps: dt 1e-12, state 1e-10), FP64 and the
FP32-mixed form (1e-5)."""
import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_batch import _state_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _gpu(ctx, n, L, q, forces, bcs_x=None, precision="fp64", art=True, mu0=0.0):
    import compactles_hip as ch
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=(bcs_x or per, per, per),
                      ic=lambda X, Y, Z: None, transport=ch.Transport(mu0=mu0),
                      sources=tuple(ch.ConstantBodyForce(g) for g in forces))
    s = ch.Solver(ctx, prob, ch.Numerics(n_global=n, art=ch.ArtParams(enabled=art), precision=precision))
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def test_body_force_known_answer_on_gpu(ctx):
    n, L = (12, 12, 12), (2 * np.pi,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved((0.5 + 0 * X, 0.25 + 0 * X, -0.1 + 0 * X, 1.0 + 0 * X, 2.0 + 0 * X))
    s = _gpu(ctx, n, L, q, [(1.0, 2.0, -3.0)], art=False)
    s.compute_rhs(False)
    for c, v in ((2, 2.0), (3, 4.0), (4, -6.0), (5, 2.6)):
        assert np.abs(s.field(f"dQ[{c}]") - v).max() < 1e-10, c


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gravity_run_matches_oracle(ctx, precision):
    import compactles_hip as ch
    n, L = (32, 12, 12), (1.0, 0.5, 0.5)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    k = 2 * np.pi / L[1]
    q = cases.conserved((0 * X, 0.05 * np.sin(k * Z), 0.02 * np.cos(k * Y), 1.0 - 0.3 * X + 0.02 * np.cos(k * Y),
                         1.0 + 0.05 * np.sin(np.pi * X)))
    forces = [(-0.3, 0.0, 0.0), (0.0, 0.05, 0.0)]
    os_ = cases.oracle_solver(n, L, q, bcs=((1, 1), (0, 0), (0, 0)), mu0=1e-3, body_forces=forces)
    gs = _gpu(ctx, n, L, q, forces, bcs_x=(ch.SlipWallBC(), ch.SlipWallBC()), precision=precision, mu0=1e-3)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    if precision == "fp64":
        for c in range(5):
            assert Hh.relerr(gs.field(f"dQ[{c + 1}]"), os_.interior(os_.dQ[..., c])) <= 1e-12, c
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    Qg = gs.allocate_state()
    gs.download(Qg)
    err = _state_err([Hh.interior(Qg[..., c], n) for c in range(5)], [os_.interior(os_.Q[..., c]) for c in range(5)])
    if precision == "fp64":
        assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
        assert err <= 1e-10
    else:
        print(f"\nFP32-mixed gravity box, 20 steps vs FP64 oracle: Q {err:.2e}")
        assert err <= 1e-5
