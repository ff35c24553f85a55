"""StiffenedGas on the GPU: the perfect-gas limit runs as the IdealMixture
(dQ and 20 steps, 1e-12), and water (gamma 4.4, p_inf 6e8) against the oracle
on the same inputs (RHS 1e-12; 20 run! steps: dt 1e-12, state 1e-10; the
FP32-mixed form 1e-5), with an isothermal no-slip wall whose energy is
rho cv T_wall + p_inf (boundary.jl:213-214)."""
import numpy as np
import pytest

import cases
import helpers as Hh
import oracle as O
from test_gpu_batch import _state_err

pytestmark = pytest.mark.gpu
WATER = (4.4, 6.0e8, 1816.0)


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _gpu(ctx, n, L, q, eos, xbcs=None, precision="fp64", art=True, mu0=0.0):
    import compactles_hip as ch
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=(xbcs or per, per, per),
                      ic=lambda X, Y, Z: None, transport=ch.Transport(mu0=mu0), eos=eos)
    s = ch.Solver(ctx, prob, ch.Numerics(n_global=n, art=ch.ArtParams(enabled=art), precision=precision))
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def _q(gs, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    return [Hh.interior(Qg[..., c], n) for c in range(5)]


def test_perfect_gas_limit_runs_as_ideal(ctx):
    import compactles_hip as ch
    n, L = (24, 20, 16), (2 * np.pi,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.smooth_prim(X, Y, Z))
    a = _gpu(ctx, n, L, q, ch.single_species(gamma=1.4, R=1.0), mu0=1e-3)
    b = _gpu(ctx, n, L, q, ch.StiffenedGas(gamma=1.4, p_inf=0.0, cv=2.5), mu0=1e-3)
    a.compute_rhs(False)
    b.compute_rhs(False)
    for c in range(5):
        assert Hh.relerr(b.field(f"dQ[{c + 1}]"), a.field(f"dQ[{c + 1}]")) <= 1e-12, c
    d1 = a.run(1e9, 20, record=20)
    d2 = b.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _state_err(_q(b, n), _q(a, n)) <= 1e-11


def _water_pair(ctx, precision="fp64", wall=False):
    import compactles_hip as ch
    eos = ch.StiffenedGas(*WATER)
    n, L = (32, 12, 12), (1.0, 0.5, 0.5)
    X, Y, Z = cases.grid(n, L, periodic=(not wall, True, True))
    k = 2 * np.pi / L[1]
    prim = ch.Prim(u=(2.0 * np.sin(np.pi * X) if wall else 1.0 * np.sin(2 * np.pi * X), 0.5 * np.cos(k * Z), 0 * X),
                   p=101325.0 * (1 + 0.5 * np.sin(2 * np.pi * X) * np.cos(k * Y)), rho=1000.0 + 0 * X)
    q = ch.conserved_from_prim(eos, prim)
    bcs = ((O.BC_NOSLIP, O.BC_NOSLIP), (0, 0), (0, 0)) if wall else None
    Tw = {(0, 0): 290.0, (0, 1): 310.0} if wall else None
    os_ = cases.oracle_solver(n, L, q, stiffened=WATER, mu0=1e-3, bcs=bcs, Twall=Tw)
    gs = _gpu(ctx, n, L, q, eos, xbcs=(ch.NoSlipWallBC(Twall=290.0), ch.NoSlipWallBC(Twall=310.0)) if wall else None,
              precision=precision, mu0=1e-3)
    return os_, gs, n


@pytest.mark.parametrize("wall", [False, True])
def test_water_matches_oracle(ctx, wall):
    os_, gs, n = _water_pair(ctx, wall=wall)
    os_.apply_bcs()
    gs.apply_bcs()
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    # the momentum RHS as one vector on one scale: a transverse component that
    # is small next to the pressure terms of the same fluxes carries their rounding
    got = [gs.field(f"dQ[{c + 1}]") for c in range(5)]
    ref = [os_.interior(os_.dQ[..., c]) for c in range(5)]
    assert _state_err(got, ref) <= 1e-12
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _state_err(_q(gs, n), [os_.interior(os_.Q[..., c]) for c in range(5)]) <= 1e-10
    if wall:
        assert np.abs(gs.field("T")[0] - 290.0).max() < 1e-6 and np.abs(gs.field("T")[-1] - 310.0).max() < 1e-6


def test_water_fp32_mixed(ctx):
    os_, gs, n = _water_pair(ctx, precision="fp32")
    os_.run(1e9, 10)
    gs.run(1e9, 10)
    err = _state_err(_q(gs, n), [os_.interior(os_.Q[..., c]) for c in range(5)])
    print(f"\nFP32-mixed water 32x12x12, 10 steps vs FP64 oracle: Q {err:.2e}")
    assert err <= 1e-5


def test_nonpositive_density_point_keeps_the_placeholder(ctx):
    """A point with rho <= 0 takes _primitives!' placeholder state
    (physics.jl:242: rho = 1); the e-sensor of compute_artificial! divides by
    that rho (artificial.jl:321), so kappa_art stays finite around the point
    and equals the oracle's."""
    import compactles_hip as ch
    n, L = (32, 12, 12), (1.0, 0.5, 0.5)
    X, Y, Z = cases.grid(n, L)
    k = 2 * np.pi / L[1]
    eos = ch.StiffenedGas(*WATER)
    prim = ch.Prim(u=(1.0 * np.sin(2 * np.pi * X), 0.5 * np.cos(k * Z), 0 * X),
                   p=101325.0 * (1 + 0.5 * np.sin(2 * np.pi * X) * np.cos(k * Y)), rho=1000.0 + 0 * X)
    q = [np.array(c) for c in ch.conserved_from_prim(eos, prim)]
    q[0][10, 5, 6] = -1.0e-3
    os_ = cases.oracle_solver(n, L, q, stiffened=WATER, mu0=1e-3)
    gs = _gpu(ctx, n, L, q, eos, mu0=1e-3)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    ref = os_.interior(os_.field("kappa_art"))
    got = gs.field("kappa_art")
    assert np.all(np.isfinite(got))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
