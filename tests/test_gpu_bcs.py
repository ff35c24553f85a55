"""NoSlipWallBC, ExtrapolationBC and SwitchableBC on the GPU against the oracle
on the same inputs: viscous runs between no-slip walls (adiabatic; isothermal
with an IdealMixture and with a Nasa9Mixture), an extrapolating outflow, and a
switchable face (bitwise equal to the wrapped condition until switched; after
a switch at step 4, by hand or from a callback, equal to the oracle doing the
same: dt 1e-12, state 1e-10)."""
import math

import numpy as np
import pytest

import cases
import helpers as Hh
import oracle as O
from test_gpu_batch import _state_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _gpu(ctx, n, L, q, xbcs, eos=None, mu0=0.0, art=True, cfl=0.5):
    import compactles_hip as ch
    per = (ch.PeriodicBC(), ch.PeriodicBC())
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=(xbcs, per, per), ic=lambda X, Y, Z: None,
                      transport=ch.Transport(mu0=mu0), **({"eos": eos} if eos is not None else {}))
    s = ch.Solver(ctx, prob, ch.Numerics(n_global=n, art=ch.ArtParams(enabled=art), cfl=cfl))
    Q = s.allocate_state()
    sl = tuple(slice(s.halo[k], s.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    s.upload(Q)
    return s


def _q_err(gs, os_, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    return _state_err([Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)],
                      [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)])


def _shear(n, L):
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    k = 2 * np.pi / L[1]
    s = np.sin(np.pi * X / L[0])
    return X, cases.conserved((0.05 * s * np.cos(k * Y), 0.2 * np.cos(np.pi * X / L[0]) + 0.05 * np.sin(k * Z),
                               0.1 * s, 1 + 0.05 * np.cos(k * Y) * s, 1 + 0.03 * s * np.sin(k * Z)))


@pytest.mark.parametrize("Twall", [math.nan, 1.3])
def test_noslip_walls_match_oracle(ctx, Twall):
    import compactles_hip as ch
    n, L = (32, 16, 12), (1.0, 0.5, 0.4)
    _, q = _shear(n, L)
    T = {} if math.isnan(Twall) else {(0, 0): Twall, (0, 1): Twall}
    os_ = cases.oracle_solver(n, L, q, bcs=((O.BC_NOSLIP, O.BC_NOSLIP), (0, 0), (0, 0)), Twall=T, mu0=2e-3)
    gs = _gpu(ctx, n, L, q, (ch.NoSlipWallBC(Twall=Twall), ch.NoSlipWallBC(Twall=Twall)), mu0=2e-3)
    os_.apply_bcs()
    gs.apply_bcs()
    Qg = gs.allocate_state()
    gs.download(Qg)
    assert np.abs(Hh.interior(Qg[..., 4], n)[0]).max() > 0
    assert _q_err(gs, os_, n) <= 1e-14
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    u = gs.field("u")
    assert np.abs(u[0]).max() == 0 and np.abs(u[-1]).max() == 0
    if T:
        assert np.abs(gs.field("T")[0] - Twall).max() < 1e-12


def test_noslip_isothermal_nasa9_matches_oracle(ctx):
    import compactles_hip as ch
    he, co2 = ch.read_nasa9(["He", "CO2"], "tests/golden/thermo_subset.inp")
    eos = ch.Nasa9Mixture([he, co2], T_guess=600.0)
    n, L = (32, 16, 12), (0.01, 0.005, 0.004)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    s = np.sin(np.pi * X / L[0])
    q = ch.conserved_from_prim(eos, ch.Prim(u=(5 * s, 30 * np.cos(np.pi * X / L[0]), 0 * X), p=1e5 + 0 * X,
                                            T_ion=700 + 200 * s, Y=(0.3 + 0 * X, 0.7 + 0 * X)))
    os_ = cases.oracle_solver(n, L, q, bcs=((O.BC_NOSLIP, O.BC_NOSLIP), (0, 0), (0, 0)),
                              Twall={(0, 0): 500.0, (0, 1): 900.0}, mu0=2e-5, **cases.nasa9_oracle_kw(eos))
    gs = _gpu(ctx, n, L, q, (ch.NoSlipWallBC(Twall=500.0), ch.NoSlipWallBC(Twall=900.0)), eos=eos, mu0=2e-5)
    d1 = os_.run(1e9, 15, record=15)
    d2 = gs.run(1e9, 15, record=15)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    T = gs.field("T")
    assert abs(T[0].mean() - 500.0) < 1e-9 and abs(T[-1].mean() - 900.0) < 1e-9


def test_extrapolation_outflow_matches_oracle(ctx):
    import compactles_hip as ch
    n, L = (32, 16, 12), (1.0, 0.5, 0.4)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    k = 2 * np.pi / L[1]
    q = cases.conserved((0.3 + 0 * X, 0.05 * np.sin(k * Y), 0 * X, 1 + 0.05 * np.exp(-40 * (X - 0.5) ** 2), 1 + 0 * X))
    os_ = cases.oracle_solver(n, L, q, bcs=((O.BC_EXTRAP, O.BC_EXTRAP), (0, 0), (0, 0)), mu0=1e-3, art=False)
    gs = _gpu(ctx, n, L, q, (ch.ExtrapolationBC(), ch.ExtrapolationBC()), mu0=1e-3, art=False)
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10


def _pulse_q(n, L):
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    return cases.conserved((0.2 + 0 * X, 0 * X, 0 * X, 1 + 0.1 * np.exp(-40 * (X - 0.5) ** 2), 1.0 + 0 * X))


def test_switchable_bc(ctx):
    import compactles_hip as ch
    n, L = (16, 12, 12), (1.0, 1.0, 1.0)
    q = _pulse_q(n, L)
    ref = _gpu(ctx, n, L, q, (ch.SlipWallBC(), ch.SlipWallBC()), art=False, cfl=0.4)
    ref.run(1e9, 12)
    sw = (ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()), ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()))
    wrap = _gpu(ctx, n, L, q, sw, art=False, cfl=0.4)
    wrap.run(1e9, 12)
    Qr, Qw = ref.allocate_state(), wrap.allocate_state()
    ref.download(Qr)
    wrap.download(Qw)
    assert np.array_equal(Qr, Qw) and not ch.switched(sw[0])
    # switched by hand at step 4, against the oracle doing the same
    sw2 = (ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()),
           ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()))
    gs = _gpu(ctx, n, L, q, sw2, art=False, cfl=0.4)
    os_ = cases.oracle_solver(n, L, q, bcs=((O.BC_SLIPWALL, O.BC_SLIPWALL), (0, 0), (0, 0)), art=False, cfl=0.4,
                              switch={(0, 0): O.BC_EXTRAP, (0, 1): O.BC_EXTRAP})
    gs.run(1e9, 4)
    os_.run(1e9, 4)
    for b in sw2:
        ch.switch_(b)
    assert ch.switch_(sw2[0]) is sw2[0] and all(ch.switched(b) for b in sw2)
    os_.switch_bc(0, 0)
    os_.switch_bc(0, 1)
    d1 = os_.run(1e9, 12, record=8)
    d2 = gs.run(1e9, 12, record=8)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    Qs = gs.allocate_state()
    gs.download(Qs)
    assert not np.array_equal(Qs, Qr)
    # switched from a callback between steps: the same trajectory, bitwise
    sw3 = (ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()),
           ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC()))
    cb_s = _gpu(ctx, n, L, q, sw3, art=False, cfl=0.4)
    cb = ch.Callback(ch.WhenState(lambda s, Q: s.step_count >= 4), lambda s, Q: ([ch.switch_(b) for b in sw3], None)[1])
    cb_s.run(1e9, 12, callback=cb)
    Qc = cb_s.allocate_state()
    cb_s.download(Qc)
    assert np.array_equal(Qc, Qs)
