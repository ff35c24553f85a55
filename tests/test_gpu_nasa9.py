"""Nasa9Mixture on the GPU (physics.jl:402-571) against the oracle on the same
inputs, through the C ABI:
 - the calorically perfect limit reproduces the IdealMixture run (dQ 1e-12,
   20 run! steps: dt history 1e-12, state 1e-11);
 - He / CO2 from the CEA thermo.inp subset (three intervals each, the state
   crossing the 1000 K join): primitives (T, p, c, cp 1e-12), one RHS (dQ
   1e-12), 20 run! steps (dt history 1e-12, state and kinetic energy 1e-10);
 - NSCBC inflow / outflow faces with a temperature-dependent cp, where the
   inflow's d(phi)/dY_k takes cv_k(T) (physics.jl:504-510, nscbc.jl:278-285);
 - batched uneven tubes of two-interval gases against the oracle per tube;
 - the FP32-mixed form against the FP64 oracle (state 1e-5)."""
import os

import numpy as np
import pytest

import cases
import helpers as Hh
from test_gpu_batch import _state_err
from test_gpu_solver import _compare_rhs
from test_oracle_nscbc import TWO_SPECIES

pytestmark = pytest.mark.gpu
THERMO = os.path.join(os.path.dirname(__file__), "golden", "thermo_subset.inp")


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _q_err(gs, os_, n):
    Qg = gs.allocate_state()
    gs.download(Qg)
    got = [Hh.interior(Qg[..., c], n) for c in range(os_.n_cons)]
    ref = [os_.interior(os_.Q[..., c]) for c in range(os_.n_cons)]
    return _state_err(got, ref)


def _prim(ch, X, Y, Z, L, T0, dT, p0, U):
    """A smooth periodic state of two species: T swings over T0 +- dT."""
    k = [2 * np.pi / L[d] for d in range(3)]
    u = U * (np.sin(k[0] * X) * np.cos(k[1] * Y) + 0.2 * np.cos(k[2] * Z))
    v = -U * np.cos(k[0] * X) * np.sin(k[1] * Y)
    w = 0.3 * U * np.sin(k[0] * X + k[2] * Z)
    T = T0 + dT * np.sin(k[0] * X + 0.5) * np.cos(k[1] * Y - k[2] * Z)
    p = p0 * (1 + 0.05 * np.cos(k[1] * Y) * np.sin(k[2] * Z))
    th = 0.5 + 0.4 * np.sin(k[0] * X) * np.sin(k[1] * Y + k[2] * Z)
    return ch.Prim(u=(u, v, w), p=p, T_ion=T, Y=(1 - th, th))


def _cea(ch):
    return ch.Nasa9Mixture(ch.read_nasa9(["He", "CO2"], THERMO), T_guess=600.0)


def _varying(ch):
    """Two gases of nondimensional temperature scale, cp rising with T; the
    second joins a constant-cp interval at T = 3 (cp and h continuous:
    cp/R = 3.5 + 0.6 - 0.045 and h/R = 10.5 + 0.9 - 0.045 there)."""
    a1 = (0.0, 0.0, 2.5, 0.3, 0.0, 0.0, 0.0)
    lo = ch.Nasa9Interval(0.01, 3.0, (0.0, 0.0, 3.5, 0.2, -0.005, 0.0, 0.0), 0.0)
    hi = ch.Nasa9Interval(3.0, 1000.0, (0.0, 0.0, 4.055, 0.0, 0.0, 0.0, 0.0), -0.81)
    return ch.Nasa9Mixture([ch.Nasa9Species("A", 1.0, a=a1, Tmin=0.01, Tmax=1000.0),
                            ch.Nasa9Species("B", 0.25, [lo, hi])], T_guess=2.0)


def test_constant_cp_nasa9_runs_as_the_ideal_mixture(ctx):
    import compactles_hip as ch
    n, L = (24, 20, 16), (1.0, 1.0, 1.0)
    X, Y, Z = cases.grid(n, L)
    ideal = ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(TWO_SPECIES)])
    poly = ch.Nasa9Mixture([ch.nasa9_constant_cp(f"s{k}", R, g * R / (g - 1)) for k, (R, g) in enumerate(TWO_SPECIES)],
                           T_guess=1.0)
    q = ch.conserved_from_prim(ideal, _prim(ch, X, Y, Z, L, 1.2, 0.3, 1.0, 0.3))
    a = cases.gpu_solver(ctx, n, L, q, species=TWO_SPECIES, mu0=2e-3)
    b = cases.gpu_solver(ctx, n, L, q, eos=poly, mu0=2e-3)
    a.compute_rhs(False)
    b.compute_rhs(False)
    for c in range(6):
        assert Hh.relerr(b.field(f"dQ[{c + 1}]"), a.field(f"dQ[{c + 1}]")) <= 1e-12, c
    d1 = a.run(1e9, 20, record=20)
    d2 = b.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    Qa, Qb = a.allocate_state(), b.allocate_state()
    a.download(Qa)
    b.download(Qb)
    assert _state_err([Qb[..., c] for c in range(6)], [Qa[..., c] for c in range(6)]) <= 1e-11


def _cea_pair(ctx, n, precision="fp64"):
    import compactles_hip as ch
    eos = _cea(ch)
    L = (0.01, 0.008, 0.006)
    X, Y, Z = cases.grid(n, L)
    q = ch.conserved_from_prim(eos, _prim(ch, X, Y, Z, L, 900.0, 350.0, 1.0e5, 60.0))
    os_ = cases.oracle_solver(n, L, q, mu0=2e-5, Pr=0.7, Sc=0.7, **cases.nasa9_oracle_kw(eos))
    gs = cases.gpu_solver(ctx, n, L, q, eos=eos, mu0=2e-5, precision=precision)
    return os_, gs


def test_cea_primitives_rhs_and_run_match_oracle(ctx):
    n = (24, 20, 16)
    os_, gs = _cea_pair(ctx, n)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    T = os_.interior(os_.field("T"))
    assert T.min() < 1000.0 < T.max()          # both sides of the CEA join
    _compare_rhs(gs, os_, n, fields=("rho", "u", "p", "T", "c", "cp", "mu_art", "beta_art", "kappa_art"))
    r1, _ = os_.max_rate()
    r2, _ = gs.max_rate()
    assert abs(r1 - r2) <= 1e-13 * r1
    d1 = os_.run(1e9, 20, record=20)
    d2 = gs.run(1e9, 20, record=20)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    assert abs(ke1 - ke2) <= 1e-10 * ke1
    assert abs(en1 - en2) <= 1e-10 * en1


def test_cea_fp32_mixed_error(ctx):
    n = (24, 20, 16)
    os_, gs = _cea_pair(ctx, n, precision="fp32")
    os_.run(1e9, 10)
    gs.run(1e9, 10)
    err = _q_err(gs, os_, n)
    print(f"\nFP32-mixed He/CO2 NASA-9 24x20x16, 10 steps vs FP64 oracle: Q {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("faces_kind", ["inflow_outflow", "outflow_only"])
def test_nscbc_with_varying_cp_matches_oracle(ctx, faces_kind):
    import compactles_hip as ch
    eos = _varying(ch)
    n, L = (40, 12, 10), (1.3, 0.5, 0.4)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    ky, kz = 2 * np.pi / L[1], 2 * np.pi / L[2]
    u = 0.3 + 0.05 * np.sin(3 * X) * np.cos(ky * Y) + 0.02 * np.cos(kz * Z)
    v = 0.04 * np.cos(2 * X + ky * Y) + 0.03 * np.sin(kz * Z)
    w = 0.03 * np.cos(X) * np.sin(ky * Y + kz * Z)
    p = 1.0 + 0.08 * np.cos(2.5 * X) * np.sin(ky * Y) + 0.03 * np.cos(kz * Z)
    rho = 1.1 + 0.1 * np.sin(1.7 * X + ky * Y) + 0.02 * np.sin(kz * Z)
    th = 0.5 * (1 + np.tanh((X - 0.65) / 0.2 + 0.3 * np.sin(ky * Y)))
    q = ch.conserved_from_prim(eos, ch.Prim(u=(u, v, w), p=p, rho=rho, Y=(1 - th, th)))
    out = {"pinf": 0.97, "sigma": 0.3, "beta_t": -1.0}
    if faces_kind == "inflow_outflow":
        faces = {(0, 0): {"u": (0.32, 0.01, 0.0), "T": 1.6, "Y": (0.9, 0.1), "eta_u": 0.28, "eta_T": 0.35,
                          "eta_t": 0.22, "eta_Y": 0.3}, (0, 1): out}
    else:
        faces = {(0, 0): dict(out, pinf=1.02), (0, 1): out}
    bcs = ((1, 1), (0, 0), (0, 0))
    os_ = cases.oracle_solver(n, L, q, art=True, bcs=cases.nscbc_codes(bcs, faces), nscbc=faces,
                              **cases.nasa9_oracle_kw(eos))
    gs = cases.gpu_solver(ctx, n, L, q, eos=eos, art=True, bcs=bcs, nscbc=faces)
    os_.compute_rhs(False)
    gs.compute_rhs(False)
    for c in range(os_.n_cons):
        ref = os_.interior(os_.dQ[..., c])
        assert Hh.relerr(gs.field(f"dQ[{c + 1}]"), ref) <= 1e-12, (c, Hh.relerr(gs.field(f"dQ[{c + 1}]"), ref))
    d1 = os_.run(1e9, 15, record=15)
    d2 = gs.run(1e9, 15, record=15)
    assert np.abs(d1 - d2).max() <= 1e-12 * d1.max()
    assert _q_err(gs, os_, n) <= 1e-10


TUBES = cases.tube_specs(4, seed=11, nmin=128, nmax=700)


def test_batched_tubes_with_nasa9_match_oracle(ctx):
    import compactles_hip as ch
    eos = _varying(ch)
    b = cases.tube_batch(ctx, TUBES, eos=eos)
    b.compute_rhs(False)
    for i, sp in enumerate(TUBES):
        o = cases.tube_oracle(sp, eos=eos)
        o.compute_rhs(False)
        ref = [o.interior(o.dQ[..., c])[:, 0, 0] for c in range(o.n_cons)]
        got = [b.field(i, f"dQ[{c + 1}]") for c in range(o.n_cons)]
        assert _state_err(got, ref) <= 1e-12, (i, sp["n"], _state_err(got, ref))
    nsteps = 20
    b = cases.tube_batch(ctx, TUBES, eos=eos, nmax=nsteps)
    _, dts = b.run(record=nsteps)
    t, steps, status = b.state()
    assert (status == 0).all() and (steps == nsteps).all()
    for i, sp in enumerate(TUBES):
        o = cases.tube_oracle(sp, eos=eos)
        d1 = o.run(1e9, nsteps, record=nsteps)
        T = o.interior(o.field("T"))
        assert T.max() > 3.0 > T.min()          # both intervals of species B in use
        assert np.abs(dts[i] - d1).max() <= 1e-12 * d1.max(), (i, sp["n"])
        ref = [o.interior(o.Q[..., c])[:, 0, 0] for c in range(o.n_cons)]
        got = [b.field(i, f"Q[{c + 1}]") for c in range(o.n_cons)]
        assert _state_err(got, ref) <= 1e-10, (i, sp["n"], _state_err(got, ref))
