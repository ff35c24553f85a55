"""Pins the oracle's whole run! pipeline (primitives, Cook sensors, fluxes,
divergence, LSRK, state filter, dt control, Dirichlet and slip-wall ends) by
the reference's own shock-capturing validation battery (test/validation.jl,
cases from test/cases.jl) at the values the reference measured on itself:

  Lax        L1 rho 5.01e-3, u 7.53e-3, p 7.57e-3 (against the exact Riemann solution)
  Shu-Osher  L1 rho 6.84e-3, wave train 2.08e-2, train peak 4.682 (against test/refs/shu_osher.csv)
  Woodward   L1 rho 3.15e-2, peak rho 6.595 at x = 0.7785 (against test/refs/woodward_colella.csv)
  Noh nu=1   plateau 3.9968 / 4, shock 0.2043 / 0.2, wall deficit 58 %

(test/validation.jl:38-46).  The reference states four significant figures
as the reproducible limit; each value is held to the digits it is quoted
with.  The stored profiles are the reference's own test fixtures, copied
unchanged by tests/golden/make_validation_refs.py.

Also the fourth-order temporal accuracy of the LSRK(5,4) tables
(timestep.jl:5-21) on a linear problem with a closed-form semi-discrete
solution.
"""
import os

import numpy as np
import pytest

import cases
import oracle as O
import riemann
import shock_cases as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def within(v, quoted, ndigits_after):
    """v rounds to `quoted` at the quoted precision (half a unit of the last digit)."""
    return abs(v - quoted) <= 0.5 * 10.0 ** (-ndigits_after) * (1 + 1e-9)


def test_lax_matches_reference_measurement():
    x, r, u, p, ok = S.oracle_run(S.lax())
    assert ok
    R, U, P = riemann.profile(x, S.LAX_T, 0.5, S.LAX_L, S.LAX_R, 1.4)
    er, eu, ep = S.l1(r, R), S.l1(u, U), S.l1(p, P)
    assert er < 1e-2 and eu < 1.5e-2 and ep < 1.5e-2          # test/validation.jl:114-116
    assert within(er * 1e3, 5.01, 2) and within(eu * 1e3, 7.53, 2) and within(ep * 1e3, 7.57, 2), (er, eu, ep)


def test_shu_osher_matches_reference_measurement():
    x, r, u, p, ok = S.oracle_run(S.shu_osher())
    assert ok
    xr, rr, _, _ = S.read_ref(os.path.join(GOLD, "shu_osher.csv"))
    ref = S.interp1(xr, rr, x)
    band = S.so_band(x)
    e, eb, peak = S.l1(r, ref), S.l1(r[band], ref[band]), r[band].max()
    assert e < 1.5e-2 and eb < 4e-2 and peak > 4.3           # test/validation.jl:131-133
    assert within(e * 1e3, 6.84, 2) and within(eb * 1e2, 2.08, 2) and within(peak, 4.682, 3), (e, eb, peak)


def test_woodward_colella_matches_reference_measurement():
    x, r, u, p, ok = S.oracle_run(S.woodward())
    assert ok
    xr, rr, _, _ = S.read_ref(os.path.join(GOLD, "woodward_colella.csv"))
    e = S.l1(r, S.interp1(xr, rr, x))
    i = int(np.argmax(r))
    assert np.all(np.isfinite(r)) and r.min() > 0 and e < 6e-2 and 0.75 < x[i] < 0.80   # :150-152
    assert within(e * 1e2, 3.15, 2) and abs(r[i] - 6.595) <= 1e-3 and within(x[i], 0.7785, 4), (e, r[i], x[i])


def test_noh_planar_matches_reference_measurement():
    x, r, u, p, ok = S.oracle_run(S.noh1())
    assert ok
    plat, deficit, shock, epre = S.noh_metrics(x, r, 1)
    assert abs(plat / 4 - 1) < 0.01 and 0 < deficit < 0.7 and abs(shock - 0.2) < 0.025 and epre < 5e-2  # :185-188
    assert within(plat, 3.9968, 4) and within(shock, 0.2043, 4) and within(100 * deficit, 58, 0), \
        (plat, shock, deficit)


# ---- LSRK(5,4) order --------------------------------------------------------
def _advect(nsteps, T=1.0, n=32, k=3, a=1.0, b=0.3, u=1.0):
    """Two identical ideal gases in a uniform stream (rho, u, p uniform, no
    viscosity, art off): the partial densities obey the linear semi-discrete
    system d(rho Y_k)/dt = -u D (rho Y_k) with the C6 operator D, whose exact
    solution for a + b sin(kx) is a + b sin(kx - u kappa t), kappa the C6
    modified wavenumber (Lele 1992)."""
    L = 2 * np.pi
    h = L / n
    x = np.arange(n) * h
    rho, p = 2.0, 1.0
    r1 = a + b * np.sin(k * x)
    q = [r1, rho - r1, rho * u + 0 * x, 0 * x, 0 * x, p / 0.4 + 0.5 * rho * u * u + 0 * x]
    s = cases.oracle_solver((n, 1, 1), (L, 1.0, 1.0), [c.reshape(n, 1, 1) for c in q],
                            species=((1.0, 1.4), (1.0, 1.4)), art=False, mu0=0.0)
    dt = T / nsteps
    for _ in range(nsteps):
        s.step(dt)
    got = s.interior(s.Q[..., 0])[:, 0, 0]
    w = k * h
    kappa = (14 / 9 * np.sin(w) + 1 / 18 * np.sin(2 * w)) / (1 + 2 / 3 * np.cos(w)) / h
    exact = a + b * np.sin(k * x - u * kappa * T)
    return np.abs(got - exact).max()


def test_lsrk_is_fourth_order_in_time():
    e1, e2, e3 = _advect(12), _advect(24), _advect(48)
    o1, o2 = np.log2(e1 / e2), np.log2(e2 / e3)
    assert 3.8 < o1 < 4.4 and 3.8 < o2 < 4.4, (e1, e2, e3, o1, o2)
    assert e3 < 1e-6


@pytest.mark.parametrize("typo", [
    # two RKB entries swapped
    ("5161836677717.0 / 13612068292357.0,\n    1720146321549.0 / 2090206949498.0",
     "1720146321549.0 / 2090206949498.0,\n    5161836677717.0 / 13612068292357.0"),
    # a digit dropped from an RKA numerator
    ("-2404267990393.0 / 2016746695238.0", "-240426799039.0 / 2016746695238.0"),
    # two leading digits transposed in an RKB numerator
    ("3134564353537.0 / 4481467310338.0", "1334564353537.0 / 4481467310338.0")])
def test_lsrk_order_check_catches_a_table_typo(tmp_path, typo):
    """A swapped or mistyped LSRK entry, planted in a scratch copy of the
    oracle, takes the observed order away from four.  (A change in the last
    digits of a 13-digit ratio moves results by ~1e-12 and is below what any
    order test, or the parity tolerance, resolves.)"""
    import shutil
    import subprocess
    here = os.path.dirname(O.__file__)
    src = open(os.path.join(here, "oracle.c")).read()
    assert src.count(typo[0]) == 1
    (tmp_path / "oracle.c").write_text(src.replace(typo[0], typo[1]))
    shutil.copy(os.path.join(here, "oracle.h"), tmp_path / "oracle.h")
    so = tmp_path / "liboracle_typo.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", "-ffp-contract=off", "-shared", "-o", str(so),
                    str(tmp_path / "oracle.c"), "-lm"], check=True)
    with O.library(str(so)):
        e1, e2 = _advect(24), _advect(48)
    o = np.log2(e1 / e2)
    assert not 3.8 < o < 4.4, o
