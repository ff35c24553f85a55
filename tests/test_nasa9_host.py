"""Nasa9Mixture on the host and in the oracle (CPU), pinned by the reference's
own known-answer tests (runtests.jl:765-900):
 - the calorically perfect limit (only a3 set) reproduces the ideal gas to
   round-off: cp, h, e, conserved_from_prim, and p, T, c, cp out of the
   primitives (where the Newton inversion runs);
 - a genuinely varying cp: dh/dT = cp, e(T) inverted by the Newton solve,
   (p, T, Y) recovered from a state initialised from them, steps stay finite;
 - the CEA reader on a subset of the reference's thermo.inp: CO2's R and
   intervals, cp(300 K) = 845.7241586606974 J/(kg K), sensible h(298.15) = 0,
   formation h(298.15) = -393510 J/mol, continuous joins, round trips."""
import math
import os

import numpy as np
import pytest

import cases

THERMO = os.path.join(os.path.dirname(__file__), "golden", "thermo_subset.inp")
FAKE_A = (1.2e4, -50.0, 3.6, 6.0e-4, -1.0e-7, 1.0e-11, -4.0e-16)


def _fake(ch):
    sp = ch.Nasa9Species("fake", 287.0, a=FAKE_A, b1=-1.0e3)
    return ch.Nasa9Mixture([sp, ch.nasa9_constant_cp("inert", 200.0, 900.0)], T_guess=500.0)


def _oracle_uniform(eos, prim, n=(12, 12, 12)):
    import compactles_hip as ch
    X = np.zeros(n)
    q = ch.conserved_from_prim(eos, ch.Prim(u=tuple(u + 0 * X for u in prim["u"]), p=prim["p"] + 0 * X,
                                            T_ion=prim["T"] + 0 * X, Y=tuple(y + 0 * X for y in prim["Y"])))
    s = cases.oracle_solver(n, (1.0, 1.0, 1.0), q, **cases.nasa9_oracle_kw(eos))
    s.max_rate()   # runs the primitives
    return s


def test_constant_cp_limit_is_the_ideal_gas():
    import compactles_hip as ch
    g, R = 1.4, 1.0
    cp = g * R / (g - 1)
    ideal = ch.single_species(gamma=g, R=R)
    poly = ch.Nasa9Mixture([ch.nasa9_constant_cp("gas", R, cp)])
    for T in (0.3, 1.0, 7.5, 300.0):
        assert math.isclose(poly.species_cp(0, T), cp, rel_tol=1e-14)
        assert math.isclose(poly.species_enthalpy(0, T), cp * T, rel_tol=1e-14)
        assert math.isclose(poly.species_energy(0, T), (cp - R) * T, rel_tol=1e-14)
    pr = ch.Prim(u=(0.3, -0.1, 0.2), p=0.8, T_ion=1.7)
    for a, b in zip(ch.conserved_from_prim(ideal, pr), ch.conserved_from_prim(poly, pr)):
        assert math.isclose(float(a), float(b), rel_tol=1e-14)
    s = _oracle_uniform(poly, {"u": (0.3, 0.0, 0.0), "p": 0.8, "T": 1.7, "Y": (1.0,)})
    I = (6, 6, 6)
    assert math.isclose(s.interior(s.field("p"))[I], 0.8, rel_tol=1e-12)
    assert math.isclose(s.interior(s.field("T"))[I], 1.7, rel_tol=1e-12)
    assert math.isclose(s.interior(s.field("c"))[I], math.sqrt(g * R * 1.7), rel_tol=1e-12)
    assert math.isclose(s.interior(s.field("cp"))[I], cp, rel_tol=1e-12)


def test_varying_cp_consistency():
    import compactles_hip as ch
    eos = _fake(ch)
    for T in (250.0, 800.0, 2500.0):
        d = 1e-4 * T
        dh = (eos.species_enthalpy(0, T + d) - eos.species_enthalpy(0, T - d)) / (2 * d)
        assert math.isclose(dh, eos.species_cp(0, T), rel_tol=1e-7)
    assert eos.species_cp(0, 2500.0) / eos.species_cp(0, 300.0) > 1.2
    assert all(eos.species_cp(0, T) > 287.0 for T in np.arange(250.0, 2501.0, 50.0))
    for Y in ((1.0, 0.0), (0.5, 0.5), (0.2, 0.8)):
        for T in (250.0, 800.0, 2500.0):
            e = sum(Y[k] * eos.species_energy(k, T) for k in range(2))
            assert math.isclose(eos.mixture_temperature(e, lambda k: Y[k]), T, rel_tol=1e-12)
    s = _oracle_uniform(eos, {"u": (120.0, 0.0, 0.0), "p": 2.5e5, "T": 1400.0, "Y": (0.35, 0.65)})
    I = (3, 4, 5)
    assert math.isclose(s.interior(s.field("p"))[I], 2.5e5, rel_tol=1e-10)
    assert math.isclose(s.interior(s.field("T"))[I], 1400.0, rel_tol=1e-10)
    assert math.isclose(s.interior(s.field("Y[2]"))[I], 0.65, rel_tol=1e-12)
    s.run(1e9, 3)
    assert np.isfinite(s.Q).all()


def test_cea_reader_and_reference_values():
    import compactles_hip as ch
    he, co2 = ch.read_nasa9(["He", "CO2"], THERMO)
    assert (he.name, co2.name) == ("He", "CO2")
    assert math.isclose(co2.R, 8.31446261815324 / 44.0095e-3, rel_tol=1e-14)
    assert [(iv.Tmin, iv.Tmax) for iv in co2.intervals] == [(200.0, 1000.0), (1000.0, 6000.0), (6000.0, 20000.0)]
    sens = ch.Nasa9Mixture([co2])
    assert abs(sens.species_enthalpy(0, 298.15)) <= 1e-6
    assert math.isclose(sens.species_cp(0, 300.0), 845.7241586606974, rel_tol=1e-13)
    for lo, hi in zip(co2.intervals, co2.intervals[1:]):
        assert math.isclose(lo.cp_over_R(lo.Tmax), hi.cp_over_R(lo.Tmax), rel_tol=5e-7)
        assert math.isclose(lo.h_over_R(lo.Tmax), hi.h_over_R(lo.Tmax), rel_tol=5e-7)
    form = ch.Nasa9Mixture([ch.read_nasa9("CO2", THERMO, reference="formation")])
    assert abs(form.species_enthalpy(0, 298.15) * 44.0095e-3 - (-393510.0)) <= 5.0
    assert form.species_cp(0, 300.0) == sens.species_cp(0, 300.0)
    for eos in (sens, form):
        for T in (220.0, 300.0, 999.0, 1001.0, 1500.0, 5999.0, 6001.0, 10000.0):
            e = eos.species_energy(0, T)
            assert math.isclose(eos.mixture_temperature(e, lambda k: 1.0), T, rel_tol=2e-13)
    Y = (0.3, 0.7)
    for ref in ("sensible", "formation"):
        mix = ch.Nasa9Mixture(ch.read_nasa9(["He", "CO2"], THERMO, reference=ref))
        for T in (300.0, 1400.0, 7000.0):
            e = sum(Y[k] * mix.species_energy(k, T) for k in range(2))
            assert math.isclose(mix.mixture_temperature(e, lambda k: Y[k]), T, rel_tol=2e-13)
    with pytest.raises(ValueError, match="reference"):
        ch.read_nasa9("CO2", THERMO, reference="unknown")
    with pytest.raises(ValueError, match="not found"):
        ch.read_nasa9("not-a-CEA-species", THERMO)


def test_oracle_primitives_match_the_frontend_inversion():
    # the oracle's Newton (C) and the frontend's (numpy) invert the same e(T)
    import compactles_hip as ch
    mix = ch.Nasa9Mixture(ch.read_nasa9(["He", "CO2"], THERMO), T_guess=300.0)
    s = _oracle_uniform(mix, {"u": (30.0, -10.0, 5.0), "p": 1.0e5, "T": 900.0, "Y": (0.3, 0.7)})
    I = (2, 3, 4)
    assert math.isclose(s.interior(s.field("T"))[I], 900.0, rel_tol=1e-12)
    cp = 0.3 * mix.species_cp(0, 900.0) + 0.7 * mix.species_cp(1, 900.0)
    assert math.isclose(s.interior(s.field("cp"))[I], cp, rel_tol=1e-13)
