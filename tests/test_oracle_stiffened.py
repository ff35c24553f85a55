"""The oracle's StiffenedGas (physics.jl:160-250), pinned by the reference's
tests (runtests.jl:908-950, 1376-1382): p_inf = 0 reproduces the perfect gas
(conserved state 1e-14 on the host; p, T, c 1e-12 out of the primitives); water
(gamma 4.4, p_inf 6e8) at one atmosphere keeps p to 1e-9 and
c = sqrt(gamma (p + p_inf) / rho) to 1e-12 (1400 < c < 1700); freestream below
1e-8 p; an acoustic pulse stays finite; the three Prim spellings agree."""
import numpy as np

import cases


def _solver(prim_fn, stiff, n=(12, 12, 12), L=(1.0, 1.0, 1.0)):
    import compactles_hip as ch
    X, Y, Z = cases.grid(n, L)
    eos = ch.StiffenedGas(gamma=stiff[0], p_inf=stiff[1], cv=stiff[2])
    q = ch.conserved_from_prim(eos, prim_fn(X, Y, Z))
    return cases.oracle_solver(n, L, q, art=False, stiffened=stiff), eos


def test_perfect_gas_limit():
    import compactles_hip as ch
    g, R = 1.4, 1.0
    cv = R / (g - 1)
    sg = ch.StiffenedGas(gamma=g, p_inf=0.0, cv=cv, name="gas")
    assert sg.nspecies() == 1
    pr = ch.Prim(u=(0.3, -0.1, 0.2), p=0.8, T_ion=1.7)
    for a, b in zip(ch.conserved_from_prim(ch.single_species(gamma=g, R=R), pr), ch.conserved_from_prim(sg, pr)):
        assert abs(a - b) <= 1e-14 * abs(a) + 1e-300
    s, _ = _solver(lambda X, Y, Z: ch.Prim(u=(0.3 + 0 * X, -0.1 + 0 * X, 0.2 + 0 * X), p=0.8 + 0 * X,
                                           T_ion=1.7 + 0 * X), (g, 0.0, cv))
    s.max_rate()
    I = (3, 4, 5)
    assert abs(s.interior(s.field("p"))[I] - 0.8) <= 1e-12 * 0.8
    assert abs(s.interior(s.field("T"))[I] - 1.7) <= 1e-12 * 1.7
    assert abs(s.interior(s.field("c"))[I] - np.sqrt(g * R * 1.7)) <= 1e-12 * np.sqrt(g * R * 1.7)


def test_water():
    import compactles_hip as ch
    water = (4.4, 6.0e8, 1816.0)
    s, eos = _solver(lambda X, Y, Z: ch.Prim(u=(0 * X, 0 * X, 0 * X), p=101325.0 + 0 * X, rho=1000.0 + 0 * X), water)
    s.max_rate()
    I = (3, 4, 5)
    p, c = s.interior(s.field("p"))[I], s.interior(s.field("c"))[I]
    assert abs(p - 101325.0) <= 1e-9 * 101325.0
    assert 1400 < c < 1700
    assert abs(c - np.sqrt(4.4 * (101325.0 + 6.0e8) / 1000.0)) <= 1e-12 * c
    s.apply_bcs()
    s.compute_rhs(False)
    assert np.abs(s.interior(s.dQ)).max() < 1e-8 * 101325.0
    s2, _ = _solver(lambda X, Y, Z: ch.Prim(u=(0 * X, 0 * X, 0 * X), p=101325.0 * (1 + 0.01 * np.sin(2 * np.pi * X)),
                                            rho=1000.0 + 0 * X), water)
    s2.run(1e9, 5)
    s2.max_rate()
    assert np.isfinite(s2.Q).all()
    assert np.abs(s2.interior(s2.field("p")) - 101325.0).max() < 0.05 * 101325.0


def test_prim_spellings_agree():
    import compactles_hip as ch
    liquid = ch.StiffenedGas(gamma=4.4, p_inf=6.0e8, cv=1816.0)
    p_liquid = 1000.0 * liquid.gas_constant() * 300.0 - liquid.p_inf
    a = ch.conserved_from_prim(liquid, ch.Prim(rho=1000.0, T_ion=300.0))
    b = ch.conserved_from_prim(liquid, ch.Prim(p=p_liquid, T_ion=300.0))
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-12 * max(abs(x), 1e-300)
