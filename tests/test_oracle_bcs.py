"""The oracle's other wall conditions, pinned by the reference's tests:
 - NoSlipWallBC: adiabatic zeroes the wall velocity, isothermal also sets
   T_ion = Twall (runtests.jl:439-461; 1e-12), for an IdealMixture and, through
   e_k(Twall), a Nasa9Mixture;
 - ExtrapolationBC copies the adjacent interior plane exactly
   (runtests.jl:463-475);
 - SwitchableBC: unswitched it is bitwise the condition it wraps; switched at
   step 4 it takes the other arm; switching twice is switching once; the
   periodicity and fold restrictions (runtests.jl:1630-1672)."""
import math

import numpy as np
import pytest

import cases
import oracle as O

NS = ((O.BC_NOSLIP, O.BC_NOSLIP), (0, 0), (0, 0))


def _uniform_solver(bcs, n=(24, 12, 12), L=(1.0, 0.4, 0.4), prim=None, **kw):
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    prim = prim or (lambda X, Y, Z: (0.3 + 0 * X, -0.2 + 0 * X, 0.1 + 0 * X, 1.0 + 0 * X, 1.0 + 0 * X))
    q = cases.conserved(prim(X, Y, Z))
    return cases.oracle_solver(n, L, q, bcs=bcs, art=False, **kw)


@pytest.mark.parametrize("Twall", [math.nan, 2.5])
def test_noslip_wall(Twall):
    T = {} if math.isnan(Twall) else {(0, 0): Twall, (0, 1): Twall}
    s = _uniform_solver(NS, Twall=T)
    s.apply_bcs()
    s.max_rate()        # primitives of the enforced state
    for f in ("u", "v", "w"):
        a = s.interior(s.field(f))
        assert np.abs(a[0]).max() < 1e-12 and np.abs(a[-1]).max() < 1e-12, f
        assert np.abs(a[5] - {"u": 0.3, "v": -0.2, "w": 0.1}[f]).max() < 1e-14   # interior untouched
    if T:
        Tf = s.interior(s.field("T"))
        assert np.abs(Tf[0] - Twall).max() < 1e-12 and np.abs(Tf[-1] - Twall).max() < 1e-12


def test_noslip_isothermal_nasa9():
    import compactles_hip as ch
    he, co2 = ch.read_nasa9(["He", "CO2"], "tests/golden/thermo_subset.inp")
    eos = ch.Nasa9Mixture([he, co2], T_guess=600.0)
    n, L = (24, 12, 12), (0.01, 0.004, 0.004)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    q = ch.conserved_from_prim(eos, ch.Prim(u=(30 + 0 * X, 0 * X, 0 * X), p=1e5 + 0 * X, T_ion=900 + 0 * X,
                                            Y=(0.3 + 0 * X, 0.7 + 0 * X)))
    s = cases.oracle_solver(n, L, q, bcs=NS, art=False, Twall={(0, 0): 400.0, (0, 1): 1300.0},
                            **cases.nasa9_oracle_kw(eos))
    s.apply_bcs()
    s.max_rate()
    Tf = s.interior(s.field("T"))
    assert np.abs(Tf[0] - 400.0).max() < 1e-9 and np.abs(Tf[-1] - 1300.0).max() < 1e-9


def test_extrapolation_copies_the_adjacent_plane():
    bcs = ((O.BC_EXTRAP, O.BC_EXTRAP), (0, 0), (0, 0))
    s = _uniform_solver(bcs, prim=lambda X, Y, Z: (0.2 + X, 0 * X, 0 * X, 1 + 0.1 * X, 1 + 0.3 * X))
    s.apply_bcs()
    Q = s.interior(s.Q)
    assert np.abs(Q[0] - Q[1]).max() == 0 and np.abs(Q[-1] - Q[-2]).max() == 0
    assert np.abs(Q[2] - Q[1]).max() > 0


def _pulse(bcs, switch=None):
    n, L = (16, 12, 12), (1.0, 1.0, 1.0)
    X, Y, Z = cases.grid(n, L, periodic=(False, True, True))
    q = cases.conserved((0.2 + 0 * X, 0 * X, 0 * X, 1 + 0.1 * np.exp(-40 * (X - 0.5) ** 2), 1.0 + 0 * X))
    return cases.oracle_solver(n, L, q, bcs=bcs, art=False, cfl=0.4, switch=switch)


def test_switchable_is_transparent_until_switched():
    slip = ((O.BC_SLIPWALL, O.BC_SLIPWALL), (0, 0), (0, 0))
    sw = {(0, 0): O.BC_EXTRAP, (0, 1): O.BC_EXTRAP}
    ref = _pulse(slip)
    ref.run(1e9, 12)
    wrap = _pulse(slip, switch=sw)
    wrap.run(1e9, 12)
    assert np.array_equal(wrap.Q, ref.Q)
    swd = _pulse(slip, switch=sw)
    swd.run(1e9, 4)
    swd.switch_bc(0, 0)
    swd.switch_bc(0, 1)
    swd.switch_bc(0, 1)          # idempotent
    swd.run(1e9, 12)
    assert not np.array_equal(swd.Q, ref.Q)
    # the switched run equals one whose faces were extrapolating from step 4 on
    ext = _pulse(slip, switch=sw)
    ext.run(1e9, 4)
    Q4 = ext.Q.copy()
    t4 = ext.time
    plain = _pulse(((O.BC_EXTRAP, O.BC_EXTRAP), (0, 0), (0, 0)))
    plain.Q[...] = Q4
    plain.set_time(*t4)
    plain.run(1e9, 12)
    assert np.array_equal(swd.Q, plain.Q)
    with pytest.raises(ValueError, match="not a SwitchableBC"):
        ref.switch_bc(0, 0)


def test_switchable_restrictions():
    slip = ((O.BC_SLIPWALL, O.BC_SLIPWALL), (0, 0), (0, 0))
    with pytest.raises(ValueError, match="periodicity"):
        _pulse(slip, switch={(0, 0): O.BC_PERIODIC})
    with pytest.raises(ValueError, match="fold"):
        _pulse(slip, switch={(0, 0): O.BC_AXIS})


def test_switchable_frontend_restrictions():
    """runtests.jl:1633-1637 on the host frontend."""
    import compactles_hip as ch
    with pytest.raises(ValueError, match="periodicity"):
        ch.SwitchableBC(ch.SlipWallBC(), ch.PeriodicBC())
    with pytest.raises(ValueError, match="fold"):
        ch.SwitchableBC(ch.AxisBC(), ch.SlipWallBC())
    with pytest.raises(ValueError, match="fold"):
        ch.SwitchableBC(ch.SlipWallBC(), ch.OriginBC())
    b = ch.SwitchableBC(ch.SlipWallBC(), ch.ExtrapolationBC())
    assert not ch.switched(b)
    assert ch.switch_(b) is b and ch.switched(b)
