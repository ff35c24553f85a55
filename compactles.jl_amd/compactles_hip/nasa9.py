"""Nasa9Mixture: multicomponent thermally ideal gas with NASA-9 heat
capacities (src/physics.jl:280-571), and a reader of NASA CEA `thermo.inp`
records (src/nasa9_data.jl, the format of NASA TP-2002-211556 Appendix A).

The host side builds the tables the device evaluates (cles_solver_desc
.nasa9) and restates the thermodynamics in numpy for initial conditions
(conserved_from_prim) and checks.  Powers are formed by multiplication, as on
the device."""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np

R_UNIVERSAL = 8.31446261815324          # J / (mol K), as the reference derives R from the molar mass
NASA9_CONTINUITY_RTOL = 2e-3            # physics.jl:312: join tolerance of cp and h
MAX_INTERVALS = 3                       # intervals per species the device tables hold


@dataclass
class Nasa9Interval:
    """physics.jl:280-288: one fit on [Tmin, Tmax]; a = (a1, ..., a7), b1 the enthalpy constant."""
    Tmin: float
    Tmax: float
    a: tuple
    b1: float

    def cp_over_R(self, T):
        a = self.a
        T2 = T * T
        T3, T4 = T2 * T, T2 * T2
        return a[0] / T2 + a[1] / T + a[2] + a[3] * T + a[4] * T2 + a[5] * T3 + a[6] * T4

    def h_over_R(self, T):
        a = self.a
        T2 = T * T
        T3, T4 = T2 * T, T2 * T2
        T5 = T4 * T
        return (-a[0] / T + a[1] * np.log(T) + a[2] * T + a[3] * T2 / 2 + a[4] * T3 / 3 + a[5] * T4 / 4 +
                a[6] * T5 / 5 + self.b1)


def _validate(name, R, intervals):
    """physics.jl:314-338 _validate_nasa9_intervals."""
    if not intervals:
        raise ValueError(f"NASA-9 species {name} has no intervals")
    if not R > 0:
        raise ValueError(f"NASA-9 species {name} has non-positive R")
    for iv in intervals:
        if not (iv.Tmin > 0 and iv.Tmin < iv.Tmax):
            raise ValueError(f"invalid NASA-9 interval for {name}")
    for lo, hi in zip(intervals, intervals[1:]):
        scale_T = max(abs(lo.Tmax), abs(hi.Tmin), 1.0)
        if abs(lo.Tmax - hi.Tmin) > 1e-10 * scale_T:
            raise ValueError(f"non-contiguous NASA-9 intervals for {name}")
        Tj = (lo.Tmax + hi.Tmin) / 2
        cl, ch = lo.cp_over_R(Tj), hi.cp_over_R(Tj)
        if abs(cl - ch) > NASA9_CONTINUITY_RTOL * max(abs(cl), abs(ch), 1.0):
            raise ValueError(f"discontinuous NASA-9 cp for {name} at {Tj} K")
        hl, hh = lo.h_over_R(Tj), hi.h_over_R(Tj)
        if abs(hl - hh) > NASA9_CONTINUITY_RTOL * max(abs(hl), abs(hh), Tj):
            raise ValueError(f"discontinuous NASA-9 h for {name} at {Tj} K")


class Nasa9Species:
    """physics.jl:340-389: Nasa9Species(name, R, intervals), or with a single
    interval from a, b1, Tmin, Tmax."""

    def __init__(self, name, R, intervals=None, *, a=None, b1=0.0, Tmin=200.0, Tmax=6000.0):
        if intervals is None:
            intervals = [Nasa9Interval(float(Tmin), float(Tmax), tuple(float(x) for x in a), float(b1))]
        intervals = list(intervals)
        _validate(name, R, intervals)
        if len(intervals) > MAX_INTERVALS:
            raise ValueError(f"NASA-9 species {name}: at most {MAX_INTERVALS} intervals are supported")
        self.name, self.R, self.intervals = str(name), float(R), intervals

    def interval(self, T):
        for iv in self.intervals[:-1]:
            if T <= iv.Tmax:
                return iv
        return self.intervals[-1]


def nasa9_constant_cp(name, R, cp):
    """physics.jl:391-400: the calorically perfect limit (only a3 = cp/R)."""
    return Nasa9Species(name, float(R), a=(0.0, 0.0, float(cp) / float(R), 0.0, 0.0, 0.0, 0.0), b1=0.0)


class Nasa9Mixture:
    """physics.jl:402-430; T_guess seeds the temperature inversion."""

    def __init__(self, sp: Sequence[Nasa9Species], T_guess=300.0):
        self.sp = list(sp)
        if not self.sp:
            raise ValueError("Nasa9Mixture needs at least one species")
        self.T_guess = float(T_guess)

    @property
    def Rk(self):
        return [s.R for s in self.sp]

    def nspecies(self):
        return len(self.sp)

    def species_names(self):
        return [s.name for s in self.sp]

    # physics.jl:438-456, scalar T
    def species_cp(self, k, T):
        s = self.sp[k]
        return s.R * s.interval(T).cp_over_R(T)

    def species_enthalpy(self, k, T):
        s = self.sp[k]
        return s.R * s.interval(T).h_over_R(T)

    def species_energy(self, k, T):
        return self.species_enthalpy(k, T) - self.Rk[k] * T

    def mixture_temperature(self, e, Yat):
        """physics.jl:459-482: bounded Newton on sum_k Y_k e_k(T) = e."""
        n = len(self.sp)
        e0 = cv0 = 0.0
        for k in range(n):
            Yk = Yat(k)
            e0 += Yk * self.species_energy(k, self.T_guess)
            cv0 += Yk * (self.species_cp(k, self.T_guess) - self.Rk[k])
        T = min(max(self.T_guess + (e - e0) / cv0, 1e-3), 1e9) if cv0 > 0 else self.T_guess
        for _ in range(30):
            f, cvm = -e, 0.0
            for k in range(n):
                Yk = Yat(k)
                f += Yk * self.species_energy(k, T)
                cvm += Yk * (self.species_cp(k, T) - self.Rk[k])
            if not cvm > 0:
                break
            dl = f / cvm
            T = max(T - dl, 0.25 * T)
            if abs(dl) <= 1e-14 * T:
                break
        return T

    def to_desc(self, d):
        """Fill the Nasa9Mixture fields of a cles_solver_desc."""
        from . import _lib as L
        d.eos_kind = L.EOS_NASA9
        d.T_guess = self.T_guess
        for k, s in enumerate(self.sp):
            d.R[k] = s.R
            d.gamma[k] = 0.0
            t = d.nasa9[k]
            t.n_intervals = len(s.intervals)
            for i, iv in enumerate(s.intervals):
                t.iv[i].Tmin, t.iv[i].Tmax, t.iv[i].b1 = iv.Tmin, iv.Tmax, iv.b1
                for j in range(7):
                    t.iv[i].a[j] = iv.a[j]


def _vectorised(eos: Nasa9Mixture, fn, k, T):
    T = np.asarray(T, dtype=np.float64)
    if T.ndim == 0:
        return fn(k, float(T))
    return np.vectorize(lambda x: fn(k, x), otypes=[np.float64])(T)


def nasa9_conserved(eos: Nasa9Mixture, pr):
    """problem.jl:76-90 for a Nasa9Mixture: (rho Y_k..., rho u, rho v, rho w,
    rho E) with E = sum_k Y_k e_k(T) + |u|^2 / 2."""
    N = eos.nspecies()
    if len(pr.Y) != N:
        raise ValueError(f"Prim carries {len(pr.Y)} mass fractions; EOS has {N} species")
    Rm = 0.0
    for k in range(N):
        Rm = Rm + pr.Y[k] * eos.Rk[k]
    rho = pr.p / (Rm * pr.T_ion) if _isnan(pr.rho) else pr.rho
    T = pr.p / (rho * Rm) if _isnan(pr.T_ion) else pr.T_ion
    e = 0.0
    for k in range(N):
        e = e + pr.Y[k] * _vectorised(eos, eos.species_energy, k, T)
    u1, u2, u3 = pr.u
    ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3)
    out = [rho * pr.Y[k] for k in range(N)]
    return out + [rho * u1, rho * u2, rho * u3, rho * (e + ke)]


def _isnan(x):
    return isinstance(x, float) and math.isnan(x)


# ---------------------------------------------------------------------------
# NASA CEA thermo.inp reader (nasa9_data.jl; record layout of NASA
# TP-2002-211556 Appendix A: name line; count / formula / phase / molar mass
# (cols 53-65) / heat of formation (cols 66-80); per interval a temperature
# line and two coefficient lines, b1 in cols 49-64 of the second)
# ---------------------------------------------------------------------------
def _f(s):
    s = s.strip().replace("D", "E").replace("d", "e")
    return float(s) if s else 0.0


def _parse(path):
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if not ln.strip() or ln.startswith("!") or ln.lower().startswith("thermo"):
            i += 1
            continue
        if ln.startswith("END"):
            i += 1
            continue
        if i + 1 >= len(lines):
            break
        head = lines[i + 1]
        try:
            nint = int(head[0:2])
        except ValueError:   # the global temperature-range line after `thermo`
            i += 1
            continue
        name = ln[0:24].split()[0]
        M = _f(head[52:65])
        if nint == 0:
            out.setdefault(name, (M, None))
            i += 3
            continue
        ivs = []
        for k in range(nint):
            tl, c1, c2 = lines[i + 2 + 3 * k], lines[i + 3 + 3 * k], lines[i + 4 + 3 * k]
            Tmin, Tmax = _f(tl[0:11]), _f(tl[11:21])
            a = [_f(c1[16 * j:16 * j + 16]) for j in range(5)] + [_f(c2[0:16]), _f(c2[16:32])]
            ivs.append(Nasa9Interval(Tmin, Tmax, tuple(a), _f(c2[48:64])))
        out.setdefault(name, (M, ivs))
        i += 2 + 3 * nint
    return out


def read_nasa9(names, path, reference="sensible", T_ref=298.15):
    """nasa9_data.jl read_nasa9: species from a CEA thermo.inp file.
    reference "sensible" shifts every interval's b1 by one constant so that
    h(T_ref) = 0; "formation" keeps the tabulated gauge."""
    if reference not in ("sensible", "formation"):
        raise ValueError("reference must be 'sensible' or 'formation'")
    if not T_ref > 0:
        raise ValueError("T_ref must be positive")
    single = isinstance(names, str)
    table = _parse(path)
    out = []
    for nm in ([names] if single else names):
        if nm not in table:
            raise ValueError(f"NASA CEA species not found: {nm}")
        M, ivs = table[nm]
        if ivs is None:
            raise ValueError(f"{nm}: reactant-only record carrying a heat of formation, no temperature intervals")
        sp = Nasa9Species(nm, R_UNIVERSAL / (M * 1e-3), ivs)
        if reference == "sensible":
            shift = -sp.interval(T_ref).h_over_R(T_ref)
            sp = Nasa9Species(nm, sp.R, [Nasa9Interval(iv.Tmin, iv.Tmax, iv.a, iv.b1 + shift) for iv in ivs])
        out.append(sp)
    return out[0] if single else out
