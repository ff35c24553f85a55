"""Compact-scheme definitions: the host-side mirror of src/kernels.jl and
src/kernels_banded.jl (same names, same fields, same presets).

A CompactScheme is a tridiagonal-LHS operator
    alpha g[i-1] + g[i] + alpha g[i+1] = a0 f[i] + sum_m coeffs[m] (f[i+m] -/+ f[i-m]),
a BandedCompactScheme has LHS half-bandwidth q.  Closure rows are given for the
low edge and mirrored at the high edge by the plan builder.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from fractions import Fraction as Fr
from typing import List, Sequence, Tuple

from ._lib import MAXCL, MAXM, MAXRHS, SchemeDesc


@dataclass
class ClosureRow:
    """kernels.jl:28-31: (sub, diag, super) LHS and RHS weights from the edge."""
    lhs: Tuple[float, float, float]
    rhs: Sequence[float]


@dataclass
class BandedClosureRow:
    """kernels_banded.jl:19-22: full centred LHS band (2q+1) and RHS weights."""
    lhs: Sequence[float]
    rhs: Sequence[float]


class AbstractCompactScheme:
    def to_desc(self) -> SchemeDesc:
        raise NotImplementedError


@dataclass
class CompactScheme(AbstractCompactScheme):
    """kernels.jl:43-50."""
    name: str
    alpha: float
    a0: float
    coeffs: List[float]
    symmetric: bool
    closures: List[ClosureRow] = field(default_factory=list)

    def nclosure(self):
        return len(self.closures)

    def halfwidth(self):
        return len(self.coeffs)

    def to_desc(self) -> SchemeDesc:
        d = SchemeDesc()
        d.q = 1
        d.lhs[0] = float(self.alpha)
        _fill_common(d, self.a0, self.coeffs, self.symmetric)
        d.nclosure = len(self.closures)
        if d.nclosure > MAXCL:
            raise ValueError("too many closure rows")
        for j, row in enumerate(self.closures):
            for t in range(3):
                d.closure_lhs[j][t] = float(row.lhs[t])
            _fill_rhs(d, j, row.rhs)
        return d


@dataclass
class BandedCompactScheme(AbstractCompactScheme):
    """kernels_banded.jl:31-39."""
    name: str
    q: int
    lhs: List[float]
    a0: float
    coeffs: List[float]
    symmetric: bool
    closures: List[BandedClosureRow] = field(default_factory=list)

    def nclosure(self):
        return len(self.closures)

    def halfwidth(self):
        return len(self.coeffs)

    def to_desc(self) -> SchemeDesc:
        d = SchemeDesc()
        d.q = int(self.q)
        for s, v in enumerate(self.lhs):
            d.lhs[s] = float(v)
        _fill_common(d, self.a0, self.coeffs, self.symmetric)
        d.nclosure = len(self.closures)
        for j, row in enumerate(self.closures):
            if len(row.lhs) != 2 * self.q + 1:
                raise ValueError("banded closure LHS must have 2q+1 entries")
            for t, v in enumerate(row.lhs):
                d.closure_lhs[j][t] = float(v)
            _fill_rhs(d, j, row.rhs)
        return d


def _fill_common(d, a0, coeffs, symmetric):
    d.a0 = float(a0)
    if len(coeffs) > MAXM:
        raise ValueError("RHS half-width too large")
    d.M = len(coeffs)
    for m, c in enumerate(coeffs):
        d.coeffs[m] = float(c)
    d.symmetric = int(bool(symmetric))


def _fill_rhs(d, j, rhs):
    if len(rhs) > MAXRHS:
        raise ValueError("closure RHS too long")
    d.closure_rhs_len[j] = len(rhs)
    for k, v in enumerate(rhs):
        d.closure_rhs[j][k] = float(v)


def _f(x) -> float:
    return float(Fr(x))


def lele_d1_6() -> CompactScheme:
    """kernels.jl:61-67: sixth-order tridiagonal first derivative (Lele 1992)."""
    return CompactScheme("Lele C6 first derivative", _f(Fr(1, 3)), 0.0,
                         [_f(Fr(7, 9)), _f(Fr(1, 36))], False,
                         [ClosureRow((0.0, 1.0, 2.0), [-2.5, 2.0, 0.5]),
                          ClosureRow((0.25, 1.0, 0.25), [-0.75, 0.0, 0.75])])


def pade_d1_4() -> CompactScheme:
    """kernels.jl:75-79: fourth-order Pade first derivative."""
    return CompactScheme("Pade C4 first derivative", 0.25, 0.0, [0.75], False,
                         [ClosureRow((0.0, 1.0, 2.0), [-2.5, 2.0, 0.5])])


def compact_filter(alphaf: float = 0.45) -> CompactScheme:
    """kernels.jl:89-111: Gaitonde-Visbal eighth-order filter with the
    order-2/4/6 closure cascade."""
    af = float(alphaf)
    a0 = (93 + 70 * af) / 128
    a1 = (7 + 18 * af) / 16
    a2 = (-7 + 14 * af) / 32
    a3 = (1 - 2 * af) / 16
    a4 = (-1 + 2 * af) / 128
    b2 = ((1 + 2 * af) / 2, (1 + 2 * af) / 2)
    b4 = ((5 + 6 * af) / 8, (1 + 2 * af) / 2, (-1 + 2 * af) / 8)
    b6 = ((11 + 10 * af) / 16, (15 + 34 * af) / 32, (-3 + 6 * af) / 16, (1 - 2 * af) / 32)

    def ctr(c):
        L = len(c)
        return [c[m] / 2 for m in range(L - 1, 0, -1)] + [c[0]] + [c[m] / 2 for m in range(1, L)]

    cl = [ClosureRow((0.0, 1.0, 0.0), [1.0]),
          ClosureRow((af, 1.0, af), ctr(b2)),
          ClosureRow((af, 1.0, af), ctr(b4)),
          ClosureRow((af, 1.0, af), ctr(b6))]
    return CompactScheme("Gaitonde-Visbal C8 filter", af, a0, [a1 / 2, a2 / 2, a3 / 2, a4 / 2], True, cl)


def lele_d1_10() -> BandedCompactScheme:
    """kernels_banded.jl:54-64: tenth-order pentadiagonal first derivative."""
    return BandedCompactScheme(
        "Lele C10 first derivative", 2, [0.5, 0.05], 0.0,
        [_f(Fr(17, 24)), _f(Fr(101, 600)), _f(Fr(1, 600))], False,
        [BandedClosureRow([0, 0, 1, 2, 0], [-2.5, 2.0, 0.5]),
         BandedClosureRow([0, 0.25, 1, 0.25, 0], [-0.75, 0.0, 0.75]),
         BandedClosureRow([0, _f(Fr(1, 3)), 1, _f(Fr(1, 3)), 0],
                          [-_f(Fr(1, 36)), -_f(Fr(7, 9)), 0.0, _f(Fr(7, 9)), _f(Fr(1, 36))])])
