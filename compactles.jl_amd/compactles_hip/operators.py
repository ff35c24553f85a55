"""Directional compact operators: host mirror of the reference's plugin API
(src/operators.jl plan_direction / apply_along! / filter_field!) on device
arrays, one rank (whole-domain lines)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import F32, F64, I64_3, INT_3, check
from .device import Context, DeviceArray
from .schemes import AbstractCompactScheme


@dataclass
class Decomp:
    """The single-rank part of src/decomposition.jl:19-37 that plans need:
    interior extents, periodicity and the per-dimension halo pad (n_halo on
    active dimensions, 0 on collapsed ones)."""
    n_local: tuple
    periodic: tuple
    n_halo: int = 4

    @property
    def active(self):
        return tuple(n > 1 for n in self.n_local)

    @property
    def n_halo_d(self):
        return tuple(self.n_halo if a else 0 for a in self.active)

    @property
    def padded(self):
        return tuple(n + 2 * h for n, h in zip(self.n_local, self.n_halo_d))


def field(ctx: Context, decomp: Decomp, dtype=np.float64) -> DeviceArray:
    """decomposition.jl:129 field(decomp): a zeroed padded device field."""
    a = DeviceArray(ctx, decomp.padded, dtype)
    a.zero()
    return a


class DirPlan:
    """Device plan for one scheme along one dimension (DirPlan / BandPlan)."""

    def __init__(self, ctx: Context, scheme: AbstractCompactScheme, dim: int, n_local, periodic: bool,
                 h: float, dtype=np.float64, lo_fold=None, hi_fold=None):
        self.ctx = ctx
        self.scheme = scheme
        self.dim = dim
        self.n_local = tuple(int(x) for x in n_local)
        self.dtype = np.dtype(dtype)
        self.lo_fold, self.hi_fold = lo_fold, hi_fold
        desc = scheme.to_desc()
        self._desc = desc
        hp = C.c_void_p()
        check(ctx.lib.cles_plan_create(ctx.handle, C.byref(desc), dim, I64_3(*self.n_local), int(bool(periodic)),
                                       float(h), F64 if self.dtype == np.float64 else F32,
                                       _fold_arg(lo_fold), _fold_arg(hi_fold), C.byref(hp)))
        self.handle = hp

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.ctx.lib.cles_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _fold_arg(v) -> int:
    """The reference's lo_fold / hi_fold keyword (nothing, +1 or -1) as the ABI's int."""
    if v is None:
        return 0
    if v not in (1, -1):
        raise ValueError(f"a fold sign must be nothing, +1 or -1, got {v!r}")
    return int(v)


def plan_direction(ctx: Context, decomp: Decomp, scheme: AbstractCompactScheme, dim: int, h: float,
                   dtype=np.float64, lo_fold=None, hi_fold=None) -> DirPlan:
    """operators.jl:44-101 / operators_banded.jl:29-125, with the parity-fold
    keywords: lo_fold / hi_fold = sigma_g (the output's sign across a
    half-offset mirror at that end) or None.  The edge rows read the mirrored
    interior with the input's sign, sigma_g for a filter and -sigma_g for a
    derivative (folds.jl:271-274), so no halo needs filling."""
    return DirPlan(ctx, scheme, dim, decomp.n_local, decomp.periodic[dim - 1], h, dtype, lo_fold, hi_fold)


def apply_along_(out: DeviceArray, plan: DirPlan, f: DeviceArray, decomp: Decomp, stream=None):
    """operators.jl:199-215 apply_along!(out, plan, f, decomp): writes the
    interior of `out`; periodic lines wrap inside the interior."""
    if out.dtype != plan.dtype or f.dtype != plan.dtype:
        raise TypeError("array dtype does not match the plan")
    if f.shape != decomp.padded or out.shape != decomp.padded:
        raise ValueError("arrays must have the decomposition's padded shape")
    check(plan.ctx.lib.cles_apply_along(plan.handle, out.ptr, f.ptr, I64_3(*decomp.padded),
                                        INT_3(*decomp.n_halo_d), stream))
    return out


def apply_along_timed(out: DeviceArray, plan: DirPlan, f: DeviceArray, decomp: Decomp, iters: int) -> float:
    ms = C.c_double()
    check(plan.ctx.lib.cles_apply_along_timed(plan.handle, out.ptr, f.ptr, I64_3(*decomp.padded),
                                              INT_3(*decomp.n_halo_d), iters, C.byref(ms)))
    return ms.value


def filter_field_(f: DeviceArray, plans, decomp: Decomp):
    """operators.jl:223-231 filter_field!: filter along every active dimension
    in place (halos are not needed: periodic lines wrap in the interior)."""
    for d in range(3):
        if decomp.active[d]:
            apply_along_(f, plans[d], f, decomp)
    return f
