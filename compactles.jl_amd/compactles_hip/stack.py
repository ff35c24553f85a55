"""Stacks of same-shape two-dimensional problems on one GPU.

The reference advances one problem per setup / run! (src/problem.jl:321,
src/timestep.jl:277).  A Stack holds `count` problems of one Problem and
Numerics on an (nx, ny, 1) grid that differ in their initial state, end time
and step cap, side by side in one solver: every kernel launch covers all of
them, and each follows run! on its own (its own dt, check_step floors, tfinal
clip and filter cadence; a failing instance stops with its status).  An
instance's result does not depend on the rest of the stack.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence, Union

import numpy as np

from . import _lib as L
from .device import Context
from .problem import Numerics, Problem, SolverSpec


class Stack:
    def __init__(self, ctx: Context, prob: Problem, num: Numerics, count: int):
        self.ctx, self.count = ctx, int(count)
        self.spec = SolverSpec(prob, num)
        if self.spec.n[2] != 1:
            raise ValueError("a stack holds two-dimensional problems (n_global[3] == 1)")
        if any(j is not None for j in self.spec.jac):
            raise ValueError("a stack holds problems on uniform grids")
        h = C.c_void_p()
        L.check(ctx.lib.cles_stack_create(ctx.handle, C.byref(self.spec.desc), self.count, C.byref(h)))
        self.handle = h

    def allocate_state(self) -> np.ndarray:
        return self.spec.allocate_state()

    def upload(self, i: int, Q: np.ndarray):
        """Instance i's padded host state (the layout of SolverSpec.allocate_state)."""
        Q = np.asfortranarray(Q, dtype=np.float64)
        assert Q.shape == tuple(self.spec.padded) + (self.spec.n_cons,)
        L.check(self.ctx.lib.cles_stack_upload(self.handle, int(i), Q.ctypes.data_as(L.DP),
                                               L.I64_3(*self.spec.padded), L.INT_3(*self.spec.halo)))

    def download(self, i: int) -> np.ndarray:
        Q = self.spec.allocate_state()
        L.check(self.ctx.lib.cles_stack_download(self.handle, int(i), Q.ctypes.data_as(L.DP),
                                                 L.I64_3(*self.spec.padded), L.INT_3(*self.spec.halo)))
        return Q

    def run(self, tfinal: Union[float, Sequence[float]], nmax: Union[int, Sequence[int]] = 2 ** 62,
            max_steps: int = 2 ** 62) -> int:
        """run!(...; tfinal[i], nmax[i]) on every instance; the number of
        stack iterations issued."""
        tf = np.broadcast_to(np.asarray(tfinal, dtype=np.float64), (self.count,)).copy()
        nm = np.broadcast_to(np.asarray(nmax, dtype=np.int64), (self.count,)).copy()
        taken = C.c_int64()
        L.check(self.ctx.lib.cles_stack_run(self.handle, tf.ctypes.data_as(L.DP),
                                            nm.ctypes.data_as(C.POINTER(C.c_int64)), int(min(max_steps, 2 ** 62)),
                                            C.byref(taken)))
        return taken.value

    def state(self):
        """(t, steps, status) arrays over the instances."""
        t = np.zeros(self.count)
        st = np.zeros(self.count, dtype=np.int64)
        status = np.zeros(self.count, dtype=np.int32)
        L.check(self.ctx.lib.cles_stack_state(self.handle, t.ctypes.data_as(L.DP),
                                              st.ctypes.data_as(C.POINTER(C.c_int64)),
                                              status.ctypes.data_as(C.POINTER(C.c_int))))
        return t, st, status

    def close(self):
        if getattr(self, "handle", None):
            L.check(self.ctx.lib.cles_stack_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
