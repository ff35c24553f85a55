"""Serving loop: independent setup + run! requests (src/problem.jl:321,
src/timestep.jl:277) admitted into device memory by their footprint under an
HBM budget and run concurrently on one GPU (cles_server in the C ABI).  A
request's result is bitwise the same alone and in any mix."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .problem import Numerics, Problem, Solver, SolverSpec, initialize_

QUEUED, RUNNING, DONE, FAILED = range(4)


class Server:
    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, runners: int = 4, run_mode: str = "host_dt",
                 stack_limit: int = 8, stack_window: float = 0.005):
        """run_mode: how each request's solver issues run! (Solver.set_run_mode);
        stack_limit: same-shape two-dimensional requests queued together run
        as one Stack of up to this many (1: each alone), a request waiting up
        to stack_window seconds after its submission for the rest."""
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.cles_server_create(int(device), C.c_size_t(int(hbm_budget)), int(runners), C.byref(h)))
        self.handle = h
        self._specs = {}
        L.check(self.lib.cles_server_set_run_mode(self.handle, Solver.RUN_MODES[run_mode]))
        L.check(self.lib.cles_server_set_stack_limit(self.handle, int(stack_limit), float(stack_window)))

    def submit(self, prob: Problem, num: Numerics, tfinal: float, nmax: int = 2 ** 62, Q: np.ndarray | None = None):
        """Queue setup(prob, num) + run!(...; tfinal, nmax); Q (padded host
        state) defaults to prob.ic on the grid.  Returns the request id."""
        spec = SolverSpec(prob, num)
        if any(j is not None for j in spec.jac):
            raise ValueError("stretched grids are not available through the server (cles_server_submit takes "
                             "a descriptor only); run them with a Solver")
        if Q is None:
            Q = initialize_(spec, spec.allocate_state(), prob.ic)
        Q = np.asfortranarray(Q, dtype=np.float64)
        rid = C.c_int64()
        L.check(self.lib.cles_server_submit(self.handle, C.byref(spec.desc), Q.ctypes.data_as(L.DP),
                                            L.I64_3(*spec.padded), L.INT_3(*spec.halo), float(tfinal),
                                            int(min(nmax, 2 ** 62)), C.byref(rid)))
        self._specs[rid.value] = spec
        return rid.value

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = C.c_int()
        L.check(self.lib.cles_server_wait(self.handle, int(rid), float(timeout), C.byref(st)))
        return st.value

    def result(self, rid: int):
        """(Q, status, t, steps) of a finished request."""
        spec = self._specs[rid]
        Q = spec.allocate_state()
        status, t, steps = C.c_int(), C.c_double(), C.c_int64()
        L.check(self.lib.cles_server_result(self.handle, int(rid), Q.ctypes.data_as(L.DP), L.I64_3(*spec.padded),
                                            L.INT_3(*spec.halo), C.byref(status), C.byref(t), C.byref(steps)))
        return Q, status.value, t.value, steps.value

    def group(self, rid: int) -> int:
        """Requests of the stack a finished request ran in (1: alone)."""
        g = C.c_int()
        L.check(self.lib.cles_server_group(self.handle, int(rid), C.byref(g)))
        return g.value

    def timing(self, rid: int):
        """(t_submit, t_admit, t_done, footprint) in seconds since the server started."""
        a, b, c, fp = C.c_double(), C.c_double(), C.c_double(), C.c_size_t()
        L.check(self.lib.cles_server_timing(self.handle, int(rid), C.byref(a), C.byref(b), C.byref(c), C.byref(fp)))
        return a.value, b.value, c.value, fp.value

    def release(self, rid: int):
        L.check(self.lib.cles_server_release(self.handle, int(rid)))
        self._specs.pop(rid, None)

    def stats(self):
        used, run, q = C.c_size_t(), C.c_int(), C.c_int()
        L.check(self.lib.cles_server_stats(self.handle, C.byref(used), C.byref(run), C.byref(q)))
        return used.value, run.value, q.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cles_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
