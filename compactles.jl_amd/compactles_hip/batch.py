"""Batches of independent one-dimensional problems on one GPU.

The reference advances one problem per setup / run! (src/problem.jl:321,
src/timestep.jl:277).  A Batch holds many one-dimensional problems
(n_global = (n, 1, 1)) that share the physics and numerics of one Problem and
Numerics and differ in length, domain, initial state, boundary states, end
time and step cap.  Each instance follows run! on its own (its own dt, floors
and tfinal clip); one that fails a check_step floor stops with its status and
the others go on.  An instance's result does not depend on the rest of the
batch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib as L
from .device import Context
from .problem import DirichletBC, Numerics, PeriodicBC, Prim, Problem, SlipWallBC, conserved_from_prim, \
    fill_common_desc


@dataclass
class Instance:
    """One problem of a batch: n nodes on `domain`, initial state ic(x, y, z)
    -> Prim, constant Dirichlet states `lo` / `hi` (Prim) when the Problem's x
    ends are DirichletBC, run!(...; tfinal, nmax)."""
    n: int
    domain: tuple
    ic: Callable
    tfinal: float = float("inf")
    nmax: int = 2 ** 62
    lo: Optional[Prim] = None
    hi: Optional[Prim] = None


class Batch:
    def __init__(self, ctx: Context, prob: Problem, num: Numerics, instances: Sequence[Instance]):
        xb = prob.bcs[0]
        for d in (1, 2):
            if not all(isinstance(bc, PeriodicBC) for bc in prob.bcs[d]):
                raise ValueError("batched instances are one-dimensional: y and z must be PeriodicBC (collapsed)")
        for bc in xb:
            if not isinstance(bc, (PeriodicBC, SlipWallBC, DirichletBC)):
                raise ValueError("batched instances take PeriodicBC, SlipWallBC or DirichletBC ends")
        self.ctx, self.prob, self.num = ctx, prob, num
        self.eos = prob.eos
        self.n_cons = self.eos.nspecies() + 4
        self.periodic = isinstance(xb[0], PeriodicBC)
        self.instances = list(instances)
        d = L.SolverDesc()
        fill_common_desc(d, prob, num)
        for k in range(3):
            d.n_global[k] = 1
            d.L_domain[k] = 1.0
            for s in range(2):
                d.bc[k][s] = prob.bcs[k][s].code
        arr = (L.BatchInstance * len(self.instances))()
        for i, ins in enumerate(self.instances):
            a = arr[i]
            a.n = int(ins.n)
            a.origin = float(ins.domain[0])
            a.L = float(ins.domain[1] - ins.domain[0])
            a.tfinal = float(ins.tfinal)
            a.nmax = int(min(ins.nmax, 2 ** 62))
            for s, (bc, prim) in enumerate(zip(xb, (ins.lo, ins.hi))):
                if isinstance(bc, DirichletBC):
                    if prim is None:
                        raise ValueError(f"instance {i}: Dirichlet ends need the `lo` and `hi` states")
                    q = conserved_from_prim(self.eos, prim)
                    for c, v in enumerate(q):
                        a.bc_state[s][c] = float(np.asarray(v))
        h = C.c_void_p()
        L.check(ctx.lib.cles_batch_create(ctx.handle, C.byref(d), len(self.instances), arr, C.byref(h)))
        self.handle = h
        self._keep = (d, arr)

    # ---- geometry and state ------------------------------------------------
    def xcoord(self, i: int) -> np.ndarray:
        ins = self.instances[i]
        L_ = ins.domain[1] - ins.domain[0]
        h = L_ / ins.n if self.periodic else L_ / (ins.n - 1)
        return ins.domain[0] + np.arange(ins.n) * h

    def initialize(self, i: int, ic: Callable | None = None, n_halo: int = 4):
        """problem.jl:146-163 initialize! for instance i, then upload."""
        ins = self.instances[i]
        x = self.xcoord(i)
        q = conserved_from_prim(self.eos, (ic or ins.ic)(x, 0 * x, 0 * x))
        Q = np.zeros((ins.n + 2 * n_halo, 1, 1, self.n_cons), order="F")
        for c in range(self.n_cons):
            Q[n_halo:n_halo + ins.n, 0, 0, c] = np.broadcast_to(q[c], x.shape)
        self.upload(i, Q, n_halo)
        return Q

    def initialize_all(self):
        for i in range(len(self.instances)):
            self.initialize(i)

    def upload(self, i: int, Q: np.ndarray, n_halo: int = 4):
        Q = np.asfortranarray(Q, dtype=np.float64)
        L.check(self.ctx.lib.cles_batch_upload(self.handle, i, Q.ctypes.data_as(L.DP),
                                               L.I64_3(*Q.shape[:3]), L.INT_3(n_halo, 0, 0)))

    def download(self, i: int, n_halo: int = 4) -> np.ndarray:
        ins = self.instances[i]
        Q = np.zeros((ins.n + 2 * n_halo, 1, 1, self.n_cons), order="F")
        L.check(self.ctx.lib.cles_batch_download(self.handle, i, Q.ctypes.data_as(L.DP),
                                                 L.I64_3(*Q.shape[:3]), L.INT_3(n_halo, 0, 0)))
        return Q

    def field(self, i: int, name: str) -> np.ndarray:
        out = np.empty(self.instances[i].n)
        L.check(self.ctx.lib.cles_batch_get_field(self.handle, i, name.encode(), out.ctypes.data_as(L.DP)))
        return out

    # ---- the per-step path ---------------------------------------------------
    def run(self, max_steps: int = 2 ** 62, record: int = 0, noalloc: bool = False):
        """run! on every instance (each to its own tfinal / nmax).  Returns the
        number of batch iterations, and with record > 0 also the (count x
        record) dt history of this call.  noalloc: through the entry point
        that allocates no device memory (cles_batch_run_noalloc)."""
        cnt = len(self.instances)
        dts = np.zeros((cnt, max(record, 1)))
        taken = C.c_int64()
        fn = self.ctx.lib.cles_batch_run_noalloc if noalloc else self.ctx.lib.cles_batch_run
        L.check(fn(self.handle, int(max_steps), int(record), dts.ctypes.data_as(L.DP) if record else None,
                   C.byref(taken)))
        return (taken.value, dts[:, :record]) if record else taken.value

    RUN_MODES = {"host_dt": 0, "device_dt": 1, "step_graph": 2}

    def set_run_mode(self, mode: str):
        """How run() issues its iterations: "host_dt" / "device_dt" launch
        every kernel from the host, "step_graph" replays a captured graph of 8
        iterations (bitwise the same results)."""
        if mode not in self.RUN_MODES:
            raise ValueError(f"run mode must be one of {sorted(self.RUN_MODES)}")
        L.check(self.ctx.lib.cles_batch_set_run_mode(self.handle, self.RUN_MODES[mode]))

    def state(self):
        """(t, steps, status) arrays over the instances."""
        cnt = len(self.instances)
        t = np.zeros(cnt)
        st = np.zeros(cnt, dtype=np.int64)
        status = np.zeros(cnt, dtype=np.int32)
        L.check(self.ctx.lib.cles_batch_state(self.handle, t.ctypes.data_as(L.DP),
                                              st.ctypes.data_as(C.POINTER(C.c_int64)),
                                              status.ctypes.data_as(C.POINTER(C.c_int))))
        return t, st, status

    def compute_rhs(self, primitives_current=False):
        L.check(self.ctx.lib.cles_batch_compute_rhs(self.handle, int(primitives_current)))

    def kinetic_energy(self) -> np.ndarray:
        ke = np.zeros(len(self.instances))
        L.check(self.ctx.lib.cles_batch_kinetic_energy(self.handle, ke.ctypes.data_as(L.DP)))
        return ke

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.cles_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
