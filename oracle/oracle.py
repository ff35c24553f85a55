"""ctypes front end of the CPU restatement (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, where it is the checker or the timed
CPU baseline -- never the thing measured or shipped.

Arrays follow the reference's padded column-major layout: a field is a numpy
array of shape ``(nx+2H, ny+2H, nz+2H)`` in Fortran order (H = 4 on active
dimensions, 0 on collapsed ones), and the conserved state is
``(..., n_cons)`` Fortran order (src/decomposition.jl:115-136).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

MAXCL, MAXRHS, MAXM, MAXSP = 8, 16, 8, 8
(BC_PERIODIC, BC_SLIPWALL, BC_DIRICHLET, BC_NSCBC_OUT, BC_NSCBC_IN, BC_AXIS, BC_ORIGIN, BC_POLE, BC_NOSLIP,
 BC_EXTRAP) = range(10)


class Scheme(C.Structure):
    _fields_ = [("q", C.c_int), ("lhs", C.c_double * 4), ("a0", C.c_double),
                ("M", C.c_int), ("coeffs", C.c_double * MAXM), ("symmetric", C.c_int),
                ("nclosure", C.c_int), ("cl_lhs", (C.c_double * 9) * MAXCL),
                ("cl_rhs_len", C.c_int * MAXCL), ("cl_rhs", (C.c_double * MAXRHS) * MAXCL)]


class Nscbc(C.Structure):
    _fields_ = [("pinf", C.c_double), ("sigma", C.c_double), ("Lref", C.c_double),
                ("beta_t", C.c_double), ("u", C.c_double * 3), ("T", C.c_double),
                ("Y", C.c_double * MAXSP), ("eta_u", C.c_double), ("eta_T", C.c_double),
                ("eta_t", C.c_double), ("eta_Y", C.c_double)]


class Nasa9Iv(C.Structure):
    _fields_ = [("Tmin", C.c_double), ("Tmax", C.c_double), ("a", C.c_double * 7), ("b1", C.c_double)]


class Nasa9Sp(C.Structure):
    _fields_ = [("n_intervals", C.c_int), ("iv", Nasa9Iv * 3)]


class Desc(C.Structure):
    _fields_ = [("n_global", C.c_int64 * 3), ("L_domain", C.c_double * 3),
                ("origin", C.c_double * 3), ("n_species", C.c_int),
                ("R", C.c_double * MAXSP), ("gamma", C.c_double * MAXSP),
                ("mu0", C.c_double), ("Pr", C.c_double), ("Sc", C.c_double),
                ("art_enabled", C.c_int), ("C_mu", C.c_double), ("C_beta", C.c_double),
                ("C_kappa", C.c_double), ("C_D", C.c_double), ("beta_sensor", C.c_int),
                ("deriv", Scheme), ("filt", Scheme), ("cfl", C.c_double),
                ("filter_interval", C.c_int), ("bc", (C.c_int * 2) * 3),
                ("bc_state", ((C.c_double * (MAXSP + 4)) * 2) * 3),
                ("nscbc", (Nscbc * 2) * 3),
                ("predict", C.c_double), ("max_growth", C.c_double),
                ("dt_min", C.c_double), ("dt_min_ratio", C.c_double),
                ("retries", C.c_int), ("cfl_backoff", C.c_double), ("savepoint_interval", C.c_int),
                ("landing_steps", C.c_int), ("eos_kind", C.c_int), ("T_guess", C.c_double),
                ("nasa9", Nasa9Sp * MAXSP), ("metric", C.c_int), ("Twall", (C.c_double * 2) * 3),
                ("switchable", (C.c_int * 2) * 3), ("bc_after", (C.c_int * 2) * 3),
                ("n_body_forces", C.c_int), ("body_force", (C.c_double * 3) * 4), ("p_inf", C.c_double)]


def _cpu_has_avx2() -> bool:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    fl = line.split()
                    return "avx2" in fl and "fma" in fl and "bmi2" in fl
    except OSError:
        pass
    return False


def _load(path=None):
    if path is None:
        name = "liboracle_v3.so" if _cpu_has_avx2() else "liboracle.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run `make -C oracle`")
    lib = C.CDLL(path)
    P, D, I64, I = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int)
    sig = {
        "or_lele_d1_6": (None, [C.POINTER(Scheme)]),
        "or_pade_d1_4": (None, [C.POINTER(Scheme)]),
        "or_lele_d1_10": (None, [C.POINTER(Scheme)]),
        "or_compact_filter": (None, [C.POINTER(Scheme), C.c_double]),
        "or_plan_create": (P, [C.POINTER(Scheme), C.c_int, I64, I, C.c_int, C.c_double, C.c_char_p, C.c_int]),
        "or_plan_create_fold": (P, [C.POINTER(Scheme), C.c_int, I64, I, C.c_int, C.c_double, C.c_int, C.c_int,
                                    C.c_char_p, C.c_int]),
        "or_fold_fill_halos": (None, [D, I64, I, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_plan_destroy": (None, [P]),
        "or_apply_along": (None, [P, D, D]),
        "or_wrap_halos": (None, [D, I64, I, I]),
        "or_solver_create": (P, [C.POINTER(Desc), C.c_char_p, C.c_int]),
        "or_solver_create_ex": (P, [C.POINTER(Desc), C.POINTER(D), C.c_char_p, C.c_int]),
        "or_solver_create_geom": (P, [C.POINTER(Desc), C.POINTER(D), C.POINTER(D), C.c_char_p, C.c_int]),
        "or_solver_destroy": (None, [P]),
        "or_solver_dims": (None, [P, I64, I, I]),
        "or_solver_Q": (D, [P]),
        "or_solver_dQ": (D, [P]),
        "or_solver_field": (D, [P, C.c_char_p]),
        "or_solver_time": (None, [P, D, I64]),
        "or_solver_set_time": (None, [P, C.c_double, C.c_int64]),
        "or_conserved_from_prim": (None, [P, D, D]),
        "or_apply_bcs": (None, [P]),
        "or_compute_rhs": (None, [P, C.c_int]),
        "or_max_rate": (None, [P, D, D]),
        "or_step": (None, [P, C.c_double, C.c_int]),
        "or_filter_state": (None, [P]),
        "or_run": (C.c_int, [P, C.c_double, C.c_int64, D, C.c_int64]),
        "or_run_ctl": (C.c_int, [P, C.c_double, C.c_int64, P, D, C.c_int64, C.POINTER(C.c_int)]),
        "or_solver_cfl": (C.c_double, [P]),
        "or_solver_set_cfl": (None, [P, C.c_double]),
        "or_diagnostics": (None, [P, D, D]),
        "or_volume_integral": (C.c_double, [P, D]),
        "or_solver_switch_bc": (C.c_int, [P, C.c_int, C.c_int]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


@contextlib.contextmanager
def library(path):
    """Within the block the oracle runs from another build of oracle.c (a test
    plants a mistake in a scratch copy to show that a check catches it).
    Objects made inside the block must be released inside it."""
    global _lib
    prev, _lib = _lib, _load(path)
    try:
        yield
    finally:
        _lib = prev


class OwnedArray(np.ndarray):
    """A numpy view of solver-owned memory that keeps the solver alive."""
    _owner = None


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


# --------------------------------------------------------------------------
# Schemes (restated in oracle.c from src/kernels.jl, src/kernels_banded.jl)
# --------------------------------------------------------------------------
def scheme(name: str, alphaf: float = 0.45) -> Scheme:
    s = Scheme()
    L = lib()
    if name == "lele_d1_6":
        L.or_lele_d1_6(C.byref(s))
    elif name == "pade_d1_4":
        L.or_pade_d1_4(C.byref(s))
    elif name == "lele_d1_10":
        L.or_lele_d1_10(C.byref(s))
    elif name == "compact_filter":
        L.or_compact_filter(C.byref(s), alphaf)
    else:
        raise ValueError(name)
    return s


def padded_shape(n_local, halo):
    return tuple(int(n_local[d]) + 2 * int(halo[d]) for d in range(3))


def halo_for(n_local, H=4):
    return tuple(H if n_local[d] > 1 else 0 for d in range(3))


def wrap_halos(f: np.ndarray, n_local, halo, periodic):
    nl = (C.c_int64 * 3)(*n_local)
    hl = (C.c_int * 3)(*halo)
    pr = (C.c_int * 3)(*[int(x) for x in periodic])
    lib().or_wrap_halos(_dptr(f), nl, hl, pr)
    return f


def fold_fill(f: np.ndarray, n_local, halo, dim: int, lo: bool, hi: bool, sigma: int) -> np.ndarray:
    """folds.jl:83-110 fold_fill! on a padded scalar field (one rank)."""
    lib().or_fold_fill_halos(_dptr(f), (C.c_int64 * 3)(*[int(x) for x in n_local]), (C.c_int * 3)(*halo),
                             int(dim), int(bool(lo)), int(bool(hi)), int(sigma))
    return f


class Plan:
    """apply_along! for one (scheme, dim) on padded arrays (one rank)."""

    def __init__(self, sch: Scheme, dim: int, n_local, periodic: bool, h: float, halo=None, lo_fold=0,
                 hi_fold=0):
        self.n_local = tuple(int(x) for x in n_local)
        self.halo = tuple(halo) if halo is not None else halo_for(self.n_local)
        self.dim, self.lo_fold, self.hi_fold = dim, int(lo_fold), int(hi_fold)
        msg = C.create_string_buffer(256)
        self._p = lib().or_plan_create_fold(C.byref(sch), dim, (C.c_int64 * 3)(*self.n_local),
                                            (C.c_int * 3)(*self.halo), int(periodic), float(h), self.lo_fold,
                                            self.hi_fold, msg, 256)
        if not self._p:
            raise ValueError(msg.value.decode())

    def __del__(self):
        if getattr(self, "_p", None):
            lib().or_plan_destroy(self._p)
            self._p = None

    def apply(self, f: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.zeros_like(f, order="F")
        lib().or_apply_along(self._p, _dptr(out), _dptr(f))
        return out


# --------------------------------------------------------------------------
# Solver
# --------------------------------------------------------------------------
class Solver:
    """The reference Solver + Workspace restated on padded arrays."""

    def __init__(self, desc: Desc, jac=None, x=None):
        """jac: per dimension None or dx/dxi at the nodes of a stretched dimension;
        x: per dimension None or the physical node coordinates of a stretched
        dimension that the metric reads (r; spherical theta)."""
        msg = C.create_string_buffer(256)
        DP = C.POINTER(C.c_double)

        def ptrs(arrs):
            if arrs is None or all(a is None for a in arrs):
                return None, None
            keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
            return keep, (DP * 3)(*[C.cast(None, DP) if a is None else a.ctypes.data_as(DP) for a in keep])
        self._jac, pj = ptrs(jac)
        self._x, px = ptrs(x)
        self._s = lib().or_solver_create_geom(C.byref(desc), px, pj, msg, 256)
        if not self._s:
            raise ValueError(msg.value.decode())
        self.desc = desc
        pad = (C.c_int64 * 3)()
        halo = (C.c_int * 3)()
        nc = C.c_int()
        lib().or_solver_dims(self._s, pad, halo, C.byref(nc))
        self.padded = tuple(pad)
        self.halo = tuple(halo)
        self.n_cons = nc.value
        self.n = tuple(int(desc.n_global[d]) for d in range(3))

    def __del__(self):
        if getattr(self, "_s", None):
            lib().or_solver_destroy(self._s)
            self._s = None

    def _view(self, ptr, ncomp=None):
        shape = self.padded if ncomp is None else self.padded + (ncomp,)
        cnt = int(np.prod(shape))
        buf = np.ctypeslib.as_array(ptr, shape=(cnt,))
        out = buf.reshape(shape, order="F").view(OwnedArray)
        out._owner = self
        return out

    @property
    def Q(self) -> np.ndarray:
        return self._view(lib().or_solver_Q(self._s), self.n_cons)

    @property
    def dQ(self) -> np.ndarray:
        return self._view(lib().or_solver_dQ(self._s), self.n_cons)

    def field(self, name: str) -> np.ndarray:
        p = lib().or_solver_field(self._s, name.encode())
        if not p:
            raise KeyError(name)
        return self._view(p)

    def switch_bc(self, dim: int, side: int):
        """boundary.jl:131-137 switch! on face (dim, side), 0-based."""
        if lib().or_solver_switch_bc(self._s, dim, side) != 0:
            raise ValueError("this face is not a SwitchableBC")

    def volume_integral(self, f: np.ndarray) -> float:
        """diagnostics.jl:60-80 volume_integral of a padded field."""
        f = np.asfortranarray(f, dtype=np.float64)
        assert f.shape == tuple(self.padded)
        return lib().or_volume_integral(self._s, f.ctypes.data_as(C.POINTER(C.c_double)))

    def interior(self, a: np.ndarray) -> np.ndarray:
        sl = tuple(slice(self.halo[d], self.halo[d] + self.n[d]) for d in range(3))
        return a[sl]

    @property
    def time(self):
        t = C.c_double()
        st = C.c_int64()
        lib().or_solver_time(self._s, C.byref(t), C.byref(st))
        return t.value, st.value

    def set_time(self, t, step):
        lib().or_solver_set_time(self._s, t, step)

    def conserved_from_prim(self, rho=math.nan, u=(0.0, 0.0, 0.0), p=math.nan, T=math.nan, Y=(1.0,)):
        prim = (C.c_double * (6 + MAXSP))(rho, *u, p, T, *Y)
        q = (C.c_double * (MAXSP + 4))()
        lib().or_conserved_from_prim(self._s, prim, q)
        return np.array(q[: self.n_cons])

    def coords(self, d):
        n = self.n[d]
        if n == 1:
            return np.array([self.desc.origin[d]])
        per = self.desc.bc[d][0] == BC_PERIODIC
        h = self.desc.L_domain[d] / (n if per else n - 1)
        return self.desc.origin[d] + np.arange(n) * h

    def initialize(self, ic):
        """initialize!(solver, Q, ic): ic(x, y, z) -> (n_cons,) conserved array
        broadcast over the interior (problem.jl:146-163)."""
        x, y, z = (self.coords(d) for d in range(3))
        X, Yc, Z = np.meshgrid(x, y, z, indexing="ij")
        q = ic(X, Yc, Z)
        Q = self.Q
        sl = tuple(slice(self.halo[d], self.halo[d] + self.n[d]) for d in range(3))
        for c in range(self.n_cons):
            Q[sl + (c,)] = q[c]

    def apply_bcs(self):
        lib().or_apply_bcs(self._s)

    def compute_rhs(self, primitives_current=False):
        lib().or_compute_rhs(self._s, int(primitives_current))

    def max_rate(self):
        r, m = C.c_double(), C.c_double()
        lib().or_max_rate(self._s, C.byref(r), C.byref(m))
        return r.value, m.value

    def step(self, dt, prepared=False):
        lib().or_step(self._s, dt, int(prepared))

    def filter_state(self):
        lib().or_filter_state(self._s)

    def run(self, tfinal, nmax, record=0, hooks=None):
        """run! (timestep.jl:277-400); hooks: a pointer to a hook table with
        the layout of or_run_hooks, or None.  Rollbacks in self.retries_used."""
        dts = (C.c_double * max(record, 1))()
        used = C.c_int(0)
        rc = lib().or_run_ctl(self._s, tfinal, nmax, hooks, dts, record, C.byref(used))
        self.retries_used = used.value
        if rc:
            raise RuntimeError(f"SolverFailure code {rc}")
        return np.array(dts[:record])

    @property
    def cfl(self):
        return lib().or_solver_cfl(self._s)

    @cfl.setter
    def cfl(self, value):
        lib().or_solver_set_cfl(self._s, float(value))

    def diagnostics(self):
        ke, en = C.c_double(), C.c_double()
        lib().or_diagnostics(self._s, C.byref(ke), C.byref(en))
        return ke.value, en.value


def make_desc(n_global, L_domain, origin=(0.0, 0.0, 0.0), species=((1.0, 1.4),),
              mu0=0.0, Pr=0.7, Sc=0.7, art=True, C_mu=0.002, C_beta=1.0, C_kappa=0.01,
              C_D=0.01, beta_sensor=0, deriv="lele_d1_6", alphaf=0.45, cfl=0.5,
              filter_interval=1, bcs=None, bc_state=None, nscbc=None,
              predict=0.0, max_growth=0.0, dt_min=0.0, dt_min_ratio=1e-8, retries=0, cfl_backoff=0.5,
              savepoint_interval=25, landing_steps=2, nasa9=None, metric=0, Twall=None, switch=None,
              body_forces=(), stiffened=None) -> Desc:
    """nasa9: (T_guess, [(R, [(Tmin, Tmax, (a1..a7), b1), ...]), ...]) selects
    the Nasa9Mixture (species then ignores gamma)."""
    d = Desc()
    for k in range(3):
        d.n_global[k] = int(n_global[k])
        d.L_domain[k] = float(L_domain[k])
        d.origin[k] = float(origin[k])
    d.n_species = len(species)
    for k, (R, g) in enumerate(species):
        d.R[k] = R
        d.gamma[k] = g
    d.mu0, d.Pr, d.Sc = mu0, Pr, Sc
    d.art_enabled = int(art)
    d.C_mu, d.C_beta, d.C_kappa, d.C_D = C_mu, C_beta, C_kappa, C_D
    d.beta_sensor = beta_sensor
    d.deriv = scheme(deriv)
    d.filt = scheme("compact_filter", alphaf)
    d.cfl = cfl
    d.filter_interval = filter_interval
    bcs = bcs or ((BC_PERIODIC, BC_PERIODIC),) * 3
    for k in range(3):
        for s in range(2):
            d.bc[k][s] = bcs[k][s]
    if bc_state:
        for (k, s), q in bc_state.items():
            for c, v in enumerate(q):
                d.bc_state[k][s][c] = v
    if nscbc:
        for (k, s), par in nscbc.items():
            f = d.nscbc[k][s]
            for key, val in par.items():
                if key in ("u", "Y"):
                    arr = getattr(f, key)
                    for i, x in enumerate(val):
                        arr[i] = x
                else:
                    setattr(f, key, val)
    d.predict, d.max_growth, d.dt_min, d.dt_min_ratio = predict, max_growth, dt_min, dt_min_ratio
    d.retries, d.cfl_backoff, d.savepoint_interval, d.landing_steps = retries, cfl_backoff, savepoint_interval, \
        landing_steps
    d.metric = {"cartesian": 0, "cylindrical": 1, "spherical": 2}.get(metric, metric)
    for (k, sd), T in (Twall or {}).items():           # NoSlipWallBC faces: {(dim, side): Twall}
        d.Twall[k][sd] = T
    if stiffened is not None:                          # StiffenedGas (gamma, p_inf, cv), one species
        g, pinf, cv = stiffened
        d.eos_kind, d.n_species, d.R[0], d.gamma[0], d.p_inf = 2, 1, (g - 1) * cv, g, pinf
    d.n_body_forces = len(body_forces)                 # ConstantBodyForce accelerations, in order
    for i, g in enumerate(body_forces):
        for k in range(3):
            d.body_force[i][k] = g[k]
    for (k, sd), after in (switch or {}).items():      # SwitchableBC faces: {(dim, side): bc_after}
        d.switchable[k][sd] = 1
        d.bc_after[k][sd] = after
    if nasa9 is not None:
        d.eos_kind = 1
        d.T_guess, table = nasa9
        d.n_species = len(table)
        for k, (R, ivs) in enumerate(table):
            d.R[k] = R
            d.gamma[k] = 0.0
            d.nasa9[k].n_intervals = len(ivs)
            for i, (Tmin, Tmax, a, b1) in enumerate(ivs):
                iv = d.nasa9[k].iv[i]
                iv.Tmin, iv.Tmax, iv.b1 = Tmin, Tmax, b1
                for j in range(7):
                    iv.a[j] = a[j]
    return d
