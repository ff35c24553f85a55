"""ctypes binding of libcompactles_hip.so (the C ABI in include/compactles_hip.h).

The library is loaded from the package directory's parent (in-tree build), or
from the path in $COMPACTLES_HIP_LIB.
There is no fallback: if the shared library is missing or no gfx950 device is
present, calls raise instead of computing anything on the host.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("COMPACTLES_HIP_LIB") or os.path.join(os.path.dirname(_PKG), "libcompactles_hip.so")

MAXCL, MAXRHS, MAXM, MAXSP = 8, 16, 8, 8

OK, ERR_INVALID, ERR_TOO_SMALL, ERR_HIP, ERR_NONFINITE, ERR_NEGATIVE_DENSITY, ERR_PLANCK, \
    ERR_DT_MIN, ERR_DT_COLLAPSE, ERR_NOMEM, ERR_UNSUPPORTED = range(11)
F64, F32 = 0, 1
(BC_PERIODIC, BC_SLIPWALL, BC_DIRICHLET, BC_NSCBC_OUT, BC_NSCBC_IN, BC_AXIS, BC_ORIGIN, BC_POLE, BC_NOSLIP,
 BC_EXTRAP) = range(10)


class SchemeDesc(C.Structure):
    _fields_ = [("q", C.c_int), ("lhs", C.c_double * 4), ("a0", C.c_double), ("M", C.c_int),
                ("coeffs", C.c_double * MAXM), ("symmetric", C.c_int), ("nclosure", C.c_int),
                ("closure_lhs", (C.c_double * 9) * MAXCL), ("closure_rhs_len", C.c_int * MAXCL),
                ("closure_rhs", (C.c_double * MAXRHS) * MAXCL)]


class NscbcDesc(C.Structure):
    _fields_ = [("pinf", C.c_double), ("sigma", C.c_double), ("Lref", C.c_double), ("beta_t", C.c_double),
                ("u", C.c_double * 3), ("T", C.c_double), ("Y", C.c_double * MAXSP),
                ("eta_u", C.c_double), ("eta_T", C.c_double), ("eta_t", C.c_double), ("eta_Y", C.c_double)]


class Nasa9IntervalDesc(C.Structure):
    _fields_ = [("Tmin", C.c_double), ("Tmax", C.c_double), ("a", C.c_double * 7), ("b1", C.c_double)]


class Nasa9SpeciesDesc(C.Structure):
    _fields_ = [("n_intervals", C.c_int), ("iv", Nasa9IntervalDesc * 3)]


EOS_IDEAL, EOS_NASA9, EOS_STIFFENED = 0, 1, 2
MAX_FORCES = 4


class SolverDesc(C.Structure):
    _fields_ = [("n_global", C.c_int64 * 3), ("L_domain", C.c_double * 3), ("origin", C.c_double * 3),
                ("n_species", C.c_int), ("R", C.c_double * MAXSP), ("gamma", C.c_double * MAXSP),
                ("mu0", C.c_double), ("Pr", C.c_double), ("Sc", C.c_double),
                ("art_enabled", C.c_int), ("C_mu", C.c_double), ("C_beta", C.c_double),
                ("C_kappa", C.c_double), ("C_D", C.c_double), ("beta_sensor", C.c_int),
                ("deriv", SchemeDesc), ("filt", SchemeDesc), ("cfl", C.c_double),
                ("filter_interval", C.c_int), ("bc", (C.c_int * 2) * 3),
                ("bc_state", ((C.c_double * (MAXSP + 4)) * 2) * 3), ("nscbc", (NscbcDesc * 2) * 3),
                ("predict", C.c_double), ("max_growth", C.c_double), ("dt_min", C.c_double),
                ("dt_min_ratio", C.c_double), ("precision", C.c_int), ("retries", C.c_int),
                ("cfl_backoff", C.c_double), ("savepoint_interval", C.c_int), ("landing_steps", C.c_int),
                ("eos_kind", C.c_int), ("T_guess", C.c_double), ("nasa9", Nasa9SpeciesDesc * MAXSP),
                ("metric", C.c_int), ("Twall", (C.c_double * 2) * 3), ("switchable", (C.c_int * 2) * 3),
                ("bc_after", (C.c_int * 2) * 3), ("n_body_forces", C.c_int), ("body_force", (C.c_double * 3) * 4),
                ("p_inf", C.c_double)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
class BatchInstance(C.Structure):
    _fields_ = [("n", C.c_int64), ("L", C.c_double), ("origin", C.c_double), ("tfinal", C.c_double),
                ("nmax", C.c_int64), ("bc_state", (C.c_double * (MAXSP + 4)) * 2)]


I64_3 = C.c_int64 * 3
INT_3 = C.c_int * 3
DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); the public ABI (every symbol of compactles_hip.h)
SIGNATURES = {
    "cles_ctx_create": (C.c_int, [C.c_int, PP]),
    "cles_ctx_destroy": (C.c_int, [P]),
    "cles_last_error": (C.c_char_p, []),
    "cles_abi_version": (C.c_int, []),
    "cles_ctx_stream": (P, [P]),
    "cles_malloc": (C.c_int, [P, C.c_size_t, PP]),
    "cles_free": (C.c_int, [P, P]),
    "cles_memcpy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "cles_memcpy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "cles_memset": (C.c_int, [P, P, C.c_int, C.c_size_t]),
    "cles_synchronize": (C.c_int, [P]),
    "cles_memcpy_d2d": (C.c_int, [P, P, P, C.c_size_t]),
    "cles_plan_create": (C.c_int, [P, C.POINTER(SchemeDesc), C.c_int, C.POINTER(C.c_int64), C.c_int,
                                   C.c_double, C.c_int, C.c_int, C.c_int, PP]),
    "cles_plan_destroy": (C.c_int, [P]),
    "cles_apply_along": (C.c_int, [P, P, P, C.POINTER(C.c_int64), C.POINTER(C.c_int), P]),
    "cles_apply_along_timed": (C.c_int, [P, P, P, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.c_int, DP]),
    "cles_solver_create": (C.c_int, [P, C.POINTER(SolverDesc), PP]),
    "cles_solver_destroy": (C.c_int, [P]),
    "cles_solver_upload": (C.c_int, [P, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_solver_download": (C.c_int, [P, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_solver_get_field": (C.c_int, [P, C.c_char_p, DP]),
    "cles_solver_get_time": (C.c_int, [P, DP, C.POINTER(C.c_int64)]),
    "cles_solver_set_time": (C.c_int, [P, C.c_double, C.c_int64]),
    "cles_apply_bcs": (C.c_int, [P]),
    "cles_compute_rhs": (C.c_int, [P, C.c_int]),
    "cles_max_rate": (C.c_int, [P, DP, DP]),
    "cles_step": (C.c_int, [P, C.c_double, C.c_int]),
    "cles_filter_state": (C.c_int, [P]),
    "cles_run": (C.c_int, [P, C.c_double, C.c_int64, DP, C.c_int64]),
    "cles_run_ctl": (C.c_int, [P, C.c_double, C.c_int64, P, DP, C.c_int64, C.POINTER(C.c_int)]),
    "cles_solver_get_cfl": (C.c_int, [P, DP]),
    "cles_solver_set_cfl": (C.c_int, [P, C.c_double]),
    "cles_diagnostics": (C.c_int, [P, DP, DP]),
    "cles_solver_shape": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_solver_profile": (C.c_int, [P, C.c_int]),
    "cles_solver_profile_read": (C.c_int, [P, C.c_int, C.c_char_p, C.c_int, DP, C.POINTER(C.c_int64), DP]),
    "cles_solver_set_run_mode": (C.c_int, [P, C.c_int]),
    "cles_solver_save_checkpoint": (C.c_int, [P, C.c_char_p, C.c_int]),
    "cles_solver_load_checkpoint": (C.c_int, [P, C.c_char_p, C.c_int]),
    "cles_solver_footprint": (C.c_int, [C.POINTER(SolverDesc), C.POINTER(C.c_size_t)]),
    "cles_solver_create_ex": (C.c_int, [P, C.POINTER(SolverDesc), C.POINTER(DP), PP]),
    "cles_solver_create_geom": (C.c_int, [P, C.POINTER(SolverDesc), C.POINTER(DP), C.POINTER(DP), PP]),
    "cles_solver_switch_bc": (C.c_int, [P, C.c_int, C.c_int]),
    "cles_solver_bc_switched": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "cles_solver_memory": (C.c_int, [P, C.POINTER(C.c_size_t)]),
    "cles_stack_create": (C.c_int, [P, C.c_void_p, C.c_int, PP]),
    "cles_stack_destroy": (C.c_int, [P]),
    "cles_stack_upload": (C.c_int, [P, C.c_int, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_stack_download": (C.c_int, [P, C.c_int, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_stack_run": (C.c_int, [P, DP, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]),
    "cles_stack_state": (C.c_int, [P, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, PP]),
    "cles_server_set_run_mode": (C.c_int, [P, C.c_int]),
    "cles_server_set_stack_limit": (C.c_int, [P, C.c_int, C.c_double]),
    "cles_server_group": (C.c_int, [P, C.c_int64, C.POINTER(C.c_int)]),
    "cles_server_destroy": (C.c_int, [P]),
    "cles_server_submit": (C.c_int, [P, C.POINTER(SolverDesc), DP, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                     C.c_double, C.c_int64, C.POINTER(C.c_int64)]),
    "cles_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(C.c_int)]),
    "cles_server_result": (C.c_int, [P, C.c_int64, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), DP, C.POINTER(C.c_int64)]),
    "cles_server_timing": (C.c_int, [P, C.c_int64, DP, DP, DP, C.POINTER(C.c_size_t)]),
    "cles_server_release": (C.c_int, [P, C.c_int64]),
    "cles_server_stats": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cles_batch_create": (C.c_int, [P, C.POINTER(SolverDesc), C.c_int64, C.POINTER(BatchInstance), PP]),
    "cles_batch_destroy": (C.c_int, [P]),
    "cles_batch_shape": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_batch_upload": (C.c_int, [P, C.c_int64, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_batch_download": (C.c_int, [P, C.c_int64, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_batch_run": (C.c_int, [P, C.c_int64, C.c_int64, DP, C.POINTER(C.c_int64)]),
    "cles_batch_run_noalloc": (C.c_int, [P, C.c_int64, C.c_int64, DP, C.POINTER(C.c_int64)]),
    "cles_batch_set_run_mode": (C.c_int, [P, C.c_int]),
    "cles_batch_state": (C.c_int, [P, DP, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "cles_batch_compute_rhs": (C.c_int, [P, C.c_int]),
    "cles_batch_get_field": (C.c_int, [P, C.c_int64, C.c_char_p, DP]),
    "cles_batch_kinetic_energy": (C.c_int, [P, DP]),
    "cles_tune_solver": (C.c_int, [P, C.POINTER(SolverDesc), C.c_int, C.POINTER(C.c_int), DP]),
    "cles_tune_lookup": (C.c_int, [C.POINTER(SolverDesc), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "cles_tune_set": (C.c_int, [C.POINTER(SolverDesc), C.POINTER(C.c_int)]),
    "cles_tune_clear": (C.c_int, []),
    "cles_tune_save": (C.c_int, [C.c_char_p]),
    "cles_tune_load": (C.c_int, [C.c_char_p]),
}


class ClesError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SolverFailure(ClesError):
    """run! failure with a reason (src/stepcontrol.jl:138-152)."""
    REASONS = {ERR_NONFINITE: "nonfinite", ERR_NEGATIVE_DENSITY: "negative_density", ERR_PLANCK: "planck",
               ERR_DT_MIN: "dt_min", ERR_DT_COLLAPSE: "dt_collapse"}

    @property
    def reason(self):
        return self.REASONS.get(self.code, "unknown")


_lib = None


def load(path: str | None = None):
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} not found: build it with `make -C compactles.jl_amd` "
                           "(or __graft_entry__.build())")
    lib = C.CDLL(p)
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    lib.missing_symbols = missing
    _lib = lib
    return lib


def check(rc: int):
    if rc != OK:
        msg = load().cles_last_error().decode(errors="replace")
        if rc in SolverFailure.REASONS:
            raise SolverFailure(rc, msg)
        raise ClesError(rc, msg)
