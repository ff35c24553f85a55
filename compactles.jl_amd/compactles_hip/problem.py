"""Host-side mirror of the CompactLES frontend (src/problem.jl, src/physics.jl,
src/artificial.jl, src/stepcontrol.jl, src/boundary.jl, src/nscbc.jl,
src/timestep.jl): the same Problem / Numerics / setup / run! vocabulary, with
the solver state resident on the GPU behind the C ABI.

Differences from the Julia frontend, all forced by running the physics on the
device: initial-condition functions are evaluated vectorised (they receive
numpy coordinate arrays), Dirichlet faces hold a constant state (their
function is evaluated once, at t = 0, and must be uniform over the face), and
NSCBC inflow targets are constants.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Sequence, Tuple

import numpy as np

from . import _lib as L
from .device import Context
from .schemes import AbstractCompactScheme, compact_filter, lele_d1_6


# --------------------------------------------------------------------------
# EOS and transport (physics.jl)
# --------------------------------------------------------------------------
@dataclass
class IdealSpecies:
    """physics.jl:86-90."""
    name: str
    R: float
    gamma: float


@dataclass
class IdealMixture:
    """physics.jl:104-115: mass-fraction averaged ideal mixture."""
    sp: Sequence[IdealSpecies]

    @property
    def Rk(self):
        return [s.R for s in self.sp]

    @property
    def cvk(self):
        return [s.R / (s.gamma - 1) for s in self.sp]

    def nspecies(self):
        return len(self.sp)


def single_species(gamma: float = 1.4, R: float = 1.0, name: str = "gas") -> IdealMixture:
    """physics.jl:125-126."""
    return IdealMixture([IdealSpecies(name, float(R), float(gamma))])


@dataclass
class Transport:
    """physics.jl:589-593."""
    mu0: float = 0.0
    Pr: float = 0.7
    Sc: float = 0.7


@dataclass
class ArtParams:
    """artificial.jl:50-57."""
    enabled: bool = True
    C_mu: float = 0.002
    C_beta: float = 1.0
    C_kappa: float = 0.01
    C_D: float = 0.01
    beta_sensor: str = "strain"


class Stretch:
    """metric.jl:45-58: a monotone map x(xi) of the computational coordinate
    xi in [0, 1] onto a non-periodic dimension, with dxdxi its derivative."""

    def __init__(self, x, dxdxi):
        self.x, self.dxdxi = x, dxdxi


def sine_cluster(lo, hi, xic, a):
    """metric.jl:60-75: x(xi) = lo + L(xi - (a/2pi)[sin 2pi(xi - xic) + sin 2pi xic]),
    dx/dxi = L(1 - a cos 2pi(xi - xic)); 0 <= a < 1."""
    if not 0 <= a < 1:
        raise ValueError("sine_cluster: need 0 ≤ a < 1")
    L = hi - lo
    return Stretch(lambda xi: lo + L * (xi - (a / (2 * math.pi)) * (np.sin(2 * math.pi * (xi - xic)) +
                                                                 math.sin(2 * math.pi * xic))),
                   lambda xi: L * (1 - a * np.cos(2 * math.pi * (xi - xic))))


@dataclass
class StepControl:
    """stepcontrol.jl:116-136: prediction, growth limit, soft landing, floors,
    and recovery (retries rollbacks to the last savepoint, each multiplying the
    CFL by cfl_backoff; a savepoint every savepoint_interval steps)."""
    predict: float = 0.0
    max_growth: float = 0.0
    landing_steps: int = 2
    dt_min: float = 0.0
    dt_min_ratio: float = 1e-8
    retries: int = 0
    cfl_backoff: float = 0.5
    savepoint_interval: int = 25

    def __post_init__(self):
        if self.landing_steps < 1:
            raise ValueError("StepControl: landing_steps must be >= 1 (1 is a hard clip onto the scheduled time)")


# --------------------------------------------------------------------------
# Primitive states (problem.jl:41-90)
# --------------------------------------------------------------------------
class Prim:
    """problem.jl:41-57: exactly two of p, rho, T_ion; the EOS derives the third.
    Entries may be numpy arrays (vectorised initial conditions)."""

    def __init__(self, u=(0.0, 0.0, 0.0), p=math.nan, T_ion=math.nan, rho=math.nan, Y=(1.0,)):
        given = [not _isnan(x) for x in (p, T_ion, rho)]
        if sum(given) != 2:
            raise ValueError("Prim: specify exactly two of p, rho, and T_ion; the EOS derives the third")
        Ysum = sum(np.asarray(y, dtype=float) for y in Y)
        if np.max(np.abs(np.asarray(Ysum) - 1.0)) >= 1e-10:
            raise ValueError("Prim: mass fractions must sum to 1")
        self.u = tuple(u)
        self.p, self.T_ion, self.rho = p, T_ion, rho
        self.Y = tuple(Y)


def _isnan(x):
    a = np.asarray(x, dtype=float)
    return a.size == 1 and bool(np.isnan(a).all())


def conserved_from_prim(eos, pr: Prim):
    """problem.jl:76-90: (rho Y_k..., rho u, rho v, rho w, rho E) for an ideal
    or a NASA-9 mixture."""
    from .nasa9 import Nasa9Mixture, nasa9_conserved
    if isinstance(eos, Nasa9Mixture):
        return nasa9_conserved(eos, pr)
    if isinstance(eos, StiffenedGas):
        # problem.jl:92-102: p + p_inf = rho R T_ion; rho e = rho cv T_ion + p_inf
        if len(pr.Y) != 1:
            raise ValueError(f"StiffenedGas is single-component; Prim carries {len(pr.Y)} mass fractions")
        R = eos.gas_constant()
        rho = (pr.p + eos.p_inf) / (R * pr.T_ion) if _isnan(pr.rho) else pr.rho
        T = (pr.p + eos.p_inf) / (rho * R) if _isnan(pr.T_ion) else pr.T_ion
        u1, u2, u3 = pr.u
        ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3)
        return [rho, rho * u1, rho * u2, rho * u3, rho * eos.cv * T + eos.p_inf + rho * ke]
    N = eos.nspecies()
    if len(pr.Y) != N:
        raise ValueError(f"Prim carries {len(pr.Y)} mass fractions; EOS has {N} species")
    Rm = 0.0
    cvm = 0.0
    for k in range(N):
        Rm = Rm + pr.Y[k] * eos.Rk[k]
        cvm = cvm + pr.Y[k] * eos.cvk[k]
    rho = pr.p / (Rm * pr.T_ion) if _isnan(pr.rho) else pr.rho
    T = pr.p / (rho * Rm) if _isnan(pr.T_ion) else pr.T_ion
    u1, u2, u3 = pr.u
    ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3)
    out = [rho * pr.Y[k] for k in range(N)]
    out += [rho * u1, rho * u2, rho * u3, rho * (cvm * T + ke)]
    return out


def tanh_blend(x, x0, delta):
    """problem.jl:172."""
    return 0.5 * (1 + np.tanh((x - x0) / delta))


# --------------------------------------------------------------------------
# Boundary conditions (boundary.jl, problem.jl:192-207, nscbc.jl)
# --------------------------------------------------------------------------
class BoundaryCondition:
    code = None


class PeriodicBC(BoundaryCondition):
    code = L.BC_PERIODIC


class SlipWallBC(BoundaryCondition):
    code = L.BC_SLIPWALL


class DirichletBC(BoundaryCondition):
    """Hard prescription of the full state on a face from fun(x1,x2,x3,t) -> Prim,
    held constant (evaluated at t = 0)."""
    code = L.BC_DIRICHLET

    def __init__(self, fun: Callable):
        self.fun = fun


class NoSlipWallBC(BoundaryCondition):
    """boundary.jl:34-43: impermeable viscous wall with zero velocity; a
    finite Twall makes it isothermal (the EOS sets the wall energy)."""
    code = L.BC_NOSLIP

    def __init__(self, Twall: float = math.nan):
        self.Twall = float(Twall)


class ExtrapolationBC(BoundaryCondition):
    """boundary.jl:45-52: copy the adjacent interior plane onto the boundary plane."""
    code = L.BC_EXTRAP


class AxisBC(BoundaryCondition):
    """boundary.jl:55-66: regularized cylindrical axis at the low end of r
    (half-offset grid, parity mirror conditions; no state enforcement)."""
    code = L.BC_AXIS


class OriginBC(BoundaryCondition):
    """boundary.jl:68-80: regularized spherical origin at the low end of r."""
    code = L.BC_ORIGIN


class PoleBC(BoundaryCondition):
    """boundary.jl:82-92: regularized spherical poles, at both ends of theta."""
    code = L.BC_POLE


_FOLDS = (AxisBC, OriginBC, PoleBC)


class SwitchableBC(BoundaryCondition):
    """boundary.jl:94-145: runs `before` until switch_(bc), then `after`.
    Both arms agree on periodicity; fold conditions cannot be wrapped."""

    def __init__(self, before: BoundaryCondition, after: BoundaryCondition):
        if isinstance(before, PeriodicBC) != isinstance(after, PeriodicBC):
            raise ValueError("SwitchableBC: both conditions must agree on periodicity")
        if isinstance(before, _FOLDS) or isinstance(after, _FOLDS):
            raise ValueError("SwitchableBC cannot wrap a fold condition (AxisBC, OriginBC, PoleBC); "
                             "setup detects those by type")
        if isinstance(before, SwitchableBC) or isinstance(after, SwitchableBC):
            raise ValueError("SwitchableBC cannot wrap a SwitchableBC")
        self.before, self.after = before, after
        self.code = before.code
        self._switched = False
        self._faces = []          # (solver, dim, side) that run this condition

    def switched(self) -> bool:
        return self._switched


def _is_periodic(bc) -> bool:
    """boundary.jl:127 isperiodic (a SwitchableBC answers for both arms)."""
    return isinstance(bc.before if isinstance(bc, SwitchableBC) else bc, PeriodicBC)


def switch_(bc: SwitchableBC) -> SwitchableBC:
    """boundary.jl:131-137 switch!: later steps of every solver running this
    condition take `after`; repeated calls have no further effect."""
    bc._switched = True
    for solver, d, side in bc._faces:
        L.check(solver.ctx.lib.cles_solver_switch_bc(solver.handle, d, side))
    return bc


def switched(bc: SwitchableBC) -> bool:
    return bc.switched()


@dataclass
class NSCBCOutflowBC(BoundaryCondition):
    """nscbc.jl:45-54."""
    pinf: float
    sigma: float = 0.25
    Lref: float = 0.0
    beta_t: float = -1.0
    code = L.BC_NSCBC_OUT


@dataclass
class NSCBCInflowBC(BoundaryCondition):
    """nscbc.jl:204-216 (constant targets)."""
    u: Tuple[float, float, float]
    T_ion: float
    Y: Sequence[float] = (1.0,)
    eta_u: float = 0.28
    eta_T: float = 0.28
    eta_t: float = 0.28
    eta_Y: float = 0.28
    Lref: float = 0.0
    code = L.BC_NSCBC_IN


class StiffenedGas:
    """physics.jl:180-205: single-component p = (gamma - 1) rho e - gamma p_inf
    with p + p_inf = rho R T_ion, R = (gamma - 1) cv."""

    def __init__(self, gamma=4.4, p_inf=6.0e8, cv=1816.0, name="liquid"):
        self.gamma, self.p_inf, self.cv, self.name = float(gamma), float(p_inf), float(cv), str(name)

    def nspecies(self):
        return 1

    def gas_constant(self):
        return (self.gamma - 1) * self.cv

    @property
    def Rk(self):
        return [self.gas_constant()]

    def species_names(self):
        return [self.name]


# --------------------------------------------------------------------------
# Sources (sources.jl): explicit terms added to the assembled RHS
# --------------------------------------------------------------------------
class ConstantBodyForce:
    """sources.jl:4-30: uniform acceleration; adds rho g to momentum and
    (rho u) . g to total energy."""

    def __init__(self, acceleration=(0.0, -9.81, 0.0)):
        if len(acceleration) != 3:
            raise ValueError("ConstantBodyForce: three acceleration components")
        self.acceleration = tuple(float(a) for a in acceleration)


# --------------------------------------------------------------------------
# Metrics (metric.jl:20-45): coordinates (x, y, z), (r, theta, z), (r, theta,
# phi) with theta the polar angle; velocities are physical components.
# --------------------------------------------------------------------------
class CartesianMetric:
    code = 0


class CylindricalMetric:
    code = 1


class SphericalMetric:
    code = 2


_METRICS = {"cartesian": 0, "cylindrical": 1, "spherical": 2}


def metric_code(metric) -> int:
    if isinstance(metric, str):
        if metric not in _METRICS:
            raise ValueError(f"unknown metric {metric!r}")
        return _METRICS[metric]
    if isinstance(metric, type):
        metric = metric()
    if not isinstance(metric, (CartesianMetric, CylindricalMetric, SphericalMetric)):
        raise ValueError(f"unknown metric {metric!r}")
    return metric.code


# --------------------------------------------------------------------------
# Problem / Numerics (problem.jl:243-306)
# --------------------------------------------------------------------------
@dataclass
class Problem:
    domain: Sequence[Tuple[float, float]]
    bcs: Sequence[Tuple[BoundaryCondition, BoundaryCondition]]
    ic: Callable
    name: str = "problem"
    eos: IdealMixture = field(default_factory=single_species)
    transport: Transport = field(default_factory=Transport)
    metric: object = "cartesian"   # a *Metric() or its name
    sources: tuple = ()


@dataclass
class Numerics:
    n_global: Tuple[int, int, int]
    deriv: AbstractCompactScheme = field(default_factory=lele_d1_6)
    filt: AbstractCompactScheme = field(default_factory=lambda: compact_filter(0.45))
    art: ArtParams = field(default_factory=ArtParams)
    cfl: float = 0.5
    control: StepControl = field(default_factory=StepControl)
    filter_interval: int = 1
    dims: object = None
    n_halo: int = 4
    stretch: object = None
    precision: str = "fp64"     # "fp64" (the drop-in) or "fp32" (mixed form)


_BETA = {"strain": 0, "gated_strain": 1, "dilatation": 2}


def fill_common_desc(d: L.SolverDesc, prob: Problem, num: Numerics):
    """The fields of a solver descriptor that do not depend on the grid:
    EOS, transport, artificial properties, schemes, cfl, filter cadence,
    step control, precision and metric."""
    from .nasa9 import Nasa9Mixture
    eos = prob.eos
    d.n_species = eos.nspecies()
    if isinstance(eos, Nasa9Mixture):
        eos.to_desc(d)
    elif isinstance(eos, StiffenedGas):
        d.eos_kind = L.EOS_STIFFENED
        d.R[0], d.gamma[0], d.p_inf = eos.gas_constant(), eos.gamma, eos.p_inf
    else:
        d.eos_kind = L.EOS_IDEAL
        for k, sp in enumerate(eos.sp):
            d.R[k] = sp.R
            d.gamma[k] = sp.gamma
    tr = prob.transport
    d.mu0, d.Pr, d.Sc = tr.mu0, tr.Pr, tr.Sc
    art = num.art
    d.art_enabled = int(art.enabled)
    d.C_mu, d.C_beta, d.C_kappa, d.C_D = art.C_mu, art.C_beta, art.C_kappa, art.C_D
    if art.beta_sensor not in _BETA:
        raise ValueError(f"art.beta_sensor must be :strain, :gated_strain or :dilatation, got :{art.beta_sensor}")
    d.beta_sensor = _BETA[art.beta_sensor]
    d.deriv = num.deriv.to_desc()
    d.filt = num.filt.to_desc()
    d.cfl = num.cfl
    d.filter_interval = num.filter_interval
    ctl = num.control
    d.predict, d.max_growth, d.dt_min, d.dt_min_ratio = ctl.predict, ctl.max_growth, ctl.dt_min, ctl.dt_min_ratio
    d.retries, d.cfl_backoff = int(ctl.retries), float(ctl.cfl_backoff)
    d.savepoint_interval, d.landing_steps = int(ctl.savepoint_interval), int(ctl.landing_steps)
    if num.precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    d.precision = L.F64 if num.precision == "fp64" else L.F32
    d.metric = metric_code(prob.metric)
    srcs = tuple(prob.sources)
    for src in srcs:
        if not isinstance(src, ConstantBodyForce):
            raise ValueError(f"source {type(src).__name__} is not available on the GPU backend (ConstantBodyForce is)")
    if len(srcs) > L.MAX_FORCES:
        raise ValueError(f"at most {L.MAX_FORCES} sources")
    d.n_body_forces = len(srcs)
    for i, src in enumerate(srcs):
        for k in range(3):
            d.body_force[i][k] = src.acceleration[k]


class SolverSpec:
    """The grid and the C-ABI descriptor of setup(prob, num), with no device
    state (what a Solver is created from, and what the server admits)."""

    def __init__(self, prob: Problem, num: Numerics):
        self.metric = metric_code(prob.metric)
        stretch = tuple(num.stretch) if num.stretch is not None else (None, None, None)
        if len(stretch) != 3:
            raise ValueError("Numerics.stretch: one entry per dimension")
        self.prob, self.num = prob, num
        self.eos = prob.eos
        self.n = tuple(int(x) for x in num.n_global)
        self.origin = tuple(float(prob.domain[d][0]) for d in range(3))
        self.L = tuple(float(prob.domain[d][1] - prob.domain[d][0]) for d in range(3))
        if not all(x > 0 for x in self.L):
            raise ValueError("domain extents must be positive")
        self.n_species = self.eos.nspecies()
        self.n_cons = self.n_species + 4
        self.periodic = tuple(_is_periodic(prob.bcs[d][0]) or self.n[d] == 1 for d in range(3))
        self.stretch = stretch
        for d, st in enumerate(stretch):
            if st is None:
                continue
            lo, hi = prob.domain[d]
            # problem.jl:324-332
            if not (abs(st.x(0.0) - lo) <= 1e-10 * self.L[d] and abs(st.x(1.0) - hi) <= 1e-10 * self.L[d]):
                raise ValueError(f"stretch mapping for dim {d + 1} does not span the domain: "
                                 f"x(0) = {st.x(0.0)}, x(1) = {st.x(1.0)}, domain = {prob.domain[d]}")
            if self.n[d] == 1 or self.periodic[d]:
                raise ValueError(f"dimension {d + 1}: stretched dimensions must be non-periodic")
        # rhs.jl:191-205: stretched dimensions are uniform in xi on [0, 1];
        # folded dimensions (axis, origin: low end of r; poles: both ends of
        # theta) sit on half-offset grids, h = L / (n - 1/2) or L / n, first node h / 2
        fold_lo = (isinstance(prob.bcs[0][0], (AxisBC, OriginBC)), isinstance(prob.bcs[1][0], PoleBC), False)
        fold_hi = (False, isinstance(prob.bcs[1][1], PoleBC), False)
        self.h = tuple(1.0 if self.n[d] == 1 else
                       (1.0 / (self.n[d] - 1) if stretch[d] is not None else
                        ((self.L[d] / self.n[d] if fold_hi[d] else self.L[d] / (self.n[d] - 0.5)) if fold_lo[d] else
                         (self.L[d] / self.n[d] if self.periodic[d] else self.L[d] / (self.n[d] - 1))))
                       for d in range(3))
        self.coord_shift = tuple(self.h[d] / 2 if fold_lo[d] and self.n[d] > 1 else 0.0 for d in range(3))
        # the mapping Jacobian at the nodes (metric.jl:100-107), handed to the device
        self.jac = tuple(None if stretch[d] is None else
                         np.ascontiguousarray(np.broadcast_to(
                             np.asarray(stretch[d].dxdxi(np.clip(np.arange(self.n[d]) * self.h[d], 0.0, 1.0)),
                                        dtype=np.float64), (self.n[d],)))
                         for d in range(3))
        # physical node coordinates of the stretched dimensions (the metric reads r, theta)
        self.xnodes = tuple(None if stretch[d] is None else np.ascontiguousarray(self.xcoord(d), dtype=np.float64)
                            for d in range(3))
        self.halo = tuple(num.n_halo if self.n[d] > 1 else 0 for d in range(3))
        self.padded = tuple(self.n[d] + 2 * self.halo[d] for d in range(3))
        self.desc = self._desc()

    def _desc(self) -> L.SolverDesc:
        prob, num = self.prob, self.num
        d = L.SolverDesc()
        for k in range(3):
            d.n_global[k] = self.n[k]
            d.L_domain[k] = self.L[k]
            d.origin[k] = self.origin[k]
        fill_common_desc(d, prob, num)
        for k in range(3):
            for s in range(2):
                bc = prob.bcs[k][s]
                arms = [bc]
                if isinstance(bc, SwitchableBC):
                    # both arms share the face's parameter slots (one per kind)
                    if type(bc.before) is type(bc.after) and not isinstance(bc.before, (SlipWallBC, PeriodicBC,
                                                                                           ExtrapolationBC)):
                        raise ValueError("SwitchableBC on the GPU backend: the two arms must be of different kinds")
                    arms = [bc.before, bc.after]
                    d.switchable[k][s] = 1
                    d.bc_after[k][s] = bc.after.code
                if arms[0].code is None or arms[-1].code is None:
                    raise ValueError(f"unsupported boundary condition {type(bc).__name__}")
                d.bc[k][s] = arms[0].code
                for arm in arms:
                    self._face_params(d, arm, k, s)
        return d

    def _face_params(self, d: L.SolverDesc, bc, k: int, s: int):
        if isinstance(bc, DirichletBC):
            q = self._dirichlet_state(bc, k, s)
            for c, v in enumerate(q):
                d.bc_state[k][s][c] = v
        elif isinstance(bc, NoSlipWallBC):
            d.Twall[k][s] = bc.Twall if bc.Twall > 0 else 0.0   # NaN / non-positive: adiabatic
        elif isinstance(bc, NSCBCOutflowBC):
            f = d.nscbc[k][s]
            f.pinf, f.sigma, f.Lref, f.beta_t = bc.pinf, bc.sigma, bc.Lref, bc.beta_t
        elif isinstance(bc, NSCBCInflowBC):
            f = d.nscbc[k][s]
            for i in range(3):
                f.u[i] = bc.u[i]
            f.T = bc.T_ion
            if len(bc.Y) != self.n_species:
                raise ValueError("NSCBCInflowBC: target composition does not match the EOS")
            for i, y in enumerate(bc.Y):
                f.Y[i] = y
            f.eta_u, f.eta_T, f.eta_t, f.eta_Y, f.Lref = bc.eta_u, bc.eta_T, bc.eta_t, bc.eta_Y, bc.Lref

    def xcoord(self, d: int) -> np.ndarray:
        """rhs.jl:320-335: node coordinates along dimension d (0-based here),
        mapped through the dimension's Stretch when it has one."""
        if self.stretch[d] is not None:
            return np.asarray(self.stretch[d].x(np.arange(self.n[d]) * self.h[d]), dtype=np.float64)
        return self.origin[d] + self.coord_shift[d] + np.arange(self.n[d]) * self.h[d]

    def _dirichlet_state(self, bc: DirichletBC, d: int, side: int):
        """The face's one conserved state.  The reference evaluates fun(x, t)
        at every stage time (problem.jl:199-204); the device holds a constant
        state, so fun is evaluated over the face at t = 0 and t = 1 and a state
        that varies over the face or in time is refused."""
        xs = [self.xcoord(k) for k in range(3)]
        xs[d] = xs[d][:1] if side == 0 else xs[d][-1:]
        X, Y, Z = np.meshgrid(*xs, indexing="ij")
        states = []
        for t in (0.0, 1.0):
            q = [np.broadcast_to(np.asarray(v, dtype=float), X.shape)
                 for v in conserved_from_prim(self.eos, bc.fun(X, Y, Z, t))]
            out = []
            for v in q:
                if np.ptp(v) != 0.0:
                    raise ValueError("DirichletBC on the GPU backend needs a state that is uniform over the face")
                out.append(float(v.flat[0]))
            states.append(out)
        if states[0] != states[1]:
            raise ValueError("DirichletBC on the GPU backend needs a state that is constant in time")
        return states[0]

    def footprint(self) -> int:
        """Device bytes the solver would allocate (cles_solver_footprint)."""
        b = C.c_size_t()
        L.check(L.load().cles_solver_footprint(C.byref(self.desc), C.byref(b)))
        return b.value

    def allocate_state(self) -> np.ndarray:
        """decomposition.jl:132: a zeroed padded host state Q[..., c] (Fortran order)."""
        return np.zeros(self.padded + (self.n_cons,), order="F")


class Solver(SolverSpec):
    """Device-resident solver (the reference's Solver + Workspace)."""

    def __init__(self, ctx: Context, prob: Problem, num: Numerics):
        super().__init__(prob, num)
        self.ctx = ctx
        h = C.c_void_p()
        if any(j is not None for j in self.jac):
            def ptrs(arrs):
                return (L.DP * 3)(*[C.cast(None, L.DP) if a is None else a.ctypes.data_as(L.DP) for a in arrs])
            L.check(ctx.lib.cles_solver_create_geom(ctx.handle, C.byref(self.desc), ptrs(self.xnodes),
                                                    ptrs(self.jac), C.byref(h)))
        else:
            L.check(ctx.lib.cles_solver_create(ctx.handle, C.byref(self.desc), C.byref(h)))
        self.handle = h
        for k in range(3):
            for side in range(2):
                bc = prob.bcs[k][side]
                if isinstance(bc, SwitchableBC):
                    bc._faces.append((self, k, side))
                    if bc.switched():
                        L.check(ctx.lib.cles_solver_switch_bc(self.handle, k, side))

    def save_checkpoint(self, prefix: str, rank: int = 0) -> str:
        """io.jl:15-41 save_checkpoint: writes prefix.rNNNN.ckpt."""
        L.check(self.ctx.lib.cles_solver_save_checkpoint(self.handle, prefix.encode(), int(rank)))
        return prefix

    def load_checkpoint(self, prefix: str, rank: int = 0):
        """io.jl:49-75 load_checkpoint!: restores Q, t and step."""
        L.check(self.ctx.lib.cles_solver_load_checkpoint(self.handle, prefix.encode(), int(rank)))

    RUN_MODES = {"host_dt": 0, "device_dt": 1, "step_graph": 2}

    def set_run_mode(self, mode: str):
        """How run! issues its steps: "host_dt" (the dt decision on the host
        every step), "device_dt" (decided on the device, read back every few
        steps) or "step_graph" (as device_dt, each step one replay of a
        captured graph).  Results are identical."""
        L.check(self.ctx.lib.cles_solver_set_run_mode(self.handle, self.RUN_MODES[mode]))

    def memory(self) -> int:
        """Device bytes held by this solver (cles_solver_memory)."""
        b = C.c_size_t()
        L.check(self.ctx.lib.cles_solver_memory(self.handle, C.byref(b)))
        return b.value

    # ---- state transfer -------------------------------------------------------
    def upload(self, Q: np.ndarray):
        Q = np.asfortranarray(Q, dtype=np.float64)
        L.check(self.ctx.lib.cles_solver_upload(self.handle, Q.ctypes.data_as(L.DP), L.I64_3(*self.padded),
                                                L.INT_3(*self.halo)))

    def download(self, Q: np.ndarray):
        assert Q.flags.f_contiguous and Q.dtype == np.float64 and Q.shape == self.padded + (self.n_cons,)
        L.check(self.ctx.lib.cles_solver_download(self.handle, Q.ctypes.data_as(L.DP), L.I64_3(*self.padded),
                                                  L.INT_3(*self.halo)))
        return Q

    def field(self, name: str) -> np.ndarray:
        """Interior of a named device field as a float64 array of shape n."""
        out = np.empty(self.n, order="F")
        L.check(self.ctx.lib.cles_solver_get_field(self.handle, name.encode(), out.ctypes.data_as(L.DP)))
        return out

    # ---- time ----------------------------------------------------------------
    @property
    def t(self):
        t, st = C.c_double(), C.c_int64()
        L.check(self.ctx.lib.cles_solver_get_time(self.handle, C.byref(t), C.byref(st)))
        return t.value

    @property
    def step_count(self):
        t, st = C.c_double(), C.c_int64()
        L.check(self.ctx.lib.cles_solver_get_time(self.handle, C.byref(t), C.byref(st)))
        return st.value

    def set_time(self, t: float, step: int):
        L.check(self.ctx.lib.cles_solver_set_time(self.handle, t, step))

    # ---- the per-step path -----------------------------------------------------
    def apply_bcs(self):
        L.check(self.ctx.lib.cles_apply_bcs(self.handle))

    def compute_rhs(self, primitives_current=False):
        L.check(self.ctx.lib.cles_compute_rhs(self.handle, int(primitives_current)))

    def max_rate(self):
        r, m = C.c_double(), C.c_double()
        L.check(self.ctx.lib.cles_max_rate(self.handle, C.byref(r), C.byref(m)))
        return r.value, m.value

    def step(self, dt, prepared=False):
        L.check(self.ctx.lib.cles_step(self.handle, float(dt), int(prepared)))

    def filter_state(self):
        L.check(self.ctx.lib.cles_filter_state(self.handle))

    def run(self, tfinal, nmax=2**62, record=0, callback=None):
        """timestep.jl:277-400 run! on the device state; returns the first
        `record` step sizes.  callback: a Callback, a tuple / list of them or
        a bare callable (callbacks.jl), run between steps; effects receive
        (solver, Q) with Q a lazy view of the device state.  The number of
        StepControl rollbacks is left in self.retries_used."""
        from .callbacks import HookAdapter, LazyState
        dts = (C.c_double * max(record, 1))()
        used = C.c_int(0)
        adapter = None
        hooks = None
        if callback is not None:
            def fetch():
                Q = self.allocate_state()
                self.download(Q)
                return Q
            adapter = HookAdapter(callback, self, LazyState(fetch), self.t, self.step_count)
            hooks = C.byref(adapter.hooks)
        rc = self.ctx.lib.cles_run_ctl(self.handle, float(tfinal), int(nmax), hooks, dts, int(record), C.byref(used))
        self.retries_used = used.value
        if adapter is not None:
            self.rewinds = adapter.rewinds
            adapter.check()
        L.check(rc)
        return np.array(dts[:record])

    @property
    def cfl(self) -> float:
        """The CFL in effect (lowered in place by StepControl rollbacks)."""
        v = C.c_double()
        L.check(self.ctx.lib.cles_solver_get_cfl(self.handle, C.byref(v)))
        return v.value

    @cfl.setter
    def cfl(self, value: float):
        L.check(self.ctx.lib.cles_solver_set_cfl(self.handle, float(value)))

    def diagnostics(self):
        ke, en = C.c_double(), C.c_double()
        L.check(self.ctx.lib.cles_diagnostics(self.handle, C.byref(ke), C.byref(en)))
        return ke.value, en.value

    def profile(self, enable: bool = True):
        """Start (fresh accumulation) or stop per-phase event timing."""
        L.check(self.ctx.lib.cles_solver_profile(self.handle, int(enable)))

    def profile_read(self):
        """{phase: (ms_total, launches, algorithmic_bytes)} for phases that ran."""
        out = {}
        name = C.create_string_buffer(64)
        ms, n, by = C.c_double(), C.c_int64(), C.c_double()
        i = 0
        while self.ctx.lib.cles_solver_profile_read(self.handle, i, name, 64, C.byref(ms), C.byref(n),
                                                    C.byref(by)) == L.OK:
            if n.value:
                out[name.value.decode()] = (ms.value, n.value, by.value)
            i += 1
        return out

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.cles_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initialize_(solver: Solver, Q: np.ndarray, ic: Callable) -> np.ndarray:
    """problem.jl:146-163 initialize!: fills the interior of Q from ic(x, y, z) -> Prim
    (vectorised over the interior nodes); halos are left unchanged."""
    X, Y, Z = np.meshgrid(*(solver.xcoord(d) for d in range(3)), indexing="ij")
    q = conserved_from_prim(solver.eos, ic(X, Y, Z))
    sl = tuple(slice(solver.halo[d], solver.halo[d] + solver.n[d]) for d in range(3))
    for c in range(solver.n_cons):
        Q[sl + (c,)] = np.broadcast_to(q[c], X.shape)
    return Q


def setup(prob: Problem, num: Numerics, ctx: Context | None = None):
    """problem.jl:321-345 setup(prob, num) -> (solver, Q); Q is also uploaded."""
    ctx = ctx or Context(0)
    solver = Solver(ctx, prob, num)
    Q = solver.allocate_state()
    initialize_(solver, Q, prob.ic)
    solver.upload(Q)
    return solver, Q


def run_(solver: Solver, Q: np.ndarray, tfinal: float, nmax: int = 2**62, upload: bool = False, callback=None):
    """timestep.jl:402-405 run!(solver, Q; tfinal, nmax, callback): advances
    the device state and writes the result back into Q.  upload=True first
    copies Q to the device (after a host-side modification)."""
    if upload:
        solver.upload(Q)
    solver.run(tfinal, nmax, callback=callback)
    solver.download(Q)
    return Q
