# CompactLESHIP.jl -- ccall binding of libcompactles_hip.so for CompactLES.jl.
#
# Drops the MI355X backend into the reference's operator plugin interface
# (src/operators.jl:11 AbstractDirPlan, :199 apply_along!) and offers the
# device-resident solver behind setup/run! (src/problem.jl:321,
# src/timestep.jl:277).  Arrays handed to apply_along! are device buffers
# (AMDGPU.jl ROCArray or a HIPBuffer from `hip_field` below) in the reference's
# padded layout; the C ABI takes plain pointers, extents and halo widths.
#
# This file cannot be exercised in the build container (no Julia toolchain);
# tests drive the same C ABI from Python ctypes.
module CompactLESHIP

using CompactLES
import CompactLES: apply_along!, AbstractDirPlan, Decomp, Problem, Numerics, nspecies,
                   conserved_from_prim, switched

const LIB = get(ENV, "COMPACTLES_HIP_LIB",
                joinpath(@__DIR__, "..", "libcompactles_hip.so"))

const CLES_OK = Cint(0)
const CLES_F64 = Cint(0)
const CLES_F32 = Cint(1)

struct ClesSchemeDesc            # cles_scheme_desc
    q::Cint
    lhs::NTuple{4,Cdouble}
    a0::Cdouble
    M::Cint
    coeffs::NTuple{8,Cdouble}
    symmetric::Cint
    nclosure::Cint
    closure_lhs::NTuple{72,Cdouble}          # [8][9] row-major
    closure_rhs_len::NTuple{8,Cint}
    closure_rhs::NTuple{128,Cdouble}         # [8][16] row-major
end

function check(rc)
    rc == CLES_OK && return nothing
    msg = unsafe_string(ccall((:cles_last_error, LIB), Cstring, ()))
    rc in (4, 5, 6, 7, 8) && throw(CompactLES.SolverFailure(
        (:nonfinite, :negative_density, :planck, :dt_min, :dt_collapse)[rc - 3], 0, NaN, NaN, NaN, msg))
    error("compactles_hip: [$rc] $msg")
end

mutable struct Context
    ptr::Ptr{Cvoid}
end
function Context(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cles_ctx_create, LIB), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r))
    c = Context(r[])
    finalizer(c -> ccall((:cles_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), c.ptr), c)
    return c
end

pad8(v, T, N) = ntuple(i -> i <= length(v) ? T(v[i]) : zero(T), N)

function scheme_desc(s::CompactLES.CompactScheme)
    cl = zeros(Cdouble, 8, 9); cr = zeros(Cdouble, 8, 16); len = zeros(Cint, 8)
    for (j, row) in enumerate(s.closures)
        cl[j, 1:3] .= row.lhs; cr[j, 1:length(row.rhs)] .= row.rhs; len[j] = length(row.rhs)
    end
    ClesSchemeDesc(1, pad8([s.alpha], Cdouble, 4), s.a0, length(s.coeffs), pad8(s.coeffs, Cdouble, 8),
                   s.symmetric, length(s.closures), Tuple(vec(permutedims(cl))), Tuple(len),
                   Tuple(vec(permutedims(cr))))
end

function scheme_desc(s::CompactLES.BandedCompactScheme)
    cl = zeros(Cdouble, 8, 9); cr = zeros(Cdouble, 8, 16); len = zeros(Cint, 8)
    for (j, row) in enumerate(s.closures)
        cl[j, 1:length(row.lhs)] .= row.lhs; cr[j, 1:length(row.rhs)] .= row.rhs; len[j] = length(row.rhs)
    end
    ClesSchemeDesc(s.q, pad8(s.lhs, Cdouble, 4), s.a0, length(s.coeffs), pad8(s.coeffs, Cdouble, 8),
                   s.symmetric, length(s.closures), Tuple(vec(permutedims(cl))), Tuple(len),
                   Tuple(vec(permutedims(cr))))
end

"GPU plan: a new AbstractDirPlan subtype, so apply_along! dispatches to the device."
mutable struct HIPDirPlan <: AbstractDirPlan
    ptr::Ptr{Cvoid}
    dim::Int
    ctx::Context
end

_fold_arg(::Nothing) = Cint(0)
function _fold_arg(s::Integer)
    s in (1, -1) || throw(ArgumentError("a fold sign must be nothing, +1 or -1, got $s"))
    return Cint(s)
end

"""
    hip_plan_direction(ctx, decomp, scheme, dim, h; lo_fold=nothing, hi_fold=nothing, T=Float64)

plan_direction(decomp, scheme, dim, h; lo_fold, hi_fold) (operators.jl:44-101,
operators_banded.jl:29-125) for one rank (whole-domain lines).  The fold
keywords keep the reference's meaning (σg, the output's parity at a half-offset
edge).  The device plan reads no halos: its edge rows read the mirrored interior
with the input sign fold_apply! pairs with σg (σf = σg for a filter, -σg for a
derivative, folds.jl:271-274), so callers skip fold_fill!.
"""
function hip_plan_direction(ctx::Context, decomp::Decomp, scheme, dim::Int, h::Real;
                            lo_fold::Union{Nothing,Int}=nothing, hi_fold::Union{Nothing,Int}=nothing,
                            T=Float64)
    desc = Ref(scheme_desc(scheme))
    n = Int64.(collect(decomp.n_local))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cles_plan_create, LIB), Cint,
                (Ptr{Cvoid}, Ref{ClesSchemeDesc}, Cint, Ptr{Int64}, Cint, Cdouble, Cint, Cint, Cint,
                 Ptr{Ptr{Cvoid}}),
                ctx.ptr, desc, dim, n, decomp.periodic[dim], h, T === Float64 ? CLES_F64 : CLES_F32,
                _fold_arg(lo_fold), _fold_arg(hi_fold), r))
    p = HIPDirPlan(r[], dim, ctx)
    finalizer(p -> ccall((:cles_plan_destroy, LIB), Cint, (Ptr{Cvoid},), p.ptr), p)
    return p
end

"apply_along!(out, plan, f, decomp) on device arrays in the padded field layout."
function apply_along!(out, plan::HIPDirPlan, f, decomp::Decomp; stream::Ptr{Cvoid}=C_NULL)
    padded = Int64.(collect(size(f)))
    halo = Cint.(collect(decomp.n_halo_d))
    check(ccall((:cles_apply_along, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cint}, Ptr{Cvoid}),
                plan.ptr, pointer(out), pointer(f), padded, halo, stream))
    return out
end

# ---------------------------------------------------------------------------
# cles_solver_desc, field for field (include/compactles_hip.h).  Julia lays out
# isbits structs with C alignment, so these records are the C structs; nested
# C arrays are flattened row-major into NTuples (C index [d][s][c] is element
# c + 1 + CLES_MAXSP4 * (s + 2d)).
# ---------------------------------------------------------------------------
const CLES_MAXSP = 8
const CLES_MAXSP4 = CLES_MAXSP + 4
const CLES_NASA9_MAXINT = 3
const CLES_MAX_FORCES = 4

struct ClesNscbcDesc             # cles_nscbc_desc
    pinf::Cdouble
    sigma::Cdouble
    Lref::Cdouble
    beta_t::Cdouble
    u::NTuple{3,Cdouble}
    T::Cdouble
    Y::NTuple{8,Cdouble}                     # [CLES_MAXSP]
    eta_u::Cdouble
    eta_T::Cdouble
    eta_t::Cdouble
    eta_Y::Cdouble
end

struct ClesNasa9Interval         # cles_nasa9_interval
    Tmin::Cdouble
    Tmax::Cdouble
    a::NTuple{7,Cdouble}
    b1::Cdouble
end

struct ClesNasa9Species          # cles_nasa9_species
    n_intervals::Cint
    iv::NTuple{3,ClesNasa9Interval}          # [CLES_NASA9_MAXINT]
end

struct ClesSolverDesc            # cles_solver_desc
    n_global::NTuple{3,Int64}
    L_domain::NTuple{3,Cdouble}
    origin::NTuple{3,Cdouble}
    n_species::Cint
    R::NTuple{8,Cdouble}                     # [CLES_MAXSP]
    gamma::NTuple{8,Cdouble}                 # [CLES_MAXSP]
    mu0::Cdouble
    Pr::Cdouble
    Sc::Cdouble
    art_enabled::Cint
    C_mu::Cdouble
    C_beta::Cdouble
    C_kappa::Cdouble
    C_D::Cdouble
    beta_sensor::Cint
    deriv::ClesSchemeDesc
    filt::ClesSchemeDesc
    cfl::Cdouble
    filter_interval::Cint
    bc::NTuple{6,Cint}                       # [3][2]
    bc_state::NTuple{72,Cdouble}             # [3][2][CLES_MAXSP + 4]
    nscbc::NTuple{6,ClesNscbcDesc}           # [3][2]
    predict::Cdouble
    max_growth::Cdouble
    dt_min::Cdouble
    dt_min_ratio::Cdouble
    precision::Cint
    retries::Cint
    cfl_backoff::Cdouble
    savepoint_interval::Cint
    landing_steps::Cint
    eos_kind::Cint
    T_guess::Cdouble
    nasa9::NTuple{8,ClesNasa9Species}        # [CLES_MAXSP]
    metric::Cint
    Twall::NTuple{6,Cdouble}                 # [3][2]
    switchable::NTuple{6,Cint}               # [3][2]
    bc_after::NTuple{6,Cint}                 # [3][2]
    n_body_forces::Cint
    body_force::NTuple{12,Cdouble}           # [CLES_MAX_FORCES][3]
    p_inf::Cdouble
end

_zero(::Type{T}) where {T<:Real} = zero(T)
_zero(::Type{NTuple{N,T}}) where {N,T} = ntuple(_ -> _zero(T), N)
_zero(::Type{T}) where {T} = T((_zero(fieldtype(T, i)) for i in 1:fieldcount(T))...)

_conv(::Type{T}, x::Union{Tuple,AbstractArray}) where {T<:Tuple} = T(Tuple(x))
_conv(::Type{T}, x) where {T} = convert(T, x)

"Field values of a ClesSolverDesc under construction, zero unless set."
struct _DescFields
    v::Dict{Symbol,Any}
end
_DescFields() = _DescFields(Dict{Symbol,Any}(n => _zero(fieldtype(ClesSolverDesc, n))
                                              for n in fieldnames(ClesSolverDesc)))
Base.setindex!(f::_DescFields, x, k::Symbol) = (f.v[k] = _conv(fieldtype(ClesSolverDesc, k), x); f)
Base.getindex(f::_DescFields, k::Symbol) = f.v[k]
_build(f::_DescFields) = ClesSolverDesc((f.v[n] for n in fieldnames(ClesSolverDesc))...)

# Boundary-condition codes (CLES_BC_*)
_bc_code(::CompactLES.PeriodicBC) = Cint(0)
_bc_code(::CompactLES.SlipWallBC) = Cint(1)
_bc_code(::CompactLES.DirichletBC) = Cint(2)
_bc_code(::CompactLES.NSCBCOutflowBC) = Cint(3)
_bc_code(::CompactLES.NSCBCInflowBC) = Cint(4)
_bc_code(::CompactLES.AxisBC) = Cint(5)
_bc_code(::CompactLES.OriginBC) = Cint(6)
_bc_code(::CompactLES.PoleBC) = Cint(7)
_bc_code(::CompactLES.NoSlipWallBC) = Cint(8)
_bc_code(::CompactLES.ExtrapolationBC) = Cint(9)
_bc_code(bc) = error("boundary condition $(typeof(bc)) is not available on the HIP backend")

_metric_code(::CompactLES.CartesianMetric) = Cint(0)
_metric_code(::CompactLES.CylindricalMetric) = Cint(1)
_metric_code(::CompactLES.SphericalMetric) = Cint(2)
_metric_code(m) = error("metric $(typeof(m)) is not available on the HIP backend")

const _BETA_SENSOR = Dict(:strain => Cint(0), :gated_strain => Cint(1), :dilatation => Cint(2))

"The grid of setup(prob, num) (rhs.jl:185-205, 286): spacing, origin and fold shift per dimension."
struct HIPGrid
    n::NTuple{3,Int}
    periodic::NTuple{3,Bool}
    h::NTuple{3,Float64}
    origin::NTuple{3,Float64}
    coord_shift::NTuple{3,Float64}
    stretch::NTuple{3,Any}
    n_halo::Int
end

function HIPGrid(prob::Problem, num::Numerics)
    n = num.n_global
    L = ntuple(d -> prob.domain[d][2] - prob.domain[d][1], 3)
    all(>(0), L) || error("domain extents must be positive")
    periodic = ntuple(d -> n[d] == 1 || prob.bcs[d][1] isa CompactLES.PeriodicBC, 3)
    stretch = num.stretch
    for d in 1:3
        st = stretch[d]
        st === nothing && continue
        # problem.jl:324-332
        isapprox(st.x(0.0), prob.domain[d][1]; atol=1e-10 * L[d]) &&
        isapprox(st.x(1.0), prob.domain[d][2]; atol=1e-10 * L[d]) ||
            error("stretch mapping for dim $d does not span the domain: " *
                  "x(0) = $(st.x(0.0)), x(1) = $(st.x(1.0)), domain = $(prob.domain[d])")
        periodic[d] && error("dimension $d: stretched dimensions must be non-periodic")
    end
    b = prob.bcs
    fold_lo = (b[1][1] isa CompactLES.AxisBC || b[1][1] isa CompactLES.OriginBC,
               b[2][1] isa CompactLES.PoleBC, false)
    fold_hi = (false, b[2][2] isa CompactLES.PoleBC, false)
    h = ntuple(3) do d
        n[d] == 1 && return 1.0
        stretch[d] === nothing || return 1.0 / (n[d] - 1)
        fold_lo[d] && fold_hi[d] && return L[d] / n[d]
        fold_lo[d] && return L[d] / (n[d] - 0.5)
        periodic[d] ? L[d] / n[d] : L[d] / (n[d] - 1)
    end
    origin = ntuple(d -> stretch[d] === nothing ? Float64(prob.domain[d][1]) : 0.0, 3)
    shift = ntuple(d -> fold_lo[d] && n[d] > 1 ? h[d] / 2 : 0.0, 3)
    return HIPGrid(n, periodic, h, origin, shift, stretch, num.n_halo)
end

"global_xcoord (rhs.jl:324-328): coordinate of one-based node g along d."
function xcoord(g::HIPGrid, d::Int, i::Int)
    ξ = g.origin[d] + g.coord_shift[d] + (i - 1) * g.h[d]
    st = g.stretch[d]
    return st === nothing ? ξ : st.x(ξ)
end

"""
DirichletBC on the device holds one conserved state per face: `fun` is
evaluated over the whole face plane at t = 0 and t = 1, and a state that
varies over the face or in time is refused (the reference evaluates it at
every stage time, problem.jl:199-204).
"""
function _dirichlet_state(bc, g::HIPGrid, eos, d::Int, side::Int)
    ranges = ntuple(k -> k == d ? (side == 1 ? (1:1) : (g.n[k]:g.n[k])) : (1:g.n[k]), 3)
    q0 = nothing
    for t in (0.0, 1.0), k in ranges[3], j in ranges[2], i in ranges[1]
        q = conserved_from_prim(eos, bc.fun(xcoord(g, 1, i), xcoord(g, 2, j), xcoord(g, 3, k), t))
        q0 === nothing && (q0 = q)
        q == q0 || error("DirichletBC on the HIP backend needs a state that is uniform over the face " *
                         "and constant in time (dim $d, side $side)")
    end
    return q0
end

function _face_params!(f::_DescFields, bc, g::HIPGrid, eos, ns::Int, d::Int, side::Int,
                       bc_state, nscbc, Twall)
    fi = side + 2 * (d - 1)                     # 1-based index of C [d-1][side-1]
    if bc isa CompactLES.DirichletBC
        q = _dirichlet_state(bc, g, eos, d, side)
        bc_state[1:length(q), side, d] .= q
    elseif bc isa CompactLES.NoSlipWallBC
        Twall[side, d] = isfinite(bc.Twall) && bc.Twall > 0 ? bc.Twall : 0.0   # NaN: adiabatic
    elseif bc isa CompactLES.NSCBCOutflowBC
        z = nscbc[fi]
        nscbc[fi] = ClesNscbcDesc(bc.pinf, bc.sigma, bc.Lref, bc.beta_t, z.u, z.T, z.Y,
                                  z.eta_u, z.eta_T, z.eta_t, z.eta_Y)
    elseif bc isa CompactLES.NSCBCInflowBC
        bc.target === nothing ||
            error("NSCBCInflowBC with a target function is not available on the HIP backend")
        length(bc.Y) == ns || error("NSCBCInflowBC: target composition does not match the EOS")
        z = nscbc[fi]
        nscbc[fi] = ClesNscbcDesc(z.pinf, z.sigma, bc.Lref, z.beta_t, Tuple(Float64.(bc.u)), bc.T_ion,
                                  ntuple(k -> k <= ns ? bc.Y[k] : 0.0, CLES_MAXSP),
                                  bc.eta_u, bc.eta_T, bc.eta_t, bc.eta_Y)
    end
    return nothing
end

function _eos_fields!(f::_DescFields, eos)
    ns = nspecies(eos)
    ns <= CLES_MAXSP || error("at most $CLES_MAXSP species on the HIP backend")
    f[:n_species] = ns
    R = zeros(Cdouble, CLES_MAXSP); γ = zeros(Cdouble, CLES_MAXSP)
    if eos isa CompactLES.IdealMixture                          # physics.jl:104-126
        f[:eos_kind] = 0
        for k in 1:ns
            R[k] = eos.sp[k].R; γ[k] = eos.sp[k].gamma
        end
    elseif eos isa CompactLES.Nasa9Mixture                      # physics.jl:402-571
        f[:eos_kind] = 1
        f[:T_guess] = eos.T_guess
        tabs = [_zero(ClesNasa9Species) for _ in 1:CLES_MAXSP]
        for k in 1:ns
            sp = eos.sp[k]
            R[k] = sp.R
            ni = length(sp.intervals)
            ni <= CLES_NASA9_MAXINT || error("NASA-9 species $(sp.name): at most $CLES_NASA9_MAXINT intervals")
            iv = ntuple(i -> i <= ni ?
                        ClesNasa9Interval(sp.intervals[i].Tmin, sp.intervals[i].Tmax,
                                          Tuple(Float64.(sp.intervals[i].a)), sp.intervals[i].b1) :
                        _zero(ClesNasa9Interval), CLES_NASA9_MAXINT)
            tabs[k] = ClesNasa9Species(ni, iv)
        end
        f[:nasa9] = tabs
    elseif eos isa CompactLES.StiffenedGas                      # physics.jl:198-250
        f[:eos_kind] = 2
        R[1] = (eos.gamma - 1) * eos.cv; γ[1] = eos.gamma
        f[:p_inf] = eos.p_inf
    else
        error("EOS $(typeof(eos)) is not available on the HIP backend")
    end
    f[:R] = R; f[:gamma] = γ
    return ns
end

"""
    solver_desc(prob::Problem, num::Numerics; precision=:fp64) -> (Ref{ClesSolverDesc}, HIPGrid)

The C-ABI descriptor of setup(prob, num) (problem.jl:243-345): grid and domain,
EOS (physics.jl:104-126, 198-250, 402-571), Transport (physics.jl:589-593),
ArtParams (artificial.jl:50-57), schemes, cfl, filter cadence, StepControl
(stepcontrol.jl:116-136), metric, boundary conditions with their parameters,
and ConstantBodyForce sources.  `precision = :fp32` selects the mixed form.
"""
function solver_desc(prob::Problem, num::Numerics; precision::Symbol=:fp64)
    g = HIPGrid(prob, num)
    f = _DescFields()
    f[:n_global] = Int64.(num.n_global)
    f[:L_domain] = ntuple(d -> prob.domain[d][2] - prob.domain[d][1], 3)
    f[:origin] = ntuple(d -> Float64(prob.domain[d][1]), 3)
    eos = prob.eos
    ns = _eos_fields!(f, eos)
    tr = prob.transport
    f[:mu0] = tr.mu0; f[:Pr] = tr.Pr; f[:Sc] = tr.Sc
    art = num.art
    haskey(_BETA_SENSOR, art.beta_sensor) ||
        error("art.beta_sensor must be :strain, :gated_strain or :dilatation, got :$(art.beta_sensor)")
    f[:art_enabled] = art.enabled ? 1 : 0
    f[:C_mu] = art.C_mu; f[:C_beta] = art.C_beta; f[:C_kappa] = art.C_kappa; f[:C_D] = art.C_D
    f[:beta_sensor] = _BETA_SENSOR[art.beta_sensor]
    f[:deriv] = scheme_desc(num.deriv)
    f[:filt] = scheme_desc(num.filt)
    f[:cfl] = num.cfl
    f[:filter_interval] = num.filter_interval
    c = num.control
    f[:predict] = c.predict; f[:max_growth] = c.max_growth
    f[:dt_min] = c.dt_min; f[:dt_min_ratio] = c.dt_min_ratio
    f[:retries] = c.retries; f[:cfl_backoff] = c.cfl_backoff
    f[:savepoint_interval] = c.savepoint_interval; f[:landing_steps] = c.landing_steps
    precision in (:fp64, :fp32) || error("precision must be :fp64 or :fp32")
    f[:precision] = precision === :fp64 ? CLES_F64 : CLES_F32
    f[:metric] = _metric_code(prob.metric)
    # faces: codes, SwitchableBC arms, parameters (one slot per kind and face)
    codes = zeros(Cint, 2, 3); after = zeros(Cint, 2, 3); sw = zeros(Cint, 2, 3)
    bc_state = zeros(Cdouble, CLES_MAXSP4, 2, 3); Twall = zeros(Cdouble, 2, 3)
    nscbc = [_zero(ClesNscbcDesc) for _ in 1:6]
    for d in 1:3, side in 1:2
        bc = prob.bcs[d][side]
        arms = Any[bc]
        if bc isa CompactLES.SwitchableBC
            typeof(bc.before) == typeof(bc.after) &&
                !(bc.before isa Union{CompactLES.SlipWallBC,CompactLES.PeriodicBC,CompactLES.ExtrapolationBC}) &&
                error("SwitchableBC on the HIP backend: the two arms must be of different kinds")
            arms = Any[bc.before, bc.after]
            sw[side, d] = 1
            after[side, d] = _bc_code(bc.after)
        end
        codes[side, d] = _bc_code(arms[1])
        _bc_code(arms[end])
        for arm in arms
            _face_params!(f, arm, g, eos, ns, d, side, bc_state, nscbc, Twall)
        end
    end
    f[:bc] = vec(codes); f[:bc_after] = vec(after); f[:switchable] = vec(sw)
    f[:bc_state] = vec(bc_state); f[:Twall] = vec(Twall); f[:nscbc] = nscbc
    srcs = prob.sources
    length(srcs) <= CLES_MAX_FORCES || error("at most $CLES_MAX_FORCES sources on the HIP backend")
    forces = zeros(Cdouble, 3, CLES_MAX_FORCES)
    for (i, src) in enumerate(srcs)
        src isa CompactLES.ConstantBodyForce ||
            error("source $(typeof(src)) is not available on the HIP backend (ConstantBodyForce is)")
        forces[:, i] .= src.acceleration
    end
    f[:n_body_forces] = length(srcs); f[:body_force] = vec(forces)
    return Ref(_build(f)), g
end

# ---------------------------------------------------------------------------
# Device-resident solver: setup(prob, num; backend = :hip) -> (HIPSolver, Q)
# ---------------------------------------------------------------------------
mutable struct HIPSolver
    ptr::Ptr{Cvoid}
    ctx::Context
    padded::NTuple{3,Int}
    halo::NTuple{3,Int}
    grid::HIPGrid
    eos::Any
    n_cons::Int
    control::CompactLES.StepControl
    desc::Base.RefValue{ClesSolverDesc}
end

"""
    HIPSolver(ctx, prob, num; precision=:fp64)

setup's Solver (problem.jl:321-345, rhs.jl:122-311) on the device: the
descriptor from `solver_desc`, and for Stretch mappings the node coordinates
and dx/dξ at ξ = (i - 1)/(n - 1) (metric.jl:56-75) through
cles_solver_create_geom.
"""
function HIPSolver(ctx::Context, prob::Problem, num::Numerics; precision::Symbol=:fp64)
    desc, g = solver_desc(prob, num; precision=precision)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    if all(st -> st === nothing, g.stretch)
        check(ccall((:cles_solver_create, LIB), Cint, (Ptr{Cvoid}, Ref{ClesSolverDesc}, Ptr{Ptr{Cvoid}}),
                    ctx.ptr, desc, r))
    else
        xs = [g.stretch[d] === nothing ? Float64[] : [xcoord(g, d, i) for i in 1:g.n[d]] for d in 1:3]
        js = [g.stretch[d] === nothing ? Float64[] :
              [Float64(g.stretch[d].dxdξ((i - 1) * g.h[d])) for i in 1:g.n[d]] for d in 1:3]
        xp = [isempty(v) ? Ptr{Cdouble}(C_NULL) : pointer(v) for v in xs]
        jp = [isempty(v) ? Ptr{Cdouble}(C_NULL) : pointer(v) for v in js]
        GC.@preserve xs js begin
            check(ccall((:cles_solver_create_geom, LIB), Cint,
                        (Ptr{Cvoid}, Ref{ClesSolverDesc}, Ptr{Ptr{Cdouble}}, Ptr{Ptr{Cdouble}}, Ptr{Ptr{Cvoid}}),
                        ctx.ptr, desc, xp, jp, r))
        end
    end
    halo = ntuple(d -> g.n[d] > 1 ? g.n_halo : 0, 3)
    padded = ntuple(d -> g.n[d] + 2 * halo[d], 3)
    s = HIPSolver(r[], ctx, padded, halo, g, prob.eos, nspecies(prob.eos) + 4, num.control, desc)
    finalizer(s -> ccall((:cles_solver_destroy, LIB), Cint, (Ptr{Cvoid},), s.ptr), s)
    for d in 1:3, side in 1:2
        bc = prob.bcs[d][side]
        bc isa CompactLES.SwitchableBC && switched(bc) &&
            check(ccall((:cles_solver_switch_bc, LIB), Cint, (Ptr{Cvoid}, Cint, Cint), s.ptr, d - 1, side - 1))
    end
    return s
end

"""
    CompactLESHIP.setup(prob, num; backend=:hip, precision=:fp64, ctx=nothing) -> (solver, Q)

setup(prob, num) (problem.jl:321-345) with a backend choice.  `backend = :cpu`
is the reference's own setup.  `backend = :hip` returns the device solver and
Q, the padded host state (allocate_state's layout, decomposition.jl:129-134,
one rank) with its interior initialised from prob.ic at the nodes
(problem.jl:146-163) and uploaded to the device.  This is a function of this
module rather than a method of CompactLES.setup: keywords do not take part in
dispatch, so a method with setup's positional signature would replace the
reference's.
"""
function setup(prob::Problem, num::Numerics; backend::Symbol=:hip, precision::Symbol=:fp64,
               ctx::Union{Nothing,Context}=nothing)
    backend === :cpu && return CompactLES.setup(prob, num)
    backend === :hip || error("backend must be :hip or :cpu, got :$backend")
    s = HIPSolver(ctx === nothing ? Context(0) : ctx, prob, num; precision=precision)
    g = s.grid
    Q = zeros(Float64, s.padded..., s.n_cons)
    o = s.halo
    for k in 1:g.n[3], j in 1:g.n[2], i in 1:g.n[1]
        q = conserved_from_prim(s.eos, prob.ic(xcoord(g, 1, i), xcoord(g, 2, j), xcoord(g, 3, k)))
        for c in 1:s.n_cons
            Q[i + o[1], j + o[2], k + o[3], c] = q[c]
        end
    end
    upload!(s, Q)
    return s, Q
end

"cles_solver_upload: the padded host state Q becomes the device state."
function upload!(s::HIPSolver, Q)
    check(ccall((:cles_solver_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), Int64.(collect(s.padded)), Cint.(collect(s.halo))))
    return Q
end

"cles_solver_download: the device state into the interior of the padded host state Q."
function download!(Q, s::HIPSolver)
    check(ccall((:cles_solver_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), Int64.(collect(s.padded)), Cint.(collect(s.halo))))
    return Q
end

"max_rate(solver, Q) -> (rate, ρ_min) (timestep.jl:125-175) of the host state Q."
function CompactLES.max_rate(s::HIPSolver, Q)
    upload!(s, Q)
    rate = Ref{Cdouble}(0); rmin = Ref{Cdouble}(0)
    check(ccall((:cles_max_rate, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), s.ptr, rate, rmin))
    return rate[], rmin[]
end

"filter_state!(solver, Q) (timestep.jl:414-426)."
function CompactLES.filter_state!(s::HIPSolver, Q)
    upload!(s, Q)
    check(ccall((:cles_filter_state, LIB), Cint, (Ptr{Cvoid},), s.ptr))
    return download!(Q, s)
end

"step!(solver, Q, dQ, du, dt, prepared) (timestep.jl:72-99) on the device's dQ and du."
function CompactLES.step!(s::HIPSolver, Q, dt::Real, prepared::Bool=false)
    upload!(s, Q)
    check(ccall((:cles_step, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cint), s.ptr, dt, prepared))
    return download!(Q, s)
end

"The device solver's time and completed steps (solver.t, solver.step)."
function solver_time(s::HIPSolver)
    t = Ref{Cdouble}(0); st = Ref{Int64}(0)
    check(ccall((:cles_solver_get_time, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}), s.ptr, t, st))
    return t[], Int(st[])
end

"Kinetic energy and enstrophy of the device state (FP64, fixed order)."
function diagnostics(s::HIPSolver)
    ke = Ref{Cdouble}(0); en = Ref{Cdouble}(0)
    check(ccall((:cles_diagnostics, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), s.ptr, ke, en))
    return ke[], en[]
end

"""
run!(solver, Q; tfinal, nmax, callback): Q (host, padded) is uploaded,
advanced on the GPU and written back.  StepControl is the one given to setup
(a different `control` here is refused: the device solver holds it).
"""
function CompactLES.run!(s::HIPSolver, Q; tfinal, nmax::Int=typemax(Int), callback=nothing,
                         control::Union{Nothing,CompactLES.StepControl}=nothing)
    control === nothing || control == s.control ||
        error("the HIP backend takes StepControl at setup (Numerics.control)")
    pd = Int64.(collect(s.padded)); hl = Cint.(collect(s.halo))
    upload!(s, Q)
    if callback === nothing
        check(ccall((:cles_run, LIB), Cint, (Ptr{Cvoid}, Cdouble, Int64, Ptr{Cdouble}, Int64),
                    s.ptr, tfinal, nmax, C_NULL, 0))
    else
        _run_hooked!(s, Q, tfinal, nmax, callback, pd, hl)
    end
    check(ccall((:cles_solver_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), pd, hl))
    return Q
end

# The reference's own callback machinery (callbacks.jl: callback_next_time,
# run_callbacks!, rewind_callbacks!) drives the GPU run through cles_run_hooks.
# Triggers read `solver.t` / `solver.step` from a clock; effects get that clock
# and Q refreshed from the device before they run.
mutable struct HIPClock
    t::Float64
    step::Int
    dt_prev::Float64
    wall_step::Float64
    wall_total::Float64
end

mutable struct _Hooked
    cb::Any
    clock::HIPClock
    s::HIPSolver
    Q::Any
    pd::Vector{Int64}
    hl::Vector{Cint}
    err::Any
end

struct ClesRunHooks              # cles_run_hooks
    user::Ptr{Cvoid}
    next_time::Ptr{Cvoid}
    after_step::Ptr{Cvoid}
    rewind::Ptr{Cvoid}
end

function _hook_next(user::Ptr{Cvoid}, t::Cdouble, step::Int64)::Cdouble
    h = unsafe_pointer_to_objref(user)::_Hooked
    h.err === nothing || return Inf
    h.clock.t = t; h.clock.step = step
    try
        return CompactLES.callback_next_time(h.cb, h.clock)
    catch e
        h.err = e
        return Inf
    end
end

function _hook_after(user::Ptr{Cvoid}, t::Cdouble, step::Int64, dt::Cdouble)::Cint
    h = unsafe_pointer_to_objref(user)::_Hooked
    h.err === nothing || return Cint(1)
    h.clock.t = t; h.clock.step = step; h.clock.dt_prev = dt
    try
        check(ccall((:cles_solver_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                    h.s.ptr, parent(h.Q), h.pd, h.hl))
        return CompactLES.run_callbacks!(h.cb, h.clock, h.Q) ? Cint(1) : Cint(0)
    catch e
        h.err = e
        return Cint(1)
    end
end

function _hook_rewind(user::Ptr{Cvoid}, t::Cdouble, step::Int64)::Cvoid
    h = unsafe_pointer_to_objref(user)::_Hooked
    CompactLES.rewind_callbacks!(h.cb, t, step)
    return nothing
end

function _run_hooked!(s::HIPSolver, Q, tfinal, nmax, callback, pd, hl)
    t = Ref{Cdouble}(0); st = Ref{Int64}(0)
    check(ccall((:cles_solver_get_time, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}), s.ptr, t, st))
    h = _Hooked(callback, HIPClock(t[], st[], 0.0, 0.0, 0.0), s, Q, pd, hl, nothing)
    used = Ref{Cint}(0)
    GC.@preserve h begin
        hooks = Ref(ClesRunHooks(pointer_from_objref(h),
                                 @cfunction(_hook_next, Cdouble, (Ptr{Cvoid}, Cdouble, Int64)),
                                 @cfunction(_hook_after, Cint, (Ptr{Cvoid}, Cdouble, Int64, Cdouble)),
                                 @cfunction(_hook_rewind, Cvoid, (Ptr{Cvoid}, Cdouble, Int64))))
        rc = ccall((:cles_run_ctl, LIB), Cint,
                   (Ptr{Cvoid}, Cdouble, Int64, Ref{ClesRunHooks}, Ptr{Cdouble}, Int64, Ref{Cint}),
                   s.ptr, tfinal, nmax, hooks, C_NULL, 0, used)
    end
    h.err === nothing || throw(h.err)
    check(rc)
    return used[]
end

# ---------------------------------------------------------------------------
# Batches of one-dimensional instances and the serving loop (no counterpart in
# the reference, which runs one problem per process): thin ccall wrappers.
# `desc` is the cles_solver_desc built from (Problem, Numerics) exactly as for
# HIPSolver; instance i is (n, L, origin, tfinal, nmax, Dirichlet states).
# ---------------------------------------------------------------------------
struct ClesBatchInstance         # cles_batch_instance
    n::Int64
    L::Cdouble
    origin::Cdouble
    tfinal::Cdouble
    nmax::Int64
    bc_state::NTuple{24,Cdouble}        # [2][CLES_MAXSP + 4], row-major
end

mutable struct HIPBatch
    ptr::Ptr{Cvoid}
    ctx::Context
end

function HIPBatch(ctx::Context, desc::Ref, instances::Vector{ClesBatchInstance})
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cles_batch_create, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{ClesBatchInstance}, Ptr{Ptr{Cvoid}}),
                ctx.ptr, desc, length(instances), instances, r))
    b = HIPBatch(r[], ctx)
    finalizer(b -> ccall((:cles_batch_destroy, LIB), Cint, (Ptr{Cvoid},), b.ptr), b)
    return b
end

"Advance every instance (each to its own tfinal / nmax); returns (t, steps, status) per instance."
function run_batch!(b::HIPBatch; max_steps::Int=typemax(Int))
    taken = Ref{Int64}(0)
    check(ccall((:cles_batch_run, LIB), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Cdouble}, Ptr{Int64}),
                b.ptr, max_steps, 0, C_NULL, taken))
    cnt = Ref{Int64}(0); npts = Ref{Int64}(0); nc = Ref{Cint}(0)
    check(ccall((:cles_batch_shape, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cint}), b.ptr, cnt, npts, nc))
    t = zeros(Cdouble, cnt[]); st = zeros(Int64, cnt[]); status = zeros(Cint, cnt[])
    check(ccall((:cles_batch_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                b.ptr, t, st, status))
    return t, st, status
end

mutable struct HIPServer
    ptr::Ptr{Cvoid}
end

function HIPServer(; device::Integer=0, hbm_budget::Integer=64 * 2^30, runners::Integer=4)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cles_server_create, LIB), Cint, (Cint, Csize_t, Cint, Ptr{Ptr{Cvoid}}),
                device, hbm_budget, runners, r))
    s = HIPServer(r[])
    finalizer(s -> ccall((:cles_server_destroy, LIB), Cint, (Ptr{Cvoid},), s.ptr), s)
    return s
end

"Queue setup + run!(...; tfinal, nmax) of the padded host state Q; returns the request id."
function submit!(s::HIPServer, desc::Ref, Q::Array{Float64}, halo::NTuple{3,Int}; tfinal, nmax::Int=typemax(Int))
    id = Ref{Int64}(0)
    check(ccall((:cles_server_submit, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}, Cdouble, Int64, Ptr{Int64}),
                s.ptr, desc, Q, Int64.(collect(size(Q)[1:3])), Cint.(collect(halo)), tfinal, nmax, id))
    return id[]
end

"Wait for request `id` and copy its final state into Q; returns (status, t, steps)."
function fetch!(s::HIPServer, id::Integer, Q::Array{Float64}, halo::NTuple{3,Int})
    st = Ref{Cint}(0)
    check(ccall((:cles_server_wait, LIB), Cint, (Ptr{Cvoid}, Int64, Cdouble, Ptr{Cint}), s.ptr, id, -1.0, st))
    status = Ref{Cint}(0); t = Ref{Cdouble}(0); steps = Ref{Int64}(0)
    check(ccall((:cles_server_result, LIB), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}, Ptr{Cint}, Ptr{Cdouble}, Ptr{Int64}),
                s.ptr, id, Q, Int64.(collect(size(Q)[1:3])), Cint.(collect(halo)), status, t, steps))
    return status[], t[], steps[]
end

# ---------------------------------------------------------------------------
# Line-tile table (chunk lengths of the solver's line plans; no counterpart in
# the reference).  Tune before setup(...; backend = :hip) so the solver uses it.
# ---------------------------------------------------------------------------
"Time the candidate chunk lengths on desc's grid; returns (chunks, ms per repetition) per direction."
function tune_line_tiles!(ctx::Context, desc::Ref; reps::Integer=5)
    chunks = zeros(Cint, 3); ms = zeros(Cdouble, 3)
    check(ccall((:cles_tune_solver, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cdouble}),
                ctx.ptr, desc, reps, chunks, ms))
    return Tuple(chunks), Tuple(ms)
end
save_line_tiles(path::AbstractString) = check(ccall((:cles_tune_save, LIB), Cint, (Cstring,), path))
load_line_tiles(path::AbstractString) = check(ccall((:cles_tune_load, LIB), Cint, (Cstring,), path))

end # module
