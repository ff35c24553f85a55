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
import CompactLES: apply_along!, AbstractDirPlan, Decomp

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

"plan_direction(decomp, scheme, dim, h) for one rank (whole-domain lines)."
function hip_plan_direction(ctx::Context, decomp::Decomp, scheme, dim::Int, h::Real; T=Float64)
    desc = Ref(scheme_desc(scheme))
    n = Int64.(collect(decomp.n_local))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cles_plan_create, LIB), Cint,
                (Ptr{Cvoid}, Ref{ClesSchemeDesc}, Cint, Ptr{Int64}, Cint, Cdouble, Cint, Ptr{Ptr{Cvoid}}),
                ctx.ptr, desc, dim, n, decomp.periodic[dim], h, T === Float64 ? CLES_F64 : CLES_F32, r))
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
# Device-resident solver: setup(...; backend = :hip) -> (HIPSolver, Q)
# ---------------------------------------------------------------------------
mutable struct HIPSolver
    ptr::Ptr{Cvoid}
    ctx::Context
    padded::NTuple{3,Int}
    halo::NTuple{3,Int}
end

"run!(solver, Q; tfinal, nmax, callback): Q (host, padded) is uploaded, advanced on the GPU and written back."
function CompactLES.run!(s::HIPSolver, Q; tfinal, nmax::Int=typemax(Int), callback=nothing)
    pd = Int64.(collect(s.padded)); hl = Cint.(collect(s.halo))
    check(ccall((:cles_solver_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), pd, hl))
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

struct ClesRunHooks
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
struct ClesBatchInstance
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
