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

"run!(solver, Q; tfinal, nmax): Q (host, padded) is uploaded, advanced on the GPU and written back."
function CompactLES.run!(s::HIPSolver, Q; tfinal, nmax::Int=typemax(Int))
    pd = Int64.(collect(s.padded)); hl = Cint.(collect(s.halo))
    check(ccall((:cles_solver_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), pd, hl))
    check(ccall((:cles_run, LIB), Cint, (Ptr{Cvoid}, Cdouble, Int64, Ptr{Cdouble}, Int64),
                s.ptr, tfinal, nmax, C_NULL, 0))
    check(ccall((:cles_solver_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Int64}, Ptr{Cint}),
                s.ptr, parent(Q), pd, hl))
    return Q
end

end # module
