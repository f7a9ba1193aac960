# BernsteinBasisAMD.jl -- the operator API of BernsteinBasis.jl on MI355X through libbbdg.so.
#
# BernsteinBasis.jl exports its operators as Julia types with LinearAlgebra.mul! methods
# (src/BernsteinBasis.jl:19-23).  This module adds methods of the same functions for
# device arrays (AMDGPU.jl ROCArray), each of which calls one entry point of the C ABI in
# include/bbdg.h.  CPU arrays keep dispatching to BernsteinBasis.jl's own methods, so the
# right-hand sides of examples/3d_wave.jl:36-84 and examples/3d_maxwell.jl:36-93 run on the
# GPU unchanged once their arrays are ROCArrays:
#
#   reference method (file:line)                                  entry point
#   mul!(out, D{r,s,t}, x)          BernsteinDerivativeMatrix_3D.jl:134,176,218,222   bb_deriv_apply
#   mul!(out, L::BernsteinLift, x)  BernsteinLift.jl:191-220 (vector), :264-269       bb_lift_apply
#   threaded_mul!(out, L::MultithreadedBernsteinLift, x, thread)  BernsteinLift.jl:233-262
#                                                                                  bb_lift_apply
#   face_lift_multiply!(out, L, x)  BernsteinLift.jl:166-179                           bb_face_lift_apply
#   L * X                           BernsteinLift.jl:271-274                           bb_lift_apply
#   ElevationMatrix{N} * x, mul!    ElevationMatrix.jl:11-24 (generic AbstractMatrix)  bb_elevate_apply
#   uM .= view(u, Fmask, :)         examples/3d_wave.jl:41                             bb_face_gather
#   interface flux loop             examples/3d_wave.jl:43-56, 3d_maxwell.jl:43-65    bb_interface_flux
#   rhs_matvec! + ODE solve         examples/3d_wave.jl:36-134, 3d_maxwell.jl:36-138  bb_solver_*
#
# Element types: Float64 / Float32 scalars, or SVector{F,Float64/Float32} (the reference's
# multi-field columns, stored field fastest).  Julia's column-major Matrix{SVector{F,T}} of
# size (n, K) is passed as-is: field stride 1, node stride F, element stride n*F.
#
# Ordering: the library works on the stream of its own context.  Each call first waits for
# AMDGPU's stream (so inputs written by Julia kernels are complete) and, unless ASYNC[] is
# set, waits for the library's stream before returning (so Julia kernels may read the
# outputs at once).
#
# Status: written against AMDGPU.jl's documented ROCArray/pointer interface; it has not been
# run here (no Julia runtime in the build environment).  tests/test_julia_shim.py checks
# every ccall below against the C declarations in include/bbdg.h (name, arity, types).
module BernsteinBasisAMD

using LinearAlgebra
using StaticArrays
using AMDGPU
using BernsteinBasis
import BernsteinBasis: BernsteinDerivativeMatrix_3D_r, BernsteinDerivativeMatrix_3D_s,
                       BernsteinDerivativeMatrix_3D_t, BernsteinLift, MultithreadedBernsteinLift,
                       ElevationMatrix, threaded_mul!, face_lift_multiply!

export BBContext, BBSolver, BBBatch, context, synchronize_library, rhs!, step!, capture!, launch!,
       autotune!, face_gather!, interface_flux!, grad!

const LIB = joinpath(@__DIR__, "..", "libbbdg.so")

# enum values of include/bbdg.h
const BB_OK = Cint(0)
const BB_F64 = Cint(0); const BB_F32 = Cint(1)
const BB_AXIS_R = Cint(0); const BB_AXIS_S = Cint(1); const BB_AXIS_T = Cint(2)
const BB_PDE_WAVE = Cint(0); const BB_PDE_MAXWELL = Cint(1)
const BB_FLUX_REFERENCE = Cint(0); const BB_FLUX_UPWIND = Cint(1)
const BB_LIFT_PLAIN = Cint(0); const BB_LIFT_ACC64 = Cint(1); const BB_LIFT_TWOSUM = Cint(2)
const BB_LIFT_ACC64_MFMA = Cint(3)
const BB_PREC_F64 = Cint(0); const BB_PREC_MIXED = Cint(1)
const BB_PATH_PLAIN = Cint(0); const BB_PATH_FUSED = Cint(1)
const BB_ERR_NONFINITE = Cint(7)

"""`bb_layout`: strides in array elements of the element, node and field indices."""
struct Layout
    elem::Int64
    node::Int64
    field::Int64
end

# ------------------------------------------------------------------------- context

"""One library context: a HIP stream, its events and graphs on one device."""
mutable struct BBContext
    ptr::Ptr{Cvoid}
    function BBContext(device::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:bb_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), Cint(device), r)
        rc == BB_OK || error("bb_ctx_create failed with status $rc")
        c = new(r[])
        finalizer(c) do c
            c.ptr == C_NULL || ccall((:bb_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), c.ptr)
            c.ptr = C_NULL
        end
        c
    end
end

const CTX = Ref{Union{Nothing,BBContext}}(nothing)
"""The module's default context (created on first use, device 0)."""
context() = (CTX[] === nothing && (CTX[] = BBContext(0)); CTX[]::BBContext)

"""Skip the wait for the library's stream at the end of each call (then call
`synchronize_library()` before Julia code reads the outputs)."""
const ASYNC = Ref(false)

version() = unsafe_string(ccall((:bb_version, LIB), Cstring, ()))
last_error(c::BBContext = context()) = unsafe_string(ccall((:bb_last_error, LIB), Cstring, (Ptr{Cvoid},), c.ptr))

function check(rc::Cint, c::BBContext = context())
    rc == BB_OK || error("libbbdg: status $rc: $(last_error(c))")
    nothing
end

function device_count()
    n = Ref{Cint}(0)
    rc = ccall((:bb_device_count, LIB), Cint, (Ref{Cint},), n)
    rc == BB_OK || error("bb_device_count failed with status $rc")
    Int(n[])
end

function device_info(c::BBContext = context())
    arch = zeros(UInt8, 64); name = zeros(UInt8, 256); cu = Ref{Cint}(0)
    check(ccall((:bb_device_info, LIB), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Cint, Ptr{UInt8}, Cint, Ref{Cint}),
                c.ptr, arch, Cint(length(arch)), name, Cint(length(name)), cu), c)
    (arch = unsafe_string(pointer(arch)), name = unsafe_string(pointer(name)), cu = Int(cu[]))
end

function stream(c::BBContext = context())
    s = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bb_ctx_stream, LIB), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), c.ptr, s), c)
    s[]
end

synchronize_library(c::BBContext = context()) = check(ccall((:bb_synchronize, LIB), Cint, (Ptr{Cvoid},), c.ptr), c)

# inputs produced on AMDGPU's stream must be complete before the library's stream reads them
enter() = AMDGPU.synchronize()
leave(c::BBContext) = ASYNC[] || synchronize_library(c)

# ------------------------------------------------------------------------- memory
# Library-owned device buffers (ROCArrays are the usual alternative).

function malloc(nbytes::Integer, c::BBContext = context())
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bb_malloc, LIB), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), c.ptr, Csize_t(nbytes), p), c)
    p[]
end
free(p::Ptr{Cvoid}, c::BBContext = context()) = check(ccall((:bb_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c.ptr, p), c)
copy_h2d!(dst::Ptr{Cvoid}, src::Ptr{Cvoid}, n::Integer, c::BBContext = context()) =
    check(ccall((:bb_memcpy_h2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), c.ptr, dst, src, Csize_t(n)), c)
copy_d2h!(dst::Ptr{Cvoid}, src::Ptr{Cvoid}, n::Integer, c::BBContext = context()) =
    check(ccall((:bb_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), c.ptr, dst, src, Csize_t(n)), c)
copy_d2d!(dst::Ptr{Cvoid}, src::Ptr{Cvoid}, n::Integer, c::BBContext = context()) =
    check(ccall((:bb_memcpy_d2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), c.ptr, dst, src, Csize_t(n)), c)
fill_zero!(dst::Ptr{Cvoid}, n::Integer, c::BBContext = context()) =
    check(ccall((:bb_fill_zero, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), c.ptr, dst, Csize_t(n)), c)
stream_copy!(dst::Ptr{Cvoid}, src::Ptr{Cvoid}, n::Integer, c::BBContext = context()) =
    check(ccall((:bb_stream_copy, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), c.ptr, dst, src, Csize_t(n)), c)

"""`count_nonfinite(x, n, T)`: the number of NaN / Inf values among n values of type T
(Float64 or Float32) at device pointer x; 0 when all are finite."""
function count_nonfinite(x::Ptr, n::Integer, ::Type{T} = Float64, c::BBContext = context()) where {T}
    cnt = Ref{Int64}(0)
    rc = ccall((:bb_count_nonfinite, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Ref{Int64}),
               c.ptr, T === Float32 ? BB_F32 : BB_F64, x, Int64(n), cnt)
    rc == BB_OK || rc == BB_ERR_NONFINITE || check(rc, c)
    Int(cnt[])
end

timer_start(c::BBContext = context()) = check(ccall((:bb_timer_start, LIB), Cint, (Ptr{Cvoid},), c.ptr), c)
function timer_stop(c::BBContext = context())
    ms = Ref{Cfloat}(0)
    check(ccall((:bb_timer_stop, LIB), Cint, (Ptr{Cvoid}, Ref{Cfloat}), c.ptr, ms), c)
    Float64(ms[])
end

# ------------------------------------------------------------------------- tables
# Bit-exact with src/ordering.jl and src/constants.jl (0-based multi-indices).

function tables(N::Integer)
    tri = zeros(Int64, 21); tet = zeros(Int64, 21); lj = zeros(Float64, 20)
    rc = ccall((:bb_tables, LIB), Cint, (Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}), Cint(N), tri, tet, lj)
    rc == BB_OK || error("bb_tables($N) failed with status $rc")
    (tri_offsets = Tuple(tri), tet_offsets = Tuple(tet), l_j = Tuple(lj))
end

function linear_to_ijkl(N::Integer)
    np = div((N + 1) * (N + 2) * (N + 3), 6)
    out = zeros(Int32, 4np)
    rc = ccall((:bb_linear_to_ijkl, LIB), Cint, (Cint, Ptr{Int32}), Cint(N), out)
    rc == BB_OK || error("bb_linear_to_ijkl($N) failed with status $rc")
    [Tuple(out[4i-3:4i]) for i in 1:np]
end

function linear_to_ijk_2d(N::Integer)
    nfp = div((N + 1) * (N + 2), 2)
    out = zeros(Int32, 3nfp)
    rc = ccall((:bb_linear_to_ijk_2d, LIB), Cint, (Cint, Ptr{Int32}), Cint(N), out)
    rc == BB_OK || error("bb_linear_to_ijk_2d($N) failed with status $rc")
    [Tuple(out[3i-2:3i]) for i in 1:nfp]
end

"""Volume index (1-based) of each face node, faces in the order the lift expects."""
function fmask(N::Integer)
    nfp = div((N + 1) * (N + 2), 2)
    out = zeros(Int32, 4nfp)
    rc = ccall((:bb_fmask, LIB), Cint, (Cint, Ptr{Int32}), Cint(N), out)
    rc == BB_OK || error("bb_fmask($N) failed with status $rc")
    reshape(Int.(out) .+ 1, nfp, 4)
end

axis(::BernsteinDerivativeMatrix_3D_r) = BB_AXIS_R
axis(::BernsteinDerivativeMatrix_3D_s) = BB_AXIS_S
axis(::BernsteinDerivativeMatrix_3D_t) = BB_AXIS_T

"""`D[m, n]` from the library (1-based, as Base.getindex)."""
function deriv_entry(D, m::Integer, n::Integer)
    v = Ref{Cdouble}(0)
    rc = ccall((:bb_deriv_entry, LIB), Cint, (Cint, Cint, Cint, Cint, Ref{Cdouble}), axis(D), Cint(D.N),
               Cint(m - 1), Cint(n - 1), v)
    rc == BB_OK || error("bb_deriv_entry failed with status $rc")
    v[]
end

function elevation_entry(::ElevationMatrix{N}, m::Integer, n::Integer) where {N}
    v = Ref{Cdouble}(0)
    rc = ccall((:bb_elevation_entry, LIB), Cint, (Cint, Cint, Cint, Ref{Cdouble}), Cint(N), Cint(m - 1), Cint(n - 1), v)
    rc == BB_OK || error("bb_elevation_entry failed with status $rc")
    v[]
end

# ------------------------------------------------------------------------- array layouts

const DevScalar = Union{Float64,Float32}
const DevVec{T} = ROCVector{T}
const DevMat{T} = ROCMatrix{T}

dtype(::Type{Float64}) = BB_F64
dtype(::Type{Float32}) = BB_F32
scalar(::Type{SVector{F,T}}) where {F,T} = T
scalar(::Type{T}) where {T<:DevScalar} = T
nfields(::Type{SVector{F,T}}) where {F,T} = F
nfields(::Type{T}) where {T<:DevScalar} = 1

"""(layout, number of columns K, fields F) of a vector or matrix of columns."""
function layout(x::ROCArray{E}) where {E}
    F = nfields(E)
    n = size(x, 1)
    K = ndims(x) == 1 ? 1 : size(x, 2)
    (Layout(Int64(n * F), Int64(F), F == 1 ? Int64(0) : Int64(1)), Int64(K), F)
end

devptr(x::ROCArray) = reinterpret(Ptr{Cvoid}, pointer(x))

function same_columns(out, x)
    lo, Ko, Fo = layout(out); lx, Kx, Fx = layout(x)
    Ko == Kx && Fo == Fx || throw(DimensionMismatch("out has $Ko columns of $Fo fields, x has $Kx of $Fx"))
    scalar(eltype(out)) === scalar(eltype(x)) || throw(ArgumentError("out and x differ in element type"))
    (lx, lo, Kx, Fx)
end

# ------------------------------------------------------------------------- operators

const DerivOp = Union{BernsteinDerivativeMatrix_3D_r,BernsteinDerivativeMatrix_3D_s,BernsteinDerivativeMatrix_3D_t}

# src/BernsteinDerivativeMatrix_3D.jl:134-136, :176-178, :218-220 (vector) and :222-231 (matrix)
function LinearAlgebra.mul!(out::ROCArray, D::DerivOp, x::ROCArray)
    lx, lo, K, F = same_columns(out, x)
    c = context(); enter()
    check(ccall((:bb_deriv_apply, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Int64, Cint, Ptr{Cvoid}, Layout, Ptr{Cvoid}, Layout),
                c.ptr, axis(D), Cint(D.N), Cint(F), K, dtype(scalar(eltype(x))), devptr(x), lx, devptr(out), lo), c)
    leave(c)
    out
end

"""`grad!(dx, dy, dz, N, x, geo)`: the three derivative matvecs of examples/3d_wave.jl:63-71
combined with the element's geometric factors as in :75-78 (geo: 9 x K Float64 device
matrix, rows rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ)."""
function grad!(dx::ROCArray, dy::ROCArray, dz::ROCArray, N::Integer, x::ROCArray, geo::ROCMatrix{Float64})
    lx, lo, K, F = same_columns(dx, x)
    layout(dy)[1] == lo && layout(dz)[1] == lo || throw(ArgumentError("dx, dy, dz must share one layout"))
    c = context(); enter()
    check(ccall((:bb_grad_apply, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Cint, Ptr{Cvoid}, Layout, Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Layout),
                c.ptr, Cint(N), Cint(F), K, dtype(scalar(eltype(x))), devptr(x), lx,
                reinterpret(Ptr{Cdouble}, pointer(geo)), devptr(dx), devptr(dy), devptr(dz), lo), c)
    leave(c)
    (dx, dy, dz)
end

# FP32 data: FP32 chain with the signed, binomially weighted accumulation in FP64
lift_mode(::Type{Float32}) = BB_LIFT_ACC64
lift_mode(::Type{Float64}) = BB_LIFT_PLAIN

function lift_apply!(out::ROCArray, N::Integer, x::ROCArray)
    lx, lo, K, F = same_columns(out, x)
    T = scalar(eltype(x))
    c = context(); enter()
    check(ccall((:bb_lift_apply, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Cint, Cint, Ptr{Cvoid}, Layout, Ptr{Cvoid}, Layout),
                c.ptr, Cint(N), Cint(F), K, dtype(T), lift_mode(T), devptr(x), lx, devptr(out), lo), c)
    leave(c)
    out
end

# src/BernsteinLift.jl:191-220 (vector) and :264-269 (matrix): x holds [f1; f2; f3; f4]
LinearAlgebra.mul!(out::ROCArray, L::BernsteinLift, x::ROCArray) = lift_apply!(out, L.N, x)

# src/BernsteinLift.jl:233-262: the thread id selects a scratch column on the CPU; the GPU
# kernel gives every column its own work, so it is accepted and ignored
threaded_mul!(out::ROCArray, L::MultithreadedBernsteinLift, x::ROCArray, thread) = lift_apply!(out, L.N, x)

# src/BernsteinLift.jl:271-274
function Base.:*(L::BernsteinLift, x::ROCMatrix{T}) where {T<:DevScalar}
    np = div((L.N + 1) * (L.N + 2) * (L.N + 3), 6)
    lift_apply!(AMDGPU.zeros(T, np, size(x, 2)), L.N, x)
end

# src/BernsteinLift.jl:166-179: one face, Nfp -> Np
function face_lift_multiply!(out::ROCArray, L::BernsteinLift, x::ROCArray)
    lx, lo, K, F = same_columns(out, x)
    T = scalar(eltype(x))
    c = context(); enter()
    check(ccall((:bb_face_lift_apply, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Layout, Ptr{Cvoid}, Layout),
                c.ptr, Cint(L.N), Cint(F), K, dtype(T), dtype(T), lift_mode(T), devptr(x), lx, devptr(out), lo), c)
    leave(c)
    out
end

# ElevationMatrix{N} (src/ElevationMatrix.jl:11-24) is an AbstractMatrix; its product with a
# device vector or matrix runs elevation_multiply! (src/BernsteinLift.jl:100-125) on the GPU
function LinearAlgebra.mul!(out::ROCArray, ::ElevationMatrix{N}, x::ROCArray) where {N}
    lx, lo, K, F = same_columns(out, x)
    c = context(); enter()
    check(ccall((:bb_elevate_apply, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Cint, Ptr{Cvoid}, Layout, Ptr{Cvoid}, Layout),
                c.ptr, Cint(N), Cint(F), K, dtype(scalar(eltype(x))), devptr(x), lx, devptr(out), lo), c)
    leave(c)
    out
end

function Base.:*(E::ElevationMatrix{N}, x::ROCArray{T}) where {N,T}
    nfp = div((N + 1) * (N + 2), 2)
    out = ndims(x) == 1 ? AMDGPU.zeros(T, nfp) : AMDGPU.zeros(T, nfp, size(x, 2))
    mul!(out, E, x)
end

# ------------------------------------------------------------------------- RHS pieces

"""`face_gather!(uM, N, u)`: uM .= view(u, Fmask, :) (examples/3d_wave.jl:41), faces in
the order the lift expects."""
function face_gather!(uM::ROCArray, N::Integer, u::ROCArray)
    lq, Kq, F = layout(u); lm, Km, Fm = layout(uM)
    Kq == Km && F == Fm || throw(DimensionMismatch("u and uM differ in columns or fields"))
    c = context(); enter()
    check(ccall((:bb_face_gather, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Cint, Ptr{Cvoid}, Layout, Ptr{Cvoid}, Layout),
                c.ptr, Cint(N), Cint(F), Kq, dtype(scalar(eltype(u))), devptr(u), lq, devptr(uM), lm), c)
    leave(c)
    uM
end

"""`interface_flux!(flux, pde, N, uM, mapP, nxJ, nyJ, nzJ; bc, bc_sign, kind)`: the flux loop
of examples/3d_wave.jl:43-56 (pde = :wave) or examples/3d_maxwell.jl:43-65 (:maxwell), FP64.
mapP holds 0-based indices into the flattened face-node array."""
function interface_flux!(flux::ROCArray, pde::Symbol, N::Integer, uM::ROCArray, mapP::ROCArray{Int32},
                         nxJ::ROCArray{Float64}, nyJ::ROCArray{Float64}, nzJ::ROCArray{Float64};
                         bc::Union{Nothing,ROCArray{Int8}} = nothing, bc_sign::Vector{Float64} = Float64[],
                         kind::Symbol = :reference)
    lm, K, F = layout(uM); lf, = layout(flux)
    c = context(); enter()
    check(ccall((:bb_interface_flux, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Int64, Ptr{Cdouble}, Layout, Ptr{Int32}, Ptr{Int8}, Ptr{Cdouble},
                 Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Layout),
                c.ptr, pde === :wave ? BB_PDE_WAVE : BB_PDE_MAXWELL,
                kind === :reference ? BB_FLUX_REFERENCE : BB_FLUX_UPWIND, Cint(N), K,
                reinterpret(Ptr{Cdouble}, pointer(uM)), lm, reinterpret(Ptr{Int32}, pointer(mapP)),
                bc === nothing ? Ptr{Int8}(C_NULL) : reinterpret(Ptr{Int8}, pointer(bc)),
                isempty(bc_sign) ? Ptr{Cdouble}(C_NULL) : pointer(bc_sign),
                reinterpret(Ptr{Cdouble}, pointer(nxJ)), reinterpret(Ptr{Cdouble}, pointer(nyJ)),
                reinterpret(Ptr{Cdouble}, pointer(nzJ)), reinterpret(Ptr{Cdouble}, pointer(flux)), lf), c)
    leave(c)
    flux
end

# ------------------------------------------------------------------------- solver
# The fused path: mesh and state resident on the device, one kernel per LSRK45 stage.

"""`BBSolver(pde, N, geo, J, nJ, nbr, nbr_face, perm, bc, bc_sign; flux = :upwind)`

Host arrays (element-major): geo 9 x K (rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ), J (K),
nJ 3 x 4 x K, nbr 4 x K (0-based neighbour), nbr_face 4 x K (0..3), perm 4 x K (0..5),
bc 4 x K (0 interior, 1 wall), bc_sign (F).  State: Np x F x K Float64."""
mutable struct BBSolver
    ptr::Ptr{Cvoid}
    ctx::BBContext
    N::Int
    F::Int
    K::Int
end

function BBSolver(pde::Symbol, N::Integer, geo::Matrix{Float64}, J::Vector{Float64}, nJ::Array{Float64,3},
                  nbr::Matrix{Int32}, nbr_face::Matrix{Int8}, perm::Matrix{Int8}, bc::Matrix{Int8},
                  bc_sign::Vector{Float64}; flux::Symbol = :upwind, ctx::BBContext = context())
    K = length(J)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bb_solver_create, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Int32},
                 Ptr{Int8}, Ptr{Int8}, Ptr{Int8}, Ptr{Cdouble}, Ref{Ptr{Cvoid}}),
                ctx.ptr, pde === :wave ? BB_PDE_WAVE : BB_PDE_MAXWELL,
                flux === :reference ? BB_FLUX_REFERENCE : BB_FLUX_UPWIND, Cint(N), Int64(K),
                geo, J, nJ, nbr, nbr_face, perm, bc, bc_sign, r), ctx)
    s = BBSolver(r[], ctx, N, pde === :wave ? 4 : 6, K)
    finalizer(s) do s
        s.ptr == C_NULL || ccall((:bb_solver_destroy, LIB), Cint, (Ptr{Cvoid},), s.ptr)
        s.ptr = C_NULL
    end
    s
end

path_code(p::Symbol) = p === :plain ? BB_PATH_PLAIN : BB_PATH_FUSED
prec_code(p::Symbol) = p === :mixed ? BB_PREC_MIXED : BB_PREC_F64

set_state!(s::BBSolver, q::Array{Float64}) =
    check(ccall((:bb_solver_set_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.ptr, q), s.ctx)
function get_state(s::BBSolver)
    np = div((s.N + 1) * (s.N + 2) * (s.N + 3), 6)
    q = zeros(Float64, np, s.F, s.K)
    check(ccall((:bb_solver_get_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.ptr, q), s.ctx)
    q
end
function state_ptr(s::BBSolver)
    p = Ref{Ptr{Cdouble}}(C_NULL)
    check(ccall((:bb_solver_state_ptr, LIB), Cint, (Ptr{Cvoid}, Ref{Ptr{Cdouble}}), s.ptr, p), s.ctx)
    p[]
end

"""`check_state(s)`: the number of NaN / Inf values in the solver's device state."""
function check_state(s::BBSolver)
    cnt = Ref{Int64}(0)
    rc = ccall((:bb_solver_check_state, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), s.ptr, cnt)
    rc == BB_OK || rc == BB_ERR_NONFINITE || check(rc, s.ctx)
    Int(cnt[])
end

"""`rhs!(du, s; path, prec)`: du = rhs_matvec!(q) of the current state (host array)."""
rhs!(du::Array{Float64}, s::BBSolver; path::Symbol = :fused, prec::Symbol = :f64) =
    (check(ccall((:bb_solver_rhs, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}), s.ptr, path_code(path),
                 prec_code(prec), du), s.ctx); du)
step!(s::BBSolver, dt::Real, nsteps::Integer; path::Symbol = :fused, prec::Symbol = :f64) =
    check(ccall((:bb_solver_step, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cint), s.ptr, path_code(path),
                prec_code(prec), Float64(dt), Cint(nsteps)), s.ctx)
capture!(s::BBSolver, dt::Real, nsteps::Integer; path::Symbol = :fused, prec::Symbol = :f64) =
    check(ccall((:bb_solver_graph_capture, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cint), s.ptr,
                path_code(path), prec_code(prec), Float64(dt), Cint(nsteps)), s.ctx)
launch!(s::BBSolver) = check(ccall((:bb_solver_graph_launch, LIB), Cint, (Ptr{Cvoid},), s.ptr), s.ctx)
function autotune!(s::BBSolver; prec::Symbol = :mixed)
    t = Ref{Cint}(0)
    check(ccall((:bb_solver_autotune, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{Cint}), s.ptr, prec_code(prec), t), s.ctx)
    Int(t[])
end
set_tuning!(s::BBSolver, tiling::Integer; prec::Symbol = :mixed) =
    check(ccall((:bb_solver_set_tuning, LIB), Cint, (Ptr{Cvoid}, Cint, Cint), s.ptr, prec_code(prec), Cint(tiling)),
          s.ctx)
function stage_bytes(s::BBSolver; prec::Symbol = :mixed)
    b = Ref{Cdouble}(0); d = Ref{Cdouble}(0)
    check(ccall((:bb_solver_stage_bytes, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{Cdouble}, Ref{Cdouble}), s.ptr,
                prec_code(prec), b, d), s.ctx)
    (bytes = b[], dofs = d[])
end

"""`step_noalloc!(s, q, dt, nsteps; prec)`: the fused steps of `step!` on the caller's device
state q (Np x F x K Float64, advanced in place), with the scratch in the library's fixed
per-GPU array: no device memory is allocated (for code under graph capture or in serving
loops).  Calls that overlap on one GPU race; sizes beyond the scratch are an error."""
step_noalloc!(s::BBSolver, q::ROCArray{Float64}, dt::Real, nsteps::Integer; prec::Symbol = :mixed) =
    check(ccall((:bb_solver_step_noalloc, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cdouble, Cint), s.ptr,
                prec_code(prec), reinterpret(Ptr{Cdouble}, pointer(q)), Float64(dt), Cint(nsteps)), s.ctx)

# ------------------------------------------------------------------------- dense nodal comparator
# extras/3d_nodal_wave.jl:57-73: dense products, flux directions and the combine

function gemm!(C::ROCMatrix{T}, A::ROCMatrix{T}, B::ROCMatrix{T}, alpha::Real = 1, beta::Real = 0) where {T<:DevScalar}
    c = context(); enter()
    check(ccall((:bb_gemm, LIB), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Cdouble, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Cdouble,
                 Ptr{Cvoid}, Int64),
                c.ptr, dtype(T), Int64(size(A, 1)), Int64(size(B, 2)), Int64(size(A, 2)), Float64(alpha), devptr(A),
                Int64(stride(A, 2)), devptr(B), Int64(stride(B, 2)), Float64(beta), devptr(C), Int64(stride(C, 2))), c)
    leave(c)
    C
end

function flux_dirs!(fx::ROCArray{Float64}, fy::ROCArray{Float64}, fz::ROCArray{Float64}, pde::Symbol, N::Integer,
                    K::Integer, q::ROCArray{Float64})
    c = context(); enter()
    check(ccall((:bb_pde_flux_dirs, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                c.ptr, pde === :wave ? BB_PDE_WAVE : BB_PDE_MAXWELL, Cint(N), Int64(K), pointer(q), pointer(fx),
                pointer(fy), pointer(fz)), c)
    leave(c)
end

function rhs_combine!(du::ROCArray{Float64}, pde::Symbol, N::Integer, K::Integer, d::NTuple{9,ROCArray{Float64}},
                      geo::ROCArray{Float64}, J::ROCArray{Float64})
    ptrs = Ptr{Cdouble}[pointer(a) for a in d]
    c = context(); enter()
    check(ccall((:bb_rhs_combine, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Cdouble}, Ptr{Ptr{Cdouble}}, Ptr{Cdouble}, Ptr{Cdouble}),
                c.ptr, pde === :wave ? BB_PDE_WAVE : BB_PDE_MAXWELL, Cint(N), Int64(K), pointer(du), ptrs,
                pointer(geo), pointer(J)), c)
    leave(c)
    du
end

# ------------------------------------------------------------------------- batches of mixed degree

"""Many meshes, each element with its own degree (wave system).  Arrays as for BBSolver,
int32; dt per element; transfer tables as packed by the Python host package."""
mutable struct BBBatch
    ptr::Ptr{Cvoid}
    ctx::BBContext
end

function BBBatch(K::Integer, deg::Vector{Int32}, geo::Matrix{Float64}, J::Vector{Float64}, nJ::Array{Float64,3},
                 nbr::Matrix{Int32}, nbr_face::Matrix{Int32}, perm::Matrix{Int32}, bc::Matrix{Int32},
                 bc_sign::Vector{Float64}, dt::Vector{Float64}, transfer::Vector{Float64},
                 transfer_off::Matrix{Int64}; flux::Symbol = :upwind, ctx::BBContext = context())
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bb_batch_create, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Int32}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Int32},
                 Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Int64}, Int64,
                 Ref{Ptr{Cvoid}}),
                ctx.ptr, BB_PDE_WAVE, flux === :reference ? BB_FLUX_REFERENCE : BB_FLUX_UPWIND, Int64(K), deg, geo,
                J, nJ, nbr, nbr_face, perm, bc, bc_sign, dt, transfer, transfer_off, Int64(length(transfer)), r), ctx)
    b = BBBatch(r[], ctx)
    finalizer(b) do b
        b.ptr == C_NULL || ccall((:bb_batch_destroy, LIB), Cint, (Ptr{Cvoid},), b.ptr)
        b.ptr = C_NULL
    end
    b
end

function state_words(b::BBBatch)
    w = Ref{Int64}(0)
    check(ccall((:bb_batch_state_words, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), b.ptr, w), b.ctx)
    Int(w[])
end
set_state!(b::BBBatch, q::Vector{Float64}) =
    check(ccall((:bb_batch_set_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), b.ptr, q), b.ctx)
function get_state(b::BBBatch)
    q = zeros(Float64, state_words(b))
    check(ccall((:bb_batch_get_state, LIB), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), b.ptr, q), b.ctx)
    q
end
function check_state(b::BBBatch)
    cnt = Ref{Int64}(0)
    rc = ccall((:bb_batch_check_state, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}), b.ptr, cnt)
    rc == BB_OK || rc == BB_ERR_NONFINITE || check(rc, b.ctx)
    Int(cnt[])
end
rhs!(du::Vector{Float64}, b::BBBatch; prec::Symbol = :f64) =
    (check(ccall((:bb_batch_rhs, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}), b.ptr, prec_code(prec), du), b.ctx); du)
step!(b::BBBatch, dt_scale::Real, nsteps::Integer; prec::Symbol = :mixed) =
    check(ccall((:bb_batch_step, LIB), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint), b.ptr, prec_code(prec),
                Float64(dt_scale), Cint(nsteps)), b.ctx)
capture!(b::BBBatch, dt_scale::Real, nsteps::Integer; prec::Symbol = :mixed) =
    check(ccall((:bb_batch_graph_capture, LIB), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cint), b.ptr, prec_code(prec),
                Float64(dt_scale), Cint(nsteps)), b.ctx)
launch!(b::BBBatch) = check(ccall((:bb_batch_graph_launch, LIB), Cint, (Ptr{Cvoid},), b.ptr), b.ctx)
function autotune!(b::BBBatch; prec::Symbol = :mixed)
    t = zeros(Cint, 11)
    check(ccall((:bb_batch_autotune, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}), b.ptr, prec_code(prec), t), b.ctx)
    t
end
set_tuning!(b::BBBatch, N::Integer, tiling::Integer; prec::Symbol = :mixed) =
    check(ccall((:bb_batch_set_tuning, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Cint), b.ptr, prec_code(prec), Cint(N),
                Cint(tiling)), b.ctx)

end # module
