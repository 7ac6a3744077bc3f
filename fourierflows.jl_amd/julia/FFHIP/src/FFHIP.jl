# FFHIP.jl -- Julia binding of libffhip.so (the C ABI in include/ffhip.h).
#
# What a FourierFlows.jl maintainer adds to use the MI355X path: a device
# tag `HIP()` beside `CPU()`/`GPU()`, a device array type whose memory libffhip
# owns, and methods of the functions FourierFlows already calls
# (`mul!`/`ldiv!` on `grid.rfftplan`, `dealias!`, the ETDRK4/RK4 substeps,
# `parsevalsum2`, `stepforward!`) that forward to `ccall`.  Every call
# checks the returned status and raises with `ffhip_last_error()`.
#
# Reference interfaces replaced (FourierFlows.jl v0.10.7):
#   plan_flows_rfft / mul! / ldiv! ......... src/domains.jl:2-5, 206-208, 347-349
#   getaliasedwavenumbers / dealias! ....... src/domains.jl:408-421, 428-476
#   ETDRK4TimeStepper, substeps, update .... src/timesteppers.jl:434-463, 501-547
#   RK4TimeStepper, substeps, update ....... src/timesteppers.jl:180-197, 225-274
#   parsevalsum2 / parsevalsum ............. src/utils.jl:113-183
#   stepforward!(prob, n) .................. src/timesteppers.jl:6-20
#   zeros(dev, T, dims), device_array ...... src/utils.jl:79-80, 329-332
module FFHIP

using LinearAlgebra

const libffhip = get(ENV, "FFHIP_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libffhip.so"))

const OK = Cint(0)
const F32 = Cint(32)
const F64 = Cint(64)
const OP_VORTICITY2D = Cint(1)
const OP_TRACER3D = Cint(2)

struct FFHIPError <: Exception
    code::Cint
    msg::String
end

function check(rc::Cint)
    rc == OK && return nothing
    msg = unsafe_string(ccall((:ffhip_last_error, libffhip), Cstring, ()))
    rc == Cint(-1) && throw(DomainError(msg))
    throw(FFHIPError(rc, msg))
end

prec(::Type{Float32}) = F32
prec(::Type{Float64}) = F64
prec(::Type{ComplexF32}) = F32
prec(::Type{ComplexF64}) = F64

"Device tag (the analogue of `GPU()` for this library)."
struct HIP
    index::Int
end
HIP() = HIP(0)

init(dev::HIP=HIP()) = check(ccall((:ffhip_init, libffhip), Cint, (Cint,), dev.index))

# ----------------------------------------------------------------------------- device arrays
"A dense device array in Julia's column-major layout; memory owned by libffhip."
mutable struct HIPArray{T,N} <: DenseArray{T,N}
    ptr::Ptr{Cvoid}
    dims::NTuple{N,Int}
    function HIPArray{T,N}(dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ffhip_malloc, libffhip), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), p, prod(dims) * sizeof(T)))
        a = new{T,N}(p[], dims)
        finalizer(x -> ccall((:ffhip_free, libffhip), Cint, (Ptr{Cvoid},), x.ptr), a)
        return a
    end
end
Base.size(a::HIPArray) = a.dims
Base.pointer(a::HIPArray) = a.ptr
Base.unsafe_convert(::Type{Ptr{Cvoid}}, a::HIPArray) = a.ptr

"`zeros(dev, T, dims)` (src/utils.jl:79-80): ffhip_malloc zero-fills."
Base.zeros(::HIP, T, dims) = HIPArray{T,length(dims)}(Tuple(dims))

function device_array(::HIP)
    return (x::AbstractArray{T,N}) where {T,N} -> begin
        a = HIPArray{T,N}(size(x))
        xc = Array(x)
        check(ccall((:ffhip_memcpy_htod, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), a, xc,
                    sizeof(xc)))
        a
    end
end

Base.similar(a::HIPArray{T,N}) where {T,N} = HIPArray{T,N}(a.dims)

function Base.copyto!(dst::HIPArray{T,N}, src::HIPArray{T,N}; stream=C_NULL) where {T,N}
    dst.dims == src.dims || throw(DimensionMismatch("copyto!: $(dst.dims) vs $(src.dims)"))
    check(ccall((:ffhip_memcpy_dtod, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                dst, src, prod(src.dims) * sizeof(T), stream))
    return dst
end
Base.copy(a::HIPArray) = copyto!(similar(a), a)

function Base.Array(a::HIPArray{T,N}) where {T,N}
    x = Array{T,N}(undef, a.dims)
    check(ccall((:ffhip_memcpy_dtoh, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), x, a, sizeof(x)))
    return x
end

# ----------------------------------------------------------------------------- FFT plans
"plan_flows_rfft / plan_flows_fft (src/domains.jl:2-5)."
mutable struct Plan{T}
    h::Ptr{Cvoid}
    ns::Vector{Int64}
    rfft::Bool
end

function plan_rfft(ns::NTuple{N,Int}, ::Type{T}; batch::Integer=1) where {N,T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    n = collect(Int64, ns)
    check(ccall((:ffhip_plan_rfft, libffhip), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Int64, Cint),
                h, N, n, batch, prec(T)))
    p = Plan{T}(h[], n, true)
    finalizer(x -> ccall((:ffhip_plan_destroy, libffhip), Cint, (Ptr{Cvoid},), x.h), p)
    return p
end

"`mul!(out, rfftplan, in)`: unnormalised forward transform."
function LinearAlgebra.mul!(out::HIPArray, p::Plan, in::HIPArray; stream=C_NULL)
    check(ccall((:ffhip_plan_forward, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.h, in, out, stream))
    return out
end

"`ldiv!(out, rfftplan, in)`: inverse scaled by 1/N; unlike FFTW the input is kept."
function LinearAlgebra.ldiv!(out::HIPArray, p::Plan, in::HIPArray; stream=C_NULL)
    check(ccall((:ffhip_plan_inverse, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.h, in, out, stream))
    return out
end

# transposes (no counterpart in the reference; used by the rrules below)
mul_adjoint!(out::HIPArray, p::Plan, Y::HIPArray; stream=C_NULL) =
    (check(ccall((:ffhip_plan_forward_adjoint, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 p.h, Y, out, stream)); out)
ldiv_adjoint!(out::HIPArray, p::Plan, phi::HIPArray; stream=C_NULL) =
    (check(ccall((:ffhip_plan_inverse_adjoint, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 p.h, phi, out, stream)); out)

# ----------------------------------------------------------------------------- dealiasing
"getaliasedwavenumbers (src/domains.jl:408-421) computed by the library."
function getaliasedwavenumbers(nk::Integer, nkr::Integer, aliased_fraction::Real)
    r = [Ref{Int64}(0) for _ in 1:4]
    check(ccall((:ffhip_aliased_wavenumbers, libffhip), Cint,
                (Int64, Int64, Cdouble, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                nk, nkr, aliased_fraction, r[1], r[2], r[3], r[4]))
    aliased_fraction == 0 && return nothing, nothing
    return r[1][]:r[2][], r[3][]:r[4][]
end

"`dealias!(fh, grid)` for a 2D/3D rfft array; `ranges` are the grid's alias ranges per dimension."
function dealias!(fh::HIPArray{T,N}, ranges; stream=C_NULL) where {T,N}
    lo = Int64[first(r) for r in ranges]
    hi = Int64[last(r) for r in ranges]
    dims = collect(Int64, size(fh))
    check(ccall((:ffhip_dealias, libffhip), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}),
                fh, prec(T), N, dims, 1, lo, hi, stream))
    return nothing
end

# ----------------------------------------------------------------------------- time steppers
"ETDRK4TimeStepper (src/timesteppers.jl:434-463); coefficients FP64 on the device."
mutable struct ETDRK4TimeStepper{T}
    h::Ptr{Cvoid}
end

function ETDRK4TimeStepper(L::AbstractArray, dt::Real, ::Type{T}) where {T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    Lre = Float64.(real.(L))
    Lim = eltype(L) <: Complex ? Float64.(imag.(L)) : nothing
    dims = collect(Int64, size(L))
    check(ccall((:ffhip_etdrk4_create, libffhip), Cint,
                (Ptr{Ptr{Cvoid}}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Int64}, Cdouble, Cint),
                h, Lre, Lim === nothing ? C_NULL : Lim, ndims(L), dims, dt, prec(T)))
    ts = ETDRK4TimeStepper{T}(h[])
    finalizer(x -> ccall((:ffhip_etdrk4_destroy, libffhip), Cint, (Ptr{Cvoid},), x.h), ts)
    return ts
end

ETDRK4substep12!(ts, out, sol, N; stream=C_NULL) =
    check(ccall((:ffhip_etdrk4_substep12, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ts.h, out, sol, N, stream))
ETDRK4substep3!(ts, out, sol1, N1, N3; stream=C_NULL) =
    check(ccall((:ffhip_etdrk4_substep3, libffhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ts.h, out, sol1, N1, N3, stream))
ETDRK4update!(ts, sol, N1, N2, N3, N4; stream=C_NULL) =
    check(ccall((:ffhip_etdrk4_update, libffhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ts.h, sol, N1, N2, N3, N4, stream))

# ----------------------------------------------------------------------------- built-in calcN! and fused runs
"A built-in `calcN!` (2D vorticity or 3D ABC tracer) that libffhip runs in fused kernels."
mutable struct BuiltinOperator{T}
    h::Ptr{Cvoid}
end

function BuiltinOperator(kind::Cint, ns::NTuple{N,Int}, Ls::NTuple{N,Float64}, ::Type{T};
                         aliased_fraction=1/3, params=Float64[]) where {N,T}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:ffhip_op_create, libffhip), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Ptr{Int64}, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}, Cint),
                h, kind, prec(T), N, collect(Int64, ns), collect(Float64, Ls), aliased_fraction,
                isempty(params) ? C_NULL : params, length(params)))
    op = BuiltinOperator{T}(h[])
    finalizer(x -> ccall((:ffhip_op_destroy, libffhip), Cint, (Ptr{Cvoid},), x.h), op)
    return op
end

"The `calcN!(N, sol, t, clock, vars, params, grid)` signature."
(op::BuiltinOperator)(N, sol, t, clock, vars, params, grid; stream=C_NULL) =
    check(ccall((:ffhip_op_calcN, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                op.h, N, sol, t, stream))

"`stepforward!(prob, nsteps)` when `prob.eqn.calcN!` is a BuiltinOperator and the stepper is ETDRK4."
stepforward!(op::BuiltinOperator, ts::ETDRK4TimeStepper, sol::HIPArray, nsteps::Integer; stream=C_NULL) =
    check(ccall((:ffhip_op_etdrk4_run, libffhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                op.h, ts.h, sol, nsteps, stream))

# ----------------------------------------------------------------------------- reductions
"parsevalsum2(uh, grid) (src/utils.jl:113-127): compensated FP64 sum."
function parsevalsum2(uh::HIPArray{T,N}, nx::Integer, norm::Real; rfft::Bool=true, stream=C_NULL) where {T,N}
    out = Ref{Cdouble}(0)
    check(ccall((:ffhip_parsevalsum2, libffhip), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Cint, Int64, Cdouble, Ref{Cdouble}, Ptr{Cvoid}),
                uh, prec(T), N, collect(Int64, size(uh)), rfft, nx, norm, out, stream))
    return out[]
end

# ----------------------------------------------------------------------------- adjoint (rrule)
"Forward run keeping `ncheck` checkpoints; `solT` may be `sol0`."
etdrk4_forward_ckpt!(op, ts, solT, sol0, nsteps, ncheck; stream=C_NULL) =
    check(ccall((:ffhip_op_etdrk4_forward_ckpt, libffhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                op.h, ts.h, sol0, solT, nsteps, ncheck, stream))

"lam0 = (d solT / d sol0)^T lamT after the matching `etdrk4_forward_ckpt!`."
etdrk4_adjoint!(op, ts, lam0, lamT, nsteps, ncheck; stream=C_NULL) =
    check(ccall((:ffhip_op_etdrk4_adjoint, libffhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                op.h, ts.h, lamT, lam0, nsteps, ncheck, stream))

"Identifier of the checkpointed forward run `op` holds (a new one per `etdrk4_forward_ckpt!`)."
function checkpoint_id(op::BuiltinOperator)
    id = Ref{Int64}(0)
    check(ccall((:ffhip_op_checkpoint_id, libffhip), Cint, (Ptr{Cvoid}, Ref{Int64}), op.h, id))
    return id[]
end

"""
    solve(op, ts, sol0, nsteps, ncheck) -> solT

`nsteps` fused ETDRK4 steps from `sol0` (left unchanged), keeping `ncheck`
checkpoints so that the reverse sweep can run.  ChainRulesCore's `rrule` for
it lives in the package extension `FFHIPChainRulesCoreExt` (ext/), loaded
when ChainRulesCore is.
"""
function solve(op::BuiltinOperator, ts::ETDRK4TimeStepper, sol0::HIPArray, nsteps::Integer, ncheck::Integer)
    solT = similar(sol0)
    etdrk4_forward_ckpt!(op, ts, solT, sol0, nsteps, ncheck)
    return solT
end

end # module
