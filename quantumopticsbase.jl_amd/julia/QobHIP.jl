"""
    QobHIP

Julia binding of libqob_hip.so: device-resident `Ket`/`Bra`/`Operator`
(data in AMDGPU.jl `ROCArray`s) get `mul!` methods that call the C ABI
declared in include/qob_hip.h, one `ccall` per `mul!`.  Operators
(`SparseOperator`, `DenseOperator` factors, `LazyTensor`, `LazySum`,
`LazyProduct`, `TimeDependentSum`) are uploaded once and cached by object
identity; `LazySum` coefficient changes are pushed with
`qob_lazysum_set_factors` before the next call, and those of a
`TimeDependentSum` after `set_time!` with the stream-ordered
`qob_lazysum_set_factors_d`.

This file is written for QuantumOpticsBase.jl v0.5 with AMDGPU.jl; it is the
binding a maintainer adds as a package extension (see INTEGRATION.md).  It
cannot be executed in this repository's CI (no Julia toolchain); the same
entry points are exercised through ctypes by tests/.
"""
module QobHIP

using LinearAlgebra, SparseArrays
using AMDGPU
using QuantumOpticsBase
using QuantumOpticsBase: LazyTensor, LazySum, LazyProduct, TimeDependentSum, SparseOpPureType, SparseOpAdjType,
                         DataOperator, EyeOpType, static_operator
import LinearAlgebra: mul!

const libqob = Ref{String}(joinpath(@__DIR__, "..", "lib", "libqob_hip.so"))

# ------------------------------------------------------------------ ABI types
struct QobC64
    re::Float64
    im::Float64
end
QobC64(z::Number) = QobC64(Float64(real(z)), Float64(imag(z)))

const QOB_C64 = Cint(0)
const QOB_C32 = Cint(1)
qob_dtype(::Type{ComplexF64}) = QOB_C64
qob_dtype(::Type{ComplexF32}) = QOB_C32

# status -> the exception the reference throws at the same place
function check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:qob_last_error, libqob[]), Cstring, ()))
    st == 1 && throw(DimensionMismatch(msg))
    st == 2 && throw(ArgumentError(msg))
    st == 3 && throw(QuantumOpticsBase.IncompatibleBases())
    st == 4 && throw(ArgumentError(msg))
    st == 5 && throw(OutOfMemoryError())
    error("libqob_hip: " * msg)
end

# ------------------------------------------------------------------ context
const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_ctx_create, libqob[]), Cint, (Cint, Ref{Ptr{Cvoid}}), AMDGPU.device_id() - 1, h))
        CTX[] = h[]
    end
    CTX[]
end
stream_ptr() = AMDGPU.stream().stream  # hipStream_t of the task-local stream

# ------------------------------------------------------------ operator handles
# Handles are cached per operator object; the library reference-counts the
# factors a LazyTensor/LazySum holds.
const HANDLES = IdDict{Any,Ptr{Cvoid}}()
const SUM_FACTORS = IdDict{Any,Vector{ComplexF64}}()

function release_handle(op)
    h = pop!(HANDLES, op, C_NULL)
    h == C_NULL || ccall((:qob_op_release, libqob[]), Cint, (Ptr{Cvoid},), h)
    nothing
end

function cache!(f, op)
    get!(HANDLES, op) do
        h = f()
        finalizer(release_handle, op)
        h
    end
end

# SparseOperator: SparseMatrixCSC handed over as Julia holds it (1-based Int64)
function handle(op::SparseOpPureType)
    cache!(op) do
        M = op.data
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_sparse_create, libqob[]), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}, Ref{Ptr{Cvoid}}),
                    context(), size(M, 1), size(M, 2), Vector{Int64}(M.colptr), Vector{Int64}(M.rowval),
                    Vector{ComplexF64}(M.nzval), h))
        h[]
    end
end

# dagger(sparse): Operator{..,Adjoint{SparseMatrixCSC}}
function handle(op::SparseOpAdjType)
    cache!(op) do
        parent_op = Operator(op.basis_r, op.basis_l, parent(op.data))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_op_adjoint, libqob[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    context(), handle(parent_op), h))
        h[]
    end
end

# dense factor (host Matrix) of a LazyTensor
function handle(op::Operator{BL,BR,<:Matrix}) where {BL,BR}
    cache!(op) do
        A = Matrix{ComplexF64}(op.data)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_dense_create, libqob[]), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF64}, Ref{Ptr{Cvoid}}),
                    context(), size(A, 1), size(A, 2), A, h))
        h[]
    end
end

# identityoperator: Operator{..,Eye}
function handle(op::EyeOpType)
    cache!(op) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_eye_create, libqob[]), Cint, (Ptr{Cvoid}, Int64, Int64, Ref{Ptr{Cvoid}}),
                    context(), size(op.data, 1), size(op.data, 2), h))
        h[]
    end
end

_shape(b) = Int64[length(x) for x in QuantumOpticsBase._comp_size(b)]

function handle(op::LazyTensor)
    cache!(op) do
        hs = Ptr{Cvoid}[handle(o) for o in op.operators]
        h = Ref{Ptr{Cvoid}}(C_NULL)
        shl, shr = _shape(op.basis_l), _shape(op.basis_r)
        check(ccall((:qob_lazytensor_create, libqob[]), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Cint, Ptr{Int64}, Ptr{Ptr{Cvoid}}, QobC64,
                     Ref{Ptr{Cvoid}}),
                    context(), length(shl), shl, shr, length(hs), Int64[i for i in op.indices], hs,
                    QobC64(op.factor), h))
        h[]
    end
end

function handle(op::LazySum)
    h = cache!(op) do
        hs = Ptr{Cvoid}[handle(o) for o in op.operators]
        fac = QobC64[QobC64(f) for f in op.factors]
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_lazysum_create, libqob[]), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Cint, Ptr{QobC64}, Ptr{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                    context(), length(op.basis_l), length(op.basis_r), length(hs), fac, hs, r))
        SUM_FACTORS[op] = ComplexF64.(op.factors)
        r[]
    end
    # coefficients rewritten in place (TimeDependentSum set_time!)
    cached = SUM_FACTORS[op]
    if cached != op.factors
        fac = QobC64[QobC64(f) for f in op.factors]
        check(ccall((:qob_lazysum_set_factors, libqob[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{QobC64}),
                    context(), h, fac))
        SUM_FACTORS[op] = ComplexF64.(op.factors)
    end
    h
end

# LazyProduct (src/operators_lazyproduct.jl:24-58): factors applied in turn
# by the library through its own device temporaries
function handle(op::LazyProduct)
    cache!(op) do
        hs = Ptr{Cvoid}[handle(o) for o in op.operators]
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:qob_lazyproduct_create, libqob[]), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, QobC64, Ref{Ptr{Cvoid}}),
                    context(), length(hs), hs, QobC64(op.factor), h))
        h[]
    end
end

# TimeDependentSum (src/time_dependent_operator.jl:274-277): mul! on the static
# LazySum.  After set_time! the evaluated coefficients go to the device and
# the library rewrites its plans there (qob_lazysum_set_factors_d), stream-
# ordered: no plan rebuild, no device synchronisation.
const DCOEF = IdDict{Any,ROCVector{ComplexF64}}()
function handle(op::TimeDependentSum)
    ls = static_operator(op)
    h = get(HANDLES, ls, C_NULL)
    h == C_NULL && return handle(ls)
    if SUM_FACTORS[ls] != ls.factors
        d = get!(() -> ROCVector{ComplexF64}(undef, length(ls.factors)), DCOEF, ls)
        copyto!(d, ComplexF64.(ls.factors))
        st = ccall((:qob_lazysum_set_factors_d, libqob[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Cvoid}),
                   context(), h, pointer(d), stream_ptr())
        st == 0 ? (SUM_FACTORS[ls] = ComplexF64.(ls.factors)) : return handle(ls)  # host path otherwise
    end
    h
end

# ------------------------------------------------------------------- mul!
const DevVec{T} = ROCVector{T}
const DevMat{T} = ROCMatrix{T}
const DevOp = Union{LazySum,LazyTensor,LazyProduct,TimeDependentSum,SparseOpPureType,SparseOpAdjType}
const CT = Union{ComplexF64,ComplexF32}

# mul!(::Ket, op, ::Ket, α, β)  — src/operators_lazysum.jl:189,
# src/operators_lazytensor.jl:539, src/operators_sparse.jl:201
function mul!(result::Ket{B1,<:DevVec{T}}, a::DevOp, b::Ket{B2,<:DevVec{T}}, alpha, beta) where {B1,B2,T<:CT}
    check(ccall((:qob_mul_ket, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{T}, Int64, Ptr{T}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(T), handle(a), pointer(b.data), length(b.data), pointer(result.data),
                length(result.data), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# mul!(::Bra, ::Bra, op, α, β)  — src/operators_lazysum.jl:202,
# src/operators_lazytensor.jl:559, src/operators_sparse.jl:202
function mul!(result::Bra{B2,<:DevVec{T}}, a::Bra{B1,<:DevVec{T}}, b::DevOp, alpha, beta) where {B1,B2,T<:CT}
    check(ccall((:qob_mul_bra, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{T}, Int64, Ptr{Cvoid}, Ptr{T}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(T), pointer(a.data), length(a.data), handle(b), pointer(result.data),
                length(result.data), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# mul!(::Operator, op, ::Operator, α, β)  — left application on rho
function mul!(result::Operator{B1,B3,<:DevMat{T}}, a::DevOp, b::Operator{B2,B3,<:DevMat{T}},
              alpha, beta) where {B1,B2,B3,T<:CT}
    check(ccall((:qob_mul_op_left, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{T}, Int64, Int64, Ptr{T}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(T), handle(a), pointer(b.data), size(b.data, 1), size(b.data, 2),
                pointer(result.data), size(result.data, 1), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# mul!(::Operator, ::Operator, op, α, β)  — right application on rho
function mul!(result::Operator{B1,B3,<:DevMat{T}}, a::Operator{B1,B2,<:DevMat{T}}, b::DevOp,
              alpha, beta) where {B1,B2,B3,T<:CT}
    check(ccall((:qob_mul_op_right, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{T}, Int64, Int64, Ptr{Cvoid}, Ptr{T}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(T), pointer(a.data), size(a.data, 1), size(a.data, 2), handle(b),
                pointer(result.data), size(result.data, 2), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# dense x dense on device matrices — src/operators_dense.jl:394
function mul!(result::Operator{B1,B3,<:DevMat{T}}, a::Operator{B1,B2,<:DevMat{T}},
              b::Operator{B2,B3,<:DevMat{T}}, alpha, beta) where {B1,B2,B3,T<:CT}
    m, k = size(a.data)
    n = size(b.data, 2)
    check(ccall((:qob_gemm, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, QobC64, Ptr{T}, Ptr{T}, QobC64, Ptr{T}, Ptr{Cvoid}),
                context(), qob_dtype(T), Cint(0), m, n, k, QobC64(alpha), pointer(a.data), pointer(b.data),
                QobC64(beta), pointer(result.data), stream_ptr()))
    result
end

# ---------------------------------------------------------------- reductions
function LinearAlgebra.norm(x::Union{Ket{B,<:DevVec{T}},Bra{B,<:DevVec{T}}}) where {B,T<:CT}
    out = Ref{Float64}(0)
    check(ccall((:qob_norm, libqob[]), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{T}, Ref{Float64}, Ptr{Cvoid}),
                context(), qob_dtype(T), length(x.data), pointer(x.data), out, stream_ptr()))
    out[]
end

function LinearAlgebra.tr(op::Operator{B,B,<:DevMat{T}}) where {B,T<:CT}
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_trace, libqob[]), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{T}, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(T), size(op.data, 1), pointer(op.data), out, stream_ptr()))
    complex(out[].re, out[].im)
end

# expect(op, rho) = tr(op rho) without forming op rho where possible
# (src/operators_sparse.jl:36-45, src/operators_dense.jl:223-230)
function QuantumOpticsBase.expect(op::DevOp, state::Operator{B,B,<:DevMat{T}}) where {B,T<:CT}
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_expect_op, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{T}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(T), handle(op), pointer(state.data), size(state.data, 1), out, stream_ptr()))
    complex(out[].re, out[].im)
end

# variance(op, ket) (src/operators.jl:139-143)
function QuantumOpticsBase.variance(op::DevOp, state::Ket{B,<:DevVec{T}}) where {B,T<:CT}
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_variance_ket, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{T}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(T), handle(op), pointer(state.data), length(state.data), out, stream_ptr()))
    complex(out[].re, out[].im)
end

# variance(op, rho) = tr(op op rho) - tr(op rho)^2, op rho formed once on the
# device (the operator form named by the docstring at src/operators.jl:132-137)
function QuantumOpticsBase.variance(op::DevOp, state::Operator{B,B,<:DevMat{T}}) where {B,T<:CT}
    x = Operator(state.basis_l, state.basis_r, similar(state.data))
    mul!(x, op, state, true, false)
    expect(op, x) - expect(op, state)^2
end

function QuantumOpticsBase.expect(op::DevOp, state::Ket{B,<:DevVec{T}}) where {B,T<:CT}
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_expect_ket, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{T}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(T), handle(op), pointer(state.data), length(state.data), out, stream_ptr()))
    complex(out[].re, out[].im)
end

end # module
