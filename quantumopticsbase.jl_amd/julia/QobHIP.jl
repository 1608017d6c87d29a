"""
    QobHIP

Julia binding of libqob_hip.so: device-resident `Ket`/`Bra`/`Operator`
(data in AMDGPU.jl `ROCArray`s) get `mul!` methods that call the C ABI
declared in include/qob_hip.h, one `ccall` per `mul!`.  Operators
(`SparseOperator`, `DenseOperator` factors, `LazyTensor`, `LazySum`,
`LazyProduct`, `TimeDependentSum`) are uploaded once and cached by object
identity.

Every `mul!` method here specialises exactly one method of the reference
(listed beside it): the operator argument has the reference's own type and
basis parameters, and the state and result arguments are narrowed to device
storage, so each method is strictly more specific than the reference method
it replaces and no call is ambiguous.  State and result may have different
element types (ComplexF64 / ComplexF32); the product then runs in the
promoted type, as the reference allows (test/test_operators_lazytensor.jl:
239-253).

This file is written for QuantumOpticsBase.jl v0.5 with AMDGPU.jl; it is the
binding a maintainer adds as a package extension (see INTEGRATION.md).  It
cannot be executed in this repository's CI (no Julia toolchain); the same
entry points are exercised through ctypes by tests/, and tests/test_julia_binding.py
checks the method table statically.
"""
module QobHIP

using LinearAlgebra, SparseArrays
using AMDGPU
using QuantumOpticsBase
using QuantumOpticsBase: LazyTensor, LazySum, LazyProduct, TimeDependentSum, SparseOpPureType, SparseOpAdjType,
                         DataOperator, EyeOpType, Basis, static_operator
import LinearAlgebra: mul!
import QuantumOpticsBase: expect, variance

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
qob_dtype(x::AbstractArray) = qob_dtype(eltype(x))

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
const DCOEF = IdDict{Any,ROCVector{ComplexF64}}()

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

# dense factor (host Matrix) of a LazyTensor; a ComplexF32 factor is widened
# exactly (the element type is promoted, src/operators_lazytensor.jl:32-33)
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
    push_factors!(op, h)
    h
end

# Coefficients rewritten in place since the last call (update_static_coefficients!,
# src/time_dependent_operator.jl:286-297): sent from device memory, stream-
# ordered on the task's stream (qob_lazysum_set_factors_d: the plans' values
# are refreshed on the device, nothing is rebuilt and nothing waits); sums the
# device form does not cover (dense terms) take the host call.
function push_factors!(op::LazySum, h::Ptr{Cvoid})
    SUM_FACTORS[op] == op.factors && return nothing
    d = get!(() -> ROCVector{ComplexF64}(undef, length(op.factors)), DCOEF, op)
    copyto!(d, ComplexF64.(op.factors))
    st = ccall((:qob_lazysum_set_factors_d, libqob[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Cvoid}),
               context(), h, pointer(d), stream_ptr())
    if st != 0
        fac = QobC64[QobC64(f) for f in op.factors]
        check(ccall((:qob_lazysum_set_factors, libqob[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{QobC64}),
                    context(), h, fac))
    end
    SUM_FACTORS[op] = ComplexF64.(op.factors)
    nothing
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

# TimeDependentSum (src/time_dependent_operator.jl:274-277): the static LazySum,
# whose coefficients set_time! has rewritten
handle(op::TimeDependentSum) = handle(static_operator(op))

# ------------------------------------------------------------ device calls
const DevVec{T} = ROCVector{T}
const DevMat{T} = ROCMatrix{T}
const CT = Union{ComplexF64,ComplexF32}

# result = alpha * op * b + beta * result (Ket), element types may differ
function _ket!(result, h::Ptr{Cvoid}, b, alpha, beta)
    check(ccall((:qob_mul_ket_x, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(b.data), h, pointer(b.data), length(b.data), qob_dtype(result.data),
                pointer(result.data), length(result.data), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# result = alpha * a * op + beta * result (Bra)
function _bra!(result, a, h::Ptr{Cvoid}, alpha, beta)
    check(ccall((:qob_mul_bra_x, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, QobC64, QobC64, Ptr{Cvoid}),
                context(), qob_dtype(a.data), pointer(a.data), length(a.data), h, qob_dtype(result.data),
                pointer(result.data), length(result.data), QobC64(alpha), QobC64(beta), stream_ptr()))
    result
end

# result = alpha * op * b + beta * result (density matrix, left)
function _left!(result, h::Ptr{Cvoid}, b, alpha, beta)
    check(ccall((:qob_mul_op_left_x, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Ptr{Cvoid}, Int64, QobC64, QobC64,
                 Ptr{Cvoid}),
                context(), qob_dtype(b.data), h, pointer(b.data), size(b.data, 1), size(b.data, 2),
                qob_dtype(result.data), pointer(result.data), size(result.data, 1), QobC64(alpha), QobC64(beta),
                stream_ptr()))
    result
end

# result = alpha * a * op + beta * result (density matrix, right)
function _right!(result, a, h::Ptr{Cvoid}, alpha, beta)
    check(ccall((:qob_mul_op_right_x, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, QobC64, QobC64,
                 Ptr{Cvoid}),
                context(), qob_dtype(a.data), pointer(a.data), size(a.data, 1), size(a.data, 2), h,
                qob_dtype(result.data), pointer(result.data), size(result.data, 2), QobC64(alpha), QobC64(beta),
                stream_ptr()))
    result
end

# C = alpha * op(A) * B + beta * C on device arrays of one element type
# (op(A) = A, or transpose(A) for the Bra form)
function _gemm!(C::AbstractArray{T}, transa::Bool, A::AbstractArray{T}, B::AbstractArray{T}, alpha, beta) where {T<:CT}
    m = transa ? size(A, 2) : size(A, 1)
    k = transa ? size(A, 1) : size(A, 2)
    n = length(B) ÷ max(k, 1)
    check(ccall((:qob_gemm, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, QobC64, Ptr{T}, Ptr{T}, QobC64, Ptr{T}, Ptr{Cvoid}),
                context(), qob_dtype(T), Cint(transa), m, n, k, QobC64(alpha), pointer(A), pointer(B),
                QobC64(beta), pointer(C), stream_ptr()))
    C
end

# the same with differing element types: the product in ComplexF64 (promoted),
# then rounded into the result once
function _gemm_promoted!(C, transa::Bool, A, B, alpha, beta)
    if eltype(C) == eltype(A) == eltype(B)
        return _gemm!(C, transa, A, B, alpha, beta)
    end
    C64 = ROCArray{ComplexF64}(C)
    _gemm!(C64, transa, ROCArray{ComplexF64}(A), ROCArray{ComplexF64}(B), alpha, beta)
    copyto!(C, C64)
    C
end

# ------------------------------------------------------------------- mul!
# LazySum — src/operators_lazysum.jl:189, 202, 215, 227
mul!(result::Ket{B1,<:DevVec{<:CT}}, a::LazySum{B1,B2}, b::Ket{B2,<:DevVec{<:CT}}, alpha, beta) where {B1,B2} =
    _ket!(result, handle(a), b, alpha, beta)
mul!(result::Bra{B2,<:DevVec{<:CT}}, a::Bra{B1,<:DevVec{<:CT}}, b::LazySum{B1,B2}, alpha, beta) where {B1,B2} =
    _bra!(result, a, handle(b), alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::LazySum{B1,B2}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = _left!(result, handle(a), b, alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::LazySum{B2,B3},
     alpha, beta) where {B1,B2,B3} = _right!(result, a, handle(b), alpha, beta)

# LazyTensor — src/operators_lazytensor.jl:539, 559, 576, 593
mul!(result::Ket{B1,<:DevVec{<:CT}}, a::LazyTensor{B1,B2,F,I,TT}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1<:Basis,B2<:Basis,F,I,TT<:Tuple{Vararg{DataOperator}}} = _ket!(result, handle(a), b, alpha, beta)
mul!(result::Bra{B2,<:DevVec{<:CT}}, a::Bra{B1,<:DevVec{<:CT}}, b::LazyTensor{B1,B2,F,I,TT}, alpha,
     beta) where {B1<:Basis,B2<:Basis,F,I,TT<:Tuple{Vararg{DataOperator}}} = _bra!(result, a, handle(b), alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::LazyTensor{B1,B2,F,I,TT}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1<:Basis,B2<:Basis,B3<:Basis,F,I,TT<:Tuple{Vararg{DataOperator}}} =
    _left!(result, handle(a), b, alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::LazyTensor{B2,B3,F,I,TT},
     alpha, beta) where {B1<:Basis,B2<:Basis,B3<:Basis,F,I,TT<:Tuple{Vararg{DataOperator}}} =
    _right!(result, a, handle(b), alpha, beta)

# SparseOperator — src/operators_sparse.jl:199-202 (the density-matrix forms
# once for the plain and once for the adjoint CSC, the two members of
# SparseOpType); dagger(sparse) on a Ket or Bra — src/operators_dense.jl:395-396
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, M::SparseOpPureType{B1,B2}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = _left!(result, handle(M), b, alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, M::SparseOpAdjType{B1,B2}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = _left!(result, handle(M), b, alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, M::SparseOpPureType{B2,B3},
     alpha, beta) where {B1,B2,B3} = _right!(result, a, handle(M), alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, M::SparseOpAdjType{B2,B3},
     alpha, beta) where {B1,B2,B3} = _right!(result, a, handle(M), alpha, beta)
mul!(result::Ket{B1,<:DevVec{<:CT}}, M::SparseOpPureType{B1,B2}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1,B2} = _ket!(result, handle(M), b, alpha, beta)
mul!(result::Bra{B2,<:DevVec{<:CT}}, b::Bra{B1,<:DevVec{<:CT}}, M::SparseOpPureType{B1,B2}, alpha,
     beta) where {B1,B2} = _bra!(result, b, handle(M), alpha, beta)
mul!(result::Ket{B1,<:DevVec{<:CT}}, M::SparseOpAdjType{B1,B2}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1,B2} = _ket!(result, handle(M), b, alpha, beta)
mul!(result::Bra{B2,<:DevVec{<:CT}}, b::Bra{B1,<:DevVec{<:CT}}, M::SparseOpAdjType{B1,B2}, alpha,
     beta) where {B1,B2} = _bra!(result, b, handle(M), alpha, beta)

# TimeDependentSum — src/time_dependent_operator.jl:274-277 (the static sum at
# the current time; its coefficients go to the device stream-ordered)
mul!(out::Operator{B1,B3,<:DevMat{<:CT}}, a::TimeDependentSum{B1,B2}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = _left!(out, handle(a), b, alpha, beta)
mul!(out::Ket{B1,<:DevVec{<:CT}}, a::TimeDependentSum{B1,B2}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1,B2} = _ket!(out, handle(a), b, alpha, beta)
mul!(out::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::TimeDependentSum{B2,B3},
     alpha, beta) where {B1,B2,B3} = _right!(out, a, handle(b), alpha, beta)
mul!(out::Bra{B2,<:DevVec{<:CT}}, a::Bra{B1,<:DevVec{<:CT}}, b::TimeDependentSum{B1,B2}, alpha,
     beta) where {B1,B2} = _bra!(out, a, handle(b), alpha, beta)

# LazyProduct — src/operators_lazyproduct.jl:103, 117, 131, 148
mul!(result::Ket{B1,<:DevVec{<:CT}}, a::LazyProduct{B1,B2}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1,B2} = _ket!(result, handle(a), b, alpha, beta)
mul!(result::Bra{B2,<:DevVec{<:CT}}, a::Bra{B1,<:DevVec{<:CT}}, b::LazyProduct{B1,B2}, alpha,
     beta) where {B1,B2} = _bra!(result, a, handle(b), alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::LazyProduct{B1,B2}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = _left!(result, handle(a), b, alpha, beta)
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::LazyProduct{B2,B3},
     alpha, beta) where {B1,B2,B3} = _right!(result, a, handle(b), alpha, beta)

# DenseOperator on device data — src/operators_dense.jl:394, 395, 396
mul!(result::Operator{B1,B3,<:DevMat{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::Operator{B2,B3,<:DevMat{<:CT}},
     alpha, beta) where {B1,B2,B3} = (_gemm_promoted!(result.data, false, a.data, b.data, alpha, beta); result)
mul!(result::Ket{B1,<:DevVec{<:CT}}, a::Operator{B1,B2,<:DevMat{<:CT}}, b::Ket{B2,<:DevVec{<:CT}}, alpha,
     beta) where {B1,B2} = (_gemm_promoted!(result.data, false, a.data, b.data, alpha, beta); result)
mul!(result::Bra{B2,<:DevVec{<:CT}}, a::Bra{B1,<:DevVec{<:CT}}, b::Operator{B1,B2,<:DevMat{<:CT}}, alpha,
     beta) where {B1,B2} = (_gemm_promoted!(result.data, true, b.data, a.data, alpha, beta); result)

# ---------------------------------------------------------------- reductions
# norm / normalize! (src/states.jl:89-90), tr (src/operators_dense.jl:189)
function LinearAlgebra.norm(x::Ket{B,<:DevVec{T}}) where {B,T<:CT}
    out = Ref{Float64}(0)
    check(ccall((:qob_norm, libqob[]), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{T}, Ref{Float64}, Ptr{Cvoid}),
                context(), qob_dtype(T), length(x.data), pointer(x.data), out, stream_ptr()))
    out[]
end

function LinearAlgebra.norm(x::Bra{B,<:DevVec{T}}) where {B,T<:CT}
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

function _expect_ket(h::Ptr{Cvoid}, state)
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_expect_ket, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(state.data), h, pointer(state.data), length(state.data), out, stream_ptr()))
    complex(out[].re, out[].im)
end

function _expect_op(h::Ptr{Cvoid}, state)
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_expect_op, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(state.data), h, pointer(state.data), size(state.data, 1), out, stream_ptr()))
    complex(out[].re, out[].im)
end

function _variance_ket(h::Ptr{Cvoid}, state)
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_variance_ket, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(state.data), h, pointer(state.data), length(state.data), out, stream_ptr()))
    complex(out[].re, out[].im)
end

# variance(op, rho): op * rho formed once into the library's workspace
function _variance_op(h::Ptr{Cvoid}, state)
    out = Ref{QobC64}(QobC64(0, 0))
    check(ccall((:qob_variance_op, libqob[]), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ref{QobC64}, Ptr{Cvoid}),
                context(), qob_dtype(state.data), h, pointer(state.data), size(state.data, 1), out, stream_ptr()))
    complex(out[].re, out[].im)
end

# expect(op, ket) — src/operators.jl:119, src/operators_dense.jl:219;
# expect(op, rho) — src/operators_dense.jl:223, src/operators_sparse.jl:36;
# variance(op, ket) — src/operators.jl:139; variance(op, rho) — the operator
# form of src/operators.jl:132-137.  One method per operator type.
expect(op::LazySum{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::LazyTensor{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::LazyProduct{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::TimeDependentSum{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::SparseOpPureType{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::SparseOpAdjType{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _expect_ket(handle(op), state)
expect(op::LazySum{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} = _expect_op(handle(op), state)
expect(op::LazyTensor{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} = _expect_op(handle(op), state)
expect(op::LazyProduct{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} = _expect_op(handle(op), state)
expect(op::TimeDependentSum{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} =
    _expect_op(handle(op), state)
expect(op::SparseOpPureType{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} =
    _expect_op(handle(op), state)
expect(op::SparseOpAdjType{B1,B2}, state::Operator{B2,B2,<:DevMat{<:CT}}) where {B1,B2} =
    _expect_op(handle(op), state)
variance(op::LazySum{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::LazyTensor{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::LazyProduct{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::TimeDependentSum{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::SparseOpPureType{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::SparseOpAdjType{B,B}, state::Ket{B,<:DevVec{<:CT}}) where {B} = _variance_ket(handle(op), state)
variance(op::LazySum{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} = _variance_op(handle(op), state)
variance(op::LazyTensor{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} = _variance_op(handle(op), state)
variance(op::LazyProduct{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} = _variance_op(handle(op), state)
variance(op::TimeDependentSum{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} =
    _variance_op(handle(op), state)
variance(op::SparseOpPureType{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} =
    _variance_op(handle(op), state)
variance(op::SparseOpAdjType{B,B}, state::Operator{B,B,<:DevMat{<:CT}}) where {B} =
    _variance_op(handle(op), state)

end # module
