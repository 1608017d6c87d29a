"""Host-side mirror of QuantumOpticsBase.jl's operator-application API.

Types and functions keep the reference's names and argument meaning:

    Ket, Bra                     src/states.jl:11-32
    Operator, DenseOperator      src/operators_dense.jl:12-60
    SparseOperator               src/operators_sparse.jl:5-23
    identityoperator             src/operators_sparse.jl:180-186
    LazyTensor                   src/operators_lazytensor.jl:15-58
    LazySum                      src/operators_lazysum.jl:41-60
    mul_(result, a, b, α, β)     5-argument mul! (src/operators_dense.jl:385-396)
    dagger, expect, norm, normalize_, tr, dot

State and density-matrix data live in device memory (``DeviceArray``);
every ``mul_`` is one call into libqob_hip.so.  Python's ``!``-less naming:
``mul!`` -> ``mul_``, ``normalize!`` -> ``normalize_``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from fractions import Fraction
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import (QOB_C32, QOB_C64, ArgumentError, DimensionMismatch, IncompatibleBases, MethodError, c64,
                   call, qob_c64)

# ------------------------------------------------------------------ context
_CTX = None


class Context:
    """A device context: device id, default stream, workspaces."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.qob_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def set_mode(self, mode: int) -> None:
        call("qob_ctx_set_mode", self.h, mode)

    def synchronize(self, stream=None) -> None:
        call("qob_stream_synchronize", self.h, stream)

    @property
    def stream(self):
        s = C.c_void_p()
        call("qob_ctx_get_stream", self.h, C.byref(s))
        return s


# Objects a HIP graph being captured uses: while a Graph (batch.py) captures,
# the entry points of this mirror add the operators, batches and device
# arrays of every call to the innermost open list; the Graph keeps them alive
# for as long as it can be replayed.
_CAPTURE_KEEP: list = []


def _note_capture(*objs) -> None:
    if _CAPTURE_KEEP:
        _CAPTURE_KEEP[-1].extend(o for o in objs if o is not None)


def context() -> Context:
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
        if os.environ.get("QOB_TUNE_FILE", "1") != "":
            from . import tune
            tune.load()
    return _CTX


def _dt(np_dtype) -> int:
    np_dtype = np.dtype(np_dtype)
    if np_dtype == np.complex128:
        return QOB_C64
    if np_dtype == np.complex64:
        return QOB_C32
    raise ArgumentError(f"unsupported element type {np_dtype}; use complex128 or complex64")


# ------------------------------------------------------------- device data
class DeviceArray:
    """A contiguous complex device array (column-major when 2-D)."""

    def __init__(self, n: int, dtype=np.complex128, shape=None):
        self.dtype = np.dtype(dtype)
        _dt(self.dtype)
        self.n = int(n)
        self.shape = tuple(shape) if shape is not None else (self.n,)
        ctx = context()
        p = C.c_void_p()
        call("qob_dmalloc", ctx.h, max(self.nbytes, 16), C.byref(p))
        self.ptr = p
        self._ctx = ctx

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    @property
    def qdt(self) -> int:
        return _dt(self.dtype)

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.asarray(a)
        dtype = np.dtype(dtype or (a.dtype if a.dtype in (np.complex64, np.complex128) else np.complex128))
        host = np.asfortranarray(a.astype(dtype, copy=False))
        d = cls(host.size, dtype, host.shape)
        if host.size:
            call("qob_memcpy_htod", d._ctx.h, d.ptr, host.ctypes.data_as(C.c_void_p), d.nbytes, None)
        return d

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.n:
            call("qob_memcpy_dtoh", self._ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, None)
        return out

    def copy(self) -> "DeviceArray":
        d = DeviceArray(self.n, self.dtype, self.shape)
        if self.n:
            call("qob_memcpy_dtod", self._ctx.h, d.ptr, self.ptr, self.nbytes, None)
            self._ctx.synchronize()
        return d

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and self.ptr.value:
                _lib.load().qob_dfree(self._ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


# -------------------------------------------------------------------- bases
class Basis:
    """A Hilbert-space basis; ``shape`` holds the per-site dimensions."""

    def __init__(self, shape: Sequence[int], name: str):
        self.shape = tuple(int(s) for s in shape)
        self._name = name
        self._len = math.prod(self.shape)

    def __len__(self) -> int:
        return self._len

    @property
    def bases(self):
        return (self,)

    def __eq__(self, other) -> bool:
        return self is other or (isinstance(other, Basis) and self._ckey() == other._ckey())

    def __hash__(self):
        return hash(self._ckey())

    def _ckey(self):
        # bases are immutable: the comparison key is built once (mul! checks
        # bases on every call)
        k = self.__dict__.get("_kc")
        if k is None:
            k = self._kc = self._key()
        return k

    def _key(self):
        return (self._name, self.shape)

    def __repr__(self):
        return f"{self._name}{self.shape}"


class GenericBasis(Basis):
    def __init__(self, n: int):
        super().__init__((n,), "GenericBasis")


class NLevelBasis(Basis):
    def __init__(self, n: int):
        super().__init__((n,), "NLevelBasis")


class FockBasis(Basis):
    """Fock space with photon numbers offset..N (QuantumInterface FockBasis)."""

    def __init__(self, N: int, offset: int = 0):
        self.N, self.offset = int(N), int(offset)
        super().__init__((self.N - self.offset + 1,), "FockBasis")

    def _key(self):
        return ("FockBasis", self.N, self.offset)


class SpinBasis(Basis):
    """Spin-j basis, m = j, j-1, ..., -j (QuantumInterface SpinBasis)."""

    def __init__(self, spin):
        self.spinnumber = Fraction(spin).limit_denominator(2)
        super().__init__((int(2 * self.spinnumber + 1),), "SpinBasis")

    def _key(self):
        return ("SpinBasis", self.spinnumber)


class CompositeBasis(Basis):
    def __init__(self, *bases: Basis):
        flat = []
        for b in bases:
            flat.extend(b.bases)
        self._bases = tuple(flat)
        super().__init__(tuple(b.shape[0] for b in flat), "CompositeBasis")

    @property
    def bases(self):
        return self._bases

    def _key(self):
        return ("CompositeBasis", tuple(b._key() for b in self._bases))


def tensor(*bases: Basis) -> Basis:
    """Tensor product of bases; site 1 is the fastest index."""
    if len(bases) == 1:
        return bases[0]
    return CompositeBasis(*bases)


# ------------------------------------------------------------------- states
class _State:
    def __init__(self, basis: Basis, data, dtype=None):
        self.basis = basis
        if isinstance(data, DeviceArray):
            self.data = data
        else:
            arr = np.asarray(data).reshape(-1)
            self.data = DeviceArray.from_numpy(arr, dtype)
        if self.data.n != len(basis):
            raise DimensionMismatch(
                f"Tried to assign data of length {self.data.n} to Hilbert space of size {len(basis)}")

    @property
    def dtype(self):
        return self.data.dtype

    def to_numpy(self) -> np.ndarray:
        return self.data.to_numpy().reshape(-1)

    def copy(self):
        return type(self)(self.basis, self.data.copy())

    def __len__(self):
        return len(self.basis)


class Ket(_State):
    """Ket state (src/states.jl:25-32) with device-resident data."""


class Bra(_State):
    """Bra state (src/states.jl:11-18) with device-resident data."""


# ----------------------------------------------------------------- operators
class AbstractOperator:
    basis_l: Basis
    basis_r: Basis

    @property
    def shape(self):
        return (len(self.basis_l), len(self.basis_r))

    def _handle(self) -> C.c_void_p:
        raise MethodError(f"{type(self).__name__} has no device operator handle")

    def __mul__(self, other):
        return _alloc_mul(self, other)

    def __rmul__(self, other):
        return _alloc_mul(other, self)


class _Handle:
    def __init__(self, h: C.c_void_p):
        self.h = h

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib.load().qob_op_release(self.h)
        except Exception:
            pass


class Operator(AbstractOperator):
    """Dense operator / density matrix with device data, m x n column-major
    (Operator{BL,BR,<:Matrix}, src/operators_dense.jl:12-20, 49-52)."""

    def __init__(self, basis_l: Basis, basis_r: Basis, data, dtype=None):
        self.basis_l, self.basis_r = basis_l, basis_r
        m, n = len(basis_l), len(basis_r)
        if isinstance(data, DeviceArray):
            if data.n != m * n:
                raise DimensionMismatch(f"Tried to assign data of size {data.n} to bases of length {m} and {n}!")
            data.shape = (m, n)
            self.data = data
        else:
            arr = np.asarray(data)
            if arr.shape != (m, n):
                raise DimensionMismatch(
                    f"Tried to assign data of size {arr.shape} to bases of length {m} and {n}!")
            self.data = DeviceArray.from_numpy(arr, dtype)
        self._h = None

    @property
    def dtype(self):
        return self.data.dtype

    def to_numpy(self) -> np.ndarray:
        return self.data.to_numpy()

    def copy(self):
        return Operator(self.basis_l, self.basis_r, self.data.copy())

    def _handle(self):
        # a dense Operator used as a LazyTensor factor: host copy uploaded once
        if self._h is None:
            a = np.asfortranarray(self.to_numpy().astype(np.complex128))
            buf = (qob_c64 * a.size).from_buffer_copy(a.tobytes(order="F"))
            h = C.c_void_p()
            call("qob_dense_create", context().h, a.shape[0], a.shape[1], buf, C.byref(h))
            self._h = _Handle(h)
        return self._h.h


def DenseOperator(basis_l: Basis, basis_r: Basis | None = None, data=None, dtype=None) -> Operator:
    """DenseOperator(b1[, b2, data]) (src/operators_dense.jl:59-60)."""
    if data is None and basis_r is not None and not isinstance(basis_r, Basis):
        data, basis_r = basis_r, basis_l
    if basis_r is None:
        basis_r = basis_l
    if data is None:
        data = np.zeros((len(basis_l), len(basis_r)))
    if hasattr(data, "toarray"):
        data = data.toarray()
    return Operator(basis_l, basis_r, np.asarray(data), dtype)


class SparseOperator(AbstractOperator):
    """SparseOperator: SparseMatrixCSC data (src/operators_sparse.jl:5-23).

    ``data`` is a scipy sparse matrix or a dense array; it is held as 1-based
    Int64 CSC arrays, exactly what Julia hands to ``qob_sparse_create``.
    """

    def __init__(self, basis_l: Basis, basis_r: Basis | None = None, data=None):
        if data is None and basis_r is not None and not isinstance(basis_r, Basis):
            data, basis_r = basis_r, basis_l
        if basis_r is None:
            basis_r = basis_l
        import scipy.sparse as sp

        self.basis_l, self.basis_r = basis_l, basis_r
        m, n = len(basis_l), len(basis_r)
        M = sp.csc_matrix(data if data is not None else (m, n), dtype=np.complex128)
        if M.shape != (m, n):
            raise DimensionMismatch(f"Tried to assign data of size {M.shape} to bases of length {m} and {n}!")
        M.sort_indices()
        self.csc = M
        self._h = None
        self._adj_of = None

    @property
    def dtype(self):
        return np.dtype(np.complex128)

    def _handle(self):
        if self._h is None:
            if self._adj_of is not None:
                h = C.c_void_p()
                call("qob_op_adjoint", context().h, self._adj_of._handle(), C.byref(h))
            else:
                M = self.csc
                colptr = (M.indptr.astype(np.int64) + 1)
                rowval = (M.indices.astype(np.int64) + 1)
                nz = np.ascontiguousarray(M.data.astype(np.complex128))
                h = C.c_void_p()
                call("qob_sparse_create", context().h, M.shape[0], M.shape[1],
                     colptr.ctypes.data_as(C.POINTER(C.c_int64)),
                     rowval.ctypes.data_as(C.POINTER(C.c_int64)),
                     nz.ctypes.data_as(C.POINTER(qob_c64)), C.byref(h))
            self._h = _Handle(h)
        return self._h.h

    def to_dense(self) -> np.ndarray:
        return self.csc.toarray()


class EyeOperator(AbstractOperator):
    """identityoperator(b1, b2) = Operator(b1, b2, Eye(m, n))."""

    def __init__(self, basis_l: Basis, basis_r: Basis | None = None):
        self.basis_l, self.basis_r = basis_l, (basis_r or basis_l)
        self._h = None

    @property
    def dtype(self):
        return np.dtype(np.complex128)

    def _handle(self):
        if self._h is None:
            h = C.c_void_p()
            call("qob_eye_create", context().h, len(self.basis_l), len(self.basis_r), C.byref(h))
            self._h = _Handle(h)
        return self._h.h


def identityoperator(basis_l: Basis, basis_r: Basis | None = None) -> EyeOperator:
    return EyeOperator(basis_l, basis_r)


def _as_tuple(x):
    if isinstance(x, (list, tuple)):
        return tuple(x)
    return (x,)


class LazyTensor(AbstractOperator):
    """LazyTensor(basis_l, basis_r, indices, operators, factor=1)
    (src/operators_lazytensor.jl:15-58).  ``indices`` are 1-based and
    sorted, as in the reference."""

    def __init__(self, basis_l: Basis, basis_r, indices=None, operators=None, factor=1.0):
        if not isinstance(basis_r, Basis):
            # LazyTensor(basis, indices, operators[, factor])
            basis_l, basis_r, indices, operators, factor = (basis_l, basis_l, basis_r, indices,
                                                            1.0 if operators is None else operators)
        self.basis_l, self.basis_r = basis_l, basis_r
        self.indices = tuple(int(i) for i in _as_tuple(indices))
        self.operators = _as_tuple(operators)
        self.factor = complex(factor)
        if len(self.indices) != len(self.operators):
            raise ArgumentError("indices and operators must have the same length")
        if list(self.indices) != sorted(self.indices):
            raise ArgumentError("LazyTensor indices must be sorted")
        self._h = None

    def _handle(self):
        if self._h is None:
            N = len(self.basis_l.shape)
            shl = np.array(self.basis_l.shape, dtype=np.int64)
            shr = np.array(self.basis_r.shape, dtype=np.int64)
            if len(shr) != N:
                raise ArgumentError("basis_l and basis_r must have the same number of sites")
            idx = np.array(self.indices, dtype=np.int64)
            hs = [op._handle() for op in self.operators]
            arr = (C.c_void_p * max(1, len(hs)))(*[h.value for h in hs])
            h = C.c_void_p()
            call("qob_lazytensor_create", context().h, N, shl.ctypes.data_as(C.POINTER(C.c_int64)),
                 shr.ctypes.data_as(C.POINTER(C.c_int64)), len(hs), idx.ctypes.data_as(C.POINTER(C.c_int64)),
                 arr, c64(self.factor), C.byref(h))
            self._h = _Handle(h)
            self._keep = self.operators
        return self._h.h

    def __rmul__(self, a):
        if isinstance(a, (int, float, complex, np.number)):
            return LazyTensor(self.basis_l, self.basis_r, self.indices, self.operators, self.factor * a)
        return _alloc_mul(a, self)

    def __mul__(self, a):
        if isinstance(a, (int, float, complex, np.number)):
            return LazyTensor(self.basis_l, self.basis_r, self.indices, self.operators, self.factor * a)
        return _alloc_mul(self, a)


class LazySum(AbstractOperator):
    """LazySum([basis_l, basis_r,] factors, operators)
    (src/operators_lazysum.jl:41-60).  An empty sum is the zero operator."""

    def __init__(self, *args):
        if len(args) == 2:
            factors, ops = args
            ops = _as_tuple(ops)
            if not ops:
                raise ArgumentError("LazySum without operators needs explicit bases")
            basis_l, basis_r = ops[0].basis_l, ops[0].basis_r
        elif len(args) == 4:
            basis_l, basis_r, factors, ops = args
            ops = _as_tuple(ops)
        else:
            raise ArgumentError("LazySum(factors, operators) or LazySum(bl, br, factors, operators)")
        factors = [complex(f) for f in factors]
        if len(factors) != len(ops):
            raise ArgumentError("LazySum `operators` and `factors` have different lengths.")
        for o in ops:
            if o.basis_l != basis_l or o.basis_r != basis_r:
                raise IncompatibleBases("LazySum terms must share bases")
        self.basis_l, self.basis_r = basis_l, basis_r
        self.factors = factors
        self.operators = ops
        self._h = None

    @classmethod
    def zero(cls, basis_l: Basis, basis_r: Basis | None = None) -> "LazySum":
        """LazySum(basis_l, basis_r): the zero operator."""
        return cls(basis_l, basis_r or basis_l, [], ())

    def _handle(self):
        if self._h is None:
            hs = [op._handle() for op in self.operators]
            arr = (C.c_void_p * max(1, len(hs)))(*[h.value for h in hs])
            fac = (qob_c64 * max(1, len(hs)))(*[c64(f) for f in self.factors])
            h = C.c_void_p()
            call("qob_lazysum_create", context().h, len(self.basis_l), len(self.basis_r), len(hs), fac, arr,
                 C.byref(h))
            self._h = _Handle(h)
            self._keep = self.operators
        return self._h.h

    def set_factors(self, factors) -> None:
        """Replace the coefficients in place (TimeDependentSum set_time!)."""
        factors = [complex(f) for f in factors]
        if len(factors) != len(self.operators):
            raise ArgumentError("wrong number of factors")
        self.factors = factors
        if self._h is not None:
            fac = (qob_c64 * max(1, len(factors)))(*[c64(f) for f in factors])
            call("qob_lazysum_set_factors", context().h, self._h.h, fac)

    def set_factors_device(self, dfactors, stream=None) -> None:
        """Take the coefficients from device memory (len(operators) complex
        doubles at ``dfactors``, a DeviceArray or a raw pointer), stream-ordered
        and capturable; the host copy ``factors`` is not updated."""
        ptr = dfactors.ptr if isinstance(dfactors, DeviceArray) else dfactors
        _note_capture(self, dfactors)
        call("qob_lazysum_set_factors_d", context().h, self._handle(), ptr, stream)


class LazyProduct(AbstractOperator):
    """LazyProduct(operators[, factor]) or LazyProduct(op1, op2, ...)
    (src/operators_lazyproduct.jl:24-58): factor * op1 * op2 * ... applied
    factor by factor (mul! :103-163)."""

    def __init__(self, *args, factor=1.0):
        if len(args) == 1 and isinstance(args[0], (list, tuple)):
            ops = tuple(args[0])
        elif len(args) == 2 and isinstance(args[0], (list, tuple)):
            ops, factor = tuple(args[0]), args[1]
        else:
            ops = tuple(args)
        if not ops:
            raise ArgumentError("LazyProduct needs at least one operator!")
        for a, b in zip(ops[:-1], ops[1:]):
            if a.basis_r != b.basis_l:
                raise IncompatibleBases("LazyProduct: operators are not multiplicable")
        self.operators = ops
        self.factor = complex(factor)
        self.basis_l, self.basis_r = ops[0].basis_l, ops[-1].basis_r
        self._h = None

    def _handle(self):
        if self._h is None:
            hs = [op._handle() for op in self.operators]
            arr = (C.c_void_p * len(hs))(*[h.value for h in hs])
            h = C.c_void_p()
            call("qob_lazyproduct_create", context().h, len(hs), arr, c64(self.factor), C.byref(h))
            self._h = _Handle(h)
            self._keep = self.operators
        return self._h.h


class _Adjoint(AbstractOperator):
    """dagger of an operator with a device handle."""

    def __init__(self, parent: AbstractOperator):
        self.parent = parent
        self.basis_l, self.basis_r = parent.basis_r, parent.basis_l
        self._h = None

    @property
    def dtype(self):
        return np.dtype(np.complex128)

    def _handle(self):
        if self._h is None:
            h = C.c_void_p()
            call("qob_op_adjoint", context().h, self.parent._handle(), C.byref(h))
            self._h = _Handle(h)
        return self._h.h


def dagger(x):
    """Hermitian conjugate (src/states.jl:96-97, src/operators_dense.jl:128)."""
    if isinstance(x, Ket):
        return Bra(x.basis, np.conj(x.to_numpy()), x.dtype)
    if isinstance(x, Bra):
        return Ket(x.basis, np.conj(x.to_numpy()), x.dtype)
    if isinstance(x, Operator):
        return Operator(x.basis_r, x.basis_l, np.conj(x.to_numpy()).T, x.dtype)
    if isinstance(x, _Adjoint):
        return x.parent
    if isinstance(x, AbstractOperator):
        return _Adjoint(x)
    raise MethodError(f"dagger not defined for {type(x).__name__}")


# ---------------------------------------------------------------------- mul!
_DT_CODE = {np.dtype(np.complex128): QOB_C64, np.dtype(np.complex64): QOB_C32}


def _check_same_dtype(*xs):
    d0 = xs[0].dtype
    for x in xs[1:]:
        if x.dtype != d0:
            raise ArgumentError(f"mixed element types {set(y.dtype for y in xs)}: states of one call share one type")
    code = _DT_CODE.get(d0)
    return _dt(d0) if code is None else code


def mul_(result, a, b, alpha=1.0, beta=0.0):
    """mul!(result, a, b, α, β): result = α a b + β result, in place.

    Dispatch mirrors the reference methods:
      (Ket, op, Ket)            op * ket
      (Bra, Bra, op)            bra * op
      (Operator, op, Operator)  left application on a density matrix
      (Operator, Operator, op)  right application
    """
    if type(a).__name__ == "SuperOperator":  # S * rho on vec(rho) (src/superoperators.jl:82-85)
        from .superop import apply
        if not isinstance(result, Operator) or not isinstance(b, Operator):
            raise MethodError("a SuperOperator acts on an Operator")
        apply(a, b, result, alpha, beta)
        return result
    ctx = context()
    al, be = c64(alpha), c64(beta)
    _note_capture(result, a, b)
    if isinstance(result, Ket):
        if not (isinstance(b, Ket) and isinstance(a, AbstractOperator)):
            raise MethodError(f"no method mul!(::Ket, ::{type(a).__name__}, ::{type(b).__name__})")
        _bases(a.basis_r, b.basis, a.basis_l, result.basis)
        if result.data.ptr.value == b.data.ptr.value:
            raise ArgumentError("output must not be aliased with input")
        if isinstance(a, Operator):
            _gemm_dense(0, len(a.basis_l), 1, len(a.basis_r), alpha, a, b.data, beta, result.data)
            return result
        dtb, dtr = _dt(b.dtype), _dt(result.dtype)
        if dtb == dtr:
            call("qob_mul_ket", ctx.h, dtr, a._handle(), b.data.ptr, b.data.n, result.data.ptr, result.data.n, al,
                 be, None)
        else:
            call("qob_mul_ket_x", ctx.h, dtb, a._handle(), b.data.ptr, b.data.n, dtr, result.data.ptr,
                 result.data.n, al, be, None)
        return result
    if isinstance(result, Bra):
        if not (isinstance(a, Bra) and isinstance(b, AbstractOperator)):
            raise MethodError(f"no method mul!(::Bra, ::{type(a).__name__}, ::{type(b).__name__})")
        _bases(a.basis, b.basis_l, result.basis, b.basis_r)
        if isinstance(b, Operator):
            _gemm_dense(1, len(b.basis_r), 1, len(b.basis_l), alpha, b, a.data, beta, result.data)
            return result
        dta, dtr = _dt(a.dtype), _dt(result.dtype)
        if dta == dtr:
            call("qob_mul_bra", ctx.h, dtr, a.data.ptr, a.data.n, b._handle(), result.data.ptr, result.data.n, al,
                 be, None)
        else:
            call("qob_mul_bra_x", ctx.h, dta, a.data.ptr, a.data.n, b._handle(), dtr, result.data.ptr,
                 result.data.n, al, be, None)
        return result
    if isinstance(result, Operator):
        if isinstance(b, Operator) and isinstance(a, AbstractOperator) and not (
                isinstance(a, Operator) and a is result):
            if isinstance(a, Operator):
                _bases(a.basis_r, b.basis_l, a.basis_l, result.basis_l)
                _gemm_dense(0, len(a.basis_l), len(b.basis_r), len(a.basis_r), alpha, a, b.data, beta,
                            result.data)
                return result
            _bases(a.basis_r, b.basis_l, a.basis_l, result.basis_l)
            _bases(b.basis_r, result.basis_r)
            dtb, dtr = _dt(b.dtype), _dt(result.dtype)
            if dtb == dtr:
                call("qob_mul_op_left", ctx.h, dtr, a._handle(), b.data.ptr, len(b.basis_l), len(b.basis_r),
                     result.data.ptr, len(result.basis_l), al, be, None)
            else:
                call("qob_mul_op_left_x", ctx.h, dtb, a._handle(), b.data.ptr, len(b.basis_l), len(b.basis_r),
                     dtr, result.data.ptr, len(result.basis_l), al, be, None)
            return result
        if isinstance(a, Operator) and isinstance(b, AbstractOperator):
            _bases(a.basis_r, b.basis_l, a.basis_l, result.basis_l)
            _bases(b.basis_r, result.basis_r)
            dta, dtr = _dt(a.dtype), _dt(result.dtype)
            if dta == dtr:
                call("qob_mul_op_right", ctx.h, dtr, a.data.ptr, len(a.basis_l), len(a.basis_r), b._handle(),
                     result.data.ptr, len(result.basis_r), al, be, None)
            else:
                call("qob_mul_op_right_x", ctx.h, dta, a.data.ptr, len(a.basis_l), len(a.basis_r), b._handle(),
                     dtr, result.data.ptr, len(result.basis_r), al, be, None)
            return result
    raise MethodError(f"no method mul!(::{type(result).__name__}, ::{type(a).__name__}, ::{type(b).__name__})")


def _bases(*pairs):
    # pairs of bases that must agree; lengths are checked by the library
    # (DimensionMismatch), basis identity here (IncompatibleBases).
    for i in range(0, len(pairs), 2):
        x, y = pairs[i], pairs[i + 1]
        if x is not y and len(x) == len(y) and x != y and not (_generic(x) or _generic(y)):
            raise IncompatibleBases(f"{x} != {y}")


def _generic(b: Basis) -> bool:
    return isinstance(b, GenericBasis)


def _widened(d: DeviceArray, copy_in: bool = True) -> DeviceArray:
    """A ComplexF64 copy of a ComplexF32 device array (exact); ComplexF64
    arrays are returned as they are."""
    if d.dtype == np.complex128:
        return d
    w = DeviceArray(d.n, np.complex128, d.shape)
    if copy_in and d.n:
        call("qob_convert", context().h, QOB_C32, d.ptr, QOB_C64, w.ptr, d.n, None)
    return w


def _gemm_dense(transa, m, n, k, alpha, A: Operator, B: DeviceArray, beta, Cd: DeviceArray):
    """C = alpha op(A) B + beta C for a device DenseOperator A
    (src/operators_dense.jl:394-396).  When the element types of A, B and C
    differ the product runs in ComplexF64 (the promoted type) and a
    ComplexF32 C is rounded once at the end."""
    if Cd.ptr.value in (A.data.ptr.value, B.ptr.value):
        raise ArgumentError("output must not be aliased with input")
    if Cd.n != m * n or B.n != k * n:
        raise DimensionMismatch(f"dense mul!: A {m}x{k}, B {B.n} elements, C {Cd.n} elements")
    h = context().h
    if A.data.dtype == B.dtype == Cd.dtype:
        call("qob_gemm", h, _dt(Cd.dtype), transa, m, n, k, c64(alpha), A.data.ptr, B.ptr, c64(beta), Cd.ptr, None)
        return
    Aw, Bw = _widened(A.data), _widened(B)
    Cw = _widened(Cd, copy_in=complex(beta) != 0)
    call("qob_gemm", h, QOB_C64, transa, m, n, k, c64(alpha), Aw.ptr, Bw.ptr, c64(beta), Cw.ptr, None)
    if Cw is not Cd:
        call("qob_convert", h, QOB_C64, Cw.ptr, QOB_C32, Cd.ptr, Cd.n, None)
    call("qob_stream_synchronize", h, None)  # the temporaries are freed on return


def _alloc_mul(a, b):
    """Allocating * (src/operators_dense.jl:101-120): similar + mul!."""
    if isinstance(a, AbstractOperator) and isinstance(b, Ket):
        r = Ket(a.basis_l, DeviceArray(len(a.basis_l), b.dtype))
        return mul_(r, a, b, 1.0, 0.0)
    if isinstance(a, Bra) and isinstance(b, AbstractOperator):
        r = Bra(b.basis_r, DeviceArray(len(b.basis_r), a.dtype))
        return mul_(r, a, b, 1.0, 0.0)
    if isinstance(a, AbstractOperator) and isinstance(b, Operator):
        r = Operator(a.basis_l, b.basis_r, DeviceArray(len(a.basis_l) * len(b.basis_r), b.dtype))
        return mul_(r, a, b, 1.0, 0.0)
    if isinstance(a, Operator) and isinstance(b, AbstractOperator):
        r = Operator(a.basis_l, b.basis_r, DeviceArray(len(a.basis_l) * len(b.basis_r), a.dtype))
        return mul_(r, a, b, 1.0, 0.0)
    return NotImplemented


# --------------------------------------------------------------- reductions
def norm(x: _State) -> float:
    """norm(ket) (src/states.jl:89), accumulated in FP64."""
    out = C.c_double()
    call("qob_norm", context().h, _dt(x.dtype), x.data.n, x.data.ptr, C.byref(out), None)
    return out.value


def normalize_(x):
    """normalize!(ket) (src/states.jl:90) / normalize!(op) (src/operators_dense.jl:217)."""
    if isinstance(x, _State):
        call("qob_normalize_ket", context().h, _dt(x.dtype), x.data.n, x.data.ptr, None)
    elif isinstance(x, Operator):
        if len(x.basis_l) != len(x.basis_r):
            raise DimensionMismatch("normalize! of a non-square operator")
        call("qob_normalize_op", context().h, _dt(x.dtype), len(x.basis_l), x.data.ptr, None)
    else:
        raise MethodError(f"normalize! not defined for {type(x).__name__}")
    return x


def dot(x: _State, y: _State) -> complex:
    """sum conj(x_i) y_i accumulated in FP64."""
    out = qob_c64()
    _check_same_dtype(x, y)
    if x.data.n != y.data.n:
        raise DimensionMismatch("dot of states of different length")
    call("qob_dot", context().h, _dt(x.dtype), x.data.n, x.data.ptr, y.data.ptr, C.byref(out), None)
    return complex(out.re, out.im)


def tr(op: Operator) -> complex:
    """tr(op) (src/operators_dense.jl:189), FP64 accumulation."""
    if len(op.basis_l) != len(op.basis_r):
        raise DimensionMismatch("trace of a non-square operator")
    out = qob_c64()
    call("qob_trace", context().h, _dt(op.dtype), len(op.basis_l), op.data.ptr, C.byref(out), None)
    return complex(out.re, out.im)


def expect(op: AbstractOperator, state) -> complex:
    """expect(op, ket) = <psi|op|psi>, expect(op, rho) = tr(op rho)
    (src/operators_dense.jl:219-230, src/operators.jl:119)."""
    out = qob_c64()
    if isinstance(state, Ket):
        if isinstance(op, Operator):
            tmp = op * state
            return dot(state, tmp)
        call("qob_expect_ket", context().h, _dt(state.dtype), op._handle(), state.data.ptr, state.data.n,
             C.byref(out), None)
    elif isinstance(state, Operator):
        if isinstance(op, Operator):
            return tr(op * state)
        call("qob_expect_op", context().h, _dt(state.dtype), op._handle(), state.data.ptr, len(state.basis_l),
             C.byref(out), None)
    else:
        raise MethodError(f"expect not defined for {type(state).__name__}")
    return complex(out.re, out.im)


def variance(op: AbstractOperator, state) -> complex:
    """variance(op, ket) = <psi|op op|psi> - <psi|op|psi>^2 with x = op psi
    formed once (src/operators.jl:139-143); on a density matrix (the
    operator form the docstring at src/operators.jl:132-137 names)
    variance(op, rho) = tr(op op rho) - tr(op rho)^2 with x = op rho formed
    once and both traces taken by expect (fused where the operator allows)."""
    if isinstance(state, Operator):
        if isinstance(op, Operator):
            x = op * state
            return expect(op, x) - expect(op, state) ** 2
        if len(state.basis_l) != len(state.basis_r):
            raise DimensionMismatch("variance on a non-square operator")
        out = qob_c64()
        call("qob_variance_op", context().h, _dt(state.dtype), op._handle(), state.data.ptr, len(state.basis_l),
             C.byref(out), None)
        return complex(out.re, out.im)
    if not isinstance(state, Ket):
        raise MethodError(f"variance not defined for {type(state).__name__}")
    if isinstance(op, Operator):
        x = op * state
        return dot(state, op * x) - dot(state, x) ** 2
    out = qob_c64()
    call("qob_variance_ket", context().h, _dt(state.dtype), op._handle(), state.data.ptr, state.data.n,
         C.byref(out), None)
    return complex(out.re, out.im)


def axpy_(y: _State, a, x: _State) -> _State:
    """y .+= a .* x on device data."""
    _check_same_dtype(x, y)
    if x.data.n != y.data.n:
        raise DimensionMismatch("axpy of arrays of different length")
    call("qob_axpy", context().h, _dt(y.dtype), y.data.n, c64(a), x.data.ptr, y.data.ptr, None)
    return y


def synchronize() -> None:
    context().synchronize()
