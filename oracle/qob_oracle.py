"""ORACLE — CPU restatement of QuantumOpticsBase.jl's mul! hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, and only as the checker.  The product path
(libqob_hip.so and the qobhip package) never imports or calls it.

Each function restates one reference function (file:line cited) in numpy,
keeping its loop structure where that structure defines the arithmetic
(CSC loops, the recursive sparse LazyTensor engine) and its data movement
where it is a reshape/permute + GEMM (_tp_matmul*).  The reference is Julia
and cannot run here, so the oracle is pinned by independent checks in
tests/test_oracle.py:
  * brute force: every mul! path against the dense Kronecker product built
    by ``dense_of`` (kron in the reference's order, site 1 fastest,
    src/operators_dense.jl:134), which shares no code with the restated
    loops;
  * closed forms (spin/Fock matrix elements, sigma_z chain energies on basis
    states, bit flips of sigma_x sigma_x, number-operator diagonals);
  * the reference's own known-answer checks (src/precompile.jl:16-18,
    :89-106);
  * invariants (linearity, Hermiticity, tr(H rho - rho H) = 0);
  * planted mistakes in this module (monkeypatched in tests) that the pins
    must catch.
Parity of every function below is pinned by at least one of these.

Conventions: operators are these small classes, not qobhip objects; index
arrays of sparse data are 1-based Int64 as Julia holds them; dense data is
numpy (column order is irrelevant for numpy arithmetic).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np


# --------------------------------------------------------------- exceptions
class DimensionMismatch(ValueError):
    pass


class ArgumentError(ValueError):
    pass


# ----------------------------------------------------------------- operators
@dataclass
class OSparse:
    """SparseMatrixCSC (1-based colptr/rowval), optionally adjoint-wrapped
    (SparseOpPureType / SparseOpAdjType, src/operators_sparse.jl:5-7)."""
    m: int
    n: int
    colptr: np.ndarray
    rowval: np.ndarray
    nzval: np.ndarray
    adj: bool = False

    @property
    def shape(self):
        return (self.n, self.m) if self.adj else (self.m, self.n)

    @classmethod
    def from_scipy(cls, M, adj=False):
        import scipy.sparse as sp
        M = sp.csc_matrix(M, dtype=np.complex128)
        M.sort_indices()
        return cls(M.shape[0], M.shape[1], M.indptr.astype(np.int64) + 1, M.indices.astype(np.int64) + 1,
                   M.data.astype(np.complex128).copy(), adj)

    def nnz(self):
        return int(self.colptr[-1] - 1)

    def parent_dense(self):
        A = np.zeros((self.m, self.n), dtype=np.complex128)
        for c in range(self.n):
            for p in range(self.colptr[c] - 1, self.colptr[c + 1] - 1):
                A[self.rowval[p] - 1, c] += self.nzval[p]
        return A


@dataclass
class ODense:
    data: np.ndarray
    adj: bool = False

    @property
    def shape(self):
        return self.data.shape[::-1] if self.adj else self.data.shape

    def mat(self):
        return self.data.conj().T if self.adj else self.data


@dataclass
class OEye:
    """FillArrays Eye(m, n) (identityoperator, src/operators_sparse.jl:180-183)."""
    m: int
    n: int

    @property
    def shape(self):
        return (self.m, self.n)


@dataclass
class OLazyTensor:
    shape_l: Tuple[int, ...]
    shape_r: Tuple[int, ...]
    indices: Tuple[int, ...]  # 1-based, sorted (src/operators_lazytensor.jl:26)
    operators: Tuple
    factor: complex = 1.0

    @property
    def shape(self):
        return (int(np.prod(self.shape_l)), int(np.prod(self.shape_r)))


@dataclass
class OLazySum:
    m: int
    n: int
    factors: List[complex]
    operators: List

    @property
    def shape(self):
        return (self.m, self.n)


@dataclass
class OLazyProduct:
    """LazyProduct (src/operators_lazyproduct.jl:24-58): factor * A1 ... An."""
    operators: List
    factor: complex = 1.0

    @property
    def shape(self):
        return (self.operators[0].shape[0], self.operators[-1].shape[1])


def dagger(op):
    """dagger (src/operators_dense.jl:128 / LazyTensor, LazySum algebra)."""
    if isinstance(op, OSparse):
        return OSparse(op.m, op.n, op.colptr, op.rowval, op.nzval, not op.adj)
    if isinstance(op, ODense):
        return ODense(op.data, not op.adj)
    if isinstance(op, OEye):
        return OEye(op.n, op.m)
    if isinstance(op, OLazyTensor):
        return OLazyTensor(op.shape_r, op.shape_l, op.indices, tuple(dagger(o) for o in op.operators),
                           np.conj(op.factor))
    if isinstance(op, OLazySum):
        return OLazySum(op.n, op.m, [np.conj(f) for f in op.factors], [dagger(o) for o in op.operators])
    if isinstance(op, OLazyProduct):
        return OLazyProduct([dagger(o) for o in reversed(op.operators)], np.conj(op.factor))
    raise TypeError(type(op))


# -------------------------------------------------------- brute-force dense
def _mat(op) -> np.ndarray:
    if isinstance(op, OSparse):
        A = op.parent_dense()
        return A.conj().T if op.adj else A
    if isinstance(op, ODense):
        return op.mat()
    if isinstance(op, OEye):
        return np.eye(op.m, op.n, dtype=np.complex128)
    return dense_of(op)


def dense_of(op) -> np.ndarray:
    """Full matrix of any operator by Kronecker products in the reference's
    order: tensor(a, b) = kron(b, a) (src/operators_dense.jl:134), so site 1
    is the fastest index.  Independent of every restated loop below."""
    if isinstance(op, OLazyTensor):
        mats = []
        for s in range(len(op.shape_l)):
            if s + 1 in op.indices:
                mats.append(_mat(op.operators[op.indices.index(s + 1)]))
            else:
                mats.append(np.eye(op.shape_l[s], op.shape_r[s], dtype=np.complex128))
        full = np.ones((1, 1), dtype=np.complex128)
        for A in mats:
            full = np.kron(A, full)
        return op.factor * full
    if isinstance(op, OLazySum):
        out = np.zeros((op.m, op.n), dtype=np.complex128)
        for f, o in zip(op.factors, op.operators):
            out = out + f * dense_of(o)
        return out
    if isinstance(op, OLazyProduct):
        out = op.factor * dense_of(op.operators[0])
        for o in op.operators[1:]:
            out = out @ dense_of(o)
        return out
    return _mat(op)


# ----------------------------------------------- src/sparsematrix.jl restated
def gemm_sp_dense_small(alpha, M: OSparse, B, result):
    """src/sparsematrix.jl:1-23: for each stored (row, col) of M,
    result[row, :] += (alpha*val) * B[col, :] (inner j loop vectorised;
    per-element accumulation order unchanged)."""
    one = alpha == 1
    for colindex in range(M.n):
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            row = M.rowval[i] - 1
            val = M.nzval[i] if one else alpha * M.nzval[i]
            result[row, :] += val * B[colindex, :]


def gemm_sp_dense_big(alpha, M: OSparse, B, result):
    """src/sparsematrix.jl:25-47: j outer; m2 = alpha*B[col, j];
    result[row, j] += nz * m2.  Vectorised over j: each element still
    accumulates in increasing (col, i) order."""
    one = alpha == 1
    for colindex in range(M.n):
        m2 = B[colindex, :] if one else alpha * B[colindex, :]
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            result[M.rowval[i] - 1, :] += M.nzval[i] * m2


def gemm_dense_adj_sp(alpha, B, M: OSparse, result):
    """src/sparsematrix.jl:49-71: result[:, row] += (alpha*conj(nz)) B[:, col]."""
    for colindex in range(M.n):
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            row = M.rowval[i] - 1
            val = np.conj(M.nzval[i]) if alpha == 1 else alpha * np.conj(M.nzval[i])
            result[:, row] += val * B[:, colindex]


def gemm_adj_sp_dense_small(alpha, M: OSparse, B, result):
    """src/sparsematrix.jl:73-96: result[col, :] += conj(nz)*alpha * B[row, :]."""
    for colindex in range(M.n):
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            mi = np.conj(M.nzval[i]) if alpha == 1 else np.conj(M.nzval[i]) * alpha
            result[colindex, :] += mi * B[M.rowval[i] - 1, :]


def _beta_prep(beta, result):
    # fill!/rmul! of src/sparsematrix.jl:103-107 and :119-123
    if beta == 0:
        result[...] = 0
    elif beta != 1:
        result *= beta


def gemm_sp_left(alpha, M: OSparse, B, beta, result):
    """gemm!(alpha, M::SparseMatrixCSC, B, beta, result), src/sparsematrix.jl:99-113,
    and the adjoint form :148-163."""
    if M.adj:
        if M.nnz() > 550:
            # :151 LinearAlgebra.mul!(result, M', B, alpha, beta) (SparseArrays)
            Md = M.parent_dense().conj().T
            if Md.shape[1] != B.shape[0] or Md.shape[0] != result.shape[0] or B.shape[1] != result.shape[1]:
                raise DimensionMismatch()
            new = alpha * (Md @ B)
            if beta != 0:
                new = new + beta * result
            result[...] = new
            return result
        if M.n != result.shape[0] or M.m != B.shape[0] or B.shape[1] != result.shape[1]:
            raise DimensionMismatch()
        _beta_prep(beta, result)
        gemm_adj_sp_dense_small(alpha, M, B, result)
        return result
    if M.n != B.shape[0] or M.m != result.shape[0] or B.shape[1] != result.shape[1]:
        raise DimensionMismatch()
    _beta_prep(beta, result)
    if M.nnz() > 550:
        gemm_sp_dense_big(alpha, M, B, result)
    else:
        gemm_sp_dense_small(alpha, M, B, result)
    return result


def gemm_sp_right(alpha, B, M: OSparse, beta, result):
    """gemm!(alpha, B, M::SparseMatrixCSC, beta, result), src/sparsematrix.jl:115-146
    (column axpys) and the adjoint form :165-175."""
    if M.adj:
        if M.n != B.shape[1] or M.m != result.shape[1] or B.shape[0] != result.shape[0]:
            raise DimensionMismatch()
        _beta_prep(beta, result)
        gemm_dense_adj_sp(alpha, B, M, result)
        return result
    if M.m != B.shape[1] or M.n != result.shape[1] or B.shape[0] != result.shape[0]:
        raise DimensionMismatch()
    _beta_prep(beta, result)
    for colindex in range(M.n):
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            mi = M.nzval[i] if alpha == 1 else M.nzval[i] * alpha
            result[:, colindex] += mi * B[:, M.rowval[i] - 1]
    return result


def gemv_sp(alpha, M: OSparse, v, beta, result):
    """gemv!(alpha, M, v, beta, result), src/sparsematrix.jl:190-214."""
    if M.n != v.shape[0] or M.m != result.shape[0]:
        raise DimensionMismatch()
    _beta_prep(beta, result)
    for colindex in range(M.n):
        vj = v[colindex] if alpha == 1 else alpha * v[colindex]
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            result[M.rowval[i] - 1] += M.nzval[i] * vj
    return result


def gemv_vec_sp(alpha, v, M: OSparse, beta, result):
    """gemv!(alpha, v, M, beta, result) (vector x CSC), src/sparsematrix.jl:216-238."""
    if M.m != v.shape[0] or M.n != result.shape[0]:
        raise DimensionMismatch()
    _beta_prep(beta, result)
    for colindex in range(M.n):
        for i in range(M.colptr[colindex] - 1, M.colptr[colindex + 1] - 1):
            if alpha == 1:
                result[colindex] += M.nzval[i] * v[M.rowval[i] - 1]
            else:
                result[colindex] += M.nzval[i] * alpha * v[M.rowval[i] - 1]
    return result


# --------------------------------------- src/operators_lazytensor.jl restated
def _strides(shape):
    """_strides (src/operators_dense.jl:296-308), column-major, 1-based use."""
    S = [1] * len(shape)
    for m in range(1, len(shape)):
        S[m] = S[m - 1] * shape[m - 1]
    return S


def _get_shape_and_strides(h: OLazyTensor):
    """src/operators_lazytensor.jl:741-746."""
    shape = [min(a, b) for a, b in zip(h.shape_l, h.shape_r)]
    return shape, _strides(h.shape_l), _strides(h.shape_r)


def _suboperator(h: OLazyTensor, i_k):
    return h.operators[h.indices.index(i_k)]


def _gemm_recursive_lazy_dense(i_k, N_k, K, J, val, shape, strides_k, strides_j, indices, h, op, result):
    """src/operators_lazytensor.jl:652-685: result[J, :] += val * op[K, :]
    (1-based K, J as in the reference)."""
    if i_k > N_k:
        result[J - 1] += val * op[K - 1]
        return
    if i_k in indices:
        h_i = _suboperator(h, i_k)
        if isinstance(h_i, OSparse) and not h_i.adj:
            for k in range(1, h_i.n + 1):
                K_ = K + strides_k[i_k - 1] * (k - 1)
                for jptr in range(h_i.colptr[k - 1], h_i.colptr[k]):
                    j = h_i.rowval[jptr - 1]
                    J_ = J + strides_j[i_k - 1] * (j - 1)
                    val_ = val * h_i.nzval[jptr - 1]
                    _gemm_recursive_lazy_dense(i_k + 1, N_k, K_, J_, val_, shape, strides_k, strides_j, indices, h,
                                               op, result)
            return
        elif not isinstance(h_i, OEye):
            raise ArgumentError(f"gemm! of LazyTensor is not implemented for {type(h_i).__name__}")
    for k in range(1, shape[i_k - 1] + 1):
        K_ = K + strides_k[i_k - 1] * (k - 1)
        J_ = J + strides_j[i_k - 1] * (k - 1)
        _gemm_recursive_lazy_dense(i_k + 1, N_k, K_, J_, val, shape, strides_k, strides_j, indices, h, op, result)


def _gemm_recursive_dense_lazy(i_k, N_k, K, J, val, shape, strides_k, strides_j, indices, h, op, result):
    """src/operators_lazytensor.jl:613-648: result[..., K] += val * op[..., J]."""
    if i_k > N_k:
        if op.ndim == 1:
            result[K - 1] += val * op[J - 1]
        else:
            result[:, K - 1] += val * op[:, J - 1]
        return
    if i_k in indices:
        h_i = _suboperator(h, i_k)
        if isinstance(h_i, OSparse) and not h_i.adj:
            for k in range(1, h_i.n + 1):
                K_ = K + strides_k[i_k - 1] * (k - 1)
                for jptr in range(h_i.colptr[k - 1], h_i.colptr[k]):
                    j = h_i.rowval[jptr - 1]
                    J_ = J + strides_j[i_k - 1] * (j - 1)
                    val_ = val * h_i.nzval[jptr - 1]
                    _gemm_recursive_dense_lazy(i_k + 1, N_k, K_, J_, val_, shape, strides_k, strides_j, indices, h,
                                               op, result)
            return
        elif not isinstance(h_i, OEye):
            raise ArgumentError(f"gemm! of LazyTensor is not implemented for {type(h_i).__name__}")
    for k in range(1, shape[i_k - 1] + 1):
        K_ = K + strides_k[i_k - 1] * (k - 1)
        J_ = J + strides_j[i_k - 1] * (k - 1)
        _gemm_recursive_dense_lazy(i_k + 1, N_k, K_, J_, val, shape, strides_k, strides_j, indices, h, op, result)


def _check_mul_dim(sizeR, sizeA, sizeB):
    """_check_mul!_dim_compatibility, src/operators_lazytensor.jl:695-703."""
    if sizeA[1] != sizeB[0]:
        raise DimensionMismatch("A and B dimensions do not match. Can't do `A*B`")
    if tuple(sizeR) != (sizeA[0],) + tuple(sizeB[1:]):
        raise DimensionMismatch("Output dimensions do not match A*B. Can't do `R.=A*B`")


def _gemm_puresparse_left(alpha, h: OLazyTensor, op, beta, result):
    """src/operators_lazytensor.jl:729-739."""
    if result is op:
        raise ArgumentError("output matrix must not be aliased with input matrix")
    _check_mul_dim(result.shape, h.shape, op.shape)
    _beta_prep(beta, result)
    shape, strides_j, strides_k = _get_shape_and_strides(h)
    N_k = len(h.shape_l)
    if op.ndim == 1:
        _gemm_recursive_lazy_dense(1, N_k, 1, 1, alpha * h.factor, shape, strides_k, strides_j, h.indices, h, op,
                                   result)
    else:  # rows of result gathered per column block: leaf is a row axpy
        _gemm_recursive_lazy_dense(1, N_k, 1, 1, alpha * h.factor, shape, strides_k, strides_j, h.indices, h,
                                   _RowView(op), _RowView(result))
    return result


def _gemm_puresparse_right(alpha, op, h: OLazyTensor, beta, result):
    """src/operators_lazytensor.jl:711-727 (op a vector = Bra, else a matrix)."""
    if result is op:
        raise ArgumentError("output matrix must not be aliased with input matrix")
    if op.ndim == 1:
        _check_mul_dim(result.shape, (h.shape[1], h.shape[0]), op.shape)
    else:
        _check_mul_dim(result.shape, op.shape, h.shape)
    _beta_prep(beta, result)
    shape, strides_j, strides_k = _get_shape_and_strides(h)
    N_k = len(h.shape_r)
    _gemm_recursive_dense_lazy(1, N_k, 1, 1, alpha * h.factor, shape, strides_k, strides_j, h.indices, h, op,
                               result)
    return result


class _RowView:
    """Index rows of a matrix so the recursive leaf `result[J] += val*op[K]`
    becomes the reference's row loop result[J, I] += val*op[K, I]."""

    def __init__(self, a):
        self.a = a

    def __getitem__(self, i):
        return self.a[i, :]

    def __setitem__(self, i, v):
        self.a[i, :] = v


def _is_pure_sparse(ops):
    """src/operators_lazytensor.jl:520."""
    return all((isinstance(o, OSparse) and not o.adj) or isinstance(o, OEye) for o in ops)


def _is_square_eye(o):
    """src/operators_lazytensor.jl:436-440."""
    return isinstance(o, OEye) and o.m == o.n


def _op_data(o, transform):
    if isinstance(o, OSparse):
        A = o.parent_dense()
        A = A.conj().T if o.adj else A
    elif isinstance(o, ODense):
        A = o.mat()
    elif isinstance(o, OEye):
        A = np.eye(o.m, o.n, dtype=np.complex128)
    else:
        raise TypeError(type(o))
    return A.T if transform == "transpose" else A


def _tpops_tuple(h: OLazyTensor, shift=0, transform=None):
    """src/operators_lazytensor.jl:523-537: (matrix, mode) pairs without
    square identities."""
    pairs = []
    for o, i in zip(h.operators, h.indices):
        if _is_square_eye(o):
            continue
        pairs.append((_op_data(o, transform), i + shift))
    return pairs


def _explicit_isometries(used_indices, shp_l, shp_r, shift=0):
    """src/operators_lazytensor.jl:491-514."""
    if tuple(shp_l) == tuple(shp_r):
        return []
    isos = []
    for i, (sl, sr) in enumerate(zip(shp_l, shp_r)):
        if sl != sr and (i + 1 + shift) not in used_indices:
            isos.append((np.eye(sl, sr, dtype=np.complex128), i + 1 + shift))
    return isos


def _tp_matmul_first(result, a, b, alpha, beta):
    """src/operators_lazytensor.jl:281-290 (column-major reshape)."""
    d_first = a.shape[1]
    br = b.reshape((d_first, -1), order="F")
    rr = result.reshape((a.shape[0], -1), order="F")
    new = alpha * (a @ br)
    if beta != 0:
        new = new + beta * rr
    rr[...] = new
    return result


def _tp_matmul_last(result, a, b, alpha, beta):
    """src/operators_lazytensor.jl:292-301."""
    d_last = a.shape[1]
    br = b.reshape((-1, d_last), order="F")
    rr = result.reshape((-1, a.shape[0]), order="F")
    new = alpha * (br @ a.T)
    if beta != 0:
        new = new + beta * rr
    rr[...] = new
    return result


def _tp_matmul_mid(result, a, loc, b, alpha, beta):
    """src/operators_lazytensor.jl:333-404: permute the mode to the front or
    back, GEMM, permute back."""
    sz1 = int(np.prod(b.shape[:loc - 1]))
    sz3 = int(np.prod(b.shape[loc:]))
    br = b.reshape((sz1, b.shape[loc - 1], sz3), order="F")
    rr = result.reshape((sz1, a.shape[0], sz3), order="F")
    move_left = sz1 < sz3
    perm = (1, 0, 2) if move_left else (0, 2, 1)
    br_p = np.asfortranarray(np.transpose(br, perm))
    rr_p = np.asfortranarray(np.transpose(rr, perm)) if beta != 0 else np.zeros(
        tuple(rr.shape[i] for i in perm), dtype=result.dtype, order="F")
    if move_left:
        _tp_matmul_first(rr_p, a, br_p, alpha, beta)
    else:
        _tp_matmul_last(rr_p, a, br_p, alpha, beta)
    rr[...] = np.transpose(rr_p, np.argsort(perm))
    return result


def _tp_matmul(result, a, loc, b, alpha, beta):
    """src/operators_lazytensor.jl:406-428."""
    if loc == 1:
        return _tp_matmul_first(result, a, b, alpha, beta)
    if loc == b.ndim:
        return _tp_matmul_last(result, a, b, alpha, beta)
    return _tp_matmul_mid(result, a, loc, b, alpha, beta)


def _tp_sum_matmul(result, ops, b, alpha, beta):
    """src/operators_lazytensor.jl:443-488 (ops already include isometries).
    result and b are column-major (Fortran-ordered) tensors."""
    n = len(ops)
    if n == 0:
        if beta == 0:
            result[...] = alpha * b
        else:
            result[...] = alpha * b + beta * result
        return result
    if n == 1:
        return _tp_matmul(result, ops[0][0], ops[0][1], b, alpha, beta)

    def tmp_like(arr, op, loc):
        shp = list(arr.shape)
        shp[loc - 1] = op.shape[0]
        return np.zeros(shp, dtype=np.complex128, order="F")

    tmp1 = tmp_like(b, ops[0][0], ops[0][1])
    _tp_matmul(tmp1, ops[0][0], ops[0][1], b, alpha, 0)
    for i in range(1, n - 1):
        tmp2 = tmp_like(tmp1, ops[i][0], ops[i][1])
        _tp_matmul(tmp2, ops[i][0], ops[i][1], tmp1, 1, 0)
        tmp1 = tmp2
    _tp_matmul(result, ops[-1][0], ops[-1][1], tmp1, 1, beta)
    return result


def _zero_op_mul(data, beta):
    """src/operators_lazysum.jl:179-186."""
    if beta == 0:
        data[...] = 0
    elif beta != 1:
        data *= beta
    return data


def _as_tensor(x, shape):
    return np.asfortranarray(x).reshape(shape, order="F")


def mul_lt_ket(result, a: OLazyTensor, b, alpha, beta):
    """mul!(::Ket, ::LazyTensor, ::Ket), src/operators_lazytensor.jl:539-557."""
    if alpha == 0:
        return _zero_op_mul(result, beta)
    if len(a.operators) > 0 and _is_pure_sparse(a.operators):
        return _gemm_puresparse_left(alpha, a, b, beta, result)
    bt = _as_tensor(b, a.shape_r)
    rt = np.asfortranarray(result.reshape(a.shape_l, order="F"))
    ops = _tpops_tuple(a) + _explicit_isometries(a.indices, a.shape_l, a.shape_r)
    _tp_sum_matmul(rt, ops, bt, alpha * a.factor, beta)
    result[...] = rt.reshape(-1, order="F")
    return result


def mul_lt_bra(result, a, b: OLazyTensor, alpha, beta):
    """mul!(::Bra, ::Bra, ::LazyTensor), src/operators_lazytensor.jl:559-574."""
    if alpha == 0:
        return _zero_op_mul(result, beta)
    if len(b.operators) > 0 and _is_pure_sparse(b.operators):
        return _gemm_puresparse_right(alpha, a, b, beta, result)
    at = _as_tensor(a, b.shape_l)
    rt = np.asfortranarray(result.reshape(b.shape_r, order="F"))
    ops = _tpops_tuple(b, transform="transpose") + _explicit_isometries(b.indices, b.shape_r, b.shape_l)
    _tp_sum_matmul(rt, ops, at, alpha * b.factor, beta)
    result[...] = rt.reshape(-1, order="F")
    return result


def mul_lt_op_left(result, a: OLazyTensor, b, alpha, beta):
    """mul!(::DenseOpType, ::LazyTensor, ::DenseOpType), src/operators_lazytensor.jl:576-591."""
    if alpha == 0:
        return _zero_op_mul(result, beta)
    if len(a.operators) > 0 and _is_pure_sparse(a.operators):
        return _gemm_puresparse_left(alpha, a, b, beta, result)
    ncol = b.shape[1]
    bt = _as_tensor(b, tuple(a.shape_r) + (ncol,))
    rt = np.asfortranarray(result).reshape(tuple(a.shape_l) + (result.shape[1],), order="F")
    ops = _tpops_tuple(a) + _explicit_isometries(a.indices, a.shape_l, a.shape_r)
    _tp_sum_matmul(rt, ops, bt, alpha * a.factor, beta)
    result[...] = rt.reshape(result.shape, order="F")
    return result


def mul_lt_op_right(result, a, b: OLazyTensor, alpha, beta):
    """mul!(::DenseOpType, ::DenseOpType, ::LazyTensor), src/operators_lazytensor.jl:593-609
    (the row index kept as one leading mode, shift = its count)."""
    if alpha == 0:
        return _zero_op_mul(result, beta)
    if len(b.operators) > 0 and _is_pure_sparse(b.operators):
        return _gemm_puresparse_right(alpha, a, b, beta, result)
    nrow = a.shape[0]
    at = _as_tensor(a, (nrow,) + tuple(b.shape_l))
    rt = np.asfortranarray(result).reshape((nrow,) + tuple(b.shape_r), order="F")
    shft = 1
    ops = _tpops_tuple(b, shift=shft, transform="transpose") + _explicit_isometries(
        tuple(i + shft for i in b.indices), b.shape_r, b.shape_l, shft)
    _tp_sum_matmul(rt, ops, at, alpha * b.factor, beta)
    result[...] = rt.reshape(result.shape, order="F")
    return result


# ------------------------------------------------------------ generic mul!

# ------------------------------------------- src/operators_lazyproduct.jl restated
def _lp_chain(result, lp, b, alpha, beta, mul, right):
    """mul! of a LazyProduct (src/operators_lazyproduct.jl:103-163): the first
    application takes the product's factor, the last (alpha, beta), the
    ones between (1, 0).  right=False runs the last factor first (Ket,
    Operator left); right=True the first factor first (Bra, Operator right)."""
    if alpha == 0:
        return _zero_op_mul(result, beta)
    ops = lp.operators
    if len(ops) == 1:
        return mul(result, ops[0], b, lp.factor * alpha, beta)
    order = list(ops) if right else list(reversed(ops))
    tmp = b
    for i, o in enumerate(order[:-1]):
        out_shape = _apply_shape(o, tmp, right)
        nxt = np.zeros(out_shape, dtype=np.complex128)
        mul(nxt, o, tmp, lp.factor if i == 0 else 1.0, 0.0)
        tmp = nxt
    return mul(result, order[-1], tmp, alpha, beta)


def _apply_shape(o, x, right):
    m, n = o.shape
    if x.ndim == 1:
        return (n,) if right else (m,)
    return (x.shape[0], n) if right else (m, x.shape[1])


def mul_ket(result, a, b, alpha=1.0, beta=0.0):
    """mul!(result::Ket, a, b::Ket, alpha, beta)."""
    if isinstance(a, OLazyProduct):
        return _lp_chain(result, a, b, alpha, beta, mul_ket, False)
    if isinstance(a, OLazySum):
        # src/operators_lazysum.jl:189-200
        if len(a.operators) == 0 or alpha == 0:
            _check_mul_dim(result.shape, a.shape, b.shape)
            return _zero_op_mul(result, beta)
        mul_ket(result, a.operators[0], b, alpha * a.factors[0], beta)
        for i in range(1, len(a.operators)):
            mul_ket(result, a.operators[i], b, alpha * a.factors[i], 1)
        return result
    if isinstance(a, OLazyTensor):
        return mul_lt_ket(result, a, b, alpha, beta)
    if isinstance(a, OSparse) and not a.adj:
        return gemv_sp(alpha, a, b, beta, result)  # src/operators_sparse.jl:201
    return _blas_mul(result, _mat(a), b, alpha, beta)  # src/operators_dense.jl:395


def mul_bra(result, a, b, alpha=1.0, beta=0.0):
    """mul!(result::Bra, a::Bra, b, alpha, beta)."""
    if isinstance(b, OLazyProduct):
        return _lp_chain(result, b, a, alpha, beta, lambda r, o, x, al, be: mul_bra(r, x, o, al, be), True)
    if isinstance(b, OLazySum):
        # src/operators_lazysum.jl:202-213
        if len(b.operators) == 0 or alpha == 0:
            _check_mul_dim(result.shape, (b.shape[1], b.shape[0]), a.shape)
            return _zero_op_mul(result, beta)
        mul_bra(result, a, b.operators[0], alpha * b.factors[0], beta)
        for i in range(1, len(b.operators)):
            mul_bra(result, a, b.operators[i], alpha * b.factors[i], 1)
        return result
    if isinstance(b, OLazyTensor):
        return mul_lt_bra(result, a, b, alpha, beta)
    if isinstance(b, OSparse) and not b.adj:
        return gemv_vec_sp(alpha, a, b, beta, result)  # src/operators_sparse.jl:202
    return _blas_mul(result, _mat(b).T, a, alpha, beta)  # src/operators_dense.jl:396


def mul_op_left(result, a, b, alpha=1.0, beta=0.0):
    """mul!(result::Operator, a, b::Operator, alpha, beta)."""
    if isinstance(a, OLazyProduct):
        return _lp_chain(result, a, b, alpha, beta, mul_op_left, False)
    if isinstance(a, OLazySum):
        # src/operators_lazysum.jl:215-226
        if len(a.operators) == 0 or alpha == 0:
            _check_mul_dim(result.shape, a.shape, b.shape)
            return _zero_op_mul(result, beta)
        mul_op_left(result, a.operators[0], b, alpha * a.factors[0], beta)
        for i in range(1, len(a.operators)):
            mul_op_left(result, a.operators[i], b, alpha * a.factors[i], 1)
        return result
    if isinstance(a, OLazyTensor):
        return mul_lt_op_left(result, a, b, alpha, beta)
    if isinstance(a, OSparse):
        return gemm_sp_left(alpha, a, b, beta, result)  # src/operators_sparse.jl:199
    return _blas_mul(result, _mat(a), b, alpha, beta)  # src/operators_dense.jl:394


def mul_op_right(result, a, b, alpha=1.0, beta=0.0):
    """mul!(result::Operator, a::Operator, b, alpha, beta)."""
    if isinstance(b, OLazyProduct):
        return _lp_chain(result, b, a, alpha, beta, lambda r, o, x, al, be: mul_op_right(r, x, o, al, be), True)
    if isinstance(b, OLazySum):
        # src/operators_lazysum.jl:227-238
        if len(b.operators) == 0 or alpha == 0:
            _check_mul_dim(result.shape, a.shape, b.shape)
            return _zero_op_mul(result, beta)
        mul_op_right(result, a, b.operators[0], alpha * b.factors[0], beta)
        for i in range(1, len(b.operators)):
            mul_op_right(result, a, b.operators[i], alpha * b.factors[i], 1)
        return result
    if isinstance(b, OLazyTensor):
        return mul_lt_op_right(result, a, b, alpha, beta)
    if isinstance(b, OSparse):
        return gemm_sp_right(alpha, a, b, beta, result)  # src/operators_sparse.jl:200
    return _blas_mul(result, a, _mat(b), alpha, beta)  # src/operators_dense.jl:394


def _blas_mul(result, A, B, alpha, beta):
    """LinearAlgebra.mul!(C, A, B, alpha, beta) (BLAS zgemm/zgemv semantics:
    beta == 0 discards C)."""
    if A.shape[1] != B.shape[0] or A.shape[0] != result.shape[0]:
        raise DimensionMismatch()
    new = alpha * (A @ B)
    if beta != 0:
        new = new + beta * result
    result[...] = new
    return result


# --------------------------------------------------------------- reductions
def norm(x):
    """src/states.jl:89."""
    return float(np.sqrt(np.sum(np.abs(np.asarray(x, dtype=np.complex128)) ** 2)))


def dot(x, y):
    return complex(np.vdot(np.asarray(x, dtype=np.complex128), np.asarray(y, dtype=np.complex128)))


def trace(A):
    """src/operators_dense.jl:189."""
    return complex(np.trace(np.asarray(A, dtype=np.complex128)))


def expect_ket(op, psi):
    """expect(op, ket) = dot(psi, op*psi), src/operators.jl:119."""
    r = np.zeros(op.shape[0], dtype=np.complex128)
    mul_ket(r, op, np.asarray(psi, dtype=np.complex128), 1, 0)
    return dot(psi, r)


def expect_op(op, rho):
    """expect(op, rho) = sum_ij op_ij rho_ji, src/operators_dense.jl:223-230."""
    r = np.zeros((op.shape[0], rho.shape[1]), dtype=np.complex128)
    mul_op_left(r, op, np.asarray(rho, dtype=np.complex128), 1, 0)
    return trace(r)


def rel_frobenius(x, ref) -> float:
    """Relative Frobenius error ||x - ref|| / ||ref||."""
    x = np.asarray(x, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    d = np.linalg.norm((x - ref).ravel())
    n = np.linalg.norm(ref.ravel())
    return float(d / n) if n > 0 else float(d)


# ---------------------------------------------- src/superoperators.jl restated
def vec(rho):
    """Column-major vec, the reshape of src/superoperators.jl:83."""
    return np.asarray(rho).reshape(-1, order="F")


def superop_times_op(S, rho):
    """SuperOperator * Operator (src/superoperators.jl:82-85)."""
    n_out = int(round(np.sqrt(S.shape[0])))
    return (S @ vec(rho)).reshape(n_out, -1, order="F")


def spre(A):
    """tensor(A, identity) = kron(1, A) (src/superoperators.jl:126-131)."""
    return np.kron(np.eye(A.shape[1]), A)


def spost(A):
    """kron(transpose(A), 1) (src/superoperators.jl:146-151)."""
    return np.kron(A.T, np.eye(A.shape[0]))


def sprepost(A, B):
    """kron(transpose(B), A) (src/superoperators.jl:165)."""
    return np.kron(B.T, A)


def liouvillian(H, J, rates=None, Jdagger=None):
    """src/superoperators.jl:205-222 with dense matrices."""
    Jd = Jdagger if Jdagger is not None else [j.conj().T for j in J]
    rates = np.ones(len(J)) if rates is None else np.asarray(rates)
    L = spre(-1j * H) + spost(1j * H)
    if rates.ndim == 2:
        pairs = [(i, j, rates[i, j]) for i in range(len(J)) for j in range(len(J))]
    else:
        pairs = [(i, i, rates[i]) for i in range(len(J))]
    for i, j, r in pairs:
        jdj = r / 2 * Jd[j] @ J[i]
        L = L - (spre(jdj) + spost(jdj))
        L = L + spre(r * J[i]) @ spost(Jd[j])
    return L
