"""Superoperators (src/superoperators.jl): SuperOperator, spre, spost,
sprepost, liouvillian, and SuperOperator * Operator on the device.

A superoperator acts on vec(rho), the column-major N*N vector of a density
matrix (src/superoperators.jl:82-85 reshapes b.data to length N^2).  The
device buffer of an Operator is exactly vec(rho), so S * rho is the sparse
(or dense) Ket path of the library at dimension N^2 (SELL gather / GEMV);
construction (Kronecker products) stays on the host, as in the reference.
Column-major vec gives vec(A B C) = (C^T kron A) vec(B), the identity the
reference's spre / spost / sprepost are built on.
"""
from __future__ import annotations

import numbers
from typing import Sequence

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, DimensionMismatch, IncompatibleBases, MethodError
from .core import (DenseOperator, DeviceArray, GenericBasis, Operator, SparseOperator, _Adjoint,
                   dagger as _dagger, mul_ as _mul_)


def _data_of(op):
    """Host matrix of a data operator (dense or sparse)."""
    if isinstance(op, _Adjoint):
        d = _data_of(op.parent)
        return d.conj().T
    if isinstance(op, SparseOperator):
        return op.csc.copy()
    if isinstance(op, Operator):
        return np.asarray(op.to_numpy())
    raise MethodError(f"superoperators need data operators, got {type(op).__name__}")


class SuperOperator:
    """SuperOperator(basis_l, basis_r, data) (src/superoperators.jl:9-25);
    basis_l and basis_r are pairs of bases, data a scipy sparse matrix or a
    dense array of size (|l1||l2|, |r1||r2|)."""

    def __init__(self, basis_l, basis_r=None, data=None):
        if data is None:
            basis_r, data = basis_l, basis_r
        basis_l, basis_r = tuple(basis_l), tuple(basis_r)
        if (len(basis_l) != 2 or len(basis_r) != 2 or len(basis_l[0]) * len(basis_l[1]) != data.shape[0]
                or len(basis_r[0]) * len(basis_r[1]) != data.shape[1]):
            raise DimensionMismatch(f"Tried to assign data of size {data.shape} to Hilbert spaces of sizes "
                                    f"{tuple(len(b) for b in basis_l)}, {tuple(len(b) for b in basis_r)}")
        self.basis_l, self.basis_r = basis_l, basis_r
        self.data = sp.csc_matrix(data, dtype=np.complex128) if sp.issparse(data) else np.asarray(data, np.complex128)
        self._dev = None

    @property
    def shape(self):
        return self.data.shape

    def is_sparse(self) -> bool:
        return sp.issparse(self.data)

    # host algebra (src/superoperators.jl:87-101) -----------------------
    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return SuperOperator(self.basis_l, self.basis_r, self.data * other)
        if isinstance(other, SuperOperator):
            if other.basis_l != self.basis_r:
                raise IncompatibleBases("SuperOperator product of incompatible bases")
            return SuperOperator(self.basis_l, other.basis_r, self.data @ other.data)
        if isinstance(other, Operator):
            return apply(self, other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return self * other
        return NotImplemented

    def __truediv__(self, b):
        return SuperOperator(self.basis_l, self.basis_r, self.data / b)

    def __add__(self, other):
        if not isinstance(other, SuperOperator) or other.basis_l != self.basis_l or other.basis_r != self.basis_r:
            raise IncompatibleBases("SuperOperator sum of incompatible bases")
        return SuperOperator(self.basis_l, self.basis_r, self.data + other.data)

    def __sub__(self, other):
        if not isinstance(other, SuperOperator) or other.basis_l != self.basis_l or other.basis_r != self.basis_r:
            raise IncompatibleBases("SuperOperator difference of incompatible bases")
        return SuperOperator(self.basis_l, self.basis_r, self.data - other.data)

    def __neg__(self):
        return SuperOperator(self.basis_l, self.basis_r, -self.data)

    def dagger(self) -> "SuperOperator":
        d = self.data.conj().T
        return SuperOperator(self.basis_r, self.basis_l, sp.csc_matrix(d) if sp.issparse(d) else np.ascontiguousarray(d))

    # device ----------------------------------------------------------------
    def _device_op(self):
        """The data as a library operator on GenericBasis(N^2) spaces."""
        if self._dev is None:
            bl = GenericBasis(self.data.shape[0])
            br = GenericBasis(self.data.shape[1])
            self._dev = (SparseOperator(bl, br, self.data) if self.is_sparse()
                         else DenseOperator(bl, br, self.data))
        return self._dev


def apply(S: SuperOperator, rho: Operator, out: Operator | None = None, alpha=1.0, beta=0.0) -> Operator:
    """S * rho (src/superoperators.jl:82-85) on the device: out = alpha S vec(rho) + beta out."""
    if rho.basis_l != S.basis_r[0] or rho.basis_r != S.basis_r[1]:
        raise IncompatibleBases("SuperOperator and Operator bases differ")
    from .core import Ket
    op = S._device_op()
    if out is None:
        out = Operator(S.basis_l[0], S.basis_l[1], DeviceArray(S.data.shape[0], rho.dtype,
                                                               (len(S.basis_l[0]), len(S.basis_l[1]))))
    vin = Ket(op.basis_r, rho.data)
    vout = Ket(op.basis_l, out.data)
    _mul_(vout, op, vin, alpha, beta)
    return out


def _kron(a, b):
    """Kronecker product, dense when either factor is dense."""
    if sp.issparse(a) and sp.issparse(b):
        return sp.kron(a, b, format="csc")
    return np.kron(a.toarray() if sp.issparse(a) else a, b.toarray() if sp.issparse(b) else b)


def _eye_like(A, n):
    return sp.identity(n, dtype=np.complex128, format="csc") if sp.issparse(A) else np.eye(n, dtype=np.complex128)


def spre(op) -> SuperOperator:
    """spre(A) B = A B (src/superoperators.jl:126-131): kron(1, A)."""
    if op.basis_l != op.basis_r:
        raise ArgumentError("It's not clear what spre of a non-square operator should be.")
    A = _data_of(op)
    return SuperOperator((op.basis_l, op.basis_l), (op.basis_r, op.basis_r), _kron(_eye_like(A, A.shape[1]), A))


def spost(op) -> SuperOperator:
    """spost(A) B = B A (src/superoperators.jl:146-151): kron(A^T, 1)."""
    if op.basis_l != op.basis_r:
        raise ArgumentError("It's not clear what spost of a non-square operator should be.")
    A = _data_of(op)
    return SuperOperator((op.basis_r, op.basis_r), (op.basis_l, op.basis_l), _kron(A.T, _eye_like(A, A.shape[0])))


def sprepost(A, B) -> SuperOperator:
    """sprepost(A, B) C = A C B (src/superoperators.jl:165): kron(B^T, A)."""
    return SuperOperator((A.basis_l, B.basis_r), (A.basis_r, B.basis_l), _kron(_data_of(B).T, _data_of(A)))


def liouvillian(H, J: Sequence = (), rates=None, Jdagger=None) -> SuperOperator:
    """liouvillian(H, J; rates, Jdagger) (src/superoperators.jl:205-222):
    S rho = -i [H, rho] + sum_ij rates_ij (J_i rho J_j^dag - 1/2 {J_j^dag J_i, rho})."""
    J = list(J)
    Jd = list(Jdagger) if Jdagger is not None else [_dagger(j) for j in J]
    if len(J) != len(Jd):
        raise ArgumentError("J and Jdagger have different lengths")
    rates = np.ones(len(J)) if rates is None else np.asarray(rates)
    Hd = sp.csc_matrix(_data_of(H))
    n = Hd.shape[0]
    I = sp.identity(n, format="csc")

    def pre(A):
        return sp.kron(I, A, format="csc")

    def post(A):
        return sp.kron(sp.csc_matrix(A).T, I, format="csc")
    L = pre(-1j * Hd) + post(1j * Hd)
    Jm = [sp.csc_matrix(_data_of(j)) for j in J]
    Jdm = [sp.csc_matrix(_data_of(j)) for j in Jd]
    pairs = ([(i, i, rates[i]) for i in range(len(J))] if rates.ndim == 1 else
             [(i, j, rates[i, j]) for i in range(len(J)) for j in range(len(J))])
    for i, j, r in pairs:
        jdj = (r / 2) * (Jdm[j] @ Jm[i])
        L = L - (pre(jdj) + post(jdj))
        L = L + pre(r * Jm[i]) @ post(Jdm[j])
    L = sp.csc_matrix(L)
    L.eliminate_zeros()
    return SuperOperator((H.basis_l, H.basis_l), (H.basis_r, H.basis_r), L)


def identitysuperoperator(b) -> SuperOperator:
    n = len(b)
    return SuperOperator((b, b), (b, b), sp.identity(n * n, dtype=np.complex128, format="csc"))
