"""Fock-space operators as SparseOperators (src/fock.jl:8-44).

  number : diag(offset .. N)
  destroy: superdiagonal sqrt(n), n = offset+1 .. N
  create : subdiagonal   sqrt(n), n = offset+1 .. N
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .core import FockBasis, SparseOperator


def number(b: FockBasis) -> SparseOperator:
    d = np.arange(b.offset, b.N + 1, dtype=np.float64).astype(np.complex128)
    return SparseOperator(b, b, sp.diags([d], [0], format="csc"))


def destroy(b: FockBasis) -> SparseOperator:
    d = np.sqrt(np.arange(b.offset + 1, b.N + 1, dtype=np.float64)).astype(np.complex128)
    n = len(b)
    return SparseOperator(b, b, sp.diags([d], [1], shape=(n, n), format="csc"))


def create(b: FockBasis) -> SparseOperator:
    d = np.sqrt(np.arange(b.offset + 1, b.N + 1, dtype=np.float64)).astype(np.complex128)
    n = len(b)
    return SparseOperator(b, b, sp.diags([d], [-1], shape=(n, n), format="csc"))
