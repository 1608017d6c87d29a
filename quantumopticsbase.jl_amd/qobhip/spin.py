"""Spin operators as SparseOperators (src/spin.jl:14-85).

Basis order m = s, s-1, ..., -s.  Matrix elements:
  sigmax: off-diagonals sqrt((s+1)*2a - a(a+1)), a = 1..N-1, both sides
  sigmay: i times the same on the subdiagonal, -i on the superdiagonal
  sigmaz: diag(2m)
  sigmap: superdiagonal sqrt(s(s+1) - m(m+1)), m = s-1 .. -s
  sigmam: subdiagonal sqrt(s(s+1) - m(m-1)), m = s .. -s+1
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .core import SparseOperator, SpinBasis


def _ms(b: SpinBasis):
    s = float(b.spinnumber)
    return s, [s - k for k in range(len(b))]


def sigmax(b: SpinBasis) -> SparseOperator:
    s, _ = _ms(b)
    N = len(b)
    d = np.array([np.sqrt((s + 1) * 2 * a - a * (a + 1)) for a in range(1, N)], dtype=np.complex128)
    return SparseOperator(b, b, sp.diags([d, d], [1, -1], shape=(N, N), format="csc"))


def sigmay(b: SpinBasis) -> SparseOperator:
    s, _ = _ms(b)
    N = len(b)
    d = np.array([1j * np.sqrt((s + 1) * 2 * a - a * (a + 1)) for a in range(1, N)], dtype=np.complex128)
    return SparseOperator(b, b, sp.diags([d, -d], [-1, 1], shape=(N, N), format="csc"))


def sigmaz(b: SpinBasis) -> SparseOperator:
    _, ms = _ms(b)
    return SparseOperator(b, b, sp.diags([np.array([2 * m for m in ms], dtype=np.complex128)], [0], format="csc"))


def sigmap(b: SpinBasis) -> SparseOperator:
    s, ms = _ms(b)
    S = (s + 1) * s
    d = np.array([np.sqrt(S - m * (m + 1)) for m in ms[1:]], dtype=np.complex128)
    N = len(b)
    return SparseOperator(b, b, sp.diags([d], [1], shape=(N, N), format="csc"))


def sigmam(b: SpinBasis) -> SparseOperator:
    s, ms = _ms(b)
    S = (s + 1) * s
    d = np.array([np.sqrt(S - m * (m - 1)) for m in ms[:-1]], dtype=np.complex128)
    N = len(b)
    return SparseOperator(b, b, sp.diags([d], [-1], shape=(N, N), format="csc"))
