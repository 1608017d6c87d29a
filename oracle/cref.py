"""ctypes binding of oracle/cpu_ref.c (built into oracle/lib/libqob_cpu_ref.so).

TEST INFRASTRUCTURE ONLY: the full-size checker and the CPU baseline used by
tests/ and bench.py's cpu_baseline leg; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB = None
_vp, _i64, _dbl, _int = C.c_void_p, C.c_int64, C.c_double, C.c_int


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(str(Path(__file__).resolve().parent / "lib" / "libqob_cpu_ref.so"))
        L.qobref_gemm_sp_dense.argtypes = [_dbl, _dbl, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _dbl, _dbl, _vp]
        L.qobref_gemm_dense_sp.argtypes = [_dbl, _dbl, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp]
        L.qobref_gemm_sp_dense_omp.argtypes = L.qobref_gemm_sp_dense.argtypes + [_int]
        L.qobref_gemm_dense_sp_omp.argtypes = L.qobref_gemm_dense_sp.argtypes + [_int]
        L.qobref_lazysum_ket.argtypes = [_dbl, _dbl, _int, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _dbl, _dbl, _vp]
        L.qobref_lazysum_ket_flat.argtypes = L.qobref_lazysum_ket.argtypes
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_vp)


class LazySumKet:
    """A LazySum of all-sparse square LazyTensors packed for qobref_lazysum_ket.

    terms: (coef, sites (1-based), [scipy sparse factors]) with the LazyTensor
    factor folded into coef."""

    def __init__(self, shape, terms):
        self.shape = np.asarray(shape, dtype=np.int64)
        self.D = int(np.prod(self.shape))
        coef, term_off, fsite, fdim, cp_off, nz_off = [], [0], [], [], [], []
        cps, rvs, nzs = [], [], []
        ncp = nnz = 0
        for c, sites, facs in terms:
            coef += [complex(c).real, complex(c).imag]
            for s, f in zip(sites, facs):
                f = f.tocsc()
                f.sort_indices()
                fsite.append(int(s))
                fdim.append(f.shape[1])
                cp_off.append(ncp)
                nz_off.append(nnz)
                cps.append(f.indptr.astype(np.int64) + 1)
                rvs.append(f.indices.astype(np.int64) + 1)
                nzs.append(f.data.astype(np.complex128))
                ncp += f.shape[1] + 1
                nnz += f.nnz
            term_off.append(len(fsite))
        self.nterms = len(terms)
        self.coef = np.asarray(coef, dtype=np.float64)
        self.term_off = np.asarray(term_off, dtype=np.int64)
        self.fsite = np.asarray(fsite, dtype=np.int64)
        self.fdim = np.asarray(fdim, dtype=np.int64)
        self.cp_off = np.asarray(cp_off, dtype=np.int64)
        self.nz_off = np.asarray(nz_off, dtype=np.int64)
        self.colptr = np.concatenate(cps) if cps else np.zeros(1, np.int64)
        self.rowval = np.concatenate(rvs) if rvs else np.zeros(1, np.int64)
        self.nzval = np.concatenate(nzs) if nzs else np.zeros(1, np.complex128)

    def mul(self, result, psi, alpha=1.0, beta=0.0):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert result.dtype == np.complex128 and result.flags.c_contiguous and result.size == self.D
        a, b = complex(alpha), complex(beta)
        if getattr(self, "_args", None) is None:
            # the operator's arrays never change: their pointers are built once,
            # so a call spends its time in C (threads calling it run in parallel)
            self._args = (len(self.shape), _p(self.shape), self.nterms, _p(self.coef), _p(self.term_off),
                          _p(self.fsite), _p(self.fdim), _p(self.cp_off), _p(self.nz_off), _p(self.colptr),
                          _p(self.rowval), _p(self.nzval))
        lib().qobref_lazysum_ket(a.real, a.imag, *self._args, psi.ctypes.data, b.real, b.imag, result.ctypes.data)
        return result

    def mul_flat(self, result, psi, alpha=1.0, beta=0.0):
        """The same product through qobref_lazysum_ket_flat (the recursion
        unrolled per term; equal to mul() bit for bit, many times faster)."""
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert result.dtype == np.complex128 and result.flags.c_contiguous and result.size == self.D
        a, b = complex(alpha), complex(beta)
        if getattr(self, "_args", None) is None:
            self._args = (len(self.shape), _p(self.shape), self.nterms, _p(self.coef), _p(self.term_off),
                          _p(self.fsite), _p(self.fdim), _p(self.cp_off), _p(self.nz_off), _p(self.colptr),
                          _p(self.rowval), _p(self.nzval))
        lib().qobref_lazysum_ket_flat(a.real, a.imag, *self._args, psi.ctypes.data, b.real, b.imag,
                                      result.ctypes.data)
        return result
