"""Seeded synthetic operators and states for the benchmark configurations.

Every generator takes a seed and uses numpy's PCG64; "complex Gaussian"
means independent N(0, 1) real and imaginary parts.  Host arrays are
returned next to the device objects so a checker can rebuild the same
inputs.

  spin_chain(L)           10-site spin-1/2 chain: LazySum of 9 sigma_x sigma_x
                          bonds + 10 sigma_z fields, coefficients U[-1, 1]
  fock_fock_spin()        a^dag a (x) 1 (x) 1 + 1 (x) b^dag b (x) 1
                          + g (a (x) b^dag (x) sigma_+ + h.c.) on
                          Fock(39) (x) Fock(39) (x) Spin(3/2), g ~ U[0, 1]
  dense_lazytensor()      LazyTensor of 3 random Hermitian dense factors
                          (16, 32, 16)
  tfim_chain(L)           transverse-field Ising chain, LazySum of
                          J_i sigma_z sigma_z + h_i sigma_x, J, h ~ U[-1, 1]
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import core, fock, spin


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def complex_gaussian(g: np.random.Generator, shape, dtype=np.complex128) -> np.ndarray:
    return (g.standard_normal(shape) + 1j * g.standard_normal(shape)).astype(dtype)


def normalized_ket(g, n, dtype=np.complex128) -> np.ndarray:
    v = complex_gaussian(g, (n,), np.complex128)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def ginibre_rho(g, n) -> np.ndarray:
    """rho = G G^dag / tr(G G^dag) (host; fine up to n ~ 2000)."""
    G = complex_gaussian(g, (n, n))
    r = G @ G.conj().T
    return r / np.trace(r)


def ginibre_rho_device(seed: int, n: int, dtype=np.complex128) -> core.Operator:
    """rho = G G^dag / tr on the device (one GEMM), for full-size states."""
    g = rng(seed)
    G = complex_gaussian(g, (n, n))
    b = core.GenericBasis(n)
    Gd = core.Operator(b, b, G)
    Gh = core.Operator(b, b, np.ascontiguousarray(G.conj().T))
    del G
    rho = Gd * Gh
    core.normalize_(rho)
    if np.dtype(dtype) != np.complex128:
        host = rho.to_numpy().astype(dtype)
        rho = core.Operator(b, b, host)
    return rho


# ------------------------------------------------------------- spin chain
@dataclass
class LazySumSpec:
    """A LazySum of LazyTensors described by host data."""
    basis: core.Basis
    terms: list = field(default_factory=list)  # (coef, indices(1-based), [scipy csc factors])

    def build(self) -> core.LazySum:
        ops, coefs = [], []
        for c, idx, facs in self.terms:
            fops = tuple(core.SparseOperator(self.basis.bases[i - 1], self.basis.bases[i - 1], f)
                         for i, f in zip(idx, facs))
            ops.append(core.LazyTensor(self.basis, self.basis, tuple(idx), fops, 1.0))
            coefs.append(c)
        return core.LazySum(self.basis, self.basis, coefs, tuple(ops))


def spin_chain(L: int = 10, seed: int = 1) -> LazySumSpec:
    """H = sum_i J_i sx_i sx_{i+1} + sum_i h_i sz_i (bonds first, then fields)."""
    g = rng(seed)
    b1 = core.SpinBasis(0.5)
    basis = core.tensor(*([b1] * L))
    sx = spin.sigmax(b1).csc
    sz = spin.sigmaz(b1).csc
    J = g.uniform(-1, 1, L - 1)
    h = g.uniform(-1, 1, L)
    spec = LazySumSpec(basis)
    for i in range(L - 1):
        spec.terms.append((complex(J[i]), (i + 1, i + 2), [sx, sx]))
    for i in range(L):
        spec.terms.append((complex(h[i]), (i + 1,), [sz]))
    return spec


def tfim_chain(L: int, seed: int) -> LazySumSpec:
    """H = sum_i J_i sz_i sz_{i+1} + sum_i h_i sx_i, J, h ~ U[-1, 1]."""
    g = rng(seed)
    b1 = core.SpinBasis(0.5)
    basis = core.tensor(*([b1] * L))
    sx = spin.sigmax(b1).csc
    sz = spin.sigmaz(b1).csc
    J = g.uniform(-1, 1, L - 1)
    h = g.uniform(-1, 1, L)
    spec = LazySumSpec(basis)
    for i in range(L - 1):
        spec.terms.append((complex(J[i]), (i + 1, i + 2), [sz, sz]))
    for i in range(L):
        spec.terms.append((complex(h[i]), (i + 1,), [sx]))
    return spec


# --------------------------------------------------- Fock x Fock x Spin(3/2)
@dataclass
class FockSpinSpec:
    basis: core.Basis
    g: float
    H: sp.csc_matrix                 # full sparse matrix, zeros dropped
    factors: dict                    # name -> scipy csc of each site operator


def fock_fock_spin(N: int = 39, spin_j=1.5, seed: int = 1) -> FockSpinSpec:
    """Full SparseOperator of the coupled Fock-Fock-spin Hamiltonian.

    tensor(a, b, c) = kron(c, kron(b, a)) (site 1 fastest)."""
    gen = rng(seed)
    gcoup = float(gen.uniform(0, 1))
    fa = core.FockBasis(N)
    fb = core.FockBasis(N)
    s = core.SpinBasis(spin_j)
    basis = core.tensor(fa, fb, s)
    a = fock.destroy(fa).csc
    ad = fock.create(fa).csc
    b = fock.destroy(fb).csc
    bd = fock.create(fb).csc
    sp_ = spin.sigmap(s).csc
    sm_ = spin.sigmam(s).csc
    na = fock.number(fa).csc
    nb = fock.number(fb).csc
    Ia = sp.identity(len(fa), dtype=np.complex128, format="csc")
    Ib = sp.identity(len(fb), dtype=np.complex128, format="csc")
    Is = sp.identity(len(s), dtype=np.complex128, format="csc")

    def t3(x, y, z):
        return sp.kron(z, sp.kron(y, x, format="csc"), format="csc")

    H = t3(na, Ib, Is) + t3(Ia, nb, Is) + gcoup * (t3(a, bd, sp_) + t3(ad, b, sm_))
    H = sp.csc_matrix(H)
    H.eliminate_zeros()
    H.sort_indices()
    return FockSpinSpec(basis, gcoup, H, dict(a=a, ad=ad, b=b, bd=bd, sp=sp_, sm=sm_, na=na, nb=nb))


def fock_fock_spin_lazy(spec: FockSpinSpec) -> core.LazySum:
    """The same Hamiltonian as a LazySum of pure-sparse LazyTensors."""
    B = spec.basis
    fa, fb, s = B.bases
    f = spec.factors

    def so(bs, m):
        return core.SparseOperator(bs, bs, m)

    t1 = core.LazyTensor(B, B, (1,), (so(fa, f["na"]),))
    t2 = core.LazyTensor(B, B, (2,), (so(fb, f["nb"]),))
    t3 = core.LazyTensor(B, B, (1, 2, 3), (so(fa, f["a"]), so(fb, f["bd"]), so(s, f["sp"])))
    t4 = core.LazyTensor(B, B, (1, 2, 3), (so(fa, f["ad"]), so(fb, f["b"]), so(s, f["sm"])))
    return core.LazySum(B, B, [1.0, 1.0, spec.g, spec.g], (t1, t2, t3, t4))


# ------------------------------------------------------- dense LazyTensor
def hermitian_factor(g, d: int) -> np.ndarray:
    X = complex_gaussian(g, (d, d))
    return (X + X.conj().T) / 2


def dense_lazytensor_factors(dims=(16, 32, 16), seed: int = 1):
    g = rng(seed)
    return [hermitian_factor(g, d) for d in dims]


def dense_lazytensor(dims=(16, 32, 16), seed: int = 1, factors=None) -> core.LazyTensor:
    facs = factors if factors is not None else dense_lazytensor_factors(dims, seed)
    bases = [core.GenericBasis(d) for d in dims]
    B = core.tensor(*bases)
    ops = tuple(core.DenseOperator(b, b, A) for b, A in zip(bases, facs))
    return core.LazyTensor(B, B, tuple(range(1, len(dims) + 1)), ops)


# ----------------------------------------------------------- batch of chains
def tfim_batch_lengths(n: int = 256, lmin: int = 8, lmax: int = 20, seed: int = 1) -> np.ndarray:
    return rng(seed).integers(lmin, lmax + 1, n)
