#!/usr/bin/env python3
"""Writes tests/golden/golden.npz: small input/output vectors of mul! computed
by dense linear algebra only.

The reference (Julia) has no golden vectors and cannot run here, so these
fixtures are made independently of both the oracle's restated loops and the
library: every operator is built as a dense matrix from the closed forms of
the reference's operator definitions, and composite operators by Kronecker
products in the reference's site order (site 1 fastest: tensor(a, b) is
kron(b, a), src/operators_dense.jl:134).  The outputs are plain matrix
products: result = alpha * op * state + beta * result (Ket, Operator left),
result = alpha * state * op + beta * result (Bra, Operator right).

Closed forms used (basis order m = s, s-1, ..., -s; Fock n = 0 .. N):
  sigmax(1/2) = [[0,1],[1,0]], sigmaz(1/2) = diag(1,-1)     (src/spin.jl)
  sigmap(s)[m+1, m] = sqrt(s(s+1) - m(m+1))                 (src/spin.jl:65-71)
  destroy(N)[n-1, n] = sqrt(n), create = destroy^T          (src/fock.jl:22-44)
  number(N) = diag(0..N)

Run from the repo root: python tests/golden/make_golden.py
"""
from pathlib import Path

import numpy as np


def kron_sites(mats):
    """Operator on sites 1..n, site 1 fastest."""
    out = np.ones((1, 1), complex)
    for A in mats:
        out = np.kron(A, out)
    return out


def cg(g, *shape):
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


def main():
    g = np.random.Generator(np.random.PCG64(20261202))
    F = {}
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.array([[1, 0], [0, -1]], complex)
    I2 = np.eye(2)

    # 5-site chain: H = sum J_i sx_i sx_{i+1} + sum h_i sz_i
    L = 5
    J, h = g.uniform(-1, 1, L - 1), g.uniform(-1, 1, L)
    H = np.zeros((2 ** L, 2 ** L), complex)
    for i in range(L - 1):
        H += J[i] * kron_sites([sx if k in (i, i + 1) else I2 for k in range(L)])
    for i in range(L):
        H += h[i] * kron_sites([sz if k == i else I2 for k in range(L)])
    D = 2 ** L
    psi, r0 = cg(g, D), cg(g, D)
    rho, R0 = cg(g, D, D), cg(g, D, D)
    al, be = 0.6 - 0.3j, 0.25 + 0.5j
    F.update(chain5_J=J, chain5_h=h, chain5_psi=psi, chain5_r0=r0, chain5_rho=rho, chain5_R0=R0,
             chain5_ab=np.array([al, be]),
             chain5_ket=al * H @ psi + be * r0, chain5_bra=al * psi @ H + be * r0,
             chain5_left=al * H @ rho + be * R0, chain5_right=al * rho @ H + be * R0)

    # Fock(3) x Fock(2) x Spin(1):  a^dag a + b^dag b + g (a b^dag s+ + h.c.)
    def destroy(N):
        return np.diag(np.sqrt(np.arange(1, N + 1)), 1).astype(complex)

    def sigmap(s2):  # spin s = s2/2
        s = s2 / 2
        ms = [s - k for k in range(s2 + 1)]
        d = [np.sqrt(s * (s + 1) - m * (m + 1)) for m in ms[1:]]
        return np.diag(d, 1).astype(complex)

    a, b, sp_ = destroy(3), destroy(2), sigmap(2)
    na, nb = a.conj().T @ a, b.conj().T @ b
    Ia, Ib, Is = np.eye(4), np.eye(3), np.eye(3)
    gc = float(g.uniform(0, 1))
    Hf = kron_sites([na, Ib, Is]) + kron_sites([Ia, nb, Is]) \
        + gc * (kron_sites([a, b.conj().T, sp_]) + kron_sites([a.conj().T, b, sp_.conj().T]))
    n = Hf.shape[0]
    rho, R0 = cg(g, n, n), cg(g, n, n)
    al, be = -0.4 + 0.2j, 0.0
    F.update(fock_g=np.array([gc]), fock_rho=rho, fock_R0=R0, fock_ab=np.array([al, be]),
             fock_left=al * Hf @ rho, fock_right=al * rho @ Hf)

    # dense LazyTensor on sites (1, 2, 3) of dims (2, 3, 4), scalar factor f
    dims = (2, 3, 4)
    facs = []
    for d in dims:
        X = cg(g, d, d)
        facs.append((X + X.conj().T) / 2)
    f = 0.7 + 0.1j
    Hl = f * kron_sites(facs)
    n = Hl.shape[0]
    rho, R0, psi = cg(g, n, n), cg(g, n, n), cg(g, n)
    al, be = 1.0, -0.5
    F.update(lt_f=np.array([f]), lt_A1=facs[0], lt_A2=facs[1], lt_A3=facs[2], lt_rho=rho, lt_R0=R0, lt_psi=psi,
             lt_ab=np.array([al, be]),
             lt_left=al * Hl @ rho + be * R0, lt_right=al * rho @ Hl + be * R0, lt_ket=Hl @ psi)

    # non-square site 2 (left dim 3, right dim 4) becomes an isometry: one
    # factor on site 1, site 2 acts as the 3 x 4 identity block
    A = cg(g, 2, 2)
    E = np.eye(3, 4)
    psi = cg(g, 8)
    F.update(iso_A=A, iso_psi=psi, iso_ket=0.5 * kron_sites([A, E]) @ psi)

    out = Path(__file__).resolve().parent / "golden.npz"
    np.savez_compressed(out, **F)
    print(f"wrote {out} ({out.stat().st_size} bytes, {len(F)} arrays)")


if __name__ == "__main__":
    main()
