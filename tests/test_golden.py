"""The oracle against the golden fixtures (dense linear algebra from closed
forms, tests/golden/make_golden.py)."""
import numpy as np

import golden_ops as GO
import qob_oracle as O

TOL = 1e-12


def test_golden_chain_ket_bra_rho():
    G = GO.load()
    L, terms = GO.chain_terms(G)
    sh = (2,) * L
    D = 2 ** L
    H = O.OLazySum(D, D, [c for c, _, _ in terms],
                   [O.OLazyTensor(sh, sh, s, tuple(O.OSparse.from_scipy(f) for f in fs), 1.0) for _, s, fs in terms])
    al, be = G["chain5_ab"]
    r = G["chain5_r0"].copy()
    O.mul_ket(r, H, G["chain5_psi"], al, be)
    assert O.rel_frobenius(r, G["chain5_ket"]) < TOL
    r = G["chain5_r0"].copy()
    O.mul_bra(r, G["chain5_psi"], H, al, be)
    assert O.rel_frobenius(r, G["chain5_bra"]) < TOL
    R = G["chain5_R0"].copy()
    O.mul_op_left(R, H, G["chain5_rho"], al, be)
    assert O.rel_frobenius(R, G["chain5_left"]) < TOL
    R = G["chain5_R0"].copy()
    O.mul_op_right(R, G["chain5_rho"], H, al, be)
    assert O.rel_frobenius(R, G["chain5_right"]) < TOL


def test_golden_fock_sparse_and_lazy():
    G = GO.load()
    al, be = G["fock_ab"]
    Hs = O.OSparse.from_scipy(GO.fock_sparse(G))
    sh = (4, 3, 3)
    Hl = O.OLazySum(36, 36, [c for c, _, _ in GO.fock_terms(G)],
                    [O.OLazyTensor(sh, sh, s, tuple(O.OSparse.from_scipy(f) for f in fs), 1.0)
                     for _, s, fs in GO.fock_terms(G)])
    for H in (Hs, Hl):
        R = np.full((36, 36), np.nan, complex)  # beta = 0 overwrites NaN
        O.mul_op_left(R, H, G["fock_rho"], al, be)
        assert O.rel_frobenius(R, G["fock_left"]) < TOL
        R = np.full((36, 36), np.nan, complex)
        O.mul_op_right(R, G["fock_rho"], H, al, be)
        assert O.rel_frobenius(R, G["fock_right"]) < TOL


def test_golden_dense_lazytensor_and_isometry():
    G = GO.load()
    dims = (2, 3, 4)
    lt = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(G[f"lt_A{i}"]) for i in (1, 2, 3)), G["lt_f"][0])
    al, be = G["lt_ab"]
    R = G["lt_R0"].copy()
    O.mul_op_left(R, lt, G["lt_rho"], al, be)
    assert O.rel_frobenius(R, G["lt_left"]) < TOL
    R = G["lt_R0"].copy()
    O.mul_op_right(R, G["lt_rho"], lt, al, be)
    assert O.rel_frobenius(R, G["lt_right"]) < TOL
    r = np.zeros(24, complex)
    O.mul_ket(r, lt, G["lt_psi"], 1, 0)
    assert O.rel_frobenius(r, G["lt_ket"]) < TOL
    iso = O.OLazyTensor((2, 3), (2, 4), (1,), (O.ODense(G["iso_A"]),), 0.5)
    r = np.zeros(6, complex)
    O.mul_ket(r, iso, G["iso_psi"], 1, 0)
    assert O.rel_frobenius(r, G["iso_ket"]) < TOL
