"""The HIP path against the golden fixtures (dense linear algebra from
closed forms, tests/golden/make_golden.py), through the C ABI."""
import numpy as np
import pytest

import golden_ops as GO
import qob_oracle as O
import qobhip as Q
from test_gpu_parity import to_q

pytestmark = pytest.mark.gpu


def _chain():
    G = GO.load()
    L, terms = GO.chain_terms(G)
    sh = (2,) * L
    D = 2 ** L
    H = O.OLazySum(D, D, [c for c, _, _ in terms],
                   [O.OLazyTensor(sh, sh, s, tuple(O.OSparse.from_scipy(f) for f in fs), 1.0) for _, s, fs in terms])
    return G, to_q(H), D


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
@pytest.mark.parametrize("qubit", [1, 0])
def test_golden_chain(dtype, tol, qubit):
    from qobhip import _lib
    G, H, D = _chain()
    al, be = G["chain5_ab"]
    _lib.call("qob_tune_set", b"qubit_enable", qubit)
    try:
        r = Q.Ket(H.basis_l, G["chain5_r0"], dtype)
        Q.mul_(r, H, Q.Ket(H.basis_r, G["chain5_psi"], dtype), al, be)
        assert O.rel_frobenius(r.to_numpy(), G["chain5_ket"]) < tol
        r = Q.Bra(H.basis_r, G["chain5_r0"], dtype)
        Q.mul_(r, Q.Bra(H.basis_l, G["chain5_psi"], dtype), H, al, be)
        assert O.rel_frobenius(r.to_numpy(), G["chain5_bra"]) < tol
        R = Q.Operator(H.basis_l, H.basis_r, G["chain5_R0"], dtype)
        Q.mul_(R, H, Q.Operator(H.basis_l, H.basis_r, G["chain5_rho"], dtype), al, be)
        assert O.rel_frobenius(R.to_numpy(), G["chain5_left"]) < tol
        R = Q.Operator(H.basis_l, H.basis_r, G["chain5_R0"], dtype)
        Q.mul_(R, Q.Operator(H.basis_l, H.basis_r, G["chain5_rho"], dtype), H, al, be)
        assert O.rel_frobenius(R.to_numpy(), G["chain5_right"]) < tol
    finally:
        _lib.call("qob_tune_set", b"qubit_enable", 1)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_golden_fock(dtype, tol):
    G = GO.load()
    al, be = G["fock_ab"]
    sh = (4, 3, 3)
    Hs = to_q(O.OSparse.from_scipy(GO.fock_sparse(G)))
    Hl = to_q(O.OLazySum(36, 36, [c for c, _, _ in GO.fock_terms(G)],
                         [O.OLazyTensor(sh, sh, s, tuple(O.OSparse.from_scipy(f) for f in fs), 1.0)
                          for _, s, fs in GO.fock_terms(G)]))
    for H in (Hs, Hl):
        b = Q.GenericBasis(36)
        rho = Q.Operator(b, b, G["fock_rho"], dtype)
        R = Q.Operator(b, b, np.full((36, 36), np.nan, complex), dtype)
        Q.mul_(R, H, rho, al, be)
        assert O.rel_frobenius(R.to_numpy(), G["fock_left"]) < tol
        R = Q.Operator(b, b, np.full((36, 36), np.nan, complex), dtype)
        Q.mul_(R, rho, H, al, be)
        assert O.rel_frobenius(R.to_numpy(), G["fock_right"]) < tol


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_golden_dense_lazytensor_and_isometry(dtype, tol):
    G = GO.load()
    dims = (2, 3, 4)
    lt = to_q(O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(G[f"lt_A{i}"]) for i in (1, 2, 3)),
                            G["lt_f"][0]))
    al, be = G["lt_ab"]
    b = Q.GenericBasis(24)
    rho = Q.Operator(b, b, G["lt_rho"], dtype)
    R = Q.Operator(b, b, G["lt_R0"], dtype)
    Q.mul_(R, lt, rho, al, be)
    assert O.rel_frobenius(R.to_numpy(), G["lt_left"]) < tol
    R = Q.Operator(b, b, G["lt_R0"], dtype)
    Q.mul_(R, rho, lt, al, be)
    assert O.rel_frobenius(R.to_numpy(), G["lt_right"]) < tol
    r = Q.Ket(lt.basis_l, np.zeros(24), dtype)
    Q.mul_(r, lt, Q.Ket(lt.basis_r, G["lt_psi"], dtype))
    assert O.rel_frobenius(r.to_numpy(), G["lt_ket"]) < tol
    iso = to_q(O.OLazyTensor((2, 3), (2, 4), (1,), (O.ODense(G["iso_A"]),), 0.5))
    r = Q.Ket(iso.basis_l, np.zeros(6), dtype)
    Q.mul_(r, iso, Q.Ket(iso.basis_r, G["iso_psi"], dtype))
    assert O.rel_frobenius(r.to_numpy(), G["iso_ket"]) < tol
