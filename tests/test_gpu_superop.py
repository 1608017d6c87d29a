"""SuperOperator * Operator on the device (src/superoperators.jl:82-85):
liouvillian / spre / spost / sprepost built on the host, applied to vec(rho)
through the sparse and dense Ket paths at dimension N^2."""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O
import qobhip as Q

pytestmark = pytest.mark.gpu


def _system():
    f = Q.FockBasis(3)
    s = Q.SpinBasis(0.5)
    B = Q.tensor(f, s)
    a = Q.destroy(f).csc
    sm = Q.sigmam(s).csc
    sz = Q.sigmaz(s).csc
    If, Is = sp.identity(4), sp.identity(2)
    Hm = sp.kron(Is, a.T.conj() @ a) + 0.5 * sp.kron(sz, If) + 0.3 * (sp.kron(sm.T.conj(), a) + sp.kron(sm, a.T.conj()))
    Jm = [sp.kron(Is, a), sp.kron(sm, If)]
    H = Q.SparseOperator(B, B, sp.csc_matrix(Hm))
    J = [Q.SparseOperator(B, B, sp.csc_matrix(j)) for j in Jm]
    return B, H, J, Hm.toarray(), [j.toarray() for j in Jm]


def _rho(g, n):
    A = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    r = A @ A.conj().T
    return r / np.trace(r)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_liouvillian_times_rho(dtype, tol):
    B, H, J, Hd, Jd = _system()
    rates = np.array([0.4, 1.3])
    L = Q.liouvillian(H, J, rates)
    Lo = O.liouvillian(Hd, Jd, rates)
    assert L.is_sparse() and abs(L.data.toarray() - Lo).max() < 1e-14
    g = np.random.default_rng(3)
    rho = _rho(g, 8)
    R = Q.Operator(B, B, rho, dtype)
    out = L * R
    want = O.superop_times_op(Lo, rho)
    assert O.rel_frobenius(out.to_numpy(), want) < tol
    assert abs(np.trace(out.to_numpy())) < 1e-5 * np.abs(want).max()
    # mul! form with alpha, beta; beta = 0 overwrites NaN
    al, be = 0.5 - 0.25j, 0.0
    res = Q.Operator(B, B, np.full((8, 8), np.nan, complex), dtype)
    Q.mul_(res, L, R, al, be)
    assert O.rel_frobenius(res.to_numpy(), al * want) < tol
    R0 = g.standard_normal((8, 8)) + 0j
    res = Q.Operator(B, B, R0, dtype)
    Q.mul_(res, L, R, al, 0.75)
    assert O.rel_frobenius(res.to_numpy(), al * want + 0.75 * R0) < tol


def test_spre_spost_sprepost_dense_and_sparse():
    B, H, J, Hd, Jd = _system()
    g = np.random.default_rng(4)
    A = g.standard_normal((8, 8)) + 1j * g.standard_normal((8, 8))
    Ad = Q.DenseOperator(B, B, A)
    C = _rho(g, 8)
    Rc = Q.Operator(B, B, C)
    for S, want in [(Q.spre(Ad), A @ C), (Q.spost(Ad), C @ A), (Q.spre(H), Hd @ C),
                    (Q.sprepost(J[0], Q.dagger(J[1])), Jd[0] @ C @ Jd[1].conj().T),
                    (Q.spre(H) * Q.spost(J[0]), Hd @ C @ Jd[0]), (2.0 * Q.identitysuperoperator(B), 2 * C)]:
        got = (S * Rc).to_numpy()
        assert O.rel_frobenius(got, want) < 1e-12
    assert not Q.spre(Ad).is_sparse()
    S = Q.spre(H) - Q.spost(H)  # commutator superoperator
    assert O.rel_frobenius((S * Rc).to_numpy(), Hd @ C - C @ Hd) < 1e-12
    Sd = S.dagger()
    assert O.rel_frobenius(Sd.data.toarray(), S.data.toarray().conj().T) < 1e-15
