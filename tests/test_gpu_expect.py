"""Fused expect / variance (src/operators.jl:119, 139-143,
src/operators_sparse.jl:36-45, src/operators_dense.jl:219-230) on the GPU:
single-pass forms against the oracle's dense algebra, against the plain
two-pass path, and the fallbacks for operators without a fused form."""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O
import qobhip as Q
from qobhip import _lib
from qobhip import workloads as W
from test_gpu_parity import to_q
from test_gpu_qubit import build_pair, random_pauli_sum

pytestmark = pytest.mark.gpu


def _ev(M, psi):
    return np.vdot(psi, M @ psi)


def _var(M, psi):
    x = M @ psi
    return np.vdot(psi, M @ x) - np.vdot(psi, x) ** 2


@pytest.mark.parametrize("qubit", [1, 0])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_expect_variance_ket_qubit_sum(qubit, dtype, tol):
    g = np.random.default_rng(90)
    L = 10
    Hq, Ho, B = build_pair(L, random_pauli_sum(g, L, 18))
    M = O.dense_of(Ho)
    psi = g.standard_normal(2 ** L) + 1j * g.standard_normal(2 ** L)
    psi /= np.linalg.norm(psi)
    if dtype == np.complex64:
        psi = psi.astype(np.complex64).astype(np.complex128)
    _lib.call("qob_tune_set", b"qubit_enable", qubit)
    try:
        k = Q.Ket(B, psi, dtype)
        e, v = Q.expect(Hq, k), Q.variance(Hq, k)
        assert abs(e - _ev(M, psi)) <= tol * max(1.0, abs(_ev(M, psi)))
        assert abs(v - _var(M, psi)) <= tol * max(1.0, abs(_var(M, psi)))
        ctx = Q.context()
        ctx.set_mode(Q.QOB_MODE_PLAIN)
        try:
            ep, vp = Q.expect(Hq, k), Q.variance(Hq, k)
        finally:
            ctx.set_mode(Q.QOB_MODE_FUSED)
        assert abs(e - ep) <= tol * max(1.0, abs(e)) and abs(v - vp) <= tol * max(1.0, abs(v))
    finally:
        _lib.call("qob_tune_set", b"qubit_enable", 1)


def test_expect_ket_fock_lazy_and_dense_fallback():
    spec = W.fock_fock_spin(4, 1.5, 3)
    lz = W.fock_fock_spin_lazy(spec)
    M = spec.H.toarray()
    g = np.random.default_rng(5)
    D = M.shape[0]
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    k = Q.Ket(spec.basis, psi)
    assert abs(Q.expect(lz, k) - _ev(M, psi)) < 1e-12 * abs(_ev(M, psi))
    assert abs(Q.variance(lz, k) - _var(M, psi)) < 1e-11 * abs(_var(M, psi))
    # a dense LazyTensor has no fused form: the two-pass path answers
    lt = W.dense_lazytensor((4, 6), 2)
    Md = O.dense_of(O.OLazyTensor((4, 6), (4, 6), (1, 2), tuple(O.ODense(f) for f in W.dense_lazytensor_factors((4, 6), 2)), 1.0))
    p2 = g.standard_normal(24) + 1j * g.standard_normal(24)
    k2 = Q.Ket(lt.basis_l, p2)
    assert abs(Q.expect(lt, k2) - _ev(Md, p2)) < 1e-12 * abs(_ev(Md, p2))
    assert abs(Q.variance(lt, k2) - _var(Md, p2)) < 1e-11 * abs(_var(Md, p2))


@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_expect_rho_sparse_and_lazy(adj, dtype, tol):
    spec = W.fock_fock_spin(5, 1.5, 4)
    n = spec.H.shape[0]
    g = np.random.default_rng(6)
    A = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    if dtype == np.complex64:
        rho = rho.astype(np.complex64).astype(np.complex128)
    Hs = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    M = spec.H.toarray()
    if adj:
        Hs = Q.dagger(Hs)
        M = M.conj().T
    R = Q.Operator(spec.basis, spec.basis, rho, dtype)
    want = np.trace(M @ rho)
    assert abs(Q.expect(Hs, R) - want) <= tol * max(1.0, abs(want))
    lz = W.fock_fock_spin_lazy(spec)
    want = np.trace(spec.H.toarray() @ rho)
    got = Q.expect(lz, R)
    assert abs(got - want) <= tol * max(1.0, abs(want))
    ctx = Q.context()
    ctx.set_mode(Q.QOB_MODE_PLAIN)
    try:
        gp = Q.expect(lz, R)
    finally:
        ctx.set_mode(Q.QOB_MODE_FUSED)
    assert abs(gp - got) <= tol * max(1.0, abs(got))


def test_expect_follows_device_coefficients():
    spec = W.spin_chain(8, 2)
    ops = spec.build().operators
    coefs = [(lambda t, k=k: np.cos(t + k)) for k in range(len(ops))]
    H = Q.TimeDependentSum(coefs, ops, init_time=0.0)
    D = 2 ** 8
    g = np.random.default_rng(2)
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    k = Q.Ket(spec.basis, psi)
    Q.expect(H, k)  # plans built
    for t in [0.3, 1.7]:
        H.set_time_(t)
        sh = spec.basis.shape
        Ho = O.OLazySum(D, D, H.eval_coefficients(t),
                        [O.OLazyTensor(sh, sh, tuple(i), tuple(O.OSparse.from_scipy(f) for f in fs), 1.0)
                         for _, i, fs in spec.terms])
        M = O.dense_of(Ho)
        assert abs(Q.expect(H, k) - _ev(M, psi)) < 1e-12 * max(1, abs(_ev(M, psi)))


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_variance_rho_sparse_lazy_dense(dtype, tol):
    """variance(op, rho) = tr(op op rho) - tr(op rho)^2 for a SparseOperator,
    its LazySum form and a dense operator, against dense algebra."""
    spec = W.fock_fock_spin(5, 1.5, 4)
    n = spec.H.shape[0]
    g = np.random.default_rng(8)
    A = g.standard_normal((n, n)) + 1j * g.standard_normal((n, n))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    if dtype == np.complex64:
        rho = rho.astype(np.complex64).astype(np.complex128)
    M = spec.H.toarray()
    R = Q.Operator(spec.basis, spec.basis, rho, dtype)
    want = np.trace(M @ M @ rho) - np.trace(M @ rho) ** 2
    scale = max(1.0, abs(np.trace(M @ M @ rho)))
    for op in (Q.SparseOperator(spec.basis, spec.basis, spec.H), W.fock_fock_spin_lazy(spec),
               Q.DenseOperator(spec.basis, spec.basis, M, dtype)):
        assert abs(Q.variance(op, R) - want) <= 10 * tol * scale, type(op).__name__


def test_time_dependent_sum_expect_variance_follow_time():
    """expect / variance of a TimeDependentSum are those of its static
    operator at the current time (src/time_dependent_operator.jl:76-80), on
    a Ket and on a density matrix."""
    from qobhip.timedependent import TimeDependentSum
    L = 6
    g = np.random.default_rng(12)
    terms = random_pauli_sum(g, L, 8)
    Hq, _, B = build_pair(L, terms)
    coefs = [lambda t, k=k: np.cos(0.3 * k * t) + 0.1j * k for k in range(len(Hq.operators))]
    tds = TimeDependentSum(list(zip(coefs, Hq.operators)))
    D = 2 ** L
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    psi /= np.linalg.norm(psi)
    rho = np.outer(psi, psi.conj())
    k, R = Q.Ket(B, psi), Q.Operator(B, B, rho)
    for t in (0.0, 0.7, 2.5):
        tds.set_time_(t)
        c = [f(t) for f in coefs]
        M = sum(ci * O.dense_of(to_o) for ci, to_o in zip(c, _oracle_terms(Hq)))
        assert abs(Q.expect(tds, k) - _ev(M, psi)) < 1e-12 * max(1, abs(_ev(M, psi)))
        assert abs(Q.variance(tds, k) - _var(M, psi)) < 1e-11 * max(1, abs(_var(M, psi)))
        want_e = np.trace(M @ rho)
        want_v = np.trace(M @ M @ rho) - want_e ** 2
        assert abs(Q.expect(tds, R) - want_e) < 1e-12 * max(1, abs(want_e))
        assert abs(Q.variance(tds, R) - want_v) < 1e-11 * max(1, abs(want_v))


def _oracle_terms(Hq):
    """The oracle LazyTensors of a qobhip LazySum of qubit LazyTensors."""
    out = []
    for lt in Hq.operators:
        sh = tuple(lt.basis_l.shape)
        out.append(O.OLazyTensor(sh, sh, tuple(lt.indices),
                                 tuple(O.OSparse.from_scipy(sp.csc_matrix(f.csc)) for f in lt.operators),
                                 lt.factor))
    return out
