"""Pins of the oracle (CPU).

Each restated reference function in oracle/qob_oracle.py is checked against
something that does not share its code: the dense Kronecker product
(dense_of), closed forms, the reference's known-answer checks, invariants,
and planted mistakes that must be caught.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O

TOL = 1e-12


def rs(g, m, n, density):
    """Random sparse complex matrix with at least one entry."""
    P = sp.random(m, n, density=density, random_state=np.random.RandomState(g.integers(1 << 30)), format="csc")
    M = sp.csc_matrix((g.standard_normal(P.nnz) + 1j * g.standard_normal(P.nnz), P.indices, P.indptr),
                      shape=(m, n))
    if M.nnz == 0:
        M = sp.csc_matrix(([1.0 + 0.5j], ([0], [0])), shape=(m, n))
    return M


def cg(g, *shape):
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


def rel(a, b):
    return O.rel_frobenius(a, b)


AB = [(1.0, 0.0), (1.5, 2.1), (complex(0.7, 1.5), complex(0.3, 2.1))]


# ------------------------------------------------------------ sparse CSC
@pytest.mark.parametrize("m,n,k,density", [(3, 5, 7, 0.5), (50, 60, 55, 0.3)])
@pytest.mark.parametrize("alpha,beta", AB)
def test_sparse_gemm_vs_dense(m, n, k, density, alpha, beta):
    g = np.random.default_rng(1)
    M = rs(g, m, n, density)
    Md = M.toarray()
    op = O.OSparse.from_scipy(M)
    if m == 50:
        assert op.nnz() > 550  # exercises gemm_sp_dense_big
    else:
        assert op.nnz() <= 550
    B = cg(g, n, k)
    R0 = cg(g, m, k)
    R = R0.copy()
    O.gemm_sp_left(alpha, op, B, beta, R)
    assert rel(R, alpha * Md @ B + beta * R0) < TOL
    # right: B2 (k x m) * M
    B2 = cg(g, k, m)
    R0 = cg(g, k, n)
    R = R0.copy()
    O.gemm_sp_right(alpha, B2, op, beta, R)
    assert rel(R, alpha * B2 @ Md + beta * R0) < TOL
    # adjoint forms (src/sparsematrix.jl:148-175)
    opa = O.dagger(op)
    B3 = cg(g, m, k)
    R0 = cg(g, n, k)
    R = R0.copy()
    O.gemm_sp_left(alpha, opa, B3, beta, R)
    assert rel(R, alpha * Md.conj().T @ B3 + beta * R0) < TOL
    B4 = cg(g, k, n)
    R0 = cg(g, k, m)
    R = R0.copy()
    O.gemm_sp_right(alpha, B4, opa, beta, R)
    assert rel(R, alpha * B4 @ Md.conj().T + beta * R0) < TOL


def test_sparse_small_and_big_branches_agree():
    g = np.random.default_rng(2)
    M = rs(g, 40, 40, 0.5)
    op = O.OSparse.from_scipy(M)
    B = cg(g, 40, 9)
    r1 = np.zeros((40, 9), complex)
    r2 = np.zeros((40, 9), complex)
    O.gemm_sp_dense_small(1.3, op, B, r1)
    O.gemm_sp_dense_big(1.3, op, B, r2)
    assert rel(r1, r2) < 1e-14


@pytest.mark.parametrize("alpha,beta", AB)
def test_gemv_vs_dense(alpha, beta):
    g = np.random.default_rng(3)
    M = rs(g, 7, 9, 0.4)
    Md = M.toarray()
    op = O.OSparse.from_scipy(M)
    v = cg(g, 9)
    r0 = cg(g, 7)
    r = r0.copy()
    O.gemv_sp(alpha, op, v, beta, r)
    assert rel(r, alpha * Md @ v + beta * r0) < TOL
    w = cg(g, 7)
    r0 = cg(g, 9)
    r = r0.copy()
    O.gemv_vec_sp(alpha, w, op, beta, r)
    assert rel(r, alpha * Md.T @ w + beta * r0) < TOL


def test_beta_zero_discards_nan_and_dim_errors():
    g = np.random.default_rng(4)
    op = O.OSparse.from_scipy(rs(g, 6, 6, 0.5))
    B = cg(g, 6, 3)
    R = np.full((6, 3), np.nan, complex)
    O.gemm_sp_left(1.0, op, B, 0.0, R)
    assert np.isfinite(R).all()
    with pytest.raises(O.DimensionMismatch):
        O.gemm_sp_left(1.0, op, cg(g, 5, 3), 0.0, np.zeros((6, 3), complex))
    with pytest.raises(O.DimensionMismatch):
        O.gemv_sp(1.0, op, cg(g, 5), 0.0, np.zeros(6, complex))


# ----------------------------------------------------------- LazyTensor
def rand_lt(g, shape_l, shape_r, indices, kinds, factor=0.7 - 0.2j):
    ops = []
    for i, kd in zip(indices, kinds):
        dl, dr = shape_l[i - 1], shape_r[i - 1]
        if kd == "sparse":
            ops.append(O.OSparse.from_scipy(rs(g, dl, dr, 0.6)))
        elif kd == "adjsparse":
            ops.append(O.dagger(O.OSparse.from_scipy(rs(g, dr, dl, 0.6))))
        elif kd == "dense":
            ops.append(O.ODense(cg(g, dl, dr)))
        elif kd == "eye":
            ops.append(O.OEye(dl, dr))
    return O.OLazyTensor(tuple(shape_l), tuple(shape_r), tuple(indices), tuple(ops), factor)


LT_CASES = [
    ((2, 1, 6), (3, 4, 5), (1, 2, 3), ("sparse", "sparse", "sparse")),  # reference test bases b_l, b_r
    ((2, 1, 6), (3, 4, 5), (1, 2, 3), ("dense", "dense", "sparse")),
    ((2, 1, 6), (3, 4, 5), (1, 2, 3), ("dense", "adjsparse", "dense")),
    ((3, 4, 2, 5), (3, 4, 2, 5), (2, 4), ("sparse", "sparse")),
    ((3, 4, 2, 5), (3, 4, 2, 5), (1, 3), ("dense", "dense")),
    ((2, 2, 2, 2, 2), (2, 2, 2, 2, 2), (2, 3), ("sparse", "sparse")),
    ((4, 3), (4, 3), (2,), ("eye",)),
    ((3, 2, 2, 1, 2), (3, 2, 2, 2, 1), (1, 3), ("sparse", "sparse")),  # implicit isometries
    ((3, 2, 2, 1, 2), (3, 2, 2, 2, 1), (1, 2, 3, 5), ("sparse", "eye", "sparse", "eye")),
    ((3, 2, 2, 1, 2), (3, 2, 2, 2, 1), (1, 2, 3, 5), ("dense", "eye", "sparse", "eye")),
]


@pytest.mark.parametrize("case", range(len(LT_CASES)))
@pytest.mark.parametrize("alpha,beta", AB)
def test_lazytensor_all_paths_vs_kron(case, alpha, beta):
    shl, shr, idx, kinds = LT_CASES[case]
    g = np.random.default_rng(10 + case)
    lt = rand_lt(g, shl, shr, idx, kinds)
    A = O.dense_of(lt)
    m, n = A.shape
    # Ket
    b = cg(g, n)
    r0 = cg(g, m)
    r = r0.copy()
    O.mul_ket(r, lt, b, alpha, beta)
    assert rel(r, alpha * A @ b + beta * r0) < TOL
    # Bra
    a = cg(g, m)
    r0 = cg(g, n)
    r = r0.copy()
    O.mul_bra(r, a, lt, alpha, beta)
    assert rel(r, alpha * A.T @ a + beta * r0) < TOL
    # Operator left / right
    B = cg(g, n, 4)
    R0 = cg(g, m, 4)
    R = R0.copy()
    O.mul_op_left(R, lt, B, alpha, beta)
    assert rel(R, alpha * A @ B + beta * R0) < TOL
    B = cg(g, 3, m)
    R0 = cg(g, 3, n)
    R = R0.copy()
    O.mul_op_right(R, B, lt, alpha, beta)
    assert rel(R, alpha * B @ A + beta * R0) < TOL


def test_lazytensor_alpha_zero_and_nan():
    g = np.random.default_rng(5)
    lt = rand_lt(g, (2, 3), (2, 3), (1, 2), ("sparse", "sparse"))
    r = np.full(6, np.nan, complex)
    O.mul_ket(r, lt, cg(g, 6), 0.0, 0.0)
    assert (r == 0).all()
    r = np.full(6, np.nan, complex)
    O.mul_ket(r, lt, cg(g, 6), 1.0, 0.0)
    assert np.isfinite(r).all()


def test_lazytensor_scaled_identity_no_factors():
    # LazyTensor(b_l, b_l, (), (), 0.5) (test_operators_lazytensor.jl:373-407)
    g = np.random.default_rng(6)
    lt = O.OLazyTensor((2, 3), (2, 3), (), (), 0.5)
    B = cg(g, 6, 3)
    R0 = cg(g, 6, 3)
    R = R0.copy()
    O.mul_op_left(R, lt, B, 1.5, 2.1)
    assert rel(R, 1.5 * 0.5 * B + 2.1 * R0) < TOL


# ---------------------------------------------------------------- LazySum
@pytest.mark.parametrize("alpha,beta", AB)
def test_lazysum_vs_dense_and_zero(alpha, beta):
    g = np.random.default_rng(7)
    shl = shr = (2, 3, 2)
    ops = [rand_lt(g, shl, shr, (1, 2), ("sparse", "sparse")),
           rand_lt(g, shl, shr, (3,), ("dense",)),
           O.OSparse.from_scipy(rs(g, 12, 12, 0.3))]
    ls = O.OLazySum(12, 12, [0.1, 0.3, 1.2], ops)
    A = O.dense_of(ls)
    b = cg(g, 12)
    r0 = cg(g, 12)
    r = r0.copy()
    O.mul_ket(r, ls, b, alpha, beta)
    assert rel(r, alpha * A @ b + beta * r0) < TOL
    zero = O.OLazySum(12, 12, [], [])
    r = r0.copy()
    O.mul_ket(r, zero, b, alpha, beta)
    assert rel(r, beta * r0) < TOL
    R = np.full((12, 12), np.nan, complex)
    O.mul_op_right(R, cg(g, 12, 12), zero, 1.0, 0.0)
    assert (R == 0).all()


# ------------------------------------------------------------ closed forms
def sx():
    return O.OSparse.from_scipy(sp.csc_matrix(np.array([[0, 1], [1, 0]], complex)))


def sz():
    return O.OSparse.from_scipy(sp.csc_matrix(np.array([[1, 0], [0, -1]], complex)))


def test_sigma_z_chain_energies_on_basis_states():
    # sum_i h_i sz_i |bits> = sum_i h_i (1 - 2 bit_i) |bits>, bit 0 <-> m=+1/2
    L = 6
    g = np.random.default_rng(8)
    h = g.uniform(-1, 1, L)
    sh = (2,) * L
    ops = [O.OLazyTensor(sh, sh, (i + 1,), (sz(),), 1.0) for i in range(L)]
    H = O.OLazySum(2 ** L, 2 ** L, list(h), ops)
    for J in [0, 5, 17, 63, 42]:
        e = np.zeros(2 ** L, complex)
        e[J] = 1
        r = np.zeros(2 ** L, complex)
        O.mul_ket(r, H, e, 1, 0)
        bits = [(J >> i) & 1 for i in range(L)]
        E = sum(h[i] * (1 - 2 * bits[i]) for i in range(L))
        want = np.zeros(2 ** L, complex)
        want[J] = E
        assert np.abs(r - want).max() < 1e-14


def test_sigma_x_sigma_x_flips_two_bits():
    L = 5
    sh = (2,) * L
    for i in range(L - 1):
        op = O.OLazyTensor(sh, sh, (i + 1, i + 2), (sx(), sx()), 1.0)
        for J in [0, 3, 9, 30]:
            e = np.zeros(2 ** L, complex)
            e[J] = 1
            r = np.zeros(2 ** L, complex)
            O.mul_ket(r, op, e, 1, 0)
            K = J ^ (3 << i)
            assert r[K] == 1 and np.count_nonzero(r) == 1


def fock_ops(N):
    n = N + 1
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(complex)
    return a, a.conj().T, np.diag(np.arange(n)).astype(complex)


def test_number_operator_diagonal_in_coupled_space():
    # H0 = n_a (x) 1 + 1 (x) n_b on Fock states |na, nb> gives (na + nb)
    N = 4
    a, ad, num = fock_ops(N)
    sh = (N + 1, N + 1)
    t1 = O.OLazyTensor(sh, sh, (1,), (O.OSparse.from_scipy(sp.csc_matrix(num)),), 1.0)
    t2 = O.OLazyTensor(sh, sh, (2,), (O.OSparse.from_scipy(sp.csc_matrix(num)),), 1.0)
    H = O.OLazySum((N + 1) ** 2, (N + 1) ** 2, [1.0, 1.0], [t1, t2])
    rho = np.diag(np.arange((N + 1) ** 2)).astype(complex)  # diagonal test state
    R = np.zeros_like(rho)
    O.mul_op_left(R, H, rho, 1, 0)
    for J in range((N + 1) ** 2):
        na, nb = J % (N + 1), J // (N + 1)
        assert abs(R[J, J] - (na + nb) * J) < 1e-13


def test_ladder_commutator_truncated():
    N = 6
    a, ad, num = fock_ops(N)
    A = O.OSparse.from_scipy(sp.csc_matrix(a))
    Ad = O.dagger(A)
    n = N + 1
    I = np.eye(n, dtype=complex)
    r1 = np.zeros((n, n), complex)
    r2 = np.zeros((n, n), complex)
    # a a^dag - a^dag a = 1 except the last diagonal entry (= -N)
    O.mul_op_left(r1, A, np.asfortranarray(a.conj().T), 1, 0)
    O.mul_op_left(r2, Ad, np.asfortranarray(a), 1, 0)
    c = r1 - r2
    want = I.copy()
    want[-1, -1] = -N
    assert np.abs(c - want).max() < 1e-13


# ------------------------------------------------ reference known answers
def test_kat_coherent_state_amplitude_and_photon_number():
    # src/precompile.jl:3-21: FockBasis(20), alpha = 1.2
    N, alpha = 20, 1.2
    n = np.arange(N + 1)
    from math import factorial
    c = np.array([alpha ** k / np.sqrt(float(factorial(k))) for k in n], dtype=complex)
    psi = c / np.linalg.norm(c)
    a, ad, num = fock_ops(N)
    A = O.OSparse.from_scipy(sp.csc_matrix(a))
    amp = O.expect_ket(A, psi)
    assert abs(amp - alpha) < 1e-10
    rho = np.outer(psi, psi.conj())
    pn = O.expect_op(O.OSparse.from_scipy(sp.csc_matrix(num)), rho)
    assert abs(pn - alpha ** 2) < 1e-9
    assert abs(O.trace(rho) - 1) < 1e-12


def test_kat_time_dependent_sum_at_t():
    # src/precompile.jl:89-106: (cos t sx + sin t sz) |up> at t = 0.25
    t = 0.25
    ls = O.OLazySum(2, 2, [np.cos(t), np.sin(t)], [sx(), sz()])
    up = np.array([1, 0], complex)
    r = np.zeros(2, complex)
    O.mul_ket(r, ls, up, 1, 0)
    expected = np.cos(t) * np.array([0, 1]) + np.sin(t) * np.array([1, 0])
    assert np.linalg.norm(r - expected) < 1e-12


# --------------------------------------------------------------- invariants
def test_hermiticity_and_commutator_trace():
    g = np.random.default_rng(9)
    sh = (3, 2, 3)
    X = rand_lt(g, sh, sh, (1, 3), ("sparse", "sparse"))
    H = O.OLazySum(18, 18, [1.0, 1.0], [X, O.dagger(X)])  # Hermitian
    psi, phi = cg(g, 18), cg(g, 18)
    Hpsi = np.zeros(18, complex)
    Hphi = np.zeros(18, complex)
    O.mul_ket(Hpsi, H, psi, 1, 0)
    O.mul_ket(Hphi, H, phi, 1, 0)
    assert abs(np.vdot(phi, Hpsi) - np.conj(np.vdot(psi, Hphi))) < 1e-12
    rho = cg(g, 18, 18)
    rho = rho @ rho.conj().T
    L = np.zeros_like(rho)
    Rr = np.zeros_like(rho)
    O.mul_op_left(L, H, rho, 1, 0)
    O.mul_op_right(Rr, rho, H, 1, 0)
    assert abs(np.trace(L - Rr)) < 1e-10 * np.abs(L).max()


def test_linearity_in_alpha_beta():
    g = np.random.default_rng(10)
    lt = rand_lt(g, (3, 4), (3, 4), (1, 2), ("sparse", "sparse"))
    b = cg(g, 12)
    r0 = cg(g, 12)
    r1 = r0.copy()
    O.mul_ket(r1, lt, b, 2.0, 0.0)
    r2 = r0.copy()
    O.mul_ket(r2, lt, b, 1.0, 0.0)
    assert rel(r1, 2 * r2) < TOL
    r3 = r0.copy()
    O.mul_ket(r3, lt, b, 1.0, 1.0)
    assert rel(r3, r2 + r0) < TOL


# ---------------------------------------------------------- planted mistakes
def test_planted_stride_mistake_is_caught(monkeypatch):
    g = np.random.default_rng(11)
    lt = rand_lt(g, (2, 3, 4), (2, 3, 4), (1, 3), ("sparse", "sparse"))
    A = O.dense_of(lt)
    b = cg(g, 24)
    monkeypatch.setattr(O, "_strides", lambda shape: list(reversed(
        [int(np.prod(shape[:i])) for i in range(len(shape))])))
    r = np.zeros(24, complex)
    O.mul_ket(r, lt, b, 1, 0)
    assert rel(r, A @ b) > 1e-3


def test_planted_dropped_alpha_is_caught(monkeypatch):
    g = np.random.default_rng(12)
    M = rs(g, 40, 40, 0.6)
    op = O.OSparse.from_scipy(M)
    assert op.nnz() > 550
    B = cg(g, 40, 5)
    orig = O.gemm_sp_dense_big
    monkeypatch.setattr(O, "gemm_sp_dense_big", lambda alpha, M, B, r: orig(1.0, M, B, r))
    R = np.zeros((40, 5), complex)
    O.gemm_sp_left(2.5, op, B, 0.0, R)
    assert rel(R, 2.5 * M.toarray() @ B) > 1e-3


def test_planted_transposed_factor_is_caught(monkeypatch):
    g = np.random.default_rng(13)
    lt = rand_lt(g, (3, 3), (3, 3), (1, 2), ("dense", "dense"))
    A = O.dense_of(lt)
    b = cg(g, 9)
    orig = O._op_data
    monkeypatch.setattr(O, "_op_data", lambda o, t: orig(o, "transpose" if t is None else None))
    r = np.zeros(9, complex)
    O.mul_ket(r, lt, b, 1, 0)
    assert rel(r, A @ b) > 1e-3


def test_c_lazysum_engine_matches_oracle_and_kron():
    """oracle/cpu_ref.c's recursive LazySum engine (the CPU baseline of the
    Ket workloads) against the numpy restatement and the Kronecker product."""
    import scipy.sparse as sp
    from cref import LazySumKet
    g = np.random.default_rng(11)
    shape = (2, 3, 2, 4)
    D = int(np.prod(shape))
    terms = []
    for _ in range(7):
        k = int(g.integers(1, 3))
        sites = sorted(g.choice(4, size=k, replace=False) + 1)
        facs = [sp.random(shape[s - 1], shape[s - 1], density=0.6, random_state=int(g.integers(1 << 30)),
                          format="csc") * (1 + 0.5j) for s in sites]
        terms.append((complex(g.standard_normal(), g.standard_normal()), tuple(int(s) for s in sites), facs))
    ops = [O.OLazyTensor(shape, shape, t[1], tuple(O.OSparse.from_scipy(f) for f in t[2]), 1.0) for t in terms]
    H = O.OLazySum(D, D, [t[0] for t in terms], ops)
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    r0 = g.standard_normal(D) + 1j * g.standard_normal(D)
    al, be = 0.3 - 0.7j, 0.2 + 0.1j
    want = r0.copy()
    O.mul_ket(want, H, psi, al, be)
    got = LazySumKet(shape, terms).mul(r0.copy(), psi, al, be)
    assert O.rel_frobenius(got, want) < 1e-13
    dense = sum(t[0] * O.dense_of(op) for t, op in zip(terms, ops))
    assert O.rel_frobenius(got, al * dense @ psi + be * r0) < 1e-12
    # empty sum with beta = 0 overwrites NaN
    r = np.full(D, np.nan + 0j)
    LazySumKet(shape, []).mul(r, psi, 1, 0)
    assert np.array_equal(r, np.zeros(D))


@pytest.mark.parametrize("shape,seed", [((2, 3, 2, 4), 5), ((3, 2, 2, 2, 3), 6), ((2,) * 9, 7), ((5,), 8)])
def test_c_flat_lazysum_engine_bit_equal_to_recursion(shape, seed):
    """qobref_lazysum_ket_flat (each term's recursion unrolled into nested
    loops; the checker of long step sequences at full size) equals the
    recursive engine bit for bit, with factors at the first and last sites,
    several stored entries per column and empty terms' identity sites; and
    both equal the Kronecker product."""
    import scipy.sparse as sp
    from cref import LazySumKet
    g = np.random.default_rng(seed)
    n = len(shape)
    D = int(np.prod(shape))
    terms = []
    for t in range(8):
        k = int(g.integers(1, min(3, n) + 1))
        sites = sorted(g.choice(n, size=k, replace=False) + 1)
        if t == 0:
            sites = [1]
        if t == 1:
            sites = [n]
        facs = [sp.random(shape[s - 1], shape[s - 1], density=0.7, random_state=int(g.integers(1 << 30)),
                          format="csc") * (1 - 0.4j) for s in sites]
        terms.append((complex(g.standard_normal(), g.standard_normal()), tuple(int(s) for s in sites), facs))
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    r0 = g.standard_normal(D) + 1j * g.standard_normal(D)
    H = LazySumKet(shape, terms)
    for al, be in ((1.0, 0.0), (0.3 - 0.7j, 0.2 + 0.1j)):
        a = H.mul(r0.copy(), psi, al, be)
        b = H.mul_flat(r0.copy(), psi, al, be)
        assert np.array_equal(a, b)
    ops = [O.OLazyTensor(shape, shape, t[1], tuple(O.OSparse.from_scipy(f) for f in t[2]), 1.0) for t in terms]
    dense = sum(t[0] * O.dense_of(op) for t, op in zip(terms, ops))
    assert O.rel_frobenius(b, al * dense @ psi + be * r0) < 1e-12
    r = np.full(D, np.nan + 0j)
    LazySumKet(shape, []).mul_flat(r, psi, 1, 0)
    assert np.array_equal(r, np.zeros(D))


def test_c_flat_engine_on_a_tfim_chain_is_bit_equal_and_fast():
    """A 16-site transverse-field Ising chain (the batch workload's term
    structure): flat equals recursive bit for bit, at a fraction of the time."""
    import time
    from cref import LazySumKet
    from qobhip import workloads as W
    spec = W.tfim_chain(16, 3)
    H = LazySumKet(spec.basis.shape, spec.terms)
    psi = W.normalized_ket(W.rng(4), 2 ** 16)
    t0 = time.perf_counter()
    a = H.mul(np.zeros(2 ** 16, complex), psi)
    t1 = time.perf_counter()
    b = H.mul_flat(np.zeros(2 ** 16, complex), psi)
    t2 = time.perf_counter()
    assert np.array_equal(a, b)
    assert (t2 - t1) < (t1 - t0)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.3 - 0.7j, 0.2 + 0.1j), (0.5, 1.0)])
def test_c_sparse_products_match_scipy(alpha, beta):
    """oracle/cpu_ref.c's sparse products (the CPU baseline of the sparse
    workload) against scipy's CSC products; beta = 0 overwrites NaN; the
    OpenMP variants sum every output element in the same order, so they
    equal the one-thread loops bit for bit."""
    import ctypes as C
    import scipy.sparse as sp
    from cref import lib
    L = lib()
    g = np.random.default_rng(19)
    m, n, k = 70, 90, 13
    M = sp.random(m, n, density=0.15, random_state=3, format="csc") * (1 - 0.5j)
    M = M + sp.random(m, n, density=0.05, random_state=4, format="csc") * 1j
    M = sp.csc_matrix(M)
    M.sort_indices()
    cp = (M.indptr.astype(np.int64) + 1)
    rv = (M.indices.astype(np.int64) + 1)
    nz = np.ascontiguousarray(M.data.astype(np.complex128))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    a, b = complex(alpha), complex(beta)
    # left: R (m x k) = a M B + b R, B (n x k)
    B = np.asfortranarray(g.standard_normal((n, k)) + 1j * g.standard_normal((n, k)))
    R0 = np.full((m, k), np.nan, complex) if b == 0 else g.standard_normal((m, k)) + 1j * g.standard_normal((m, k))
    want = a * (M @ B) + (0 if b == 0 else b * R0)
    outs = []
    for omp in (0, 1):
        R = np.asfortranarray(R0.copy())
        if omp:
            L.qobref_gemm_sp_dense_omp(a.real, a.imag, m, n, P(cp), P(rv), P(nz), P(B), k, b.real, b.imag, P(R), 4)
        else:
            L.qobref_gemm_sp_dense(a.real, a.imag, m, n, P(cp), P(rv), P(nz), P(B), k, b.real, b.imag, P(R))
        assert rel(R, want) < 1e-14
        outs.append(R)
    assert np.array_equal(outs[0], outs[1])
    # right: R (r x n) = a B M + b R, B (r x m)
    r = 11
    B = np.asfortranarray(g.standard_normal((r, m)) + 1j * g.standard_normal((r, m)))
    R0 = np.full((r, n), np.nan, complex) if b == 0 else g.standard_normal((r, n)) + 1j * g.standard_normal((r, n))
    want = a * (M.T @ B.T).T + (0 if b == 0 else b * R0)
    outs = []
    for omp in (0, 1):
        R = np.asfortranarray(R0.copy())
        if omp:
            L.qobref_gemm_dense_sp_omp(a.real, a.imag, r, m, n, P(cp), P(rv), P(nz), P(B), b.real, b.imag, P(R), 4)
        else:
            L.qobref_gemm_dense_sp(a.real, a.imag, r, m, n, P(cp), P(rv), P(nz), P(B), b.real, b.imag, P(R))
        assert rel(R, want) < 1e-14
        outs.append(R)
    assert np.array_equal(outs[0], outs[1])


def _lp_case(g):
    """A LazyProduct of a LazyTensor, a dense operator, a LazySum and a
    sparse operator on 2 x 3 sites (all 6 x 6), factor 0.5 - 0.25i."""
    lt = rand_lt(g, (2, 3), (2, 3), (1, 2), ("sparse", "dense"))
    dn = O.ODense(cg(g, 6, 6))
    ls = O.OLazySum(6, 6, [0.3, -1.1j], [rand_lt(g, (2, 3), (2, 3), (2,), ("sparse",)),
                                          rand_lt(g, (2, 3), (2, 3), (1,), ("dense",))])
    spm = O.OSparse.from_scipy(rs(g, 6, 6, 0.4))
    return O.OLazyProduct([lt, dn, ls, spm], 0.5 - 0.25j)


@pytest.mark.parametrize("alpha,beta", AB)
def test_lazyproduct_vs_dense_product(alpha, beta):
    g = np.random.default_rng(77)
    lp = _lp_case(g)
    M = O.dense_of(lp)
    # brute force: the product of the factors' own dense matrices
    P = lp.factor * O.dense_of(lp.operators[0])
    for o in lp.operators[1:]:
        P = P @ O.dense_of(o)
    assert O.rel_frobenius(M, P) < 1e-14
    x, r0 = cg(g, 6), cg(g, 6)
    B, R0 = cg(g, 6, 4), cg(g, 6, 4)
    Bl, Rl = cg(g, 3, 6), cg(g, 3, 6)
    r = r0.copy()
    O.mul_ket(r, lp, x, alpha, beta)
    assert O.rel_frobenius(r, alpha * M @ x + beta * r0) < 1e-12
    r = r0.copy()
    O.mul_bra(r, x, lp, alpha, beta)
    assert O.rel_frobenius(r, alpha * x @ M + beta * r0) < 1e-12
    R = R0.copy()
    O.mul_op_left(R, lp, B, alpha, beta)
    assert O.rel_frobenius(R, alpha * M @ B + beta * R0) < 1e-12
    R = Rl.copy()
    O.mul_op_right(R, Bl, lp, alpha, beta)
    assert O.rel_frobenius(R, alpha * Bl @ M + beta * Rl) < 1e-12
    # dagger reverses the factors and conjugates the factor
    assert O.rel_frobenius(O.dense_of(O.dagger(lp)), M.conj().T) < 1e-14


def test_lazyproduct_single_factor_alpha_zero_and_in_sum():
    g = np.random.default_rng(78)
    lp1 = O.OLazyProduct([O.ODense(cg(g, 4, 5))], 2.0)
    x = cg(g, 5)
    r = np.zeros(4, complex)
    O.mul_ket(r, lp1, x, 0.5, 0)
    assert O.rel_frobenius(r, O.dense_of(lp1) @ x * 0.5) < 1e-14
    r = np.full(4, np.nan + 0j)
    O.mul_ket(r, lp1, x, 0, 0)  # alpha = 0, beta = 0 overwrites NaN
    assert np.array_equal(r, np.zeros(4))
    lp = _lp_case(g)
    s = O.OLazySum(6, 6, [1.0, 2.0j], [lp, O.ODense(cg(g, 6, 6))])
    y = cg(g, 6)
    r = np.zeros(6, complex)
    O.mul_ket(r, s, y)
    assert O.rel_frobenius(r, O.dense_of(s) @ y) < 1e-12


def test_planted_lazyproduct_order_mistake_is_caught(monkeypatch):
    g = np.random.default_rng(79)
    lp = _lp_case(g)
    x = cg(g, 6)
    orig = O._lp_chain

    def reversed_chain(result, lpx, b, alpha, beta, mul, right):
        return orig(result, lpx, b, alpha, beta, mul, not right)  # wrong order
    monkeypatch.setattr(O, "_lp_chain", reversed_chain)
    r = np.zeros(6, complex)
    with pytest.raises(Exception):
        O.mul_ket(r, lp, x, 1, 0)
        assert O.rel_frobenius(r, O.dense_of(lp) @ x) < 1e-12


def test_superoperator_identities_and_lindblad_form():
    """spre / spost / sprepost act as A B, B A, A C B on vec (closed forms),
    the liouvillian equals the Lindblad right-hand side written out, and is
    trace preserving (src/precompile.jl:84 checks the same at 1e-11)."""
    g = np.random.default_rng(81)
    n = 5
    A, B, C = cg(g, n, n), cg(g, n, n), cg(g, n, n)
    assert O.rel_frobenius(O.superop_times_op(O.spre(A), B), A @ B) < 1e-14
    assert O.rel_frobenius(O.superop_times_op(O.spost(A), B), B @ A) < 1e-14
    assert O.rel_frobenius(O.superop_times_op(O.sprepost(A, B), C), A @ C @ B) < 1e-14
    H = A + A.conj().T
    J = [cg(g, n, n), cg(g, n, n)]
    rates = np.array([0.3, 1.7])
    rho = B @ B.conj().T
    rho /= np.trace(rho)
    L = O.liouvillian(H, J, rates)
    want = -1j * (H @ rho - rho @ H)
    for r, j in zip(rates, J):
        jd = j.conj().T
        want = want + r * (j @ rho @ jd - 0.5 * (jd @ j @ rho + rho @ jd @ j))
    got = O.superop_times_op(L, rho)
    assert O.rel_frobenius(got, want) < 1e-13
    assert abs(np.trace(got)) < 1e-11 * np.abs(got).max()
    # matrix rates: the diagonal matrix equals the vector form
    assert O.rel_frobenius(O.liouvillian(H, J, np.diag(rates)), L) < 1e-14
