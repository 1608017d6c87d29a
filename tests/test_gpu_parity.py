"""GPU parity: every mul! path through libqob_hip.so against the oracle.

Bar: relative Frobenius error <= 1e-12 for ComplexF64, <= 1e-5 for
ComplexF32, on the same seeded inputs.  Cases follow the reference tests
(test/test_operators_sparse.jl:318-473, test_operators_lazytensor.jl:237-496,
test_operators_lazysum.jl:255-347, test_operators_dense.jl:342-419).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O
import qobhip as Q
from test_oracle import AB, LT_CASES, cg, rand_lt, rs

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel(a, b):
    return O.rel_frobenius(a, b)


def gb(d):
    return Q.GenericBasis(d)


def comp(shape):
    return Q.tensor(*[gb(d) for d in shape]) if len(shape) > 1 else gb(shape[0])


def to_q(op, bl=None, br=None):
    """Build the qobhip operator equal to an oracle operator."""
    if isinstance(op, O.OSparse):
        M = sp.csc_matrix((op.nzval, op.rowval - 1, op.colptr - 1), shape=(op.m, op.n))
        if op.adj:
            return Q.dagger(Q.SparseOperator(br or gb(op.m), bl or gb(op.n), M))
        return Q.SparseOperator(bl or gb(op.m), br or gb(op.n), M)
    if isinstance(op, O.ODense):
        return Q.DenseOperator(bl or gb(op.shape[0]), br or gb(op.shape[1]), op.mat())
    if isinstance(op, O.OEye):
        return Q.identityoperator(bl or gb(op.m), br or gb(op.n))
    if isinstance(op, O.OLazyTensor):
        bl2, br2 = comp(op.shape_l), comp(op.shape_r)
        ops = tuple(to_q(o, gb(op.shape_l[i - 1]), gb(op.shape_r[i - 1])) for o, i in zip(op.operators, op.indices))
        return Q.LazyTensor(bl2, br2, op.indices, ops, op.factor)
    if isinstance(op, O.OLazySum):
        bl2 = bl or gb(op.m)
        br2 = br or gb(op.n)
        for o in op.operators:
            if isinstance(o, O.OLazyTensor):
                bl2, br2 = comp(o.shape_l), comp(o.shape_r)
        return Q.LazySum(bl2, br2, op.factors, tuple(to_q(o, bl2, br2) for o in op.operators))
    raise TypeError(type(op))


def run_all_paths(oop, qop, g, alpha, beta, dtype=np.complex128, tol=TOL64, ncols=4):
    A_m, A_n = oop.shape
    bl, br = qop.basis_l, qop.basis_r
    # Ket
    b = cg(g, A_n)
    r0 = cg(g, A_m)
    want = r0.copy()
    O.mul_ket(want, oop, b, alpha, beta)
    r = Q.Ket(bl, r0, dtype)
    Q.mul_(r, qop, Q.Ket(br, b, dtype), alpha, beta)
    assert rel(r.to_numpy(), want) < tol, "ket"
    # Bra
    a = cg(g, A_m)
    r0 = cg(g, A_n)
    want = r0.copy()
    O.mul_bra(want, a, oop, alpha, beta)
    r = Q.Bra(br, r0, dtype)
    Q.mul_(r, Q.Bra(bl, a, dtype), qop, alpha, beta)
    assert rel(r.to_numpy(), want) < tol, "bra"
    # Operator left
    cb = gb(ncols)
    B = cg(g, A_n, ncols)
    R0 = cg(g, A_m, ncols)
    want = R0.copy()
    O.mul_op_left(want, oop, B, alpha, beta)
    R = Q.Operator(bl, cb, R0, dtype)
    Q.mul_(R, qop, Q.Operator(br, cb, B, dtype), alpha, beta)
    assert rel(R.to_numpy(), want) < tol, "op-left"
    # Operator right
    B = cg(g, ncols + 1, A_m)
    R0 = cg(g, ncols + 1, A_n)
    want = R0.copy()
    O.mul_op_right(want, B, oop, alpha, beta)
    rb = gb(ncols + 1)
    R = Q.Operator(rb, br, R0, dtype)
    Q.mul_(R, Q.Operator(rb, bl, B, dtype), qop, alpha, beta)
    assert rel(R.to_numpy(), want) < tol, "op-right"


# ------------------------------------------------------------------ sparse
@pytest.mark.parametrize("m,n,density", [(3, 5, 0.5), (50, 60, 0.3), (4, 4, 0.5)])
@pytest.mark.parametrize("alpha,beta", AB)
@pytest.mark.parametrize("adj", [False, True])
def test_sparse_operator(m, n, density, alpha, beta, adj):
    g = np.random.default_rng(100 + m)
    oop = O.OSparse.from_scipy(rs(g, m, n, density))
    if adj:
        oop = O.dagger(oop)
    run_all_paths(oop, to_q(oop), g, alpha, beta, ncols=7)


def test_sparse_operator_c32():
    g = np.random.default_rng(5)
    oop = O.OSparse.from_scipy(rs(g, 50, 60, 0.3))
    run_all_paths(oop, to_q(oop), g, 1.5, 2.1, np.complex64, TOL32, ncols=7)


# -------------------------------------------------------------- LazyTensor
@pytest.mark.parametrize("case", range(len(LT_CASES)))
@pytest.mark.parametrize("alpha,beta", AB)
def test_lazytensor(case, alpha, beta):
    shl, shr, idx, kinds = LT_CASES[case]
    g = np.random.default_rng(10 + case)
    lt = rand_lt(g, shl, shr, idx, kinds)
    run_all_paths(lt, to_q(lt), g, alpha, beta)


@pytest.mark.parametrize("case", range(len(LT_CASES)))
@pytest.mark.parametrize("materialize", [1, 0])
def test_lazytensor_on_rho_materialized(case, materialize):
    """Ten columns (left) and eleven rows (right): an all-sparse LazyTensor
    runs as its sparse matrix (Eye factors and implicit isometries
    included), or through the gather with materialisation off."""
    import ctypes as C
    from qobhip import _lib
    shl, shr, idx, kinds = LT_CASES[case]
    g = np.random.default_rng(50 + case)
    lt = rand_lt(g, shl, shr, idx, kinds)
    v = C.c_int64()
    _lib.call("qob_tune_get", b"ls_materialize", C.byref(v))
    try:
        _lib.call("qob_tune_set", b"ls_materialize", materialize)
        run_all_paths(lt, to_q(lt), g, 0.3 - 0.6j, 0.5 + 0.1j, ncols=10)
    finally:
        _lib.call("qob_tune_set", b"ls_materialize", v.value)


@pytest.mark.parametrize("case", [0, 1, 4, 7])
def test_lazytensor_c32(case):
    shl, shr, idx, kinds = LT_CASES[case]
    g = np.random.default_rng(10 + case)
    lt = rand_lt(g, shl, shr, idx, kinds)
    run_all_paths(lt, to_q(lt), g, 1.5, 2.1, np.complex64, TOL32)


def test_lazytensor_alpha_zero_beta_zero_nan():
    g = np.random.default_rng(21)
    lt = rand_lt(g, (2, 3), (2, 3), (1, 2), ("sparse", "sparse"))
    q = to_q(lt)
    r = Q.Ket(q.basis_l, np.full(6, np.nan, complex))
    Q.mul_(r, q, Q.Ket(q.basis_r, cg(g, 6)), 0.0, 0.0)
    assert (r.to_numpy() == 0).all()
    r = Q.Ket(q.basis_l, np.full(6, np.nan, complex))
    Q.mul_(r, q, Q.Ket(q.basis_r, cg(g, 6)), 1.0, 0.0)
    assert np.isfinite(r.to_numpy()).all()


def test_lazytensor_larger_spin_chain_shapes():
    # 12 qubits, mixed terms: exercises digit extraction on many sites
    g = np.random.default_rng(22)
    sh = (2,) * 12
    lt = rand_lt(g, sh, sh, (1, 5, 12), ("sparse", "sparse", "sparse"))
    run_all_paths(lt, to_q(lt), g, 0.3 + 0.1j, 0.5, ncols=3)


# ---------------------------------------------------------------- LazySum
def lazysum_case(g):
    shl = shr = (2, 3, 2)
    ops = [rand_lt(g, shl, shr, (1, 2), ("sparse", "sparse")),
           rand_lt(g, shl, shr, (3,), ("dense",)),
           rand_lt(g, shl, shr, (2, 3), ("sparse", "sparse")),
           O.OSparse.from_scipy(rs(g, 12, 12, 0.3)),
           O.OSparse.from_scipy(rs(g, 12, 12, 0.2))]
    return O.OLazySum(12, 12, [0.1, 0.3, 1.2 - 0.4j, -0.7, 0.25j], ops)


@pytest.mark.parametrize("mode", [Q.QOB_MODE_FUSED, Q.QOB_MODE_PLAIN])
@pytest.mark.parametrize("alpha,beta", AB)
def test_lazysum(mode, alpha, beta):
    g = np.random.default_rng(30)
    ls = lazysum_case(g)
    ctx = Q.context()
    ctx.set_mode(mode)
    try:
        run_all_paths(ls, to_q(ls), g, alpha, beta)
    finally:
        ctx.set_mode(Q.QOB_MODE_FUSED)


def _materialize_case(g):
    """LazySum on shl = (2, 3, 2) -> shr = (2, 4, 2): two all-sparse terms on
    the same sites (their entries overlap and are summed), a term whose
    unfactored middle site is a truncating identity (3 x 4), a SparseOperator
    term, and a dense-factor term that stays outside the summed matrix."""
    shl, shr = (2, 3, 2), (2, 4, 2)
    ops = [rand_lt(g, shl, shr, (2, 3), ("sparse", "sparse")),
           rand_lt(g, shl, shr, (2, 3), ("sparse", "sparse")),
           rand_lt(g, shl, shr, (1,), ("sparse",)),
           O.OSparse.from_scipy(rs(g, 12, 16, 0.2)),
           rand_lt(g, shl, shr, (3,), ("dense",))]
    return O.OLazySum(12, 16, [0.3, -1.1 + 0.2j, 0.7j, 0.5, -0.4], ops)


@pytest.mark.parametrize("alpha,beta", AB)
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_lazysum_materialized_on_rho(alpha, beta, dtype, tol):
    """With ten columns (left) or nine rows (right) an all-sparse LazySum runs
    as its summed sparse matrix; against the oracle, and against the gather
    (materialisation off) to rounding."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(33)
    ls = _materialize_case(g)
    v = C.c_int64()
    _lib.call("qob_tune_get", b"ls_materialize", C.byref(v))
    got = {}
    try:
        for mat in (1, 0):
            _lib.call("qob_tune_set", b"ls_materialize", mat)
            q = to_q(ls)
            bl, br = q.basis_l, q.basis_r
            B = cg(np.random.default_rng(5), 16, 10)
            R0 = np.full((12, 10), np.nan, complex) if beta == 0 else cg(np.random.default_rng(6), 12, 10)
            want = np.zeros((12, 10), complex) if beta == 0 else R0.copy()
            O.mul_op_left(want, ls, B, alpha, beta)
            R = Q.Operator(bl, gb(10), R0, dtype)
            Q.mul_(R, q, Q.Operator(br, gb(10), B, dtype), alpha, beta)
            assert rel(R.to_numpy(), want) < tol, "left"
            B2 = cg(np.random.default_rng(7), 9, 12)
            R2 = np.full((9, 16), np.nan, complex) if beta == 0 else cg(np.random.default_rng(8), 9, 16)
            want2 = np.zeros((9, 16), complex) if beta == 0 else R2.copy()
            O.mul_op_right(want2, B2, ls, alpha, beta)
            R = Q.Operator(gb(9), br, R2, dtype)
            Q.mul_(R, Q.Operator(gb(9), bl, B2, dtype), q, alpha, beta)
            assert rel(R.to_numpy(), want2) < tol, "right"
            got[mat] = R.to_numpy()
    finally:
        _lib.call("qob_tune_set", b"ls_materialize", v.value)
    assert rel(got[1], got[0]) < tol


def test_lazysum_materialized_follows_set_factors():
    """The summed matrix is rebuilt when the coefficients change on the host."""
    g = np.random.default_rng(34)
    ls = _materialize_case(g)
    q = to_q(ls)
    B = cg(g, 16, 10)
    for factors in ([0.3, -1.1 + 0.2j, 0.7j, 0.5, -0.4], [1.0, 0.0, -2.0, 0.25j, 0.1]):
        q.set_factors(factors)
        ls2 = O.OLazySum(12, 16, factors, ls.operators)
        want = np.zeros((12, 10), complex)
        O.mul_op_left(want, ls2, B, 1.0, 0.0)
        R = Q.Operator(q.basis_l, gb(10), np.zeros((12, 10), complex))
        Q.mul_(R, q, Q.Operator(q.basis_r, gb(10), B))
        assert rel(R.to_numpy(), want) < TOL64


def test_lazysum_zero_and_nan():
    g = np.random.default_rng(31)
    B = comp((2, 3))
    z = Q.LazySum.zero(B, B)
    r0 = cg(g, 6)
    r = Q.Ket(B, r0)
    Q.mul_(r, z, Q.Ket(B, cg(g, 6)), 1.5, 2.1)
    assert rel(r.to_numpy(), 2.1 * r0) < TOL64
    R = Q.Operator(B, B, np.full((6, 6), np.nan, complex))
    Q.mul_(R, Q.Operator(B, B, cg(g, 6, 6)), z, 1.0, 0.0)
    assert (R.to_numpy() == 0).all()
    ls = lazysum_case(g)
    q = to_q(ls)
    R = Q.Operator(q.basis_l, q.basis_r, np.full((12, 12), np.nan, complex))
    Bm = cg(g, 12, 12)
    Q.mul_(R, q, Q.Operator(q.basis_r, q.basis_r, Bm), 1.0, 0.0)
    want = np.zeros((12, 12), complex)
    O.mul_op_left(want, ls, Bm, 1.0, 0.0)
    assert rel(R.to_numpy(), want) < TOL64


def test_lazysum_set_factors():
    g = np.random.default_rng(32)
    ls = lazysum_case(g)
    q = to_q(ls)
    b = cg(g, 12)
    r = Q.Ket(q.basis_l, np.zeros(12, complex))
    kb = Q.Ket(q.basis_r, b)
    Q.mul_(r, q, kb)
    new = [0.5, -1.0, 0.2, 0.3j, 1.0]
    q.set_factors(new)
    ls.factors = new
    Q.mul_(r, q, kb)
    want = np.zeros(12, complex)
    O.mul_ket(want, ls, b)
    assert rel(r.to_numpy(), want) < TOL64


# ------------------------------------------------------------------ dense
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (complex(0.7, 1.5), complex(0.3, 2.1))])
def test_dense_operator_paths(alpha, beta):
    g = np.random.default_rng(40)
    b1, b2, b3 = gb(37), gb(53), gb(41)
    A = cg(g, 37, 53)
    opq = Q.DenseOperator(b1, b2, A)
    x = cg(g, 53)
    r0 = cg(g, 37)
    r = Q.Ket(b1, r0)
    Q.mul_(r, opq, Q.Ket(b2, x), alpha, beta)
    assert rel(r.to_numpy(), alpha * A @ x + beta * r0) < TOL64
    y = cg(g, 37)
    r0 = cg(g, 53)
    r = Q.Bra(b2, r0)
    Q.mul_(r, Q.Bra(b1, y), opq, alpha, beta)
    assert rel(r.to_numpy(), alpha * A.T @ y + beta * r0) < TOL64
    B = cg(g, 53, 41)
    R0 = cg(g, 37, 41)
    R = Q.Operator(b1, b3, R0)
    Q.mul_(R, opq, Q.Operator(b2, b3, B), alpha, beta)
    assert rel(R.to_numpy(), alpha * A @ B + beta * R0) < TOL64


# ------------------------------------------------------------------ errors
def test_dimension_mismatch_and_alias():
    g = np.random.default_rng(50)
    b1, b2, b3 = Q.NLevelBasis(2), Q.NLevelBasis(3), Q.NLevelBasis(4)
    S = Q.SparseOperator(b2, b2, rs(g, 3, 3, 0.6))
    with pytest.raises(Q.DimensionMismatch):
        Q.mul_(Q.Ket(b1, cg(g, 2)), S, Q.Ket(b3, cg(g, 4)))
    with pytest.raises(Q.DimensionMismatch):
        Q.mul_(Q.Operator(b1, b1, cg(g, 2, 2)), S, Q.Operator(b3, b3, cg(g, 4, 4)))
    k = Q.Ket(b2, cg(g, 3))
    with pytest.raises(Q.ArgumentError):
        Q.mul_(k, S, k)


# --------------------------------------------------------------- reductions
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 1e-6)])
def test_reductions(dtype, tol):
    g = np.random.default_rng(60)
    n = 1000
    x, y = cg(g, n).astype(dtype), cg(g, n).astype(dtype)
    b = gb(n)
    kx, ky = Q.Ket(b, x), Q.Ket(b, y)
    xd, yd = x.astype(np.complex128), y.astype(np.complex128)
    assert abs(Q.norm(kx) - np.linalg.norm(xd)) < tol * np.linalg.norm(xd)
    assert abs(Q.dot(kx, ky) - np.vdot(xd, yd)) < tol * np.linalg.norm(xd) * np.linalg.norm(yd)
    A = cg(g, 70, 70).astype(dtype)
    op = Q.Operator(gb(70), gb(70), A)
    assert abs(Q.tr(op) - np.trace(A.astype(np.complex128))) < tol * np.abs(A).sum()
    Q.normalize_(kx)
    assert abs(np.linalg.norm(kx.to_numpy().astype(np.complex128)) - 1) < 10 * tol
    Q.normalize_(op)
    assert abs(np.trace(op.to_numpy().astype(np.complex128)) - 1) < 10 * tol


def test_expect_matches_oracle():
    g = np.random.default_rng(61)
    lt = rand_lt(g, (3, 4), (3, 4), (1, 2), ("sparse", "sparse"))
    q = to_q(lt)
    psi = cg(g, 12)
    want = O.expect_ket(lt, psi)
    got = Q.expect(q, Q.Ket(q.basis_r, psi))
    assert abs(got - want) < 1e-12 * abs(want)
    rho = cg(g, 12, 12)
    want = O.expect_op(lt, rho)
    got = Q.expect(q, Q.Operator(q.basis_r, q.basis_r, rho))
    assert abs(got - want) < 1e-12 * abs(want)


# ------------------------------------------------ fused dense LazyTensor
FUSED_CASES = [
    ((16,), (1,)),
    ((16, 32), (1, 2)),
    ((32, 16), (1, 2)),
    ((16, 2, 16), (1, 3)),      # identity site between factors
    ((2, 16, 16), (2, 3)),      # identity fastest site: n1 = 2
    ((16, 32, 16), (1, 2, 3)),
    ((64, 16), (1, 2)),
    ((16, 16, 32), (2, 3)),     # no site-1 factor: columns as tiles of 16 rows
    ((16, 32, 16), (2, 3)),
    ((16, 16, 16), (2, 3)),
]


def _tune(name, v):
    from qobhip import _lib
    _lib.call("qob_tune_set", name.encode(), v)


@pytest.mark.parametrize("case", range(len(FUSED_CASES)))
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_fused_dense_lazytensor(case, dtype, tol):
    shape, idx = FUSED_CASES[case]
    g = np.random.default_rng(70 + case)
    lt = rand_lt(g, shape, shape, idx, ("dense",) * len(idx))
    q = to_q(lt)
    run_all_paths(lt, q, g, 0.8 - 0.3j, 0.4 + 0.1j, dtype, tol, ncols=5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_fused_equals_chain(dtype):
    g = np.random.default_rng(80)
    lt = rand_lt(g, (16, 32, 16), (16, 32, 16), (1, 2, 3), ("dense",) * 3)
    q = to_q(lt)
    B = cg(g, 8192, 3)
    outs = []
    for fused in (1, 0):
        _tune("fused_lt", fused)
        try:
            R = Q.Operator(q.basis_l, gb(3), np.zeros((8192, 3)), dtype)
            Q.mul_(R, q, Q.Operator(q.basis_r, gb(3), B, dtype), 1.0, 0.0)
            outs.append(R.to_numpy().astype(np.complex128))
        finally:
            _tune("fused_lt", 1)
    tol = TOL64 if dtype == np.complex128 else TOL32
    assert rel(outs[0], outs[1]) < tol


@pytest.mark.parametrize("shape,idx", [((16, 16), (1, 2)),            # D = 256: lanes past D in the column
                                       ((16, 3, 16), (1, 3)),         # D = 768: second element per lane partly past D
                                       ((16, 32, 16), (1, 2, 3)),     # c64: planes + factors fill the 160 KiB LDS
                                       ((64, 16), (1, 2))])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_fused_resident_factors(shape, idx, dtype, tol):
    """Factors resident in LDS for the whole launch against the per-mode
    staging, with two fibre groups per wave in both: bitwise equal (same
    products in the same order), left on columns and right on 16-row tiles.
    The automatic launch shape (ComplexF64: one fibre group per wave and
    Gauss's three-product form when resident; ComplexF32 at D = 8192: 1024
    threads) against the oracle."""
    g = np.random.default_rng(90 + sum(shape))
    lt = rand_lt(g, shape, shape, idx, ("dense",) * len(idx))
    D = int(np.prod(shape))
    B = cg(g, D, 5)
    Bt = (g.standard_normal((64, D)) + 1j * g.standard_normal((64, D))) / np.sqrt(D)
    R0 = g.standard_normal((D, 5)) + 0j
    want = R0.copy()
    O.mul_op_left(want, lt, B, 0.6 + 0.2j, -0.5j)
    want_r = np.zeros((64, D), complex)
    O.mul_op_right(want_r, Bt, lt, 1.0, 0.0)
    got = {}
    try:
        for gpw, res in ((2, 1), (2, 0), (0, 1)):
            _tune("fused_gpw", gpw)
            _tune("fused_resident", res)
            q = to_q(lt)
            R = Q.Operator(q.basis_l, gb(5), R0, dtype)
            Q.mul_(R, q, Q.Operator(q.basis_r, gb(5), B, dtype), 0.6 + 0.2j, -0.5j)
            Rt = Q.Operator(gb(64), q.basis_r, np.full((64, D), np.nan, complex), dtype)
            Q.mul_(Rt, Q.Operator(gb(64), q.basis_l, Bt, dtype), q, 1.0, 0.0)
            got[gpw, res] = (R.to_numpy(), Rt.to_numpy())
    finally:
        _tune("fused_resident", 1)
        _tune("fused_gpw", 0)
    for key in ((2, 1), (0, 1)):
        assert rel(got[key][0], want) < tol
        assert rel(got[key][1], want_r) < tol
    assert np.array_equal(got[2, 1][0], got[2, 0][0])
    assert np.array_equal(got[2, 1][1], got[2, 0][1])
    if dtype == np.complex64:
        # ComplexF32 runs the same three-product tiles in every shape (the
        # automatic one is 1024 threads, one group per wave at D = 8192)
        assert np.array_equal(got[0, 1][0], got[2, 1][0])
        assert np.array_equal(got[0, 1][1], got[2, 1][1])


@pytest.mark.parametrize("gpw", [0, 1, 2, 4])
@pytest.mark.parametrize("resident", [1, 0])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_fused_every_launch_shape(gpw, resident, dtype, tol):
    """Every fibre-groups-per-wave setting with and without resident
    factors (four or three real products per complex k-step, 512 or 1024
    threads) on the cfg3 factor shape, left on columns and right on tiles,
    against the oracle."""
    g = np.random.default_rng(95)
    shape, idx = (16, 32, 16), (1, 2, 3)
    lt = rand_lt(g, shape, shape, idx, ("dense",) * 3)
    D = 8192
    B = cg(g, D, 3)
    Bt = (g.standard_normal((64, D)) + 1j * g.standard_normal((64, D))) / np.sqrt(D)  # 64 rows: tiled path
    want = np.zeros((D, 3), complex)
    O.mul_op_left(want, lt, B, 1.0, 0.0)
    want_r = np.zeros((64, D), complex)
    O.mul_op_right(want_r, Bt, lt, 0.5j, 0.0)
    try:
        _tune("fused_gpw", gpw)
        _tune("fused_resident", resident)
        q = to_q(lt)
        R = Q.Operator(q.basis_l, gb(3), np.full((D, 3), np.nan, complex), dtype)
        Q.mul_(R, q, Q.Operator(q.basis_r, gb(3), B, dtype), 1.0, 0.0)
        Rt = Q.Operator(gb(64), q.basis_r, np.full((64, D), np.nan, complex), dtype)
        Q.mul_(Rt, Q.Operator(gb(64), q.basis_l, Bt, dtype), q, 0.5j, 0.0)
    finally:
        _tune("fused_resident", 1)
        _tune("fused_gpw", 0)
    assert rel(R.to_numpy(), want) < tol
    assert rel(Rt.to_numpy(), want_r) < tol


@pytest.mark.parametrize("rows_tunable,waves,sweep", [(16, 1, 0), (32, 1, 0), (64, 1, 0), (64, 1, 1), (64, 2, 0),
                                                      (64, 4, 1), (64, 4, 2)])
@pytest.mark.parametrize("cc", [1, 3])
@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_sparse_right_stripes(rows_tunable, waves, sweep, cc, adj, dtype, tol):
    """rho * SparseOperator through every stripe height, waves per workgroup
    and sweep of the CSC right kernel (rows >= 32, so the wide kernels run),
    beta = 0 over NaN and beta != 0."""
    from qobhip import _lib
    import ctypes as C
    g = np.random.default_rng(31 + rows_tunable + cc)
    m, n, rows = 150, 97, 77
    M = sp.random(m, n, density=0.05, random_state=int(g.integers(1 << 30)), format="csc")
    M = M + 1j * sp.random(m, n, density=0.05, random_state=int(g.integers(1 << 30)), format="csc")
    oop = O.OSparse.from_scipy(sp.csc_matrix(M))
    if adj:
        oop = O.dagger(oop)
    qop = to_q(oop)
    mo, no = oop.shape
    B = g.standard_normal((rows, mo)) + 1j * g.standard_normal((rows, mo))
    saved = {}
    for k in ("csc_rows", "csc_cc", "csc_waves", "csc_sweep"):
        v = C.c_int64()
        _lib.call("qob_tune_get", k.encode(), C.byref(v))
        saved[k] = v.value
    try:
        _lib.call("qob_tune_set", b"csc_rows", rows_tunable)
        _lib.call("qob_tune_set", b"csc_cc", cc)
        _lib.call("qob_tune_set", b"csc_waves", waves)
        _lib.call("qob_tune_set", b"csc_sweep", sweep)
        for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
            R0 = np.full((rows, no), np.nan, complex) if beta == 0 else g.standard_normal((rows, no)) + 0j
            want = np.zeros((rows, no), complex) if beta == 0 else R0.copy()
            O.mul_op_right(want, B, oop, alpha, beta)
            bl, br = gb(rows), gb(mo)
            R = Q.Operator(bl, gb(no), R0, dtype)
            Q.mul_(R, Q.Operator(bl, br, B, dtype), qop, alpha, beta)
            assert rel(R.to_numpy(), want) < tol
    finally:
        for k, v in saved.items():
            _lib.call("qob_tune_set", k.encode(), v)


@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_sparse_right_reordered_columns(adj, dtype, tol):
    """rho * X for a banded X with rows and columns relabelled at random, so
    the right kernel sweeps columns in its reverse Cuthill-McKee order (the
    order is checked to be in use)."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(57)
    n, rows = 900, 70
    vals = [g.standard_normal(n - k) + 1j * g.standard_normal(n - k) for k in (2, 0, 1)]
    Xb = sp.diags(vals, [-2, 0, 1], format="csc")
    p = g.permutation(n)
    X = sp.csc_matrix(Xb[p][:, p])
    X.sort_indices()
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    # the order is built from the columns the kernel sweeps: those of op
    cols = sp.csc_matrix(X.conj().T) if adj else X
    cols.sort_indices()
    order, used = np.empty(n, np.int32), C.c_int()
    _lib.call("qob_csc_column_order", n, n, np.ascontiguousarray(cols.indptr, np.int32).ctypes.data,
              np.ascontiguousarray(cols.indices, np.int32).ctypes.data, order.ctypes.data, C.byref(used))
    assert used.value == 1
    qop = to_q(oop)
    B = g.standard_normal((rows, n)) + 1j * g.standard_normal((rows, n))
    for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
        R0 = np.full((rows, n), np.nan, complex) if beta == 0 else g.standard_normal((rows, n)) + 0j
        want = np.zeros((rows, n), complex) if beta == 0 else R0.copy()
        O.mul_op_right(want, B, oop, alpha, beta)
        R = Q.Operator(gb(rows), gb(n), R0, dtype)
        Q.mul_(R, Q.Operator(gb(rows), gb(n), B, dtype), qop, alpha, beta)
        assert rel(R.to_numpy(), want) < tol


@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_sparse_left_diagonal_aligned(adj, dtype, tol):
    """X * B for a banded X with a third of its band entries removed at
    random: 64-row slices hold few distinct diagonals, so the left kernel's
    table puts each diagonal in its own slot with holes where a row lacks it.
    Ket, Bra and Operator, beta = 0 over NaN and beta != 0, against the
    oracle; aligned and packed tables agree bit for bit."""
    from qobhip import _lib
    import ctypes as C
    g = np.random.default_rng(71)
    n, cols = 1000, 9
    offs = [-130, -7, 0, 3, 64]
    X = sp.diags([g.standard_normal(n - abs(k)) + 1j * g.standard_normal(n - abs(k)) for k in offs], offs,
                 format="csr")
    keep = g.random(X.nnz) > 0.33
    X = sp.csr_matrix((X.data * keep, X.indices, X.indptr), shape=X.shape)
    X.eliminate_zeros()
    X = sp.csc_matrix(X)
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    B = g.standard_normal((n, cols)) + 1j * g.standard_normal((n, cols))
    v = C.c_int64()
    _lib.call("qob_tune_get", b"sell_align", C.byref(v))
    outs = []
    try:
        for align in (1, 0):
            _lib.call("qob_tune_set", b"sell_align", align)
            qop = to_q(oop)   # tables are built on first use, under this setting
            for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
                R0 = np.full((n, cols), np.nan, complex) if beta == 0 else g.standard_normal((n, cols)) + 0j
                want = np.zeros((n, cols), complex) if beta == 0 else R0.copy()
                O.mul_op_left(want, oop, B, alpha, beta)
                R = Q.Operator(gb(n), gb(cols), R0, dtype)
                Q.mul_(R, qop, Q.Operator(gb(n), gb(cols), B, dtype), alpha, beta)
                assert rel(R.to_numpy(), want) < tol
                outs.append(R.to_numpy())
            psi = B[:, 0].copy()
            want = np.zeros(n, complex)
            O.mul_ket(want, oop, psi, 1.0, 0.0)
            r = Q.Ket(gb(n), np.full(n, np.nan, complex), dtype)
            Q.mul_(r, qop, Q.Ket(gb(n), psi, dtype))
            assert rel(r.to_numpy(), want) < tol
            want = np.zeros(n, complex)
            O.mul_bra(want, psi, oop, 1.0, 0.0)
            r = Q.Bra(gb(n), np.zeros(n, complex), dtype)
            Q.mul_(r, Q.Bra(gb(n), psi, dtype), qop)
            assert rel(r.to_numpy(), want) < tol
    finally:
        _lib.call("qob_tune_set", b"sell_align", v.value)
    # beta = 0 outputs do not depend on R0; the beta != 0 ones used different R0 draws
    assert np.array_equal(outs[0], outs[2])


def _block_matrix(g, sizes, n_extra_isolated=0):
    """Square sparse matrix whose pattern splits into connected blocks of the
    given sizes (each block a random connected pattern), rows and columns
    relabelled by one random permutation, plus isolated empty rows."""
    blocks = []
    for k in sizes:
        M = np.zeros((k, k), complex)
        for i in range(k):  # a path keeps the block connected
            M[i, i] = g.standard_normal() + 1j * g.standard_normal()
            if i + 1 < k:
                M[i, i + 1] = g.standard_normal() + 1j * g.standard_normal()
        extra = g.random((k, k)) < min(0.3, 3.0 / k)
        M[extra] = g.standard_normal(extra.sum()) + 1j * g.standard_normal(extra.sum())
        blocks.append(sp.csr_matrix(M))
    if n_extra_isolated:
        blocks.append(sp.csr_matrix((n_extra_isolated, n_extra_isolated), dtype=complex))
    X = sp.block_diag(blocks, format="csr")
    p = g.permutation(X.shape[0])
    X = sp.csc_matrix(X[p][:, p])
    X.sort_indices()
    return X


@pytest.mark.parametrize("sizes", [[4, 1, 3, 2, 4] * 40, [2, 1] * 90, [64, 1, 63, 30, 40, 2], [256, 3, 200, 60] * 2,
                                   [5] * 30 + [300]])
@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_sparse_left_closed_slices(sizes, adj, dtype, tol):
    """X * B, X * psi and psi' * X for block-structured X: blocks of at most
    256 rows (and at most 16 entries per row) are packed whole into 256-row
    workgroup blocks and read B once through LDS; a block of 300 rows makes
    the table fall back to contiguous slices.  Each layout is built in turn
    and checked against the oracle with beta = 0 over NaN and beta != 0,
    and bit for bit against the contiguous-slice table."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(sum(sizes) + 7 * len(sizes))
    X = _block_matrix(g, sizes, n_extra_isolated=5)
    n, cols = X.shape[0], 11
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    B = g.standard_normal((n, cols)) + 1j * g.standard_normal((n, cols))
    v = C.c_int64()
    _lib.call("qob_tune_get", b"sell_closed", C.byref(v))
    outs = []
    try:
        for closed in (1, 0):
            _lib.call("qob_tune_set", b"sell_closed", closed)
            qop = to_q(oop)   # tables are built on first use, under these settings
            for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
                R0 = np.full((n, cols), np.nan, complex) if beta == 0 else g.standard_normal((n, cols)) + 0j
                want = np.zeros((n, cols), complex) if beta == 0 else R0.copy()
                O.mul_op_left(want, oop, B, alpha, beta)
                R = Q.Operator(gb(n), gb(cols), R0, dtype)
                Q.mul_(R, qop, Q.Operator(gb(n), gb(cols), B, dtype), alpha, beta)
                assert rel(R.to_numpy(), want) < tol
                if beta == 0:
                    outs.append(R.to_numpy())
            psi = B[:, 0].copy()
            want = np.zeros(n, complex)
            O.mul_ket(want, oop, psi, 0.7, 0.0)
            r = Q.Ket(gb(n), np.full(n, np.nan, complex), dtype)
            Q.mul_(r, qop, Q.Ket(gb(n), psi, dtype), 0.7, 0.0)
            assert rel(r.to_numpy(), want) < tol
            want = np.zeros(n, complex)
            O.mul_bra(want, psi, oop, 1.0, 0.0)
            r = Q.Bra(gb(n), np.zeros(n, complex), dtype)
            Q.mul_(r, Q.Bra(gb(n), psi, dtype), qop)
            assert rel(r.to_numpy(), want) < tol
    finally:
        _lib.call("qob_tune_set", b"sell_closed", v.value)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def _line_groups_matrix(g, n=1280, shift=643):
    """Square sparse X whose rows, joined within aligned groups of 8 as
    well as by the pattern, form components that pack 320-row blocks at
    least 90% full: a path inside every aligned group of 8 rows and a
    coupling of row i with row i + shift (an offset that is not a multiple
    of 8) for the first row of every group of the first half, so each
    component is two groups of 8."""
    rows, cols = [], []
    for i in range(n):
        rows.append(i); cols.append(i)
        if i % 8 != 7:
            rows.append(i); cols.append(i + 1)
    for i in range(0, n // 2, 8):
        rows.append(i); cols.append(i + shift)
    v = g.standard_normal(len(rows)) + 1j * g.standard_normal(len(rows))
    X = sp.csc_matrix((v, (rows, cols)), shape=(n, n))
    X.sort_indices()
    return X


def _line_layout_fill(X, grp=8, blk=320):
    """Largest component and block fill of the line-aligned closed layout."""
    from scipy.sparse.csgraph import connected_components
    n = X.shape[0]
    A = (abs(X) + abs(X.T)).tocsr()
    i = np.arange(n)
    G = sp.csr_matrix((np.ones(n), (i, i - i % grp)), shape=(n, n))
    _, lab = connected_components((A + G + G.T).tocsr(), directed=False)
    sizes = np.bincount(lab)
    first = {}
    for r in range(n):
        first.setdefault(lab[r], r)
    blocks, used = 0, blk
    for c in sorted(first, key=first.get):
        if used + sizes[c] > blk:
            blocks, used = blocks + 1, 0
        used += sizes[c]
    return sizes.max(), n / (blocks * blk)


@pytest.mark.parametrize("adj", [False, True])
def test_sparse_left_closed_line_aligned(adj):
    """ComplexF64 closed blocks whose rows are also joined within aligned
    8-row groups (every run of a block is whole 128-byte lines of R), packed
    into 320-row blocks: against the oracle, and bit for bit against the
    256-row closed form and the contiguous slices."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(23)
    X = _line_groups_matrix(g)
    big, fill = _line_layout_fill(X)
    assert big <= 320 and fill >= 0.9, (big, fill)  # the line-aligned form applies
    n, cols = X.shape[0], 9
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    B = g.standard_normal((n, cols)) + 1j * g.standard_normal((n, cols))
    want = np.zeros((n, cols), complex)
    O.mul_op_left(want, oop, B, 1.0, 0.0)
    saved = {}
    for name in (b"sell_closed", b"sell_cline"):
        v = C.c_int64()
        _lib.call("qob_tune_get", name, C.byref(v))
        saved[name] = v.value
    outs = []
    try:
        for closed, cline in ((1, 1), (1, 0), (0, 0)):
            _lib.call("qob_tune_set", b"sell_closed", closed)
            _lib.call("qob_tune_set", b"sell_cline", cline)
            qop = to_q(oop)   # tables are built on first use, under these settings
            R = Q.Operator(gb(n), gb(cols), np.full((n, cols), np.nan, complex))
            Q.mul_(R, qop, Q.Operator(gb(n), gb(cols), B), 1.0, 0.0)
            assert rel(R.to_numpy(), want) < TOL64
            outs.append(R.to_numpy())
    finally:
        for name, v in saved.items():
            _lib.call("qob_tune_set", name, v)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def _chains_matrix(g, n, length, stride):
    """Square sparse X made of chains of `length` rows: row i couples with
    row i + stride inside its chain of rows {c, c + stride, ...}; chain c
    starts at c, so aligned pairs of rows fall into different chains unless
    they are joined as pairs.  Up to 4 entries per row (diagonal, both
    neighbours, and a coupling two steps along)."""
    rows, cols = [], []
    for i in range(n):
        rows.append(i); cols.append(i)
        for d in (stride, -stride, 2 * stride):
            j = i + d
            if 0 <= j < n and (j // (stride * length)) == (i // (stride * length)):
                rows.append(i); cols.append(j)
    v = g.standard_normal(len(rows)) + 1j * g.standard_normal(len(rows))
    X = sp.csc_matrix((v, (rows, cols)), shape=(n, n))
    X.sort_indices()
    return X


@pytest.mark.parametrize("case", ["pairs256", "pairs320", "odd"])
@pytest.mark.parametrize("adj", [False, True])
def test_sparse_left_closed_row_pairs(case, adj):
    """ComplexF32 closed blocks with rows also joined in aligned pairs, so a
    thread moves rows 2i and 2i + 1 with 16-byte accesses: 256-row blocks
    (components of 16 rows, the last block partly idle), 320-row blocks
    (components of 160 rows, which fill 256-row blocks to 62%) and an odd row
    count (no pairs, one row per lane).  Against the oracle with beta = 0
    over NaN and beta != 0, and bit for bit against the one-row-per-lane
    closed form and the contiguous slices."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng({"pairs256": 31, "pairs320": 32, "odd": 33}[case])
    if case == "pairs256":
        X = _line_groups_matrix(g, n=1232, shift=611)
        big, fill = _line_layout_fill(X, grp=2, blk=256)
        assert big <= 256 and fill >= 0.9 and X.shape[0] % 256, (big, fill)
    elif case == "pairs320":
        X = _chains_matrix(g, 1600, 80, 2)   # pairs join two chains of 80 rows
        assert _line_layout_fill(X, grp=2, blk=256)[1] < 0.9
        big, fill = _line_layout_fill(X, grp=2, blk=320)
        assert big == 160 and fill >= 0.9, (big, fill)
    else:
        X = _chains_matrix(g, 1001, 7, 3)
    n, cols = X.shape[0], 11
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    B = g.standard_normal((n, cols)) + 1j * g.standard_normal((n, cols))
    names = (b"sell_closed", b"sell_cpair")
    saved = {}
    for name in names:
        v = C.c_int64()
        _lib.call("qob_tune_get", name, C.byref(v))
        saved[name] = v.value
    outs = []
    try:
        for closed, cpair in ((1, 1), (1, 0), (0, 0)):
            _lib.call("qob_tune_set", b"sell_closed", closed)
            _lib.call("qob_tune_set", b"sell_cpair", cpair)
            qop = to_q(oop)   # tables are built on first use, under these settings
            for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
                R0 = np.full((n, cols), np.nan, complex) if beta == 0 else g.standard_normal((n, cols)) + 0j
                want = np.zeros((n, cols), complex) if beta == 0 else R0.copy()
                O.mul_op_left(want, oop, B, alpha, beta)
                R = Q.Operator(gb(n), gb(cols), R0, np.complex64)
                Q.mul_(R, qop, Q.Operator(gb(n), gb(cols), B, np.complex64), alpha, beta)
                assert rel(R.to_numpy(), want) < TOL32
                if beta == 0:
                    outs.append(R.to_numpy())
    finally:
        for name, v in saved.items():
            _lib.call("qob_tune_set", name, v)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("sizes", [[1, 2, 3, 4] * 50, [8, 7, 1, 5] * 20, [2] * 60, [9] + [3] * 40])
@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
@pytest.mark.parametrize("rows", [70, 133, 200])
def test_sparse_right_component_groups(sizes, adj, dtype, tol, rows):
    """rho * X for block-structured X: components of at most 8 columns run
    as groups (one load of each column of rho per row block); a 9-column
    component leaves the table without groups.  ComplexF32 with an even row
    count runs two rows per lane (70: one partly filled 128-row block, 200:
    two blocks); 133 rows keep one row per lane.  Against the oracle with
    beta = 0 over NaN and beta != 0, and bit for bit against the CSC
    kernel."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(sum(sizes) + 3 * len(sizes))
    X = _block_matrix(g, sizes, n_extra_isolated=3)
    n = X.shape[0]
    oop = O.OSparse.from_scipy(X)
    if adj:
        oop = O.dagger(oop)
    qop = to_q(oop)
    B = g.standard_normal((rows, n)) + 1j * g.standard_normal((rows, n))
    v = C.c_int64()
    _lib.call("qob_tune_get", b"csc_gkernel", C.byref(v))
    outs = []
    try:
        for gk in (1, 0):
            _lib.call("qob_tune_set", b"csc_gkernel", gk)
            for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
                R0 = np.full((rows, n), np.nan, complex) if beta == 0 else g.standard_normal((rows, n)) + 0j
                want = np.zeros((rows, n), complex) if beta == 0 else R0.copy()
                O.mul_op_right(want, B, oop, alpha, beta)
                R = Q.Operator(gb(rows), gb(n), R0, dtype)
                Q.mul_(R, Q.Operator(gb(rows), gb(n), B, dtype), qop, alpha, beta)
                assert rel(R.to_numpy(), want) < tol
                if beta == 0:
                    outs.append(R.to_numpy())
    finally:
        _lib.call("qob_tune_set", b"csc_gkernel", v.value)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("kind", ["zero", "diagonal", "one_block"])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_sparse_rho_degenerate_components(kind, dtype, tol):
    """Left and right products where the component layouts degenerate: an
    all-zero matrix (singleton components, no entries: alpha * 0 + beta R,
    NaN overwritten when beta = 0), a diagonal matrix (singletons with one
    entry) and one dense 300-row block (too large for either layout)."""
    g = np.random.default_rng(91)
    n, k = 300, 40
    if kind == "zero":
        X = sp.csc_matrix((n, n), dtype=complex)
    elif kind == "diagonal":
        X = sp.diags(g.standard_normal(n) + 1j * g.standard_normal(n), format="csc")
    else:
        X = sp.csc_matrix(g.standard_normal((n, n)) * (g.random((n, n)) < 0.02) + 0j)
        X = X + sp.diags(np.ones(n - 1), 1, format="csc")  # keeps it one component
    oop = O.OSparse.from_scipy(sp.csc_matrix(X))
    qop = to_q(oop)
    for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
        B = g.standard_normal((n, k)) + 1j * g.standard_normal((n, k))
        R0 = np.full((n, k), np.nan, complex) if beta == 0 else g.standard_normal((n, k)) + 0j
        want = np.zeros((n, k), complex) if beta == 0 else R0.copy()
        O.mul_op_left(want, oop, B, alpha, beta)
        R = Q.Operator(gb(n), gb(k), R0, dtype)
        Q.mul_(R, qop, Q.Operator(gb(n), gb(k), B, dtype), alpha, beta)
        got = R.to_numpy()
        assert np.isfinite(got).all()
        assert rel(got, want) < tol or np.abs(got - want).max() < tol
        B2 = g.standard_normal((k, n)) + 1j * g.standard_normal((k, n))
        R2 = np.full((k, n), np.nan, complex) if beta == 0 else g.standard_normal((k, n)) + 0j
        want2 = np.zeros((k, n), complex) if beta == 0 else R2.copy()
        O.mul_op_right(want2, B2, oop, alpha, beta)
        R = Q.Operator(gb(k), gb(n), R2, dtype)
        Q.mul_(R, Q.Operator(gb(k), gb(n), B2, dtype), qop, alpha, beta)
        got = R.to_numpy()
        assert np.isfinite(got).all()
        assert rel(got, want2) < tol or np.abs(got - want2).max() < tol


@pytest.mark.parametrize("shape,idx", [((16, 16), (1, 2)), ((16, 16, 16, 16), (1, 2, 3, 4)),
                                       ((16, 16, 16, 16), (2, 4)), ((16, 32, 16), (1, 2, 3)),
                                       ((32, 16, 16), (1, 2, 3))])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_dense_lazytensor_right_tiles(shape, idx, dtype, tol):
    """rho * LazyTensor of dense factors with 64 rows: the leading sites run
    fused over 16-row tiles, the others through the chain first (two of them
    in the 4-site cases, one an identity site inside the tile in the third).
    Tiles of two factors (16 x 16, 16 x 32, 32 x 16) run the register-chained
    tile kernel, the rest the fused column kernel on tiles; against the
    oracle, against the chain alone (tiles off) and against the fused column
    kernel on the tiles (register chain off)."""
    import ctypes as C
    from qobhip import _lib
    g = np.random.default_rng(sum(shape) + len(idx))
    lt = rand_lt(g, shape, shape, idx, ("dense",) * len(idx))
    D = int(np.prod(shape))
    rows = 64
    B = (g.standard_normal((rows, D)) + 1j * g.standard_normal((rows, D))) / np.sqrt(D)
    v, vr = C.c_int64(), C.c_int64()
    _lib.call("qob_tune_get", b"fused_tiles", C.byref(v))
    _lib.call("qob_tune_get", b"fused_reg", C.byref(vr))
    got = {}
    try:
        for tiles, reg in ((1, 1), (0, 1), (1, 0)):
            _lib.call("qob_tune_set", b"fused_tiles", tiles)
            _lib.call("qob_tune_set", b"fused_reg", reg)
            q = to_q(lt)
            for alpha, beta in [(1.0, 0.0), (0.3 - 0.8j, 0.5 + 0.25j)]:
                R0 = np.full((rows, D), np.nan, complex) if beta == 0 else g.standard_normal((rows, D)) + 0j
                want = np.zeros((rows, D), complex) if beta == 0 else R0.copy()
                O.mul_op_right(want, B, lt, alpha, beta)
                R = Q.Operator(gb(rows), q.basis_r, R0, dtype)
                Q.mul_(R, Q.Operator(gb(rows), q.basis_l, B, dtype), q, alpha, beta)
                assert rel(R.to_numpy(), want) < tol
                if beta == 0:
                    got[(tiles, reg)] = R.to_numpy()
    finally:
        _lib.call("qob_tune_set", b"fused_tiles", v.value)
        _lib.call("qob_tune_set", b"fused_reg", vr.value)
    assert rel(got[(1, 1)], got[(0, 1)]) < tol
    assert rel(got[(1, 1)], got[(1, 0)]) < tol


def _q_lazyproduct(lp):
    B = comp((2, 3))
    ops = []
    for o in lp.operators:
        ops.append(to_q(o) if isinstance(o, (O.OLazyTensor, O.OLazySum)) else to_q(o, B, B))
    return Q.LazyProduct(ops, lp.factor)


@pytest.mark.parametrize("alpha,beta", AB)
@pytest.mark.parametrize("dtype,tol", [(np.complex128, TOL64), (np.complex64, TOL32)])
def test_lazyproduct(alpha, beta, dtype, tol):
    from test_oracle import _lp_case
    g = np.random.default_rng(177)
    lp = _lp_case(g)
    qlp = _q_lazyproduct(lp)
    run_all_paths(lp, qlp, g, alpha, beta, dtype=dtype, tol=tol)
    # adjoint, inside a LazySum, and nested in another LazyProduct
    run_all_paths(O.dagger(lp), Q.dagger(qlp), g, alpha, beta, dtype=dtype, tol=tol)
    B = comp((2, 3))
    D = cg(g, 6, 6)
    s_o = O.OLazySum(6, 6, [0.5, 1j], [lp, O.ODense(D)])
    s_q = Q.LazySum(B, B, [0.5, 1j], (qlp, Q.DenseOperator(B, B, D)))
    run_all_paths(s_o, s_q, g, alpha, beta, dtype=dtype, tol=tol)
    n_o = O.OLazyProduct([lp, O.ODense(D), lp], -0.5)
    n_q = Q.LazyProduct([qlp, Q.DenseOperator(B, B, D), qlp], -0.5)
    run_all_paths(n_o, n_q, g, alpha, beta, dtype=dtype, tol=tol)
