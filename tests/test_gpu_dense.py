"""GPU parity of the hand-written dense kernels (k_mode.hip) against the
oracle: DenseOperator mul! on Kets, Bras and density matrices from both
sides (src/operators_dense.jl:394-396), dense mode products of LazyTensors
with factors of any width (src/operators_lazytensor.jl:281-404, through the
chain of :443-488), and qob_gemm.

The shapes are chosen to reach every kernel and branch:
  k_zgemv       Kets and Bras (vectors), with and without split partial sums
  k_mode_chunk  factor widths <= 64 with fewer than 16 fibres along the
                leading index (n1 = 1, 2, 3, ...)
  k_mode_fib    factor widths <= 64 with n1 >= 16 (exact and ragged multiples)
  k_zgemm       factor widths > 64: n1 = 1, n1 >= 64, and 1 < n1 < 64
Widths are off the multiple-of-16 grid (12, 20, 37, 53, 100) as well as on it.

Bar: relative Frobenius error <= 1e-12 (ComplexF64), <= 1e-5 (ComplexF32)
against the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

import qob_oracle as O
import qobhip as Q
from qobhip import _lib
from test_gpu_parity import comp, gb, rel, run_all_paths, to_q
from test_oracle import cg, rand_lt

pytestmark = pytest.mark.gpu

TOLS = {np.complex128: 1e-12, np.complex64: 1e-5}
DTS = [np.complex128, np.complex64]


def _tune(name, v):
    old = C.c_int64()
    _lib.call("qob_tune_get", name.encode(), C.byref(old))
    _lib.call("qob_tune_set", name.encode(), int(v))
    return int(old.value)


@pytest.mark.parametrize("m,n", [(37, 53), (16, 16), (64, 48), (100, 70), (200, 200), (130, 257)])
@pytest.mark.parametrize("ncols", [4, 40, 130])
@pytest.mark.parametrize("dtype", DTS)
def test_dense_operator_every_state_form(m, n, ncols, dtype):
    """A device DenseOperator of the states' element type on a Ket and a Bra
    (matrix-vector kernels), and on a density matrix from the left and from
    the right (tiled GEMM), with (alpha, beta) = (1, 0) and general."""
    g = np.random.default_rng(m * 1000 + n + ncols)
    A = cg(g, m, n).astype(dtype).astype(np.complex128)
    oop = O.ODense(A)
    qop = Q.DenseOperator(gb(m), gb(n), A, dtype)
    for alpha, beta in ((1.0, 0.0), (complex(0.7, 1.5), complex(0.3, -2.1))):
        run_all_paths(oop, qop, g, alpha, beta, dtype, TOLS[dtype], ncols=ncols)


@pytest.mark.parametrize("m,n,ncols", [(37, 53, 4), (100, 70, 40)])
@pytest.mark.parametrize("op_dtype,dtype", [(np.complex128, np.complex64), (np.complex64, np.complex128)])
def test_dense_operator_element_type_differs_from_states(m, n, ncols, op_dtype, dtype):
    """A DenseOperator of one element type on states of the other: the
    product runs in ComplexF64 (the promoted type)."""
    g = np.random.default_rng(m + n)
    A = cg(g, m, n).astype(op_dtype).astype(np.complex128)
    run_all_paths(O.ODense(A), Q.DenseOperator(gb(m), gb(n), A, op_dtype), g, complex(0.7, 1.5), 0.5, dtype,
                  TOLS[dtype], ncols=ncols)


@pytest.mark.parametrize("m,n", [(90, 75), (16, 130)])
@pytest.mark.parametrize("ncols", [5, 33, 70])
@pytest.mark.parametrize("dtype", DTS)
def test_dense_handle_every_state_form(m, n, ncols, dtype):
    """A DenseOperator uploaded as an operator handle (qob_dense_create, as a
    LazyTensor factor or LazySum term is) applied on its own: the mode-product
    dispatch over the whole view (gemv for vectors, chunk / fibre / GEMM
    kernels for density matrices), through its adjoint as well."""
    g = np.random.default_rng(m * 7 + n + ncols)
    A = cg(g, m, n)
    oop = O.OLazySum(m, n, [complex(0.5, 0.5)], [O.ODense(A)])
    qop = Q.LazySum(gb(m), gb(n), (complex(0.5, 0.5),), (Q.DenseOperator(gb(m), gb(n), A),))
    run_all_paths(oop, qop, g, complex(0.2, -0.4), 0.5, dtype, TOLS[dtype], ncols=ncols)
    run_all_paths(O.dagger(oop), Q.dagger(qop), g, 1.0, 0.0, dtype, TOLS[dtype], ncols=ncols)


@pytest.mark.parametrize("dtype", DTS)
def test_dense_operator_adjoint_and_nan(dtype):
    g = np.random.default_rng(5)
    A = cg(g, 90, 75).astype(dtype).astype(np.complex128)
    opq = Q.dagger(Q.DenseOperator(gb(90), gb(75), A, dtype))
    oop = O.dagger(O.ODense(A))
    run_all_paths(oop, opq, g, complex(0.2, -0.4), 0.5, dtype, TOLS[dtype], ncols=33)
    # beta = 0 discards NaN in the result (src/operators_lazytensor.jl:540 semantics)
    x = cg(g, 90, 17)
    r = Q.Operator(gb(75), gb(17), np.full((75, 17), np.nan + 0j), dtype)
    Q.mul_(r, opq, Q.Operator(gb(90), gb(17), x, dtype), 1.0, 0.0)
    assert rel(r.to_numpy(), A.conj().T @ x) < TOLS[dtype]


@pytest.mark.parametrize("N", [1024, 2048])
@pytest.mark.parametrize("dtype", DTS)
def test_dense_operator_on_large_density_matrix(N, dtype):
    """N x N DenseOperator on an N x N density matrix from both sides: the
    tiled GEMM over many tiles."""
    g = np.random.default_rng(N)
    A = cg(g, N, N).astype(dtype).astype(np.complex128)
    B = cg(g, N, N).astype(dtype).astype(np.complex128)
    b = gb(N)
    opq = Q.DenseOperator(b, b, A, dtype)
    Bq = Q.Operator(b, b, B, dtype)
    R = Q.Operator(b, b, Q.DeviceArray(N * N, dtype, (N, N)))
    Q.mul_(R, opq, Bq, 1.0, 0.0)
    assert rel(R.to_numpy(), A @ B) < TOLS[dtype]
    Q.mul_(R, Bq, opq, 0.5j, 0.0)
    assert rel(R.to_numpy(), 0.5j * B @ A) < TOLS[dtype]


@pytest.mark.parametrize("N", [4096, 777])
@pytest.mark.parametrize("dtype", DTS)
def test_dense_operator_large_ket_and_bra(N, dtype):
    """gemv with the column range split over workgroups (deterministic
    second-stage sum), on a Ket and a Bra."""
    g = np.random.default_rng(N + 1)
    A = cg(g, N, N).astype(dtype).astype(np.complex128)
    x = cg(g, N).astype(dtype).astype(np.complex128)
    y0 = cg(g, N).astype(dtype).astype(np.complex128)
    b = gb(N)
    opq = Q.DenseOperator(b, b, A, dtype)
    r = Q.Ket(b, y0, dtype)
    Q.mul_(r, opq, Q.Ket(b, x, dtype), 2.0, -1.0)
    assert rel(r.to_numpy(), 2.0 * A @ x - y0) < TOLS[dtype]
    r = Q.Bra(b, y0, dtype)
    Q.mul_(r, Q.Bra(b, x, dtype), opq, 1.0, 0.0)
    assert rel(r.to_numpy(), A.T @ x) < TOLS[dtype]


# dense factors of every width class at sites with n1 = 1, small n1, n1 >= 16
# and n1 >= 64 (shape_l, shape_r, indices, kinds)
WIDE_CASES = [
    ((12, 3), (12, 3), (1,), ("dense",)),                   # n1 = 1, d = 12 (chunk)
    ((3, 20, 2), (3, 20, 2), (2,), ("dense",)),            # n1 = 3, d = 20 (chunk)
    ((2, 37), (2, 53), (2,), ("dense",)),                  # non-square, n1 = 2
    ((24, 20), (24, 20), (2,), ("dense",)),                # n1 = 24 (fibre, ragged groups)
    ((80, 12, 2), (80, 12, 2), (2,), ("dense",)),          # n1 = 80 (fibre)
    ((96, 3), (96, 3), (1,), ("dense",)),                  # d = 96, n1 = 1 (GEMM)
    ((2, 128), (2, 128), (2,), ("dense",)),                # d = 128, n1 = 2 (GEMM over i1)
    ((70, 100), (70, 100), (2,), ("dense",)),              # d = 100, n1 = 70 (GEMM over i3)
    ((5, 96, 3), (5, 96, 3), (1, 2, 3), ("dense", "dense", "dense")),
    ((100,), (100,), (1,), ("dense",)),                    # one site: Ket is a gemv
    ((16, 32, 16), (16, 32, 16), (1, 2, 3), ("dense", "dense", "dense")),
]


@pytest.mark.parametrize("case", range(len(WIDE_CASES)))
@pytest.mark.parametrize("dtype", DTS)
def test_lazytensor_dense_factors_any_width(case, dtype):
    shl, shr, idx, kinds = WIDE_CASES[case]
    g = np.random.default_rng(300 + case)
    oop = rand_lt(g, shl, shr, idx, kinds)
    qop = to_q(oop)
    fused = _tune("fused_lt", 0)  # the mode-product chain, not the fused column kernel
    try:
        run_all_paths(oop, qop, g, complex(0.3, 0.8), complex(-0.5, 0.25), dtype, TOLS[dtype], ncols=5)
        run_all_paths(oop, qop, g, 1.0, 0.0, dtype, TOLS[dtype], ncols=19)
    finally:
        _tune("fused_lt", fused)


@pytest.mark.parametrize("case", [0, 1, 3, 4, 10])
@pytest.mark.parametrize("dtype", DTS)
def test_small_mode_kernels_equal_gemm_path(case, dtype):
    """The fibre and LDS-chunk kernels (mode_kernel = 0) and the fibre
    kernel also for leading dimensions below 16 (mode_kernel = 2) against
    the tiled GEMM on the same mode products (mode_kernel = 1 sends every
    dense mode product to the GEMM)."""
    shl, shr, idx, kinds = WIDE_CASES[case]
    g = np.random.default_rng(400 + case)
    oop = rand_lt(g, shl, shr, idx, kinds)
    qop = to_q(oop)
    D = int(np.prod(shr))
    B = cg(g, D, 21)
    outs = []
    fused = _tune("fused_lt", 0)
    try:
        for mk in (0, 1, 2):
            old = _tune("mode_kernel", mk)
            try:
                R = Q.Operator(qop.basis_l, gb(21), Q.DeviceArray(int(np.prod(shl)) * 21, dtype, (int(np.prod(shl)), 21)))
                Q.mul_(R, qop, Q.Operator(qop.basis_r, gb(21), B, dtype), 1.0, 0.0)
                outs.append(R.to_numpy())
            finally:
                _tune("mode_kernel", old)
    finally:
        _tune("fused_lt", fused)
    assert rel(outs[0], outs[1]) < TOLS[dtype]
    assert rel(outs[2], outs[1]) < TOLS[dtype]


@pytest.mark.parametrize("transa", [0, 1])
@pytest.mark.parametrize("dtype", DTS)
def test_qob_gemm(transa, dtype):
    g = np.random.default_rng(7 + transa)
    m, n, k = 150, 70, 90
    A = cg(g, k, m) if transa else cg(g, m, k)
    B = cg(g, k, n)
    C0 = cg(g, m, n)
    ctx = Q.context()
    dA, dB = Q.DeviceArray.from_numpy(A, dtype), Q.DeviceArray.from_numpy(B, dtype)
    dC = Q.DeviceArray.from_numpy(C0, dtype)
    _lib.call("qob_gemm", ctx.h, dA.qdt, transa, m, n, k, _lib.c64(0.5 - 1j), dA.ptr, dB.ptr, _lib.c64(2.0), dC.ptr,
              None)
    opA = A.T if transa else A
    assert rel(dC.to_numpy(), (0.5 - 1j) * opA @ B + 2.0 * C0) < TOLS[dtype]
    # k = 0: C = beta C, and beta = 0 writes zeros over NaN
    _lib.call("qob_gemm", ctx.h, dA.qdt, transa, m, n, 0, _lib.c64(1.0), dA.ptr, dB.ptr, _lib.c64(0.5), dC.ptr, None)
    want = 0.5 * ((0.5 - 1j) * opA @ B + 2.0 * C0)
    assert rel(dC.to_numpy(), want) < TOLS[dtype]
    dN = Q.DeviceArray.from_numpy(np.full((m, n), np.nan + 0j), dtype)
    _lib.call("qob_gemm", ctx.h, dA.qdt, transa, m, n, 0, _lib.c64(1.0), dA.ptr, dB.ptr, _lib.c64(0.0), dN.ptr, None)
    assert np.array_equal(dN.to_numpy(), np.zeros((m, n)))


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_mixed_element_types_like_reference(kind):
    """test/test_operators_lazytensor.jl:239-253: a LazyTensor on the reference
    test bases (2, 1, 6) <- (3, 4, 5) with a ComplexF32 factor and two
    adjoint factors, times 0.1, applied to a ComplexF32 Ket into a ComplexF64
    result (the reference holds 1e-6).  The product runs in the promoted
    type, so against the oracle on the same values it holds 1e-12; the other
    state forms, and a ComplexF64 state into a ComplexF32 result (rounded
    once: 1e-6), are checked as well."""
    g = np.random.default_rng(239)
    s1 = cg(g, 2, 3).astype(np.complex64).astype(np.complex128)
    s2 = O.dagger(O.ODense(cg(g, 4, 1)))
    s3 = O.dagger(O.ODense(cg(g, 5, 6)))
    facs = (O.ODense(s1), s2, s3)
    if kind == "sparse":
        import scipy.sparse as sp
        facs = tuple(O.OSparse.from_scipy(sp.csc_matrix(O._mat(f))) for f in facs)
    oop = O.OLazyTensor((2, 1, 6), (3, 4, 5), (1, 2, 3), facs, 0.1)
    qop = to_q(oop)
    bl, br = qop.basis_l, qop.basis_r
    x32 = cg(g, 60).astype(np.complex64)
    want = np.zeros(12, complex)
    O.mul_ket(want, oop, x32.astype(np.complex128), 1.0, 0.0)
    r = Q.Ket(bl, cg(g, 12), np.complex128)
    Q.mul_(r, qop, Q.Ket(br, x32, np.complex64), complex(1.0), complex(0.0))
    assert rel(r.to_numpy(), want) < 1e-12
    assert rel(r.to_numpy(), want) < 1e-6
    # ComplexF64 state into a ComplexF32 result, with beta != 0
    x64 = cg(g, 60)
    r0 = cg(g, 12).astype(np.complex64)
    want = r0.astype(np.complex128)
    O.mul_ket(want, oop, x64, 0.5, 2.0)
    r = Q.Ket(bl, r0, np.complex64)
    Q.mul_(r, qop, Q.Ket(br, x64, np.complex128), 0.5, 2.0)
    assert r.data.dtype == np.complex64 and rel(r.to_numpy(), want) < 1e-6
    # Bra, and a density matrix from the left and from the right
    y32 = cg(g, 12).astype(np.complex64)
    want = np.zeros(60, complex)
    O.mul_bra(want, y32.astype(np.complex128), oop, 1.0, 0.0)
    r = Q.Bra(br, np.zeros(60, complex), np.complex128)
    Q.mul_(r, Q.Bra(bl, y32, np.complex64), qop)
    assert rel(r.to_numpy(), want) < 1e-12
    B = cg(g, 60, 7).astype(np.complex64)
    want = np.zeros((12, 7), complex)
    O.mul_op_left(want, oop, B.astype(np.complex128), 1.0, 0.0)
    R = Q.Operator(bl, gb(7), np.zeros((12, 7), complex), np.complex128)
    Q.mul_(R, qop, Q.Operator(br, gb(7), B, np.complex64))
    assert rel(R.to_numpy(), want) < 1e-12
    B = cg(g, 7, 12)
    R0 = cg(g, 7, 60).astype(np.complex64)
    want = R0.astype(np.complex128)
    O.mul_op_right(want, B, oop, 1.0, -1.0)
    R = Q.Operator(gb(7), br, R0, np.complex64)
    Q.mul_(R, Q.Operator(gb(7), bl, B, np.complex128), qop, 1.0, -1.0)
    assert rel(R.to_numpy(), want) < 1e-6


# factors on sites 1, 2 (and 3): the register-chained kernel (k_ltreg.hip)
REG_CASES = [
    ((16, 16), (1, 2)),           # no site-3 factor, d2 = 16
    ((16, 32), (1, 2)),           # no site-3 factor, d2 = 32
    ((16, 32, 4), (1, 2)),        # no site-3 factor, 4 blocks per column
    ((16, 16, 16), (1, 2, 3)),
    ((16, 32, 16), (1, 2, 3)),    # the 8192-element column of the dense benchmark
    ((16, 16, 32), (1, 2, 3)),
    ((16, 32, 32), (1, 2, 3)),    # d2 = d3 = 32: falls back to the fused column kernel
]


@pytest.mark.parametrize("case", range(len(REG_CASES)))
@pytest.mark.parametrize("dtype", DTS)
def test_register_chain_kernel(case, dtype):
    """Modes 1 and 2 chained in registers, mode 3 through one LDS exchange:
    against the oracle, and against the fused column kernel (fused_reg = 0),
    with (alpha, beta) general and beta = 0 over NaN."""
    dims, idx = REG_CASES[case]
    g = np.random.default_rng(500 + case)
    facs = [cg(g, d, d) for d in dims[:len(idx)]]
    lt = Q.LazyTensor(comp(dims), comp(dims), idx, tuple(Q.DenseOperator(gb(d), gb(d), A)
                                                         for d, A in zip(dims, facs)))
    olt = O.OLazyTensor(dims, dims, idx, tuple(O.ODense(A) for A in facs), 1.0)
    D = int(np.prod(dims))
    ncol = 37
    B = cg(g, D, ncol)
    R0 = cg(g, D, ncol)
    al, be = complex(0.4, -0.9), complex(0.25, 0.5)
    want = R0.copy()
    O.mul_op_left(want, olt, B, al, be)
    outs = []
    for reg in (1, 0):
        old = _tune("fused_reg", reg)
        try:
            R = Q.Operator(lt.basis_l, gb(ncol), R0, dtype)
            Q.mul_(R, lt, Q.Operator(lt.basis_r, gb(ncol), B, dtype), al, be)
            outs.append(R.to_numpy())
            assert rel(outs[-1], want) < TOLS[dtype], reg
            Rn = Q.Operator(lt.basis_l, gb(ncol), np.full((D, ncol), np.nan + 0j), dtype)
            Q.mul_(Rn, lt, Q.Operator(lt.basis_r, gb(ncol), B, dtype), 1.0, 0.0)
            w0 = np.zeros((D, ncol), complex)
            O.mul_op_left(w0, olt, B, 1.0, 0.0)
            assert np.isfinite(Rn.to_numpy()).all() and rel(Rn.to_numpy(), w0) < TOLS[dtype]
            # a Ket is a single column
            k = Q.Ket(lt.basis_l, np.zeros(D, complex), dtype)
            Q.mul_(k, lt, Q.Ket(lt.basis_r, B[:, 0], dtype))
            assert rel(k.to_numpy(), w0[:, 0]) < TOLS[dtype]
        finally:
            _tune("fused_reg", old)
    assert rel(outs[0], outs[1]) < TOLS[dtype]
