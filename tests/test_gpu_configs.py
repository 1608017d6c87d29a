"""GPU parity at the benchmark configurations' full sizes.

  spin chain   : L = 10 LazySum (19 pure-sparse terms) on a Ket, vs oracle
  Fock-Fock-spin: 6400 x 6400 SparseOperator from the left and the right on
                 a trace-1 Ginibre density matrix, and the same Hamiltonian
                 as a LazySum of LazyTensors, vs oracle; tr(H rho - rho H) = 0
  dense LT     : LazyTensor of Hermitian (16, 32, 16) factors on 8192^2, CF64
                 vs oracle at 1e-12 and CF32 at 1e-5
"""
import numpy as np
import pytest

import qob_oracle as O
import qobhip as Q
from qobhip import workloads as W

pytestmark = pytest.mark.gpu


def oracle_lazysum(spec):
    """Oracle operator for a LazySumSpec."""
    sh = spec.basis.shape
    ops = []
    for c, idx, facs in spec.terms:
        ops.append(O.OLazyTensor(sh, sh, tuple(idx), tuple(O.OSparse.from_scipy(f) for f in facs), 1.0))
    D = len(spec.basis)
    return O.OLazySum(D, D, [c for c, _, _ in spec.terms], ops)


@pytest.mark.parametrize("L,seed", [(10, 1), (10, 2), (13, 3)])
def test_spin_chain_ket(L, seed):
    spec = W.spin_chain(L, seed)
    H = spec.build()
    psi = W.normalized_ket(W.rng(seed + 100), len(spec.basis))
    out = Q.Ket(spec.basis, np.full(len(spec.basis), np.nan, complex))
    Q.mul_(out, H, Q.Ket(spec.basis, psi), 1, 0)
    want = np.zeros(len(spec.basis), complex)
    O.mul_ket(want, oracle_lazysum(spec), psi, 1, 0)
    assert O.rel_frobenius(out.to_numpy(), want) < 1e-12


def test_spin_chain_plain_equals_fused():
    spec = W.spin_chain(10, 4)
    H = spec.build()
    psi = W.normalized_ket(W.rng(5), 1024)
    k = Q.Ket(spec.basis, psi)
    a = Q.Ket(spec.basis, np.zeros(1024, complex))
    b = Q.Ket(spec.basis, np.zeros(1024, complex))
    Q.mul_(a, H, k)
    Q.context().set_mode(Q.QOB_MODE_PLAIN)
    try:
        Q.mul_(b, H, k)
    finally:
        Q.context().set_mode(Q.QOB_MODE_FUSED)
    assert O.rel_frobenius(a.to_numpy(), b.to_numpy()) < 1e-13


@pytest.fixture(scope="module")
def fock_spin():
    spec = W.fock_fock_spin(seed=1)
    n = spec.H.shape[0]
    rho = W.ginibre_rho_device(seed=2, n=n)
    return spec, rho


def test_fock_spin_sparse_left_right(fock_spin):
    spec, rho = fock_spin
    n = spec.H.shape[0]
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    rho_h = rho.to_numpy()
    oH = O.OSparse.from_scipy(spec.H)
    out = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    # left
    Q.mul_(out, H, rho, 1, 0)
    want = np.zeros((n, n), complex)
    O.gemm_sp_left(1.0, oH, rho_h, 0.0, want)
    left = out.to_numpy()
    assert O.rel_frobenius(left, want) < 1e-12
    # right
    out2 = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    Q.mul_(out2, rho, H, 1, 0)
    want = np.zeros((n, n), complex)
    O.gemm_sp_right(1.0, rho_h, oH, 0.0, want)
    right = out2.to_numpy()
    assert O.rel_frobenius(right, want) < 1e-12
    # tr(H rho - rho H) = 0 and (H rho)^dag = rho H for Hermitian H, rho
    assert abs(np.trace(left - right)) < 1e-12 * np.abs(left).sum()
    assert O.rel_frobenius(left.conj().T, right) < 1e-12


def test_fock_spin_sparse_left_right_c32(fock_spin):
    """The same products in ComplexF32 at full size: the left product runs
    on closed blocks of row pairs (320-row blocks, two rows per thread), the
    right product on component groups with two rows per lane.  Against the
    oracle in ComplexF64 at the ComplexF32 tolerance, and (H rho)^dag =
    rho H for the Hermitian H and rho."""
    spec, rho = fock_spin
    n = spec.H.shape[0]
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    rho_h = rho.to_numpy()
    r32 = Q.Operator(rho.basis_l, rho.basis_r, rho_h, np.complex64)
    oH = O.OSparse.from_scipy(spec.H)
    rho_q = r32.to_numpy().astype(complex)   # the rounded input the device sees
    out = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n, np.complex64, (n, n)))
    Q.mul_(out, H, r32, 1, 0)
    want = np.zeros((n, n), complex)
    O.gemm_sp_left(1.0, oH, rho_q, 0.0, want)
    left = out.to_numpy().astype(complex)
    assert O.rel_frobenius(left, want) < 1e-5
    out2 = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n, np.complex64, (n, n)))
    Q.mul_(out2, r32, H, 1, 0)
    want = np.zeros((n, n), complex)
    O.gemm_sp_right(1.0, rho_q, oH, 0.0, want)
    right = out2.to_numpy().astype(complex)
    assert O.rel_frobenius(right, want) < 1e-5
    assert O.rel_frobenius(left.conj().T, right) < 1e-5


def test_fock_spin_right_column_order_bit_exact(fock_spin):
    """The right product runs on component groups, or sweeps columns in
    reverse Cuthill-McKee order by row stripes or column-major; each output
    column is still summed in the same order, so every form equals the
    natural sweep bit for bit."""
    import ctypes as C
    from qobhip import _lib
    spec, rho = fock_spin
    n = spec.H.shape[0]
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    outs = []
    v = C.c_int64()
    _lib.call("qob_tune_get", b"csc_order", C.byref(v))
    w, gk = C.c_int64(), C.c_int64()
    _lib.call("qob_tune_get", b"csc_sweep", C.byref(w))
    _lib.call("qob_tune_get", b"csc_gkernel", C.byref(gk))
    try:
        for gkern, flag, sweep in ((1, 1, 2), (0, 1, 2), (0, 1, 1), (0, 1, 0), (0, 0, 0)):
            _lib.call("qob_tune_set", b"csc_gkernel", gkern)
            _lib.call("qob_tune_set", b"csc_order", flag)
            _lib.call("qob_tune_set", b"csc_sweep", sweep)
            out = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
            Q.mul_(out, rho, H, 0.5 - 0.25j, 0)
            outs.append(out.to_numpy())
    finally:
        _lib.call("qob_tune_set", b"csc_order", v.value)
        _lib.call("qob_tune_set", b"csc_sweep", w.value)
        _lib.call("qob_tune_set", b"csc_gkernel", gk.value)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_fock_spin_left_aligned_table_bit_exact(fock_spin):
    """Closed-block, diagonal-aligned and packed tables of the left kernel
    sum each row in the same (column) order: equal results bit for bit."""
    import ctypes as C
    from qobhip import _lib
    spec, rho = fock_spin
    n = spec.H.shape[0]
    outs = []
    v = C.c_int64()
    _lib.call("qob_tune_get", b"sell_align", C.byref(v))
    w = C.c_int64()
    _lib.call("qob_tune_get", b"sell_closed", C.byref(w))
    try:
        for closed, flag in ((1, 1), (0, 1), (0, 0)):
            _lib.call("qob_tune_set", b"sell_closed", closed)
            _lib.call("qob_tune_set", b"sell_align", flag)
            H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
            out = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
            Q.mul_(out, H, rho, 0.5 - 0.25j, 0)
            outs.append(out.to_numpy())
    finally:
        _lib.call("qob_tune_set", b"sell_align", v.value)
        _lib.call("qob_tune_set", b"sell_closed", w.value)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_fock_spin_lazy_form_equals_sparse(fock_spin):
    spec, rho = fock_spin
    n = spec.H.shape[0]
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    HL = W.fock_fock_spin_lazy(spec)
    for side in ("left", "right"):
        a = Q.Operator(spec.basis, spec.basis, Q.DeviceArray(n * n))
        b = Q.Operator(spec.basis, spec.basis, Q.DeviceArray(n * n))
        if side == "left":
            Q.mul_(a, H, rho, 1, 0)
            Q.mul_(b, HL, rho, 1, 0)
        else:
            Q.mul_(a, rho, H, 1, 0)
            Q.mul_(b, rho, HL, 1, 0)
        assert O.rel_frobenius(b.to_numpy(), a.to_numpy()) < 1e-12, side


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_dense_lazytensor_full(dtype, tol):
    dims = (16, 32, 16)
    facs = W.dense_lazytensor_factors(dims, seed=3)
    lt = W.dense_lazytensor(dims, factors=facs)
    n = int(np.prod(dims))
    g = W.rng(4)
    rho_h = W.complex_gaussian(g, (n, n)) / n  # direction-agnostic random state
    rho = Q.Operator(lt.basis_r, lt.basis_r, rho_h, dtype)
    out = Q.Operator(lt.basis_l, lt.basis_r, Q.DeviceArray(n * n, dtype))
    Q.mul_(out, lt, rho, 1, 0)
    olt = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(A) for A in facs), 1.0)
    want = np.zeros((n, n), complex)
    O.mul_lt_op_left(want, olt, rho_h, 1.0, 0.0)
    assert O.rel_frobenius(out.to_numpy(), want) < tol
