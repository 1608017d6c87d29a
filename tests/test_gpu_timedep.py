"""TimeDependentSum on the GPU: mul! at several times against the oracle,
device-held coefficients through the qubit and gather paths, and a captured
graph of time steps (coefficients from a device table) against eager steps."""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O
import qobhip as Q
from qobhip import _lib
from qobhip import workloads as W

pytestmark = pytest.mark.gpu

COEFS = [np.cos, lambda t: 0.5 * np.exp(1j * t), 0.25, lambda t: t * t - 1]


def _chain(L):
    spec = W.spin_chain(L, 3)  # J sx sx bonds, then h sz fields
    return spec


def _pieces(spec, coefs):
    """4 LazyTensors (two bonds, two fields) of the spec with the given
    time-dependent coefficients, and the oracle sum at time t."""
    terms = [spec.terms[0], spec.terms[1], spec.terms[-1], spec.terms[-2]]
    qops = []
    for c, idx, facs in terms:
        fops = tuple(Q.SparseOperator(spec.basis.bases[i - 1], spec.basis.bases[i - 1], f) for i, f in zip(idx, facs))
        qops.append(Q.LazyTensor(spec.basis, spec.basis, tuple(idx), fops, c))
    sh = spec.basis.shape
    D = len(spec.basis)

    def oracle_at(t):
        fs = [complex(c(t)) if callable(c) else complex(c) for c in coefs]
        return O.OLazySum(D, D, fs, [O.OLazyTensor(sh, sh, tuple(idx), tuple(O.OSparse.from_scipy(f) for f in facs), c)
                                      for c, idx, facs in terms])
    return tuple(qops), oracle_at


@pytest.mark.parametrize("qubit", [1, 0])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_mul_at_times(qubit, dtype, tol):
    spec = _chain(7)
    ops, oracle_at = _pieces(spec, COEFS)
    H = Q.TimeDependentSum(COEFS, ops, init_time=0.0)
    g = np.random.default_rng(4)
    D = len(spec.basis)
    psi = g.standard_normal(D) + 1j * g.standard_normal(D)
    rho = g.standard_normal((D, D)) + 1j * g.standard_normal((D, D))
    _lib.call("qob_tune_set", b"qubit_enable", qubit)
    try:
        for t in [0.0, 0.4, 1.3, -2.0]:
            H.set_time_(t)
            Ho = oracle_at(t)
            out = Q.Ket(spec.basis, np.zeros(D), dtype)
            Q.mul_(out, H, Q.Ket(spec.basis, psi, dtype))
            want = np.zeros(D, complex)
            O.mul_ket(want, Ho, psi, 1, 0)
            assert O.rel_frobenius(out.to_numpy(), want) < tol, t
            R = Q.Operator(spec.basis, spec.basis, np.zeros((D, D)), dtype)
            Q.mul_(R, Q.Operator(spec.basis, spec.basis, rho, dtype), H, 0.5, 0.0)
            want = np.zeros((D, D), complex)
            O.mul_op_right(want, rho, Ho, 0.5, 0)
            assert O.rel_frobenius(R.to_numpy(), want) < tol, t
    finally:
        _lib.call("qob_tune_set", b"qubit_enable", 1)


def test_plain_mode_refuses_device_coefficients_until_host_update():
    spec = _chain(5)
    ops, oracle_at = _pieces(spec, COEFS)
    H = Q.TimeDependentSum(COEFS, ops, init_time=0.2)
    D = len(spec.basis)
    psi = np.arange(D) + 1j
    out = Q.Ket(spec.basis, np.zeros(D))
    Q.mul_(out, H, Q.Ket(spec.basis, psi))  # builds the handle and plans
    H.set_time_(0.9)  # device-held coefficients from here on
    ctx = Q.context()
    ctx.set_mode(Q.QOB_MODE_PLAIN)
    try:
        with pytest.raises(Q.QobError):
            Q.mul_(out, H, Q.Ket(spec.basis, psi))
        H.static_op.set_factors(H.eval_coefficients(0.9))
        Q.mul_(out, H, Q.Ket(spec.basis, psi))
        want = np.zeros(D, complex)
        O.mul_ket(want, oracle_at(0.9), psi, 1, 0)
        assert O.rel_frobenius(out.to_numpy(), want) < 1e-12
    finally:
        ctx.set_mode(Q.QOB_MODE_FUSED)


def test_non_sparse_terms_fall_back_to_host_coefficients():
    b = Q.GenericBasis(4)
    B = Q.tensor(b, b)
    A = np.arange(16).reshape(4, 4) + 1j
    dense_lt = Q.LazyTensor(B, B, (1,), (Q.DenseOperator(b, b, A),))
    sp_lt = Q.LazyTensor(B, B, (2,), (Q.SparseOperator(b, b, sp.csc_matrix(A.T)),))
    H = Q.TimeDependentSum([np.sin, 2.0], (dense_lt, sp_lt), init_time=0.0)
    psi = np.linspace(0, 1, 16) + 0.5j
    out = Q.Ket(B, np.zeros(16))
    Q.mul_(out, H, Q.Ket(B, psi))
    with pytest.raises(Q.QobError):
        H.static_op.set_factors_device(Q.DeviceArray.from_numpy(np.array([1.0, 2.0], complex)))
    H.set_time_(0.7)
    Q.mul_(out, H, Q.Ket(B, psi))
    M = np.sin(0.7) * np.kron(np.eye(4), A) + 2.0 * np.kron(A.T, np.eye(4))
    assert O.rel_frobenius(out.to_numpy(), M @ psi) < 1e-12


@pytest.mark.parametrize("qubit", [1, 0])
def test_graph_replayed_time_steps_match_eager(qubit):
    """psi <- psi + c H(t_k) psi for k = 0..K-1, coefficients from a device
    table row per step, captured once and replayed, against eager steps with
    host-side set_time! and against the oracle."""
    import ctypes as C
    spec = _chain(9)
    ops, oracle_at = _pieces(spec, COEFS)
    D = len(spec.basis)
    K, dt = 8, 0.05
    times = [j * dt for j in range(K)]
    c = -1j * dt
    g = np.random.default_rng(8)
    psi0 = W.normalized_ket(g, D)
    want = psi0.astype(complex)
    for t in times:
        kk = np.zeros(D, complex)
        O.mul_ket(kk, oracle_at(t), want, 1, 0)
        want = want + c * kk
    _lib.call("qob_tune_set", b"qubit_enable", qubit)
    try:
        H = Q.TimeDependentSum(COEFS, ops, init_time=0.0)
        ctx = Q.context()
        st = ctx.stream
        table = H.coefficient_table(times)
        xg, kg = Q.DeviceArray.from_numpy(psi0), Q.DeviceArray(D)
        x, k = Q.Ket(spec.basis, xg), Q.Ket(spec.basis, kg)

        def step(j):
            H.set_time_row(table, j, st)
            Q.mul_(k, H, x)
            _lib.call("qob_axpy", ctx.h, Q.QOB_C64, D, _lib.c64(c), kg.ptr, xg.ptr, st)

        step(0)  # plans and coefficient maps are built outside capture
        ctx.synchronize()
        with Q.Graph(st) as gr:
            for j in range(K):
                step(j)
        # eager steps with host set_time! on an independent sum
        Hh = Q.TimeDependentSum(COEFS, ops, init_time=0.0)
        xe, ke = Q.Ket(spec.basis, psi0), Q.Ket(spec.basis, np.zeros(D))
        for t in times:
            Hh.set_time_(t)
            Q.mul_(ke, Hh, xe)
            Q.axpy_(xe, c, ke)
        assert O.rel_frobenius(xe.to_numpy(), want) < 1e-12
        for rep in range(3):
            host = np.ascontiguousarray(psi0)
            _lib.call("qob_memcpy_htod", ctx.h, xg.ptr, host.ctypes.data_as(C.c_void_p), 16 * D, None)
            gr.launch()
            ctx.synchronize()
            got = xg.to_numpy()
            assert O.rel_frobenius(got, want) < 1e-12, rep
            assert O.rel_frobenius(got, xe.to_numpy()) < 1e-14, rep
    finally:
        _lib.call("qob_tune_set", b"qubit_enable", 1)


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_density_matrix_steps_with_device_coefficients(side, dtype, tol):
    """rho <- rho + c H(t_k) rho (or rho H(t_k)) with coefficients from a
    device table row per step.  On a density matrix an all-sparse sum runs
    as its summed sparse matrix, whose tables take their values from the
    device coefficients (refreshed on the device at each set_time_row): a
    captured graph of the steps is replayed three times against eager steps
    with host set_time!, against the oracle, and against the same steps with
    the summed matrix turned off (the one-launch gather)."""
    import ctypes as C
    spec = _chain(7)
    ops, oracle_at = _pieces(spec, COEFS)
    D = len(spec.basis)
    K, dt = 6, 0.07
    times = [0.1 + j * dt for j in range(K)]
    c = -1j * dt
    g = np.random.default_rng(17)
    rho0 = g.standard_normal((D, D)) + 1j * g.standard_normal((D, D))
    want = rho0.copy()
    for t in times:
        kk = np.zeros((D, D), complex)
        if side == "left":
            O.mul_op_left(kk, oracle_at(t), want, 1, 0)
        else:
            O.mul_op_right(kk, want, oracle_at(t), 1, 0)
        want = want + c * kk
    qdt = Q.QOB_C64 if dtype == np.complex128 else Q.QOB_C32
    esz = 16 if dtype == np.complex128 else 8
    _lib.call("qob_tune_set", b"qubit_enable", 0)
    results = {}
    try:
        for mat in (1, 0):
            _lib.call("qob_tune_set", b"ls_materialize", mat)
            H = Q.TimeDependentSum(COEFS, ops, init_time=0.0)
            ctx = Q.context()
            st = ctx.stream
            table = H.coefficient_table(times)
            xg, kg = Q.DeviceArray.from_numpy(rho0, dtype), Q.DeviceArray(D * D, dtype, (D, D))
            x, k = Q.Operator(spec.basis, spec.basis, xg), Q.Operator(spec.basis, spec.basis, kg)

            def step(j):
                H.set_time_row(table, j, st)
                if side == "left":
                    Q.mul_(k, H, x)
                else:
                    Q.mul_(k, x, H)
                _lib.call("qob_axpy", ctx.h, qdt, D * D, _lib.c64(c), kg.ptr, xg.ptr, st)

            step(0)  # tables and coefficient maps are built outside capture
            ctx.synchronize()
            with Q.Graph(st) as gr:
                for j in range(K):
                    step(j)
            host = np.asfortranarray(rho0.astype(dtype))
            for rep in range(3):
                _lib.call("qob_memcpy_htod", ctx.h, xg.ptr, host.ctypes.data_as(C.c_void_p), esz * D * D, None)
                gr.launch()
                ctx.synchronize()
                got = xg.to_numpy()
                assert O.rel_frobenius(got, want) < tol, (mat, rep)
            results[mat] = got
            del gr
            if mat == 1:
                # eager steps with host set_time! on an independent sum
                Hh = Q.TimeDependentSum(COEFS, ops, init_time=0.0)
                xe = Q.Operator(spec.basis, spec.basis, rho0, dtype)
                ke = Q.Operator(spec.basis, spec.basis, np.zeros((D, D)), dtype)
                for t in times:
                    Hh.set_time_(t)
                    if side == "left":
                        Q.mul_(ke, Hh, xe)
                    else:
                        Q.mul_(ke, xe, Hh)
                    Q.axpy_(xe, c, ke)
                assert O.rel_frobenius(xe.to_numpy(), want) < tol
        assert O.rel_frobenius(results[1], results[0]) < tol
    finally:
        _lib.call("qob_tune_set", b"ls_materialize", 1)
        _lib.call("qob_tune_set", b"qubit_enable", 1)
