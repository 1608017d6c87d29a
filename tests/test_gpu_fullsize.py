"""GPU parity at the full sizes the benchmark runs, every result checked.

  batch256  : the 256 transverse-field Ising chains of 8..20 sites of the
              batch benchmark (lengths tfim_batch_lengths(256, 8, 20, seed=1)),
              x' = x + c H x captured as a 32-step HIP graph and replayed;
              every state against the C oracle stepped the same way
              (qobref_lazysum_ket_flat, bit-equal to the recursive
              restatement of src/operators_lazysum.jl:189-200 +
              src/operators_lazytensor.jl:652-685)
  dense8192 : LazyTensor of Hermitian (16, 32, 16) factors applied from the
              left and from the right to the 8192 x 8192 Ginibre trace-1
              density matrix (src/operators_lazytensor.jl:576-609), both
              element types, against the numpy restatement
  Re/Im     : ComplexF64 dense factors and states whose imaginary parts are
              1e-6 of their real parts; the real and the imaginary part of
              the result are each held to 1e-12 (the reference's BLAS forms
              every complex product from four real ones)

Tolerances: 1e-12 relative Frobenius for ComplexF64, 1e-5 for ComplexF32,
both against the FP64 oracle.
"""
import concurrent.futures as cf
import ctypes as C
import os

import numpy as np
import pytest

import qob_oracle as O
import qobhip as Q
from qobhip import _lib, workloads as W

pytestmark = pytest.mark.gpu

C_STEP = -0.01j
NSTEPS = 32


def _log(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ batch256
@pytest.fixture(scope="module")
def batch256_oracle():
    from cref import LazySumKet
    Ls = [int(L) for L in W.tfim_batch_lengths(256, 8, 20, seed=1)]
    specs = [W.tfim_chain(L, 1000 + i) for i, L in enumerate(Ls)]
    g = W.rng(5)
    psi0 = [W.normalized_ket(g, 2 ** L) for L in Ls]

    def run(i):
        H = LazySumKet(specs[i].basis.shape, specs[i].terms)
        x = psi0[i].copy()
        k = np.empty_like(x)
        for _ in range(NSTEPS):
            H.mul_flat(k, x, 1.0, 0.0)
            x += C_STEP * k
        return x

    _log(f"batch256: oracle, {NSTEPS} steps over 256 chains ({sum(2 ** L for L in Ls)} elements)")
    order = sorted(range(len(Ls)), key=lambda i: -Ls[i])
    nthr = max(1, min(16, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(nthr) as ex:
        res = dict(zip(order, ex.map(run, order)))
    return Ls, specs, psi0, [res[i] for i in range(len(Ls))]


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_batch256_graph32_every_state(batch256_oracle, dtype, tol):
    Ls, specs, psi0, want = batch256_oracle
    _log(f"batch256: {np.dtype(dtype).name}, 32-step graph")
    ops = [s.build() for s in specs]
    ctx = Q.context()
    xa = [Q.DeviceArray.from_numpy(p, dtype) for p in psi0]
    xb = [Q.DeviceArray(2 ** L, dtype) for L in Ls]
    bt = Q.Batch(ops, dtype)
    bt.bind(0, xa, xb)
    bt.bind(1, xb, xa)
    # the captured calls run once first (plans and tables exist), on states
    # that are then reset
    bt.apply(0, C_STEP, 0.0, 1.0)
    bt.apply(1, C_STEP, 0.0, 1.0)
    ctx.synchronize()
    for d, p in zip(xa, psi0):
        h = np.ascontiguousarray(p.astype(dtype))
        _lib.call("qob_memcpy_htod", ctx.h, d.ptr, h.ctypes.data_as(C.c_void_p), d.nbytes, None)
        ctx.synchronize()
    with Q.Graph(ctx.stream) as gr:
        for s in range(NSTEPS):
            bt.apply(s & 1, C_STEP, 0.0, 1.0, ctx.stream)
    gr.launch()
    ctx.synchronize()
    worst = 0.0
    for i, (d, w) in enumerate(zip(xa, want)):
        err = O.rel_frobenius(d.to_numpy().astype(np.complex128), w)
        worst = max(worst, err)
        assert err < tol, (i, Ls[i], err)
    _log(f"batch256: {np.dtype(dtype).name} worst relative error {worst:.2e}")


# ----------------------------------------------------------------- dense8192
@pytest.fixture(scope="module")
def dense8192():
    dims = (16, 32, 16)
    facs = W.dense_lazytensor_factors(dims, 1)
    lt = W.dense_lazytensor(dims, 1, facs)
    _log("dense8192: Ginibre trace-1 rho on the device")
    rho = W.ginibre_rho_device(seed=3, n=8192)
    rho_h = rho.to_numpy()
    olt = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(f) for f in facs), 1.0)
    _log("dense8192: oracle left and right")
    wl = np.zeros((8192, 8192), complex, order="F")
    O.mul_lt_op_left(wl, olt, rho_h, 1.0, 0.0)
    wr = np.zeros((8192, 8192), complex, order="F")
    O.mul_lt_op_right(wr, rho_h, olt, 1.0, 0.0)
    return lt, rho, rho_h, wl, wr


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_dense8192_ginibre_left_and_right(dense8192, dtype, tol):
    lt, rho64, rho_h, wl, wr = dense8192
    n = 8192
    _log(f"dense8192: {np.dtype(dtype).name} left and right")
    rho = rho64 if dtype == np.complex128 else Q.Operator(rho64.basis_l, rho64.basis_r, rho_h, dtype)
    out = Q.Operator(lt.basis_l, lt.basis_r, Q.DeviceArray(n * n, dtype, (n, n)))
    Q.mul_(out, lt, rho, 1, 0)
    el = O.rel_frobenius(out.to_numpy().astype(np.complex128), wl)
    assert el < tol, el
    Q.mul_(out, rho, lt, 1, 0)
    er = O.rel_frobenius(out.to_numpy().astype(np.complex128), wr)
    assert er < tol, er
    _log(f"dense8192: {np.dtype(dtype).name} left {el:.2e} right {er:.2e}")


# --------------------------------------------------------------------- Re/Im
def _small_imag(g, shape, ratio=1e-6):
    return g.standard_normal(shape) + 1j * ratio * g.standard_normal(shape)


def _componentwise(got, want):
    er = np.linalg.norm(got.real - want.real) / np.linalg.norm(want.real)
    ei = np.linalg.norm(got.imag - want.imag) / np.linalg.norm(want.imag)
    return er, ei


@pytest.mark.parametrize("side", ["left", "right"])
def test_dense_lt_c64_small_imaginary_parts_componentwise(side):
    """Factors and state with Im = 1e-6 Re: the imaginary part of the result
    is ~1e-6 of the real part, so a complex product formed with cancellation
    (Gauss's three-product form, Im = (a+b)(c+d) - ac - bd) would lose about
    six digits there while the Frobenius error of the whole stays small."""
    dims = (16, 32, 16)
    g = np.random.default_rng(77)
    facs = []
    for d in dims:
        X = _small_imag(g, (d, d))
        facs.append((X + X.conj().T) / 2)
    lt = W.dense_lazytensor(dims, factors=facs)
    n = 8192
    m = 128
    olt = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(f) for f in facs), 1.0)
    bm = Q.GenericBasis(m)
    if side == "left":
        x = np.asfortranarray(_small_imag(g, (n, m)))
        want = np.zeros((n, m), complex, order="F")
        O.mul_lt_op_left(want, olt, x, 1.0, 0.0)
        out = Q.Operator(lt.basis_l, bm, Q.DeviceArray(n * m, np.complex128, (n, m)))
        Q.mul_(out, lt, Q.Operator(lt.basis_r, bm, x), 1, 0)
    else:
        x = np.asfortranarray(_small_imag(g, (m, n)))
        want = np.zeros((m, n), complex, order="F")
        O.mul_lt_op_right(want, x, olt, 1.0, 0.0)
        out = Q.Operator(bm, lt.basis_r, Q.DeviceArray(m * n, np.complex128, (m, n)))
        Q.mul_(out, Q.Operator(bm, lt.basis_l, x), lt, 1, 0)
    got = out.to_numpy()
    assert O.rel_frobenius(got, want) < 1e-12
    er, ei = _componentwise(got, want)
    _log(f"Re/Im {side}: real part {er:.2e}, imaginary part {ei:.2e}")
    assert er < 1e-12 and ei < 1e-12, (er, ei)
