"""Serving loop on the GPU: every job's result is the same alone and in the
mix, equals the oracle's step sequence, and admission respects the HBM budget."""
import copy
import os

import numpy as np
import pytest

import qob_oracle as O
from qobhip import serve as S
from qobhip import workloads as W

pytestmark = pytest.mark.gpu

STEPS, CHUNK, C_STEP = 8, 4, -0.01j


def _oracle_ops(classes):
    out = {}
    for cl in classes:
        if cl.kind == "ket":
            out[cl.name] = {}
            for L, op in cl.ops.items():
                sh = (2,) * L
                spec = W.tfim_chain(L, 100 + L) if cl.name == "tfim" else W.spin_chain(L, 1)
                out[cl.name][L] = O.OLazySum(2 ** L, 2 ** L, [c for c, _, _ in spec.terms],
                                             [O.OLazyTensor(sh, sh, tuple(i), tuple(O.OSparse.from_scipy(f) for f in fs),
                                                            1.0) for _, i, fs in spec.terms])
        elif cl.name == "fock_rho":
            out[cl.name] = O.OSparse.from_scipy(W.fock_fock_spin(5, 1.5, 2).H)
        else:
            facs = W.dense_lazytensor_factors((16, 16), 3)
            out[cl.name] = O.OLazyTensor((16, 16), (16, 16), (1, 2), tuple(O.ODense(f) for f in facs), 1.0)
    return out


def _oracle_result(cl, oop, req):
    if cl.kind == "ket":
        x = req.payload.astype(np.complex128)
        for _ in range(STEPS):
            k = np.zeros_like(x)
            O.mul_ket(k, oop[req.L], x, 1, 0)
            x = x + C_STEP * k
        return x / np.linalg.norm(x)
    x = cl.template.to_numpy().astype(np.complex128).reshape(cl.n, cl.n, order="F")
    x = x * (1.0 + 0.5 * np.random.Generator(np.random.PCG64(req.seed)).uniform())
    for _ in range(STEPS):
        k = np.zeros_like(x)
        O.mul_op_left(k, oop, x, 1, 0)
        x = x + C_STEP * k
    return x / np.trace(x)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_alone_equals_in_mix_and_oracle(dtype, tol):
    classes = S.make_classes(dtype, small=True)
    reqs = S.make_requests(classes, 16, rate=0, seed=3, dtype=dtype)
    srv = S.Server(classes, dtype, steps=STEPS, chunk=CHUNK, c=C_STEP, keep_states=True, rho_slots=3)
    mix = srv.run(copy.deepcopy(reqs))
    assert all(r.status == "done" for r in mix)
    oops = _oracle_ops(classes)
    byname = {c.name: c for c in classes}
    for r, m in zip(reqs, mix):
        alone = srv.run([copy.deepcopy(r)])[0]
        assert alone.status == "done"
        assert O.rel_frobenius(m.state, alone.state) <= tol, (r.rid, r.cls)
        want = _oracle_result(byname[r.cls], oops[r.cls], r)
        assert O.rel_frobenius(m.state.astype(np.complex128).reshape(want.shape), want) < 10 * tol, (r.rid, r.cls)


def test_budget_admission_and_rejection():
    dtype = np.complex128
    classes = S.make_classes(dtype, small=True)
    reqs = S.make_requests(classes, 24, rate=2000, seed=9, dtype=dtype)
    item = 16
    big = max(2 * (2 ** r.L if r.L else {c.name: c for c in classes}[r.cls].n ** 2) * item for r in reqs)
    budget = 2 * big
    srv = S.Server(classes, dtype, budget_bytes=budget, steps=STEPS, chunk=CHUNK, rho_slots=4)
    res = srv.run(reqs)
    assert all(r.status == "done" for r in res)
    assert 0 < srv.peak <= budget
    # a job larger than the budget is rejected, the others still complete
    srv2 = S.Server(classes, dtype, budget_bytes=big - 1, steps=STEPS, chunk=CHUNK)
    res2 = srv2.run(S.make_requests(classes, 12, rate=0, seed=9, dtype=dtype))
    sizes = [srv2._footprint(r) for r in S.make_requests(classes, 12, rate=0, seed=9, dtype=dtype)]
    for r, sz in zip(res2, sizes):
        assert r.status == ("rejected" if sz > big - 1 else "done")
    s = S.summarize(res)
    assert s["done"] == 24 and s["p99_ms"] >= s["p50_ms"] > 0


def _full_oracle(cl, req, steps, c):
    """The job's final state by the oracle at the full shapes: Ket jobs through
    the C LazySum engine (bit-equal to the recursive restatement), the Fock
    density matrix through the C CSC loop, the dense LazyTensor through the
    numpy restatement of the mode chain."""
    import ctypes as Cc
    import cref
    if cl.kind == "ket":
        spec = W.tfim_chain(req.L, 100 + req.L) if cl.name.startswith("tfim") else W.spin_chain(req.L, 1)
        H = cref.LazySumKet(spec.basis.shape, spec.terms)
        x = req.payload.astype(np.complex128)
        k = np.empty_like(x)
        for _ in range(steps):
            H.mul_flat(k, x)
            x = x + c * k
        return x / np.linalg.norm(x)
    x = np.asfortranarray(cl.template.to_numpy().astype(np.complex128).reshape(cl.n, cl.n, order="F"))
    x = x * (1.0 + 0.5 * np.random.Generator(np.random.PCG64(req.seed)).uniform())
    k = np.zeros_like(x)
    if cl.name == "fock6400":
        M = W.fock_fock_spin(39, 1.5, 1).H.tocsc()
        M.sort_indices()
        cp = (M.indptr + 1).astype(np.int64)
        rv = (M.indices + 1).astype(np.int64)
        nz = M.data.astype(np.complex128)
        p = lambda a: a.ctypes.data_as(Cc.c_void_p)
        for _ in range(steps):
            cref.lib().qobref_gemm_sp_dense_omp(1.0, 0.0, cl.n, cl.n, p(cp), p(rv), p(nz), p(x), cl.n, 0.0, 0.0,
                                                 p(k), min(16, os.cpu_count() or 1))
            x = np.asfortranarray(x + c * k)
    else:
        facs = W.dense_lazytensor_factors((16, 32, 16), 1)
        olt = O.OLazyTensor((16, 32, 16), (16, 32, 16), (1, 2, 3), tuple(O.ODense(f) for f in facs), 1.0)
        for _ in range(steps):
            O.mul_lt_op_left(k, olt, x, 1.0, 0.0)
            x = np.asfortranarray(x + c * k)
    return x / np.trace(x)


def _full_requests(classes, dtype):
    """One job of every full shape: the 10-site chain, TFIM chains of 8 and
    20 sites, the 6400 x 6400 Fock density matrix, the 8192 x 8192 dense
    LazyTensor density matrix; plus a 14-site TFIM chain."""
    reqs = []
    for rid, (name, L) in enumerate((("chain10", 10), ("tfim8_20", 8), ("tfim8_20", 20), ("fock6400", 0),
                                     ("dense8192", 0), ("tfim8_20", 14))):
        r = S.Request(rid, 0.0, name, 7000 + rid, L)
        if L:
            gg = np.random.Generator(np.random.PCG64(r.seed))
            v = gg.standard_normal(2 ** L) + 1j * gg.standard_normal(2 ** L)
            r.payload = (v / np.linalg.norm(v)).astype(dtype)
        reqs.append(r)
    return reqs


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_full_shapes_every_final_state(dtype, tol):
    """Every full job shape admitted together; each final state (the
    normalised Ket, or the whole trace-normalised density matrix) equals the
    same job run alone and the oracle's step sequence."""
    classes = S.make_classes(dtype, small=False)
    byname = {c.name: c for c in classes}
    reqs = _full_requests(classes, dtype)
    srv = S.Server(classes, dtype, steps=STEPS, chunk=CHUNK, c=C_STEP, keep_states=True, rho_slots=2)
    print(f"serve full shapes {np.dtype(dtype).name}: {len(reqs)} jobs in one run", flush=True)
    mix = srv.run(copy.deepcopy(reqs))
    assert all(r.status == "done" for r in mix)
    for r, m in zip(reqs, mix):
        cl = byname[r.cls]
        alone = srv.run([copy.deepcopy(r)])[0]
        assert alone.status == "done"
        assert O.rel_frobenius(m.state, alone.state) <= tol, (r.rid, r.cls)
        want = _full_oracle(cl, r, STEPS, C_STEP)
        got = m.state.astype(np.complex128).reshape(want.shape, order="F")
        err = O.rel_frobenius(got, want)
        print(f"serve full shapes {np.dtype(dtype).name}: job {r.rid} {r.cls} L={r.L} error {err:.2e}", flush=True)
        assert err < 10 * tol, (r.rid, r.cls, err)
