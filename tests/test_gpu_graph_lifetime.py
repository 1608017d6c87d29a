"""Captured graphs and the memory they use, through the C ABI.

A HIP graph replays the device addresses its captured launches had.  The
library therefore keeps everything of its own that a live graph may read:
  * a per-stream workspace that grows after capture (the old buffer),
  * the plans a host coefficient change replaces (qob_lazysum_set_factors),
  * an operator whose last handle is released,
  * a batch that is destroyed,
until the last graph of the context is destroyed.  Each case below captures a
graph, triggers the replacement, replays the graph and checks the replay
against the result the captured calls gave eagerly (the replay reads the
kept memory, which still holds what it held at capture).  The state arrays
themselves are the caller's: the tests keep them alive.
"""
import ctypes as C
import gc

import numpy as np
import pytest

import qob_oracle as O
import qobhip as Q
from qobhip import _lib
from test_oracle import cg

pytestmark = pytest.mark.gpu


def _dense_lt_handle(ctx, g, dims):
    """A LazyTensor of dense factors on every site, as raw handles."""
    hs = []
    facs = []
    for d in dims:
        A = cg(g, d, d)
        facs.append(A)
        flat = np.asfortranarray(A).ravel(order="F").astype(np.complex128)
        h = C.c_void_p()
        _lib.call("qob_dense_create", ctx.h, d, d, flat.ctypes.data_as(C.POINTER(_lib.qob_c64)), C.byref(h))
        hs.append(h)
    sh = (C.c_int64 * len(dims))(*dims)
    idx = (C.c_int64 * len(dims))(*range(1, len(dims) + 1))
    arr = (C.c_void_p * len(dims))(*[h.value for h in hs])
    lt = C.c_void_p()
    _lib.call("qob_lazytensor_create", ctx.h, len(dims), sh, sh, len(dims), idx, arr, _lib.c64(1.0), C.byref(lt))
    for h in hs:
        _lib.call("qob_op_release", h)
    return lt, facs


def _tune(name, v):
    old = C.c_int64()
    _lib.call("qob_tune_get", name.encode(), C.byref(old))
    _lib.call("qob_tune_set", name.encode(), int(v))
    return int(old.value)


def test_workspace_growth_and_release_while_graph_alive():
    ctx = Q.context()
    s = C.c_void_p()
    _lib.call("qob_stream_create", ctx.h, C.byref(s))
    g = np.random.default_rng(1)
    dims = (3, 5, 4)
    lt, facs = _dense_lt_handle(ctx, g, dims)
    D = int(np.prod(dims))
    fused = _tune("fused_lt", 0)  # the chain: its temporaries live in the stream's workspace
    try:
        ncol = 3
        x = Q.DeviceArray.from_numpy(cg(g, D, ncol))
        y = Q.DeviceArray(D * ncol)
        print("graph lifetime: eager apply, then capture", flush=True)
        _lib.call("qob_mul_op_left", ctx.h, 0, lt, x.ptr, D, ncol, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        _lib.call("qob_stream_synchronize", ctx.h, s)
        want = y.to_numpy().reshape(D, ncol, order="F")
        ref = np.zeros((D, ncol), complex)
        olt = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(A) for A in facs), 1.0)
        O.mul_op_left(ref, olt, x.to_numpy().reshape(D, ncol, order="F"), 1.0, 0.0)
        assert O.rel_frobenius(want, ref) < 1e-12
        _lib.call("qob_scale", ctx.h, 0, D * ncol, _lib.c64(0.0), y.ptr, s)
        _lib.call("qob_graph_begin", ctx.h, s)
        _lib.call("qob_mul_op_left", ctx.h, 0, lt, x.ptr, D, ncol, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        gr = C.c_void_p()
        _lib.call("qob_graph_end", ctx.h, s, C.byref(gr))
        # grow the stream's workspace: a wider input than the captured one
        big = 400
        xb = Q.DeviceArray.from_numpy(cg(g, D, big))
        yb = Q.DeviceArray(D * big)
        _lib.call("qob_mul_op_left", ctx.h, 0, lt, xb.ptr, D, big, yb.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        # release the operator's last handle
        _lib.call("qob_op_release", lt)
        lt = None
        gc.collect()
        for rep in range(3):
            _lib.call("qob_scale", ctx.h, 0, D * ncol, _lib.c64(0.0), y.ptr, s)
            _lib.call("qob_graph_launch", ctx.h, gr, s)
            _lib.call("qob_stream_synchronize", ctx.h, s)
            assert np.array_equal(y.to_numpy().reshape(D, ncol, order="F"), want), rep
        _lib.call("qob_graph_destroy", gr)
        # the library works as before once the graph is gone
        lt2, facs2 = _dense_lt_handle(ctx, np.random.default_rng(1), dims)
        _lib.call("qob_mul_op_left", ctx.h, 0, lt2, x.ptr, D, ncol, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        _lib.call("qob_stream_synchronize", ctx.h, s)
        assert O.rel_frobenius(y.to_numpy().reshape(D, ncol, order="F"), ref) < 1e-12
        _lib.call("qob_op_release", lt2)
    finally:
        _tune("fused_lt", fused)
        _lib.call("qob_stream_destroy", ctx.h, s)


def test_host_coefficient_change_and_batch_destroy_while_graph_alive():
    """A graph captured before qob_lazysum_set_factors replays the captured
    coefficients (the replaced plans are kept for it); the next eager call
    uses the new ones.  A batch destroyed while a graph that applies it is
    alive keeps its tables until the graph goes."""
    from qobhip import workloads as W
    ctx = Q.context()
    s = ctx.stream
    spec = W.tfim_chain(9, 4)
    H = spec.build()
    D = 2 ** 9
    psi = W.normalized_ket(W.rng(2), D)
    x = Q.DeviceArray.from_numpy(psi)
    y = Q.DeviceArray(D)
    k = Q.Ket(spec.basis, y)
    Q.mul_(k, H, Q.Ket(spec.basis, x))
    ctx.synchronize()
    want_old = y.to_numpy()
    with Q.Graph(s) as gr:
        Q.mul_(k, H, Q.Ket(spec.basis, x))
    newf = [2.0 * complex(f) for f in H.factors]
    H.set_factors(newf)
    y2 = Q.DeviceArray(D)
    Q.mul_(Q.Ket(spec.basis, y2), H, Q.Ket(spec.basis, x))
    ctx.synchronize()
    assert O.rel_frobenius(y2.to_numpy(), 2.0 * want_old) < 1e-12
    for rep in range(3):
        _lib.call("qob_scale", ctx.h, 0, D, _lib.c64(0.0), y.ptr, s)
        gr.launch()
        ctx.synchronize()
        assert np.array_equal(y.to_numpy(), want_old), rep
    # batch destroyed while a graph applying it is alive
    ops = [W.tfim_chain(L, 10 + L).build() for L in (6, 9, 11)]
    ins = [Q.DeviceArray.from_numpy(W.normalized_ket(W.rng(L), 2 ** L)) for L in (6, 9, 11)]
    outs = [Q.DeviceArray(2 ** L) for L in (6, 9, 11)]
    bt = Q.Batch(ops)
    bt.bind(0, ins, outs)
    bt.apply(0, 1.0, 0.0, 0.0)
    ctx.synchronize()
    want = [o.to_numpy() for o in outs]
    gb = C.c_void_p()
    _lib.call("qob_graph_begin", ctx.h, s)
    _lib.call("qob_batch_apply", ctx.h, bt.h, 0, _lib.c64(1.0), _lib.c64(0.0), _lib.c64(0.0), s)
    _lib.call("qob_graph_end", ctx.h, s, C.byref(gb))
    _lib.call("qob_batch_destroy", bt.h)
    bt.h = None
    for o in outs:
        _lib.call("qob_scale", ctx.h, 0, o.n, _lib.c64(0.0), o.ptr, s)
    _lib.call("qob_graph_launch", ctx.h, gb, s)
    ctx.synchronize()
    for o, w in zip(outs, want):
        assert np.array_equal(o.to_numpy(), w)
    _lib.call("qob_graph_destroy", gb)
    del gr


def test_workspace_growth_under_capture_is_refused():
    """A call whose workspace would have to grow while its stream is being
    captured is refused (QOB_EUNSUPPORTED) instead of allocating inside the
    capture; the capture is then ended and discarded."""
    ctx = Q.context()
    s = C.c_void_p()
    _lib.call("qob_stream_create", ctx.h, C.byref(s))
    g = np.random.default_rng(3)
    dims = (4, 6, 5)
    lt, _ = _dense_lt_handle(ctx, g, dims)
    D = int(np.prod(dims))
    fused = _tune("fused_lt", 0)
    try:
        x = Q.DeviceArray.from_numpy(cg(g, D, 64))
        y = Q.DeviceArray(D * 64)
        lib = _lib.load()
        # the factors' device copies are made by a first call on the context's
        # stream; stream s has no workspace yet
        _lib.call("qob_mul_op_left", ctx.h, 0, lt, x.ptr, D, 64, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), None)
        ctx.synchronize()
        _lib.call("qob_graph_begin", ctx.h, s)
        st = lib.qob_mul_op_left(ctx.h, 0, lt, x.ptr, D, 64, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        gr = C.c_void_p()
        end = lib.qob_graph_end(ctx.h, s, C.byref(gr))
        if end == 0:
            _lib.call("qob_graph_destroy", gr)
        assert st == 4, _lib.load().qob_last_error()
        # outside capture the same call runs
        _lib.call("qob_mul_op_left", ctx.h, 0, lt, x.ptr, D, 64, y.ptr, D, _lib.c64(1.0), _lib.c64(0.0), s)
        _lib.call("qob_stream_synchronize", ctx.h, s)
    finally:
        _tune("fused_lt", fused)
        _lib.call("qob_op_release", lt)
        _lib.call("qob_stream_destroy", ctx.h, s)
