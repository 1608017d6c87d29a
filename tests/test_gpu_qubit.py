"""GPU parity of the qubit Pauli-sum path, batches and graph capture.

The qubit path (k_qubit.hip) serves LazySums / LazyTensors whose sites are
all qubits and whose factors are monomial 2x2 matrices.  It is checked
against the oracle (the reference's recursive algorithm restated), against
the generic gather path of the same library, and as a batch of states of
different lengths in one launch, eager and replayed from a HIP graph.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import qob_oracle as O
import qobhip as Q
from qobhip import _lib
from qobhip._lib import call
from qobhip.serve import new_stream
from qobhip import workloads as W

pytestmark = pytest.mark.gpu

PAULI = {
    "x": np.array([[0, 1], [1, 0]], complex),
    "y": np.array([[0, -1j], [1j, 0]], complex),
    "z": np.array([[1, 0], [0, -1]], complex),
    "p": np.array([[0, 1], [0, 0]], complex),   # sigma_+ (row 1 empty)
    "m": np.array([[0, 0], [1, 0]], complex),   # sigma_-
    "n": np.array([[0, 0], [0, 1]], complex),   # projector |1><1|
}


def random_pauli_sum(g, L, nterms, maxspan=3):
    """(coef, sites(1-based), [2x2 matrices]) for random monomial strings."""
    terms = []
    for _ in range(nterms):
        span = int(g.integers(1, min(maxspan, L) + 1))
        start = int(g.integers(0, L - span + 1))
        k = int(g.integers(1, span + 1))
        sites = sorted(g.choice(np.arange(start, start + span), size=k, replace=False) + 1)
        mats = [PAULI[g.choice(list(PAULI))] for _ in sites]
        coef = complex(g.standard_normal() + 1j * g.standard_normal())
        terms.append((coef, tuple(int(s) for s in sites), mats))
    return terms


def build_pair(L, terms):
    sh = (2,) * L
    b1 = Q.SpinBasis(0.5)
    B = Q.tensor(*([b1] * L))
    ops_q, ops_o, coefs = [], [], []
    for c, sites, mats in terms:
        ops_q.append(Q.LazyTensor(B, B, sites, tuple(Q.SparseOperator(b1, b1, sp.csc_matrix(m)) for m in mats)))
        ops_o.append(O.OLazyTensor(sh, sh, sites, tuple(O.OSparse.from_scipy(sp.csc_matrix(m)) for m in mats), 1.0))
        coefs.append(c)
    D = 2 ** L
    return Q.LazySum(B, B, coefs, tuple(ops_q)), O.OLazySum(D, D, coefs, ops_o), B


def cg(g, *s):
    return g.standard_normal(s) + 1j * g.standard_normal(s)


@pytest.mark.parametrize("L,nterms", [(1, 3), (3, 6), (9, 14), (13, 20), (15, 16)])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_qubit_sum_vs_oracle(L, nterms, dtype, tol):
    g = np.random.default_rng(100 + L)
    terms = random_pauli_sum(g, L, nterms)
    Hq, Ho, B = build_pair(L, terms)
    D = 2 ** L
    al, be = 0.7 - 0.2j, 0.3 + 0.4j
    # Ket
    psi, r0 = cg(g, D), cg(g, D)
    want = r0.copy()
    O.mul_ket(want, Ho, psi, al, be)
    r = Q.Ket(B, r0, dtype)
    Q.mul_(r, Hq, Q.Ket(B, psi, dtype), al, be)
    assert O.rel_frobenius(r.to_numpy(), want) < tol
    # Bra (transposed factors)
    a, r0 = cg(g, D), cg(g, D)
    want = r0.copy()
    O.mul_bra(want, a, Ho, al, be)
    r = Q.Bra(B, r0, dtype)
    Q.mul_(r, Q.Bra(B, a, dtype), Hq, al, be)
    assert O.rel_frobenius(r.to_numpy(), want) < tol
    # density matrix columns (left)
    if L <= 9:
        Bm, R0 = cg(g, D, 3), cg(g, D, 3)
        want = R0.copy()
        O.mul_op_left(want, Ho, Bm, al, be)
        cb = Q.GenericBasis(3)
        R = Q.Operator(B, cb, R0, dtype)
        Q.mul_(R, Hq, Q.Operator(B, cb, Bm, dtype), al, be)
        assert O.rel_frobenius(R.to_numpy(), want) < tol


def test_qubit_equals_generic_gather():
    g = np.random.default_rng(7)
    terms = random_pauli_sum(g, 14, 25)
    Hq, _, B = build_pair(14, terms)
    psi = cg(g, 2 ** 14)
    outs = []
    for en in (1, 0):
        _lib.call("qob_tune_set", b"qubit_enable", en)
        try:
            r = Q.Ket(B, np.zeros(2 ** 14, complex))
            Q.mul_(r, Hq, Q.Ket(B, psi))
            outs.append(r.to_numpy())
        finally:
            _lib.call("qob_tune_set", b"qubit_enable", 1)
    assert O.rel_frobenius(outs[0], outs[1]) < 1e-13


def test_qubit_zero_weight_rows_do_not_read_nan():
    # sigma_+ on site 1 never reads basis states with bit 1 set... of the input
    # index J ^ F when its weight is zero (rows absent in the reference CSC)
    L = 4
    b1 = Q.SpinBasis(0.5)
    B = Q.tensor(*([b1] * L))
    H = Q.LazySum(B, B, [1.0], (Q.LazyTensor(B, B, (1,), (Q.SparseOperator(b1, b1, sp.csc_matrix(PAULI["p"])),)),))
    psi = np.arange(16, dtype=complex)
    psi[0::2] = np.nan  # sigma_+ maps |1> -> |0>: reads only odd J
    r = Q.Ket(B, np.zeros(16, complex))
    Q.mul_(r, H, Q.Ket(B, psi))
    out = r.to_numpy()
    assert np.isfinite(out).all()
    assert np.array_equal(out[0::2], psi[1::2])


def _tfim_states(Ls, seed):
    specs = [W.tfim_chain(L, seed + i) for i, L in enumerate(Ls)]
    ops = [s.build() for s in specs]
    g = np.random.default_rng(seed)
    psis = [W.normalized_ket(g, 2 ** L) for L in Ls]
    return specs, ops, psis


def _oracle_of(spec):
    sh = spec.basis.shape
    D = len(spec.basis)
    return O.OLazySum(D, D, [c for c, _, _ in spec.terms],
                      [O.OLazyTensor(sh, sh, tuple(i), tuple(O.OSparse.from_scipy(f) for f in fs), 1.0)
                       for _, i, fs in spec.terms])


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_batch_apply_and_axpy(dtype, tol):
    Ls = [3, 8, 12, 13, 14, 16]
    specs, ops, psis = _tfim_states(Ls, 11)
    bt = Q.Batch(ops, dtype)
    tot, n = bt.info()
    assert n == len(Ls) and tot == sum(2 ** L for L in Ls)
    ins = [Q.DeviceArray.from_numpy(p, dtype) for p in psis]
    outs = [Q.DeviceArray.from_numpy(np.ones(2 ** L) * (1 + 1j), dtype) for L in Ls]
    bt.bind(0, ins, outs)
    al, be, ga = 0.5 - 0.1j, 0.25, -0.01j
    bt.apply(0, al, be, ga)
    for spec, p, o, L in zip(specs, psis, outs, Ls):
        want = np.ones(2 ** L) * (1 + 1j)
        O.mul_ket(want, _oracle_of(spec), p, al, be)
        want = want + ga * p
        assert O.rel_frobenius(o.to_numpy(), want) < tol
    # axpy: ins += a * outs (slot 1 binds outs as input, ins as output)
    before = [i.to_numpy().copy() for i in ins]
    bt.bind(1, outs, ins)
    bt.axpy(1, 0.3j)
    for b0, i, o in zip(before, ins, outs):
        assert O.rel_frobenius(i.to_numpy(), b0 + 0.3j * o.to_numpy()) < tol


def test_batch_rebind_while_busy_on_another_stream():
    """Re-binding a slot while applies that read it are queued on another
    stream: the queued applies use the old buffers, later ones the new."""
    Ls = [10, 14, 16]
    specs, ops, psis = _tfim_states(Ls, 31)
    bt = Q.Batch(ops)
    s = new_stream()
    ins = [Q.DeviceArray.from_numpy(p) for p in psis]
    outs_a = [Q.DeviceArray(2 ** L) for L in Ls]
    outs_b = [Q.DeviceArray(2 ** L) for L in Ls]
    bt.bind(0, ins, outs_a)
    for _ in range(20):
        bt.apply(0, 1.0, 0.0, 0.0, s)
    bt.bind(0, ins, outs_b)
    bt.apply(0, 2.0, 0.0, 0.0, s)
    call("qob_stream_synchronize", Q.context().h, s)
    call("qob_stream_destroy", Q.context().h, s)
    for spec, p, oa, ob, L in zip(specs, psis, outs_a, outs_b, Ls):
        want = np.zeros(2 ** L, complex)
        O.mul_ket(want, _oracle_of(spec), p, 1.0, 0.0)
        assert O.rel_frobenius(oa.to_numpy(), want) < 1e-12
        assert O.rel_frobenius(ob.to_numpy(), 2 * want) < 1e-12


def test_batch_bind_checks_buffers():
    Ls = [4, 6]
    specs, ops, psis = _tfim_states(Ls, 3)
    bt = Q.Batch(ops)
    ins = [Q.DeviceArray.from_numpy(p) for p in psis]
    outs = [Q.DeviceArray(2 ** L) for L in Ls]
    with pytest.raises(ValueError):
        bt.bind(0, ins[:1], outs[:1])
    with pytest.raises(ValueError):
        bt.bind(0, ins, [outs[0], Q.DeviceArray(8)])
    with pytest.raises(TypeError):
        bt.bind(0, ins, [outs[0], Q.DeviceArray(2 ** 6, np.complex64)])
    bt.bind(0, ins, outs)


def test_batch_rejects_non_qubit_operator():
    sp1 = Q.SpinBasis(1)
    B = Q.tensor(sp1, sp1)
    op = Q.LazyTensor(B, B, (1,), (Q.sigmax(sp1),))
    with pytest.raises(Q.MethodError):
        Q.Batch([op])


def test_batch_graph_replays_match_eager():
    """32-step-style sequence (apply into k, axpy into psi) captured once and
    replayed several times, each replay checked against eager execution."""
    Ls = [5, 9, 12, 14]
    specs, ops, psis = _tfim_states(Ls, 21)
    c = -0.01j
    nsteps = 4
    bt = Q.Batch(ops)
    psi_e = [Q.DeviceArray.from_numpy(p) for p in psis]
    k_e = [Q.DeviceArray(2 ** L) for L in Ls]
    psi_g = [Q.DeviceArray.from_numpy(p) for p in psis]
    k_g = [Q.DeviceArray(2 ** L) for L in Ls]
    bt.bind(0, psi_e, k_e)
    bt.bind(1, k_e, psi_e)
    bt.bind(2, psi_g, k_g)
    bt.bind(3, k_g, psi_g)
    # warm-up of the captured calls on the graph buffers, then reset them
    bt.apply(2, 1.0, 0.0, 0.0)
    bt.axpy(3, 0.0)
    Q.synchronize()
    ctx = Q.context()
    with Q.Graph(ctx.stream) as gr:
        for _ in range(nsteps):
            bt.apply(2, 1.0, 0.0, 0.0, ctx.stream)
            bt.axpy(3, c, ctx.stream)
    for rep in range(3):
        for _ in range(nsteps):
            bt.apply(0, 1.0, 0.0, 0.0)
            bt.axpy(1, c)
        Q.synchronize()
        gr.launch()
        Q.synchronize()
        for a, b in zip(psi_e, psi_g):
            assert O.rel_frobenius(b.to_numpy(), a.to_numpy()) < 1e-14, rep
    # and the eager result matches the oracle sequence
    for spec, p0, a in zip(specs, psis, psi_e):
        if len(p0) > 2 ** 9:
            continue  # larger states: equal to eager above, eager pinned in test_batch_apply_and_axpy
        H = _oracle_of(spec)
        x = p0.copy()
        for _ in range(3 * nsteps):
            k = np.zeros_like(x)
            O.mul_ket(k, H, x, 1, 0)
            x = x + c * k
        assert O.rel_frobenius(a.to_numpy(), x) < 1e-12


def _tune_get(name):
    import ctypes as C
    v = C.c_int64()
    _lib.call("qob_tune_get", name.encode(), C.byref(v))
    return v.value


@pytest.mark.parametrize("split,tb,nth", [(1, 10, 256), (1, 12, 512), (0, 10, 1024), (0, 13, 512)])
def test_qubit_tile_configs_vs_oracle(split, tb, nth):
    """Every tile size / thread count, with the high-bit flips either in a
    second strided pass (split=1) or read from partner tiles (split=0)."""
    saved = {k: _tune_get(k) for k in ("qubit_split", "qubit_tbits", "qubit_nth")}
    try:
        _lib.call("qob_tune_set", b"qubit_split", split)
        _lib.call("qob_tune_set", b"qubit_tbits", tb)
        _lib.call("qob_tune_set", b"qubit_nth", nth)
        g = np.random.default_rng(300 + tb)
        L = 15
        Hq, Ho, B = build_pair(L, random_pauli_sum(g, L, 24, maxspan=4))
        D = 2 ** L
        al, be = 0.4 + 0.3j, -0.2j
        psi, r0 = cg(g, D), cg(g, D)
        want = r0.copy()
        O.mul_ket(want, Ho, psi, al, be)
        r = Q.Ket(B, r0)
        Q.mul_(r, Hq, Q.Ket(B, psi), al, be)
        assert O.rel_frobenius(r.to_numpy(), want) < 1e-12
        # batch of chains straddling the tile size
        Ls = [tb - 2, tb + 1, 2 * tb - 1 if 2 * tb - 1 <= 16 else 16]
        specs, ops, psis = _tfim_states(Ls, 40 + tb)
        bt = Q.Batch(ops)
        ins = [Q.DeviceArray.from_numpy(p) for p in psis]
        outs = [Q.DeviceArray.from_numpy(np.full(2 ** L_, 0.5 - 0.5j)) for L_ in Ls]
        bt.bind(0, ins, outs)
        bt.apply(0, al, be, 0.1)
        for spec, p, o, L_ in zip(specs, psis, outs, Ls):
            want = np.full(2 ** L_, 0.5 - 0.5j)
            O.mul_ket(want, _oracle_of(spec), p, al, be)
            want = want + 0.1 * p
            assert O.rel_frobenius(o.to_numpy(), want) < 1e-12
    finally:
        for k, v in saved.items():
            _lib.call("qob_tune_set", k.encode(), v)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_batch_schedules_bitwise_equal(dtype, tol):
    """Tile walk order (index / Gray code / Gray code above 64 tiles), one or
    two LDS tile buffers, tile sizes 2^11 / 2^12 and persistent grids small
    enough that every workgroup walks several items: the same bits, and the
    oracle's values."""
    names = ("qubit_order", "qubit_dbuf", "qubit_tbits", "qubit_grid")
    saved = {k: _tune_get(k) for k in names}
    Ls = [9, 11, 12, 13, 15, 17, 17]
    specs, ops, psis = _tfim_states(Ls, 77)
    al, be, ga = 0.3 - 0.2j, 0.5, -0.05j
    r0 = [np.full(2 ** L, 0.25 + 0.5j) for L in Ls]
    got = []
    try:
        for order, dbuf, tb, grid in [(0, 0, 11, 16), (1, 1, 12, 16), (2, 1, 11, 24), (1, 0, 12, 0), (1, 1, 0, 8)]:
            for k, v in zip(names, (order, dbuf, tb, grid)):
                _lib.call("qob_tune_set", k.encode(), v)
            bt = Q.Batch(ops, dtype)
            ins = [Q.DeviceArray.from_numpy(p, dtype) for p in psis]
            outs = [Q.DeviceArray.from_numpy(r, dtype) for r in r0]
            bt.bind(0, ins, outs)
            bt.apply(0, al, be, ga)
            got.append([o.to_numpy() for o in outs])
            del bt
    finally:
        for k, v in saved.items():
            _lib.call("qob_tune_set", k.encode(), v)
    for other in got[1:]:
        for a, b in zip(got[0], other):
            assert np.array_equal(a, b)
    for spec, p, r, o in zip(specs, psis, r0, got[0]):
        want = r.copy()
        O.mul_ket(want, _oracle_of(spec), p, al, be)
        want = want + ga * p
        assert O.rel_frobenius(o, want) < tol


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 1e-5)])
def test_qubit_columns_persistent_walks(dtype, tol):
    """A qubit sum on the 16 columns of a density-matrix slice with grids
    small enough that each workgroup walks several (column, tile) items, with
    one and two LDS tile buffers and with the XCD ranges: against the oracle,
    and the same bits every time."""
    names = ("qubit_grid", "qubit_dbuf", "qubit_remap")
    saved = {k: _tune_get(k) for k in names}
    g = np.random.default_rng(913)
    L, ncol = 13, 16
    Hq, Ho, B = build_pair(L, random_pauli_sum(g, L, 18, maxspan=13))
    D = 2 ** L
    Bm, R0 = cg(g, D, ncol), cg(g, D, ncol)
    al, be = 0.6 + 0.1j, -0.3j
    want = R0.copy()
    O.mul_op_left(want, Ho, Bm, al, be)
    cb = Q.GenericBasis(ncol)
    got = []
    try:
        for grid, dbuf, remap in [(0, 1, 2), (8, 1, 2), (8, 0, 2), (3, 1, 2), (5, 1, 1)]:
            for k, v in zip(names, (grid, dbuf, remap)):
                _lib.call("qob_tune_set", k.encode(), v)
            R = Q.Operator(B, cb, R0, dtype)
            Q.mul_(R, Hq, Q.Operator(B, cb, Bm, dtype), al, be)
            got.append(R.to_numpy())
    finally:
        for k, v in saved.items():
            _lib.call("qob_tune_set", k.encode(), v)
    assert O.rel_frobenius(got[0], want) < tol
    for other in got[1:]:
        assert np.array_equal(got[0], other)


@pytest.mark.parametrize("L,nterms", [(1, 3), (3, 6), (9, 14), (12, 18), (15, 16)])
def test_complexf32_pair_lanes_bitwise(L, nterms):
    """ComplexF32 with lanes holding pairs of adjacent elements (16-byte
    accesses) against single elements: the same bits on a Ket, a Bra and
    density-matrix columns (random Pauli sums with diagonal windows, flips of
    bit 0, register bits and partner tiles), and against the oracle."""
    dtype, tol = np.complex64, 1e-5
    g = np.random.default_rng(500 + L)
    Hq, Ho, B = build_pair(L, random_pauli_sum(g, L, nterms, maxspan=min(L, 12)))
    D = 2 ** L
    al, be, ncol = 0.7 - 0.2j, 0.3 + 0.4j, 3
    psi, r0 = cg(g, D), cg(g, D)
    Bm, R0 = cg(g, D, ncol), cg(g, D, ncol)
    cb = Q.GenericBasis(ncol)
    saved = _tune_get("qubit_pair")
    got = {}
    try:
        for pair in (0, 1):
            _lib.call("qob_tune_set", b"qubit_pair", pair)
            r = Q.Ket(B, r0, dtype)
            Q.mul_(r, Hq, Q.Ket(B, psi, dtype), al, be)
            rb = Q.Bra(B, r0, dtype)
            Q.mul_(rb, Q.Bra(B, psi, dtype), Hq, al, be)
            R = Q.Operator(B, cb, R0, dtype)
            Q.mul_(R, Hq, Q.Operator(B, cb, Bm, dtype), al, be)
            got[pair] = (r.to_numpy(), rb.to_numpy(), R.to_numpy())
    finally:
        _lib.call("qob_tune_set", b"qubit_pair", saved)
    for a, b in zip(got[0], got[1]):
        assert np.array_equal(a, b)
    want = r0.copy()
    O.mul_ket(want, Ho, psi, al, be)
    assert O.rel_frobenius(got[1][0], want) < tol
    want = R0.copy()
    O.mul_op_left(want, Ho, Bm, al, be)
    assert O.rel_frobenius(got[1][2], want) < tol


def test_complexf32_pair_lanes_batch_bitwise():
    """The 256-chain-like batch path (partner tiles, per-XCD walks, Gray
    order, fused gamma) with pair lanes against single elements."""
    Ls = [9, 11, 12, 13, 15, 17, 17]
    specs, ops, psis = _tfim_states(Ls, 78)
    al, be, ga = 0.3 - 0.2j, 0.5, -0.05j
    r0 = [np.full(2 ** L, 0.25 + 0.5j) for L in Ls]
    saved = {k: _tune_get(k) for k in ("qubit_pair", "qubit_grid")}
    got = []
    try:
        for pair, grid in [(0, 0), (1, 0), (1, 16)]:
            _lib.call("qob_tune_set", b"qubit_pair", pair)
            _lib.call("qob_tune_set", b"qubit_grid", grid)
            bt = Q.Batch(ops, np.complex64)
            ins = [Q.DeviceArray.from_numpy(p, np.complex64) for p in psis]
            outs = [Q.DeviceArray.from_numpy(r, np.complex64) for r in r0]
            bt.bind(0, ins, outs)
            bt.apply(0, al, be, ga)
            got.append([o.to_numpy() for o in outs])
            del bt
    finally:
        for k, v in saved.items():
            _lib.call("qob_tune_set", k.encode(), v)
    for other in got[1:]:
        for a, b in zip(got[0], other):
            assert np.array_equal(a, b)
