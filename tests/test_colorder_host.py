"""Column traversal order of the right sparse product (host code of the
library, no GPU): qob_csc_column_order returns a permutation, keeps the
natural order where it is already local, and on the Fock-Fock-spin
Hamiltonian brings the distance between output columns that read the same
column of B down to the reverse Cuthill-McKee bandwidth scipy finds for the
pattern of X^T X."""
import ctypes as C
import threading

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

from qobhip import _lib
from qobhip import workloads as W


def column_order(X):
    X = sp.csc_matrix(X)
    X.sort_indices()
    n = X.shape[1]
    colptr = np.ascontiguousarray(X.indptr, dtype=np.int32)
    rowind = np.ascontiguousarray(X.indices, dtype=np.int32)
    order = np.empty(n, np.int32)
    used = C.c_int(-1)
    _lib.call("qob_csc_column_order", n, X.shape[0], colptr.ctypes.data, rowind.ctypes.data,
              order.ctypes.data, C.byref(used))
    return order, used.value


def reuse_distance(X, order):
    """Largest distance in the order between two columns sharing a row of X."""
    P = (abs(sp.csc_matrix(X)) > 0).astype(np.int32)
    G = (P.T @ P).tocoo()
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    off = G.row != G.col
    if not off.any():
        return 0
    return int(np.abs(pos[G.row[off]] - pos[G.col[off]]).max())


def test_fock_spin_order_is_a_local_permutation():
    X = W.fock_fock_spin().H
    order, used = column_order(X)
    assert used == 1
    assert np.array_equal(np.sort(order), np.arange(X.shape[1]))
    natural = reuse_distance(X, np.arange(X.shape[1]))
    ours = reuse_distance(X, order)
    P = (abs(X) > 0).astype(np.int32)
    scipy_rcm = reuse_distance(X, reverse_cuthill_mckee((P.T @ P).tocsr(), symmetric_mode=True))
    assert natural == 2 * 1561
    assert ours <= max(scipy_rcm, 2)


def test_banded_matrix_keeps_natural_order():
    n = 500
    X = sp.diags([np.ones(n - 3), np.ones(n), np.ones(n - 3)], [-3, 0, 3], format="csc")
    order, used = column_order(X)
    assert used == 0
    assert np.array_equal(order, np.arange(n))


def test_scrambled_band_is_recovered():
    """A banded matrix with its columns and rows relabelled by a random
    permutation: the order brings the reuse distance back to the band's."""
    g = np.random.default_rng(3)
    n = 2000
    Xb = sp.diags([np.ones(n - 1), np.ones(n), np.ones(n - 1)], [-1, 0, 1], format="csc")
    p = g.permutation(n)
    X = Xb[p][:, p].tocsc()
    order, used = column_order(X)
    assert used == 1
    assert np.array_equal(np.sort(order), np.arange(n))
    assert reuse_distance(X, np.arange(n)) > 1000
    assert reuse_distance(X, order) <= 4


def test_rectangular_empty_columns_and_components():
    # 3 components of different sizes, empty columns, more rows than columns
    blocks = [sp.random(40, 30, density=0.2, random_state=s, format="csc") for s in range(3)]
    X = sp.block_diag(blocks + [sp.csc_matrix((10, 5))], format="csc")
    g = np.random.default_rng(9)
    pc = g.permutation(X.shape[1])
    X = X[:, pc].tocsc()
    order, used = column_order(X)
    assert np.array_equal(np.sort(order), np.arange(X.shape[1]))
    if used:
        assert 2 * reuse_distance(X, order) <= reuse_distance(X, np.arange(X.shape[1]))


def test_rejects_bad_tables():
    """Run in a thread of its own: the library's last-error message is per
    thread, so the main thread's stays clean for the other tests."""
    def bad_calls():
        order = np.empty(2, np.int32)
        used = C.c_int()
        msgs = []
        for colptr, rowind in (([0, 2, 1], [0, 1]), ([0, 1, 2], [0, 5])):
            cp, ri = np.array(colptr, np.int32), np.array(rowind, np.int32)
            try:
                _lib.call("qob_csc_column_order", 2, 2, cp.ctypes.data, ri.ctypes.data, order.ctypes.data,
                          C.byref(used))
                msgs.append(None)
            except _lib.QobError as e:
                msgs.append(str(e))
        out.extend(msgs)

    out = []
    t = threading.Thread(target=bad_calls)
    t.start()
    t.join()
    assert len(out) == 2 and all(m is not None for m in out), out
    assert "monotone" in out[0] and "out of range" in out[1]
