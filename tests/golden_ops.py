"""Operators of the golden fixtures (tests/golden/golden.npz) built through
the library's host classes and through the oracle's classes."""
from pathlib import Path

import numpy as np
import scipy.sparse as sp

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.npz"


def load():
    return dict(np.load(GOLDEN))


def chain_terms(G):
    """(coef, sites, [scipy factors]) of sum J sx sx + sum h sz."""
    sx = sp.csc_matrix(np.array([[0, 1], [1, 0]], complex))
    sz = sp.csc_matrix(np.array([[1, 0], [0, -1]], complex))
    J, h = G["chain5_J"], G["chain5_h"]
    L = len(h)
    return L, [(complex(J[i]), (i + 1, i + 2), [sx, sx]) for i in range(L - 1)] + \
              [(complex(h[i]), (i + 1,), [sz]) for i in range(L)]


def fock_factors(G):
    a = sp.csc_matrix(np.diag(np.sqrt(np.arange(1, 4)), 1).astype(complex))
    b = sp.csc_matrix(np.diag(np.sqrt(np.arange(1, 3)), 1).astype(complex))
    s = 1.0
    ms = [s - k for k in range(3)]
    spl = sp.csc_matrix(np.diag([np.sqrt(s * (s + 1) - m * (m + 1)) for m in ms[1:]], 1).astype(complex))
    g = float(G["fock_g"][0])
    return a, b, spl, g


def fock_sparse(G):
    a, b, spl, g = fock_factors(G)

    def k3(x, y, z):  # site 1 fastest
        return sp.kron(z, sp.kron(y, x))
    Ia, Ib, Is = sp.identity(4), sp.identity(3), sp.identity(3)
    H = k3(a.T.conj() @ a, Ib, Is) + k3(Ia, b.T.conj() @ b, Is) + g * (k3(a, b.T.conj(), spl) +
                                                                    k3(a.T.conj(), b, spl.T.conj()))
    H = sp.csc_matrix(H)
    H.eliminate_zeros()
    return H


def fock_terms(G):
    a, b, spl, g = fock_factors(G)
    ad, bd, sm = [sp.csc_matrix(x.T.conj()) for x in (a, b, spl)]
    return [(1.0, (1,), [sp.csc_matrix(ad @ a)]), (1.0, (2,), [sp.csc_matrix(bd @ b)]),
            (g, (1, 2, 3), [a, bd, spl]), (g, (1, 2, 3), [ad, b, sm])]
