"""TimeDependentSum host semantics (src/time_dependent_operator.jl): clock,
coefficient evaluation, op(t) copies, dagger conjugates (no GPU needed)."""
import numpy as np
import scipy.sparse as sp

import qobhip as Q


def _ops():
    b = Q.SpinBasis(0.5)
    B = Q.tensor(b, b)
    sx = sp.csc_matrix(np.array([[0, 1], [1, 0]], complex))
    sz = sp.csc_matrix(np.array([[1, 0], [0, -1]], complex))
    t1 = Q.LazyTensor(B, B, (1,), (Q.SparseOperator(b, b, sx),))
    t2 = Q.LazyTensor(B, B, (2,), (Q.SparseOperator(b, b, sz),))
    return t1, t2


def test_clock_and_coefficients():
    t1, t2 = _ops()
    H = Q.TimeDependentSum([np.cos, 0.5 + 1j], (t1, t2), init_time=0.3)
    assert H.current_time == 0.3
    assert np.allclose(H.static_operator().factors, [np.cos(0.3), 0.5 + 1j])
    H.set_time_(1.1)
    assert H.current_time == 1.1 and np.allclose(H.static_op.factors, [np.cos(1.1), 0.5 + 1j])
    assert not H.is_const()
    G = H(2.0)  # op(t) = set_time!(copy(op), t)
    assert G.current_time == 2.0 and H.current_time == 1.1 and G is not H
    assert np.allclose(G.static_op.factors, [np.cos(2.0), 0.5 + 1j])
    D = H.dagger()
    assert np.allclose(D.static_op.factors, np.conj(H.static_op.factors))
    assert np.allclose(D.eval_coefficients(0.7), np.conj(H.eval_coefficients(0.7)))
    P = Q.TimeDependentSum([(lambda t: t, t1), (2.0, t2)], init_time=4.0)
    assert np.allclose(P.static_op.factors, [4.0, 2.0]) and Q.TimeDependentSum(P.static_op).is_const()
