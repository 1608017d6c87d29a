"""The autotuner on the GPU: a coordinate search over qubit launch shapes on
a small batch; the chosen shapes keep oracle parity."""
import numpy as np
import pytest

import qob_oracle as O
import qobhip as Q
from qobhip import tune as T
from qobhip import workloads as W

pytestmark = pytest.mark.gpu


def test_tune_qubit_shapes_then_parity():
    Ls = [6, 9, 12, 13]
    specs = [W.tfim_chain(L, 70 + L) for L in Ls]
    ops = [s.build() for s in specs]
    g = np.random.default_rng(5)
    psis = [W.normalized_ket(g, 2 ** L) for L in Ls]
    xa = [Q.DeviceArray.from_numpy(p) for p in psis]
    xb = [Q.DeviceArray(2 ** L) for L in Ls]
    space = {"qubit_tbits": [10, 11], "qubit_nth": [256, 512], "qubit_split": [0, 1]}
    saved = T.snapshot(space)
    try:
        def make_run():
            bt = Q.Batch(ops)
            bt.bind(0, xa, xb)
            return lambda: bt.apply(0, 1.0, 0.0, 0.0)

        r = T.tune("qubit", make_run, reps=3, warmup=1, space=space)
        assert set(r["best"]) == set(space) and r["best_ms"] <= r["start_ms"]
        assert T.snapshot(space) == r["best"]
        assert 2 <= len(r["trials"]) <= 1 + 2 * sum(len(v) - 1 for v in space.values())
        bt = Q.Batch(ops)
        bt.bind(0, xa, xb)
        bt.apply(0, 0.5, 0.0, 0.0)
        for s, p, o in zip(specs, psis, xb):
            D = len(p)
            sh = s.basis.shape
            H = O.OLazySum(D, D, [c for c, _, _ in s.terms],
                           [O.OLazyTensor(sh, sh, tuple(i), tuple(O.OSparse.from_scipy(f) for f in fs), 1.0)
                            for _, i, fs in s.terms])
            want = np.zeros(D, complex)
            O.mul_ket(want, H, p, 0.5, 0)
            assert O.rel_frobenius(o.to_numpy(), want) < 1e-12
    finally:
        T.apply(saved)
