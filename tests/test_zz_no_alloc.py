"""qob_normalize_ket_noalloc through the C ABI: the no-allocation
normalize!(ket) (src/states.jl:90) must give bitwise the result of
qob_normalize_ket, which runs the same kernels on the per-stream workspace,
on two calls in a row with the second on a smaller Ket (the fixed scratch
array keeps the first call's partial sums), and must agree with numpy's
x / norm(x) to the dtype's tolerance."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 1e-5}


def _ket(rng, n, dtype):
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * 3.0
    return x.astype(dtype)


def _normalize(name, x_host):
    import qobhip as Q
    from qobhip import _lib
    ctx = Q.context()
    d = Q.DeviceArray.from_numpy(x_host)
    print(f"no-alloc test: {name} n={x_host.size} dtype={x_host.dtype}", flush=True)
    _lib.call(name, ctx.h, d.qdt, d.n, d.ptr, None)
    ctx.synchronize()
    return d.to_numpy()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_equals_workspace_variant(dtype):
    rng = np.random.default_rng(11)
    # the first size needs the full 1024 partial sums, the second only 4
    for n in (3_000_017, 777):
        x = _ket(rng, n, dtype)
        got = _normalize("qob_normalize_ket_noalloc", x)
        want = _normalize("qob_normalize_ket", x)
        assert np.array_equal(got.view(np.float64 if dtype == np.complex128 else np.float32),
                              want.view(np.float64 if dtype == np.complex128 else np.float32)), n
        ref = x.astype(np.complex128) / np.linalg.norm(x.astype(np.complex128))
        err = np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref)
        assert err < TOL[dtype], (n, err)


def test_noalloc_empty_and_bad_arguments():
    import qobhip as Q
    from qobhip import _lib
    ctx = Q.context()
    lib = _lib.load()
    assert lib.qob_normalize_ket_noalloc(ctx.h, Q.QOB_C64, 0, None, None) == 0
    assert lib.qob_normalize_ket_noalloc(ctx.h, Q.QOB_C64, -1, None, None) == 1  # QOB_EDIM
    assert lib.qob_normalize_ket_noalloc(None, Q.QOB_C64, 4, None, None) == 7    # QOB_EARG
    assert lib.qob_normalize_ket_noalloc(ctx.h, Q.QOB_C64, 4, None, None) == 7   # null state
