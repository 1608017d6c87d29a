"""The C ABI on its own: tests/capi_check.c, compiled as C against
include/qob_hip.h and linked to libqob_hip.so, applies a SparseOperator to a
density matrix from both sides and a LazySum of LazyTensors to a Ket, checks
each against a dense product it computes itself, and checks that dimension
mismatch and aliasing come back as QOB_EDIM / QOB_EALIAS."""
import subprocess

import pytest

from test_abi import _build_capi

pytestmark = pytest.mark.gpu


def test_c_program_through_the_abi(tmp_path):
    exe = _build_capi(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK"), r.stdout
