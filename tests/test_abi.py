"""C ABI checks that need no GPU: the library loads, exports every entry
point declared in include/qob_hip.h, the ctypes prototypes match the
header, and the host half of the fast integer division is exact."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "qob_hip.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qob_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_core_entry_points():
    names = header_functions()
    for n in ["qob_mul_ket", "qob_mul_bra", "qob_mul_op_left", "qob_mul_op_right", "qob_sparse_create",
              "qob_lazytensor_create", "qob_lazysum_create", "qob_gemm"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from qobhip import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_prototypes_cover_header():
    from qobhip import _lib
    assert sorted(_lib.exported_names()) == header_functions()


def test_version_string_without_gpu():
    from qobhip import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.qob_version()
    assert lib.qob_last_error() == b""


def test_fastdiv_exact(tmp_path):
    exe = tmp_path / "fastdiv_check"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", str(ROOT / "tests" / "fastdiv_check.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr


def _build_capi(tmp_path):
    """Compile tests/capi_check.c as C11 against include/qob_hip.h and link
    it to the built library (no Python involved)."""
    from qobhip import _lib
    libdir = Path(_lib.LIB_PATH).parent
    exe = tmp_path / "capi_check"
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
                        str(ROOT / "tests" / "capi_check.c"), "-o", str(exe), "-L", str(libdir), "-lqob_hip",
                        f"-Wl,-rpath,{libdir}", "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_header_is_c_and_a_c_program_links(tmp_path):
    assert _build_capi(tmp_path).exists()
