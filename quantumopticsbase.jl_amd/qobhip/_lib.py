"""ctypes binding of libqob_hip.so (declarations in include/qob_hip.h).

The library is the only compute path: if it cannot be loaded, every call
raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "libqob_hip.so"

QOB_C64, QOB_C32 = 0, 1
QOB_MODE_FUSED, QOB_MODE_PLAIN = 0, 1


class qob_c64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def c64(z) -> qob_c64:
    z = complex(z)
    return qob_c64(z.real, z.imag)


# --------------------------------------------------------------- exceptions
class QobError(RuntimeError):
    """Base class of errors reported by libqob_hip."""


class DimensionMismatch(QobError, ValueError):
    """Julia's DimensionMismatch."""


class ArgumentError(QobError, ValueError):
    """Julia's ArgumentError (aliasing, bad arguments)."""


class IncompatibleBases(QobError):
    """QuantumInterface.IncompatibleBases."""


class MethodError(QobError, TypeError):
    """Julia's MethodError: no method for this operand combination."""


class OutOfMemoryError(QobError, MemoryError):
    pass


class HIPError(QobError):
    pass


_STATUS = {
    1: DimensionMismatch,
    2: ArgumentError,
    3: IncompatibleBases,
    4: MethodError,
    5: OutOfMemoryError,
    6: HIPError,
    7: ArgumentError,
}

# name -> (restype, argtypes)
_vp, _i64, _i32, _sz, _dbl = C.c_void_p, C.c_int64, C.c_int, C.c_size_t, C.c_double
_P = C.POINTER
_PROTOS = {
    "qob_ctx_create": (_i32, [_i32, _P(_vp)]),
    "qob_ctx_destroy": (_i32, [_vp]),
    "qob_ctx_set_mode": (_i32, [_vp, _i32]),
    "qob_ctx_get_stream": (_i32, [_vp, _P(_vp)]),
    "qob_last_error": (C.c_char_p, []),
    "qob_version": (C.c_char_p, []),
    "qob_dmalloc": (_i32, [_vp, _sz, _P(_vp)]),
    "qob_dfree": (_i32, [_vp, _vp]),
    "qob_memcpy_htod": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "qob_memcpy_dtoh": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "qob_memcpy_dtod": (_i32, [_vp, _vp, _vp, _sz, _vp]),
    "qob_stream_create": (_i32, [_vp, _P(_vp)]),
    "qob_stream_destroy": (_i32, [_vp, _vp]),
    "qob_stream_synchronize": (_i32, [_vp, _vp]),
    "qob_event_create": (_i32, [_vp, _P(_vp)]),
    "qob_event_record": (_i32, [_vp, _vp, _vp]),
    "qob_event_query": (_i32, [_vp, _P(_i32)]),
    "qob_event_synchronize": (_i32, [_vp]),
    "qob_event_elapsed": (_i32, [_vp, _vp, _P(C.c_float)]),
    "qob_event_destroy": (_i32, [_vp]),
    "qob_sparse_create": (_i32, [_vp, _i64, _i64, _P(_i64), _P(_i64), _P(qob_c64), _P(_vp)]),
    "qob_dense_create": (_i32, [_vp, _i64, _i64, _P(qob_c64), _P(_vp)]),
    "qob_eye_create": (_i32, [_vp, _i64, _i64, _P(_vp)]),
    "qob_op_adjoint": (_i32, [_vp, _vp, _P(_vp)]),
    "qob_lazytensor_create": (_i32, [_vp, _i32, _P(_i64), _P(_i64), _i32, _P(_i64), _P(_vp), qob_c64, _P(_vp)]),
    "qob_lazysum_create": (_i32, [_vp, _i64, _i64, _i32, _P(qob_c64), _P(_vp), _P(_vp)]),
    "qob_lazyproduct_create": (_i32, [_vp, _i32, _P(_vp), qob_c64, _P(_vp)]),
    "qob_lazysum_set_factors": (_i32, [_vp, _vp, _P(qob_c64)]),
    "qob_lazysum_set_factors_d": (_i32, [_vp, _vp, _vp, _vp]),
    "qob_op_shape": (_i32, [_vp, _P(_i64), _P(_i64)]),
    "qob_op_release": (_i32, [_vp]),
    "qob_mul_ket": (_i32, [_vp, _i32, _vp, _vp, _i64, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_bra": (_i32, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_op_left": (_i32, [_vp, _i32, _vp, _vp, _i64, _i64, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_op_right": (_i32, [_vp, _i32, _vp, _i64, _i64, _vp, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_gemm": (_i32, [_vp, _i32, _i32, _i64, _i64, _i64, qob_c64, _vp, _vp, qob_c64, _vp, _vp]),
    "qob_mul_ket_x": (_i32, [_vp, _i32, _vp, _vp, _i64, _i32, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_bra_x": (_i32, [_vp, _i32, _vp, _i64, _vp, _i32, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_op_left_x": (_i32, [_vp, _i32, _vp, _vp, _i64, _i64, _i32, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_mul_op_right_x": (_i32, [_vp, _i32, _vp, _i64, _i64, _vp, _i32, _vp, _i64, qob_c64, qob_c64, _vp]),
    "qob_scale": (_i32, [_vp, _i32, _i64, qob_c64, _vp, _vp]),
    "qob_axpy": (_i32, [_vp, _i32, _i64, qob_c64, _vp, _vp, _vp]),
    "qob_copy": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp]),
    "qob_tune_set": (_i32, [C.c_char_p, _i64]),
    "qob_tune_get": (_i32, [C.c_char_p, _P(_i64)]),
    "qob_csc_column_order": (_i32, [_i64, _i64, _vp, _vp, _vp, _P(C.c_int)]),
    "qob_convert": (_i32, [_vp, _i32, _vp, _i32, _vp, _i64, _vp]),
    "qob_dot": (_i32, [_vp, _i32, _i64, _vp, _vp, _P(qob_c64), _vp]),
    "qob_norm": (_i32, [_vp, _i32, _i64, _vp, _P(_dbl), _vp]),
    "qob_trace": (_i32, [_vp, _i32, _i64, _vp, _P(qob_c64), _vp]),
    "qob_normalize_ket": (_i32, [_vp, _i32, _i64, _vp, _vp]),
    "qob_normalize_op": (_i32, [_vp, _i32, _i64, _vp, _vp]),
    "qob_normalize_ket_noalloc": (_i32, [_vp, _i32, _i64, _vp, _vp]),
    "qob_expect_ket": (_i32, [_vp, _i32, _vp, _vp, _i64, _P(qob_c64), _vp]),
    "qob_expect_op": (_i32, [_vp, _i32, _vp, _vp, _i64, _P(qob_c64), _vp]),
    "qob_variance_ket": (_i32, [_vp, _i32, _vp, _vp, _i64, _P(qob_c64), _vp]),
    "qob_variance_op": (_i32, [_vp, _i32, _vp, _vp, _i64, _P(qob_c64), _vp]),
    "qob_batch_create": (_i32, [_vp, _i32, _P(_vp), _i32, _P(_vp)]),
    "qob_batch_bind": (_i32, [_vp, _vp, _i32, _P(_vp), _P(_vp)]),
    "qob_batch_apply": (_i32, [_vp, _vp, _i32, qob_c64, qob_c64, qob_c64, _vp]),
    "qob_batch_axpy": (_i32, [_vp, _vp, _i32, qob_c64, _vp]),
    "qob_batch_info": (_i32, [_vp, _P(_i64), _P(_i32)]),
    "qob_batch_destroy": (_i32, [_vp]),
    "qob_graph_begin": (_i32, [_vp, _vp]),
    "qob_graph_end": (_i32, [_vp, _vp, _P(_vp)]),
    "qob_graph_launch": (_i32, [_vp, _vp, _vp]),
    "qob_graph_destroy": (_i32, [_vp]),
    "qob_dot_d": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp, _vp]),
    "qob_norm_d": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp]),
    "qob_trace_d": (_i32, [_vp, _i32, _i64, _vp, _vp, _vp]),
}

_lib = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load the library once; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path or os.environ.get("QOB_HIP_LIB", LIB_PATH))
    if not p.exists():
        raise QobError(f"libqob_hip.so not found at {p}; run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names() -> list[str]:
    return list(_PROTOS)


def check(status: int) -> None:
    if status == 0:
        return
    msg = load().qob_last_error().decode(errors="replace")
    raise _STATUS.get(status, QobError)(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
