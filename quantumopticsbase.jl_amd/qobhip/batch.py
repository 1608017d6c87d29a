"""Batched application of many qubit LazySums in one launch, and HIP graph
capture of library call sequences (C entry points qob_batch_* and
qob_graph_*).  No reference counterpart: the reference applies one operator
per mul! call; these exist for time stepping many states at once.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import c64, call
from . import core as _core
from .core import DeviceArray, Ket, _dt, context


class Batch:
    """ops[j] acts on states of length 2^L_j; one launch applies all of them."""

    def __init__(self, ops: Sequence, dtype=np.complex128):
        self.ops = list(ops)
        self.dtype = np.dtype(dtype)
        ctx = context()
        hs = (C.c_void_p * len(ops))(*[op._handle().value for op in ops])
        h = C.c_void_p()
        call("qob_batch_create", ctx.h, len(ops), hs, _dt(self.dtype), C.byref(h))
        self.h = h
        self._bufs = {}

    def info(self):
        tot, n = C.c_int64(), C.c_int()
        call("qob_batch_info", self.h, C.byref(tot), C.byref(n))
        return tot.value, n.value

    def bind(self, slot: int, ins: Sequence[DeviceArray], outs: Sequence[DeviceArray]) -> None:
        """Bind per-state input/output device buffers to a slot.

        Re-binding a bound slot waits for the whole device first."""
        n = len(self.ops)
        if len(ins) != n or len(outs) != n:
            raise ValueError(f"bind needs {n} input and {n} output buffers, got {len(ins)} and {len(outs)}")
        for j, (x, y) in enumerate(zip(ins, outs)):
            if x.dtype != self.dtype or y.dtype != self.dtype:
                raise TypeError(f"state {j}: buffers must be {self.dtype}")
            d = self.ops[j].shape[1]
            if x.n < d or y.n < d:
                raise ValueError(f"state {j}: buffers hold fewer than {d} elements")
        pin = (C.c_void_p * len(ins))(*[d.ptr.value for d in ins])
        pout = (C.c_void_p * len(outs))(*[d.ptr.value for d in outs])
        call("qob_batch_bind", context().h, self.h, slot, pin, pout)
        self._bufs[slot] = (list(ins), list(outs))

    def apply(self, slot: int, alpha=1.0, beta=0.0, gamma=0.0, stream=None) -> None:
        """out_j = alpha op_j in_j + gamma in_j + beta out_j for every j."""
        _core._note_capture(self, self._bufs.get(slot))
        call("qob_batch_apply", context().h, self.h, slot, c64(alpha), c64(beta), c64(gamma), stream)

    def axpy(self, slot: int, a, stream=None) -> None:
        """out_j += a in_j for every j."""
        _core._note_capture(self, self._bufs.get(slot))
        call("qob_batch_axpy", context().h, self.h, slot, c64(a), stream)

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib.load().qob_batch_destroy(self.h)
        except Exception:
            pass


class Graph:
    """Capture library calls issued inside ``with Graph() as g:`` into a HIP
    graph; ``g.launch()`` replays them.  Every captured call must have run
    once before capture.

    A replay reads the operators, batches and device arrays that the captured
    calls used, at the addresses they had during capture.  The Graph keeps a
    reference to every one of them that went through this package (mul_,
    Batch.apply / axpy, set_factors_device) for as long as it lives; arrays
    reached only through raw library calls can be added with ``keep``.  The
    library itself keeps its own workspaces, plans, operator tables and
    batch tables alive while any graph of the context exists."""

    def __init__(self, stream=None, keep=()):
        self.stream = stream
        self.h = None
        self.keep = list(keep)

    def __enter__(self):
        call("qob_graph_begin", context().h, self.stream)
        _core._CAPTURE_KEEP.append(self.keep)
        return self

    def __exit__(self, exc_type, exc, tb):
        _core._CAPTURE_KEEP.pop()
        h = C.c_void_p()
        call("qob_graph_end", context().h, self.stream, C.byref(h))
        self.h = h
        return False

    def launch(self) -> None:
        call("qob_graph_launch", context().h, self.h, self.stream)

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                _lib.load().qob_graph_destroy(self.h)
        except Exception:
            pass
