"""Request loop: mixed operator-application jobs admitted by HBM budget,
batched as they arrive, with latency percentiles and throughput.

No reference counterpart: the reference's mul! runs one call at a time on
the caller's task.  A job here is the stepping pattern of a downstream
time-evolution caller on a device-resident state:
    steps x ( mul!(k, H, x, 1, 0) ; x += c k )  then  normalize!(x)
(normalize! of a Ket: src/states.jl:90; of a density matrix, x / tr(x):
src/operators_dense.jl:217).

Scheduling (one process, one GPU):
  * Ket jobs on qubit Hamiltonians share ONE batched launch per step
    (qob_batch_*, all states of the batch in one persistent kernel; the
    step is fused as x' = x + c H x).  The batch is re-formed at every
    cycle boundary: finished jobs leave, waiting jobs join (continuous
    batching), a cycle being ``chunk`` steps on the Ket stream.
  * Density-matrix jobs each take one of ``rho_slots`` streams and run
    there ``rho_chunk`` steps at a time, concurrently with the Ket batch.  With the default of 2 slots the Ket
    stream, the auxiliary stream (uploads, retirements) and the slots get
    hardware queues of their own (4 per process by default): a Ket step
    never waits behind a queued density-matrix sequence.
  * Up to ``depth`` Ket cycles are in flight, so the Ket stream does not
    idle while the host retires and admits jobs.
  * Admission is FIFO per engine and requires the job's device footprint
    (state + step buffer) to fit the HBM budget; a job larger than the
    budget is rejected, never queued forever.
  * Completion is polled with events (qob_event_query); latency runs from
    the scheduled arrival to the moment the result is on the host (the
    normalised Ket itself; for a density matrix the FP64 trace that
    normalised it, the matrix staying resident on the device).
"""
from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import c64, call
from .core import DeviceArray, _dt, context


# ------------------------------------------------------------------ events
class Event:
    def __init__(self):
        h = C.c_void_p()
        call("qob_event_create", context().h, C.byref(h))
        self.h = h

    def record(self, stream) -> None:
        call("qob_event_record", context().h, self.h, stream)

    def done(self) -> bool:
        d = C.c_int()
        call("qob_event_query", self.h, C.byref(d))
        return bool(d.value)

    def synchronize(self) -> None:
        call("qob_event_synchronize", self.h)

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib.load().qob_event_destroy(self.h)
        except Exception:
            pass


def new_stream():
    s = C.c_void_p()
    call("qob_stream_create", context().h, C.byref(s))
    return s


# ---------------------------------------------------------------- job model
@dataclass
class KetClass:
    """Ket jobs on qubit LazySums; ``ops[L]`` is the Hamiltonian of a chain of L sites."""
    name: str
    ops: Dict[int, object]
    lengths: Sequence[int]

    kind = "ket"


@dataclass
class RhoClass:
    """Density-matrix jobs: left mul! of ``op`` on an n x n state started
    from ``template`` (device-resident) scaled by a per-job factor."""
    name: str
    op: object
    n: int
    template: DeviceArray  # n x n initial state, column-major, in the serving dtype

    kind = "rho"


@dataclass
class Request:
    rid: int
    t_arrival: float          # seconds after the start of the run
    cls: str
    seed: int
    L: int = 0                # chain length for Ket jobs
    payload: Optional[np.ndarray] = None   # initial Ket (host)


@dataclass
class JobResult:
    rid: int
    cls: str
    status: str               # "done" or "rejected"
    t_arrival: float
    t_admit: float = 0.0
    t_done: float = 0.0
    value: float = 0.0        # norm (Ket) or trace (rho) before normalisation, FP64
    state: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


def make_requests(classes: Sequence, n: int, rate: float, seed: int, dtype=np.complex128) -> List[Request]:
    """n requests with Poisson arrivals at ``rate`` per second, classes drawn
    uniformly; Ket jobs draw their chain length uniformly from the class."""
    g = np.random.Generator(np.random.PCG64(seed))
    gaps = g.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    t = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(n):
        c = classes[int(g.integers(0, len(classes)))]
        r = Request(i, float(t[i]), c.name, int(g.integers(0, 2 ** 31)))
        if c.kind == "ket":
            r.L = int(c.lengths[int(g.integers(0, len(c.lengths)))])
            gg = np.random.Generator(np.random.PCG64(r.seed))
            v = gg.standard_normal(2 ** r.L) + 1j * gg.standard_normal(2 ** r.L)
            r.payload = (v / np.linalg.norm(v)).astype(dtype)
        reqs.append(r)
    return reqs


class _Pool:
    """Device buffers reused across jobs, keyed by (elements, dtype)."""

    def __init__(self):
        self.free: Dict[tuple, list] = {}

    def get(self, n: int, dtype) -> DeviceArray:
        lst = self.free.get((n, np.dtype(dtype).str))
        return lst.pop() if lst else DeviceArray(n, dtype)

    def put(self, d: DeviceArray) -> None:
        self.free.setdefault((d.n, d.dtype.str), []).append(d)


@dataclass
class _Job:
    req: Request
    res: JobResult
    bytes: int
    bufs: list = field(default_factory=list)
    steps_done: int = 0
    last_cycle: int = -1


class Server:
    """Runs a list of requests in real time; see the module docstring."""

    def __init__(self, classes: Sequence, dtype=np.complex128, budget_bytes: int = 32 << 30, rho_slots: int = 2,
                 steps: int = 32, chunk: int = 4, c: complex = -0.01j, keep_states: bool = False,
                 max_batch: int = 1024, depth: int = 2, rho_chunk: int = 8):
        if chunk % 2 or steps % chunk:
            raise ValueError("chunk must be even and divide steps")
        self.classes = {c.name: c for c in classes}
        self.dtype = np.dtype(dtype)
        self.dt = _dt(self.dtype)
        self.item = self.dtype.itemsize
        self.budget = int(budget_bytes)
        self.steps, self.chunk, self.c = steps, chunk, c
        self.keep = keep_states
        self.max_batch = max_batch
        self.depth = max(1, depth)
        self.rho_chunk = max(1, rho_chunk)
        self.ctx = context()
        self.pool = _Pool()
        self.ket_stream = new_stream()
        self.aux_stream = new_stream()
        self.rho = [{"stream": new_stream(), "ev": Event(), "job": None} for _ in range(rho_slots)]
        self.used = 0
        self.peak = 0

    # ------------------------------------------------------------ helpers
    def _footprint(self, req: Request) -> int:
        cl = self.classes[req.cls]
        n = 2 ** req.L if cl.kind == "ket" else cl.n * cl.n
        return 2 * n * self.item

    def _clock(self) -> float:
        return time.perf_counter() - self._t0

    # --------------------------------------------------------- Ket engine
    def _ket_cycle(self) -> bool:
        """Keep up to ``depth`` cycles of the Ket batch in flight.  A cycle is
        ``chunk`` fused steps of every member; membership changes between
        cycles (finished jobs leave, admitted jobs join).  Uploads and
        retirements run on the auxiliary stream, so the Ket stream never
        waits for the host."""
        worked = False
        # cycles that finished: retire the jobs whose last cycle it was
        while self._ket_inflight and self._ket_inflight[0][0].done():
            ev, cid, _ = self._ket_inflight.popleft()
            self._ket_events.append(ev)
            for j in [j for j in self._ket_pending if j.last_cycle == cid]:
                self._ket_pending.remove(j)
                x = j.bufs[0]
                nrm = C.c_double()
                call("qob_norm", self.ctx.h, self.dt, x.n, x.ptr, C.byref(nrm), self.aux_stream)
                call("qob_normalize_ket", self.ctx.h, self.dt, x.n, x.ptr, self.aux_stream)
                host = np.empty(x.n, self.dtype)
                call("qob_memcpy_dtoh", self.ctx.h, host.ctypes.data_as(C.c_void_p), x.ptr, x.nbytes,
                     self.aux_stream)
                j.res.value = nrm.value
                j.res.state = host if self.keep else None
                self._finish(j)
                worked = True
        if len(self._ket_inflight) >= self.depth:
            return worked
        # admit waiting jobs, FIFO, while the budget allows
        while self._ket_wait and len(self._ket_active) < self.max_batch:
            j = self._ket_wait[0]
            if self.used + j.bytes > self.budget:
                break
            self._ket_wait.popleft()
            self._admit(j)
            n = 2 ** j.req.L
            xa, xb = self.pool.get(n, self.dtype), self.pool.get(n, self.dtype)
            p = np.ascontiguousarray(j.req.payload, dtype=self.dtype)
            call("qob_memcpy_htod", self.ctx.h, xa.ptr, p.ctypes.data_as(C.c_void_p), xa.nbytes, self.aux_stream)
            j.bufs = [xa, xb]
            self._ket_active.append(j)
            worked = True
        if not self._ket_active:
            return worked
        members = tuple(j.req.rid for j in self._ket_active)
        if members != self._ket_members:
            ops = [self.classes[j.req.cls].ops[j.req.L] for j in self._ket_active]
            hs = (C.c_void_p * len(ops))(*[op._handle().value for op in ops])
            h = C.c_void_p()
            call("qob_batch_create", self.ctx.h, len(ops), hs, self.dt, C.byref(h))
            # a replaced batch may still be read by a cycle in flight, and
            # freeing device memory waits for the whole device: keep it
            # until the end of the run
            if self._batch is not None:
                self._graveyard.append(self._batch)
            self._batch = _BatchHandle(h)
            self._ket_members = members
            xa = (C.c_void_p * len(ops))(*[j.bufs[0].ptr.value for j in self._ket_active])
            xb = (C.c_void_p * len(ops))(*[j.bufs[1].ptr.value for j in self._ket_active])
            call("qob_batch_bind", self.ctx.h, h, 0, xa, xb)
            call("qob_batch_bind", self.ctx.h, h, 1, xb, xa)
        # chunk steps, x' = x + c H x, alternating the two buffers (chunk is
        # even: every member's state is back in its first buffer afterwards)
        for s in range(self.chunk):
            call("qob_batch_apply", self.ctx.h, self._batch.h, s % 2, c64(self.c), c64(0.0), c64(1.0),
                 self.ket_stream)
        ev = self._ket_events.pop() if self._ket_events else Event()
        ev.record(self.ket_stream)
        self._cycle_id += 1
        self._ket_inflight.append((ev, self._cycle_id, members))
        keep = []
        for j in self._ket_active:
            j.steps_done += self.chunk
            if j.steps_done >= self.steps:
                j.last_cycle = self._cycle_id
                self._ket_pending.append(j)
            else:
                keep.append(j)
        self._ket_active = keep
        return True

    # --------------------------------------------------------- rho engine
    def _rho_steps(self, slot, j) -> None:
        """Enqueue the job's next ``rho_chunk`` steps on the slot's stream."""
        cl = self.classes[j.req.cls]
        x, k = j.bufs
        s = slot["stream"]
        h = cl.op._handle()
        nn = cl.n * cl.n
        n = min(self.rho_chunk, self.steps - j.steps_done)
        for _ in range(n):
            call("qob_mul_op_left", self.ctx.h, self.dt, h, x.ptr, cl.n, cl.n, k.ptr, cl.n, c64(1.0), c64(0.0), s)
            call("qob_axpy", self.ctx.h, self.dt, nn, c64(self.c), k.ptr, x.ptr, s)
        j.steps_done += n
        slot["ev"].record(s)

    def _rho_cycle(self) -> bool:
        """Each slot runs one job, ``rho_chunk`` steps at a time: short
        queues let the Ket batch's launches interleave with long sequences."""
        worked = False
        for slot in self.rho:
            j = slot["job"]
            if j is not None and slot["ev"].done():
                if j.steps_done < self.steps:
                    self._rho_steps(slot, j)
                    worked = True
                    continue
                x = j.bufs[0]
                cl = self.classes[j.req.cls]
                tr = _lib.qob_c64()
                call("qob_trace", self.ctx.h, self.dt, cl.n, x.ptr, C.byref(tr), slot["stream"])
                call("qob_normalize_op", self.ctx.h, self.dt, cl.n, x.ptr, slot["stream"])
                j.res.value = tr.re
                if self.keep:
                    host = np.empty(x.n, self.dtype)
                    call("qob_memcpy_dtoh", self.ctx.h, host.ctypes.data_as(C.c_void_p), x.ptr, x.nbytes,
                         slot["stream"])
                    j.res.state = host.reshape(cl.n, cl.n, order="F")
                else:
                    call("qob_stream_synchronize", self.ctx.h, slot["stream"])
                self._finish(j)
                slot["job"] = None
                worked = True
            if slot["job"] is None and self._rho_wait:
                j = self._rho_wait[0]
                if self.used + j.bytes > self.budget:
                    continue
                self._rho_wait.popleft()
                self._admit(j)
                cl = self.classes[j.req.cls]
                s = slot["stream"]
                nn = cl.n * cl.n
                x, k = self.pool.get(nn, self.dtype), self.pool.get(nn, self.dtype)
                j.bufs = [x, k]
                scale = 1.0 + 0.5 * np.random.Generator(np.random.PCG64(j.req.seed)).uniform()
                call("qob_copy", self.ctx.h, self.dt, nn, cl.template.ptr, x.ptr, s)
                call("qob_scale", self.ctx.h, self.dt, nn, c64(scale), x.ptr, s)
                self._rho_steps(slot, j)
                slot["job"] = j
                worked = True
        return worked

    def _admit(self, j: _Job) -> None:
        j.res.t_admit = self._clock()
        self.used += j.bytes
        self.peak = max(self.peak, self.used)

    def _finish(self, j: _Job) -> None:
        j.res.t_done = self._clock()
        j.res.status = "done"
        for b in j.bufs:
            self.pool.put(b)
        j.bufs = []
        self.used -= j.bytes
        self._results.append(j.res)

    # ---------------------------------------------------------------- run
    def run(self, requests: Sequence[Request]) -> List[JobResult]:
        reqs = deque(sorted(requests, key=lambda r: r.t_arrival))
        self._results: List[JobResult] = []
        self._ket_wait: deque = deque()
        self._rho_wait: deque = deque()
        self._ket_active: List[_Job] = []
        self._ket_pending: List[_Job] = []
        self._ket_inflight: deque = deque()
        self._ket_events: list = []
        self._ket_members: tuple = ()
        self._cycle_id = 0
        self._batch = None
        self._graveyard: list = []
        total = len(reqs)
        self.used = self.peak = 0
        self._t0 = time.perf_counter()
        while len(self._results) < total:
            now = self._clock()
            while reqs and reqs[0].t_arrival <= now:
                r = reqs.popleft()
                res = JobResult(r.rid, r.cls, "queued", r.t_arrival)
                j = _Job(r, res, self._footprint(r))
                if j.bytes > self.budget:
                    res.status, res.t_done = "rejected", now
                    self._results.append(res)
                    continue
                (self._ket_wait if self.classes[r.cls].kind == "ket" else self._rho_wait).append(j)
            worked = self._ket_cycle()
            worked = self._rho_cycle() or worked
            if not worked:
                if reqs and not (self._ket_active or self._ket_wait or self._ket_inflight or self._rho_wait
                                 or any(s["job"] for s in self.rho)):
                    time.sleep(max(0.0, min(1e-3, reqs[0].t_arrival - self._clock())))
                else:
                    time.sleep(20e-6)
        call("qob_stream_synchronize", self.ctx.h, self.ket_stream)
        self._graveyard.clear()
        self._batch = None
        return sorted(self._results, key=lambda r: r.rid)


class _BatchHandle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib.load().qob_batch_destroy(self.h)
        except Exception:
            pass


def summarize(results: Sequence[JobResult]) -> dict:
    """p50 / p99 latency (ms) overall and per class, and throughput (jobs/s)."""
    done = [r for r in results if r.status == "done"]
    out = {"jobs": len(results), "done": len(done), "rejected": len(results) - len(done)}
    if not done:
        return out
    lat = np.array([r.latency for r in done]) * 1e3
    span = max(r.t_done for r in done) - min(r.t_arrival for r in done)
    out.update({"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "throughput_jobs_s": len(done) / span if span > 0 else float("inf"), "span_s": span})
    per = {}
    for name in sorted({r.cls for r in done}):
        l = np.array([r.latency for r in done if r.cls == name]) * 1e3
        per[name] = {"n": int(l.size), "p50_ms": float(np.percentile(l, 50)), "p99_ms": float(np.percentile(l, 99))}
    out["per_class"] = per
    return out


# ---------------------------------------------------------------- job mixes
def _template(seed: int, n: int, dtype) -> DeviceArray:
    from . import workloads as W
    rho = W.ginibre_rho_device(seed, n)
    if np.dtype(dtype) == np.complex128:
        return rho.data
    d = DeviceArray(n * n, dtype, (n, n))
    call("qob_convert", context().h, _dt(np.complex128), rho.data.ptr, _dt(dtype), d.ptr, n * n, None)
    call("qob_stream_synchronize", context().h, None)
    return d


def make_classes(dtype=np.complex128, small: bool = False) -> list:
    """Job shapes of the four workloads: the 10-site spin chain (Ket), the
    transverse-field Ising chains of 8..20 sites (Ket), the Fock x Fock x
    Spin(3/2) SparseOperator on a 6400 x 6400 density matrix and the dense
    (16, 32, 16) LazyTensor on an 8192 x 8192 density matrix.  small=True
    gives the same kinds at test sizes."""
    from . import workloads as W
    from .core import SparseOperator
    if small:
        chain = KetClass("chain", {6: W.spin_chain(6, 1).build()}, [6])
        tfim = KetClass("tfim", {L: W.tfim_chain(L, 100 + L).build() for L in (5, 7, 9, 11)}, [5, 7, 9, 11])
        fs = W.fock_fock_spin(5, 1.5, 2)
        sparse = RhoClass("fock_rho", SparseOperator(fs.basis, fs.basis, fs.H), len(fs.basis),
                          _template(7, len(fs.basis), dtype))
        lt = W.dense_lazytensor((16, 16), 3)
        dense = RhoClass("dense_rho", lt, 256, _template(8, 256, dtype))
        return [chain, tfim, sparse, dense]
    chain = KetClass("chain10", {10: W.spin_chain(10, 1).build()}, [10])
    tfim = KetClass("tfim8_20", {L: W.tfim_chain(L, 100 + L).build() for L in range(8, 21)}, list(range(8, 21)))
    fs = W.fock_fock_spin(39, 1.5, 1)
    sparse = RhoClass("fock6400", SparseOperator(fs.basis, fs.basis, fs.H), 6400, _template(5, 6400, dtype))
    dense = RhoClass("dense8192", W.dense_lazytensor((16, 32, 16), 1), 8192, _template(6, 8192, dtype))
    return [chain, tfim, sparse, dense]
