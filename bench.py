#!/usr/bin/env python3
"""Benchmark of the mul! hot path on one MI355X.  Prints ONE JSON line.

Default workload (--workload dense8192), the largest single-GPU
configuration: a LazyTensor of three dense random Hermitian factors
(dims 16, 32, 16; entries complex Gaussian, symmetrised) applied from the
left to an 8192 x 8192 Ginibre trace-1 density matrix in ComplexF64
(--dtype c32 for the ComplexF32 FP32-MFMA form); one step = one
mul!(out, LT, rho, 1, 0).  vs_baseline is the ratio to the plain per-call
GPU path on the same workload (the mode-by-mode chain).  The line also
carries the 6400 x 6400 SparseOperator figures under config.sparse6400.

Other workloads (--workload):
  sparse6400 SparseOperator on Fock(39) (x) Fock(39) (x) Spin(3/2), N = 6400,
             nnz = 15522, mul! from the left and the right onto a Ginibre
             trace-1 density matrix; one step = 2 applies
  ket10      10-site spin chain, H = sum J sx sx + sum h sz (19 LazyTensors),
             mul!(out, H, psi, 1, 0) on a normalised complex-Gaussian Ket
  batch256   256 transverse-field Ising chains of 8..20 sites, one launch per
             step x' = x + c H x for all of them
  denseop    DenseOperator (8192 x 8192) on a density matrix (MFMA GEMM), on a
             Ket and on a Bra (gemv)
  serve      Poisson request stream of mixed jobs (32 x (mul! + axpy) +
             normalize!) over the four shapes above; p50/p99 latency

Kernel times come from events on the stream the kernels run on; the
CPU baseline is the C restatement of the reference loops (oracle/cpu_ref.c)
or the numpy restatement (oracle/qob_oracle.py) on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "quantumopticsbase.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP64_MFMA_TFLOPS = 78.6
FP32_MFMA_TFLOPS = 157.3
METRIC = "applies/sec + HBM GB/s (% of roofline) vs plain FP64 GPU; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_of(key):
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            return json.loads(tf.read_text()).get(key)
        except Exception:
            return None
    return None


def timed(step, steps: int, warmup: int, sync) -> float:
    """W untimed warmup steps, then exactly K steps between synchronisations; seconds."""
    for _ in range(warmup):
        step()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sync()
    return time.perf_counter() - t0


def probe(ctx, src, dst, step):
    """Fixed launch sequence for counter runs (rocprofv3 --pmc): 3 plain
    copies of ``src`` (a calibration kernel of known bytes), then 5 steps."""
    from qobhip import _lib
    for _ in range(3):
        _lib.call("qob_copy", ctx.h, src.qdt, src.n, src.ptr, dst.ptr, None)
    ctx.synchronize()
    step()
    ctx.synchronize()
    for _ in range(5):
        step()
    ctx.synchronize()
    log(f"bench: probe done (copy bytes per launch {2 * src.nbytes})")
    return None


def host_cpu() -> str:
    """Model name and logical CPU count of the host the CPU baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def base_result(args, value, unit, ms_per_step, dtype, data, config, roofline, cpu):
    if cpu is not None:
        cpu = {**cpu, "host": host_cpu()}
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config, "roofline": roofline,
            "cpu_baseline": cpu}


def blas_threads() -> int:
    try:
        from threadpoolctl import threadpool_info
        n = [i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"]
        return int(max(n)) if n else 1
    except Exception:
        return 1


# --------------------------------------------------------------- sparse6400
def cpu_sparse(spec, rho_h, budget_s: float = 12.0):
    """Reference loops restated in C (oracle/cpu_ref.c), one thread."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import cref
    lib = cref.lib()
    H = spec.H
    n = H.shape[0]
    colptr = (H.indptr.astype(np.int64) + 1)
    rowval = (H.indices.astype(np.int64) + 1)
    nz = np.ascontiguousarray(H.data.astype(np.complex128))
    B = np.asfortranarray(rho_h)
    out = np.empty((n, n), dtype=np.complex128, order="F")
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    applies = 0
    t0 = time.perf_counter()
    while True:
        lib.qobref_gemm_sp_dense(1.0, 0.0, n, n, P(colptr), P(rowval), P(nz), P(B), n, 0.0, 0.0, P(out))
        lib.qobref_gemm_dense_sp(1.0, 0.0, n, n, n, P(colptr), P(rowval), P(nz), P(B), 0.0, 0.0, P(out))
        applies += 2
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    # all-cores upper bound: the same loops with the outer column loop over
    # OpenMP threads (the reference's loops are single-threaded)
    nth = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1))
    omp_applies = 0
    t0 = time.perf_counter()
    while True:
        lib.qobref_gemm_sp_dense_omp(1.0, 0.0, n, n, P(colptr), P(rowval), P(nz), P(B), n, 0.0, 0.0, P(out), nth)
        lib.qobref_gemm_dense_sp_omp(1.0, 0.0, n, n, n, P(colptr), P(rowval), P(nz), P(B), 0.0, 0.0, P(out), nth)
        omp_applies += 2
        omp_el = time.perf_counter() - t0
        if omp_el > budget_s / 3:
            break
    return {"value": applies / el, "unit": "applies/s", "cores": 1, "kind": "port",
            "sample": f"{applies} full applies ({applies // 2} left + {applies // 2} right) of the same "
                      f"6400x6400 workload, C restatement of src/sparsematrix.jl:25-47,115-146, 1 thread, "
                      f"{el:.1f} s",
            "all_cores": {"value": omp_applies / omp_el, "cores": nth,
                          "sample": f"{omp_applies} applies, the same loops with the column loop over {nth} "
                                    f"OpenMP threads, {omp_el:.1f} s"}}


def run_sparse6400(args):
    import qobhip as Q
    from qobhip import workloads as W
    from qobhip.tune import EventTimer

    log("bench: building workload (Fock(39) x Fock(39) x Spin(3/2), N=6400)")
    spec = W.fock_fock_spin(seed=1)
    n = spec.H.shape[0]
    nnz = spec.H.nnz
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    rho = W.ginibre_rho_device(seed=2, n=n)
    outL = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    outR = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    ctx = Q.context()
    stream = ctx.stream

    def left():
        Q.mul_(outL, H, rho, 1.0, 0.0)

    def right():
        Q.mul_(outR, rho, H, 1.0, 0.0)

    if args.probe:
        # fixed launch sequence for counter collection (rocprofv3 --pmc):
        # 5 x plain copy of N^2 elements, 5 x left, 5 x right
        from qobhip import _lib
        for _ in range(5):
            _lib.call("qob_copy", ctx.h, 0, n * n, rho.data.ptr, outL.data.ptr, None)
        ctx.synchronize()
        for _ in range(5):
            left()
        ctx.synchronize()
        for _ in range(5):
            right()
        ctx.synchronize()
        log(f"bench: probe done (copy bytes per launch {32 * n * n})")
        return None

    log(f"bench: warmup {args.warmup}, timing {args.steps} steps")
    el = timed(lambda: (left(), right()), args.steps, args.warmup, ctx.synchronize)
    ms_per_step = el * 1e3 / args.steps
    applies_per_s = 2 * args.steps / el

    # Dominant-kernel roofline: average kernel duration from events on the
    # library's stream; algorithmic bytes per launch (beta = 0): read rho
    # (16 N^2) + write out (16 N^2) + CSR/CSC (nnz*(16+4) + (N+1)*4).
    ev = EventTimer(ctx.h, stream)
    reps = max(20, min(200, args.steps))
    from qobhip import tune as T
    ms_left = ev(left, reps)
    ms_right = ev(right, reps)
    # the same bytes moved by the library's streaming copy kernel (read rho,
    # write out): the practical ceiling for a kernel that reads and writes
    # N^2 elements once
    from qobhip import _lib
    copy = lambda: _lib.call("qob_copy", ctx.h, 0, n * n, rho.data.ptr, outL.data.ptr, None)
    ms_copy = ev(copy, reps)
    bytes_per_launch = 32 * n * n + nnz * 20 + (n + 1) * 4
    dom = "left" if ms_left >= ms_right else "right"
    ms_dom = max(ms_left, ms_right)
    achieved = bytes_per_launch / (ms_dom * 1e-3) / 1e9

    # the same Hamiltonian as a LazySum of LazyTensors: fused (one gather) vs
    # the plain per-term path
    # ("fused" runs it as its summed sparse matrix on rho, "gather" evaluates
    # every term per element in one launch)
    lz = W.fock_fock_spin_lazy(spec)
    outZ = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    lz_ms = {}
    for mode, name in ((Q.QOB_MODE_FUSED, "fused"), (Q.QOB_MODE_PLAIN, "plain"))[:1 if args.no_plain else 2]:
        ctx.set_mode(mode)
        f = lambda: Q.mul_(outZ, lz, rho, 1.0, 0.0)
        f()
        ctx.synchronize()
        lz_ms[name] = ev(f, 10)
    ctx.set_mode(Q.QOB_MODE_FUSED)
    T.set_("ls_materialize", 0)
    lzg = W.fock_fock_spin_lazy(spec)
    f = lambda: Q.mul_(outZ, lzg, rho, 1.0, 0.0)
    f()
    ctx.synchronize()
    lz_ms["gather"] = ev(f, 10)
    T.set_("ls_materialize", 1)

    # expect(H, rho) = tr(H rho): one pass over the nonzeros (fused) against
    # the plain form (apply into a temporary, then trace)
    ex_ms = {}
    for mode, name in ((Q.QOB_MODE_FUSED, "fused"), (Q.QOB_MODE_PLAIN, "apply_then_trace")):
        ctx.set_mode(mode)
        Q.expect(H, rho)
        t0 = time.perf_counter()
        for _ in range(20):
            Q.expect(H, rho)
        ex_ms[name] = (time.perf_counter() - t0) * 1e3 / 20
    ctx.set_mode(Q.QOB_MODE_FUSED)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (C restatement, 1 thread)")
        cpu = cpu_sparse(spec, rho.to_numpy())

    config = {
        "workload": "SparseOperator mul! left + right on a 6400x6400 ComplexF64 density matrix "
                    "(Fock(39) x Fock(39) x Spin(3/2), nnz 15522); 1 step = 2 applies",
        "N": n, "nnz": nnz, "element": "ComplexF64",
        "kernel_ms": {"left": ms_left, "right": ms_right, "copy_same_bytes": ms_copy},
        "hbm_gbs": {"left": bytes_per_launch / (ms_left * 1e-3) / 1e9,
                    "right": bytes_per_launch / (ms_right * 1e-3) / 1e9,
                    "copy": 32 * n * n / (ms_copy * 1e-3) / 1e9},
        "lazysum_form_left_ms": {**lz_ms, "fused_speedup": lz_ms["plain"] / lz_ms["fused"] if "plain" in lz_ms else None},
        "expect_rho_ms_host_clock": ex_ms,
    }
    # the Hamiltonian's pattern splits into components of at most 4 rows, so
    # with closed blocks enabled the left product runs k_sell_left_closed
    left_kernel = "k_sell_left_closed" if T.get("sell_closed") else "k_sell_left_rr"
    # and into components of at most 4 columns: the right product runs the
    # component-group kernel when groups are built and enabled
    right_kernel = "k_csc_right_groups" if T.get("csc_groups") and T.get("csc_gkernel") else "k_csc_right"
    roof = {"bound": "hbm", "kernel": left_kernel if dom == "left" else right_kernel,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_of(f"csr_{dom}"), "algorithmic_bytes_per_launch": bytes_per_launch}
    return base_result(args, applies_per_s, "applies/s", ms_per_step, "f64",
                       "synthetic (seeded Fock-Fock-spin Hamiltonian, Ginibre trace-1 rho)", config, roof, cpu)


# ------------------------------------------------------------------- ket10
def cpu_lazysum(specs_psis, budget_s: float):
    """C restatement of the reference's recursive LazySum engine, 1 thread:
    applies of each (spec, psi) in turn until the budget is used."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from cref import LazySumKet
    engines = [(LazySumKet(s.basis.shape, s.terms), p) for s, p in specs_psis]
    outs = [np.empty(e.D, np.complex128) for e, _ in engines]
    n = 0
    elems = 0
    t0 = time.perf_counter()
    while True:
        for (e, p), o in zip(engines, outs):
            e.mul(o, p, 1.0, 0.0)
            n += 1
            elems += e.D
        el = time.perf_counter() - t0
        if el > budget_s:
            return n, elems, el


def cpu_lazysum_threads(specs_psis, budget_s: float):
    """All-cores upper bound of cpu_lazysum: the same C engine on a pool of
    threads (the foreign calls release the GIL), thread w applying states
    w, w + T, ... (every state when there is only one) into its own output
    buffers until the budget is used.  Returns (applies, elements, seconds,
    threads)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, str(ROOT / "oracle"))
    from cref import LazySumKet
    nth = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1))
    engines = [(LazySumKet(s.basis.shape, s.terms), p) for s, p in specs_psis]
    t0 = time.perf_counter()

    def worker(w):
        mine = engines[w::nth] if len(engines) >= nth else engines
        outs = [np.empty(e.D, np.complex128) for e, _ in mine]
        n = elems = 0
        while True:
            for (e, p), o in zip(mine, outs):
                e.mul(o, p, 1.0, 0.0)
                n += 1
                elems += e.D
            if time.perf_counter() - t0 > budget_s:
                return n, elems

    with ThreadPoolExecutor(nth) as ex:
        res = list(ex.map(worker, range(nth)))
    el = time.perf_counter() - t0
    return sum(r[0] for r in res), sum(r[1] for r in res), el, nth


def run_ket10(args):
    import qobhip as Q
    from qobhip import workloads as W
    from qobhip.tune import EventTimer

    L = 10
    D = 2 ** L
    spec = W.spin_chain(L, 1)
    H = spec.build()
    ctx = Q.context()
    g = W.rng(2)
    psi0 = W.normalized_ket(g, D)
    psi = Q.Ket(spec.basis, psi0)
    out = Q.Ket(spec.basis, Q.DeviceArray(D))
    step = lambda: Q.mul_(out, H, psi, 1.0, 0.0)
    if args.probe:
        return probe(ctx, psi.data, out.data, step)
    log(f"bench: ket10 warmup {args.warmup}, timing {args.steps} steps")
    el = timed(step, args.steps, args.warmup, ctx.synchronize)
    ev = EventTimer(ctx.h, ctx.stream)
    # per-apply device time with the host out of the way: a captured graph
    with Q.Graph(ctx.stream) as gr:
        for _ in range(100):
            step()
    gr.launch()
    ctx.synchronize()
    graph_ms = ev(gr.launch, 20) / 100
    # the plain per-call path: one launch and one output pass per term
    plain_rate = None
    if not args.no_plain:
        ctx.set_mode(Q.QOB_MODE_PLAIN)
        step()
        ctx.synchronize()
        el_plain = timed(step, max(100, args.steps // 4), 10, ctx.synchronize)
        plain_rate = max(100, args.steps // 4) / el_plain
        ctx.set_mode(Q.QOB_MODE_FUSED)
    # 1024 such Kets in one batched launch
    nb = 1024
    bt = Q.Batch([H] * nb)
    ins = [Q.DeviceArray.from_numpy(psi0) for _ in range(nb)]
    outs = [Q.DeviceArray(D) for _ in range(nb)]
    bt.bind(0, ins, outs)
    bt.apply(0)
    ctx.synchronize()
    batch_ms = ev(lambda: bt.apply(0, 1.0, 0.0, 0.0), 50)
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (C recursive engine, 1 thread)")
        n, _, t = cpu_lazysum([(spec, psi0)], 10.0)
        na, _, ta, nth = cpu_lazysum_threads([(spec, psi0)], 3.0)
        cpu = {"value": n / t, "unit": "applies/s", "cores": 1, "kind": "port",
               "sample": f"{n} applies of the same H and psi, C restatement of src/operators_lazysum.jl:189-200 + "
                         f"src/operators_lazytensor.jl:652-685, 1 thread, {t:.1f} s",
               "all_cores": {"value": na / ta, "cores": nth,
                             "sample": f"{na} applies, {nth} threads applying the same H and psi into their own "
                                       f"outputs, {ta:.1f} s"}}
    byts = 32 * D
    roof = {"bound": "hbm", "kernel": "k_qubit (1024-Ket batch)", "achieved": byts * nb / (batch_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": byts * nb / (batch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": byts * nb,
            "single_apply_kernel_us": graph_ms * 1e3}
    config = {"workload": "LazySum (9 J sx sx + 10 h sz LazyTensors) mul! on one 1024-dim Ket, per call",
              "L": L, "D": D, "terms": 19,
              "per_call_applies_s": args.steps / el, "graph_replayed_applies_s": 1e3 / graph_ms,
              "plain_per_term_applies_s": plain_rate,
              "fused_vs_plain": (args.steps / el) / plain_rate if plain_rate else None,
              "batch1024_applies_s": nb / (batch_ms * 1e-3), "batch1024_ms": batch_ms}
    return base_result(args, args.steps / el, "applies/s", el * 1e3 / args.steps, "f64",
                       "synthetic (seeded J, h ~ U[-1,1], complex-Gaussian Ket)", config, roof, cpu)


# --------------------------------------------------------------- dense8192
def _events(ctx, n):
    from qobhip import _lib
    evs = []
    for _ in range(n):
        e = C.c_void_p()
        _lib.call("qob_event_create", ctx.h, C.byref(e))
        evs.append(e)
    return evs


def _elapsed(a, b):
    from qobhip import _lib
    v = C.c_float()
    _lib.call("qob_event_elapsed", a, b, C.byref(v))
    return float(v.value)


def timed_events(ctx, step, steps: int, warmup: int):
    """W untimed warmup steps, then exactly K steps between synchronisations.
    Events on the library's stream before the first and after the last timed
    step give the kernel time over the timed region itself (nothing else is
    enqueued between the steps).  Returns (host seconds, event ms over the K
    steps)."""
    from qobhip import _lib
    for _ in range(warmup):
        step()
    ctx.synchronize()
    e0, e1 = _events(ctx, 2)
    _lib.call("qob_event_record", ctx.h, e0, ctx.stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    _lib.call("qob_event_record", ctx.h, e1, ctx.stream)
    ctx.synchronize()
    el = time.perf_counter() - t0
    total = _elapsed(e0, e1)
    for e in (e0, e1):
        _lib.load().qob_event_destroy(e)
    return el, total


def latency_events(ctx, step, n: int):
    """Per-apply device latency: an event after every one of n steps."""
    from qobhip import _lib
    evs = _events(ctx, n + 1)
    _lib.call("qob_event_record", ctx.h, evs[0], ctx.stream)
    for i in range(n):
        step()
        _lib.call("qob_event_record", ctx.h, evs[i + 1], ctx.stream)
    ctx.synchronize()
    per = [_elapsed(evs[i], evs[i + 1]) for i in range(n)]
    for e in evs:
        _lib.load().qob_event_destroy(e)
    return per


def sparse6400_figures(steps: int = 20):
    """The 6400x6400 SparseOperator density-matrix workload, kernels only:
    left and right apply times from events, applies/s of a left + right step,
    and the dominant kernel's fraction of HBM peak."""
    import qobhip as Q
    from qobhip import _lib, workloads as W
    from qobhip.tune import EventTimer

    spec = W.fock_fock_spin(seed=1)
    n, nnz = spec.H.shape[0], spec.H.nnz
    H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
    rho = W.ginibre_rho_device(seed=2, n=n)
    outL = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    outR = Q.Operator(rho.basis_l, rho.basis_r, Q.DeviceArray(n * n))
    ctx = Q.context()
    left = lambda: Q.mul_(outL, H, rho, 1.0, 0.0)
    right = lambda: Q.mul_(outR, rho, H, 1.0, 0.0)
    el = timed(lambda: (left(), right()), steps, 3, ctx.synchronize)
    ev = EventTimer(ctx.h, ctx.stream)
    ms_l, ms_r = ev(left, steps), ev(right, steps)
    ms_copy = ev(lambda: _lib.call("qob_copy", ctx.h, 0, n * n, rho.data.ptr, outL.data.ptr, None), steps)
    b = 32 * n * n + nnz * 20 + (n + 1) * 4
    return {"workload": "SparseOperator mul! left + right on a 6400x6400 ComplexF64 density matrix "
                        "(Fock(39) x Fock(39) x Spin(3/2), nnz 15522)",
            "applies_per_s": 2 * steps / el, "kernel_ms": {"left": ms_l, "right": ms_r, "copy_same_bytes": ms_copy},
            "algorithmic_bytes_per_launch": b,
            "hbm_frac": {"left": b / (ms_l * 1e-3) / 1e9 / HBM_PEAK_GBS, "right": b / (ms_r * 1e-3) / 1e9 / HBM_PEAK_GBS}}


def run_dense8192(args):
    import qobhip as Q
    from qobhip import _lib, tune as T, workloads as W
    from qobhip.core import _dt
    from qobhip.tune import EventTimer

    dt = np.complex64 if args.dtype == "c32" else np.complex128
    N = 8192
    dims = (16, 32, 16)
    facs = W.dense_lazytensor_factors(dims, 1)
    LT = W.dense_lazytensor(dims, 1, facs)
    log("bench: dense8192 building rho")
    rho64 = W.ginibre_rho_device(seed=3, n=N)
    ctx = Q.context()
    b = rho64.basis_l
    if dt == np.complex64:
        d32 = Q.DeviceArray(N * N, np.complex64, (N, N))
        _lib.call("qob_convert", ctx.h, Q.QOB_C64, rho64.data.ptr, Q.QOB_C32, d32.ptr, N * N, None)
        rho = Q.Operator(b, b, d32)
    else:
        rho = rho64
    out = Q.Operator(b, b, Q.DeviceArray(N * N, dt, (N, N)))
    outR = Q.Operator(b, b, Q.DeviceArray(N * N, dt, (N, N)))
    left = lambda: Q.mul_(out, LT, rho, 1.0, 0.0)
    right = lambda: Q.mul_(outR, rho, LT, 1.0, 0.0)
    if args.probe:
        return probe(ctx, rho.data, out.data, left)
    log(f"bench: dense8192 {args.dtype} warmup {args.warmup}, timing {args.steps} steps")
    # one step is one left apply, which is one launch of the fused kernel:
    # the events around the timed steps give that kernel's average duration
    el, ev_ms = timed_events(ctx, left, args.steps, args.warmup)
    ms_fused = ev_ms / args.steps
    per = latency_events(ctx, left, max(20, min(200, args.steps)))
    ev = EventTimer(ctx.h, ctx.stream)
    right()
    ctx.synchronize()
    ms_right = ev(right, 5)
    # the plain per-call path: the reference's mode-by-mode chain
    # (_tp_sum_matmul!, one pass over the state per factor), same stream
    _lib.call("qob_tune_set", b"fused_lt", 0)
    plain_steps = max(3, min(10, args.steps))
    el_plain, ev_plain = timed_events(ctx, left, plain_steps, 2)
    ms_chain = ev_plain / plain_steps
    _lib.call("qob_tune_set", b"fused_lt", 1)
    plain_aps = plain_steps / el_plain
    flops = 8.0 * N * N * sum(dims)
    peak = FP32_MFMA_TFLOPS if dt == np.complex64 else FP64_MFMA_TFLOPS
    tflops = flops / (ms_fused * 1e-3) / 1e12
    # FP64 reductions of the result, and the ComplexF32 error against ComplexF64
    left()
    tr = _lib.qob_c64()
    _lib.call("qob_trace", ctx.h, _dt(dt), N, out.data.ptr, C.byref(tr), None)
    extra = {"trace_fp64": [tr.re, tr.im]}
    if dt == np.complex64:
        out64 = Q.Operator(b, b, Q.DeviceArray(N * N))
        Q.mul_(out64, LT, rho64, 1.0, 0.0)
        up = Q.DeviceArray(N * N)
        _lib.call("qob_convert", ctx.h, Q.QOB_C32, out.data.ptr, Q.QOB_C64, up.ptr, N * N, None)
        diff = Q.DeviceArray(N * N)
        _lib.call("qob_copy", ctx.h, Q.QOB_C64, N * N, out64.data.ptr, diff.ptr, None)
        _lib.call("qob_axpy", ctx.h, Q.QOB_C64, N * N, _lib.c64(-1.0), up.ptr, diff.ptr, None)
        nd, nr = C.c_double(), C.c_double()
        _lib.call("qob_norm", ctx.h, Q.QOB_C64, N * N, diff.ptr, C.byref(nd), None)
        _lib.call("qob_norm", ctx.h, Q.QOB_C64, N * N, out64.data.ptr, C.byref(nr), None)
        extra["rel_frobenius_vs_c64"] = nd.value / nr.value
        del out64, up, diff
    cpu = None
    if not args.no_cpu:
        nthr = blas_threads()
        log(f"bench: CPU baseline (numpy restatement of _tp_sum_matmul!, {nthr} BLAS threads)")
        sys.path.insert(0, str(ROOT / "oracle"))
        import qob_oracle as O
        rho_h = rho64.to_numpy()
        oLT = O.OLazyTensor(dims, dims, (1, 2, 3), tuple(O.ODense(f) for f in facs), 1.0)
        res = np.zeros((N, N), np.complex128, order="F")
        n = 0
        t0 = time.perf_counter()
        while True:
            O.mul_op_left(res, oLT, rho_h, 1.0, 0.0)
            n += 1
            t = time.perf_counter() - t0
            if t > 10.0:
                break
        cpu = {"value": n / t, "unit": "applies/s", "cores": nthr, "kind": "port",
               "sample": f"{n} full left applies of the same 8192x8192 workload, numpy restatement of "
                         f"src/operators_lazytensor.jl:281-488 (reshape/permute + BLAS GEMM per mode), ComplexF64, "
                         f"numpy BLAS on {nthr} threads, {t:.1f} s"}
        del rho_h, res
    # factors on sites 1, 2, 3 of widths (16, 32, 16): the register-chained
    # kernel k_lt_reg, four real MFMAs per complex product in both element
    # types, so the MFMAs execute exactly the algorithmic 8 flops per complex
    # multiply-add.  With fused_reg = 0 the fused column kernel runs; its
    # ComplexF32 form uses three real products (Gauss's form).
    reg = T.get("fused_reg") == 1
    mfma_flops = flops if reg or dt == np.complex128 else flops * 3 / 4
    roof = {"bound": "mfma", "kernel": "k_lt_reg" if reg else "k_lt_fused", "achieved": tflops, "peak": peak,
            "unit": "TFLOP/s",
            "frac": tflops / peak, "traffic": traffic_of(f"{'lt_reg' if reg else 'lt_fused'}_{args.dtype}"),
            "algorithmic_flops_per_launch": flops, "mfma_flops_per_launch": mfma_flops,
            "executed_tflops": mfma_flops / (ms_fused * 1e-3) / 1e12,
            "executed_frac": mfma_flops / (ms_fused * 1e-3) / 1e12 / peak,
            "kernel_ms_avg": ms_fused, "timing": "events on the library stream around the timed steps",
            "algorithmic_bytes_per_launch": 2 * N * N * np.dtype(dt).itemsize}
    ps = np.percentile(per, [50, 99])
    config = {"workload": f"config 3: dense LazyTensor (16, 32, 16) of Hermitian factors, left mul! on an "
                          f"8192x8192 {args.dtype} Ginibre trace-1 density matrix; 1 step = 1 apply",
              "N": N, "dims": dims,
              "kernel_ms": {"fused_left": ms_fused, "chain_left": ms_chain, "right_tiles": ms_right},
              "latency_ms_per_apply": {"p50": float(ps[0]), "p99": float(ps[1]), "applies": len(per)},
              "plain_gpu_path": {"applies_per_s": plain_aps,
                                 "what": "mode-by-mode chain on the GPU (_tp_sum_matmul!, one pass per factor)"},
              **extra}
    del out, outR, rho, rho64, LT
    if not args.no_plain:
        log("bench: 6400x6400 sparse figures")
        config["sparse6400"] = sparse6400_figures()
    res = base_result(args, args.steps / el, "applies/s", el * 1e3 / args.steps,
                      "f32" if dt == np.complex64 else "f64",
                      "synthetic (seeded Hermitian factors, Ginibre trace-1 rho)", config, roof, cpu)
    res["vs_baseline"] = (args.steps / el) / plain_aps
    res["vs_baseline_what"] = "value / applies/s of the plain per-call GPU path (mode-by-mode chain), same workload"
    return res


# ---------------------------------------------------------------- denseop
def run_denseop(args):
    """DenseOperator mul! (src/operators_dense.jl:394-396) on an 8192 x 8192
    density matrix (k_zgemm), on a Ket (k_zgemv) and on a Bra (k_zgemv_t)."""
    import qobhip as Q
    from qobhip import workloads as W
    from qobhip.tune import EventTimer

    dt = np.complex64 if args.dtype == "c32" else np.complex128
    N = 8192
    g = W.rng(11)
    b = Q.GenericBasis(N)
    A_h = W.complex_gaussian(g, (N, N)).astype(dt)
    A = Q.Operator(b, b, A_h)
    X = Q.Operator(b, b, W.complex_gaussian(g, (N, N)).astype(dt))
    Y = Q.Operator(b, b, Q.DeviceArray(N * N, dt, (N, N)))
    k = Q.Ket(b, Q.DeviceArray.from_numpy(W.complex_gaussian(g, (N,)).astype(dt), dt))
    ko = Q.Ket(b, Q.DeviceArray(N, dt))
    br = Q.Bra(b, Q.DeviceArray.from_numpy(W.complex_gaussian(g, (N,)).astype(dt), dt))
    bo = Q.Bra(b, Q.DeviceArray(N, dt))
    ctx = Q.context()
    rho_op = lambda: Q.mul_(Y, A, X, 1.0, 0.0)
    ket_op = lambda: Q.mul_(ko, A, k, 1.0, 0.0)
    bra_op = lambda: Q.mul_(bo, br, A, 1.0, 0.0)
    if args.probe:
        return probe(ctx, X.data, Y.data, rho_op)
    steps = max(3, min(args.steps, 20))
    log(f"bench: denseop {args.dtype} warmup {args.warmup}, timing {steps} steps")
    el, ev_ms = timed_events(ctx, rho_op, steps, max(1, min(args.warmup, 3)))
    ms_gemm = ev_ms / steps
    ev = EventTimer(ctx.h, ctx.stream)
    ket_op()
    bra_op()
    ctx.synchronize()
    ms_gemv = ev(ket_op, 50)
    ms_gemv_t = ev(bra_op, 50)
    esz = np.dtype(dt).itemsize
    flops = 8.0 * N ** 3
    peak = FP32_MFMA_TFLOPS if dt == np.complex64 else FP64_MFMA_TFLOPS
    tflops = flops / (ms_gemm * 1e-3) / 1e12
    gemv_bytes = (N * N + 2 * N) * esz
    cpu = None
    if not args.no_cpu:
        nthr = blas_threads()
        log(f"bench: CPU baseline (numpy BLAS, {nthr} threads)")
        X_h = X.to_numpy()
        n, t0 = 0, time.perf_counter()
        while True:
            _ = A_h @ X_h
            n += 1
            t = time.perf_counter() - t0
            if t > 10.0:
                break
        cpu = {"value": n / t, "unit": "applies/s", "cores": nthr, "kind": "port",
               "sample": f"{n} products A rho of the same 8192x8192 operands, numpy BLAS (the reference's "
                         f"mul! calls BLAS gemm, src/operators_dense.jl:394), {nthr} threads, {t:.1f} s"}
    roof = {"bound": "mfma", "kernel": "k_zgemm", "achieved": tflops, "peak": peak, "unit": "TFLOP/s",
            "frac": tflops / peak, "traffic": None, "algorithmic_flops_per_launch": flops,
            "kernel_ms_avg": ms_gemm, "timing": "events on the library stream around the timed steps"}
    config = {"workload": f"DenseOperator mul! on an 8192x8192 {args.dtype} density matrix (left); 1 step = 1 apply",
              "N": N, "kernel_ms": {"zgemm_rho": ms_gemm, "zgemv_ket": ms_gemv, "zgemv_t_bra": ms_gemv_t},
              "gemv_gbs": {"ket": gemv_bytes / (ms_gemv * 1e-3) / 1e9, "bra": gemv_bytes / (ms_gemv_t * 1e-3) / 1e9}}
    return base_result(args, steps / el, "applies/s", el * 1e3 / steps, "f32" if dt == np.complex64 else "f64",
                       "synthetic (complex Gaussian operator and matrices)", config, roof, cpu)


# ---------------------------------------------------------------- batch256
def run_batch256(args):
    import qobhip as Q
    from qobhip import workloads as W
    from qobhip.core import _dt
    from qobhip.tune import EventTimer

    dt = np.complex64 if args.dtype == "c32" else np.complex128
    Ls = W.tfim_batch_lengths(256, 8, 20, seed=1)
    specs = [W.tfim_chain(int(L), 1000 + i) for i, L in enumerate(Ls)]
    ops = [s.build() for s in specs]
    ctx = Q.context()
    g = W.rng(5)
    psi0 = [W.normalized_ket(g, 2 ** int(L)) for L in Ls]
    xa = [Q.DeviceArray.from_numpy(p, dt) for p in psi0]
    xb = [Q.DeviceArray(2 ** int(L), dt) for L in Ls]
    bt = Q.Batch(ops, dt)
    tot, nst = bt.info()
    bt.bind(0, xa, xb)
    bt.bind(1, xb, xa)
    c = -0.01j
    state = {"k": 0}

    def step():  # x' = x + c H x, alternating the two buffers
        bt.apply(state["k"] & 1, c, 0.0, 1.0)
        state["k"] += 1

    if args.probe:
        big = max(range(len(Ls)), key=lambda i: Ls[i])
        return probe(ctx, xa[big], xb[big], step)
    steps = args.steps + (args.steps & 1)
    log(f"bench: batch256 {args.dtype} warmup {args.warmup}, timing {steps} steps")
    el = timed(step, steps, args.warmup + (args.warmup & 1), ctx.synchronize)
    ev = EventTimer(ctx.h, ctx.stream)
    ms_step = ev(step, 20)
    # apply and axpy as separate launches (mul!(k, H, x) ; x += c k)
    ks = [Q.DeviceArray(2 ** int(L), dt) for L in Ls]
    bt.bind(2, xa, ks)
    bt.bind(3, ks, xa)
    ms_apply = ev(lambda: bt.apply(2, 1.0, 0.0, 0.0), 10)
    ms_axpy = ev(lambda: bt.axpy(3, c), 10)
    # 32 captured steps replayed from a HIP graph
    with Q.Graph(ctx.stream) as gr:
        for s in range(32):
            bt.apply(s & 1, c, 0.0, 1.0, ctx.stream)
    gr.launch()
    ctx.synchronize()
    ms_graph32 = ev(gr.launch, 5)
    # the plain per-call ComplexF64 path: per state, one launch per term + axpy
    ctx.set_mode(Q.QOB_MODE_PLAIN)
    kets = [Q.Ket(s.basis, x) for s, x in zip(specs, xa)]
    kk = [Q.Ket(s.basis, k) for s, k in zip(specs, ks)]
    from qobhip import _lib

    def plain_step():
        for H, x, k in zip(ops, kets, kk):
            Q.mul_(k, H, x, 1.0, 0.0)
            _lib.call("qob_axpy", ctx.h, _dt(dt), x.data.n, _lib.c64(c), k.data.ptr, x.data.ptr, None)

    t_plain = None
    if not args.no_plain:
        plain_step()
        ctx.synchronize()
        t_plain = timed(plain_step, 3, 0, ctx.synchronize) / 3
    ctx.set_mode(Q.QOB_MODE_FUSED)
    byts = 2 * np.dtype(dt).itemsize * tot  # read x, write x' once per element
    achieved = byts / (ms_step * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (C recursive engine, 1 thread, whole batch)")
        n, elems, t = cpu_lazysum(list(zip(specs, psi0)), 1.0)
        na, ea, ta, nth = cpu_lazysum_threads(list(zip(specs, psi0)), 1.0)
        cpu = {"value": n / t, "unit": "state-applies/s", "cores": 1, "kind": "port",
               "sample": f"{n} applies ({elems} elements) over the 256 chains in order, C restatement of "
                         f"src/operators_lazysum.jl:189-200 + src/operators_lazytensor.jl:652-685, 1 thread, "
                         f"{t:.1f} s (axpy not included)",
               "all_cores": {"value": na / ta, "cores": nth,
                             "sample": f"{na} applies ({ea} elements), the chains split over {nth} threads, "
                                       f"{ta:.1f} s"}}
    roof = {"bound": "hbm", "kernel": "k_qubit", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(f"qubit_batch_{args.dtype}"),
            "algorithmic_bytes_per_launch": byts}
    config = {"workload": "256 transverse-field Ising chains of 8..20 sites, one fused launch per step "
                          "x' = x + c H x (c = -0.01i)",
              "states": nst, "elements": tot, "step_kernel_ms": ms_step, "apply_ms": ms_apply, "axpy_ms": ms_axpy,
              "graph32_ms": ms_graph32, "graph32_per_step_ms": ms_graph32 / 32,
              "plain_per_call_step_ms": t_plain * 1e3 if t_plain else None,
              "vs_plain_per_call": t_plain * 1e3 / ms_step if t_plain else None}
    return base_result(args, nst * steps / el, "state-applies/s", el * 1e3 / steps,
                       "f32" if dt == np.complex64 else "f64",
                       "synthetic (seeded chain lengths, J, h ~ U[-1,1], complex-Gaussian Kets)", config, roof, cpu)


# ------------------------------------------------------------------- serve
def run_serve(args):
    from qobhip import serve as S

    rates = [10, 25, 50, 100, 200, 500, 1000, 2000, 5000]
    dtypes = {"c64": [np.complex128], "c32": [np.complex64]}.get(args.dtype, [np.complex128, np.complex64])
    out = {}
    values = {}
    for dt in dtypes:
        name = "c32" if dt == np.complex64 else "c64"
        log(f"bench: serve {name}: building job classes")
        classes = S.make_classes(dt)
        srv = S.Server(classes, dt, budget_bytes=args.budget_gb << 30)
        srv.run(S.make_requests(classes, 16, rate=0, seed=99, dtype=dt))  # warm pools and plans
        per = []
        for r in rates:
            reqs = S.make_requests(classes, args.requests, rate=r, seed=r, dtype=dt)
            res = srv.run(reqs)
            s = S.summarize(res)
            s["rate"] = r
            s["peak_hbm_gb"] = srv.peak / 2 ** 30
            vals = {x.rid: x.value for x in res}
            per.append((s, vals))
            log(f"bench: serve {name} rate {r}: p50 {s.get('p50_ms', 0):.1f} ms p99 {s.get('p99_ms', 0):.1f} ms "
                f"{s.get('throughput_jobs_s', 0):.1f} jobs/s")
        out[name] = [p[0] for p in per]
        values[name] = [p[1] for p in per]
    err = None
    if "c64" in values and "c32" in values:
        errs = []
        for v64, v32 in zip(values["c64"], values["c32"]):
            for k in v64:
                if v64[k]:
                    errs.append(abs(v32[k] - v64[k]) / abs(v64[k]))
        err = {"max_rel_err_c32_vs_c64_result_scalar": max(errs) if errs else None}
    main = out.get("c64", out.get("c32"))
    value = max(s.get("throughput_jobs_s", 0) for s in main)
    config = {"workload": "Poisson request stream, jobs uniform over {10-site chain Ket, TFIM 8..20-site Ket, "
                          "Fock 6400 rho sparse, dense LazyTensor 8192 rho}; job = 32 x (mul! + axpy) + normalize!",
              "requests_per_rate": args.requests, "rates": rates, "budget_gb": args.budget_gb,
              "serving": out, "c32_error": err}
    return base_result(args, value, "jobs/s", None, "f64" if "c64" in out else "f32",
                       "synthetic (seeded arrivals, states and Hamiltonians)", config, None,
                       {"value": None, "unit": "jobs/s", "cores": 0, "kind": "port",
                        "sample": "not measured: the reference has no serving loop; per-shape CPU rates are in the "
                                  "ket10 / sparse6400 / dense8192 / batch256 lines"})


WORKLOADS = {"sparse6400": run_sparse6400, "ket10": run_ket10, "dense8192": run_dense8192,
             "batch256": run_batch256, "denseop": run_denseop, "serve": run_serve}
DEFAULT_STEPS = {"sparse6400": (200, 10), "ket10": (2000, 100), "dense8192": (100, 20), "batch256": (200, 10),
                 "denseop": (10, 2), "serve": (0, 0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="dense8192")
    ap.add_argument("--dtype", choices=["c64", "c32", "both"], default=None)
    ap.add_argument("--requests", type=int, default=48, help="serve: requests per arrival rate")
    ap.add_argument("--budget-gb", type=int, default=64, help="serve: HBM budget for admitted jobs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-call comparison runs")
    ap.add_argument("--probe", action="store_true",
                    help="fixed short launch sequence for rocprofv3 counter runs (copy, then the workload's kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the library has no multi-GPU "
              f"path: run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    ds, dw = DEFAULT_STEPS[args.workload]
    args.steps = ds if args.steps is None else args.steps
    args.warmup = dw if args.warmup is None else args.warmup
    if args.dtype is None:
        args.dtype = "both" if args.workload == "serve" else "c64"
    if args.dtype == "both" and args.workload != "serve":
        args.dtype = "c64"
    res = WORKLOADS[args.workload](args)
    if res is not None:
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
