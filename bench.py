#!/usr/bin/env python3
"""Benchmark: SparseOperator mul! on a density matrix, from the left and
the right, ComplexF64, on one MI355X.

Workload (the ComplexF64 density-matrix configuration):
  H = a^dag a (x) 1 (x) 1 + 1 (x) b^dag b (x) 1 + g (a (x) b^dag (x) s+ + h.c.)
  on Fock(39) (x) Fock(39) (x) Spin(3/2): N = 6400, nnz = 15522, g ~ U[0,1];
  rho = G G^dag / tr, G complex Gaussian 6400 x 6400 (seeded);
  one step = mul!(out, H, rho, 1, 0) + mul!(out2, rho, H, 1, 0) = 2 applies.

Prints one JSON line: applies/s (value), the dominant kernel's achieved HBM
GB/s against the 8 TB/s peak (timed live with HIP events on the stream the
kernels run on), and the C restatement of the reference loops timed
single-threaded on the host as the CPU baseline.
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
METRIC = "applies/sec + HBM GB/s (% of roofline) vs plain FP64 GPU; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an explicit stream (the library's context stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        for fn in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime",
                   "hipEventDestroy"):
            getattr(self.hip, fn).restype = C.c_int

    def time(self, stream, fn, reps: int) -> float:
        """Average ms per call of fn() over reps calls, measured on stream."""
        e0, e1 = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e0)) == 0 and self.hip.hipEventCreate(C.byref(e1)) == 0
        assert self.hip.hipEventRecord(e0, stream) == 0
        for _ in range(reps):
            fn()
        assert self.hip.hipEventRecord(e1, stream) == 0
        assert self.hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        self.hip.hipEventDestroy(e0)
        self.hip.hipEventDestroy(e1)
        return ms.value / reps


def cpu_baseline(spec, rho_h, budget_s: float = 12.0):
    """Reference loops restated in C (oracle/cpu_ref.c), one thread."""
    lib = C.CDLL(str(ROOT / "oracle" / "lib" / "libqob_cpu_ref.so"))
    H = spec.H
    n = H.shape[0]
    colptr = (H.indptr.astype(np.int64) + 1)
    rowval = (H.indices.astype(np.int64) + 1)
    nz = np.ascontiguousarray(H.data.astype(np.complex128))
    B = np.asfortranarray(rho_h)
    out = np.empty((n, n), dtype=np.complex128, order="F")
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.qobref_gemm_sp_dense.argtypes = [C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_void_p]
    lib.qobref_gemm_dense_sp.argtypes = [C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
    applies = 0
    t0 = time.perf_counter()
    while True:
        lib.qobref_gemm_sp_dense(1.0, 0.0, n, n, P(colptr), P(rowval), P(nz), P(B), n, 0.0, 0.0, P(out))
        lib.qobref_gemm_dense_sp(1.0, 0.0, n, n, n, P(colptr), P(rowval), P(nz), P(B), 0.0, 0.0, P(out))
        applies += 2
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return {"value": applies / el, "unit": "applies/s", "cores": 1, "kind": "port",
            "sample": f"{applies} full applies ({applies // 2} left + {applies // 2} right) of the same "
                      f"6400x6400 workload, C restatement of src/sparsematrix.jl:25-47,115-146, 1 thread, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--probe", action="store_true",
                    help="only launch copy/left/right 5 times each (for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU "
              f"path: run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import qobhip as Q
    from qobhip import workloads as W

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
        return

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        left()
        right()
    ctx.synchronize()

    log(f"bench: timing {args.steps} steps")
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        left()
        right()
    ctx.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    applies_per_s = 2 * args.steps / el

    # Dominant-kernel roofline: average kernel duration from HIP events on
    # the library's stream, algorithmic bytes per launch (beta = 0):
    # read rho (16 N^2) + write out (16 N^2) + CSR/CSC (nnz*(16+4) + (N+1)*4).
    ev = HipEvents()
    reps = max(20, min(200, args.steps))
    ms_left = ev.time(stream, left, reps)
    ms_right = ev.time(stream, right, reps)
    bytes_per_launch = 32 * n * n + nnz * 20 + (n + 1) * 4
    dom = "left" if ms_left >= ms_right else "right"
    ms_dom = max(ms_left, ms_right)
    achieved = bytes_per_launch / (ms_dom * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get(f"csr_{dom}")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (C restatement, 1 thread)")
        cpu = cpu_baseline(spec, rho.to_numpy())

    res = {
        "metric": METRIC,
        "value": applies_per_s,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fock-Fock-spin Hamiltonian, Ginibre trace-1 rho)",
        "config": {
            "workload": "SparseOperator mul! left + right on a 6400x6400 ComplexF64 density matrix "
                        "(Fock(39) x Fock(39) x Spin(3/2), nnz 15522); 1 step = 2 applies",
            "N": n,
            "nnz": nnz,
            "element": "ComplexF64",
            "kernel_ms": {"left": ms_left, "right": ms_right},
            "hbm_gbs": {"left": bytes_per_launch / (ms_left * 1e-3) / 1e9,
                        "right": bytes_per_launch / (ms_right * 1e-3) / 1e9},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_sell_left" if dom == "left" else "k_csc_right",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
