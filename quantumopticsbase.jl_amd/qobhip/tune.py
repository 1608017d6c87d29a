"""Autotuner for the library's launch shapes.

No reference counterpart.  The kernels' launch shapes are runtime tunables
of libqob_hip.so (qob_tune_set / qob_tune_get):

  mode products     fused_gpw (fibre groups in flight per wave in the
                    fused MFMA kernel; 0 = by precision and factor width),
                    fused_resident (factors kept in LDS for the launch);
                    mode_kernel (1 = tiled GEMM for every dense mode product, 2 = fibre-group
                    kernel for leading dimensions below 16 as well,
                    a test and timing switch, not tuned)
  qubit sums        qubit_tbits (tile = 2^tbits states; 0 = 2^12 for
                    ComplexF64, 2^11 for ComplexF32), qubit_nth (0 = 4
                    elements per lane in ComplexF64, 8 in ComplexF32),
                    qubit_grid (persistent workgroups; 0 = 1024 / nth per CU),
                    qubit_remap (0 plain, 1 XCD-contiguous rounds, 2 per-XCD
                    item ranges), qubit_ntfar (non-temporal loads of partner
                    tiles outside the run in flight), qubit_order (tile walk
                    order of batches created after it is set: 0 index, 1 Gray
                    code, 2 Gray code above 64 tiles), qubit_dbuf (two LDS tile
                    buffers in persistent walks), qubit_pair (ComplexF32
                    lanes on pairs of adjacent elements), qubit_split
  sparse products   sell_variant, sell_cj, sell_jchunk, sell_unr, csc_rows,
                    csc_rb, csc_cc, csc_remap, csc_order (column sweep order),
                    csc_sweep (row stripes / column-major / by reuse
                    distance), csc_waves, csc_gkernel / csc_gremap
                    (component-group right kernel, its XCD remap),
                    sell_cjchunk / sell_cremap (closed-block left kernel),
                    nt_store; sell_align (diagonal-aligned slots),
                    sell_closed (closed blocks), sell_cline (closed blocks
                    of whole 128-byte lines, 320 rows), sell_cpair
                    (ComplexF32 closed blocks of row pairs) and csc_groups
                    (component groups) apply to tables built after they are
                    set

``tune`` runs a coordinate search: starting from the current values it
tries every value of one tunable with the others fixed, keeps the fastest,
and sweeps the tunables again until a full sweep changes nothing (or
``passes`` sweeps).  Each trial is timed with events on the stream the work
runs on.  The result is written as JSON (device name, library version,
chosen values, timings) and ``apply`` / ``load`` set the values again in a
later process.  ``context()`` applies the file named by $QOB_TUNE_FILE, or
the tuned.json shipped next to the package when it was tuned on the same
device model.  Run ``python -m qobhip.tune`` to tune on the benchmark shapes.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import Callable, Dict, Iterable, List, Optional, Sequence

from . import _lib
from ._lib import call

SPACES: Dict[str, Dict[str, List[int]]] = {
    "fused": {"fused_gpw": [0, 1, 2, 4], "fused_resident": [1, 0]},
    "qubit": {"qubit_tbits": [0, 10, 11, 12, 13], "qubit_nth": [0, 256, 512, 1024], "qubit_grid": [0, 256, 512, 1024],
              "qubit_remap": [2, 1, 0], "qubit_ntfar": [1, 0], "qubit_order": [1, 0, 2], "qubit_dbuf": [1, 0], "qubit_pair": [1, 0], "qubit_split": [0, 1]},
    "sparse_left": {"sell_variant": [0, 1], "sell_cj": [1, 2, 4, 8], "sell_jchunk": [4, 8, 16, 32],
                    "sell_unr": [2, 4, 8], "sell_cjchunk": [4, 2, 8], "sell_cremap": [1, 0]},
    "sparse_right": {"csc_rows": [16, 32, 64], "csc_rb": [64, 128, 256], "csc_cc": [1, 2, 4, 8], "csc_remap": [0, 1],
                     "csc_order": [1, 0], "csc_sweep": [2, 0, 1], "csc_waves": [1, 2, 4], "csc_gkernel": [1, 0],
                     "csc_gremap": [0, 1]},
}


def get(name: str) -> int:
    v = C.c_int64()
    call("qob_tune_get", name.encode(), C.byref(v))
    return int(v.value)


def set_(name: str, value: int) -> None:
    call("qob_tune_set", name.encode(), int(value))


def apply(values: Dict[str, int]) -> None:
    for k, v in values.items():
        set_(k, v)


def snapshot(names: Iterable[str]) -> Dict[str, int]:
    return {n: get(n) for n in names}


class EventTimer:
    """Average milliseconds per call of fn(), by events on ``stream``."""

    def __init__(self, ctx_h, stream):
        self.ctx_h, self.stream = ctx_h, stream
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        call("qob_event_create", ctx_h, C.byref(self.e0))
        call("qob_event_create", ctx_h, C.byref(self.e1))

    def __call__(self, fn: Callable[[], None], reps: int) -> float:
        call("qob_event_record", self.ctx_h, self.e0, self.stream)
        for _ in range(reps):
            fn()
        call("qob_event_record", self.ctx_h, self.e1, self.stream)
        call("qob_event_synchronize", self.e1)
        ms = C.c_float()
        call("qob_event_elapsed", self.e0, self.e1, C.byref(ms))
        return ms.value / reps

    def __del__(self):
        try:
            lib = _lib.load()
            lib.qob_event_destroy(self.e0)
            lib.qob_event_destroy(self.e1)
        except Exception:
            pass


def coordinate_search(space: Dict[str, Sequence[int]], start: Dict[str, int], cost: Callable[[Dict[str, int]], float],
                      passes: int = 2, min_gain: float = 0.0) -> tuple:
    """Minimise cost(config) one coordinate at a time; returns (best, best_cost, trials).

    A move is taken only when it lowers the cost by more than the fraction
    ``min_gain`` (timing noise between runs and devices must not flip a
    setting).  Every distinct configuration is evaluated once (cached)."""
    seen: Dict[tuple, float] = {}
    trials = []

    def ev(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = float(cost(dict(cfg)))
            trials.append((dict(cfg), seen[key]))
        return seen[key]

    best = dict(start)
    best_cost = ev(best)
    for _ in range(max(1, passes)):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                c = ev(cand)
                if c < best_cost * (1.0 - min_gain):
                    best, best_cost, changed = cand, c, True
        if not changed:
            break
    return best, best_cost, trials


def tune(family: str, make_run: Callable[[], Callable[[], None]], stream=None, reps: int = 10, warmup: int = 2,
         passes: int = 2, space: Optional[Dict[str, Sequence[int]]] = None, repeats: int = 3,
         min_gain: float = 0.02) -> dict:
    """Tune one family of tunables on a workload.

    ``make_run()`` builds the workload under the current tunables (it is
    called again for each trial, since plans and batches depend on them) and
    returns a callable that enqueues one unit of work on ``stream``.  A
    trial's cost is the fastest of ``repeats`` timings of ``reps`` calls; a
    setting changes only for a gain above ``min_gain``.  The tunables are
    left at the best values found."""
    from .core import context
    space = dict(space or SPACES[family])
    ctx = context()
    timer = EventTimer(ctx.h, stream if stream is not None else ctx.stream)
    start = snapshot(space)

    def cost(cfg):
        apply(cfg)
        run = make_run()
        for _ in range(warmup):
            run()
        return min(timer(run, reps) for _ in range(max(1, repeats)))

    try:
        best, best_ms, trials = coordinate_search(space, start, cost, passes, min_gain)
    except Exception:
        apply(start)
        raise
    apply(best)
    return {"family": family, "best": best, "best_ms": best_ms, "start": start,
            "start_ms": next(t for c, t in trials if c == start),
            "trials": [{"config": c, "ms": t} for c, t in trials]}


def device_name() -> str:
    try:
        hip = C.CDLL("libamdhip64.so.7")
    except OSError:
        return "unknown"
    dev = C.c_int()
    if hip.hipGetDevice(C.byref(dev)) != 0:
        return "unknown"
    hip.hipDeviceGetName.argtypes = [C.c_char_p, C.c_int, C.c_int]
    name = C.create_string_buffer(256)
    if hip.hipDeviceGetName(name, 256, dev.value) != 0:
        return "unknown"
    return name.value.decode(errors="replace")


def save(path, results: Sequence[dict]) -> dict:
    lib = _lib.load()
    lib.qob_version.restype = C.c_char_p
    values: Dict[str, int] = {}
    for r in results:
        values.update(r["best"])
    doc = {"device": device_name(), "library": lib.qob_version().decode(), "tunables": values,
           "results": list(results)}
    Path(path).write_text(json.dumps(doc, indent=1))
    return doc


def load(path=None, strict_device: bool = True) -> Optional[dict]:
    """Apply the tunables stored at ``path`` (default $QOB_TUNE_FILE, else
    the tuned.json shipped next to the package).

    With strict_device, a file tuned on another device model is ignored."""
    path = path or os.environ.get("QOB_TUNE_FILE") or str(PACKAGED)
    if not path or not Path(path).exists():
        return None
    doc = json.loads(Path(path).read_text())
    if strict_device and doc.get("device") not in (None, "unknown", device_name()):
        return None
    apply(doc.get("tunables", {}))
    return doc


PACKAGED = Path(__file__).resolve().parent.parent / "tuned.json"


def _workloads(families):
    """Standard tuning workloads (the benchmark shapes) per family."""
    from . import core as Q
    from . import workloads as W
    from .batch import Batch
    runs = {}
    if "fused" in families:
        N = 8192
        LT = W.dense_lazytensor((16, 32, 16), 1)
        b = Q.GenericBasis(N)
        rho_d = Q.Operator(b, b, Q.DeviceArray(N * N))
        # any contents: only the launch shapes are timed
        call("qob_scale", Q.context().h, Q.QOB_C64, N * N, _lib.c64(0.0), rho_d.data.ptr, None)
        out_d = Q.Operator(b, b, Q.DeviceArray(N * N))
        # the left apply runs the fused MFMA kernel
        runs["fused"] = lambda: (lambda: Q.mul_(out_d, LT, rho_d, 1.0, 0.0))
    if "qubit" in families:
        Ls = W.tfim_batch_lengths(256, 8, 20, seed=1)
        ops = [W.tfim_chain(int(L), 1000 + i).build() for i, L in enumerate(Ls)]
        xa = [Q.DeviceArray(2 ** int(L)) for L in Ls]
        xb = [Q.DeviceArray(2 ** int(L)) for L in Ls]

        def make_q():
            bt = Batch(ops)
            bt.bind(0, xa, xb)
            return lambda: bt.apply(0, -0.01j, 0.0, 1.0)
        runs["qubit"] = make_q
    if "sparse_left" in families or "sparse_right" in families:
        spec = W.fock_fock_spin(seed=1)
        H = Q.SparseOperator(spec.basis, spec.basis, spec.H)
        n = spec.H.shape[0]
        b = spec.basis
        rho_s = Q.Operator(b, b, Q.DeviceArray(n * n))
        call("qob_scale", Q.context().h, Q.QOB_C64, n * n, _lib.c64(0.0), rho_s.data.ptr, None)
        out_s = Q.Operator(b, b, Q.DeviceArray(n * n))
        runs["sparse_left"] = lambda: (lambda: Q.mul_(out_s, H, rho_s, 1.0, 0.0))
        runs["sparse_right"] = lambda: (lambda: Q.mul_(out_s, rho_s, H, 1.0, 0.0))
    return runs


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="tune libqob_hip launch shapes on the benchmark workloads")
    ap.add_argument("--families", default="fused,qubit,sparse_left,sparse_right")
    ap.add_argument("--out", default=str(PACKAGED))
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    fams = [f for f in args.families.split(",") if f]
    runs = _workloads(fams)
    results = []
    for f in fams:
        r = tune(f, runs[f], reps=args.reps)
        print(json.dumps({"family": f, "best": r["best"], "best_ms": r["best_ms"], "start": r["start"],
                          "start_ms": r["start_ms"], "trials": len(r["trials"])}), flush=True)
        results.append(r)
    save(args.out, results)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
