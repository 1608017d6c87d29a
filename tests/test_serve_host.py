"""Host-side logic of the serving loop (no GPU): the seeded Poisson request
stream and the latency summary."""
import numpy as np

from qobhip import serve as S


def _classes():
    return [S.KetClass("a", {}, [3, 4]), S.KetClass("b", {}, [5]),
            S.RhoClass("r", None, 8, None)]


def test_requests_are_seeded_poisson():
    a = S.make_requests(_classes(), 4000, rate=500.0, seed=1)
    b = S.make_requests(_classes(), 4000, rate=500.0, seed=1)
    assert [(r.t_arrival, r.cls, r.L) for r in a] == [(r.t_arrival, r.cls, r.L) for r in b]
    t = np.array([r.t_arrival for r in a])
    gaps = np.diff(t)
    assert t[0] == 0 and (gaps >= 0).all()
    # exponential gaps: mean 1/rate, coefficient of variation 1
    assert abs(gaps.mean() * 500.0 - 1) < 0.06
    assert abs(gaps.std() / gaps.mean() - 1) < 0.08
    counts = {n: sum(r.cls == n for r in a) for n in ("a", "b", "r")}
    assert all(abs(c / 4000 - 1 / 3) < 0.04 for c in counts.values())
    for r in a:
        if r.cls == "a":
            assert r.L in (3, 4) and abs(np.linalg.norm(r.payload) - 1) < 1e-12
        if r.cls == "r":
            assert r.payload is None
    # rate 0: everything arrives at once
    assert all(r.t_arrival == 0 for r in S.make_requests(_classes(), 10, rate=0, seed=2))


def test_summary_percentiles():
    res = [S.JobResult(i, "a" if i % 2 else "b", "done", t_arrival=0.0, t_done=(i + 1) * 1e-3) for i in range(100)]
    res.append(S.JobResult(100, "a", "rejected", 0.0))
    s = S.summarize(res)
    assert s["jobs"] == 101 and s["done"] == 100 and s["rejected"] == 1
    assert abs(s["p50_ms"] - np.percentile(np.arange(1, 101), 50)) < 1e-9
    assert abs(s["p99_ms"] - np.percentile(np.arange(1, 101), 99)) < 1e-9
    assert abs(s["throughput_jobs_s"] - 100 / 0.1) < 1e-6
    assert set(s["per_class"]) == {"a", "b"}
