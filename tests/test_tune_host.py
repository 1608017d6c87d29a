"""Autotuner search and persistence (host only: tunables are host state)."""
import json

import pytest

from qobhip import tune as T


def test_coordinate_search_separable_finds_optimum_once_per_config():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": [0, 1]}
    calls = []

    def cost(cfg):
        calls.append(tuple(sorted(cfg.items())))
        return (cfg["a"] - 3) ** 2 + (cfg["b"] - 20) ** 2 / 100 + (0.5 if cfg["c"] else 0)

    best, best_cost, trials = T.coordinate_search(space, {"a": 1, "b": 10, "c": 1}, cost, passes=3)
    assert best == {"a": 3, "b": 20, "c": 0} and best_cost == 0
    assert len(calls) == len(set(calls)) == len(trials)


def test_coordinate_search_interaction_reaches_coordinate_optimum():
    # cost minimal at (2, 2) but (1, 1) is better than any single move from (0, 0)
    table = {(0, 0): 5, (0, 1): 6, (1, 0): 6, (1, 1): 2, (2, 2): 0, (0, 2): 7, (2, 0): 7, (1, 2): 3, (2, 1): 3}
    space = {"x": [0, 1, 2], "y": [0, 1, 2]}
    best, c, _ = T.coordinate_search(space, {"x": 1, "y": 0}, lambda g: table[(g["x"], g["y"])], passes=4)
    # from (1,0): y -> 1 gives 2; then x -> ... (2,1)=3, then y->2 from (1,1)=3; stays (1,1) unless improving
    for name, vals in space.items():
        for v in vals:
            cand = dict(best)
            cand[name] = v
            assert table[(cand["x"], cand["y"])] >= c


def test_save_load_roundtrip(tmp_path):
    names = ["qubit_tbits", "qubit_nth", "mode_kernel", "fused_gpw"]
    saved = T.snapshot(names)
    try:
        res = [{"family": "qubit", "best": {"qubit_tbits": 12, "qubit_nth": 256}, "best_ms": 1.0, "trials": []},
               {"family": "mode", "best": {"mode_kernel": 1}, "best_ms": 2.0, "trials": []}]
        path = tmp_path / "tuned.json"
        doc = T.save(path, res)
        assert json.loads(path.read_text())["tunables"] == {"qubit_tbits": 12, "qubit_nth": 256, "mode_kernel": 1}
        T.apply(saved)
        T.load(path)
        assert T.get("qubit_tbits") == 12 and T.get("qubit_nth") == 256 and T.get("mode_kernel") == 1
        # a file tuned on another device model is ignored
        doc["device"] = "some other GPU"
        doc["tunables"] = {"qubit_tbits": 13}
        path.write_text(json.dumps(doc))
        assert T.load(path) is None and T.get("qubit_tbits") == 12
    finally:
        T.apply(saved)


def test_invalid_tunable_values_are_rejected():
    from qobhip import ArgumentError
    for name, bad in [("mode_kernel", 3), ("fused_gpw", 3), ("qubit_split", 2), ("qubit_tbits", 9), ("qubit_remap", 3), ("qubit_ntfar", 9), ("qubit_order", 3), ("qubit_dbuf", 2), ("qubit_pair", 2), ("sell_cpair", 2)]:
        with pytest.raises(ArgumentError):
            T.set_(name, bad)


def test_every_documented_and_searched_tunable_exists():
    """Each name the header lists and each name the tuner searches is known
    to the library: get returns a value and setting it back is accepted."""
    import re
    from pathlib import Path
    text = Path(__file__).resolve().parents[1].joinpath("include", "qob_hip.h").read_text()
    block = text[text.index("Launch-shape knobs"):text.index("The project's")]
    names = set(re.findall(r'"([a-z0-9_]+)"', block))
    names |= {n for space in T.SPACES.values() for n in space}
    assert {"fused_resident", "fused_tiles", "sell_closed", "csc_groups"} <= names
    for n in sorted(names):
        v = T.get(n)
        T.set_(n, v)
        assert T.get(n) == v


def test_min_gain_keeps_start_against_noise():
    space = {"a": [1, 2, 3]}
    costs = {1: 1.0, 2: 0.99, 3: 1.2}  # 1% better only
    best, c, _ = T.coordinate_search(space, {"a": 1}, lambda g: costs[g["a"]], min_gain=0.02)
    assert best == {"a": 1} and c == 1.0
    best, _, _ = T.coordinate_search(space, {"a": 1}, lambda g: costs[g["a"]], min_gain=0.0)
    assert best == {"a": 2}
