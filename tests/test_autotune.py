"""Autotuner: the coordinate descent, the persisted file's lookup, and (GPU) a tuning run whose
chosen configuration still counts exactly."""
from __future__ import annotations

import json

import pytest

from gary_amd import autotune, graph, graph_query as gq


def test_descent_finds_separable_minimum():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": [0, 1]}
    target = {"a": 3, "b": 20, "c": 1}
    seen = []

    def measure(c):
        seen.append(dict(c))
        return 1.0 + sum(abs(c[k] - target[k]) for k in target)

    best, t, evals = autotune._descend(measure, space, {"a": 1, "b": 10, "c": 0})
    assert best == target and t == 1.0
    assert evals == len(seen)


def test_descent_skips_infeasible():
    space = {"a": [1, 2, 3]}

    def measure(c):
        return float("inf") if c["a"] == 3 else 5.0 - c["a"]

    best, _, _ = autotune._descend(measure, space, {"a": 1})
    assert best == {"a": 2}


def test_cfg_for_lookup(tmp_path):
    p = tmp_path / "tuned.json"
    p.write_text(json.dumps({"entries": {
        "w/bitmap": {"cfg": {"block": 512, "grid": 4096, "bin_small": 32, "bin_hub": 1024}},
        "w/merge": {"cfg": {"block": 128, "group": 8, "cache_depth": 16, "partition": 1}}}}))
    c = autotune.cfg_for("w", "bitmap", p)
    assert (c.block, c.grid, c.bin_small, c.bin_hub) == (512, 4096, 32, 1024)
    h = autotune.cfg_for("w", "learned", p)  # routed strategies share the pivot knobs
    assert (h.block, h.bin_small) == (512, 32)
    m = autotune.cfg_for("w", "merge", p)
    assert (m.group, m.partition, m.bin_small) == (8, 1, 0)
    assert autotune.cfg_for("w", "cached_bsearch", p) is None
    assert autotune.cfg_for("other", "bitmap", p) is None
    assert autotune.cfg_for("w", "bitmap", tmp_path / "missing.json") is None


def test_shipped_file_is_well_formed():
    if not autotune.TUNED_PATH.exists():
        pytest.skip("no tuned configuration shipped")
    doc = json.loads(autotune.TUNED_PATH.read_text())
    for key, rec in doc["entries"].items():
        wl, st = key.split("/")
        if st == "heuristic_table":  # route thresholds of the heuristic, not launch knobs
            assert set(rec["thresholds"]) == set(autotune.HEURISTIC_SPACE)
            for k, v in rec["thresholds"].items():
                assert v in autotune.HEURISTIC_SPACE[k], key
            assert autotune.heuristic_for(wl) == rec["thresholds"]
        else:
            assert st in gq.STRATEGIES
            assert set(rec["cfg"]) <= set(autotune.FIELDS)
        assert rec["ms"] <= rec["default_ms"] * 1.05, key


@pytest.mark.gpu
def test_tune_er4096_counts_exact():
    from gary_amd.device import DeviceGraphCtx
    import oracle
    g = graph.oriented(*graph.er_4096())
    ctx = DeviceGraphCtx(g)
    ctx.populateDevice()
    want = oracle.tri_count(g.indPtr, g.colIdx)
    for st in ("bitmap", "cached_bsearch"):
        rec = autotune.tune_graph(ctx, st, repeats=2)
        assert rec["count"] == want
        assert gq.triangleCountHom(ctx, gq.STRATEGIES[st], autotune.to_cfg(rec["cfg"])) == want
