"""4-cycle vertex routing (gary_hom4_count_routed): the fixed table and argument checks on the
CPU; on the GPU, counts under every route table against the oracle, the per-bin vertex counts
against a host binning, and the learned table's choice rule on models with constant outputs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from gary_amd import c4_selector as cs, graph, graph_query as gq
from gary_amd._lib import DeviceGraph, GaryError, check, lib, ptr


def test_bin_of_closed_forms():
    assert cs.bin_of(2, 1) == 16
    assert cs.bin_of(3, 3) == 16           # log2 3 -> 1; log2 3 / 2 -> 0
    assert cs.bin_of(255, 1 << 20) == 7 * 16 + 10
    assert cs.bin_of(256, (1 << 31) - 1) == 8 * 16 + 15
    assert cs.bin_of(1 << 20, 1 << 40) == 15 * 16 + 15


def test_fixed_table_matches_library():
    """The library's FIXED table (empty graph: no device work) is the degree split."""
    t = np.zeros(256, np.uint8)
    g = DeviceGraph()
    assert lib.gary_hom4_c4_table(C.byref(g), gq.C4_FIXED, ptr(t)) == 0
    np.testing.assert_array_equal(t, cs.fixed_table())
    assert set(np.nonzero(t == cs.WINDOW)[0]) == set(range(128, 256))


def test_routed_arguments_checked():
    g = DeviceGraph()
    bad = np.full(256, 4, np.uint8)
    with pytest.raises(GaryError):
        gq.hom4Count(g, gq.HOM_C4, c4=bad)
    with pytest.raises(GaryError):
        gq.hom4Count(g, gq.HOM_C4, c4=7)
    with pytest.raises(ValueError):
        gq.hom4Count(g, gq.HOM_C4, c4=np.zeros(255, np.uint8))
    with pytest.raises(GaryError):
        check(lib.gary_hom4_c4_load_model(b"/nonexistent/c4_model.json"))
    with pytest.raises(FileNotFoundError):
        cs.load("/nonexistent/c4_model.json")
    assert cs.load("/nonexistent/c4_model.json", required=False) is None
    # an empty graph counts nothing under any table
    assert gq.hom4Count(g, gq.HOM_C4, c4=cs.fixed_table())["c4"] == 0


def test_shipped_model_file_is_well_formed():
    import json
    p = cs.MODEL_PATH
    if not p.exists():
        pytest.skip("no trained 4-cycle model shipped")
    doc = json.loads(p.read_text())
    assert len(doc["params"]) == cs.NPARAMS and all(np.isfinite(doc["params"]))
    assert 0 <= doc["margin"] < 1
    assert doc["routes"] == ["light", "window", "dense"]


# ---- GPU ---------------------------------------------------------------------------------------

def _ranked(n, s, d):
    from gary_amd.device import DeviceGraphCtx
    g = graph.csr_from_edges(n, np.asarray(s, np.int32), np.asarray(d, np.int32),
                             graph.SYMMETRIC_RANKED)
    return g, DeviceGraphCtx(g)


def _host_bins(g):
    """Per-vertex bins from the host CSR (nl < 2: -1). S is the sum of d(v) over the lower
    neighbours for nl <= 64, else nl * (Σ over the 64 neighbours at positions i * nl // 64) // 64."""
    ip, ci = g.indPtr, g.colIdx
    deg = np.diff(ip).astype(np.int64)
    rows = np.repeat(np.arange(g.m), deg)
    low = ci < rows
    nl = np.bincount(rows[low], minlength=g.m)
    S = np.bincount(rows[low], weights=deg[ci[low]], minlength=g.m).astype(np.int64)
    for u in np.nonzero(nl > 64)[0]:
        k = (np.arange(64) * int(nl[u])) // 64
        S[u] = int(deg[ci[ip[u] + k]].sum()) * int(nl[u]) // 64
    b = np.full(g.m, -1, np.int64)
    for u in np.nonzero(nl >= 2)[0]:
        b[u] = cs.bin_of(int(nl[u]), int(S[u]))
    return b, nl, S


def _constant_model(y):
    """Parameters whose outputs are y for every input (zero weights, b3 = y)."""
    p = np.zeros(cs.NPARAMS, np.float32)
    p[-3:] = y
    return p


@pytest.fixture
def light_paths_ctx():
    from test_gpu_hom4 import _light_paths_graph
    n, s, d = _light_paths_graph()
    g, ctx = _ranked(n, s, d)
    want = oracle.hom4(g.indPtr, g.colIdx)
    return g, ctx, want


@pytest.mark.gpu
def test_counts_equal_under_every_table(light_paths_ctx):
    """The graph with every light path (wave hash, workgroup hash, dense overflow) and heavy
    vertices: each single-route table, the fixed table and random tables count the oracle's
    4-cycles; all-skip counts none."""
    g, ctx, want = light_paths_ctx
    rng = np.random.default_rng(7)
    tables = [np.full(256, r, np.uint8) for r in (cs.LIGHT, cs.WINDOW, cs.DENSE)]
    tables += [cs.fixed_table()] + [rng.integers(0, 3, 256).astype(np.uint8) for _ in range(6)]
    for t in tables:
        got = gq.hom4Count(ctx, gq.HOM_ALL, c4=t)
        assert got == want, t
    assert gq.hom4Count(ctx, gq.HOM_C4, c4=gq.C4_FIXED)["c4"] == want["c4"]
    none = gq.hom4Count(ctx, gq.HOM_C4, c4=np.full(256, cs.SKIP, np.uint8))
    assert none["c4"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["er", "rmat14"])
def test_skip_table_parts_add_up(which):
    """Skipping bins removes exactly their vertices' 4-cycles: the counts of a table and of its
    complement (the skipped bins routed, the rest skipped) add up to the full count."""
    n, s, d = graph.er_4096() if which == "er" else graph.rmat(14, 16, seed=99)
    g, ctx = _ranked(n, s, d)
    want = oracle.hom4(g.indPtr, g.colIdx)["c4"]
    rng = np.random.default_rng(11)
    for _ in range(4):
        keep = rng.random(256) < 0.5
        a = np.where(keep, rng.integers(0, 3, 256), cs.SKIP).astype(np.uint8)
        b = np.where(keep, cs.SKIP, rng.integers(0, 3, 256)).astype(np.uint8)
        ca = gq.hom4Count(ctx, gq.HOM_C4, c4=a)["c4"]
        cb = gq.hom4Count(ctx, gq.HOM_C4, c4=b)["c4"]
        assert ca + cb == want


@pytest.mark.gpu
def test_features_bin_counts_match_host():
    """Vertices per bin equal a host binning by (log2 nl, log2 S / 2), S sampled above 64 lower
    neighbours as the kernel does, on a graph with such vertices; the mean-log features
    agree with the host's means within the fixed-point rounding (1/256 per vertex)."""
    n, s, d = graph.rmat(14, 16, seed=99)
    g, ctx = _ranked(n, s, d)
    X, nv = cs.features(ctx)
    b, nl, S = _host_bins(g)
    assert (nl > 64).sum() > 100
    want = np.bincount(b[b >= 0], minlength=256)
    np.testing.assert_array_equal(nv.astype(np.int64), want)
    for k in np.nonzero(want)[0]:
        sel = b == k
        assert abs(X[k, 0] * 8 - np.log2(nl[sel]).mean()) < 2.5e-3
        assert abs(X[k, 1] * 16 - np.log2(S[sel].astype(np.float64)).mean()) < 2.5e-3
        assert abs(X[k, 3] * 16 - np.log2(g.m)) < 1e-5
    assert np.all(X[want == 0] == 0)


@pytest.mark.gpu
def test_learned_table_choice_rule():
    """With constant predicted costs the learned table is the cheapest feasible route: light only
    below 256 lower neighbours, dense only in S bins <= 11, empty bins fixed; the margin keeps
    the fixed route unless the predicted saving exceeds it. Counts stay exact."""
    n, s, d = graph.rmat(14, 16, seed=99)
    g, ctx = _ranked(n, s, d)
    want = oracle.hom4(g.indPtr, g.colIdx)
    _, nv = cs.features(ctx)
    full = nv > 0
    fixed = cs.fixed_table()
    lok, dok = cs.light_ok(), cs.dense_ok()
    try:
        cs.set_model(_constant_model([5.0, 1.0, 3.0]))  # window cheapest
        t = cs.table(ctx, gq.C4_LEARNED)
        np.testing.assert_array_equal(t[full], cs.WINDOW)
        np.testing.assert_array_equal(t[~full], fixed[~full])
        cs.set_model(_constant_model([1.0, 5.0, 3.0]))  # light, else dense, else window
        t = cs.table(ctx, gq.C4_LEARNED)
        exp = np.where(lok, cs.LIGHT, np.where(dok, cs.DENSE, cs.WINDOW))
        np.testing.assert_array_equal(t[full], exp[full])
        assert gq.hom4Count(ctx, gq.HOM_ALL, c4=gq.C4_LEARNED) == want
        assert gq.hom4Count(ctx, gq.HOM_ALL) == want  # AUTO follows the model
        cs.set_model(_constant_model([1.1, 1.0, 5.0]), margin=0.5)  # 7% saving < margin
        np.testing.assert_array_equal(cs.table(ctx, gq.C4_LEARNED), fixed)
        cs.set_model(_constant_model([1.1, 1.0, 5.0]), margin=0.0)
        t = cs.table(ctx, gq.C4_LEARNED)
        np.testing.assert_array_equal(t[full], cs.WINDOW)
        p, mg = cs.model()
        np.testing.assert_array_equal(p, _constant_model([1.1, 1.0, 5.0]))
        assert mg == 0.0
    finally:
        cs.clear_model()
    assert cs.model() is None
    np.testing.assert_array_equal(cs.table(ctx, gq.C4_AUTO), fixed)
    with pytest.raises(GaryError):
        gq.hom4Count(ctx, gq.HOM_C4, c4=gq.C4_LEARNED)


@pytest.mark.gpu
def test_shipped_model_counts_exact():
    """The shipped model loads, routes AUTO, and leaves the R-MAT 16 counts equal to the oracle."""
    if not cs.MODEL_PATH.exists():
        pytest.skip("no trained 4-cycle model shipped")
    n, s, d = graph.rmat(16, 16, seed=99)
    g, ctx = _ranked(n, s, d)
    want = oracle.hom4_ranked(g.indPtr, g.colIdx)
    try:
        cs.load()
        assert cs.model() is not None
        assert gq.hom4Count(ctx) == want
        t = cs.table(ctx, gq.C4_LEARNED)
        assert np.all(cs.light_ok()[t == cs.LIGHT]) and np.all(cs.dense_ok()[t == cs.DENSE])
    finally:
        cs.clear_model()
