"""GPU parity of the 4-vertex homomorphism counts against the oracle and the golden fixtures."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gary_amd import _lib, graph, graph_query as gq
from gary_amd.device import DeviceGraphCtx

pytestmark = pytest.mark.gpu

KEYS = ("tri", "k4", "hom_k4", "hom_diamond", "hom_c4", "c4")


def _ranked_ctx(n, s, d):
    g = graph.csr_from_edges(n, np.asarray(s, np.int32), np.asarray(d, np.int32),
                             graph.SYMMETRIC_RANKED)
    return g, DeviceGraphCtx(g)


def test_hom4_golden(golden_graphs):
    for rec in golden_graphs:
        g, ctx = _ranked_ctx(rec["n"], rec["src"], rec["dst"])
        h = gq.hom4Count(ctx)
        want = {"tri": rec["triangles"], "k4": rec["k4"], "hom_k4": rec["hom_k4"],
                "hom_diamond": rec["hom_diamond"], "hom_c4": rec["hom_c4"], "c4": rec["c4"]}
        assert h == want, rec["name"]


@pytest.mark.parametrize("which", ["er", "rmat12", "rmat14"])
def test_hom4_vs_oracle(which):
    if which == "er":
        n, s, d = graph.er_4096()
    else:
        n, s, d = graph.rmat(int(which[4:]), 16, seed=99)
    g, ctx = _ranked_ctx(n, s, d)
    want = oracle.hom4(g.indPtr, g.colIdx)
    h = gq.hom4Count(ctx)
    assert h == want
    # pattern subsets give the same values for what they compute
    assert gq.hom4Count(ctx, gq.HOM_C4)["hom_c4"] == want["hom_c4"]
    assert gq.rectCountHom(ctx) == want["hom_c4"]
    assert gq.hom4Count(ctx, gq.HOM_K4)["k4"] == want["k4"]


@pytest.mark.parametrize("scale", [20, 22])
def test_hom4_full_size_fixture(scale):
    """R-MAT scales 20 and 22 (edgefactor 16, seeded): all six values equal the oracle's
    (tests/golden/hom4_full.json, written by make_hom4_full.py), and the triangle count equals
    the triangle-counting path's on the degree-oriented graph."""
    import json
    from conftest import GOLDEN
    doc = json.loads((GOLDEN / "hom4_full.json").read_text())
    rec = doc.get(f"rmat{scale}")
    if rec is None:
        pytest.skip(f"no fixture for rmat{scale}")
    n, s, d = graph.rmat(scale, 16, seed=rec["seed"])
    g, ctx = _ranked_ctx(n, s, d)
    assert g.nnz == rec["nnz"]
    h = gq.hom4Count(ctx)
    assert h == {k: rec[k] for k in KEYS}
    del ctx
    go = graph.oriented(n, s, d)
    octx = DeviceGraphCtx(go)
    assert gq.triangleCountHom(octx, gq.BITMAP) == rec["tri"]


def test_hom4_rejects_unranked_or_oriented():
    n, s, d = graph.rmat(10, 8, seed=5)
    sym = graph.csr_from_edges(n, s, d, graph.SYMMETRIC)  # not ranked
    ctx = DeviceGraphCtx(sym)
    with pytest.raises(_lib.GaryError) as ei:
        gq.hom4Count(ctx)
    assert ei.value.code == _lib.GARY_ERR_INVALID
    g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
    # keep only the upper parts (an oriented graph with ranked ids): not symmetric
    up_ptr = [0]
    up_col = []
    for u in range(g.m):
        row = g.N(u)
        up = row[row > u]
        up_col.extend(up.tolist())
        up_ptr.append(len(up_col))
    og = graph.GraphCSR(np.array(up_ptr), np.array(up_col, np.int32))
    if np.all(np.diff(og.deg) >= 0):  # only meaningful if the degree order still holds
        with pytest.raises(_lib.GaryError):
            gq.hom4Count(DeviceGraphCtx(og))


def _light_paths_graph():
    """u with 255 neighbours v_i and 300 hub neighbours (a 300-clique, each
    hub with 600 leaves); every v_i is joined to the same 100 vertices w_j. Ranked by degree, the
    w_j and u are light (fewer than 256 lower neighbours) with thousands to tens of thousands of
    wedges each, and the hubs are heavy."""
    u, V, Wn, H = 0, np.arange(1, 256), np.arange(256, 356), np.arange(356, 656)
    leaves = 656 + np.arange(300 * 600)
    src, dst = [], []
    src += [np.full(255, u)]; dst += [V]
    vv, ww = np.meshgrid(V, Wn, indexing="ij")
    src += [vv.ravel()]; dst += [ww.ravel()]
    src += [np.full(300, u)]; dst += [H]
    hi, hj = np.triu_indices(300, 1)
    src += [H[hi]]; dst += [H[hj]]
    src += [np.repeat(H, 600)]; dst += [leaves]
    n = 656 + 300 * 600
    return n, np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32)


def test_hom4_light_vertex_paths():
    """Every light-vertex 4-cycle path (wave hash, workgroup hash, HBM counters) and the window
    path on one graph, against the oracle."""
    n, s, d = _light_paths_graph()
    g, ctx = _ranked_ctx(n, s, d)
    ip, ci = g.indPtr, g.colIdx
    rows = np.repeat(np.arange(g.m), np.diff(ip))
    lo = np.bincount(rows[ci < rows], minlength=g.m)

    def wedges(x):  # Σ over lower neighbours v of |N(v) ∩ [0, x)|
        return sum(int(np.searchsorted(ci[ip[v]:ip[v + 1]], x)) for v in ci[ip[x]:ip[x + 1]]
                   if v < x)

    w = np.array([wedges(x) for x in np.where((lo >= 2) & (lo < 256))[0]])
    # wave hash (<= 512 wedges), workgroup hash (<= 16,384), HBM counters (more); heavy windows
    assert (w <= 512).sum() > 0 and ((w > 512) & (w <= 16384)).sum() > 0 and (w > 16384).sum() > 0
    assert (lo >= 256).sum() > 0
    assert g.nnz == 2 * (255 + 25500 + 300 + 300 * 299 // 2 + 300 * 600)
    want = oracle.hom4(g.indPtr, g.colIdx)
    assert want["c4"] > 0
    assert gq.hom4Count(ctx) == want
    assert gq.hom4Count(ctx, gq.HOM_C4)["c4"] == want["c4"]


def test_hom4_window_halves_complete_bipartite():
    """K_{3,70000}: each hub has 70,000 lower neighbours, more than 16-bit counters hold, so its
    windows are counted in two halves with 32-bit counters. #C4 = C(3,2) * C(70000,2) (closed
    form), no triangles, and the oracle's wedge counts agree."""
    a, b = 3, 70_000
    hubs = np.arange(a)
    leaves = a + np.arange(b)
    hh, ll = np.meshgrid(hubs, leaves, indexing="ij")
    g, ctx = _ranked_ctx(a + b, hh.ravel().astype(np.int32), ll.ravel().astype(np.int32))
    got = gq.hom4Count(ctx)
    assert got["c4"] == (a * (a - 1) // 2) * (b * (b - 1) // 2)
    assert got["tri"] == 0 and got["k4"] == 0 and got["hom_diamond"] == 0
    assert got == oracle.hom4_ranked(g.indPtr, g.colIdx)
