"""k-clique counts on the GPU (gary_kclique_count) against the oracle's or_kclique_count (pinned
to networkx clique enumeration, closed forms and brute force) and, for k = 3, 4, against the
triangle and 4-vertex kernels. No reference counterpart: the reference counts triangles only."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from gary_amd import _lib, graph
from gary_amd import graph_query as gq
from gary_amd.device import DeviceGraphCtx

pytestmark = pytest.mark.gpu


def _ranked_ctx(n, s, d):
    return DeviceGraphCtx(graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED))


def test_kclique_fixture_all_k():
    for rec in json.loads((GOLDEN / "kcliques.json").read_text()):
        n = rec["n"]
        s, d = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        ctx = _ranked_ctx(n, s, d)
        got = [gq.kcliqueCount(ctx, k) for k in rec["k"]]
        assert got == rec["count"], rec["name"]
        g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC)
        assert gq.kcliqueCount(ctx, 1) == n and gq.kcliqueCount(ctx, 2) == g.nnz // 2


def test_kclique_complete_graphs():
    from math import comb
    # the kernel enumerates every (k-1)-clique, so k is kept where C(n, k-1) is at most ~1e6;
    # 70 > 64: local rows span two 64-bit words
    for n, kmax in ((2, 10), (7, 10), (20, 8), (70, 5)):
        s, d = np.triu_indices(n, 1)
        ctx = _ranked_ctx(n, s.astype(np.int32), d.astype(np.int32))
        for k in range(3, kmax + 1):
            assert gq.kcliqueCount(ctx, k) == comb(n, k), (n, k)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7])
def test_kclique_seeded_equals_oracle(k):
    for n, s, d in (graph.er_4096(), graph.rmat(12, 16, seed=21),
                    (3000,) + graph.gen_chung_lu(3000, 60000, seed=4, max_w=600.0)):
        ctx = _ranked_ctx(n, s, d)
        o = graph.oriented(n, s, d)
        assert gq.kcliqueCount(ctx, k) == oracle.kclique_count(o.indPtr, o.colIdx, k), (n, k)


def test_kclique_rmat16_k5_equals_oracle():
    n, s, d = graph.rmat(16, 16, seed=graph.SEEDS["rmat20"])
    o = graph.oriented(n, s, d)
    assert gq.kcliqueCount(_ranked_ctx(n, s, d), 5) == oracle.kclique_count(o.indPtr, o.colIdx, 5)


def test_kclique_rmat18_consistency():
    """R-MAT 18 (the oracle needs minutes for K5 here): K3 and K4 equal to the triangle and hom4
    kernels; a disjoint union of two copies counts twice for K5."""
    n, s, d = graph.rmat(18, 16, seed=graph.SEEDS["rmat20"])
    ctx = _ranked_ctx(n, s, d)
    o = graph.oriented(n, s, d)
    h = gq.hom4Count(ctx, gq.HOM_K4)
    assert gq.kcliqueCount(ctx, 3) == h["tri"] == oracle.tri_count(o.indPtr, o.colIdx)
    assert gq.kcliqueCount(ctx, 4) == h["k4"]
    two = _ranked_ctx(2 * n, np.concatenate([s, s + n]), np.concatenate([d, d + n]))
    assert gq.kcliqueCount(two, 5) == 2 * gq.kcliqueCount(ctx, 5)


def test_kclique_rejects_unranked_and_bad_k():
    n, s, d = graph.rmat(10, 8, seed=1)
    ctx = DeviceGraphCtx(graph.csr_from_edges(n, s, d, graph.SYMMETRIC))
    with pytest.raises(_lib.GaryError) as ei:
        gq.kcliqueCount(ctx, 5)
    assert ei.value.code == _lib.GARY_ERR_INVALID
    ok = _ranked_ctx(n, s, d)
    for k in (0, 11):
        with pytest.raises(_lib.GaryError):
            gq.kcliqueCount(ok, k)
