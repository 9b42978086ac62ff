"""Triangle listing on the GPU (gary_tri_list) against the oracle's listing (the emit variant of
the triangle query, restated from include/ops/Intersection.cuh:18-50 and
test/TriQueryTest.cu:17-37): the same set of (u, v, w) triples; at full size an order-independent
checksum of the triples and their count."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gary_amd import _lib, graph
from gary_amd import graph_query as gq
from gary_amd.device import DeviceBuffer, DeviceGraphCtx

pytestmark = pytest.mark.gpu


def _sorted_rows(t: np.ndarray) -> np.ndarray:
    return t[np.lexsort((t[:, 2], t[:, 1], t[:, 0]))]


def _check(g):
    ctx = DeviceGraphCtx(g)
    got = gq.triangleList(ctx)
    want = oracle.tri_list(g.indPtr, g.colIdx)
    assert got.shape == want.shape
    assert np.array_equal(_sorted_rows(got), _sorted_rows(want))


def test_listing_golden(golden_graphs):
    for rec in golden_graphs:
        g = graph.oriented(rec["n"], np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32))
        _check(g)


@pytest.mark.parametrize("mode", [graph.ORIENT_DEGREE, graph.ORIENT_ID])
def test_listing_seeded(mode):
    # ER, R-MAT with hubs above the LDS-cached list length (searched in HBM), Chung-Lu
    for n, s, d in (graph.er_4096(), graph.rmat(16, 16, seed=9),
                    (20000,) + graph.gen_chung_lu(20000, 400000, seed=3, max_w=8000.0)):
        _check(graph.csr_from_edges(n, s, d, mode))


def test_listing_capacity_and_empty():
    g = graph.oriented(*graph.er_4096())
    ctx = DeviceGraphCtx(g)
    n = gq.triangleCountHom(ctx, gq.BITMAP)
    out = DeviceBuffer(12 * 100)
    with pytest.raises(_lib.GaryError) as ei:
        gq.triangleListAsync(ctx, out, 100)
    assert ei.value.code == _lib.GARY_ERR_OVERFLOW and str(n) in str(ei.value)
    part = out.to_array(np.int32, 300).reshape(100, 3)
    full = {tuple(r) for r in oracle.tri_list(g.indPtr, g.colIdx).tolist()}
    assert all(tuple(r) in full for r in part.tolist())
    star = graph.oriented(9, np.zeros(8, np.int32), np.arange(1, 9, dtype=np.int32))
    assert gq.triangleList(DeviceGraphCtx(star), 4).shape == (0, 3)


def test_listing_rmat20_full_size_checksum():
    """R-MAT 20: 424,325,219 triangles (5.1 GB of triples); count and order-independent checksum
    equal to the oracle's."""
    n, s, d = graph.rmat(20, 16)
    g = graph.oriented(n, s, d)
    ctx = DeviceGraphCtx(g)
    want_n = oracle.tri_count(g.indPtr, g.colIdx)
    out = DeviceBuffer(12 * want_n)
    assert gq.triangleListAsync(ctx, out, want_n) == want_n
    total = np.uint64(0)
    chunk = 1 << 25
    for at in range(0, want_n, chunk):
        k = min(chunk, want_n - at)
        t = np.empty(3 * k, np.int32)
        _lib.check(_lib.lib.gary_memcpy_d2h(0, t.ctypes.data, out.ptr + 12 * at, t.nbytes))
        with np.errstate(over="ignore"):
            total += oracle.triple_hash_np(t.reshape(k, 3)).sum(dtype=np.uint64)
    assert int(total) == oracle.tri_list_checksum(g.indPtr, g.colIdx)
