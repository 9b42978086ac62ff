"""Graph construction on the device (gary_gen_rmat_device, gary_csr_from_edges_device,
gary_ctx_build_from_edges) against the oracle's COO -> CSR build (src/Graph.cpp:19-91 restated
with dedup and degree orientation) and the host product build: bit-exact indPtr and colIdx."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gary_amd import _lib, graph
from gary_amd import graph_query as gq
from gary_amd.device import (DeviceGraphCtx, csr_from_edges_device, gen_chung_lu_device,
                             gen_rmat_device)

pytestmark = pytest.mark.gpu
MODES = [0, 1, 2, 3]


def _same(g, ip, ci, what=""):
    assert np.array_equal(g.indPtr, ip), f"indPtr differs {what}"
    assert np.array_equal(g.colIdx, ci), f"colIdx differs {what}"


@pytest.mark.parametrize("mode", MODES)
def test_device_build_equals_oracle_golden(golden_graphs, mode):
    for rec in golden_graphs:
        n = rec["n"]
        src, dst = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        g = csr_from_edges_device(n, src, dst, mode)
        ip, ci = oracle.csr_from_edges(n, src, dst, mode)
        _same(g, ip, ci, rec["name"])


@pytest.mark.parametrize("mode", MODES)
def test_device_build_edge_cases(mode):
    cases = {
        "empty": (5, [], []),
        "no vertices": (0, [], []),
        "one vertex, loops": (1, [0, 0], [0, 0]),
        "only loops": (4, [0, 1, 2, 3], [0, 1, 2, 3]),
        "one edge repeated": (3, [2, 1, 2, 1, 2], [1, 2, 1, 2, 1]),
        "star": (9, [0] * 8, list(range(1, 9))),
        "isolated tail": (1000, [0, 1, 2], [1, 2, 0]),
        "exact power of two": (8, [7, 6, 0, 7], [6, 7, 7, 0]),
    }
    for name, (n, s, d) in cases.items():
        s, d = np.array(s, np.int32), np.array(d, np.int32)
        g = csr_from_edges_device(n, s, d, mode)
        ip, ci = oracle.csr_from_edges(n, s, d, mode)
        _same(g, ip, ci, name)


@pytest.mark.parametrize("mode", MODES)
def test_device_build_equals_oracle_seeded(mode):
    # ER with one tile, R-MAT spanning many tiles and sort passes, Chung-Lu with heavy hubs
    for n, s, d in (graph.er_4096(), graph.rmat(16, 16, seed=5),
                    (20000,) + graph.gen_chung_lu(20000, 300000, seed=8, max_w=3000.0)):
        g = csr_from_edges_device(n, s, d, mode)
        ip, ci = oracle.csr_from_edges(n, s, d, mode)
        _same(g, ip, ci, f"n={n}")


def test_device_build_errors():
    with pytest.raises(_lib.GaryError) as ei:
        csr_from_edges_device(3, np.array([0, 5], np.int32), np.array([1, 1], np.int32))
    assert ei.value.code == _lib.GARY_ERR_INVALID and "out of" in _lib.last_error()
    with pytest.raises(_lib.GaryError):
        csr_from_edges_device(3, np.array([0, -1], np.int32), np.array([1, 1], np.int32))


def test_device_rmat_tuples_equal_host():
    for scale, nt, seed in ((5, 1000, 3), (16, 1 << 20, graph.SEEDS["rmat20"])):
        hs, hd = graph.gen_rmat(scale, nt, seed=seed)
        ds, dd = gen_rmat_device(scale, nt, seed=seed)
        assert np.array_equal(ds.to_array(np.int32, nt), hs)
        assert np.array_equal(dd.to_array(np.int32, nt), hd)


def test_device_chung_lu_tuples_equal_host():
    for n, nt, w, seed in ((5000, 40000, 500.0, 2), (4_194_304, 1 << 22, 50_000.0,
                                                      graph.SEEDS["chung_lu"])):
        hs, hd = graph.gen_chung_lu(n, nt, seed=seed, max_w=w)
        ds, dd = gen_chung_lu_device(n, nt, seed=seed, max_w=w)
        assert np.array_equal(ds.to_array(np.int32, nt), hs)
        assert np.array_equal(dd.to_array(np.int32, nt), hd)


def test_ctx_build_rmat20_full_size():
    """RMAT-20 (16.8 M tuples) generated and built on the device: CSR equal to the host build of
    the host tuples, and the triangle count equal to the oracle's."""
    n, s, d = graph.rmat(20, 16)
    host = graph.oriented(n, s, d)
    ds, dd = gen_rmat_device(20, 16 * n, seed=graph.SEEDS["rmat20"])
    ctx = DeviceGraphCtx()
    ctx.buildFromEdges(n, ds, dd, 16 * n, graph.ORIENT_DEGREE)
    g = ctx.getHostCSR()
    _same(g, host.indPtr, host.colIdx, "rmat20")
    assert np.array_equal(g.deg, host.deg)
    ctx.populateDevice()
    want = oracle.tri_count(host.indPtr, host.colIdx)
    assert gq.triangleCountHom(ctx, gq.BITMAP) == want


def test_device_build_rmat22_ranked_full_size():
    """RMAT-22 ranked symmetric CSR (the 4-vertex counting input, 134 M keys) equal to the host
    build."""
    n, s, d = graph.rmat(22, 16, seed=graph.SEEDS["rmat22"])
    host = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
    g = csr_from_edges_device(n, s, d, graph.SYMMETRIC_RANKED)
    _same(g, host.indPtr, host.colIdx, "rmat22 ranked")
