"""GPU parity: triangle counts of every strategy and launch configuration equal the oracle's
(bit-exact uint64), through the C ABI."""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from gary_amd import graph, selector
from gary_amd import graph_query as gq
from gary_amd.device import DeviceBuffer, DeviceGraphCtx, gen_chung_lu_device
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

ALL = [gq.BITMAP, gq.HEURISTIC, gq.AUTO_LEARNED, gq.CACHED_BSEARCH, gq.BSEARCH, gq.MERGE]
TRI_FULL = json.loads((GOLDEN / "tri_full.json").read_text())


@pytest.fixture(scope="module", autouse=True)
def learned_model():
    """Every AUTO_LEARNED case here runs the shipped model, whatever ran before in the process."""
    doc = selector.load()
    assert selector.trained()
    return doc


def _ctx(g: graph.GraphCSR) -> DeviceGraphCtx:
    ctx = DeviceGraphCtx(g)
    ctx.populateDevice()
    return ctx


@pytest.fixture(scope="module")
def golden_ctxs(golden_graphs):
    out = []
    for rec in golden_graphs:
        g = graph.csr_from_edges(rec["n"], np.array(rec["src"], np.int32),
                                 np.array(rec["dst"], np.int32), graph.ORIENT_DEGREE)
        out.append((rec, _ctx(g)))
    return out


@pytest.fixture(scope="module")
def er_ctx():
    n, s, d = graph.er_4096()
    g = graph.oriented(n, s, d)
    return g, _ctx(g), oracle.tri_count(g.indPtr, g.colIdx)


@pytest.mark.parametrize("strategy", ALL)
def test_golden_all_strategies(golden_ctxs, strategy):
    for rec, ctx in golden_ctxs:
        assert gq.triangleCountHom(ctx, strategy) == rec["triangles"], rec["name"]


def test_golden_id_orientation(golden_graphs):
    # id orientation (the reference's GraphCOO::orientation) gives the same counts
    for rec in golden_graphs:
        g = graph.csr_from_edges(rec["n"], np.array(rec["src"], np.int32),
                                 np.array(rec["dst"], np.int32), graph.ORIENT_ID)
        ctx = _ctx(g)
        for s in (gq.BITMAP, gq.CACHED_BSEARCH, gq.MERGE):
            assert gq.triangleCountHom(ctx, s) == rec["triangles"], rec["name"]


@pytest.mark.parametrize("strategy", ALL)
def test_er_4096(er_ctx, strategy):
    g, ctx, want = er_ctx
    assert gq.triangleCountHom(ctx, strategy) == want


@pytest.mark.parametrize("group", [8, 16, 32, 64])
@pytest.mark.parametrize("partition", [gq.INTERLEAVED, gq.CHUNKED, gq.DYNAMIC])
@pytest.mark.parametrize("method", [gq.CACHED_BSEARCH, gq.BSEARCH, gq.MERGE])
def test_er_4096_edge_configs(er_ctx, group, partition, method):
    g, ctx, want = er_ctx
    for cache in (1, 16, 64, 256):
        cfg = gq.launch_cfg(block=256, grid=97, group=group, cache_depth=cache,
                            partition=partition)
        assert gq.triangleCountHom(ctx, method, cfg) == want, (group, partition, method, cache)


@pytest.mark.parametrize("bins", [(2, 2), (8, 64), (128, 1024), (128, 128), (16, 512)])
def test_er_4096_pivot_bins(er_ctx, bins):
    g, ctx, want = er_ctx
    cfg = gq.launch_cfg(grid=50, bin_small=bins[0], bin_hub=bins[1])
    assert gq.triangleCountHom(ctx, gq.BITMAP, cfg) == want


@pytest.mark.parametrize("block", [128, 256, 512])
@pytest.mark.parametrize("strategy", [gq.BITMAP, gq.HEURISTIC, gq.AUTO_LEARNED])
def test_pivot_waves(rmat16, block, strategy):
    """The main pivot kernel with 2, 4 and 8 waves per workgroup, with the per-wave small-pivot
    sets at their largest and smallest bins."""
    g, ctx, want = rmat16
    for bins in ((2048 // (block // 64), 2048), (2, 2), (16, 256)):
        cfg = gq.launch_cfg(block=block, grid=333, bin_small=bins[0], bin_hub=bins[1])
        assert gq.triangleCountHom(ctx, strategy, cfg) == want, (block, bins)


def test_pivot_cfg_limits(er_ctx):
    g, ctx, _ = er_ctx
    from gary_amd._lib import GaryError
    with pytest.raises(GaryError):  # per-wave sets of 512 words hold at most 256 ids
        gq.triangleCountHom(ctx, gq.BITMAP, gq.launch_cfg(block=512, bin_small=512))
    with pytest.raises(GaryError):  # edge-centric kernels are bounded to 256 threads
        gq.triangleCountHom(ctx, gq.CACHED_BSEARCH, gq.launch_cfg(block=512))


@pytest.fixture(scope="module")
def rmat16():
    n, s, d = graph.rmat(16, 16, seed=12345)
    g = graph.oriented(n, s, d)
    return g, _ctx(g), oracle.tri_count(g.indPtr, g.colIdx)


@pytest.mark.parametrize("strategy", ALL)
def test_rmat16(rmat16, strategy):
    g, ctx, want = rmat16
    assert gq.triangleCountHom(ctx, strategy) == want


@pytest.mark.parametrize("method", [gq.CACHED_BSEARCH, gq.MERGE])
def test_rmat16_dynamic_partition(rmat16, method):
    """WorkPartition::DYNAMIC (declared but unsupported in the reference): groups claim chunks of
    16 edges from a device-wide cursor; the count does not depend on grid or group size, and a
    second launch restarts the cursor."""
    g, ctx, want = rmat16
    for grid, group in ((1, 64), (7, 16), (4096, 8), (0, 32)):
        cfg = gq.launch_cfg(block=256, grid=grid, group=group, partition=gq.DYNAMIC)
        assert gq.triangleCountHom(ctx, method, cfg) == want, (grid, group)
        assert gq.triangleCountHom(ctx, method, cfg) == want, (grid, group)


def test_rmat16_hub_kernel_paths(rmat16):
    # force every pivot through the workgroup kernel, and the wave kernels' small tables
    g, ctx, want = rmat16
    assert gq.triangleCountHom(ctx, gq.BITMAP, gq.launch_cfg(bin_small=2, bin_hub=2)) == want
    assert gq.triangleCountHom(ctx, gq.BITMAP, gq.launch_cfg(bin_small=128, bin_hub=1024)) == want


def test_accumulate_and_legacy(er_ctx):
    g, ctx, want = er_ctx
    buf = DeviceBuffer(8)
    buf.zero()
    gq.triangleCountHomAsync(ctx, buf, gq.BITMAP)
    gq.triangleCountHomAsync(ctx, buf, gq.CACHED_BSEARCH)
    gq.synchronize()
    assert int(buf.to_array(np.uint64, 1)[0]) == 2 * want  # accumulates like the reference
    b32 = DeviceBuffer(4)
    b32.zero()
    gq.triangleCountLegacy(ctx, b32)
    assert int(b32.to_array(np.uint32, 1)[0]) == want


def test_empty_and_edgeless():
    g = graph.csr_from_edges(10, np.zeros(0, np.int32), np.zeros(0, np.int32))
    ctx = _ctx(g)
    for s in ALL:
        assert gq.triangleCountHom(ctx, s) == 0
    g = graph.csr_from_edges(0, np.zeros(0, np.int32), np.zeros(0, np.int32))
    ctx = _ctx(g)
    for s in ALL:
        assert gq.triangleCountHom(ctx, s) == 0


def test_ctx_roundtrip(er_ctx):
    g, ctx, _ = er_ctx
    ctx.detachFromHost()
    h = ctx.getHostCSR()
    assert np.array_equal(h.indPtr, g.indPtr) and np.array_equal(h.colIdx, g.colIdx)
    assert np.array_equal(h.deg, g.deg)


@pytest.fixture(scope="module")
def rmat20():
    n, s, d = graph.rmat(20, 16, seed=graph.SEEDS["rmat20"])
    g = graph.oriented(n, s, d)
    fx = TRI_FULL["rmat20"]
    assert (g.m, g.nnz, int(g.deg.max())) == (fx["n"], fx["m"], fx["max_out_degree"])
    return g, _ctx(g), fx["triangles"]


@pytest.mark.parametrize("strategy", ALL)
def test_rmat20_full(rmat20, strategy):
    """R-MAT scale 20, edgefactor 16 (the headline workload): every strategy, the learned one
    with the shipped model, equals the oracle's count (tests/golden/tri_full.json), bit-exact."""
    g, ctx, want = rmat20
    assert gq.triangleCountHom(ctx, strategy) == want


@pytest.fixture(scope="module")
def chung_lu_4m():
    """Chung-Lu, 4,194,304 vertices, 67,108,864 tuples, exponent 2.1, weights capped at 50,000:
    generated and built on the device (the same tuples and CSR as the host build that wrote the
    fixture)."""
    fx = TRI_FULL["chung_lu_4m"]
    n, nt = fx["n"], fx["tuples"]
    ds, dd = gen_chung_lu_device(n, nt, 2.1, 50_000.0, graph.SEEDS["chung_lu"])
    ctx = DeviceGraphCtx()
    ctx.buildFromEdges(n, ds, dd, nt, graph.ORIENT_DEGREE)
    del ds, dd
    ctx.populateDevice()
    dg = ctx.deviceGraph()
    assert (dg.csr.m, dg.csr.nnz) == (n, fx["m"])
    return ctx, fx["triangles"]


@pytest.mark.parametrize("strategy", [gq.AUTO_LEARNED, gq.HEURISTIC, gq.BITMAP,
                                      gq.CACHED_BSEARCH])
def test_chung_lu_4m(chung_lu_4m, strategy):
    """The skewed Chung-Lu workload: the learned selector, the plain heuristic, the bitmap and the
    reference's cached two-phase search all give the oracle's count."""
    ctx, want = chung_lu_4m
    assert gq.triangleCountHom(ctx, strategy) == want


def test_chung_lu_4m_forward(chung_lu_4m):
    ctx, want = chung_lu_4m
    assert gq.triangleCountHom(ctx, gq.BITMAP, gq.launch_cfg(direction=gq.FORWARD)) == want


def test_learned_table_is_the_models(rmat20):
    """AUTO_LEARNED reads the loaded model: its table on R-MAT 20 is the fused-inference table,
    and a model that forbids the bitmap route yields a table without it."""
    g, ctx, want = rmat20
    doc = selector.load()
    routes, _ = selector.learned_table(ctx)
    assert set(np.unique(routes)) <= {selector.BITMAP, selector.CACHED, selector.MERGE}
    try:
        selector.set_params(np.array(doc["params"], np.float32), True, 0.0, allowed=6,
                            margin=0.0)
        routes, _ = selector.learned_table(ctx)
        assert selector.BITMAP not in set(np.unique(routes))
        assert gq.triangleCountHom(ctx, gq.AUTO_LEARNED) == want
    finally:
        selector.load()


def _decide(costs, arcs, allowed, margin, fixed):
    """The fused inference's decision restated: the route subset R (within `allowed`) with the
    least sum of fixed costs plus sum over bins of arcs * min_{r in R} cost, leaving the
    all-bitmap routing only for a predicted gain above `margin`; each bin its cheapest route in
    R. Returns (routes, margin of the subset choice)."""
    tot = {}
    for R in range(1, 8):
        if R & allowed != R:
            continue
        sel = [o for o in range(3) if R >> o & 1]
        tot[R] = float((arcs * costs[:, sel].min(axis=1)).sum()) + sum(max(0.0, fixed[o]) for o in sel)
    bestR = min(tot, key=tot.get)
    if allowed & 1 and bestR != 1 and tot[bestR] > (1 - margin) * tot[1]:
        bestR = 1
    sel = [o for o in range(3) if bestR >> o & 1]
    routes = np.array([sel[int(np.argmin(costs[b, sel]))] for b in range(costs.shape[0])])
    others = sorted(tot.values())
    gap = (others[1] - others[0]) / others[0] if len(others) > 1 else 1.0
    return routes, gap


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fused_inference_matches_forward_kernel(er_ctx, seed):
    """The fused table kernel (features -> MLP split over threads -> subset choice) decides as
    the standalone forward kernel's outputs on the same features do, for random models that
    may leave the bitmap (no fixed costs). Bins whose two cheapest routes are within 1% are not
    compared (the sampled features are float sums whose order varies between runs)."""
    g, ctx, want = er_ctx
    rng = np.random.default_rng(seed)
    try:
        p = (rng.standard_normal(selector.NPARAMS) * 0.5).astype(np.float32)
        selector.set_params(p, True, 0.0, allowed=7, margin=0.0)
        selector.set_fixed_costs([1.0], [[0.0, 0.0, 0.0]])
        X, arcs = selector.bin_features(ctx)
        costs = np.exp2(np.minimum(selector.mlp_forward(p, X), 30.0)).astype(np.float64)
        want_routes, gap = _decide(costs, arcs.astype(np.float64), 7, 0.0, [0.0, 0.0, 0.0])
        routes, _ = selector.learned_table(ctx)
        srt = np.sort(costs, axis=1)
        clear = (srt[:, 1] - srt[:, 0]) > 0.01 * srt[:, 0]
        if gap > 0.01:
            assert np.array_equal(routes[clear], want_routes[clear])
        assert gq.triangleCountHom(ctx, gq.AUTO_LEARNED) == want
    finally:
        selector.load()


def _reverse_graph():
    """Id orientation (u -> v iff u > v). Three pivots receive many reversed arcs (sources of
    out-degree 2 or 3 far below theirs): A (out-degree 5,000 over ids spread across 10^6: too
    large for any LDS set, the edge-by-edge fallback), B (3,000: the hub kernel) and C (300: a
    workgroup pivot). Each source also points at one or two out-neighbours of its pivot, so every
    triangle (source, pivot, w) is counted through a reversed arc."""
    src, dst = [], []
    A, B, C = 1_000_000, 900_000, 800_000
    outs = {A: [200 * i for i in range(5000)], B: [150 * i + 7 for i in range(3000)],
            C: [1000 * i + 3 for i in range(300)]}
    for p, ws in outs.items():
        src += [p] * len(ws)
        dst += ws
    for j in range(400):  # into A: two out-neighbours each
        u = 1_000_001 + j
        src += [u, u, u]
        dst += [A, 200 * ((7 * j) % 5000), 200 * ((7 * j + 1) % 5000)]
    for j in range(300):  # into B
        u = 1_050_000 + j
        src += [u, u]
        dst += [B, 150 * ((3 * j) % 3000) + 7]
    for j in range(100):  # into C
        u = 1_060_000 + j
        src += [u, u]
        dst += [C, 1000 * ((11 * j) % 300) + 3]
    return graph.csr_from_edges(1_100_000, np.array(src, np.int32), np.array(dst, np.int32),
                                graph.ORIENT_ID)


@pytest.mark.parametrize("direction", [gq.MIN_DIRECTION, gq.FORWARD])
def test_reverse_rows_every_pivot_path(direction):
    g = _reverse_graph()
    want = oracle.tri_count(g.indPtr, g.colIdx)
    assert want == 800 + 300 + 100
    ctx = _ctx(g)
    for s in (gq.BITMAP, gq.HEURISTIC, gq.AUTO_LEARNED):
        for bins in ((0, 0), (512, 2048), (2, 2)):  # default; big pivots per wave; all per group
            cfg = gq.launch_cfg(bin_small=bins[0], bin_hub=bins[1], direction=direction)
            assert gq.triangleCountHom(ctx, s, cfg) == want, (s, bins)


@pytest.mark.parametrize("direction", [gq.MIN_DIRECTION, gq.FORWARD, gq.AUTO_DIRECTION])
@pytest.mark.parametrize("strategy", [gq.BITMAP, gq.HEURISTIC, gq.AUTO_LEARNED])
def test_pivot_directions(golden_ctxs, rmat16, er_ctx, direction, strategy):
    """Every arc direction of the pivot path gives the oracle's count: forward (every N+(v) at
    pivot u), min-direction (arcs into pivots with much longer out-lists streamed there) and
    auto (min-direction from 2^20 arcs on: forward on these graphs)."""
    cfg = gq.launch_cfg(direction=direction)
    for rec, ctx in golden_ctxs:
        assert gq.triangleCountHom(ctx, strategy, cfg) == rec["triangles"], rec["name"]
    for _, ctx, want in (rmat16, er_ctx):
        assert gq.triangleCountHom(ctx, strategy, cfg) == want
        # per-wave pivots up to out-degree 512, which then receive reversed arcs too
        big = gq.launch_cfg(bin_small=512, bin_hub=2048, direction=direction)
        assert gq.triangleCountHom(ctx, strategy, big) == want


def test_reverse_rows_huge_source():
    """Id orientation. U = 2,000,000 points at V = 1,999,999 and at 2^19 + 6 low ids; V points at
    2^19 + 200 low ids, so the arc (U, V) has d+(V) > d+(U) + 128 but its source's out-degree is
    2^19 or more: it stays forward (the grouped word holds 19 bits of out-degree). Both hubs
    exceed every LDS set and go through the edge-by-edge fallback."""
    U, V = 2_000_000, 1_999_999
    k = 1 << 19
    out_u = np.arange(0, 2 * (k + 6), 2, dtype=np.int32)  # even ids
    out_v = np.arange(0, 3 * (k + 200), 3, dtype=np.int32)  # multiples of 3
    src = np.concatenate([np.full(len(out_u) + 1, U, np.int32), np.full(len(out_v), V, np.int32)])
    dst = np.concatenate([out_u, [V], out_v]).astype(np.int32)
    g = graph.csr_from_edges(U + 1, src, dst, graph.ORIENT_ID)
    want = oracle.tri_count(g.indPtr, g.colIdx)
    assert want == len(np.intersect1d(out_u, out_v))  # triangles (U, V, w), w = 0 mod 6
    ctx = _ctx(g)
    for direction in (gq.MIN_DIRECTION, gq.FORWARD):
        cfg = gq.launch_cfg(direction=direction)
        for s in (gq.BITMAP, gq.AUTO_LEARNED):
            assert gq.triangleCountHom(ctx, s, cfg) == want, (s, direction)


def test_bitmap_without_coo(rmat16):
    """Min-direction lists the reversed arcs from the COO; a graph whose COO was never populated
    streams every arc forward and still counts exactly."""
    g, _, want = rmat16
    ctx = DeviceGraphCtx(g)  # no populateDevice(): rowIdx absent
    assert gq.triangleCountHom(ctx, gq.BITMAP) == want
    with pytest.raises(Exception):  # the routed strategies need the COO
        gq.triangleCountHom(ctx, gq.HEURISTIC)


def test_rmat20_full_forward(rmat20):
    """The headline workload with every arc streamed forward (no reversed-arc list)."""
    g, ctx, want = rmat20
    for s in (gq.BITMAP, gq.AUTO_LEARNED):
        assert gq.triangleCountHom(ctx, s, gq.launch_cfg(direction=gq.FORWARD)) == want


def test_hub_fallback_star():
    """A pivot with out-degree 5000 (id orientation keeps src > dst) exceeds every LDS set and
    goes through the edge-by-edge fallback; 2,500 triangles (center, 2i, 2i+1)."""
    c = 5999
    src = [c] * 5000 + list(range(0, 5000, 2))
    dst = list(range(5000)) + list(range(1, 5001, 2))
    g = graph.csr_from_edges(6000, np.array(src, np.int32), np.array(dst, np.int32),
                             graph.ORIENT_ID)
    assert int(g.deg.max()) == 5000
    ctx = _ctx(g)
    want = oracle.tri_count(g.indPtr, g.colIdx)
    assert want == 2500
    for s in ALL:
        assert gq.triangleCountHom(ctx, s) == want, s
