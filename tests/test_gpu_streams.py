"""Counts on two streams of one GPU at once: the routed, heuristic and learned strategies keep
their per-call route table, bin sums and arc lists in the stream's own workspace, so two graphs
whose tables differ (ER-4096: a mixed learned table; R-MAT 18: all bitmap) count exactly when
their launches interleave with no synchronisation between them."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from gary_amd import graph, selector
from gary_amd import graph_query as gq
from gary_amd._lib import check, lib, ptr
from gary_amd.device import DeviceBuffer, DeviceGraphCtx, Stream

pytestmark = pytest.mark.gpu

ROUNDS = 8


@pytest.fixture(scope="module")
def two_graphs():
    selector.load()
    out = []
    for n, s, d in (graph.er_4096(), graph.rmat(18, 16, seed=graph.SEEDS["rmat20"] + 18)):
        g = graph.oriented(n, s, d)
        ctx = DeviceGraphCtx(g)
        ctx.populateDevice()
        out.append((ctx, oracle.tri_count(g.indPtr, g.colIdx)))
    return out


def _interleave(two_graphs, launch):
    streams = [Stream(), Stream()]
    counts = [DeviceBuffer(8), DeviceBuffer(8)]
    for c in counts:
        c.zero()
    gq.synchronize()  # the null-stream clears are done before the non-blocking streams start
    for _ in range(ROUNDS):
        for (ctx, _), c, st in zip(two_graphs, counts, streams):
            launch(ctx, c, st.handle)
    for st in streams:
        st.synchronize()
    got = [int(c.to_array(np.uint64, 1)[0]) for c in counts]
    for st in streams:
        st.close()
    return got


@pytest.mark.parametrize("strategy", [gq.AUTO_LEARNED, gq.HEURISTIC, gq.BITMAP])
def test_two_streams_strategy(two_graphs, strategy):
    tabs = [selector.learned_table(ctx)[0] for ctx, _ in two_graphs]
    if strategy == gq.AUTO_LEARNED:  # the case this guards: the two tables differ
        assert not np.array_equal(tabs[0], tabs[1])
    got = _interleave(two_graphs,
                      lambda ctx, c, st: gq.triangleCountHomAsync(ctx, c, strategy, None, st))
    assert got == [ROUNDS * w for _, w in two_graphs]


def test_two_streams_routed_tables(two_graphs):
    """Caller tables: all arcs by cached binary search on one stream, all by the bitmap on the
    other."""
    tables = [np.full(selector.NBINS, selector.CACHED, np.uint8),
              np.full(selector.NBINS, selector.BITMAP, np.uint8)]

    def launch(ctx, c, st):
        k = 0 if ctx is two_graphs[0][0] else 1
        dg = ctx.deviceGraph()
        check(lib.gary_tc_count_routed(C.byref(dg), ptr(tables[k]), None, 0.0, None, c.ptr, st))

    got = _interleave(two_graphs, launch)
    assert got == [ROUNDS * w for _, w in two_graphs]
