"""The largest graph the tests build: R-MAT scale 23 (8,388,608 vertices, 134,217,728 tuples,
edgefactor 16), generated and built on the device. The oracle would need minutes here, so the
counts are pinned by agreement of independent algorithms on the same graph: the pivot-bitmap
kernel, the learned routing, the edge-centric cached search and merge (DYNAMIC partition), the
k-clique kernel's K3 (local bitset graphs on the ranked symmetric CSR) and the 4-vertex
kernel's triangle count; and the device build equals the host build."""
from __future__ import annotations

import numpy as np
import pytest

from gary_amd import graph, selector
from gary_amd import graph_query as gq
from gary_amd.device import DeviceGraphCtx, gen_rmat_device

pytestmark = pytest.mark.gpu

SCALE = 23


@pytest.fixture(scope="module")
def rmat23():
    n, nt = 1 << SCALE, 16 << SCALE
    src, dst = gen_rmat_device(SCALE, nt, seed=graph.SEEDS["rmat22"] + 1)
    ori = DeviceGraphCtx()
    ori.buildFromEdges(n, src, dst, nt, graph.ORIENT_DEGREE)
    ori.populateDevice()
    ranked = DeviceGraphCtx()
    ranked.buildFromEdges(n, src, dst, nt, graph.SYMMETRIC_RANKED)
    return n, nt, src, dst, ori, ranked


def test_rmat23_device_build_equals_host(rmat23):
    n, nt, src, dst, ori, _ = rmat23
    s = src.to_array(np.int32, nt)
    d = dst.to_array(np.int32, nt)
    host = graph.oriented(n, s, d)
    dev = ori.getHostCSR()
    assert np.array_equal(dev.indPtr, host.indPtr) and np.array_equal(dev.colIdx, host.colIdx)


def test_rmat23_counts_agree(rmat23):
    _, _, _, _, ori, ranked = rmat23
    want = gq.triangleCountHom(ori, gq.BITMAP)
    assert want > 0
    selector.load()
    assert gq.triangleCountHom(ori, gq.AUTO_LEARNED) == want
    assert gq.triangleCountHom(ori, gq.CACHED_BSEARCH) == want
    cfg = gq.launch_cfg(block=256, group=64, partition=gq.DYNAMIC)
    assert gq.triangleCountHom(ori, gq.MERGE, cfg) == want
    assert gq.kcliqueCount(ranked, 3) == want
    assert gq.hom4Count(ranked, gq.HOM_K4)["tri"] == want
