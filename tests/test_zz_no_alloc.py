"""gary_tc_count_noalloc: the pivot-bitmap count with its scratch in a fixed device array, called
through the C ABI, equal to gary_tc_count (GARY_TC_BITMAP) and to the oracle on two calls in a
row (the second on a smaller graph, so it runs over what the first left in the array), with
accumulation into the caller's counter, and a too-large graph rejected before any device work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from gary_amd import _lib, graph
from gary_amd import graph_query as gq
from gary_amd._lib import DeviceGraph, GaryError, LaunchCfg, check, lib
from gary_amd.device import DeviceBuffer, DeviceGraphCtx

pytestmark = pytest.mark.gpu

_lib.bind("gary_tc_count_noalloc", C.c_int, [C.POINTER(DeviceGraph), C.POINTER(LaunchCfg),
                                             C.c_void_p, C.c_void_p])
MAX_VERTICES = 1_397_568


def _noalloc(ctx, d_count, cfg=None):
    dg = ctx.deviceGraph()
    check(lib.gary_tc_count_noalloc(C.byref(dg), C.byref(cfg) if cfg else None, d_count.ptr,
                                    None))


def _count(ctx, fn):
    d = DeviceBuffer(8)
    d.zero()
    fn(ctx, d)
    gq.synchronize()
    return int(d.to_array(np.uint64, 1)[0])


def _graph(n, s, d):
    g = graph.oriented(n, s, d)
    ctx = DeviceGraphCtx(g)
    return ctx, oracle.tri_count(g.indPtr, g.colIdx)


def test_noalloc_equals_tc_count_twice():
    big = _graph(*graph.rmat(18, 16, seed=graph.SEEDS["rmat20"] + 7))
    small = _graph(*graph.er_4096())
    for ctx, want in (big, small, big):
        got = _count(ctx, _noalloc)
        ref = _count(ctx, lambda c, d: gq.triangleCountHomAsync(c, d, gq.BITMAP))
        assert got == ref == want


def test_noalloc_cfg_and_accumulation():
    ctx, want = _graph(*graph.rmat(16, 16, seed=12345))
    d = DeviceBuffer(8)
    d.zero()
    for cfg in (None, gq.launch_cfg(block=128, grid=333, bin_small=16, bin_hub=256),
                gq.launch_cfg(block=512, bin_small=2, bin_hub=2)):
        _noalloc(ctx, d, cfg)
    gq.synchronize()
    assert int(d.to_array(np.uint64, 1)[0]) == 3 * want


def test_noalloc_limits():
    # the largest vertex count the array holds is accepted, one more is rejected on the host
    src = np.array([0, 0, 1], np.int32)
    dst = np.array([1, 2, 2], np.int32)
    ok = DeviceGraphCtx(graph.csr_from_edges(MAX_VERTICES, src, dst))
    assert _count(ok, _noalloc) == 1
    too_big = DeviceGraphCtx(graph.csr_from_edges(MAX_VERTICES + 1, src, dst))
    d = DeviceBuffer(8)
    with pytest.raises(GaryError):
        _noalloc(too_big, d)
    empty = DeviceGraphCtx(graph.csr_from_edges(5, np.zeros(0, np.int32), np.zeros(0, np.int32)))
    assert _count(empty, _noalloc) == 0
