"""graph_query — triangle counting entry points (include/kernels/GraphQueryHom.cuh).

The reference launches `triangleCountHom<useCache, N, Policy><<<grid, block, smem>>>(DeviceGraph,
size_t*)` directly (test/TriQueryTest.cu:99-128); here the same choice is made through a
strategy and a launch configuration, and the kernels run on the device behind libgary's C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceGraph, LaunchCfg, check, lib
from .device import DeviceBuffer, DeviceGraphCtx

AUTO_LEARNED = 0
HEURISTIC = 1
MERGE = 2
CACHED_BSEARCH = 3
BSEARCH = 4
BITMAP = 5
STRATEGIES = {"learned": AUTO_LEARNED, "heuristic": HEURISTIC, "merge": MERGE,
              "cached_bsearch": CACHED_BSEARCH, "bsearch": BSEARCH, "bitmap": BITMAP}

INTERLEAVED = 0
CHUNKED = 1
DYNAMIC = -1  # WorkPartition::DYNAMIC: chunks of edges claimed from an atomic cursor

# Arc direction of the pivot strategies: each arc streamed from the endpoint with the shorter
# out-list (reverse rows grouped per call), every N+(v) at pivot u, or AUTO (the former from 2^20
# arcs on).
AUTO_DIRECTION = 0
FORWARD = 1
MIN_DIRECTION = 2

_P = C.c_void_p
_bind = _lib.bind
_bind("gary_tc_count", C.c_int, [C.POINTER(DeviceGraph), C.c_int, C.POINTER(LaunchCfg), _P, _P])
_bind("gary_tc_count_sync", C.c_int, [C.POINTER(DeviceGraph), C.c_int, C.POINTER(LaunchCfg),
                                      C.POINTER(C.c_uint64)])
_bind("gary_tc_count_legacy", C.c_int, [C.POINTER(DeviceGraph), _P, _P])
_bind("gary_index_cache_smem", C.c_int64, [C.c_int32, C.c_int64])


def launch_cfg(block=0, grid=0, group=0, cache_depth=0, partition=INTERLEAVED, bin_small=0,
               bin_hub=0, direction=AUTO_DIRECTION) -> LaunchCfg:
    return LaunchCfg(block, grid, group, cache_depth, partition, bin_small, bin_hub, direction)


def _graph(g) -> DeviceGraph:
    return g.deviceGraph() if isinstance(g, DeviceGraphCtx) else g


def getIndexCacheSMem(block: int, cacheSize: int) -> int:
    """Elements of LDS index cache per workgroup (src/GraphQueryHom.cu:11-13, 64-lane waves)."""
    return int(lib.gary_index_cache_smem(block, cacheSize))


def triangleCountHom(g, strategy: int = HEURISTIC, cfg: LaunchCfg | None = None) -> int:
    """Triangle count of an oriented graph; returns the uint64 count."""
    dg = _graph(g)
    out = C.c_uint64(0)
    check(lib.gary_tc_count_sync(C.byref(dg), strategy, C.byref(cfg) if cfg else None,
                                 C.byref(out)))
    return out.value


def triangleCountHomAsync(g, d_count: DeviceBuffer, strategy: int = HEURISTIC,
                          cfg: LaunchCfg | None = None, stream: int | None = None) -> None:
    """Accumulate into a caller-zeroed device counter (the reference's `size_t *count`)."""
    dg = _graph(g)
    check(lib.gary_tc_count(C.byref(dg), strategy, C.byref(cfg) if cfg else None, d_count.ptr,
                            stream))


def triangleCountLegacy(g, d_count: DeviceBuffer, stream: int | None = None) -> None:
    """triangleCountLegacy(DeviceGraph, unsigned*): 32-bit counter, raises on overflow."""
    dg = _graph(g)
    check(lib.gary_tc_count_legacy(C.byref(dg), d_count.ptr, stream))


def synchronize(device: int = 0) -> None:
    check(lib.gary_device_synchronize(device))


HOM_K4, HOM_DIAMOND, HOM_C4, HOM_ALL = 1, 2, 4, 7


class Hom4Counts(C.Structure):
    _fields_ = [("tri", C.c_uint64), ("k4", C.c_uint64), ("hom_k4", C.c_uint64),
                ("hom_diamond", C.c_uint64), ("hom_c4", C.c_uint64), ("c4", C.c_uint64)]


_bind("gary_hom4_count", C.c_int, [C.POINTER(DeviceGraph), C.c_int, C.POINTER(Hom4Counts), _P])
_bind("gary_hom4_count_routed", C.c_int, [C.POINTER(DeviceGraph), C.c_int, C.c_int, _P,
                                          C.POINTER(Hom4Counts), _P])

# 4-cycle routing modes (gary_hom4_count_routed); a route table is GARY_C4_BINS route bytes
C4_AUTO, C4_FIXED, C4_LEARNED, C4_TABLE = 0, 1, 2, 3


def hom4Count(g, patterns: int = HOM_ALL, stream: int | None = None, c4=None) -> dict:
    """4-vertex homomorphism counts of a ranked symmetric graph (graph.SYMMETRIC_RANKED):
    dict with tri, k4, hom_k4, hom_diamond, hom_c4, c4. `c4` routes the 4-cycle vertices: None
    (AUTO: the learned model when one is set, else the fixed degree split), C4_FIXED,
    C4_LEARNED, or a table of 256 route bytes (gary_amd.c4_selector)."""
    dg = _graph(g)
    out = Hom4Counts()
    if c4 is None:
        check(lib.gary_hom4_count(C.byref(dg), patterns, C.byref(out), stream))
    elif isinstance(c4, int):
        check(lib.gary_hom4_count_routed(C.byref(dg), patterns, c4, None, C.byref(out), stream))
    else:
        t = np.ascontiguousarray(c4, np.uint8)
        if t.shape != (256,):
            raise ValueError("a 4-cycle route table has 256 entries")
        check(lib.gary_hom4_count_routed(C.byref(dg), patterns, C4_TABLE,
                                         t.ctypes.data_as(C.c_void_p), C.byref(out), stream))
    return {k: int(getattr(out, k)) for k, _ in Hom4Counts._fields_}


def rectCountHom(g) -> int:
    """hom(C4) (the reference declares graph_query::rectCountHom without defining it)."""
    return hom4Count(g, HOM_C4)["hom_c4"]


_bind("gary_kclique_count", C.c_int, [C.POINTER(DeviceGraph), C.c_int, C.POINTER(C.c_uint64), _P])


def kcliqueCount(g, k: int, stream: int | None = None) -> int:
    """Number of k-cliques (1 <= k <= 10) of a graph loaded as a ranked symmetric CSR
    (graph.SYMMETRIC_RANKED, the input of hom4Count). No reference counterpart."""
    dg = _graph(g)
    out = C.c_uint64(0)
    check(lib.gary_kclique_count(C.byref(dg), int(k), C.byref(out), stream))
    return int(out.value)


_bind("gary_tri_list", C.c_int, [C.POINTER(DeviceGraph), _P, C.c_uint64, C.POINTER(C.c_uint64),
                                 _P])


def triangleListAsync(g, out: DeviceBuffer, cap: int, stream: int | None = None) -> int:
    """Every triangle of a degree-oriented graph as int32 triples (u, v, w) in the device buffer
    `out` (3 * cap int32); returns the number of triangles (GaryError with GARY_ERR_OVERFLOW when
    it exceeds cap). No reference counterpart: the reference counts triangles only."""
    dg = _graph(g)
    total = C.c_uint64(0)
    check(lib.gary_tri_list(C.byref(dg), out.ptr, int(cap), C.byref(total), stream))
    return int(total.value)


def triangleList(g, cap: int | None = None) -> np.ndarray:
    """triangleListAsync into a fresh buffer, copied back as an (n, 3) int32 array (order set
    by scheduling)."""
    if cap is None:
        cap = triangleCountHom(g, BITMAP)
    out = DeviceBuffer(max(12 * cap, 4))
    n = triangleListAsync(g, out, cap)
    return out.to_array(np.int32, 3 * n).reshape(n, 3)
