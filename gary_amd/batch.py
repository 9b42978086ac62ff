"""Batched triangle counting: many graphs of uneven size in one persistent launch (hipGraph).

The batch is the disjoint union of its graphs (graph g owns vertex ids [vbase[g], vbase[g+1])),
built once on the host; counting replays one captured hipGraph and yields a uint64 count per
graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, graph
from ._lib import DeviceGraph, LaunchCfg, check, lib, ptr

_P = C.c_void_p
_lib.bind("gary_batch_create", C.c_int, [C.c_int, C.c_int32, _P, C.c_int64, _P, _P,
                                         C.POINTER(_P)])
_lib.bind("gary_batch_destroy", C.c_int, [_P])
_lib.bind("gary_batch_set_cfg", C.c_int, [_P, C.POINTER(LaunchCfg)])
_lib.bind("gary_batch_device_graph", C.c_int, [_P, C.POINTER(DeviceGraph)])
_lib.bind("gary_batch_tc_launch", C.c_int, [_P, C.c_int, _P])
_lib.bind("gary_batch_counts", C.c_int, [_P, _P, _P])
_lib.bind("gary_batch_tc_count", C.c_int, [_P, _P])

_M64 = (1 << 64) - 1


def _mix(z: int) -> int:
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


@dataclass
class GraphSpec:
    kind: str   # "er" or "rmat"
    n: int
    m: int      # ER: distinct edges; R-MAT: tuples drawn before rejecting ids >= n
    seed: int


def batch_specs(count: int = 2000, seed: int = graph.SEEDS["batch"], n_range=(64, 20000),
                deg_range=(4, 64)) -> list[GraphSpec]:
    """Graph i: n ~ U[n_range], average degree ~ U[deg_range], m = min(n*deg/2, C(n,2));
    even i Erdős–Rényi, odd i R-MAT(.57,.19,.19,.05) at scale ceil(log2 n) with ids >= n
    rejected. Every draw is a splitmix64 function of (seed, i)."""
    out = []
    for i in range(count):
        x = _mix(seed ^ _mix(i + 1))
        y = _mix(x)
        n = n_range[0] + x % (n_range[1] - n_range[0] + 1)
        dg = deg_range[0] + y % (deg_range[1] - deg_range[0] + 1)
        m = min(n * dg // 2, n * (n - 1) // 2)
        out.append(GraphSpec("er" if i % 2 == 0 else "rmat", n, m, _mix(seed + i)))
    return out


def spec_edges(sp: GraphSpec) -> tuple[np.ndarray, np.ndarray]:
    if sp.kind == "er":
        return graph.gen_er(sp.n, sp.m, sp.seed)
    scale = max(1, int(np.ceil(np.log2(sp.n))))
    s, d = graph.gen_rmat(scale, sp.m, 0.57, 0.19, 0.19, sp.seed)
    keep = (s < sp.n) & (d < sp.n)
    return s[keep], d[keep]


def build_union(specs: list[GraphSpec]) -> tuple[graph.GraphCSR, np.ndarray]:
    """Degree-oriented CSR of the disjoint union and the vertex offsets vbase[len+1]."""
    vbase = np.zeros(len(specs) + 1, np.int64)
    vbase[1:] = np.cumsum([sp.n for sp in specs])
    srcs, dsts = [], []
    for sp, b in zip(specs, vbase[:-1]):
        s, d = spec_edges(sp)
        srcs.append(s + np.int32(b))
        dsts.append(d + np.int32(b))
    src = np.concatenate(srcs)
    dst = np.concatenate(dsts)
    del srcs, dsts
    g = graph.csr_from_edges(int(vbase[-1]), src, dst, graph.ORIENT_DEGREE)
    return g, vbase


class Batch:
    """Device-resident batch of graphs (gary_batch)."""

    def __init__(self, g: graph.GraphCSR, vbase: np.ndarray, device: int = 0):
        self.ngraphs = len(vbase) - 1
        self.vbase = np.asarray(vbase, np.int64)
        self.nnz = g.nnz
        counts = np.diff(self.vbase).astype(np.int32)
        h = _P(None)
        check(lib.gary_batch_create(device, self.ngraphs, ptr(counts), g.nnz, ptr(g.indPtr),
                                    ptr(g.colIdx), C.byref(h)))
        self._h = h

    def set_cfg(self, cfg: LaunchCfg | None) -> None:
        check(lib.gary_batch_set_cfg(self._h, C.byref(cfg) if cfg else None))

    def launch(self, use_graph: bool = True, stream: int | None = None) -> None:
        check(lib.gary_batch_tc_launch(self._h, 1 if use_graph else 0, stream))

    def counts(self, stream: int | None = None) -> np.ndarray:
        out = np.zeros(self.ngraphs, np.uint64)
        check(lib.gary_batch_counts(self._h, ptr(out), stream))
        return out

    def count(self) -> np.ndarray:
        out = np.zeros(self.ngraphs, np.uint64)
        check(lib.gary_batch_tc_count(self._h, ptr(out)))
        return out

    def deviceGraph(self) -> DeviceGraph:
        dg = DeviceGraph()
        check(lib.gary_batch_device_graph(self._h, C.byref(dg)))
        return dg

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.gary_batch_destroy(self._h)
            self._h = _P(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
