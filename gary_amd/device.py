"""Device graph context and device buffers over the C ABI.

DeviceGraphCtx mirrors the reference's owner class (include/DeviceGraph.cuh:75-133,
src/DeviceGraph.cu:14-96): loadFromHost, populateDevice, getDeviceCSR / getDeviceCOO,
getHostCSR, clear. Device memory is owned by libgary; nothing here touches a GPU API directly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceGraph, check, lib, ptr
from .graph import GraphCSR

_P = C.c_void_p
_bind = _lib.bind
_bind("gary_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)])
_bind("gary_ctx_destroy", C.c_int, [_P])
_bind("gary_ctx_load_csr", C.c_int, [_P, C.c_int32, C.c_int64, _P, _P])
_bind("gary_ctx_populate", C.c_int, [_P])
_bind("gary_ctx_device_graph", C.c_int, [_P, C.POINTER(DeviceGraph)])
_bind("gary_ctx_copy_to_host", C.c_int, [_P, _P, _P, _P])
_bind("gary_ctx_clear", C.c_int, [_P])
_bind("gary_ctx_build_from_edges", C.c_int, [_P, C.c_int32, C.c_int64, _P, _P, C.c_int, _P])
_bind("gary_gen_rmat_device", C.c_int, [C.c_int32, C.c_int64, C.c_double, C.c_double,
                                        C.c_double, C.c_uint64, _P, _P, _P])
_bind("gary_gen_chung_lu_device", C.c_int, [C.c_int32, C.c_int64, C.c_double, C.c_double,
                                            C.c_uint64, _P, _P, _P])
_bind("gary_csr_from_edges_device", C.c_int, [C.c_int32, C.c_int64, _P, _P, C.c_int, _P, _P,
                                              C.c_int64, _P, _P])


class DeviceBuffer:
    """A device allocation made through gary_malloc (freed on close / garbage collection)."""

    def __init__(self, nbytes: int, device: int = 0):
        self.device = device
        self.nbytes = int(nbytes)
        p = C.c_void_p(None)
        check(lib.gary_malloc(device, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1), device)
        if a.nbytes:
            check(lib.gary_memcpy_h2d(device, b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        if out.nbytes:
            check(lib.gary_memcpy_d2h(self.device, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self) -> None:
        check(lib.gary_memset(self.device, self.ptr, 0, self.nbytes))

    def close(self) -> None:
        if getattr(self, "ptr", 0):
            lib.gary_free(self.device, self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream of `device` (default priority), created by the library; pass
    `.handle` as the stream argument of the asynchronous calls."""

    def __init__(self, device: int = 0):
        self.device = device
        h = C.c_void_p(None)
        check(lib.gary_stream_create(device, C.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        check(lib.gary_stream_synchronize(self.device, self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.gary_stream_destroy(self.device, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraphCtx:
    """Owner of one graph's device buffers (DeviceGraphCtx)."""

    def __init__(self, graph: GraphCSR | None = None, device: int = 0):
        h = C.c_void_p(None)
        check(lib.gary_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device
        self.hostCSR: GraphCSR | None = None
        if graph is not None:
            self.loadFromHost(graph)

    def loadFromHost(self, graph: GraphCSR) -> None:
        """H2D of indPtr and colIdx; deg is derived on the device (src/DeviceGraph.cu:18-28)."""
        check(lib.gary_ctx_load_csr(self._h, graph.m, graph.nnz, ptr(graph.indPtr),
                                    ptr(graph.colIdx)))
        self.hostCSR = graph

    def buildFromEdges(self, n: int, src, dst, ntuples: int | None = None,
                       mode: int = 0) -> None:
        """Build the CSR on the device from tuple arrays (host numpy arrays are uploaded first;
        DeviceBuffers are used in place): symmetrize, drop loops, dedup, orient (`mode` as in
        graph.csr_from_edges). Replaces the reference's host GraphCOO -> GraphCSR build
        (src/Graph.cpp:19-91) followed by loadFromHost."""
        src_b, dst_b, nt = _tuple_buffers(src, dst, ntuples, self.device)
        check(lib.gary_ctx_build_from_edges(self._h, n, nt, src_b.ptr, dst_b.ptr, mode, None))
        self.hostCSR = None

    def populateDevice(self) -> None:
        """Derive the COO rowIdx on the device (src/DeviceGraph.cu:36-47)."""
        check(lib.gary_ctx_populate(self._h))

    def deviceGraph(self) -> DeviceGraph:
        g = DeviceGraph()
        check(lib.gary_ctx_device_graph(self._h, C.byref(g)))
        return g

    def getDeviceCSR(self):
        return self.deviceGraph().csr

    def getDeviceCOO(self):
        return self.deviceGraph().coo

    def detachFromHost(self) -> None:
        self.hostCSR = None

    def getHostCSR(self) -> GraphCSR:
        """Host CSR; after detach it is copied back from the device (m+1 row pointers)."""
        if self.hostCSR is None:
            g = self.deviceGraph()
            indPtr = np.empty(g.csr.m + 1, np.int64)
            colIdx = np.empty(max(g.csr.nnz, 1), np.int32)
            deg = np.empty(max(g.csr.m, 1), np.int32)
            check(lib.gary_ctx_copy_to_host(self._h, ptr(indPtr), ptr(colIdx), ptr(deg)))
            self.hostCSR = GraphCSR(indPtr, colIdx[:g.csr.nnz].copy(), deg[:g.csr.m].copy())
        return self.hostCSR

    def clear(self) -> None:
        check(lib.gary_ctx_clear(self._h))
        self.hostCSR = None

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib.gary_ctx_destroy(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _tuple_buffers(src, dst, ntuples, device):
    if isinstance(src, DeviceBuffer):
        if ntuples is None:
            raise ValueError("ntuples is required with device buffers")
        return src, dst, int(ntuples)
    src_b = DeviceBuffer.from_array(np.ascontiguousarray(src, np.int32), device)
    dst_b = DeviceBuffer.from_array(np.ascontiguousarray(dst, np.int32), device)
    return src_b, dst_b, int(np.asarray(src).shape[0])


def gen_rmat_device(scale: int, ntuples: int, a=0.57, b=0.19, c=0.19, seed: int = 1,
                    device: int = 0) -> tuple[DeviceBuffer, DeviceBuffer]:
    """R-MAT tuples generated on the device, equal to graph.gen_rmat's bit for bit."""
    src = DeviceBuffer(max(4 * ntuples, 4), device)
    dst = DeviceBuffer(max(4 * ntuples, 4), device)
    check(lib.gary_gen_rmat_device(scale, ntuples, a, b, c, seed & (2**64 - 1), src.ptr,
                                   dst.ptr, None))
    check(lib.gary_device_synchronize(device))
    return src, dst


def gen_chung_lu_device(n: int, ntuples: int, gamma=2.1, max_w=50_000.0, seed: int = 1,
                        device: int = 0) -> tuple[DeviceBuffer, DeviceBuffer]:
    """Chung-Lu tuples generated on the device, equal to graph.gen_chung_lu's bit for bit."""
    src = DeviceBuffer(max(4 * ntuples, 4), device)
    dst = DeviceBuffer(max(4 * ntuples, 4), device)
    check(lib.gary_gen_chung_lu_device(n, ntuples, gamma, max_w, seed & (2**64 - 1), src.ptr,
                                       dst.ptr, None))
    return src, dst


def csr_from_edges_device(n: int, src, dst, mode: int = 0, ntuples: int | None = None,
                          device: int = 0) -> GraphCSR:
    """graph.csr_from_edges computed on the device (gary_csr_from_edges_device), copied back."""
    src_b, dst_b, nt = _tuple_buffers(src, dst, ntuples, device)
    cap = 2 * nt if mode in (1, 3) else nt
    ip = DeviceBuffer(8 * (n + 1), device)
    ci = DeviceBuffer(max(4 * cap, 4), device)
    nnz = C.c_int64(0)
    check(lib.gary_csr_from_edges_device(n, nt, src_b.ptr, dst_b.ptr, mode, ip.ptr, ci.ptr, cap,
                                         C.byref(nnz), None))
    return GraphCSR(ip.to_array(np.int64, n + 1), ci.to_array(np.int32, nnz.value))
