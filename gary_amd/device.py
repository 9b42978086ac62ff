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
