"""Search and intersection operators on the device (include/ops/Search.cuh, Intersection.cuh).

Host-side mirror of the reference's operator vocabulary: binary_search (plain or two-phase with
an index cache of N entries), intersection count and intersection with output. Inputs are
numpy arrays copied to the device, or DeviceBuffers already there.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .device import DeviceBuffer

SEARCH_PLAIN = 0
SEARCH_2PHASE = 1
ISECT_BSEARCH = 0
ISECT_CACHED_BSEARCH = 1
ISECT_MERGE = 2
ISECT_BITMAP = 3

_P, _I64 = C.c_void_p, C.c_int64
_lib.bind("gary_search", C.c_int, [_P, _I64, _P, _I64, _P, C.c_int, C.c_int, _P])
_lib.bind("gary_search_count", C.c_int, [_P, _I64, _P, _I64, _P, C.c_int, C.c_int, _P])
_lib.bind("gary_search_batched", C.c_int, [C.c_int32, _P, _P, _P, _P, _P, C.c_int, C.c_int, _P])
_lib.bind("gary_intersect_count", C.c_int, [_P, _I64, _P, _I64, C.c_int, C.c_int, _P, _P])
_lib.bind("gary_intersect_emit", C.c_int, [_P, _I64, _P, _I64, C.c_int, C.c_int, _P, _P, _P])
_lib.bind("gary_block_queue_exercise", C.c_int, [C.c_int, _P, C.c_uint32, _P, _P, _P])

QUEUE_EXCLUSIVE = 0
QUEUE_CONCURRENT = 1
QUEUE_SLOTS = 128
QUEUE_WAVES = 16


def _dev(a) -> tuple[DeviceBuffer, int]:
    if isinstance(a, DeviceBuffer):
        return a, a.nbytes // 4
    a = np.ascontiguousarray(a, dtype=np.int32)
    return DeviceBuffer.from_array(a), int(a.size)


def binary_search(lst, keys, mode: int = SEARCH_PLAIN, cache_depth: int = 32) -> np.ndarray:
    """Membership of every key in the sorted list (binarySearchKernel / 2-phase variant)."""
    dl, n = _dev(lst)
    dk, nk = _dev(keys)
    out = DeviceBuffer(max(nk, 1))
    check(lib.gary_search(dl.ptr, n, dk.ptr, nk, out.ptr, mode, cache_depth, None))
    return out.to_array(np.uint8, nk).astype(bool)


def binary_search_count(lst, keys, mode: int = SEARCH_PLAIN, cache_depth: int = 32) -> int:
    dl, n = _dev(lst)
    dk, nk = _dev(keys)
    cnt = DeviceBuffer(8)
    cnt.zero()
    check(lib.gary_search_count(dl.ptr, n, dk.ptr, nk, cnt.ptr, mode, cache_depth, None))
    return int(cnt.to_array(np.uint64, 1)[0])


def batched_binary_search(lists, keys, mode: int = SEARCH_PLAIN, cache_depth: int = 32):
    """One (list, keys) pair per item (batchedBinarySearch{,2Phase}Kernel)."""
    assert len(lists) == len(keys)
    bufs = [(_dev(l), _dev(k)) for l, k in zip(lists, keys)]
    outs = [DeviceBuffer(max(nk, 1)) for _, (_, nk) in bufs]
    lp = DeviceBuffer.from_array(np.array([b[0][0].ptr for b in bufs], np.uint64))
    ls = DeviceBuffer.from_array(np.array([b[0][1] for b in bufs], np.int64))
    kp = DeviceBuffer.from_array(np.array([b[1][0].ptr for b in bufs], np.uint64))
    ks = DeviceBuffer.from_array(np.array([b[1][1] for b in bufs], np.int64))
    rp = DeviceBuffer.from_array(np.array([o.ptr for o in outs], np.uint64))
    check(lib.gary_search_batched(len(lists), lp.ptr, ls.ptr, kp.ptr, ks.ptr, rp.ptr, mode,
                                  cache_depth, None))
    return [o.to_array(np.uint8, b[1][1]).astype(bool) for o, b in zip(outs, bufs)]


def intersect_count(a, b, method: int = ISECT_BSEARCH, cache_depth: int = 32) -> int:
    da, na = _dev(a)
    db, nb = _dev(b)
    cnt = DeviceBuffer(8)
    cnt.zero()
    check(lib.gary_intersect_count(da.ptr, na, db.ptr, nb, method, cache_depth, cnt.ptr, None))
    return int(cnt.to_array(np.uint64, 1)[0])


def intersect(a, b, method: int = ISECT_BSEARCH, cache_depth: int = 32) -> np.ndarray:
    """Sorted intersection values (intersectBinarySearch with output, OutSorted)."""
    da, na = _dev(a)
    db, nb = _dev(b)
    out = DeviceBuffer(max(min(na, nb), 1) * 4)
    cnt = DeviceBuffer(8)
    check(lib.gary_intersect_emit(da.ptr, na, db.ptr, nb, method, cache_depth, out.ptr, cnt.ptr,
                                  None))
    k = int(cnt.to_array(np.uint64, 1)[0])
    return out.to_array(np.int32, k)


def block_queue_exercise(mode: int, items) -> tuple[np.ndarray, np.ndarray]:
    """The block-scope queue (CTAQueue, include/ops/DeviceQueue.cuh:33-182) driven on the device:
    (out, status) as gary_block_queue_exercise describes them."""
    a = np.ascontiguousarray(items, dtype=np.uint32)
    d_items = DeviceBuffer.from_array(a)
    n_out = QUEUE_WAVES * QUEUE_SLOTS if mode == QUEUE_EXCLUSIVE else max(int(a.size), 1)
    d_out = DeviceBuffer(4 * n_out)
    d_status = DeviceBuffer(4 * QUEUE_WAVES)
    check(lib.gary_block_queue_exercise(int(mode), d_items.ptr, int(a.size), d_out.ptr,
                                        d_status.ptr, None))
    check(lib.gary_device_synchronize(0))
    return d_out.to_array(np.uint32, n_out), d_status.to_array(np.int32, QUEUE_WAVES)
