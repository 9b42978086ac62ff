"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker / CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build() -> Path:
    r = subprocess.run(["make", "-B", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P, I32, I64, U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    sigs = {
        "or_binary_search": (C.c_int, [P, I64, I32]),
        "or_binary_search_pos": (C.c_int, [P, I64, I32, C.POINTER(I32)]),
        "or_build_cache": (None, [P, I64, P, I32]),
        "or_binary_search_2phase": (C.c_int, [P, I64, P, I32, I32]),
        "or_intersect_count": (U64, [P, I64, P, I64]),
        "or_intersect_bsearch_lanes": (U64, [P, I64, P, I64, I32, I32, P]),
        "or_intersect_emit": (I64, [P, I64, P, I64, P]),
        "or_csr_from_edges": (C.c_int, [I32, I64, P, P, I32, P, P, I64, P]),
        "or_tri_count": (U64, [I32, P, P, I32]),
        "or_tri_count_range": (U64, [I32, I32, P, P, I32]),
        "or_tri_count_batch": (U64, [I32, P, P, P, I32, P]),
        "or_omp_threads": (I32, [I32]),
        "or_hom4": (C.c_int, [I32, P, P, I32, P]),
        "or_hom4_ranked": (C.c_int, [I32, P, P, I32, P]),
        "or_kclique_count": (C.c_int, [I32, P, P, I32, I32, P]),
        "or_tri_list": (I64, [I32, P, P, P, I64]),
        "or_tri_list_checksum": (U64, [I32, P, P, I32]),
        "or_triple_hash": (U64, [I32, I32, I32]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class Hom4(C.Structure):
    _fields_ = [("tri", C.c_uint64), ("k4", C.c_uint64), ("hom_k4", C.c_uint64),
                ("hom_diamond", C.c_uint64), ("hom_c4", C.c_uint64), ("c4", C.c_uint64)]


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def binary_search(lst, key) -> bool:
    lst = _a(lst, np.int32)
    return bool(lib.or_binary_search(lst.ctypes.data, lst.size, int(key)))


def binary_search_pos(lst, key) -> tuple[bool, int]:
    lst = _a(lst, np.int32)
    pos = C.c_int32(0)
    f = lib.or_binary_search_pos(lst.ctypes.data, lst.size, int(key), C.byref(pos))
    return bool(f), pos.value


def build_cache(lst, n_cache: int) -> np.ndarray:
    lst = _a(lst, np.int32)
    out = np.empty(n_cache, np.int32)
    lib.or_build_cache(lst.ctypes.data, lst.size, out.ctypes.data, n_cache)
    return out


def binary_search_2phase(lst, cache, key) -> bool:
    lst = _a(lst, np.int32)
    cache = _a(cache, np.int32)
    return bool(lib.or_binary_search_2phase(lst.ctypes.data, lst.size, cache.ctypes.data,
                                            cache.size, int(key)))


def intersect_count(a, b) -> int:
    a, b = _a(a, np.int32), _a(b, np.int32)
    return int(lib.or_intersect_count(a.ctypes.data, a.size, b.ctypes.data, b.size))


def intersect_bsearch_lanes(a, b, n_cache: int = 0, lanes: int = 32):
    a, b = _a(a, np.int32), _a(b, np.int32)
    part = np.zeros(lanes, np.uint64)
    tot = lib.or_intersect_bsearch_lanes(a.ctypes.data, a.size, b.ctypes.data, b.size, n_cache,
                                         lanes, part.ctypes.data)
    return int(tot), part


def intersect_emit(a, b) -> np.ndarray:
    a, b = _a(a, np.int32), _a(b, np.int32)
    out = np.empty(max(min(a.size, b.size), 1), np.int32)
    n = lib.or_intersect_emit(a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data)
    return out[:n].copy()


def csr_from_edges(n: int, src, dst, mode: int = 0) -> tuple[np.ndarray, np.ndarray]:
    src, dst = _a(src, np.int32), _a(dst, np.int32)
    cap = 2 * src.size if mode in (1, 3) else src.size
    indPtr = np.empty(n + 1, np.int64)
    colIdx = np.empty(max(cap, 1), np.int32)
    nnz = C.c_int64(0)
    rc = lib.or_csr_from_edges(n, src.size, src.ctypes.data, dst.ctypes.data, mode,
                               indPtr.ctypes.data, colIdx.ctypes.data, cap, C.byref(nnz))
    if rc != 0:
        raise ValueError("or_csr_from_edges failed")
    return indPtr, colIdx[:nnz.value].copy()


def tri_count(indPtr, colIdx, threads: int = 0) -> int:
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    return int(lib.or_tri_count(indPtr.size - 1, indPtr.ctypes.data, colIdx.ctypes.data, threads))


def tri_count_range(indPtr, colIdx, vbegin: int, vend: int, threads: int = 0) -> int:
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    return int(lib.or_tri_count_range(vbegin, vend, indPtr.ctypes.data, colIdx.ctypes.data,
                                      threads))


def tri_count_batch(vbase, indPtr, colIdx, threads: int = 0) -> np.ndarray:
    """Per-graph triangle counts of a disjoint union (graph g on [vbase[g], vbase[g+1])),
    parallel over graphs, each graph sequential."""
    vbase = _a(vbase, np.int32)
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    out = np.zeros(max(vbase.size - 1, 1), np.uint64)
    lib.or_tri_count_batch(vbase.size - 1, vbase.ctypes.data, indPtr.ctypes.data,
                           colIdx.ctypes.data, threads, out.ctypes.data)
    return out[:vbase.size - 1]


def omp_threads(threads: int = 0) -> int:
    return int(lib.or_omp_threads(threads))


def hom4(indPtr, colIdx, threads: int = 0) -> dict:
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    r = Hom4()
    rc = lib.or_hom4(indPtr.size - 1, indPtr.ctypes.data, colIdx.ctypes.data, threads,
                     C.byref(r))
    if rc != 0:
        raise OverflowError("hom4 count overflows 64 bits")
    return {k: int(getattr(r, k)) for k, _ in Hom4._fields_}


def hom4_ranked(indPtr, colIdx, threads: int = 0) -> dict:
    """hom4 by orientation and wedges, for ranked symmetric CSRs (graph.SYMMETRIC_RANKED) of any
    size; equal to hom4() wherever both run."""
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    r = Hom4()
    rc = lib.or_hom4_ranked(indPtr.size - 1, indPtr.ctypes.data, colIdx.ctypes.data, threads,
                            C.byref(r))
    if rc == -1:
        raise ValueError("hom4_ranked: ids are not ranked by degree")
    if rc != 0:
        raise OverflowError("hom4 count overflows 64 bits")
    return {k: int(getattr(r, k)) for k, _ in Hom4._fields_}


def kclique_count(indPtr, colIdx, k: int, threads: int = 0) -> int:
    """Number of k-cliques from an acyclically oriented CSR (e.g. csr_from_edges mode 0)."""
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    out = C.c_uint64(0)
    rc = lib.or_kclique_count(indPtr.size - 1, indPtr.ctypes.data, colIdx.ctypes.data, int(k),
                              threads, C.byref(out))
    if rc == -4:
        raise OverflowError("k-clique count overflows 64 bits")
    if rc != 0:
        raise ValueError(f"kclique_count: bad k={k}")
    return int(out.value)


def tri_list(indPtr, colIdx, limit: int | None = None) -> np.ndarray:
    """The triangles as an (n, 3) int32 array of (u, v, w), v in N+(u), w in N+(u) ∩ N+(v), in
    the sequential order; only the first `limit` when given."""
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    m = indPtr.size - 1
    n = int(limit) if limit is not None else int(lib.or_tri_list(m, indPtr.ctypes.data,
                                                                   colIdx.ctypes.data, None, 0))
    out = np.empty((max(n, 1), 3), np.int32)
    n = int(lib.or_tri_list(m, indPtr.ctypes.data, colIdx.ctypes.data, out.ctypes.data, n))
    return out[:n]


def tri_list_checksum(indPtr, colIdx, threads: int = 0) -> int:
    """Wrapping uint64 sum of triple_hash over the listing (order-independent)."""
    indPtr, colIdx = _a(indPtr, np.int64), _a(colIdx, np.int32)
    return int(lib.or_tri_list_checksum(indPtr.size - 1, indPtr.ctypes.data, colIdx.ctypes.data,
                                        threads))


def triple_hash_np(t: np.ndarray) -> np.ndarray:
    """or_triple_hash over the rows of an (n, 3) int32 array, vectorised (uint64, wrapping)."""
    u, v, w = (t[:, i].astype(np.uint32).astype(np.uint64) for i in range(3))
    with np.errstate(over="ignore"):
        x = (u * np.uint64(0x9E3779B97F4A7C15)) ^ (v * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
            (w * np.uint64(0x165667B19E3779F9))
        x ^= x >> np.uint64(29)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(32)
    return x
