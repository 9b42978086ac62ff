"""Host graphs: seeded generators and the COO -> CSR build (through libgary's host code).

Mirrors the reference's host containers (include/Graph.h:157-292 GraphCSR, src/Graph.cpp:19-91
symmetrize / orientation / toCSR_) for what the hot path consumes: an int64 row-pointer array,
int32 sorted adjacency lists and int32 degrees.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, lib, ptr

ORIENT_DEGREE = _lib.CSR_ORIENT_DEGREE
SYMMETRIC = _lib.CSR_SYMMETRIC
ORIENT_ID = _lib.CSR_ORIENT_ID
SYMMETRIC_RANKED = 3


@dataclass
class GraphCSR:
    """CSR graph on the host (GraphCSR, include/Graph.h:157-292)."""
    indPtr: np.ndarray  # int64 [m+1]
    colIdx: np.ndarray  # int32 [nnz]
    deg: np.ndarray = field(default=None)  # int32 [m]

    def __post_init__(self):
        self.indPtr = np.ascontiguousarray(self.indPtr, dtype=np.int64)
        self.colIdx = np.ascontiguousarray(self.colIdx, dtype=np.int32)
        if self.deg is None:  # GraphCSR ctor, include/Graph.h:242-248
            self.deg = np.diff(self.indPtr).astype(np.int32)

    @property
    def m(self) -> int:
        return int(self.indPtr.shape[0] - 1)

    def numV(self) -> int:
        return self.m

    @property
    def nnz(self) -> int:
        return int(self.colIdx.shape[0])

    def numE(self) -> int:
        return self.nnz

    def N(self, v: int) -> np.ndarray:
        return self.colIdx[self.indPtr[v]:self.indPtr[v + 1]]

    def toCOO(self) -> tuple[np.ndarray, np.ndarray]:
        """GraphCSR::toCOO (src/Graph.cpp:141-155): (rowIdx, colIdx)."""
        row = np.repeat(np.arange(self.m, dtype=np.int32), self.deg)
        return row, self.colIdx


def gen_er(n: int, m: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Erdős–Rényi G(n, m): m distinct unordered pairs (no loops) in draw order."""
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    check(lib.gary_gen_er(n, m, seed & (2**64 - 1), ptr(src), ptr(dst)))
    return src, dst


def gen_rmat(scale: int, ntuples: int, a=0.57, b=0.19, c=0.19, seed: int = 1):
    """R-MAT tuples (no noise, no permutation); may contain loops and duplicates."""
    src = np.empty(ntuples, np.int32)
    dst = np.empty(ntuples, np.int32)
    check(lib.gary_gen_rmat(scale, ntuples, a, b, c, seed & (2**64 - 1), ptr(src), ptr(dst)))
    return src, dst


def gen_chung_lu(n: int, ntuples: int, gamma=2.1, max_w=50_000.0, seed: int = 1):
    """Chung–Lu tuples with power-law expected degrees (exponent gamma, capped at max_w)."""
    src = np.empty(ntuples, np.int32)
    dst = np.empty(ntuples, np.int32)
    check(lib.gary_gen_chung_lu(n, ntuples, gamma, max_w, seed & (2**64 - 1), ptr(src),
                                ptr(dst)))
    return src, dst


def csr_from_edges(n: int, src: np.ndarray, dst: np.ndarray, mode: int = ORIENT_DEGREE) -> GraphCSR:
    """Symmetrize + drop loops + dedup, then a CSR oriented by (degree, id), symmetric, or
    oriented by id (src > dst, the reference's GraphCOO::orientation)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    nt = int(src.shape[0])
    cap = 2 * nt if mode in (SYMMETRIC, SYMMETRIC_RANKED) else nt
    indPtr = np.empty(n + 1, np.int64)
    colIdx = np.empty(max(cap, 1), np.int32)
    nnz = C.c_int64(0)
    check(lib.gary_csr_from_edges(n, nt, ptr(src), ptr(dst), mode, ptr(indPtr), ptr(colIdx),
                                  cap, C.byref(nnz)))
    return GraphCSR(indPtr, colIdx[:nnz.value].copy())


# ---- workloads (seeded; the seeds are part of the workload definition) -----------------------

SEEDS = {"er": 0x6A7E0001, "rmat20": 0x6A7E0002, "chung_lu": 0x6A7E0003, "batch": 0x6A7E0004,
         "rmat22": 0x6A7E0005}


def er_4096(seed: int = SEEDS["er"]):
    """G(n=4,096, m=65,536) degree-oriented int32 CSR."""
    s, d = gen_er(4096, 65536, seed)
    return 4096, s, d


def rmat(scale: int, edgefactor: int = 16, seed: int = SEEDS["rmat20"]):
    n = 1 << scale
    s, d = gen_rmat(scale, edgefactor * n, 0.57, 0.19, 0.19, seed)
    return n, s, d


def chung_lu_4m(seed: int = SEEDS["chung_lu"], n: int = 4_194_304, ntuples: int = 64 * 1_048_576):
    s, d = gen_chung_lu(n, ntuples, 2.1, 50_000.0, seed)
    return n, s, d


def oriented(n, s, d) -> GraphCSR:
    return csr_from_edges(n, s, d, ORIENT_DEGREE)


def symmetric(n, s, d) -> GraphCSR:
    return csr_from_edges(n, s, d, SYMMETRIC)


# ---- graph files (the reference's saveToBinary / loadFromBinary / loadFromSNAP) -------------

_lib.bind("gary_graph_file_info", C.c_int, [C.c_char_p, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int32), C.POINTER(C.c_int64)])
_lib.bind("gary_graph_file_read_csr", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p])
_lib.bind("gary_graph_file_read_coo", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p])
_lib.bind("gary_graph_file_write_csr", C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_void_p,
                                                 C.c_void_p, C.c_void_p])
_lib.bind("gary_graph_file_write_coo", C.c_int, [C.c_char_p, C.c_int32, C.c_int64, C.c_void_p,
                                                 C.c_void_p])
_lib.bind("gary_snap_read", C.c_int, [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int32)])
_lib.bind("gary_coo_to_csr", C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p])

FILE_CSR, FILE_COO = 0, 1


def _path(p) -> bytes:
    return str(p).encode()


def file_info(path) -> tuple[int, int, int]:
    """(format, m, nnz) of a binary graph file (cereal binary archive, SerDesVersion V1)."""
    fmt, m, nnz = C.c_int(0), C.c_int32(0), C.c_int64(0)
    check(lib.gary_graph_file_info(_path(path), C.byref(fmt), C.byref(m), C.byref(nnz)))
    return fmt.value, m.value, nnz.value


def load_from_binary(path):
    """loadFromBinary<GraphCSR | GraphCOO> (include/Graph.h:301-309): a GraphCSR for a CSR file,
    (m, row, col) for a COO file."""
    fmt, m, nnz = file_info(path)
    if fmt == FILE_CSR:
        indPtr = np.empty(m + 1, np.int64)
        colIdx = np.empty(max(nnz, 1), np.int32)
        deg = np.empty(max(m, 1), np.int32)
        check(lib.gary_graph_file_read_csr(_path(path), ptr(indPtr), ptr(colIdx), ptr(deg)))
        return GraphCSR(indPtr, colIdx[:nnz].copy(), deg[:m].copy())
    row = np.empty(max(nnz, 1), np.int32)
    col = np.empty(max(nnz, 1), np.int32)
    check(lib.gary_graph_file_read_coo(_path(path), ptr(row), ptr(col)))
    return m, row[:nnz].copy(), col[:nnz].copy()


def save_to_binary(graph, path) -> None:
    """saveToBinary (include/Graph.h:311-316) of a GraphCSR, or of a COO given as (m, row, col)."""
    if isinstance(graph, GraphCSR):
        check(lib.gary_graph_file_write_csr(_path(path), graph.m, graph.nnz, ptr(graph.indPtr),
                                            ptr(graph.colIdx), ptr(graph.deg)))
    else:
        m, row, col = graph
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        check(lib.gary_graph_file_write_coo(_path(path), int(m), int(row.shape[0]), ptr(row),
                                            ptr(col)))


def load_from_snap(path, remove_loops: bool = True) -> tuple[int, np.ndarray, np.ndarray]:
    """loadFromSNAP (src/Graph.cpp:157-196): (number of vertices = max id + 1, src, dst)."""
    nt, n = C.c_int64(0), C.c_int32(0)
    check(lib.gary_snap_read(_path(path), int(remove_loops), None, None, 0, C.byref(nt),
                             C.byref(n)))
    src = np.empty(max(nt.value, 1), np.int32)
    dst = np.empty(max(nt.value, 1), np.int32)
    check(lib.gary_snap_read(_path(path), int(remove_loops), ptr(src), ptr(dst), nt.value,
                             C.byref(nt), C.byref(n)))
    return n.value, src[:nt.value].copy(), dst[:nt.value].copy()


def coo_to_csr(m: int, row, col) -> GraphCSR:
    """GraphCOO::toCSR_ (src/Graph.cpp:63-91): rows by source id, each sorted ascending,
    duplicates and loops kept (no dedup, unlike csr_from_edges)."""
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    nnz = int(row.shape[0])
    indPtr = np.empty(int(m) + 1, np.int64)
    colIdx = np.empty(max(nnz, 1), np.int32)
    check(lib.gary_coo_to_csr(int(m), nnz, ptr(row), ptr(col), ptr(indPtr), ptr(colIdx)))
    return GraphCSR(indPtr, colIdx[:nnz].copy())
