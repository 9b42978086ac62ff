"""Graph files: the reference's binary GraphCSR / GraphCOO files (saveToBinary / loadFromBinary
through a cereal binary archive, include/Graph.h:16-23,133-148,270-285,301-316) and SNAP edge lists
(loadFromSNAP, src/Graph.cpp:157-196). Host code only.

The reference ships no graph file, so the byte layout is pinned by files assembled here from the
published cereal binary-archive rules: arithmetic values as raw little-endian bytes, enums as their
4-byte underlying type, a shared_ptr as a uint32 id (0x80000000 | k for the k-th new pointer,
k from 1) followed by the pointee, a vector as a uint64 element count and its raw elements."""
from __future__ import annotations

import struct

import numpy as np
import pytest

import oracle
from gary_amd import _lib, graph


def _vec(pid: int, fmt: str, vals) -> bytes:
    return struct.pack("<IQ", pid, len(vals)) + struct.pack(f"<{len(vals)}{fmt}", *vals)


def _csr_bytes(m, nnz, ptr, idx, deg) -> bytes:
    return (struct.pack("<IIiq", 1, 0, m, nnz) + _vec(0x80000001, "q", ptr) +
            _vec(0x80000002, "i", idx) + _vec(0x80000003, "i", deg))


def _coo_bytes(m, row, col) -> bytes:
    return (struct.pack("<IIiq", 1, 1, m, len(row)) + _vec(0x80000001, "i", row) +
            _vec(0x80000002, "i", col))


# the triangle 0-1-2 oriented low -> high, plus an isolated vertex 3
TRI = dict(m=4, nnz=3, ptr=[0, 2, 3, 3, 3], idx=[1, 2, 2], deg=[2, 1, 0, 0])


def test_csr_file_layout_known_answer(tmp_path):
    g = graph.GraphCSR(np.array(TRI["ptr"]), np.array(TRI["idx"]))
    p = tmp_path / "tri.bin"
    graph.save_to_binary(g, p)
    want = _csr_bytes(**TRI)
    assert p.read_bytes() == want
    assert len(want) == 20 + (12 + 40) + (12 + 12) + (12 + 16)  # header 4+4+4+8
    q = tmp_path / "hand.bin"
    q.write_bytes(want)
    h = graph.load_from_binary(q)
    assert h.indPtr.tolist() == TRI["ptr"] and h.colIdx.tolist() == TRI["idx"]
    assert h.deg.tolist() == TRI["deg"]
    assert graph.file_info(q) == (graph.FILE_CSR, 4, 3)


def test_coo_file_layout_known_answer(tmp_path):
    row, col = [0, 0, 1, 2], [1, 2, 2, 2]
    p = tmp_path / "coo.bin"
    graph.save_to_binary((3, row, col), p)
    assert p.read_bytes() == _coo_bytes(3, row, col)
    m, r, c = graph.load_from_binary(p)
    assert m == 3 and r.tolist() == row and c.tolist() == col


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_round_trip_golden_and_rmat(tmp_path, golden_graphs, mode):
    cases = [(rec["n"], np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32))
             for rec in golden_graphs[:8]] + [graph.rmat(12, 16, seed=3)]
    for i, (n, s, d) in enumerate(cases):
        g = graph.csr_from_edges(n, s, d, mode)
        p = tmp_path / f"g{i}.bin"
        graph.save_to_binary(g, p)
        h = graph.load_from_binary(p)
        assert np.array_equal(h.indPtr, g.indPtr) and np.array_equal(h.colIdx, g.colIdx)
        assert np.array_equal(h.deg, g.deg)
        c = tmp_path / f"c{i}.bin"
        graph.save_to_binary((n, s, d), c)
        m, r, cc = graph.load_from_binary(c)
        assert m == n and np.array_equal(r, s) and np.array_equal(cc, d)


def test_loaded_file_counts_like_the_oracle(tmp_path):
    n, s, d = graph.er_4096()
    g = graph.oriented(n, s, d)
    p = tmp_path / "er.bin"
    graph.save_to_binary(g, p)
    h = graph.load_from_binary(p)
    assert oracle.tri_count(h.indPtr, h.colIdx) == oracle.tri_count(g.indPtr, g.colIdx) > 0


def test_bad_files_are_rejected(tmp_path):
    good = _csr_bytes(**TRI)
    bad = {
        "version 2": struct.pack("<I", 2) + good[4:],
        "format 7": good[:4] + struct.pack("<I", 7) + good[8:],
        "truncated": good[:-3],
        "size mismatch": _csr_bytes(4, 3, [0, 2, 3, 3, 3], [1, 2, 2], [2, 1, 0]),
        "null pointer": good[:20] + struct.pack("<I", 0) + good[24:],
        "back reference": good[:20] + struct.pack("<I", 1) + good[24:],
        "not a csr": _csr_bytes(4, 3, [0, 2, 1, 3, 3], [1, 2, 2], [2, 1, 0, 0]),
    }
    for name, blob in bad.items():
        p = tmp_path / "bad.bin"
        p.write_bytes(blob)
        with pytest.raises(_lib.GaryError):
            graph.load_from_binary(p)
    p = tmp_path / "coo.bin"
    p.write_bytes(_coo_bytes(3, [0], [1]))
    with pytest.raises(_lib.GaryError) as ei:
        _lib.check(_lib.lib.gary_graph_file_read_csr(str(p).encode(), None, None, None))
    assert "not a CSR" in str(ei.value)
    with pytest.raises(_lib.GaryError):
        graph.load_from_binary(tmp_path / "missing.bin")


def test_snap_reader(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# Directed graph: example\n# Nodes: 6 Edges: 6\n"
                 "0\t1\n  1 2\n\n2 2\n   # indented comment\n5\t0\r\n3 4\n")
    n, s, d = graph.load_from_snap(p)
    assert n == 6 and s.tolist() == [0, 1, 5, 3] and d.tolist() == [1, 2, 0, 4]
    n, s, d = graph.load_from_snap(p, remove_loops=False)
    assert n == 6 and s.tolist() == [0, 1, 2, 5, 3] and d.tolist() == [1, 2, 2, 0, 4]
    e = tmp_path / "empty.txt"
    e.write_text("# nothing\n")
    assert graph.load_from_snap(e)[0] == 1  # numVertex = maxId + 1 with maxId starting at 0
    for body in ("0 x\n", "7\n", "-1 2\n", "1 99999999999\n"):
        b = tmp_path / "bad.txt"
        b.write_text(body)
        with pytest.raises(_lib.GaryError):
            graph.load_from_snap(b)


def test_snap_to_csr_matches_generator(tmp_path):
    n, s, d = graph.rmat(10, 8, seed=11)
    p = tmp_path / "rmat.txt"
    p.write_text("# rmat\n" + "".join(f"{a} {b}\n" for a, b in zip(s.tolist(), d.tolist())))
    n2, s2, d2 = graph.load_from_snap(p, remove_loops=False)
    assert np.array_equal(s2, s) and np.array_equal(d2, d) and n2 == int(max(s.max(), d.max())) + 1
    g = graph.oriented(n2, s2, d2)
    ip, ci = oracle.csr_from_edges(n2, s, d, 0)
    assert np.array_equal(g.indPtr, ip) and np.array_equal(g.colIdx, ci)
