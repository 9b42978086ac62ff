"""The reference's conversion tool (convert.cpp:13-112) as gary_amd/gary_convert, and its COO -> CSR
step gary_coo_to_csr (GraphCOO::toCSR_, src/Graph.cpp:63-91). Host code only.

toCSR_ keeps duplicates and loops and sorts each row, so its output is fixed by mathematics: the
tuples in lexicographic (row, col) order (numpy lexsort here), rows cut by a histogram of the row
ids. The chained tool (SNAP -> COO file -> symmetric CSR -> id-oriented CSR) is pinned by the
networkx triangle counts of the golden graphs, and the files by their byte layout."""
from __future__ import annotations

import subprocess
import struct

import numpy as np
import pytest

import oracle
from gary_amd import _lib, build, graph

BIN = build.CONVERT_BIN


def _expected_csr(m, row, col):
    row, col = np.asarray(row, np.int64), np.asarray(col, np.int64)
    order = np.lexsort((col, row))
    ptr = np.zeros(m + 1, np.int64)
    np.add.at(ptr, row + 1, 1)
    return np.cumsum(ptr), col[order].astype(np.int32)


def _run(*args):
    r = subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def test_convert_binary_built():
    assert BIN.exists(), "gary_amd/gary_convert is built by __graft_entry__.build()"


def test_coo_to_csr_known_answer():
    # duplicates (2,0) twice and a loop (1,1) stay; rows sorted; vertex 3 empty
    g = graph.coo_to_csr(4, [2, 0, 1, 2, 0, 2], [0, 3, 1, 0, 1, 1])
    assert g.indPtr.tolist() == [0, 2, 3, 6, 6]
    assert g.colIdx.tolist() == [1, 3, 1, 0, 0, 1]
    assert g.deg.tolist() == [2, 1, 3, 0]


@pytest.mark.parametrize("m,nnz,seed", [(1, 0, 0), (7, 50, 1), (1000, 20000, 2), (3, 5000, 3)])
def test_coo_to_csr_equals_lexsort(m, nnz, seed):
    rng = np.random.default_rng(seed)
    row = rng.integers(0, m, nnz, dtype=np.int32)
    col = rng.integers(0, m, nnz, dtype=np.int32)
    g = graph.coo_to_csr(m, row, col)
    ptr, idx = _expected_csr(m, row, col)
    assert np.array_equal(g.indPtr, ptr) and np.array_equal(g.colIdx, idx)


def test_coo_to_csr_rejects_out_of_range():
    for row, col in (([0, 3], [1, 1]), ([0, 1], [-1, 1])):
        with pytest.raises(_lib.GaryError) as ei:
            graph.coo_to_csr(3, row, col)
        assert ei.value.code == _lib.GARY_ERR_INVALID


def test_snap2bin_file_bytes(tmp_path):
    snap = tmp_path / "g.txt"
    snap.write_text("# comment\n0 1\n2 2\n\n4 0\n1 0\n")
    for keep, row, col in ((0, [0, 4, 1], [1, 0, 0]), (1, [0, 2, 4, 1], [1, 2, 0, 0])):
        out = tmp_path / f"g{keep}.bin"
        rc, so, se = _run(f"-input={snap}", f"-output={out}", "-convert=snap2bin",
                          f"-keepLoops={keep}")
        assert rc == 0, se
        assert f"5 nodes, {len(row)} edges" in so
        want = (struct.pack("<IIiq", 1, 1, 5, len(row)) +
                struct.pack("<IQ", 0x80000001, len(row)) + struct.pack(f"<{len(row)}i", *row) +
                struct.pack("<IQ", 0x80000002, len(col)) + struct.pack(f"<{len(col)}i", *col))
        assert out.read_bytes() == want


def test_coo2csr_symmetrize_csr2csc_orient(tmp_path):
    rng = np.random.default_rng(11)
    m, nnz = 300, 4000
    row = rng.integers(0, m, nnz, dtype=np.int32)
    col = rng.integers(0, m, nnz, dtype=np.int32)
    coo = tmp_path / "coo.bin"
    graph.save_to_binary((m, row, col), coo)
    for sym in (0, 1):
        out = tmp_path / f"csr{sym}.bin"
        rc, so, se = _run(f"--input={coo}", f"--output={out}", "--convert=coo2csr",
                          f"--symmetrize={sym}")
        assert rc == 0, se
        r, c = (np.concatenate([row, col]), np.concatenate([col, row])) if sym else (row, col)
        ptr, idx = _expected_csr(m, r, c)
        g = graph.load_from_binary(out)
        assert np.array_equal(g.indPtr, ptr) and np.array_equal(g.colIdx, idx)
        assert f"d_max={int(np.diff(ptr).max())}" in so
    csr = tmp_path / "csr0.bin"
    g = graph.load_from_binary(csr)
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(g.indPtr))
    # transpose: the CSR of the swapped tuples; twice gives back the sorted original
    rc, _, se = _run(f"-input={csr}", f"-output={tmp_path / 'csc.bin'}", "-convert=csr2csc")
    assert rc == 0, se
    t = graph.load_from_binary(tmp_path / "csc.bin")
    ptr, idx = _expected_csr(m, g.colIdx, rows)
    assert np.array_equal(t.indPtr, ptr) and np.array_equal(t.colIdx, idx)
    rc, _, se = _run(f"-input={tmp_path / 'csc.bin'}", f"-output={tmp_path / 'cc.bin'}",
                     "-convert=csr2csc")
    assert rc == 0, se
    tt = graph.load_from_binary(tmp_path / "cc.bin")
    assert np.array_equal(tt.indPtr, g.indPtr) and np.array_equal(tt.colIdx, g.colIdx)
    # orient: the arcs u -> v with u > v, duplicates kept
    rc, _, se = _run(f"-input={csr}", f"-output={tmp_path / 'o.bin'}", "-convert=orient")
    assert rc == 0, se
    o = graph.load_from_binary(tmp_path / "o.bin")
    keep = rows > g.colIdx
    ptr, idx = _expected_csr(m, rows[keep], g.colIdx[keep])
    assert np.array_equal(o.indPtr, ptr) and np.array_equal(o.colIdx, idx)
    assert np.array_equal(o.deg, np.diff(ptr).astype(np.int32))


def test_snap_to_oriented_chain_counts_triangles(tmp_path, golden_graphs):
    """SNAP text -> snap2bin -> coo2csr -symmetrize=1 -> orient, then the oracle's triangle count
    of the file equals networkx's, for every golden graph listed with each edge once."""
    done = 0
    for rec in golden_graphs:
        s, d = np.array(rec["src"], np.int64), np.array(rec["dst"], np.int64)
        pairs = {(min(a, b), max(a, b)) for a, b in zip(s.tolist(), d.tolist()) if a != b}
        if len(pairs) != len(s) or rec["n"] < 1:
            continue  # loops or repeated edges: the reference's chain keeps duplicates
        snap = tmp_path / "g.txt"
        snap.write_text("".join(f"{a}\t{b}\n" for a, b in zip(s.tolist(), d.tolist())))
        steps = (("snap2bin", snap, "coo.bin", []), ("coo2csr", "coo.bin", "sym.bin",
                                                    ["-symmetrize=1"]),
                 ("orient", "sym.bin", "dag.bin", []))
        for conv, src, dst, extra in steps:
            src = src if conv == "snap2bin" else tmp_path / src
            rc, _, se = _run(f"-input={src}", f"-output={tmp_path / dst}", f"-convert={conv}",
                             *extra)
            assert rc == 0, (rec["name"], se)
        g = graph.load_from_binary(tmp_path / "dag.bin")
        assert g.nnz == len(s), rec["name"]
        assert oracle.tri_count(g.indPtr, g.colIdx) == rec["triangles"], rec["name"]
        done += 1
    assert done >= 5


def test_convert_errors(tmp_path):
    coo = tmp_path / "coo.bin"
    graph.save_to_binary((3, [0, 1], [1, 2]), coo)
    rc, _, se = _run(f"-input={coo}", f"-output={tmp_path / 'x.bin'}", "-convert=orient")
    assert rc == 1 and "not a binary GraphCSR file" in se
    rc, _, se = _run(f"-input={coo}", f"-output={tmp_path / 'x.bin'}", "-convert=bogus")
    assert rc == 1 and "-convert must be one of" in se
    rc, _, se = _run(f"-output={tmp_path / 'x.bin'}", "-convert=coo2csr")
    assert rc == 1 and "missing -input" in se
    rc, _, se = _run(f"-input={tmp_path / 'none.txt'}", f"-output={tmp_path / 'x.bin'}",
                     "-convert=snap2bin")
    assert rc == 1 and "none.txt" in se
