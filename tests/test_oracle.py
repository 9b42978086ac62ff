"""Pins the CPU oracle against independent answers: numpy, networkx/brute-force fixtures and
closed forms. No GPU."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


def _sorted_unique(rng, n, lo, hi):
    return np.unique(rng.integers(lo, hi, size=n)).astype(np.int32)


# ---- search (Search.cuh:33-138) --------------------------------------------------------------

@pytest.mark.parametrize("size", list(range(1, 70)) + [97, 128, 129, 200])
def test_binary_search_exhaustive(size):
    rng = np.random.default_rng(size)
    lst = np.sort(rng.choice(np.arange(-5, 3 * size + 5), size=size, replace=False)).astype(np.int32)
    for key in range(int(lst[0]) - 1, int(lst[-1]) + 2):
        want = bool(np.isin(key, lst))
        assert oracle.binary_search(lst, key) == want
        found, pos = oracle.binary_search_pos(lst, key)
        assert found == want
        if found:
            assert lst[pos] == key
        else:  # index of the last element < key
            assert pos == int(np.searchsorted(lst, key, side="left")) - 1


@pytest.mark.parametrize("n_cache", [1, 2, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("size", [1, 2, 3, 7, 31, 32, 33, 63, 64, 65, 100, 200, 1000])
def test_binary_search_2phase_exhaustive(n_cache, size):
    rng = np.random.default_rng(1000 * n_cache + size)
    lst = np.sort(rng.choice(np.arange(0, 4 * size + 8), size=size, replace=False)).astype(np.int32)
    cache = oracle.build_cache(lst, n_cache)
    # cache[i] = list[i*size/N] (Search.cuh:93-100): closed form, 64-bit products
    idx = (np.arange(n_cache, dtype=np.int64) * size) // n_cache
    assert np.array_equal(cache, lst[idx])
    keys = np.arange(int(lst[0]) - 1, int(lst[-1]) + 2)
    want = np.isin(keys, lst)
    got = np.array([oracle.binary_search_2phase(lst, cache, k) for k in keys])
    assert np.array_equal(got, want)


def test_search_reference_shape_seed0():
    # the shape of test/SearchTest.cu:196-210 (4 Mi sorted ints in [2^29, 2^30], 32 Ki keys),
    # scaled down; membership must equal std::binary_search == numpy isin.
    rng = np.random.default_rng(0)
    lst = np.sort(rng.integers(2**29, 2**30, size=1 << 16)).astype(np.int32)
    lst = np.unique(lst)
    keys = rng.integers(2**29, 2**30, size=4096).astype(np.int32)
    keys[:1000] = rng.choice(lst, 1000)
    cache = oracle.build_cache(lst, 32)
    want = np.isin(keys, lst)
    assert np.array_equal([oracle.binary_search(lst, k) for k in keys], want)
    assert np.array_equal([oracle.binary_search_2phase(lst, cache, k) for k in keys], want)


# ---- intersection (Intersection.cuh) ---------------------------------------------------------

def test_intersections_random():
    rng = np.random.default_rng(5)
    for trial in range(300):
        na, nb = rng.integers(0, 300, size=2)
        hi = int(rng.integers(1, 1000))
        a = _sorted_unique(rng, na, 0, hi)
        b = _sorted_unique(rng, nb, 0, hi)
        want = np.intersect1d(a, b)
        assert oracle.intersect_count(a, b) == want.size
        assert np.array_equal(oracle.intersect_emit(a, b), want)
        for n_cache in (0, 8, 32, 64):
            tot, part = oracle.intersect_bsearch_lanes(a, b, n_cache, lanes=32)
            assert tot == want.size and int(part.sum()) == want.size


def test_intersections_exhaustive_tiny():
    universe = range(6)
    subsets = [np.array(c, np.int32) for k in range(0, 5) for c in itertools.combinations(universe, k)]
    for a in subsets[::3]:
        for b in subsets:
            want = np.intersect1d(a, b)
            assert oracle.intersect_count(a, b) == want.size
            assert oracle.intersect_bsearch_lanes(a, b, 8, 4)[0] == want.size


def test_intersect_lane_partials_follow_shorter_list():
    # per-lane partials: lanes stride over the shorter list (Intersection.cuh:130)
    a = np.arange(0, 100, 2, dtype=np.int32)     # 50 items
    b = np.arange(0, 1000, 1, dtype=np.int32)    # superset
    tot, part = oracle.intersect_bsearch_lanes(b, a, 0, lanes=32)  # swapped internally
    assert tot == 50
    assert [int(x) for x in part] == [2 if l < 18 else 1 for l in range(32)]


# ---- CSR build (Graph.cpp:19-91) ---------------------------------------------------------------

def test_csr_build_matches_networkx(golden_graphs):
    for rec in golden_graphs:
        n = rec["n"]
        src, dst = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        indPtr, colIdx = oracle.csr_from_edges(n, src, dst, mode=1)
        got = sorted([u, int(v)] for u in range(n) for v in colIdx[indPtr[u]:indPtr[u + 1]] if u < v)
        assert got == rec["edges_sorted"], rec["name"]
        deg_full = np.diff(indPtr)
        for mode in (0, 2):
            ip, ci = oracle.csr_from_edges(n, src, dst, mode=mode)
            assert ip[-1] == rec["m"]
            for u in range(n):
                row = ci[ip[u]:ip[u + 1]]
                assert np.all(np.diff(row) > 0)
                for v in row:
                    if mode == 0:
                        assert (deg_full[u], u) < (deg_full[v], v)
                    else:
                        assert u > v


def test_tri_count_golden(golden_graphs):
    for rec in golden_graphs:
        n = rec["n"]
        src, dst = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        for mode in (0, 2):  # orientation invariance (degree vs id)
            indPtr, colIdx = oracle.csr_from_edges(n, src, dst, mode=mode)
            assert oracle.tri_count(indPtr, colIdx) == rec["triangles"], rec["name"]


def test_tri_count_closed_forms():
    for k in (3, 10, 50):
        s, d = zip(*itertools.combinations(range(k), 2))
        ip, ci = oracle.csr_from_edges(k, np.array(s), np.array(d), 0)
        assert oracle.tri_count(ip, ci) == k * (k - 1) * (k - 2) // 6
    # star and bipartite: no triangles
    ip, ci = oracle.csr_from_edges(101, np.zeros(100, np.int32), np.arange(1, 101), 0)
    assert oracle.tri_count(ip, ci) == 0


def test_hom4_golden(golden_graphs):
    for rec in golden_graphs:
        n = rec["n"]
        ip, ci = oracle.csr_from_edges(n, np.array(rec["src"], np.int32),
                                       np.array(rec["dst"], np.int32), 1)
        h = oracle.hom4(ip, ci)
        assert h["tri"] == rec["triangles"], rec["name"]
        assert h["k4"] == rec["k4"], rec["name"]
        assert h["hom_k4"] == rec["hom_k4"], rec["name"]
        assert h["hom_diamond"] == rec["hom_diamond"], rec["name"]
        assert h["hom_c4"] == rec["hom_c4"], rec["name"]
        assert h["c4"] == rec["c4"], rec["name"]


def test_hom4_ranked_equals_trace_oracle(golden_graphs):
    """The orientation/wedge restatement against the trace-based one (pinned above by brute
    force and networkx) on every golden graph and on seeded ER / R-MAT graphs."""
    from gary_amd import graph
    cases = [(rec["n"], np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32))
             for rec in golden_graphs]
    cases.append(graph.er_4096())
    for sc in (8, 11, 13):
        cases.append(graph.rmat(sc, 16, seed=77 + sc))
    for n, s, d in cases:
        g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
        assert oracle.hom4_ranked(g.indPtr, g.colIdx) == oracle.hom4(g.indPtr, g.colIdx), n


def test_hom4_ranked_rejects_unranked():
    # a star whose centre has the smallest id: degrees not non-decreasing in id
    ip, ci = oracle.csr_from_edges(6, np.zeros(5, np.int32), np.arange(1, 6), 1)
    with pytest.raises(ValueError):
        oracle.hom4_ranked(ip, ci)


def test_hom4_full_fixture_triangles():
    """The full-size hom4 fixture's triangle counts equal the triangle oracle's (a different
    code path) on the same seeded graphs; only scale 20 here, to keep the CPU suite short."""
    import json
    from gary_amd import graph
    path = GOLDEN / "hom4_full.json"
    if not path.exists():
        pytest.skip("hom4_full.json not generated")
    doc = json.loads(path.read_text())
    rec = doc["rmat20"]
    n, s, d = graph.rmat(20, 16, seed=rec["seed"])
    g = graph.oriented(n, s, d)
    assert oracle.tri_count(g.indPtr, g.colIdx) == rec["tri"]
    sym = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
    assert sym.nnz == rec["nnz"]
    assert rec["hom_c4"] >= 8 * rec["c4"] and rec["hom_k4"] == 24 * rec["k4"]


def test_golden_brute_force_consistency(golden_graphs):
    # fixtures with brute-force hom counts must agree with their closed forms
    n_brute = 0
    for rec in golden_graphs:
        if "brute_hom_c4" in rec:
            n_brute += 1
            assert rec["brute_hom_k4"] == rec["hom_k4"]
            assert rec["brute_hom_diamond"] == rec["hom_diamond"]
            assert rec["brute_hom_c4"] == rec["hom_c4"]
    assert n_brute >= 8


def test_oracle_plants_fail():
    # a planted mistake (dropping one element of a list) must change the count
    indPtr, colIdx = oracle.csr_from_edges(5, np.array([0, 0, 1, 0, 1, 2]), np.array([1, 2, 2, 3, 3, 3]), 0)
    t = oracle.tri_count(indPtr, colIdx)
    assert t == 4
    ci = colIdx.copy()
    ci[2] = 4  # replace arc 0->3 by 0->4 (vertex 4 has no arcs): two triangles disappear
    assert oracle.tri_count(indPtr, ci) == t - 2


# ---- triangle listing and per-range counts -----------------------------------------------------

@pytest.mark.parametrize("mode", [0, 2])
def test_tri_list_matches_networkx_sets(golden_graphs, mode):
    """or_tri_list (the emit variants of include/ops/Intersection.cuh:18-50,76-112 applied to the
    triangle query) lists each triangle of the networkx fixture exactly once, as (u, v, w) with
    v in N+(u) and w in N+(u) and N+(v)."""
    for rec in golden_graphs:
        n = rec["n"]
        src, dst = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        ip, ci = oracle.csr_from_edges(n, src, dst, mode=mode)
        t = oracle.tri_list(ip, ci)
        assert t.shape == (rec["triangles"], 3), rec["name"]
        got = sorted(sorted(r) for r in t.tolist())
        assert got == rec["triangle_list"], rec["name"]
        arcs = {(u, int(v)) for u in range(n) for v in ci[ip[u]:ip[u + 1]]}
        for u, v, w in t.tolist():
            assert (u, v) in arcs and (u, w) in arcs and (v, w) in arcs, rec["name"]


@pytest.mark.parametrize("mode", [0, 2])
def test_tri_list_checksum_matches_networkx_sets(golden_graphs, mode):
    """The order-independent checksum equals the sum of or_triple_hash over the fixture's
    triangles, each put in orientation order (lowest, middle, highest): with degree orientation
    by (degree, id), with id orientation by decreasing id (u -> v iff u > v)."""
    for rec in golden_graphs:
        n = rec["n"]
        src, dst = np.array(rec["src"], np.int32), np.array(rec["dst"], np.int32)
        ip, ci = oracle.csr_from_edges(n, src, dst, mode=mode)
        deg = np.zeros(n, np.int64)
        for a, b in rec["edges_sorted"]:
            deg[a] += 1
            deg[b] += 1
        key = (lambda v: (int(deg[v]), v)) if mode == 0 else (lambda v: -v)
        want = 0
        for tri in rec["triangle_list"]:
            u, v, w = sorted(tri, key=key)
            want = (want + int(oracle.lib.or_triple_hash(u, v, w))) % (1 << 64)
        assert oracle.tri_list_checksum(ip, ci) == want, rec["name"]


def test_tri_list_checksum_detects_a_changed_triple():
    """The checksum is not blind to which triangles are listed: moving one arc of K4 changes it."""
    ip, ci = oracle.csr_from_edges(5, np.array([0, 0, 0, 1, 1, 2]), np.array([1, 2, 3, 2, 3, 3]), 2)
    c0 = oracle.tri_list_checksum(ip, ci)
    ip2, ci2 = oracle.csr_from_edges(5, np.array([0, 0, 0, 1, 1, 2, 4, 4]),
                                     np.array([1, 2, 3, 2, 3, 3, 0, 1]), 2)
    assert oracle.tri_list_checksum(ip2, ci2) != c0


def test_tri_count_range_partitions_and_unions(golden_graphs):
    """or_tri_count_range counts the triangles whose lowest vertex lies in [vbegin, vend): the
    counts of any partition of the vertex range add up to the whole count, and on the disjoint
    union of the golden graphs (graph g on ids [base_g, base_g + n_g)) the range of each graph
    gives that graph's networkx count."""
    srcs, dsts, bases, base = [], [], [], 0
    for rec in golden_graphs:
        bases.append(base)
        srcs.append(np.array(rec["src"], np.int64) + base)
        dsts.append(np.array(rec["dst"], np.int64) + base)
        base += rec["n"]
    bases.append(base)
    ip, ci = oracle.csr_from_edges(base, np.concatenate(srcs).astype(np.int32),
                                   np.concatenate(dsts).astype(np.int32), 0)
    total = oracle.tri_count(ip, ci)
    assert total == sum(r["triangles"] for r in golden_graphs)
    for k, rec in enumerate(golden_graphs):
        assert oracle.tri_count_range(ip, ci, bases[k], bases[k + 1]) == rec["triangles"], \
            rec["name"]
    rng = np.random.default_rng(5)
    for _ in range(5):
        cuts = np.sort(rng.integers(0, base + 1, 7))
        edges = [0, *cuts.tolist(), base]
        parts = [oracle.tri_count_range(ip, ci, a, b) for a, b in zip(edges, edges[1:])]
        assert sum(parts) == total
    assert oracle.tri_count_range(ip, ci, 7, 7) == 0
    # the batch form (parallel over graphs, each graph sequential) gives every graph's count
    per = oracle.tri_count_batch(np.array(bases, np.int32), ip, ci)
    assert per.tolist() == [r["triangles"] for r in golden_graphs]


# ---- k-cliques (no reference counterpart; pinned by networkx, closed forms, brute force) -------

def _kc_fixture():
    import json
    return json.loads((GOLDEN / "kcliques.json").read_text())


def test_kclique_oracle_matches_networkx_fixture():
    for rec in _kc_fixture():
        n = rec["n"]
        ip, ci = oracle.csr_from_edges(n, np.array(rec["src"], np.int32),
                                       np.array(rec["dst"], np.int32), 0)
        got = [oracle.kclique_count(ip, ci, k) for k in rec["k"]]
        assert got == rec["count"], rec["name"]


def test_kclique_oracle_closed_forms():
    from math import comb

    from gary_amd import graph
    # complete t-partite graph with parts of size s: #K_j = C(t, j) * s^j
    for t, s in ((5, 3), (7, 2), (4, 4)):
        part = np.repeat(np.arange(t), s)
        n = t * s
        src, dst = [], []
        for a in range(n):
            for b in range(a + 1, n):
                if part[a] != part[b]:
                    src.append(a)
                    dst.append(b)
        ip, ci = oracle.csr_from_edges(n, np.array(src, np.int32), np.array(dst, np.int32), 0)
        for j in range(1, t + 2):
            assert oracle.kclique_count(ip, ci, j) == comb(t, j) * s ** j, (t, s, j)
    # id orientation gives the same counts as degree orientation
    n, s, d = graph.rmat(9, 8, seed=4)
    a = oracle.csr_from_edges(n, s, d, 0)
    b = oracle.csr_from_edges(n, s, d, 2)
    for k in range(3, 7):
        assert oracle.kclique_count(*a, k) == oracle.kclique_count(*b, k)


def test_kclique_oracle_brute_force():
    import itertools
    rng = np.random.default_rng(5)
    for _ in range(6):
        n = 9
        adj = np.triu(rng.random((n, n)) < 0.6, 1)
        src, dst = np.nonzero(adj)
        A = adj | adj.T
        ip, ci = oracle.csr_from_edges(n, src.astype(np.int32), dst.astype(np.int32), 0)
        for k in range(3, 8):
            want = sum(all(A[a, b] for a, b in itertools.combinations(q, 2))
                       for q in itertools.combinations(range(n), k))
            assert oracle.kclique_count(ip, ci, k) == want
