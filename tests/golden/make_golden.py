"""Writes tests/golden/graphs.json: small graphs with independently computed answers.

Answers come from networkx (triangles, the triangle vertex sets, clique enumeration), numpy
integer matrix powers (trace(A^3)/6, trace(A^4)) and brute force over all vertex maps (homomorphism counts, n <= 8).
None of them uses this repository's code, so the fixtures pin the oracle and the GPU path.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import json
import random
from pathlib import Path

import networkx as nx
import numpy as np

OUT = Path(__file__).resolve().parent / "graphs.json"

K4 = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
DIAMOND = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3)]  # K4 minus edge (2, 3)
C4 = [(0, 1), (1, 2), (2, 3), (3, 0)]


def brute_hom(pattern, n, adj) -> int:
    cnt = 0
    for phi in itertools.product(range(n), repeat=4):
        if all(adj[phi[a]][phi[b]] for a, b in pattern):
            cnt += 1
    return cnt


def describe(name: str, n: int, edges: list[tuple[int, int]], raw=None, brute: bool = False):
    G = nx.Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from((u, v) for u, v in edges if u != v)
    A = nx.to_numpy_array(G, nodelist=range(n), dtype=np.int64).astype(object)
    A2 = A.dot(A)
    tr3 = int(np.trace(A2.dot(A)))
    tr4 = int(np.trace(A2.dot(A2)))
    tri_nx = sum(nx.triangles(G).values()) // 3
    assert tr3 == 6 * tri_nx, name
    k4 = 0
    for c in nx.enumerate_all_cliques(G):  # yields cliques in non-decreasing size
        if len(c) > 4:
            break
        k4 += len(c) == 4
    # diamond homs: sum over arcs of codeg^2, by networkx common neighbours
    diam = 0
    for u, v in G.edges():
        t = len(list(nx.common_neighbors(G, u, v)))
        diam += 2 * t * t
    # every triangle as its sorted vertex triple, by networkx clique enumeration
    tri_list = []
    for c in nx.enumerate_all_cliques(G):  # non-decreasing size: stop past the triangles
        if len(c) > 3:
            break
        if len(c) == 3:
            tri_list.append(sorted(int(x) for x in c))
    tri_list.sort()
    assert len(tri_list) == tri_nx, name
    m = G.number_of_edges()
    degs = [d for _, d in G.degree()]
    c4 = (tr4 - 4 * sum(d * (d - 1) // 2 for d in degs) - 2 * m)
    assert c4 % 8 == 0, name
    rec = {
        "name": name, "n": n,
        "src": [int(u) for u, _ in (raw or edges)], "dst": [int(v) for _, v in (raw or edges)],
        "m": m, "triangles": tri_nx, "trace_a3": tr3, "k4": k4, "hom_k4": 24 * k4,
        "hom_diamond": diam, "hom_c4": tr4, "c4": c4 // 8,
        "edges_sorted": sorted([min(u, v), max(u, v)] for u, v in G.edges()),
        "triangle_list": tri_list,
    }
    if brute:
        adj = [[bool(A[i][j]) for j in range(n)] for i in range(n)]
        rec["brute_hom_k4"] = brute_hom(K4, n, adj)
        rec["brute_hom_diamond"] = brute_hom(DIAMOND, n, adj)
        rec["brute_hom_c4"] = brute_hom(C4, n, adj)
        assert rec["brute_hom_k4"] == rec["hom_k4"], name
        assert rec["brute_hom_diamond"] == rec["hom_diamond"], name
        assert rec["brute_hom_c4"] == rec["hom_c4"], name
    return rec


def main():
    rng = random.Random(20240601)
    recs = []
    recs.append(describe("empty_graph_8", 8, []))
    recs.append(describe("single_edge", 2, [(0, 1)], brute=True))
    recs.append(describe("triangle", 3, [(0, 1), (1, 2), (2, 0)], brute=True))
    for k in (4, 5, 6, 8):
        recs.append(describe(f"complete_{k}", k, list(itertools.combinations(range(k), 2)),
                             brute=k <= 6))
    recs.append(describe("complete_30", 30, list(itertools.combinations(range(30), 2))))
    recs.append(describe("star_20", 21, [(0, i) for i in range(1, 21)]))
    recs.append(describe("cycle_4", 4, C4, brute=True))
    recs.append(describe("diamond", 4, DIAMOND, brute=True))
    recs.append(describe("bipartite_10_12", 22, [(i, 10 + j) for i in range(10) for j in range(12)
                                                 if rng.random() < 0.5]))
    recs.append(describe("wheel_12", 13, [(0, i) for i in range(1, 13)] +
                         [(i, i % 12 + 1) for i in range(1, 13)]))
    # raw tuple lists with loops and duplicates (exercise symmetrize + dedup + loop removal)
    raw = [(0, 1), (1, 0), (1, 2), (2, 2), (2, 0), (0, 1), (3, 3), (3, 4), (4, 5), (5, 3),
           (5, 4)]
    recs.append(describe("dups_and_loops", 7, [(u, v) for u, v in raw], raw=raw, brute=True))
    for i, (n, p) in enumerate([(7, 0.5), (8, 0.6), (8, 0.3)]):
        e = [(u, v) for u, v in itertools.combinations(range(n), 2) if rng.random() < p]
        recs.append(describe(f"gnp_small_{i}", n, e, brute=True))
    for i, (n, p) in enumerate([(60, 0.2), (120, 0.1), (200, 0.05), (300, 0.08)]):
        e = [(u, v) for u, v in itertools.combinations(range(n), 2) if rng.random() < p]
        rng.shuffle(e)
        e = [(v, u) if rng.random() < 0.5 else (u, v) for u, v in e]
        recs.append(describe(f"gnp_{n}_{p}", n, e))
    # a skewed graph: powerlaw cluster with hubs
    G = nx.powerlaw_cluster_graph(400, 6, 0.4, seed=7)
    recs.append(describe("powerlaw_cluster_400", 400, list(G.edges())))
    G = nx.barabasi_albert_graph(500, 10, seed=11)
    recs.append(describe("barabasi_albert_500", 500, list(G.edges())))
    OUT.write_text(json.dumps(recs, separators=(",", ":")))
    print(f"wrote {len(recs)} graphs to {OUT}")


if __name__ == "__main__":
    main()
