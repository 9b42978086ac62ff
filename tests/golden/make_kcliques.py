"""Writes tests/golden/kcliques.json: k-clique counts (k = 3..8) from networkx's clique
enumeration for every graph of graphs.json and for a few denser seeded graphs, plus their
tuples. networkx is independent of this repository, so the fixture pins the oracle's
or_kclique_count and the GPU kernel.

Run: python tests/golden/make_kcliques.py
"""
from __future__ import annotations

import json
import random
from pathlib import Path

import networkx as nx

HERE = Path(__file__).resolve().parent
OUT = HERE / "kcliques.json"
KMAX = 8


def counts(n: int, src, dst) -> list[int]:
    G = nx.Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from((u, v) for u, v in zip(src, dst) if u != v)
    c = [0] * (KMAX + 1)
    for q in nx.enumerate_all_cliques(G):
        if len(q) > KMAX:
            break
        c[len(q)] += 1
    return c[3:]


def main() -> None:
    recs = []
    for g in json.loads((HERE / "graphs.json").read_text()):
        recs.append({"name": g["name"], "n": g["n"], "src": g["src"], "dst": g["dst"],
                     "k": list(range(3, KMAX + 1)), "count": counts(g["n"], g["src"], g["dst"])})
    rng = random.Random(0x6A7E00C1)
    for name, n, p in (("gnp_60_0.5", 60, 0.5), ("gnp_120_0.3", 120, 0.3),
                       ("gnp_300_0.12", 300, 0.12)):
        src, dst = [], []
        for u in range(n):
            for v in range(u + 1, n):
                if rng.random() < p:
                    src.append(u)
                    dst.append(v)
        recs.append({"name": name, "n": n, "src": src, "dst": dst, "k": list(range(3, KMAX + 1)),
                     "count": counts(n, src, dst)})
    OUT.write_text(json.dumps(recs))
    print(f"wrote {len(recs)} graphs to {OUT}")


if __name__ == "__main__":
    main()
