"""Writes tri_full.json: the oracle's triangle counts (oracle.tri_count, cpuTriCount restated) of
the full-size single-graph workloads, for GPU tests that cannot afford the CPU count at run time.

    python tests/golden/make_tri_full.py [--graphs rmat20,chung_lu_4m]

Graphs (host generators and host CSR build; the device build is tested equal to it):
  rmat20       R-MAT(20, edgefactor 16, a=.57 b=.19 c=.19), seed SEEDS["rmat20"];
  chung_lu_4m  Chung-Lu, 4,194,304 vertices, 67,108,864 tuples, exponent 2.1, weights capped at
               50,000, seed SEEDS["chung_lu"];
both symmetrized, dedup, no loops, degree-oriented. tri_count itself is pinned against networkx,
trace(A^3)/6 and closed forms on the golden graphs (tests/test_oracle.py).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from gary_amd import graph  # noqa: E402

GRAPHS = {
    "rmat20": lambda: graph.rmat(20, 16, seed=graph.SEEDS["rmat20"]),
    "chung_lu_4m": lambda: graph.chung_lu_4m(),
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", default=",".join(GRAPHS))
    a = ap.parse_args()
    out = HERE / "tri_full.json"
    doc = json.loads(out.read_text()) if out.exists() else {}
    for name in a.graphs.split(","):
        t0 = time.time()
        n, s, d = GRAPHS[name]()
        g = graph.oriented(n, s, d)
        tri = oracle.tri_count(g.indPtr, g.colIdx)
        doc[name] = {"n": n, "tuples": int(s.shape[0]), "m": g.nnz,
                     "max_out_degree": int(g.deg.max()), "triangles": int(tri)}
        print(f"{name}: m={g.nnz} triangles={tri} ({time.time() - t0:.0f}s)", flush=True)
        out.write_text(json.dumps(doc, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
