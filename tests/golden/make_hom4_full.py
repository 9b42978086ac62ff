"""Writes hom4_full.json: the oracle's 4-vertex counts (oracle.hom4_ranked) of the full-size
R-MAT graphs, for GPU tests that cannot afford the CPU count at run time.

    python tests/golden/make_hom4_full.py [--scales 20,22]

Graphs: R-MAT(scale, edgefactor 16, a=.57 b=.19 c=.19), seeds as in gary_amd.graph.SEEDS
("rmat20" for scale 20, "rmat22" for scale 22), symmetrized, dedup, no loops, ranked by
(degree, id). hom4_ranked is pinned against the trace-based oracle.hom4 up to scale 16
(tests/test_oracle.py).
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


def seed_for(scale: int) -> int:
    return graph.SEEDS["rmat22"] if scale == 22 else graph.SEEDS["rmat20"]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scales", default="20,22")
    a = ap.parse_args()
    out = HERE / "hom4_full.json"
    doc = json.loads(out.read_text()) if out.exists() else {}
    for sc in map(int, a.scales.split(",")):
        t0 = time.time()
        n, s, d = graph.rmat(sc, 16, seed=seed_for(sc))
        g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
        h = oracle.hom4_ranked(g.indPtr, g.colIdx)
        doc[f"rmat{sc}"] = {"scale": sc, "edgefactor": 16, "seed": seed_for(sc), "n": n,
                            "nnz": g.nnz, **h}
        print(f"rmat{sc}: {h} ({time.time() - t0:.0f}s)", flush=True)
        out.write_text(json.dumps(doc, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
