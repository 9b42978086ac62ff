"""Trains the learned strategy selector on timings measured on this GPU and saves it.

    python -m gary_amd.train_selector [--out gary_amd/selector_model.json] [--epochs 4000]

Training graphs (default): Erdős–Rényi G(4096, 65536), the union of the first 100 graphs of the
batched workload (small ER and R-MAT graphs, where the per-bin best route is mixed) and R-MAT
scales 18 and 20 (edgefactor 16), degree-oriented. Edge routes (cached binary search, merge) are timed bin by bin against the
all-skip table; the bitmap route bin by bin as the saving of skipping the bin from the all-bitmap
table, its remainder being its fixed cost (selector.time_bins). The per-arc costs train the MLP on
the device; the fixed costs are kept per training graph and interpolated in log2(nnz).
Then the decision margin (the predicted gain needed to leave the all-bitmap routing) and the
pivot-split threshold cost_big are chosen on the training graphs, and the learned, heuristic and
bitmap-only strategies are timed there and on held-out graphs (default: the Chung–Lu graph of
4,194,304 vertices), with the regret of each table against the exhaustive per-bin best.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import graph, graph_query as gq, selector
from .device import DeviceGraphCtx


def log(msg: str) -> None:
    print(f"[train] {msg}", flush=True)


def prepare(name: str):
    if name == "er4096":
        n, s, d = graph.er_4096()
    elif name == "c3":
        n, s, d = graph.chung_lu_4m()
    elif name.startswith("rmat"):
        n, s, d = graph.rmat(int(name[4:]))
    elif name.startswith("batch"):  # the first N graphs of the batched workload, as one union
        from . import batch
        g, _ = batch.build_union(batch.batch_specs(int(name[5:])))
        ctx = DeviceGraphCtx(g)
        ctx.populateDevice()
        return g, ctx
    else:
        raise ValueError(name)
    g = graph.oriented(n, s, d)
    ctx = DeviceGraphCtx(g)
    ctx.populateDevice()
    return g, ctx


def evaluate(nm, g, ctx, ns, arcs, margin, cost_big) -> dict:
    """Times the heuristic, bitmap-only and learned strategies (gary_tc_count, so the learned
    time includes its features and inference) and the regret of both tables."""
    heur = selector.heuristic_table()
    learned, _ = selector.learned_table(ctx)
    rec = {
        "m": int(g.nnz),
        "ms_heuristic": selector.time_strategy(ctx, gq.HEURISTIC),
        "ms_bitmap_only": selector.time_strategy(ctx, gq.BITMAP),
        "ms_learned": selector.time_strategy(ctx, gq.AUTO_LEARNED),
        "learned_routes": {selector.ROUTE_NAMES[r]: int(arcs[learned == r].sum()) for r in (0, 1, 2)},
        "heuristic_routes": {selector.ROUTE_NAMES[r]: int(arcs[heur == r].sum()) for r in (0, 1, 2)},
    }
    if ns is not None:
        rec["regret_learned"] = selector.regret(ns, arcs, learned)
        rec["regret_heuristic"] = selector.regret(ns, arcs, heur)
    rec["learned_vs_heuristic"] = rec["ms_heuristic"] / rec["ms_learned"]
    log(f"{nm}: heuristic {rec['ms_heuristic']:.3f} ms, bitmap-only {rec['ms_bitmap_only']:.3f} ms, "
        f"learned {rec['ms_learned']:.3f} ms (x{rec['learned_vs_heuristic']:.2f} vs heuristic; "
        f"margin {margin}, cost_big {cost_big})"
        + (f", regret learned {rec['regret_learned']:.3f} heuristic {rec['regret_heuristic']:.3f}"
           if ns is not None else ""))
    return rec


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", default="er4096,batch100,rmat18,rmat20")
    ap.add_argument("--eval", default="c3", help="held-out graphs (timed, not trained on)")
    ap.add_argument("--out", default=str(selector.MODEL_PATH))
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--min-arcs", type=int, default=2048)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args(argv)
    names = a.graphs.split(",")
    data = {}
    Xs, Ts, Ms = [], [], []
    fixed_pts = []
    for nm in names:
        t0 = time.time()
        g, ctx = prepare(nm)
        X, _ = selector.bin_features(ctx)
        arcs = selector.exact_bin_arcs(g)
        ns, fixed = selector.time_bins(ctx, arcs, min_arcs=min(a.min_arcs, max(64, g.nnz // 200)),
                                       repeats=a.repeats)
        rows = selector.training_rows(X, ns)
        Xs.append(rows[0]); Ts.append(rows[1]); Ms.append(rows[2])
        data[nm] = (g, ctx, arcs, ns)
        fixed_pts.append((float(g.nnz), fixed))
        log(f"{nm}: m={g.nnz} timed bins={rows[0].shape[0]} fixed ns={np.round(fixed).tolist()} "
            f"({time.time() - t0:.1f}s)")
    X = np.concatenate(Xs); T = np.concatenate(Ts); M = np.concatenate(Ms)
    loss0 = selector.train(X, T, M, epochs=1, lr=a.lr, seed=0x5E1EC7)
    loss = selector.train(X, T, M, epochs=a.epochs, lr=a.lr, seed=0)
    log(f"trained on {X.shape[0]} bins: loss {loss0:.4f} -> {loss:.5f}")
    params = selector.get_params()
    fixed_pts.sort(key=lambda p: p[0])
    fx = [p[0] for p in fixed_pts][:4]
    fc = [p[1] for p in fixed_pts][:4]
    selector.set_fixed_costs(fx, fc)

    # decision margin and pivot-split cost: least Σ log(time) of the learned strategy over the
    # training graphs
    grid = {}
    for margin in (0.0, 0.03, 0.1, 0.2):
        for cb in (0.0, 5.0, 20.0, 80.0):
            selector.set_params(params, trained=True, cost_big=cb, margin=margin)
            grid[(margin, cb)] = float(sum(
                np.log(selector.time_strategy(ctx, gq.AUTO_LEARNED, repeats=3))
                for (_, ctx, _, _) in data.values()))
    margin, cost_big = min(grid, key=grid.get)
    selector.set_params(params, trained=True, cost_big=cost_big, margin=margin)
    log(f"margin {margin}, cost_big {cost_big}")

    report = {nm: evaluate(nm, g, ctx, ns, arcs, margin, cost_big)
              for nm, (g, ctx, arcs, ns) in data.items()}
    held_out = {}
    for nm in [x for x in a.eval.split(",") if x]:
        g, ctx = prepare(nm)
        arcs = selector.exact_bin_arcs(g)
        ns, _ = selector.time_bins(ctx, arcs, min_arcs=max(64, g.nnz // 200), repeats=2)
        held_out[nm] = evaluate(nm, g, ctx, ns, arcs, margin, cost_big)
        del ctx
    selector.save(a.out, fixed_nnz=fx, fixed_costs=fc, cost_big=cost_big, margin=margin, allowed=7,
                  loss=loss, loss_initial=loss0, samples=int(X.shape[0]), graphs=names,
                  epochs=a.epochs, lr=a.lr,
                  grid={f"{k[0]}/{k[1]}": v for k, v in grid.items()},
                  report=report, held_out=held_out)
    log(f"saved {a.out}")
    print(json.dumps({"report": report, "held_out": held_out}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
