"""Trains the learned strategy selector on timings measured on this GPU and saves it.

    python -m gary_amd.train_selector [--out gary_amd/selector_model.json] [--epochs 4000]

Training graphs (default): Erdős–Rényi G(4096, 65536) and R-MAT scales 18 and 20 (edgefactor
16), degree-oriented. Edge routes (cached binary search, merge) are timed bin by bin against the
all-skip table; the bitmap route bin by bin as the saving of skipping the bin from the all-bitmap
table, its remainder being its fixed cost (selector.time_bins). The per-arc costs train the MLP on
the device; the fixed costs are kept per training graph and interpolated in log2(nnz).
Then the pivot-split threshold cost_big is tuned for the learned table, and the regret against
the exhaustive per-bin best is reported for the learned and the heuristic tables.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import graph, selector
from .device import DeviceGraphCtx


def log(msg: str) -> None:
    print(f"[train] {msg}", flush=True)


def prepare(name: str):
    if name == "er4096":
        n, s, d = graph.er_4096()
    elif name.startswith("rmat"):
        n, s, d = graph.rmat(int(name[4:]))
    else:
        raise ValueError(name)
    g = graph.oriented(n, s, d)
    ctx = DeviceGraphCtx(g)
    ctx.populateDevice()
    return g, ctx


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", default="er4096,rmat18,rmat20")
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
    selector.set_params(params, trained=True, cost_big=0.0)
    selector.set_fixed_costs(fx, fc)

    report = {}
    best_cost_big = 0.0
    for nm, (g, ctx, arcs, ns) in data.items():
        heur = selector.heuristic_table()
        learned, costs = selector.learned_table(ctx)
        t_h = selector.time_table(ctx, heur, repeats=5)
        t_b = selector.time_table(ctx, np.zeros(256, np.uint8), repeats=5)
        sweep = {}
        for cb in (0.0, 5.0, 20.0, 80.0, 320.0, 1280.0):
            sweep[cb] = selector.time_table(ctx, learned, costs, cost_big=cb, repeats=5)
        cb_best = min(sweep, key=sweep.get)
        if nm == names[-1]:
            best_cost_big = cb_best
        report[nm] = {
            "m": int(g.nnz),
            "ms_heuristic": t_h, "ms_bitmap_only": t_b, "ms_learned": sweep[cb_best],
            "ms_learned_by_cost_big": {str(k): v for k, v in sweep.items()},
            "regret_learned": selector.regret(ns, arcs, learned),
            "regret_heuristic": selector.regret(ns, arcs, heur),
            "learned_routes": {selector.ROUTE_NAMES[r]: int(arcs[learned == r].sum())
                               for r in (0, 1, 2)},
            "heuristic_routes": {selector.ROUTE_NAMES[r]: int(arcs[heur == r].sum())
                                 for r in (0, 1, 2)},
        }
        log(f"{nm}: heuristic {t_h:.3f} ms, bitmap-only {t_b:.3f} ms, learned {sweep[cb_best]:.3f} ms "
            f"(cost_big={cb_best}), regret learned {report[nm]['regret_learned']:.3f} "
            f"heuristic {report[nm]['regret_heuristic']:.3f}")
    selector.set_params(params, trained=True, cost_big=best_cost_big)
    selector.set_fixed_costs(fx, fc)
    selector.save(a.out, fixed_nnz=fx, fixed_costs=fc, cost_big=best_cost_big, allowed=7, loss=loss,
                  loss_initial=loss0,
                  samples=int(X.shape[0]), graphs=names, epochs=a.epochs, lr=a.lr,
                  report=report)
    log(f"saved {a.out}")
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
