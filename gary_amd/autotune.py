"""Autotuner: sweeps the launch knobs of the counting kernels on this GPU and persists the
fastest exact configuration per workload.

    python -m gary_amd.autotune [--workloads rmat20,batch2000,c3,er4096] [--out gary_amd/tuned.json]

Knobs (gary_launch_cfg):
- pivot path (bitmap, heuristic, learned, batched): `block` = threads of the main pivot kernel's
  workgroup (128/256/512, i.e. its waves), `grid`, `bin_small` (pivots up to this out-degree
  run one per wave with a per-wave LDS set) and `bin_hub` (pivots above it get the 8-wave hub
  kernel with a 32 KiB set);
- edge-centric strategies (cached_bsearch, merge, bsearch): `block`, `group` (lanes per arc),
  `cache_depth` (LDS index cache entries per group), `partition` and `grid`;
- the heuristic's route thresholds (`heuristic_table`, gary_selector_set_heuristic): the log2
  out-degree from which a pivot's arcs go to the bitmap, the log2 degree below which both lists
  count as small (merge), and the log2 length ratio from which the cached binary search is used.
The search is coordinate descent from the library defaults (two passes over the knobs); every
candidate's count must equal the default configuration's count, or the tuner stops with an
error. The result is a JSON file keyed "<workload>/<strategy>"; `cfg_for` reads it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import graph, graph_query as gq
from ._lib import GaryError, LaunchCfg
from .device import DeviceBuffer, DeviceGraphCtx
from .timing import EventTimer

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"

FIELDS = ("block", "grid", "group", "cache_depth", "partition", "bin_small", "bin_hub")
PIVOT_SPACE = {"block": [128, 256, 512], "grid": [1024, 2048, 4096, 8192],
               "bin_small": [16, 32, 64, 128, 256], "bin_hub": [256, 512, 1024, 2048]}
EDGE_SPACE = {"block": [128, 256], "group": [8, 16, 32, 64], "cache_depth": [16, 32, 64, 128],
              "partition": [gq.INTERLEAVED, gq.CHUNKED, gq.DYNAMIC], "grid": [0, 3072, 6144, 12288]}
PIVOT_STRATEGIES = ("bitmap", "heuristic", "learned")
HEURISTIC_SPACE = {"bitmap_from": list(range(0, 11)) + [16], "small_max": [2, 3, 4, 5, 6, 7],
                   "skew": [1, 2, 3, 4, 5, 6]}


def to_cfg(d: dict | None) -> LaunchCfg | None:
    if not d:
        return None
    return gq.launch_cfg(**{k: int(d.get(k, 0)) for k in FIELDS})


def cfg_for(workload: str, strategy: str, path: Path = TUNED_PATH) -> LaunchCfg | None:
    """The tuned configuration of `strategy` on `workload`, or None (library defaults). The
    routed strategies (heuristic, learned) use the bitmap entry's pivot knobs when they have no
    entry of their own."""
    p = Path(path)
    if not p.exists():
        return None
    ent = json.loads(p.read_text()).get("entries", {})
    rec = ent.get(f"{workload}/{strategy}")
    if rec is None and strategy in PIVOT_STRATEGIES:
        rec = ent.get(f"{workload}/bitmap")
    return to_cfg(rec["cfg"]) if rec else None


def log(msg: str) -> None:
    print(f"[autotune] {msg}", flush=True)


def _descend(measure, space: dict, start: dict, passes: int = 2) -> tuple[dict, float, int]:
    """Coordinate descent: each knob in turn takes its best value with the others held."""
    best = dict(start)
    best_t = measure(best)
    evals = 1
    for _ in range(passes):
        changed = False
        for k, values in space.items():
            for v in values:
                if v == best.get(k):
                    continue
                cand = dict(best)
                cand[k] = v
                t = measure(cand)
                evals += 1
                if t < best_t * 0.99:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return best, best_t, evals


def tune_graph(ctx, strategy: str, repeats: int = 5) -> dict:
    sid = gq.STRATEGIES[strategy]
    d = DeviceBuffer(8)
    timer = EventTimer()

    def run(cfg: LaunchCfg | None, reps: int) -> tuple[float, int]:
        d.zero()
        gq.triangleCountHomAsync(ctx, d, sid, cfg)  # warm-up, and the count check
        gq.synchronize()
        got = int(d.to_array(np.uint64, 1)[0])
        timer.start()
        for _ in range(reps):
            gq.triangleCountHomAsync(ctx, d, sid, cfg)
        return timer.stop() / reps, got

    t_def, want = run(None, repeats)

    def measure(c: dict) -> float:
        try:
            t, got = run(to_cfg(c), repeats)
        except GaryError:
            return float("inf")  # outside the kernel's limits (e.g. bin_small for the waves)
        if got != want:
            raise SystemExit(f"autotune: {strategy} cfg {c} counted {got} != {want}")
        return t

    space = PIVOT_SPACE if strategy in PIVOT_STRATEGIES else EDGE_SPACE
    start = {"block": 256, "grid": 2048, "bin_small": 128, "bin_hub": 2048} \
        if strategy in PIVOT_STRATEGIES else \
        {"block": 256, "group": 16 if strategy == "merge" else 64, "cache_depth": 64,
         "partition": gq.INTERLEAVED, "grid": 0}
    best, t, evals = _descend(measure, space, start)
    return {"cfg": best, "ms": t, "default_ms": t_def, "evals": evals, "count": want}


def heuristic_for(workload: str, path: Path = TUNED_PATH) -> dict | None:
    """The tuned heuristic thresholds of `workload` ({bitmap_from, small_max, skew}), or None."""
    p = Path(path)
    if not p.exists():
        return None
    rec = json.loads(p.read_text()).get("entries", {}).get(f"{workload}/heuristic_table")
    return dict(rec["thresholds"]) if rec else None


def tune_heuristic(ctx, repeats: int = 5, cfg: LaunchCfg | None = None) -> dict:
    """Coordinate descent over the heuristic's route thresholds (counts must not change); the
    library defaults are restored afterwards."""
    from . import selector
    d = DeviceBuffer(8)
    timer = EventTimer()

    def run(th: dict, reps: int) -> tuple[float, int]:
        selector.set_heuristic(**th)
        d.zero()
        gq.triangleCountHomAsync(ctx, d, gq.HEURISTIC, cfg)
        gq.synchronize()
        got = int(d.to_array(np.uint64, 1)[0])
        timer.start()
        for _ in range(reps):
            gq.triangleCountHomAsync(ctx, d, gq.HEURISTIC, cfg)
        return timer.stop() / reps, got

    try:
        t_def, want = run(selector.HEURISTIC_DEFAULT, repeats)

        def measure(th: dict) -> float:
            t, got = run(th, repeats)
            if got != want:
                raise SystemExit(f"autotune: heuristic thresholds {th} counted {got} != {want}")
            return t

        best, t, evals = _descend(measure, HEURISTIC_SPACE, dict(selector.HEURISTIC_DEFAULT))
    finally:
        selector.set_heuristic(**selector.HEURISTIC_DEFAULT)
    return {"thresholds": best, "ms": t, "default_ms": t_def, "evals": evals, "count": want}


def tune_batch(b, repeats: int = 5) -> dict:
    timer = EventTimer()
    b.set_cfg(None)
    want = b.count().copy()

    def run(reps: int) -> float:
        b.launch(use_graph=True)
        gq.synchronize()
        timer.start()
        for _ in range(reps):
            b.launch(use_graph=True)
        return timer.stop() / reps

    t_def = run(repeats)

    def measure(c: dict) -> float:
        try:
            b.set_cfg(to_cfg(c))
        except GaryError:
            return float("inf")
        t = run(repeats)
        got = b.counts()
        if not np.array_equal(got, want):
            raise SystemExit(f"autotune: batch cfg {c} changed per-graph counts")
        return t

    best, t, evals = _descend(measure, PIVOT_SPACE,
                              {"block": 256, "grid": 2048, "bin_small": 128, "bin_hub": 2048})
    b.set_cfg(None)
    return {"cfg": best, "ms": t, "default_ms": t_def, "evals": evals}


def _graph_for(name: str):
    if name == "er4096":
        return graph.oriented(*graph.er_4096())
    if name == "c3":
        return graph.oriented(*graph.chung_lu_4m())
    if name.startswith("rmat"):
        return graph.oriented(*graph.rmat(int(name[4:])))
    raise ValueError(name)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="rmat20,batch2000,c3,er4096")
    ap.add_argument("--strategies", default="bitmap,cached_bsearch,merge,heuristic_table")
    ap.add_argument("--out", default=str(TUNED_PATH))
    ap.add_argument("--repeats", type=int, default=5)
    a = ap.parse_args(argv)
    out = Path(a.out)
    doc = json.loads(out.read_text()) if out.exists() else {"entries": {}}
    doc["arch"] = "gfx950"
    from . import selector
    selector.load()
    for wl in a.workloads.split(","):
        t0 = time.time()
        if wl.startswith("batch"):
            from . import batch
            g, vbase = batch.build_union(batch.batch_specs(int(wl[5:] or 2000)))
            b = batch.Batch(g, vbase)
            rec = tune_batch(b, a.repeats)
            b.close()
            doc["entries"][f"{wl}/bitmap"] = rec
            log(f"{wl}/bitmap: {rec['default_ms']:.3f} -> {rec['ms']:.3f} ms {rec['cfg']} "
                f"({rec['evals']} configs, {time.time() - t0:.1f}s)")
            continue
        g = _graph_for(wl)
        ctx = DeviceGraphCtx(g)
        ctx.populateDevice()
        for st in a.strategies.split(","):
            t1 = time.time()
            if st == "heuristic_table":
                rec = tune_heuristic(ctx, a.repeats, cfg_for(wl, "heuristic"))
                doc["entries"][f"{wl}/{st}"] = rec
                log(f"{wl}/{st}: {rec['default_ms']:.3f} -> {rec['ms']:.3f} ms "
                    f"{rec['thresholds']} ({rec['evals']} tables, {time.time() - t1:.1f}s)")
                continue
            rec = tune_graph(ctx, st, a.repeats)
            doc["entries"][f"{wl}/{st}"] = rec
            log(f"{wl}/{st}: {rec['default_ms']:.3f} -> {rec['ms']:.3f} ms {rec['cfg']} "
                f"({rec['evals']} configs, {time.time() - t1:.1f}s)")
        del ctx
    out.write_text(json.dumps(doc, indent=1))
    log(f"saved {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
