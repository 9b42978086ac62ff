#!/usr/bin/env python3
"""Exact triangle counting on R-MAT scale 20 (edgefactor 16, a=.57 b=.19 c=.19) on one MI355X.

A step is one full count of the degree-oriented graph already resident in HBM (every device
launch of the counting path, including the per-call pivot binning). Prints one JSON line:
edges/s of the headline strategy, per-strategy times, the HBM roofline fraction of the counting
pass under the merge byte model B_alg = 4*W + 32*m (W = sum over arcs (u,v) of d+(u) + d+(v)),
and the reference's CPU algorithm (cpuTriCount restated in the oracle) timed on this host.
Beside the headline: the learned selector against the plain degree-threshold heuristic on R-MAT
20 and on the Chung-Lu graph (4,194,304 vertices, ~64M edges), and the batched form (2,000
graphs of uneven size, one hipGraph replay per step, per-graph counts). Launch knobs come from
the autotuner's file (gary_amd/tuned.json) where it has an entry for the workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 20] [--no-cpu] [--no-batch]
                    [--no-c3] [--strategy learned|heuristic|bitmap|...] [--only]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--strategy", default="learned")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--only", action="store_true",
                    help="time the --strategy only (for profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="also run one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2,000-graph leg")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the selector leg on the Chung-Lu graph (4,194,304 vertices)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported (each GPU box "
              "here has one GPU; the counting path shards by edge ranges with one u64 sum)",
              file=sys.stderr)
        return 2

    import numpy as np

    from gary_amd import autotune, graph, selector
    from gary_amd import graph_query as gq
    from gary_amd.device import DeviceBuffer, DeviceGraphCtx
    from gary_amd.profiling import read_traffic
    from gary_amd.timing import EventTimer, kernel_timing, kernel_timing_read

    model = selector.load()
    if model is None and a.strategy == "learned":
        raise SystemExit(f"no trained selector at {selector.MODEL_PATH}")
    wl = f"rmat{a.scale}" if a.edgefactor == 16 else None

    t0 = time.time()
    n, s, d = graph.rmat(a.scale, a.edgefactor)
    g = graph.oriented(n, s, d)
    del s, d
    log(f"rmat{a.scale} ef{a.edgefactor}: n={n} m={g.nnz} max_out_deg={int(g.deg.max())} "
        f"host build {time.time() - t0:.2f}s")
    row, col = g.toCOO()
    dp = g.deg.astype(np.int64)
    per_arc = 4 * (dp[row] + dp[col]) + 32  # merge-model bytes of each arc (u -> v)
    B_alg = int(per_arc.sum())
    # arcs whose pivot the dominant (main pivot) kernel handles: out-degree <= 2048 (larger
    # pivots go to the hub kernel)
    B_alg_main = int(per_arc[dp[row] <= 2048].sum())
    del row, col, per_arc

    ctx = DeviceGraphCtx(g)
    ctx.populateDevice()
    cnt = DeviceBuffer(8)
    timer = EventTimer()

    kernel_ms = {}

    def run(strategy: int, steps: int, warmup: int, cfg=None, name=""):
        if cfg is None and wl:
            cfg = autotune.cfg_for(wl, name)
        for _ in range(warmup):
            cnt.zero()
            gq.triangleCountHomAsync(ctx, cnt, strategy, cfg)
        gq.synchronize()
        cnt.zero()
        gq.synchronize()
        kernel_timing(True)
        timer.start()
        for _ in range(steps):
            gq.triangleCountHomAsync(ctx, cnt, strategy, cfg)
        ms = timer.stop() / steps
        gq.synchronize()
        launches, kms = kernel_timing_read()
        kernel_timing(False)
        if launches:
            kernel_ms[name] = kms / launches
        total = int(cnt.to_array(np.uint64, 1)[0])
        assert total % steps == 0
        return ms, total // steps

    names = {"learned": gq.AUTO_LEARNED, "heuristic": gq.HEURISTIC, "bitmap": gq.BITMAP,
             "cached_bsearch": gq.CACHED_BSEARCH, "merge": gq.MERGE, "bsearch": gq.BSEARCH}
    head = names[a.strategy]
    ms, tri = run(head, a.steps, a.warmup, name=a.strategy)
    log(f"{a.strategy}: {ms:.3f} ms/step, triangles={tri}")
    results = {a.strategy: ms}
    for nm, sid in names.items():
        if nm == a.strategy or a.only:
            continue
        k = max(1, min(a.steps, 3)) if nm in ("cached_bsearch", "merge", "bsearch") else a.steps
        ms_s, tri_s = run(sid, k, 1, name=nm)
        if tri_s != tri:
            raise SystemExit(f"strategy {nm} counted {tri_s} != {tri}")
        results[nm] = ms_s
        log(f"{nm}: {ms_s:.3f} ms/step")

    if a.calibrate_copy:
        from gary_amd._lib import check as _check, lib as _glib
        nbytes = a.calibrate_copy << 20
        src_b, dst_b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        src_b.zero()
        _check(_glib.gary_memcpy_d2d(0, dst_b.ptr, src_b.ptr, nbytes))
        gq.synchronize()
        log(f"calibration copy: {nbytes} bytes device to device")

    def pivot_bytes(gx):
        """Σ over arcs (u,v) of 4*(d+(u) + d+(v)) + 32, and the part whose pivot has
        out-degree <= 2048 (the main pivot kernel's share)."""
        dpx = gx.deg.astype(np.int64)
        rowx = np.repeat(np.arange(gx.m, dtype=np.int64), dpx)
        du = dpx[rowx]
        per = 4 * (du + dpx[gx.colIdx]) + 32
        return int(per.sum()), int(per[du <= 2048].sum())

    selector_legs = {}
    if not a.only:
        selector_legs[f"rmat{a.scale}"] = {
            "heuristic_ms": results.get("heuristic"), "learned_ms": results.get("learned"),
            "bitmap_ms": results.get("bitmap")}
    if not a.no_c3 and not a.only:
        t0 = time.time()
        gc = graph.oriented(*graph.chung_lu_4m())
        log(f"chung_lu_4m: m={gc.nnz} host build {time.time() - t0:.2f}s")
        cctx = DeviceGraphCtx(gc)
        cctx.populateDevice()
        leg = {}
        ref = None
        for nm in ("heuristic", "learned", "bitmap"):
            cfg = autotune.cfg_for("c3", nm)
            for _ in range(2):
                cnt.zero()
                gq.triangleCountHomAsync(cctx, cnt, names[nm], cfg)
            gq.synchronize()
            cnt.zero()
            k = max(1, min(a.steps, 10))
            timer.start()
            for _ in range(k):
                gq.triangleCountHomAsync(cctx, cnt, names[nm], cfg)
            leg[f"{nm}_ms"] = timer.stop() / k
            tot = int(cnt.to_array(np.uint64, 1)[0]) // k
            if ref is not None and tot != ref:
                raise SystemExit(f"chung_lu_4m: {nm} counted {tot} != {ref}")
            ref = tot
            log(f"chung_lu_4m {nm}: {leg[f'{nm}_ms']:.3f} ms")
        leg["triangles"] = ref
        leg["m"] = gc.nnz
        selector_legs["chung_lu_4m"] = leg
        del cctx, gc
    for leg in selector_legs.values():
        if leg.get("heuristic_ms") and leg.get("learned_ms"):
            leg["learned_vs_plain"] = leg["heuristic_ms"] / leg["learned_ms"]

    batched = None
    if not a.no_batch and not a.only:
        from gary_amd import batch as gb
        t0 = time.time()
        gB, vbase = gb.build_union(gb.batch_specs(2000))
        log(f"batch2000: m={gB.nnz} host build {time.time() - t0:.2f}s")
        bB_all, _ = pivot_bytes(gB)
        B = gb.Batch(gB, vbase)
        B.set_cfg(autotune.cfg_for("batch2000", "bitmap"))
        per_graph = B.count()
        one = gq.triangleCountHom(B.deviceGraph(), gq.BITMAP)
        if int(per_graph.sum()) != one:
            raise SystemExit(f"batch2000: per-graph sum {int(per_graph.sum())} != union count {one}")
        for _ in range(a.warmup):
            B.launch(use_graph=True)
        gq.synchronize()
        timer.start()
        for _ in range(a.steps):
            B.launch(use_graph=True)
        ms_b = timer.stop() / a.steps
        if not np.array_equal(B.counts(), per_graph):
            raise SystemExit("batch2000: replayed counts differ")
        ach_b = bB_all / (ms_b * 1e-3) / 1e9
        batched = {"workload": "2,000 seeded graphs, n ~ U[64, 20000], avg degree ~ U[4, 64], "
                               "half ER half R-MAT(.57,.19,.19,.05), degree-oriented int32 CSR, "
                               "one hipGraph replay per step, uint64 count per graph",
                   "graphs": 2000, "m": gB.nnz, "ms_per_launch": ms_b,
                   "edges_per_s": gB.nnz / (ms_b * 1e-3), "triangles": one,
                   "roofline": {"achieved": ach_b, "frac": ach_b / HBM_PEAK_GBS,
                                "algorithmic_bytes": bB_all}}
        log(f"batch2000: {ms_b:.3f} ms per launch, {gB.nnz / (ms_b * 1e-3):.3e} edges/s")
        B.close()
        del gB

    cpu = None
    if not a.no_cpu:
        import oracle
        threads = oracle.omp_threads(a.cpu_threads)
        oracle.tri_count(g.indPtr, g.colIdx, threads)  # warm-up
        times = []
        for _ in range(3):
            c0 = time.perf_counter()
            tri_cpu = oracle.tri_count(g.indPtr, g.colIdx, threads)
            times.append(time.perf_counter() - c0)
        if tri_cpu != tri:
            raise SystemExit(f"CPU count {tri_cpu} != GPU count {tri}")
        t_cpu = sorted(times)[1]
        cpu = {"value": g.nnz / t_cpu, "unit": "edges/s", "cores": threads, "kind": "port",
               "sample": f"full rmat{a.scale} count (cpuTriCount restated, OpenMP dynamic, "
                         f"median of 3 after 1 warm-up, {t_cpu:.2f}s each)"}
        log(f"cpu baseline: {t_cpu:.3f}s with {threads} threads")

    ms = results[a.strategy]
    pivot_path = a.strategy in ("learned", "heuristic", "bitmap")
    k_ms = kernel_ms.get(a.strategy, ms)
    k_bytes = B_alg_main if pivot_path else B_alg
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    traffic = None
    if pivot_path and a.scale == 20 and a.edgefactor == 16:
        traffic = read_traffic(ROOT / "profiles" / "tc_pivot_rmat20_traffic.md")
    out = {
        "metric": "edges/s triangle count RMAT-20 (exact); learned-vs-plain GPU speedup, %HBM",
        "value": g.nnz / (ms * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"rmat{a.scale}-ef{a.edgefactor} (a=.57,b=.19,c=.19) "
                               "symmetrized, dedup, no loops, degree-oriented int32 CSR",
                   "n": n, "m": g.nnz, "triangles": tri, "strategy": a.strategy},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "tc_pivot_kernel<4,4096>" if pivot_path else "tc_edge_kernel",
                     "kernel_ms": k_ms, "bytes_model": "sum over arcs (u,v) of 4*(d+(u)+d+(v))+32",
                     "algorithmic_bytes": k_bytes, "step_B_alg": B_alg},
        "kernel_ms": kernel_ms,
        "strategies_ms": results,
        "learned_vs_plain": results["heuristic"] / results[a.strategy]
        if "heuristic" in results and a.strategy != "heuristic" else 1.0,
        "selector": selector_legs,
        "batched": batched,
        "tuned_cfg": {k: (None if v is None else {f: getattr(v, f) for f in autotune.FIELDS})
                      for k, v in ((nm, autotune.cfg_for(wl, nm) if wl else None)
                                   for nm in results)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
