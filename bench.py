#!/usr/bin/env python3
"""Exact triangle counting on R-MAT scale 20 (edgefactor 16, a=.57 b=.19 c=.19) on one MI355X.

A step is one full count of the degree-oriented graph already resident in HBM (every device
launch of the counting path, including the per-call pivot binning and, for the learned strategy,
its features and inference). Prints one JSON line: edges/s of the headline strategy, per-strategy
times, the HBM roofline fraction of the dominant kernel under the merge byte model
B_alg = 4*W + 32*m (W = sum over arcs (u,v) of d+(u) + d+(v)), and the reference's CPU algorithm
(cpuTriCount restated in the oracle) timed on this host, with the CPU model and thread count.

ms_per_step is the mean of the K back-to-back steps; ms_per_step_median the median of K counts
each timed alone. Further legs (--legs, default all):
  c1     Erdos-Renyi G(4096, 65536): every strategy and the CPU count;
  c3     the learned selector against the plain degree-threshold heuristic (and bitmap-only, and
         the reference dispatcher's cases as routes) on the Chung-Lu graph (4,194,304 vertices,
         ~64M edges, exponent 2.1, max weight 50,000);
  batch  the batched form: 2,000 graphs of uneven size, one hipGraph replay per step,
         per-graph uint64 counts;
  hom4   K4 / diamond / 4-cycle homomorphism counts on R-MAT scale 22 (wall clock per call:
         each call returns its counts to the host).
  ops    the search and intersection operators on the reference's SearchTest / SetOpTest shapes
         (a sorted list of 4 Mi ints, 32 Ki keys) and with 4 Mi keys, HIP events around repeated
         device-resident calls, beside the oracle's sequential set intersection;
  patterns  triangle listing (every triangle written as an int32 triple) on R-MAT 20, and
         k-clique counts (k = 4, 5) on R-MAT 20 (ranked symmetric CSR built on the device);
  build  graph construction on the device (R-MAT tuples, symmetrize, dedup, orient, CSR) for
         R-MAT 20 degree-oriented and R-MAT 22 ranked symmetric, against the host build of the
         same tuples (equal arrays required).
Launch knobs come from the autotuner's file (gary_amd/tuned.json) where it has an entry for the
workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 20] [--no-cpu]
                    [--legs headline,c3,batch,hom4] [--strategy learned|heuristic|bitmap|...]
                    [--only]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "edges/s triangle count RMAT-20 (exact); learned-vs-plain GPU speedup, %HBM"


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


class Bench:
    def __init__(self, a):
        import numpy as np

        from gary_amd import autotune, graph, selector
        from gary_amd import graph_query as gq
        from gary_amd.device import DeviceBuffer, DeviceGraphCtx
        from gary_amd.timing import EventTimer
        self.np, self.autotune, self.graph, self.selector, self.gq = np, autotune, graph, selector, gq
        self.DeviceGraphCtx = DeviceGraphCtx
        self.a = a
        self.cnt = DeviceBuffer(8)
        self.timer = EventTimer()
        self.names = {"learned": gq.AUTO_LEARNED, "heuristic": gq.HEURISTIC, "bitmap": gq.BITMAP,
                      "cached_bsearch": gq.CACHED_BSEARCH, "merge": gq.MERGE,
                      "bsearch": gq.BSEARCH}

    # ---- helpers ---------------------------------------------------------------------------
    def pivot_bytes(self, g):
        """Σ over arcs (u,v) of 4*(d+(u) + d+(v)) + 32, and the part whose pivot has out-degree
        <= 2048 (the main pivot kernel's share; larger pivots go to the hub kernel)."""
        np = self.np
        dp = g.deg.astype(np.int64)
        row = np.repeat(np.arange(g.m, dtype=np.int64), dp)
        du = dp[row]
        per = 4 * (du + dp[g.colIdx]) + 32
        return int(per.sum()), int(per[du <= 2048].sum())

    def time_count(self, ctx, strategy: int, steps: int, warmup: int, cfg=None):
        """ms per count (HIP events over `steps` back-to-back counts), the count, and the main
        pivot kernel's average launch time where the strategy launches it."""
        from gary_amd.timing import kernel_timing, kernel_timing_read
        gq, cnt = self.gq, self.cnt
        for _ in range(warmup):
            cnt.zero()
            gq.triangleCountHomAsync(ctx, cnt, strategy, cfg)
        gq.synchronize()
        cnt.zero()
        gq.synchronize()
        kernel_timing(True)
        self.timer.start()
        for _ in range(steps):
            gq.triangleCountHomAsync(ctx, cnt, strategy, cfg)
        ms = self.timer.stop() / steps
        gq.synchronize()
        launches, kms = kernel_timing_read()
        kernel_timing(False)
        total = int(cnt.to_array(self.np.uint64, 1)[0])
        assert total % steps == 0
        return ms, total // steps, (kms / launches if launches else None)

    def forward_cfg(self, workload, strategy: str):
        """The tuned launch configuration of `strategy` with every arc streamed forward."""
        c = self.autotune.cfg_for(workload, strategy) if workload else None
        f = {k: getattr(c, k) for k in self.autotune.FIELDS} if c is not None else {}
        return self.gq.launch_cfg(**f, direction=self.gq.FORWARD)

    def step_median(self, ctx, strategy: int, steps: int, cfg=None) -> float:
        """Median ms of `steps` counts each timed alone (HIP events around one count, the host
        waiting between counts): the per-step figure beside the back-to-back mean."""
        gq, cnt = self.gq, self.cnt
        ts = []
        for _ in range(max(steps, 10)):
            self.timer.start()
            gq.triangleCountHomAsync(ctx, cnt, strategy, cfg)
            ts.append(self.timer.stop())
        ts.sort()
        return ts[len(ts) // 2]

    def time_heuristic_tuned(self, ctx, workload: str, want: int, steps: int, cfg=None):
        """ms per count of the heuristic strategy with the autotuner's thresholds for this
        workload (None when the tuner has no entry), the thresholds restored afterwards."""
        th = self.autotune.heuristic_for(workload)
        if th is None:
            return None, None
        try:
            self.selector.set_heuristic(**th)
            ms, t, _ = self.time_count(ctx, self.gq.HEURISTIC, steps, 1, cfg)
        finally:
            self.selector.set_heuristic(**self.selector.HEURISTIC_DEFAULT)
        if t != want:
            raise SystemExit(f"tuned heuristic counted {t} != {want}")
        return ms, th

    def time_dispatch(self, ctx, want: int) -> float:
        """ms per count with the reference dispatcher's cases as routes (no bitmap route)."""
        tab = self.selector.dispatch_table()
        got = self.selector.count_with_table(ctx, tab)
        if got != want:
            raise SystemExit(f"dispatch table counted {got} != {want}")
        ms = self.selector.time_table(ctx, tab, repeats=3)
        log(f"reference-dispatch table: {ms:.3f} ms")
        return ms

    # ---- legs ------------------------------------------------------------------------------
    def headline(self) -> dict:
        a, gq, graph = self.a, self.gq, self.graph
        self.selector.load()  # the learned strategy runs the shipped model (raises if absent)
        wl = f"rmat{a.scale}" if a.edgefactor == 16 else None
        t0 = time.time()
        n, s, d = graph.rmat(a.scale, a.edgefactor)
        g = graph.oriented(n, s, d)
        del s, d
        log(f"rmat{a.scale} ef{a.edgefactor}: n={n} m={g.nnz} max_out_deg={int(g.deg.max())} "
            f"host build {time.time() - t0:.2f}s")
        B_alg, B_alg_main = self.pivot_bytes(g)
        ctx = self.DeviceGraphCtx(g)
        ctx.populateDevice()
        results, kernel_ms = {}, {}
        order = [a.strategy] + ([] if a.only else [x for x in self.names if x != a.strategy])
        tri = None
        for nm in order:
            slow = nm in ("cached_bsearch", "merge", "bsearch")
            steps = max(1, min(a.steps, 3)) if slow and nm != a.strategy else a.steps
            warm = a.warmup if nm == a.strategy else 1
            cfg = self.autotune.cfg_for(wl, nm) if wl else None
            ms, t, kms = self.time_count(ctx, self.names[nm], steps, warm, cfg)
            if tri is not None and t != tri:
                raise SystemExit(f"strategy {nm} counted {t} != {tri}")
            tri = t
            results[nm] = ms
            if kms is not None:
                kernel_ms[nm] = kms
            log(f"{nm}: {ms:.3f} ms/step" + (f", triangles={tri}" if nm == a.strategy else ""))
        direction = None
        if not a.only:
            # the pivot path with every arc streamed forward (no reversed-arc list), beside the
            # default min-direction bitmap
            fms, ft, fkms = self.time_count(ctx, self.names["bitmap"], a.steps, 1,
                                            self.forward_cfg(wl, "bitmap"))
            if ft != tri:
                raise SystemExit(f"forward bitmap counted {ft} != {tri}")
            direction = {"forward_ms": fms, "forward_kernel_ms": fkms,
                         "min_direction_ms": results.get("bitmap"),
                         "min_direction_kernel_ms": kernel_ms.get("bitmap")}
            log(f"bitmap forward: {fms:.3f} ms/step (main kernel {fkms:.3f} ms)")
        self.dispatch_ms = None if a.only else self.time_dispatch(ctx, tri)
        medians = {nm: self.step_median(ctx, self.names[nm], a.steps,
                                        self.autotune.cfg_for(wl, nm) if wl else None)
                   for nm in results if nm in ("learned", "heuristic", "bitmap")}
        self.heur_tuned = (None, None) if a.only or not wl else \
            self.time_heuristic_tuned(ctx, wl, tri, a.steps, self.autotune.cfg_for(wl, "heuristic"))
        if self.heur_tuned[0] is not None:
            log(f"heuristic (tuned thresholds {self.heur_tuned[1]}): {self.heur_tuned[0]:.3f} ms/step")
        pivot_path = a.strategy in ("learned", "heuristic", "bitmap")
        k_ms = kernel_ms.get(a.strategy, results[a.strategy])
        k_bytes = B_alg_main if pivot_path else B_alg
        achieved = k_bytes / (k_ms * 1e-3) / 1e9
        traffic, traffic_src = None, None
        if pivot_path and a.scale == 20 and a.edgefactor == 16:
            # not measured by this run: the calibrated FETCH_SIZE / WRITE_SIZE passes of the
            # committed profile, named with the commit they profiled
            from gary_amd.profiling import read_commit, read_traffic
            prof = ROOT / "profiles" / "tc_pivot_rmat20_traffic.md"
            traffic = read_traffic(prof)
            if traffic is not None:
                traffic_src = f"profiles/{prof.name} (source commit {read_commit(prof)})"
        self.g, self.tri = g, tri
        ms = results[a.strategy]
        return {
            "value": g.nnz / (ms * 1e-3), "ms_per_step": ms,
            "ms_per_step_median": medians.get(a.strategy),
            "strategies_ms_median": medians,
            "config": {"workload": f"rmat{a.scale}-ef{a.edgefactor} (a=.57,b=.19,c=.19) "
                                   "symmetrized, dedup, no loops, degree-oriented int32 CSR",
                       "n": n, "m": g.nnz, "triangles": tri, "strategy": a.strategy},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "traffic_source": traffic_src,
                         "kernel": "tc_pivot_kernel<4,4096>" if pivot_path else "tc_edge_kernel",
                         "kernel_ms": k_ms,
                         "bytes_model": "sum over arcs (u,v) of 4*(d+(u)+d+(v))+32",
                         "algorithmic_bytes": k_bytes, "step_B_alg": B_alg,
                         # the measured HBM-side bytes over the same launch time
                         "achieved_traffic": traffic / (k_ms * 1e-3) / 1e9 if traffic else None,
                         "frac_traffic": traffic / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                         if traffic else None},
            "kernel_ms": kernel_ms,
            "strategies_ms": results,
            "pivot_direction": direction,
            "learned_vs_plain": results["heuristic"] / results[a.strategy]
            if "heuristic" in results and a.strategy != "heuristic" else 1.0,
            "tuned_cfg": {nm: (None if c is None else
                               {f: getattr(c, f) for f in self.autotune.FIELDS})
                          for nm, c in ((nm, self.autotune.cfg_for(wl, nm) if wl else None)
                                        for nm in results)},
        }

    def c3(self) -> dict:
        a = self.a
        self.selector.load()
        from gary_amd.device import gen_chung_lu_device
        t0 = time.time()
        n, nt = 4_194_304, 64 * 1_048_576  # graph.chung_lu_4m(), generated and built on the device
        ds, dd = gen_chung_lu_device(n, nt, 2.1, 50_000.0, self.graph.SEEDS["chung_lu"])
        ctx = self.DeviceGraphCtx()
        ctx.buildFromEdges(n, ds, dd, nt, self.graph.ORIENT_DEGREE)
        del ds, dd
        ctx.populateDevice()
        nnz = ctx.deviceGraph().csr.nnz
        log(f"chung_lu_4m: m={nnz} device generation and build {time.time() - t0:.2f}s")
        leg, ref = {"m": nnz}, None
        for nm in ("heuristic", "learned", "bitmap"):
            ms, t, _ = self.time_count(ctx, self.names[nm], max(1, min(a.steps, 10)), 2,
                                       self.autotune.cfg_for("c3", nm))
            if ref is not None and t != ref:
                raise SystemExit(f"chung_lu_4m: {nm} counted {t} != {ref}")
            ref = t
            leg[f"{nm}_ms"] = ms
            log(f"chung_lu_4m {nm}: {ms:.3f} ms")
        leg["triangles"] = ref
        fms, ft, _ = self.time_count(ctx, self.names["bitmap"], max(1, min(a.steps, 10)), 2,
                                     self.forward_cfg("c3", "bitmap"))
        if ft != ref:
            raise SystemExit(f"chung_lu_4m: forward bitmap counted {ft} != {ref}")
        leg["bitmap_forward_ms"] = fms
        log(f"chung_lu_4m bitmap forward: {fms:.3f} ms")
        leg["reference_dispatch_ms"] = self.time_dispatch(ctx, ref)
        ht_ms, ht = self.time_heuristic_tuned(ctx, "c3", ref, max(1, min(a.steps, 10)),
                                              self.autotune.cfg_for("c3", "heuristic"))
        leg["heuristic_tuned_ms"], leg["heuristic_tuned_thresholds"] = ht_ms, ht
        leg["learned_vs_plain"] = leg["heuristic_ms"] / leg["learned_ms"]
        leg["learned_vs_tuned_heuristic"] = ht_ms / leg["learned_ms"] if ht_ms else None
        leg["learned_vs_bitmap"] = leg["bitmap_ms"] / leg["learned_ms"]
        leg["learned_vs_reference_dispatch"] = leg["reference_dispatch_ms"] / leg["learned_ms"]
        leg["edges_per_s"] = nnz / (leg["learned_ms"] * 1e-3)
        if not a.no_cpu:  # cpuTriCount restated, on the same CSR copied back to the host
            import oracle
            g = ctx.getHostCSR()
            threads = oracle.omp_threads(a.cpu_threads)
            ts = []
            for _ in range(2):
                c0 = time.perf_counter()
                tri = oracle.tri_count(g.indPtr, g.colIdx, threads)
                ts.append(time.perf_counter() - c0)
            if tri != ref:
                raise SystemExit(f"chung_lu_4m: CPU count {tri} != GPU count {ref}")
            t = min(ts)
            leg["cpu_baseline"] = self.cpu_record(nnz / t, threads, f"full Chung-Lu count, best of 2, {t:.2f}s each")
            log(f"chung_lu_4m cpu baseline: {t:.2f}s with {threads} threads")
        return leg

    def batch(self) -> dict:
        from gary_amd import batch as gb
        a, np, gq = self.a, self.np, self.gq
        t0 = time.time()
        g, vbase = gb.build_union(gb.batch_specs(2000))
        log(f"batch2000: m={g.nnz} host build {time.time() - t0:.2f}s")
        b_all, _ = self.pivot_bytes(g)
        B = gb.Batch(g, vbase)
        B.set_cfg(self.autotune.cfg_for("batch2000", "bitmap"))
        per_graph = B.count()
        one = gq.triangleCountHom(B.deviceGraph(), gq.BITMAP)
        if int(per_graph.sum()) != one:
            raise SystemExit(f"batch2000: per-graph sum {int(per_graph.sum())} != union count {one}")
        for _ in range(a.warmup):
            B.launch(use_graph=True)
        gq.synchronize()
        self.timer.start()
        for _ in range(a.steps):
            B.launch(use_graph=True)
        ms = self.timer.stop() / a.steps
        if not np.array_equal(B.counts(), per_graph):
            raise SystemExit("batch2000: replayed counts differ")
        ach = b_all / (ms * 1e-3) / 1e9
        log(f"batch2000: {ms:.3f} ms per launch, {g.nnz / (ms * 1e-3):.3e} edges/s")
        B.close()
        leg = {"workload": "2,000 seeded graphs, n ~ U[64, 20000], avg degree ~ U[4, 64], half ER "
                           "half R-MAT(.57,.19,.19,.05), degree-oriented int32 CSR, one hipGraph "
                           "replay per step, uint64 count per graph",
               "graphs": 2000, "m": g.nnz, "ms_per_launch": ms,
               "edges_per_s": g.nnz / (ms * 1e-3), "triangles": one,
               "roofline": {"achieved": ach, "frac": ach / HBM_PEAK_GBS,
                            "algorithmic_bytes": b_all}}
        if not a.no_cpu:
            # cpuTriCount per graph, graphs in parallel, on every 10th graph of the union
            import oracle
            threads = oracle.omp_threads(a.cpu_threads)
            picks = list(range(0, 2000, 10))
            ptrs, cols, base = [np.zeros(1, np.int64)], [], [0]
            for i in picks:  # graph i's rows of the union CSR, its ids moved to base[-1]
                lo, hi = int(vbase[i]), int(vbase[i + 1])
                e0, e1 = int(g.indPtr[lo]), int(g.indPtr[hi])
                ptrs.append(g.indPtr[lo + 1:hi + 1] - e0 + ptrs[-1][-1])
                cols.append((g.colIdx[e0:e1].astype(np.int64) - lo + base[-1]).astype(np.int32))
                base.append(base[-1] + hi - lo)
            sub = self.graph.GraphCSR(np.concatenate(ptrs), np.concatenate(cols))
            c0 = time.perf_counter()
            per = oracle.tri_count_batch(np.array(base, np.int32), sub.indPtr, sub.colIdx, threads)
            t = time.perf_counter() - c0
            if not np.array_equal(per, per_graph[picks]):
                raise SystemExit("batch2000: CPU per-graph counts differ from the GPU's")
            leg["cpu_baseline"] = self.cpu_record(
                sub.nnz / t, threads, f"every 10th graph (200 graphs, {sub.nnz} arcs), "
                                      f"parallel over graphs, one run, {t:.2f}s")
            log(f"batch2000 cpu baseline: 200 graphs in {t:.2f}s with {threads} threads")
        return leg

    def hom4(self) -> dict:
        gq, graph, np = self.gq, self.graph, self.np
        t0 = time.time()
        n, s, d = graph.rmat(22, 16, seed=graph.SEEDS["rmat22"])
        g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
        del s, d
        log(f"rmat22 ranked symmetric: nnz={g.nnz} host build {time.time() - t0:.2f}s")
        ctx = self.DeviceGraphCtx(g)
        leg = {"workload": "rmat22-ef16 (a=.57,b=.19,c=.19) seed 0x6A7E0005, symmetrized, dedup, "
                           "no loops, ids ranked by (degree, id), int32 CSR",
               "n": n, "undirected_edges": g.nnz // 2}
        from gary_amd import c4_selector
        model = c4_selector.load(required=False)  # AUTO routes the 4-cycle vertices by it
        leg["c4_routes"] = "learned (gary_amd/c4_model.json)" if model else "fixed degree split"
        vals = None
        for nm, pat in (("k4_diamond", gq.HOM_K4 | gq.HOM_DIAMOND), ("c4", gq.HOM_C4),
                        ("all", gq.HOM_ALL)):
            gq.hom4Count(ctx, pat)  # warm-up
            ts = []
            for _ in range(5):
                t1 = time.perf_counter()
                vals = gq.hom4Count(ctx, pat)
                ts.append((time.perf_counter() - t1) * 1e3)
            leg[f"{nm}_ms"] = float(np.median(ts))
            log(f"rmat22 hom4 {nm}: {leg[f'{nm}_ms']:.1f} ms (median of 5)")
        leg["counts"] = vals
        if model and not self.a.no_c4_compare:
            leg["c4_learned_vs_fixed"] = {"rmat22": self.c4_routing(ctx)}
            del ctx
            n20, s20, d20 = graph.rmat(20, 16, seed=graph.SEEDS["rmat22"])
            g20 = graph.csr_from_edges(n20, s20, d20, graph.SYMMETRIC_RANKED)
            leg["c4_learned_vs_fixed"]["rmat20"] = self.c4_routing(self.DeviceGraphCtx(g20))
        leg["edges_per_s"] = (g.nnz // 2) / (leg["all_ms"] * 1e-3)
        # per-pattern algorithmic bytes (m undirected edges, d full degree, d+ upper degree):
        #   diamond 4*sum d^2 + 32m; 4-cycles 4*sum d^2 + 16n; 4-cliques at least 4*W + 32m
        #   (W = sum over arcs (u,v) of d+(u) + d+(v); the full K4 formula adds the common
        #   out-neighbour terms, which need per-arc triangle counts)
        d = np.diff(g.indPtr).astype(np.int64)
        row = np.repeat(np.arange(n, dtype=np.int64), d)
        upper = g.colIdx > row
        dplus = np.bincount(row[upper], minlength=n).astype(np.int64)
        W = int((dplus[row[upper]] + dplus[g.colIdx[upper]]).sum())
        m_und = g.nnz // 2
        sd2 = int((d * d).sum())
        patterns = {"diamond": (4 * sd2 + 32 * m_und, leg["k4_diamond_ms"]),
                    "c4": (4 * sd2 + 16 * n, leg["c4_ms"]),
                    "k4_lower_bound": (4 * W + 32 * m_und, leg["k4_diamond_ms"])}
        leg["roofline"] = {k: {"algorithmic_bytes": b, "ms": t,
                               "achieved": b / (t * 1e-3) / 1e9,
                               "frac": b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
                           for k, (b, t) in patterns.items()}
        if not self.a.no_cpu:
            leg["cpu_sample"] = self.hom4_cpu_sample()
        return leg

    def c4_routing(self, ctx) -> dict:
        """Learned vs fixed 4-cycle vertex routing on one graph: both tables' counts (all six
        values must agree) and the medians of 7 alternating C4 calls of each."""
        from gary_amd import c4_selector
        gq = self.gq
        a = gq.hom4Count(ctx, gq.HOM_ALL, c4=gq.C4_FIXED)
        b = gq.hom4Count(ctx, gq.HOM_ALL, c4=gq.C4_LEARNED)
        if a != b:
            raise SystemExit(f"learned 4-cycle routing changed the counts: {a} != {b}")
        tf, tl = c4_selector.compare(ctx, gq.HOM_C4, 7)
        _, nv = c4_selector.features(ctx)
        lt = c4_selector.table(ctx, gq.C4_LEARNED)
        same = bool(self.np.all((lt == c4_selector.fixed_table()) | (nv == 0)))
        moved = int(nv[(lt != c4_selector.fixed_table())].sum())
        log(f"c4 routing: fixed {tf:.1f} ms, learned {tl:.1f} ms (x{tf / tl:.2f}), "
            f"{moved} vertices off the fixed split")
        return {"c4_ms_fixed": tf, "c4_ms_learned": tl, "learned_vs_fixed": tf / tl,
                "counts_equal": True, "same_routes": same, "vertices_moved": moved}

    def hom4_cpu_sample(self) -> dict:
        """The CPU hom4 path (the oracle's oriented-wedge counts, OpenMP) on R-MAT 20 beside the
        GPU call on the same graph: RMAT-22 takes minutes on the CPU, so the sample is the same
        generator two scales down."""
        import oracle
        gq, graph = self.gq, self.graph
        n, s, d = graph.rmat(20, 16, seed=graph.SEEDS["rmat20"])
        g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
        ctx = self.DeviceGraphCtx(g)
        want = gq.hom4Count(ctx, gq.HOM_ALL)
        t1 = time.perf_counter()
        for _ in range(3):
            gq.hom4Count(ctx, gq.HOM_ALL)
        gpu_ms = (time.perf_counter() - t1) * 1e3 / 3
        threads = oracle.omp_threads(self.a.cpu_threads)
        c0 = time.perf_counter()
        h = oracle.hom4_ranked(g.indPtr, g.colIdx, threads)
        t = time.perf_counter() - c0
        if h != want:
            raise SystemExit(f"rmat20 hom4: CPU {h} != GPU {want}")
        m_und = g.nnz // 2
        log(f"rmat20 hom4: GPU {gpu_ms:.1f} ms, CPU {t:.2f}s with {threads} threads")
        return {"workload": "rmat20-ef16 ranked symmetric, all six counts", "undirected_edges": m_und,
                "gpu_ms": gpu_ms, "gpu_edges_per_s": m_und / (gpu_ms * 1e-3),
                "cpu_baseline": self.cpu_record(m_und / t, threads,
                                                f"full R-MAT 20 hom4 (oracle.hom4_ranked), one run, "
                                                f"{t:.1f}s")}

    def build(self) -> dict:
        """Device graph construction against the host build of the same tuples (wall clock,
        median of 5 after a warm-up; the device build returns after reading its edge count)."""
        from gary_amd.device import gen_rmat_device
        np, graph, gq = self.np, self.graph, self.gq

        def med(f, k=5):
            f()
            ts = []
            for _ in range(k):
                t0 = time.perf_counter()
                f()
                gq.synchronize()
                ts.append(time.perf_counter() - t0)
            return sorted(ts)[k // 2] * 1e3

        leg = {}
        for name, scale, seed, mode in (("rmat20_degree", 20, graph.SEEDS["rmat20"],
                                         graph.ORIENT_DEGREE),
                                        ("rmat22_ranked", 22, graph.SEEDS["rmat22"],
                                         graph.SYMMETRIC_RANKED)):
            if name not in self.a.build_graphs.split(","):
                continue
            n, nt = 1 << scale, 16 << scale
            host = {}

            def host_build():
                s, d = graph.gen_rmat(scale, nt, seed=seed)
                host["g"] = graph.csr_from_edges(n, s, d, mode)

            host_ms = med(host_build, 3)
            gen_ms = med(lambda: gen_rmat_device(scale, nt, seed=seed))
            ds, dd = gen_rmat_device(scale, nt, seed=seed)
            ctx = self.DeviceGraphCtx()
            build_ms = med(lambda: ctx.buildFromEdges(n, ds, dd, nt, mode))
            g = ctx.getHostCSR()
            if not (np.array_equal(g.indPtr, host["g"].indPtr)
                    and np.array_equal(g.colIdx, host["g"].colIdx)):
                raise SystemExit(f"device build of {name} differs from the host build")
            leg[name] = {"tuples": nt, "stored_edges": g.nnz, "device_gen_ms": gen_ms,
                         "device_build_ms": build_ms, "host_gen_build_ms": host_ms,
                         "speedup": host_ms / (gen_ms + build_ms)}
            log(f"build {name}: device gen {gen_ms:.2f} ms + build {build_ms:.2f} ms, "
                f"host {host_ms:.1f} ms")
            ctx.close()
            del ds, dd, host
        return leg

    def patterns(self) -> dict:
        """Triangle listing and k-cliques on R-MAT 20 (wall clock per call, median of 5 after a
        warm-up; each call returns its count to the host)."""
        from gary_amd.device import DeviceBuffer, gen_rmat_device
        gq, graph = self.gq, self.graph

        def med(f, k=5):
            f()
            ts = []
            for _ in range(k):
                t0 = time.perf_counter()
                r = f()
                ts.append(time.perf_counter() - t0)
            return sorted(ts)[k // 2] * 1e3, r

        n, nt = 1 << 20, 16 << 20
        ds, dd = gen_rmat_device(20, nt, seed=graph.SEEDS["rmat20"])
        ctx = self.DeviceGraphCtx()
        ctx.buildFromEdges(n, ds, dd, nt, graph.ORIENT_DEGREE)
        tri = gq.triangleCountHom(ctx, gq.BITMAP)
        out = DeviceBuffer(12 * tri)
        list_ms, listed = med(lambda: gq.triangleListAsync(ctx, out, tri))
        if listed != tri:
            raise SystemExit(f"listing found {listed} triangles, count {tri}")
        g = ctx.getHostCSR()
        b_alg, _ = self.pivot_bytes(g)
        leg = {"workload": "rmat20-ef16 (a=.57,b=.19,c=.19), degree-oriented int32 CSR; every "
                           "triangle written as an int32 triple",
               "triangles": tri, "list_ms": list_ms, "triangles_per_s": tri / (list_ms * 1e-3),
               "output_bytes": 12 * tri,
               "roofline": {"algorithmic_bytes": b_alg + 12 * tri,
                            "achieved": (b_alg + 12 * tri) / (list_ms * 1e-3) / 1e9,
                            "frac": (b_alg + 12 * tri) / (list_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
        log(f"patterns: listed {tri} triangles in {list_ms:.2f} ms")
        del out
        rctx = self.DeviceGraphCtx()
        rctx.buildFromEdges(n, ds, dd, nt, graph.SYMMETRIC_RANKED)
        for k in (4, 5):
            ms, c = med(lambda: gq.kcliqueCount(rctx, k), 3)
            leg[f"k{k}_ms"] = ms
            leg[f"k{k}"] = c
            log(f"patterns: {c} {k}-cliques in {ms:.1f} ms")
        return leg

    def ops(self) -> dict:
        """Search and intersection operators (include/ops/Search.cuh, Intersection.cuh) on the
        reference's test shapes (test/SearchTest.cu:201-208, test/SetOpTest.cu:62-68): sorted
        int32 lists generated here from a fixed seed."""
        import oracle
        from gary_amd import ops
        from gary_amd._lib import check, lib
        from gary_amd.device import DeviceBuffer
        np = self.np
        rng = np.random.default_rng(0x6A7E00F0)
        lst = np.sort(rng.integers(1 << 29, 1 << 30, 4 << 20, dtype=np.int64)).astype(np.int32)
        dl = DeviceBuffer.from_array(lst)
        res = DeviceBuffer(4 << 20)
        leg = {"workload": "sorted list of 4,194,304 int32 in [2^29, 2^30]; keys uniform in the "
                           "same range (SearchTest); sets A of 4,194,304 in [2^26, 2^30] and B "
                           "(SetOpTest)"}

        def timed(f, reps):
            f()
            self.gq.synchronize()
            self.timer.start()
            for _ in range(reps):
                f()
            return self.timer.stop() / reps

        for nk in (32 << 10, 4 << 20):
            keys = np.sort(rng.integers(1 << 29, 1 << 30, nk, dtype=np.int64)).astype(np.int32)
            dk = DeviceBuffer.from_array(keys)
            want = np.isin(keys, lst)
            # two_phase: the reference's 32-entry index cache (test/TriQueryTest.cu:42);
            # two_phase_1024: a 1,024-entry cache (4 KiB of LDS per workgroup)
            for name, mode, depth in (("plain", ops.SEARCH_PLAIN, 32),
                                      ("two_phase", ops.SEARCH_2PHASE, 32),
                                      ("two_phase_1024", ops.SEARCH_2PHASE, 1024)):
                check(lib.gary_search(dl.ptr, lst.size, dk.ptr, nk, res.ptr, mode, depth, None))
                got = res.to_array(np.uint8, nk).astype(bool)
                if not np.array_equal(got, want):
                    raise SystemExit(f"search {name} with {nk} keys differs from numpy")
                ms = timed(lambda: check(lib.gary_search(dl.ptr, lst.size, dk.ptr, nk, res.ptr,
                                                         mode, depth, None)), 20)
                leg[f"search_{name}_{nk}_us"] = ms * 1e3
                leg[f"search_{name}_{nk}_keys_per_s"] = nk / (ms * 1e-3)
            log(f"ops: search {nk} keys: plain {leg[f'search_plain_{nk}_us']:.1f} us, two-phase "
                f"{leg[f'search_two_phase_{nk}_us']:.1f} us (1,024-entry cache "
                f"{leg[f'search_two_phase_1024_{nk}_us']:.1f} us)")
        a = np.sort(rng.integers(1 << 26, 1 << 30, 4 << 20, dtype=np.int64)).astype(np.int32)
        a = np.unique(a)
        da = DeviceBuffer.from_array(a)
        cnt = DeviceBuffer(8)
        for nb in (32 << 10, 4 << 20):
            b = np.unique(np.sort(rng.integers(1 << 26, 1 << 30, nb, dtype=np.int64)).astype(np.int32))
            db = DeviceBuffer.from_array(b)
            want = oracle.intersect_count(a, b)
            c0 = time.perf_counter()
            for _ in range(5):
                oracle.intersect_count(a, b)
            cpu_us = (time.perf_counter() - c0) / 5 * 1e6
            leg[f"isect_{nb}_cpu_us"] = cpu_us
            for name, mth in (("bsearch", ops.ISECT_BSEARCH), ("cached", ops.ISECT_CACHED_BSEARCH),
                              ("merge", ops.ISECT_MERGE), ("bitmap", ops.ISECT_BITMAP)):
                cnt.zero()
                check(lib.gary_intersect_count(da.ptr, a.size, db.ptr, b.size, mth, 32, cnt.ptr,
                                               None))
                if int(cnt.to_array(np.uint64, 1)[0]) != want:
                    raise SystemExit(f"intersection {name} with |B|={nb} differs from the oracle")
                ms = timed(lambda: check(lib.gary_intersect_count(
                    da.ptr, a.size, db.ptr, b.size, mth, 32, cnt.ptr, None)), 20)
                bytes_alg = 4 * (a.size + b.size)
                leg[f"isect_{nb}_{name}_us"] = ms * 1e3
                leg[f"isect_{nb}_{name}_frac"] = bytes_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
            best = min(("bsearch", "cached", "merge", "bitmap"),
                       key=lambda x: leg[f"isect_{nb}_{x}_us"])
            log(f"ops: intersect |A|={a.size} |B|={b.size}: best {best} "
                f"{leg[f'isect_{nb}_{best}_us']:.1f} us (cpu {cpu_us:.0f} us)")
        return leg

    @staticmethod
    def cpu_model() -> str:
        try:
            for line in Path("/proc/cpuinfo").read_text().splitlines():
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
        except OSError:
            pass
        return "unknown"

    def cpu_record(self, value: float, threads: int, sample: str) -> dict:
        return {"value": value, "unit": "edges/s", "cores": threads, "kind": "port",
                "cpu_model": self.cpu_model(), "sample": sample}

    def cpu(self) -> dict:
        import oracle
        a, g = self.a, self.g
        threads = oracle.omp_threads(a.cpu_threads)
        oracle.tri_count(g.indPtr, g.colIdx, threads)  # warm-up
        times = []
        for _ in range(9):
            c0 = time.perf_counter()
            tri = oracle.tri_count(g.indPtr, g.colIdx, threads)
            times.append(time.perf_counter() - c0)
        if tri != self.tri:
            raise SystemExit(f"CPU count {tri} != GPU count {self.tri}")
        t = sorted(times)[4]
        log(f"cpu baseline: {t:.3f}s with {threads} threads ({self.cpu_model()})")
        return self.cpu_record(g.nnz / t, threads,
                               f"full rmat{a.scale} count (cpuTriCount restated, OpenMP dynamic), "
                               f"median of 9 after 1 warm-up, {t:.2f}s each")

    def c1(self) -> dict:
        """Erdős–Rényi G(4096, 65536), degree-oriented: every strategy on the GPU (median of
        per-count times, each count being launch-bound at this size) and the CPU count."""
        import oracle
        a, gq, graph = self.a, self.gq, self.graph
        self.selector.load()
        g = graph.oriented(*graph.er_4096())
        ctx = self.DeviceGraphCtx(g)
        ctx.populateDevice()
        want = oracle.tri_count(g.indPtr, g.colIdx)
        leg = {"workload": "Erdos-Renyi G(n=4096, m=65536), seed 0x6A7E0001, degree-oriented "
                           "int32 CSR", "m": g.nnz, "triangles": want, "strategies_ms": {}}
        for nm, sid in self.names.items():
            cfg = self.autotune.cfg_for("er4096", nm)
            ms, t, _ = self.time_count(ctx, sid, max(a.steps, 10), 2, cfg)
            if t != want:
                raise SystemExit(f"er4096: {nm} counted {t} != {want}")
            leg["strategies_ms"][nm] = self.step_median(ctx, sid, max(a.steps, 10), cfg)
        best = min(leg["strategies_ms"], key=leg["strategies_ms"].get)
        leg["best"] = best
        leg["edges_per_s"] = g.nnz / (leg["strategies_ms"][best] * 1e-3)
        log(f"er4096: " + ", ".join(f"{k} {v * 1e3:.0f} us" for k, v in leg["strategies_ms"].items()))
        if not a.no_cpu:
            threads = oracle.omp_threads(a.cpu_threads)
            oracle.tri_count(g.indPtr, g.colIdx, threads)
            ts = []
            for _ in range(21):
                c0 = time.perf_counter()
                oracle.tri_count(g.indPtr, g.colIdx, threads)
                ts.append(time.perf_counter() - c0)
            t = sorted(ts)[10]
            leg["cpu_baseline"] = self.cpu_record(g.nnz / t, threads,
                                                  f"full count, median of 21, {t * 1e3:.2f} ms each")
            log(f"er4096 cpu baseline: {t * 1e3:.2f} ms with {threads} threads")
        return leg

    def calibrate_copy(self, mib: int) -> None:
        from gary_amd._lib import check, lib
        from gary_amd.device import DeviceBuffer
        nbytes = mib << 20
        src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        src.zero()
        check(lib.gary_memcpy_d2d(0, dst.ptr, src.ptr, nbytes))
        self.gq.synchronize()
        log(f"calibration copy: {nbytes} bytes device to device")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--strategy", default="learned")
    ap.add_argument("--legs", default="headline,c1,c3,batch,hom4,build,patterns,ops")
    ap.add_argument("--build-graphs", default="rmat20_degree,rmat22_ranked",
                    help="graphs of the build leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-c4-compare", action="store_true",
                    help="hom4 leg: skip the learned vs fixed 4-cycle routing comparison (its "
                         "R-MAT 20 calls would mix into a profile of the R-MAT 22 kernels)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--only", action="store_true",
                    help="headline: time --strategy only, and run no other leg (profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="also run one device-to-device copy of this size (counter calibration)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported (each GPU box "
              "here has one GPU; the counting path shards by pivot ranges with one u64 sum)",
              file=sys.stderr)
        return 2
    legs = [x for x in a.legs.split(",") if x]
    if a.only:
        legs = [x for x in legs if x == "headline"]
    b = Bench(a)
    out = {"metric": METRIC, "value": None, "unit": "edges/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32", "data": "synthetic"}
    if "headline" in legs:
        out.update(b.headline())
    selector_legs = {}
    if "headline" in legs and not a.only:
        r = out["strategies_ms"]
        ht_ms, ht = b.heur_tuned
        selector_legs[f"rmat{a.scale}"] = {
            "heuristic_ms": r["heuristic"], "learned_ms": r["learned"],
            "bitmap_ms": r["bitmap"], "reference_dispatch_ms": b.dispatch_ms,
            "heuristic_tuned_ms": ht_ms, "heuristic_tuned_thresholds": ht,
            "learned_vs_plain": r["heuristic"] / r["learned"],
            "learned_vs_tuned_heuristic": ht_ms / r["learned"] if ht_ms else None,
            "learned_vs_bitmap": r["bitmap"] / r["learned"],
            "learned_vs_reference_dispatch": b.dispatch_ms / r["learned"]}
    if "c1" in legs:
        out["c1"] = b.c1()
    if "c3" in legs:
        selector_legs["chung_lu_4m"] = b.c3()
    if selector_legs:
        out["selector"] = selector_legs
    if "batch" in legs:
        out["batched"] = b.batch()
    if "hom4" in legs:
        out["hom4"] = b.hom4()
    if "build" in legs:
        out["build"] = b.build()
    if "patterns" in legs:
        out["patterns"] = b.patterns()
    if "ops" in legs:
        out["ops"] = b.ops()
    if a.calibrate_copy:
        b.calibrate_copy(a.calibrate_copy)
    out["cpu_baseline"] = b.cpu() if "headline" in legs and not a.no_cpu else None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
