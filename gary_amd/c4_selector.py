"""Routing of the 4-cycle count's vertices by predicted cost (gary_hom4_count_routed).

4-cycles are counted at their highest-ranked vertex u from its wedges u-v-w (v, w < u). Vertex u
with nl >= 2 lower neighbours falls in bin 16 * min(15, log2 nl) + min(15, log2 S / 2), where S
is the sum of d(v) over its lower neighbours (a bound on its wedges; above 64 lower neighbours,
estimated from 64 evenly spaced ones). A table of 256 route bytes
sends each bin to the light kernels (LDS hash tables, nl < 256), the LDS endpoint windows, or the
dense per-workgroup counters. The fixed split is the degree threshold nl >= 256 -> windows; the
learned table is the route with the least MLP-predicted ns per vertex, where the MLP (the
selector's 6 -> 32 -> 32 -> 3 network, trained on the device) maps a bin's features to log2 ns
per vertex of each route, from per-bin timings measured here (time_bins). Counts never depend on
the table.

    python -m gary_amd.c4_selector [--out gary_amd/c4_model.json] [--graphs rmat18,rmat20,rmat22]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib, graph, graph_query as gq
from ._lib import DeviceGraph, check, lib, ptr

LIGHT, WINDOW, DENSE, SKIP = 0, 1, 2, 3
ROUTE_NAMES = {LIGHT: "light", WINDOW: "window", DENSE: "dense", SKIP: "skip"}
NBINS, IN, OUT = 256, 6, 3
NPARAMS = 32 * IN + 32 + 32 * 32 + 32 + OUT * 32 + OUT
HEAVY_LO = 256  # light kernels hold fewer lower neighbours than this
DENSE_MAX_S_BIN = 11  # learned tables use the dense counters only in bins (*, <= 11)
MODEL_PATH = Path(__file__).resolve().parent / "c4_model.json"

_P, _F = C.c_void_p, C.c_float
_lib.bind("gary_hom4_c4_features", C.c_int, [C.POINTER(DeviceGraph), _P, _P])
_lib.bind("gary_hom4_c4_table", C.c_int, [C.POINTER(DeviceGraph), C.c_int, _P])
_lib.bind("gary_hom4_c4_set_model", C.c_int, [_P, _F])
_lib.bind("gary_hom4_c4_load_model", C.c_int, [C.c_char_p])
_lib.bind("gary_hom4_c4_model", C.c_int, [C.POINTER(C.c_int), _P, C.POINTER(C.c_float)])
_lib.bind("gary_mlp_train", C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, _F, C.c_uint64, _P, _P])


def _dg(g):
    return g.deviceGraph() if hasattr(g, "deviceGraph") else g


def bin_of(nl: int, S: int) -> int:
    return min(15, int(nl).bit_length() - 1) * 16 + min(15, (int(max(S, 1)).bit_length() - 1) >> 1)


def fixed_table() -> np.ndarray:
    """The degree split: nl >= 256 (bins 128..255) -> windows, the rest light."""
    t = np.full(NBINS, LIGHT, np.uint8)
    t[(1 << (np.arange(NBINS) >> 4)) >= HEAVY_LO] = WINDOW
    return t


def light_ok() -> np.ndarray:
    """Bins whose vertices the light kernels can hold (nl < 256)."""
    return (1 << (np.arange(NBINS) >> 4)) < HEAVY_LO


def dense_ok() -> np.ndarray:
    """Bins a learned table may send to the dense counters (S < 2^24 per vertex)."""
    return (np.arange(NBINS) & 15) <= DENSE_MAX_S_BIN


def features(g) -> tuple[np.ndarray, np.ndarray]:
    """(X[256][6] MLP inputs, vertices per bin) of a ranked symmetric graph."""
    X = np.zeros((NBINS, IN), np.float32)
    nv = np.zeros(NBINS, np.float32)
    dg = _dg(g)
    check(lib.gary_hom4_c4_features(C.byref(dg), ptr(X), ptr(nv)))
    return X, nv


def table(g, mode: int = gq.C4_AUTO) -> np.ndarray:
    """The route table a mode (C4_AUTO, C4_FIXED, C4_LEARNED) gives g."""
    t = np.zeros(NBINS, np.uint8)
    dg = _dg(g)
    check(lib.gary_hom4_c4_table(C.byref(dg), int(mode), ptr(t)))
    return t


def set_model(params, margin: float = 0.0) -> None:
    p = np.ascontiguousarray(params, np.float32)
    assert p.size == NPARAMS
    check(lib.gary_hom4_c4_set_model(ptr(p), float(margin)))


def clear_model() -> None:
    check(lib.gary_hom4_c4_set_model(None, 0.0))


def model() -> tuple[np.ndarray, float] | None:
    """(params, margin) of the current device's model, or None."""
    t = C.c_int(0)
    p = np.zeros(NPARAMS, np.float32)
    mg = C.c_float(0)
    check(lib.gary_hom4_c4_model(C.byref(t), ptr(p), C.byref(mg)))
    return (p, float(mg.value)) if t.value else None


def load(path: Path = MODEL_PATH, required: bool = True) -> dict | None:
    """Loads a saved model (gary_hom4_c4_load_model; AUTO then routes by it) and returns the
    file's contents; a missing file raises FileNotFoundError, or returns None if not required."""
    p = Path(path)
    if not p.exists():
        if required:
            raise FileNotFoundError(f"no 4-cycle routing model at {p}")
        return None
    check(lib.gary_hom4_c4_load_model(str(p).encode()))
    return json.loads(p.read_text())


def train(X, T, M, epochs: int = 4000, lr: float = 3e-3, seed: int = 0xC4C4) -> tuple[np.ndarray, float]:
    """Full-batch Adam on the device from a fresh initialisation: (params, final loss)."""
    X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
    p = np.zeros(NPARAMS, np.float32)
    loss = C.c_float(0)
    check(lib.gary_mlp_train(ptr(X), ptr(T), ptr(M), X.shape[0], epochs, lr, seed, ptr(p),
                             C.byref(loss)))
    return p, float(loss.value)


def time_routes(g, routes, patterns: int = gq.HOM_C4, repeats: int = 3) -> float:
    """Median milliseconds of a synchronous routed count (after one warm-up)."""
    gq.hom4Count(g, patterns, c4=routes)
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        gq.hom4Count(g, patterns, c4=routes)
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def time_bins(g, X: np.ndarray, nv: np.ndarray, min_vertices: int = 64, repeats: int = 3,
              dense_max_wedge_bound: float = 4e9, floor_ms: float = 0.02,
              log=None) -> tuple[np.ndarray, dict]:
    """ns per vertex of every route on every bin with >= min_vertices vertices (NaN elsewhere).

    Each bin is timed alone against the all-skip table. The window route has a fixed cost (the
    run index, built once when any vertex takes the windows): with b1, b2 the two timed bins of
    least wedge bound n * 2^mean(log2 S), it is t(b1) + t(b2) - t(b1 and b2) - t(all skip) on the
    window route, and is subtracted from every window time. The dense route is timed only where
    the wedge bound is below dense_max_wedge_bound (dense counters take one device-memory atomic
    per wedge). A bin's time is at least floor_ms, the run-to-run spread of a call."""
    out = np.full((NBINS, OUT), np.nan, np.float64)
    timed = [b for b in range(NBINS) if nv[b] >= min_vertices]
    skip = np.full(NBINS, SKIP, np.uint8)
    t_skip = min(time_routes(g, skip, repeats=repeats) for _ in range(2))
    wedge_bound = nv * np.exp2(X[:, 1] * 16.0)
    fixed_w = 0.0
    if len(timed) >= 2:
        b1, b2 = sorted(timed, key=lambda b: wedge_bound[b])[:2]
        t = []
        for sel in ((b1,), (b2,), (b1, b2)):
            tab = skip.copy()
            tab[list(sel)] = WINDOW
            t.append(min(time_routes(g, tab, repeats=repeats) for _ in range(2)))
        fixed_w = max(0.0, t[0] + t[1] - t[2] - t_skip)
    ok, dok = light_ok(), dense_ok()
    for b in timed:
        for r in (LIGHT, WINDOW, DENSE):
            if r == LIGHT and not ok[b]:
                continue
            if r == DENSE and (not dok[b] or wedge_bound[b] > dense_max_wedge_bound):
                continue
            tab = skip.copy()
            tab[b] = r
            t = time_routes(g, tab, repeats=repeats) - t_skip - (fixed_w if r == WINDOW else 0.0)
            out[b, r] = max(t, floor_ms) * 1e6 / nv[b]
        if log:
            log(f"bin {b:3d} ({nv[b]:.0f} vertices): " +
                ", ".join(f"{ROUTE_NAMES[r]} {out[b, r]:.0f}" for r in range(OUT)
                          if not np.isnan(out[b, r])) + " ns/vertex")
    return out, {"t_skip_ms": t_skip, "window_fixed_ms": fixed_w}


def training_rows(X: np.ndarray, ns: np.ndarray, nv: np.ndarray | None = None):
    """(features, targets log2(ns per vertex), mask) rows for the timed bins. With nv, a bin of
    >= 1,000 vertices appears twice and one of >= 10,000 three times (the loss weighs the bins
    that hold the time)."""
    rows = np.nonzero(~np.all(np.isnan(ns), axis=1))[0]
    if nv is not None:
        rows = np.repeat(rows, 1 + (nv[rows] >= 1000) + (nv[rows] >= 10000))
    T = np.where(np.isnan(ns), 0.0, np.log2(np.maximum(ns, 1e-3)))
    M = (~np.isnan(ns)).astype(np.float32)
    return X[rows].astype(np.float32), T[rows].astype(np.float32), M[rows]


def regret(ns: np.ndarray, nv: np.ndarray, routes) -> float:
    """Σ_bins n * (t(chosen) - t(best)) / Σ_bins n * t(best) over the bins timed on every
    route the table could choose."""
    timed = ~np.all(np.isnan(ns), axis=1)
    best = np.nanmin(np.where(timed[:, None], ns, np.inf), axis=1)
    num = den = 0.0
    for b in np.nonzero(timed)[0]:
        r = int(routes[b])
        if r >= OUT or np.isnan(ns[b, r]):
            continue
        num += nv[b] * (ns[b, r] - best[b])
        den += nv[b] * best[b]
    return float(num / den) if den else 0.0


# ---- training ------------------------------------------------------------------------------------

def _log(msg: str) -> None:
    print(f"[c4] {msg}", flush=True)


def prepare(name: str):
    """Ranked symmetric device graph of a named workload: rmatS (edgefactor 16, the seed of the
    hom4 benchmark graph) or clN (Chung-Lu, N * 2^20 vertices, 16 tuples per vertex)."""
    from .device import DeviceGraphCtx
    if name.startswith("rmat"):
        n, s, d = graph.rmat(int(name[4:]), 16, seed=graph.SEEDS["rmat22"])
    elif name.startswith("cl"):
        n = int(name[2:]) << 20
        s, d = graph.gen_chung_lu(n, 16 * n, seed=graph.SEEDS["chung_lu"])
    else:
        raise ValueError(name)
    g = graph.csr_from_edges(n, s, d, graph.SYMMETRIC_RANKED)
    return g, DeviceGraphCtx(g)


def compare(ctx, patterns: int = gq.HOM_C4, repeats: int = 7) -> tuple[float, float]:
    """Median ms of the fixed and the learned routing, timed alternately (the window kernel's
    call-to-call spread is several percent on R-MAT 22)."""
    tf, tl = [], []
    gq.hom4Count(ctx, patterns, c4=gq.C4_FIXED)
    gq.hom4Count(ctx, patterns, c4=gq.C4_LEARNED)
    for _ in range(repeats):
        for mode, acc in ((gq.C4_FIXED, tf), (gq.C4_LEARNED, tl)):
            t0 = time.perf_counter()
            gq.hom4Count(ctx, patterns, c4=mode)
            acc.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(tf)), float(np.median(tl))


def evaluate(name, g, ctx, ns=None, nv=None, repeats: int = 7) -> dict:
    """Fixed vs learned 4-cycle routing (c4 alone and all patterns), equal counts checked;
    same_routes says whether the learned table routes every vertex as the fixed split does."""
    rec = {"vertices": int(g.m), "nnz": int(g.nnz)}
    c_fixed = gq.hom4Count(ctx, gq.HOM_ALL, c4=gq.C4_FIXED)
    c_learned = gq.hom4Count(ctx, gq.HOM_ALL, c4=gq.C4_LEARNED)
    if c_fixed != c_learned:
        raise SystemExit(f"{name}: learned routing changed the counts: {c_fixed} != {c_learned}")
    rec["c4"] = c_fixed["c4"]
    for pat, nm in ((gq.HOM_C4, "c4"), (gq.HOM_ALL, "all")):
        rec[f"{nm}_ms_fixed"], rec[f"{nm}_ms_learned"] = compare(ctx, pat, repeats)
        rec[f"{nm}_learned_vs_fixed"] = rec[f"{nm}_ms_fixed"] / rec[f"{nm}_ms_learned"]
    lt = table(ctx, gq.C4_LEARNED)
    _, nvb = features(ctx)
    fx = fixed_table()
    rec["same_routes"] = bool(np.all((lt == fx) | (nvb == 0)))
    rec["learned_vertices"] = {ROUTE_NAMES[r]: int(nvb[lt == r].sum()) for r in (LIGHT, WINDOW, DENSE)}
    if ns is not None:
        rec["regret_learned"] = regret(ns, nv, lt)
        rec["regret_fixed"] = regret(ns, nv, fixed_table())
    _log(f"{name}: c4 fixed {rec['c4_ms_fixed']:.1f} ms, learned {rec['c4_ms_learned']:.1f} ms "
         f"(x{rec['c4_learned_vs_fixed']:.2f}); all patterns x{rec['all_learned_vs_fixed']:.2f}"
         + ("; learned routes = fixed routes" if rec["same_routes"] else ""))
    return rec


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(MODEL_PATH))
    ap.add_argument("--graphs", default="rmat18,rmat20,rmat22,cl1")
    ap.add_argument("--holdout", default="rmat19,rmat21")
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--margin", type=float, default=0.2)
    ap.add_argument("--min-vertices", type=int, default=64)
    a = ap.parse_args(argv)
    t0 = time.time()
    train_names = a.graphs.split(",")
    hold_names = [h for h in a.holdout.split(",") if h]
    rows_X, rows_T, rows_M, per_graph = [], [], [], {}
    for name in train_names + hold_names:  # held-out graphs are timed for their regret only
        g, ctx = prepare(name)
        X, nv = features(ctx)
        ns, info = time_bins(ctx, X, nv, a.min_vertices, log=_log)
        if name in train_names:
            Xr, Tr, Mr = training_rows(X, ns, nv)
            rows_X.append(Xr), rows_T.append(Tr), rows_M.append(Mr)
        per_graph[name] = (g, ctx, ns, nv, info)
        _log(f"{name}: {int((~np.all(np.isnan(ns), axis=1)).sum())} timed bins, skip "
             f"{info['t_skip_ms']:.2f} ms, window fixed {info['window_fixed_ms']:.2f} ms "
             f"({time.time() - t0:.0f}s)")
    X = np.concatenate(rows_X)
    T = np.concatenate(rows_T)
    M = np.concatenate(rows_M)
    params, loss = train(X, T, M, epochs=a.epochs)
    _log(f"trained on {X.shape[0]} rows: loss {loss:.4f}")
    set_model(params, a.margin)
    report = {}
    for name, (g, ctx, ns, nv, info) in per_graph.items():
        report[name] = evaluate(name, g, ctx, ns, nv)
        report[name].update(info)
        report[name]["held_out"] = name in hold_names
        best = np.nanargmin(np.where(np.isnan(ns), np.inf, ns), axis=1)
        timed = ~np.all(np.isnan(ns), axis=1)
        lt = table(ctx, gq.C4_LEARNED)
        report[name]["per_bin"] = {
            int(b): {"vertices": int(nv[b]), "best": ROUTE_NAMES[int(best[b])],
                     "learned": ROUTE_NAMES[int(lt[b])],
                     "ns_per_vertex": [None if np.isnan(x) else round(float(x), 1) for x in ns[b]]}
            for b in np.nonzero(timed)[0]}
        del ctx
    per_graph.clear()
    doc = {"params": [float(x) for x in params], "margin": a.margin, "in": IN, "out": OUT,
           "routes": [ROUTE_NAMES[r] for r in (LIGHT, WINDOW, DENSE)], "loss": loss,
           "rows": int(X.shape[0]), "graphs": train_names, "held_out": hold_names,
           "report": report}
    Path(a.out).write_text(json.dumps(doc, indent=1))
    _log(f"saved {a.out} ({time.time() - t0:.0f}s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
