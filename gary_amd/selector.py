"""Strategy selector for triangle counting: route tables, per-bin timing, MLP training.

Arcs (u, v) fall in one of 256 bins (min(15, log2 d+(u)), min(15, log2 d+(v))); a route table
sends each bin to the pivot bitmap, the two-phase cached binary search or the merge path.
The plain heuristic is a degree-threshold table; the learned selector is an FP32 MLP trained on
the device, by the library's own kernels, on per-bin timings measured here, and it predicts the
per-arc cost of every route. Counts never depend on the table.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib
from . import graph_query as _gq  # noqa: F401  (binds gary_tc_count)
from ._lib import DeviceGraph, LaunchCfg, check, lib, ptr
from .device import DeviceBuffer, DeviceGraphCtx
from .timing import EventTimer

BITMAP, CACHED, MERGE, SKIP = 0, 1, 2, 3
ROUTE_NAMES = {BITMAP: "bitmap", CACHED: "cached_bsearch", MERGE: "merge", SKIP: "skip"}
NBINS = 256
IN, HID, OUT = 6, 32, 3
NPARAMS = HID * IN + HID + HID * HID + HID + OUT * HID + OUT
MODEL_PATH = Path(__file__).resolve().parent / "selector_model.json"

_P, _F = C.c_void_p, C.c_float
_lib.bind("gary_tc_count_routed", C.c_int, [C.POINTER(DeviceGraph), _P, _P, _F,
                                            C.POINTER(LaunchCfg), _P, _P])
_lib.bind("gary_selector_heuristic_table", C.c_int, [_P])
_lib.bind("gary_selector_set_heuristic", C.c_int, [C.c_int32, C.c_int32, C.c_int32])
_lib.bind("gary_selector_bin_features", C.c_int, [C.POINTER(DeviceGraph), _P, _P])
_lib.bind("gary_mlp_forward", C.c_int, [_P, _P, C.c_int32, _P])
_lib.bind("gary_mlp_loss_grad", C.c_int, [_P, _P, _P, _P, C.c_int32, _P, _P])
_lib.bind("gary_selector_train", C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, _F, C.c_uint64, _P])
_lib.bind("gary_selector_set_params", C.c_int, [_P, C.c_int, _F, C.c_uint32, _F])
_lib.bind("gary_selector_get_params", C.c_int, [_P])
_lib.bind("gary_selector_set_fixed_costs", C.c_int, [C.c_int32, _P, _P])
_lib.bind("gary_selector_learned_table", C.c_int, [C.POINTER(DeviceGraph), _P, _P])
_lib.bind("gary_selector_load", C.c_int, [C.c_char_p])
_lib.bind("gary_selector_trained", C.c_int, [C.POINTER(C.c_int)])


def _dg(g):
    return g.deviceGraph() if hasattr(g, "deviceGraph") else g


def heuristic_table() -> np.ndarray:
    t = np.zeros(NBINS, np.uint8)
    check(lib.gary_selector_heuristic_table(ptr(t)))
    return t


HEURISTIC_DEFAULT = {"bitmap_from": 6, "small_max": 5, "skew": 3}


def set_heuristic(bitmap_from: int = 6, small_max: int = 5, skew: int = 3) -> None:
    """Thresholds of the heuristic table (log2 out-degrees): log2 d+(u) >= bitmap_from -> bitmap;
    both logs <= small_max -> merge; log2 d+(v) - log2 d+(u) >= skew -> cached; else merge."""
    check(lib.gary_selector_set_heuristic(int(bitmap_from), int(small_max), int(skew)))


def dispatch_table() -> np.ndarray:
    """The reference dispatcher's four cases as a route table, without the bitmap route
    (include/ops/Intersection.cuh:182-187): both lists small (< 64) -> merge; one small, one
    large, or both large and skewed (>= 8x) -> binary search (cached); both large and even ->
    merge."""
    t = np.zeros(NBINS, np.uint8)
    for b in range(NBINS):
        i, j = b >> 4, b & 15
        if i <= 5 and j <= 5:
            t[b] = MERGE
        elif abs(j - i) >= 3:
            t[b] = CACHED
        else:
            t[b] = MERGE
    return t


def bin_of(du: int, dv: int) -> int:
    return min(15, int(du).bit_length() - 1) * 16 + min(15, int(dv).bit_length() - 1)


def count_with_table(g, routes, costs=None, cost_big: float = 0.0, cfg: LaunchCfg | None = None,
                     d_count: DeviceBuffer | None = None) -> int:
    """Exact triangle count with every arc routed by `routes` (256 entries)."""
    routes = np.ascontiguousarray(routes, np.uint8)
    costs = None if costs is None else np.ascontiguousarray(costs, np.float32)
    own = d_count is None
    if own:
        d_count = DeviceBuffer(8)
        d_count.zero()
    dg = _dg(g)
    check(lib.gary_tc_count_routed(C.byref(dg), ptr(routes), ptr(costs), cost_big,
                                   C.byref(cfg) if cfg else None, d_count.ptr, None))
    if own:
        return int(d_count.to_array(np.uint64, 1)[0])
    return -1


def bin_features(g) -> tuple[np.ndarray, np.ndarray]:
    X = np.zeros((NBINS, IN), np.float32)
    arcs = np.zeros(NBINS, np.float32)
    dg = _dg(g)
    check(lib.gary_selector_bin_features(C.byref(dg), ptr(X), ptr(arcs)))
    return X, arcs


def exact_bin_arcs(host_csr) -> np.ndarray:
    """Exact arcs per bin from a host CSR (arcs without work are not counted)."""
    dp = host_csr.deg.astype(np.int64)
    row = np.repeat(np.arange(host_csr.m), host_csr.deg)
    a, b = dp[row], dp[host_csr.colIdx]
    keep = (a >= 2) & (b >= 1)
    la = np.minimum(15, np.floor(np.log2(np.maximum(a[keep], 1))).astype(np.int64))
    lb = np.minimum(15, np.floor(np.log2(np.maximum(b[keep], 1))).astype(np.int64))
    return np.bincount(la * 16 + lb, minlength=NBINS).astype(np.int64)


def time_table(g, routes, costs=None, cost_big=0.0, repeats=5, cfg=None) -> float:
    """Milliseconds per routed count (HIP events, after one warm-up)."""
    d = DeviceBuffer(8)
    d.zero()
    count_with_table(g, routes, costs, cost_big, cfg, d)
    t = EventTimer()
    t.start()
    for _ in range(repeats):
        count_with_table(g, routes, costs, cost_big, cfg, d)
    return t.stop() / repeats


def time_strategy(g, strategy: int, repeats: int = 5, cfg: LaunchCfg | None = None) -> float:
    """Milliseconds per count with a gary_tc_count strategy (features and table included)."""
    dg = _dg(g)
    d = DeviceBuffer(8)
    d.zero()
    pc = C.byref(cfg) if cfg else None
    check(lib.gary_tc_count(C.byref(dg), strategy, pc, d.ptr, None))
    t = EventTimer()
    t.start()
    for _ in range(repeats):
        check(lib.gary_tc_count(C.byref(dg), strategy, pc, d.ptr, None))
    return t.stop() / repeats


def time_bins(g, arcs: np.ndarray, min_arcs: int = 4096, repeats: int = 3,
              routes=(BITMAP, CACHED, MERGE)) -> tuple[np.ndarray, np.ndarray]:
    """ns per arc of every route on every bin with >= min_arcs arcs (NaN elsewhere), and the
    fixed cost (ns) of each route.

    Edge routes (cached, merge) cost the same per arc whatever else runs, so each bin is timed
    alone against the all-skip table. The bitmap route shares each pivot's set between the
    pivot's bins, so it is timed in context: cost(b) = t(all bitmap) - t(all bitmap, b skipped);
    its fixed cost is t(all bitmap) - t(all skip) - Σ_b arcs[b] * cost(b)."""
    out = np.full((NBINS, OUT), np.nan, np.float64)
    fixed = np.zeros(OUT, np.float64)
    timed = np.nonzero(arcs >= min_arcs)[0]
    skip = np.full(NBINS, SKIP, np.uint8)
    t_skip = min(time_table(g, skip, repeats=repeats) for _ in range(2))
    for r in routes:
        if r == BITMAP:
            base = np.full(NBINS, BITMAP, np.uint8)
            t_all = min(time_table(g, base, repeats=repeats) for _ in range(2))
            acc = 0.0
            for b in timed:
                tab = base.copy()
                tab[b] = SKIP
                out[b, r] = max(t_all - time_table(g, tab, repeats=repeats), 1e-4) * 1e6 / arcs[b]
                acc += out[b, r] * arcs[b]
            fixed[r] = max(0.0, (t_all - t_skip) * 1e6 - acc)
        else:
            for b in timed:
                tab = skip.copy()
                tab[b] = r
                out[b, r] = max(time_table(g, tab, repeats=repeats) - t_skip, 1e-4) * 1e6 / arcs[b]
    return out, fixed


def set_fixed_costs(nnz_points, costs) -> None:
    """Fixed route costs (ns): costs[k][route] at graph size nnz_points[k] (increasing)."""
    x = np.ascontiguousarray(nnz_points, np.float32)
    c = np.ascontiguousarray(costs, np.float32).reshape(-1, OUT)
    assert x.size == c.shape[0] <= 4
    check(lib.gary_selector_set_fixed_costs(int(x.size), ptr(x), ptr(c)))


def training_rows(X: np.ndarray, ns_per_arc: np.ndarray):
    """(features, targets log2(ns/arc), mask) rows for the bins that were timed."""
    rows = ~np.all(np.isnan(ns_per_arc), axis=1)
    T = np.where(np.isnan(ns_per_arc), 0.0, np.log2(np.maximum(ns_per_arc, 1e-6)))
    M = (~np.isnan(ns_per_arc)).astype(np.float32)
    return X[rows].astype(np.float32), T[rows].astype(np.float32), M[rows]


def train(X, T, M, epochs: int = 3000, lr: float = 3e-3, seed: int = 0x5E1EC7) -> float:
    """Full-batch Adam on the device; returns the final loss. The model is not yet enabled."""
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    M = np.ascontiguousarray(M, np.float32)
    loss = C.c_float(0)
    check(lib.gary_selector_train(ptr(X), ptr(T), ptr(M), X.shape[0], epochs, lr, seed,
                                  C.byref(loss)))
    return float(loss.value)


def get_params() -> np.ndarray:
    p = np.zeros(NPARAMS, np.float32)
    check(lib.gary_selector_get_params(ptr(p)))
    return p


def set_params(p, trained: bool = True, cost_big: float = 0.0, allowed: int = 7,
               margin: float = 0.0) -> None:
    p = np.ascontiguousarray(p, np.float32)
    assert p.size == NPARAMS
    check(lib.gary_selector_set_params(ptr(p), 1 if trained else 0, cost_big, allowed, margin))


def learned_table(g) -> tuple[np.ndarray, np.ndarray]:
    r = np.zeros(NBINS, np.uint8)
    c = np.zeros(NBINS, np.float32)
    dg = _dg(g)
    check(lib.gary_selector_learned_table(C.byref(dg), ptr(r), ptr(c)))
    return r, c


def mlp_forward(p, X) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.zeros((X.shape[0], OUT), np.float32)
    check(lib.gary_mlp_forward(ptr(p), ptr(X), X.shape[0], ptr(Y)))
    return Y


def mlp_loss_grad(p, X, T, M) -> tuple[float, np.ndarray]:
    p = np.ascontiguousarray(p, np.float32)
    X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
    g = np.zeros(NPARAMS, np.float32)
    loss = C.c_float(0)
    check(lib.gary_mlp_loss_grad(ptr(p), ptr(X), ptr(T), ptr(M), X.shape[0], C.byref(loss),
                                 ptr(g)))
    return float(loss.value), g


def regret(ns_per_arc: np.ndarray, arcs: np.ndarray, routes) -> float:
    """Σ_bins arcs * (t(chosen) - t(best)) / Σ_bins arcs * t(best) over the timed bins."""
    timed = ~np.all(np.isnan(ns_per_arc), axis=1)
    best = np.nanmin(np.where(timed[:, None], ns_per_arc, np.inf), axis=1)
    chosen = np.array([ns_per_arc[b, routes[b]] if timed[b] and routes[b] < OUT else np.nan
                       for b in range(NBINS)])
    ok = timed & ~np.isnan(chosen)
    return float((arcs[ok] * (chosen[ok] - best[ok])).sum() / (arcs[ok] * best[ok]).sum())


def save(path: Path = MODEL_PATH, fixed_nnz=(), fixed_costs=(), **meta) -> None:
    doc = {"params": [float(x) for x in get_params()], "in": IN, "hidden": HID, "out": OUT,
           "fixed_nnz": [float(x) for x in fixed_nnz],
           "fixed_costs": [[float(v) for v in row] for row in fixed_costs]}
    doc.update(meta)
    Path(path).write_text(json.dumps(doc, indent=1))


def load(path: Path = MODEL_PATH, required: bool = True) -> dict | None:
    """Loads a saved model into the library (gary_selector_load; enables the learned strategy)
    and returns the file's contents. A missing file raises FileNotFoundError, or returns None
    when required=False."""
    p = Path(path)
    if not p.exists():
        if required:
            raise FileNotFoundError(f"no selector model at {p}")
        return None
    check(lib.gary_selector_load(str(p).encode()))
    if not trained():
        raise RuntimeError(f"{p} loaded but the learned strategy is not enabled")
    return json.loads(p.read_text())


def trained() -> bool:
    """Whether GARY_TC_AUTO_LEARNED runs a model on the current device (else the heuristic)."""
    t = C.c_int(0)
    check(lib.gary_selector_trained(C.byref(t)))
    return bool(t.value)
