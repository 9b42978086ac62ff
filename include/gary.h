/*
 * gary.h — C ABI of the MI355X (gfx950) graph pattern-matching library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int status
 * (GARY_OK = 0, negative = error class) and never throws or exits. A thread-local message for
 * the last failure is available from gary_last_error().
 *
 * Counts are always uint64_t. Device pointers are HIP device pointers of the device the
 * context / call is bound to; `stream` is a hipStream_t passed as void* (NULL = null stream).
 * Scratch memory is owned by the library and kept between calls: per (device, stream) for the
 * triangle counts and the device build, per device for the synchronous calls (4-vertex counts,
 * k-cliques, listing, selector training), so the caller serializes those on one device.
 *
 * Reference interfaces replaced (hirayaku/Gary, paths relative to the reference root):
 *   DeviceCSR / DeviceCOO / DeviceGraph ........ include/DeviceGraph.cuh:15-72
 *   DeviceGraphCtx ............................. include/DeviceGraph.cuh:75-133, src/DeviceGraph.cu:14-96
 *   binary_search / build_cache / 2-phase ...... include/ops/Search.cuh:33-138 (test kernels test/SearchTest.cu:14-183)
 *   intersectBinarySearch / intersect .......... include/ops/Intersection.cuh:18-199, 239-300
 *   triangleCountHom / triangleCountLegacy ..... include/kernels/GraphQueryHom.cuh:15-43, src/GraphQueryHom.cu:15-38
 *   getIndexCacheSMem .......................... src/GraphQueryHom.cu:11-13
 *   rectCountHom (declared only there) ......... include/kernels/GraphQueryHom.cuh:48
 *   cpuTriCount's CSR input (COO -> CSR) ....... src/Graph.cpp:19-91, 141-155
 */
#ifndef GARY_H_
#define GARY_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- type ABI (include/Defines.h:9-10) --------------------------------------------------- */
typedef int32_t gary_vid_t; /* vidT */
typedef int64_t gary_eid_t; /* eidT */

/* ---- status codes ------------------------------------------------------------------------- */
enum {
  GARY_OK = 0,
  GARY_ERR_INVALID = -1,     /* bad argument (null pointer, bad size, bad enum, invalid CSR) */
  GARY_ERR_HIP = -2,         /* a HIP runtime call failed */
  GARY_ERR_OOM = -3,         /* device or host allocation failed */
  GARY_ERR_OVERFLOW = -4,    /* result does not fit the requested output type */
  GARY_ERR_UNSUPPORTED = -5, /* valid request this build does not handle */
  GARY_ERR_STATE = -6        /* object not initialised / wrong state */
};

/* ---- device graph views (layout-identical to DeviceCSR / DeviceCOO / DeviceGraph,
 *      include/DeviceGraph.cuh:15-72) ----------------------------------------------------- */
typedef struct gary_device_csr {
  gary_vid_t m;       /* number of vertices */
  gary_eid_t nnz;     /* number of stored (directed) edges */
  gary_eid_t *indPtr; /* [m+1] row pointers */
  gary_vid_t *colIdx; /* [nnz] sorted adjacency lists */
  gary_vid_t *deg;    /* [m] list lengths */
} gary_device_csr;

typedef struct gary_device_coo {
  gary_vid_t m;
  gary_eid_t nnz;
  gary_vid_t *rowIdx; /* [nnz] source of each edge */
  gary_vid_t *colIdx; /* [nnz] shared with the CSR */
  gary_vid_t *deg;    /* [m]   shared with the CSR */
} gary_device_coo;

typedef struct gary_device_graph {
  gary_device_csr csr;
  gary_device_coo coo;
} gary_device_graph;

/* ---- errors / device ----------------------------------------------------------------------- */
const char *gary_last_error(void);
int gary_version(void);
int gary_device_count(int *count);
int gary_device_synchronize(int device);

/* ---- host graph construction (src/Graph.cpp:19-91, 141-155 semantics, plus dedup and
 *      degree orientation) -------------------------------------------------------------------- */

/* Graph generators. All are deterministic functions of their arguments (splitmix64 streams,
 * integer-only sampling). Output arrays hold `ntuples` directed tuples (u, v); loops and
 * duplicates are left in for gary_csr_from_edges to remove. */
int gary_gen_er(int32_t n, int64_t m, uint64_t seed, int32_t *src, int32_t *dst);
int gary_gen_rmat(int32_t scale, int64_t ntuples, double a, double b, double c, uint64_t seed,
                  int32_t *src, int32_t *dst);
/* Chung-Lu: expected degrees w_i ~ (i + i0)^(-1/(gamma-1)), capped at max_w, `ntuples`
 * endpoint pairs drawn with probability proportional to w. */
int gary_gen_chung_lu(int32_t n, int64_t ntuples, double gamma, double max_w, uint64_t seed,
                      int32_t *src, int32_t *dst);

enum {
  GARY_CSR_ORIENT_DEGREE = 0, /* u -> v iff (deg u, u) < (deg v, v); one arc per undirected edge */
  GARY_CSR_SYMMETRIC = 1,     /* both arcs of every undirected edge */
  GARY_CSR_ORIENT_ID = 2,     /* u -> v iff u > v (the reference's GraphCOO::orientation) */
  GARY_CSR_SYMMETRIC_RANKED = 3 /* both arcs; vertex ids relabeled by (degree, old id) rank, so
                                   id order is degree order and the upper part of each row is
                                   the degree orientation (input of gary_hom4_count) */
};

/* Symmetrize, drop self-loops, dedup, then build a CSR in `mode`. Lists are sorted ascending.
 * indPtr has n+1 entries; colIdx must hold `cap` entries (ntuples suffices for the oriented
 * modes, 2*ntuples for the symmetric ones). *nnz_out receives the stored edge count. */
int gary_csr_from_edges(int32_t n, int64_t ntuples, const int32_t *src, const int32_t *dst,
                        int mode, int64_t *indPtr, int32_t *colIdx, int64_t cap,
                        int64_t *nnz_out);

/* ---- graph files (the reference's saveToBinary / loadFromBinary and loadFromSNAP,
 *      include/Graph.h:16-23,133-148,270-285,301-320, src/Graph.cpp:157-196) ------------------ */
enum { GARY_GRAPH_FILE_CSR = 0, GARY_GRAPH_FILE_COO = 1 }; /* GraphFormat */

/* Header of a binary graph file (cereal binary archive, SerDesVersion V1): its format, m and nnz.
 * Files written by the reference's saveToBinary<GraphCSR / GraphCOO> read unchanged. */
int gary_graph_file_info(const char *path, int *format, int32_t *m, int64_t *nnz);
/* A GraphCSR file into indPtr[m+1], colIdx[nnz], deg[m] (a NULL array is skipped). */
int gary_graph_file_read_csr(const char *path, int64_t *indPtr, int32_t *colIdx, int32_t *deg);
/* A GraphCOO file into row[nnz], col[nnz]. */
int gary_graph_file_read_coo(const char *path, int32_t *row, int32_t *col);
/* Write a GraphCSR file the reference's loadFromBinary<GraphCSR> reads (deg NULL: derived). */
int gary_graph_file_write_csr(const char *path, int32_t m, int64_t nnz, const int64_t *indPtr,
                              const int32_t *colIdx, const int32_t *deg);
/* Write a GraphCOO file the reference's loadFromBinary<GraphCOO> reads. */
int gary_graph_file_write_coo(const char *path, int32_t m, int64_t nnz, const int32_t *row,
                              const int32_t *col);
/* SNAP edge list: one "src dst" pair per line, '#' comments and blank lines skipped, loops
 * dropped when remove_loops. Call with src = dst = NULL for *ntuples and *n (max id + 1, at
 * least 1), then with arrays of cap >= *ntuples entries. */
int gary_snap_read(const char *path, int remove_loops, int32_t *src, int32_t *dst, int64_t cap,
                   int64_t *ntuples, int32_t *n);
/* GraphCOO::toCSR_ (src/Graph.cpp:63-91) on the host: indPtr[m+1] and colIdx[nnz] with rows by
 * source id and each row sorted ascending; duplicates and loops are kept (no dedup, unlike
 * gary_csr_from_edges). Ids outside [0, m) fail with GARY_ERR_INVALID. The reference's `convert`
 * tool (convert.cpp:13-78) is `gary_amd/gary_convert`, built on these file functions. */
int gary_coo_to_csr(int32_t m, int64_t nnz, const int32_t *row, const int32_t *col,
                    int64_t *indPtr, int32_t *colIdx);

/* ---- device graph construction (the same builds on the GPU; the reference has host code only:
 *      src/Graph.cpp:19-91, common/ParRadixSort.h:218-306, include/Graph.h:242-248) -------- */

/* gary_gen_rmat on the device: the same tuples bit for bit, written to device arrays src/dst
 * (ntuples each) on `stream`. Asynchronous. */
int gary_gen_rmat_device(int32_t scale, int64_t ntuples, double a, double b, double c,
                         uint64_t seed, int32_t *src, int32_t *dst, void *stream);

/* gary_gen_chung_lu on the device: the same tuples bit for bit (the cumulative weight table is
 * computed on the host and uploaded). Returns when the tuples are written. */
int gary_gen_chung_lu_device(int32_t n, int64_t ntuples, double gamma, double max_w,
                             uint64_t seed, int32_t *src, int32_t *dst, void *stream);

/* gary_csr_from_edges with device arrays: src/dst (ntuples each) in, indPtr[n+1] and colIdx
 * (capacity cap) out, all device pointers; identical output for every mode. Both arcs of every
 * tuple are sorted, so 2*ntuples is limited to 2^31 - 1. Returns when the CSR is complete (it
 * reads the deduplicated edge count back to the host for *nnz_out). Scratch (about 2 x 16 bytes
 * per tuple) is a library buffer per (device, stream), kept for later builds and separate from
 * the counting calls' one. */
int gary_csr_from_edges_device(int32_t n, int64_t ntuples, const int32_t *src,
                               const int32_t *dst, int mode, int64_t *indPtr, int32_t *colIdx,
                               int64_t cap, int64_t *nnz_out, void *stream);

/* ---- device graph context (DeviceGraphCtx) ------------------------------------------------- */
typedef struct gary_ctx gary_ctx;

int gary_ctx_create(int device, gary_ctx **out);
int gary_ctx_destroy(gary_ctx *ctx);
/* loadFromHost(shared_ptr<GraphCSR>): H2D of indPtr and colIdx; deg is derived on the device. */
int gary_ctx_load_csr(gary_ctx *ctx, int32_t m, int64_t nnz, const int64_t *indPtr,
                      const int32_t *colIdx);
/* Builds the ctx's CSR on its device from device tuple arrays (gary_csr_from_edges_device), so
 * nothing crosses PCIe; replaces loadFromHost(GraphCOO -> GraphCSR) of the reference's host
 * path (src/Graph.cpp:19-91, src/DeviceGraph.cu:18-28). */
int gary_ctx_build_from_edges(gary_ctx *ctx, int32_t n, int64_t ntuples, const int32_t *src,
                              const int32_t *dst, int mode, void *stream);
/* populateDevice(): derive the COO rowIdx on the device. */
int gary_ctx_populate(gary_ctx *ctx);
/* getDeviceCSR() + getDeviceCOO() in one view; valid while the ctx lives and is not reloaded. */
int gary_ctx_device_graph(gary_ctx *ctx, gary_device_graph *out);
/* getHostCSR() after detach: D2H copy of m+1 row pointers, nnz column indices, m degrees. */
int gary_ctx_copy_to_host(gary_ctx *ctx, int64_t *indPtr, int32_t *colIdx, int32_t *deg);
/* clear(): release device buffers. */
int gary_ctx_clear(gary_ctx *ctx);

/* ---- event timer (utils::CUDATimer, common/Timer.h:38-64, common/Timer.cpp:8-36) ---------- */
typedef struct gary_timer gary_timer;
int gary_timer_create(gary_timer **out);
int gary_timer_destroy(gary_timer *t);
/* Records the start event on `stream`. */
int gary_timer_start(gary_timer *t, void *stream);
/* Records the stop event on `stream`, waits for it and returns the elapsed milliseconds. */
int gary_timer_stop(gary_timer *t, void *stream, float *ms);

/* Opt-in HIP-event timing of the dominant counting kernel of each gary_tc_count call (the main
 * pivot kernel for the pivot strategies, the edge kernel for the edge-centric ones): events are
 * recorded on the launch stream around that kernel only. enable != 0 turns it on and clears the
 * record; gary_kernel_timing_read waits for the recorded events, returns the number of launches
 * and their summed milliseconds, and clears the record. */
int gary_kernel_timing(int enable);
int gary_kernel_timing_read(int *launches, float *total_ms);

/* ---- device memory helpers (so hosts without a HIP toolchain can drive the ABI) --------- */
int gary_malloc(int device, size_t bytes, void **ptr);
int gary_free(int device, void *ptr);
int gary_memcpy_h2d(int device, void *dst, const void *src, size_t bytes);
int gary_memcpy_d2h(int device, void *dst, const void *src, size_t bytes);
int gary_memcpy_d2d(int device, void *dst, const void *src, size_t bytes);
int gary_memset(int device, void *dst, int value, size_t bytes);
/* Streams for callers without a HIP toolchain: a non-blocking stream (default priority) of
 * `device`, its destruction, and a wait for its queued work. */
int gary_stream_create(int device, void **stream);
int gary_stream_destroy(int device, void *stream);
int gary_stream_synchronize(int device, void *stream);

/* ---- search and intersection operators (include/ops/Search.cuh, Intersection.cuh) ------- */
enum {
  GARY_SEARCH_PLAIN = 0,  /* binary_search, Search.cuh:48-61 */
  GARY_SEARCH_2PHASE = 1  /* build_cache + binary_search_2phase, Search.cuh:93-138 */
};
/* Per-key membership of keys[i] in the sorted list (binarySearchKernel /
 * binarySearch2PhaseKernel, test/SearchTest.cu:14-67). results[i] is 0/1. cache_depth is N of
 * Array<T,N> (power of two, 8..1024) for GARY_SEARCH_2PHASE and ignored otherwise. */
int gary_search(const int32_t *d_list, int64_t len, const int32_t *d_keys, int64_t nkeys,
                uint8_t *d_results, int mode, int cache_depth, void *stream);
/* Number of keys found (binarySearchCountKernel / binarySearch2PhaseCountKernel), accumulated
 * into *d_count. */
int gary_search_count(const int32_t *d_list, int64_t len, const int32_t *d_keys, int64_t nkeys,
                      uint64_t *d_count, int mode, int cache_depth, void *stream);
/* Batched search, one list per item (batchedBinarySearch{,2Phase}Kernel,
 * test/SearchTest.cu:69-136). All pointer arrays are device arrays of device pointers. */
int gary_search_batched(int32_t batch, const int32_t *const *d_lists, const int64_t *d_list_sizes,
                        const int32_t *const *d_keys, const int64_t *d_key_sizes,
                        uint8_t *const *d_results, int mode, int cache_depth, void *stream);

enum {
  GARY_ISECT_BSEARCH = 0,        /* intersectBinarySearch(sA,sB), Intersection.cuh:56-71 */
  GARY_ISECT_CACHED_BSEARCH = 1, /* intersectBinarySearch(sA,sB,cache), :118-135 */
  GARY_ISECT_MERGE = 2,          /* wave merge-path (the reference's intersectMerge is a stub) */
  GARY_ISECT_BITMAP = 3          /* LDS bitmap over the shorter list's id window */
};
/* |A ∩ B| of two sorted duplicate-free device lists, accumulated into *d_count. */
int gary_intersect_count(const int32_t *d_a, int64_t na, const int32_t *d_b, int64_t nb,
                         int method, int cache_depth, uint64_t *d_count, void *stream);
/* Intersection with output (Intersection.cuh:18-50, 76-112, 239-271). Writes |A ∩ B| values
 * to d_out (capacity min(na, nb)) in ascending order, and the count to *d_count (overwritten). */
int gary_intersect_emit(const int32_t *d_a, int64_t na, const int32_t *d_b, int64_t nb,
                        int method, int cache_depth, int32_t *d_out, uint64_t *d_count,
                        void *stream);

/* ---- triangle counting (include/kernels/GraphQueryHom.cuh:15-43, src/GraphQueryHom.cu) --- */
enum {
  GARY_TC_AUTO_LEARNED = 0,   /* MLP strategy/cost selector (falls back to HEURISTIC untrained) */
  GARY_TC_HEURISTIC = 1,      /* degree-binned, plain degree-threshold strategy choice */
  GARY_TC_MERGE = 2,          /* every edge intersected by wave merge-path */
  GARY_TC_CACHED_BSEARCH = 3, /* triangleCountHom<true,N,Policy>: two-phase LDS-cached search */
  GARY_TC_BSEARCH = 4,        /* triangleCountHom<false>: plain binary search */
  GARY_TC_BITMAP = 5,         /* pivot-vertex LDS bitmap/hash of N+(u), probes N+(v) */
  GARY_TC_NUM_STRATEGIES = 6
};
enum {
  GARY_PARTITION_DYNAMIC = -1,    /* WorkPartition::DYNAMIC (DeviceIterator.cuh:47-51, declared
                                     there but rejected by a static_assert at :141): groups claim
                                     chunks of edges from a device-wide atomic cursor */
  GARY_PARTITION_INTERLEAVED = 0, /* WorkPartition::INTERLEAVED, DeviceIterator.cuh:117-131 */
  GARY_PARTITION_CHUNKED = 1      /* WorkPartition::CHUNKED */
};

/* Arc direction of the pivot strategies (BITMAP and the bitmap route of HEURISTIC /
 * AUTO_LEARNED). MIN_DIRECTION streams each arc (u, v) from the endpoint with the shorter
 * out-list: N+(u) through the set of N+(v) at pivot v when d+(u) is shorter by more than a small
 * slack, which needs a per-call grouping of those arcs by v (16 bytes per arc and 9 per vertex
 * of scratch; the graph's COO rowIdx; m <= 2^24 and nnz < 2^32). FORWARD streams every N+(v) at pivot u, with
 * no grouping. AUTO (the default) is MIN_DIRECTION from 2^20 arcs on and FORWARD below, where the
 * grouping's fixed cost of a few launches outweighs what it saves. Counts are the same in every
 * mode. */
enum { GARY_PIVOT_AUTO = 0, GARY_PIVOT_FORWARD = 1, GARY_PIVOT_MIN_DIRECTION = 2 };

/* Launch knobs. Zero fields take the library default (or the autotuned value). */
typedef struct gary_launch_cfg {
  int32_t block;       /* threads per workgroup of the edge-centric kernels (64, 128, 192, 256) */
  int32_t grid;        /* workgroups (0 = library default) */
  int32_t group;       /* lanes per edge in the edge-centric kernels: 8, 16, 32 or 64 */
  int32_t cache_depth; /* N of the per-group LDS index cache (2-phase search), power of two */
  int32_t partition;   /* GARY_PARTITION_* for the edge-centric strategies */
  int32_t bin_small;   /* pivots with out-degree <= bin_small use the small-table wave kernel */
  int32_t bin_hub;     /* pivots with out-degree > bin_hub get a whole workgroup */
  int32_t direction;   /* GARY_PIVOT_* for the pivot strategies (0 = AUTO) */
} gary_launch_cfg;

/* Elements (not bytes) of LDS index cache for a workgroup of `block` threads with depth
 * `cache_size` (getIndexCacheSMem, src/GraphQueryHom.cu:11-13; waves of 64 lanes). */
int64_t gary_index_cache_smem(int32_t block, int64_t cache_size);

/* Triangle count of an acyclically oriented graph: Σ over arcs (u,v) of |N+(u) ∩ N+(v)|.
 * Accumulates into the device counter *d_count (the caller zeroes it), on `stream`. cfg may be
 * NULL. The graph's COO (populate) is required by the edge-centric strategies. Scratch comes
 * from a library buffer per (device, stream) that grows for larger graphs: a hipGraph captured
 * from this call points into it, so re-capture after a call on that stream with a larger graph. */
int gary_tc_count(const gary_device_graph *g, int strategy, const gary_launch_cfg *cfg,
                  uint64_t *d_count, void *stream);
/* Synchronous convenience: returns the count in *h_count. */
int gary_tc_count_sync(const gary_device_graph *g, int strategy, const gary_launch_cfg *cfg,
                       uint64_t *h_count);
/* triangleCountLegacy(DeviceGraph, unsigned*): accumulates into a 32-bit device counter.
 * Unlike the reference it counts in 64 bits and fails with GARY_ERR_OVERFLOW instead of
 * wrapping when the sum would not fit. Synchronizes `stream`. */
int gary_tc_count_legacy(const gary_device_graph *g, uint32_t *d_count, void *stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, hosts that forbid hipMalloc): gary_tc_count with
 * strategy GARY_TC_BITMAP (cfg: block, grid, bin_small, bin_hub as there; may be NULL), whose
 * pivot lists and counters live in a fixed 16 MiB device array that each GPU holds until the
 * process ends, so the call allocates nothing. Graphs of more than 1,397,568 vertices fail with
 * GARY_ERR_INVALID. Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race. Accumulates into *d_count on `stream`, like gary_tc_count. */
int gary_tc_count_noalloc(const gary_device_graph *g, const gary_launch_cfg *cfg,
                          uint64_t *d_count, void *stream);

/* ---- strategy selection: routed counting, heuristic, learned MLP selector ------------------ */
/* No counterpart in the reference, whose dispatcher intersect() names four size cases but always
 * binary-searches (include/ops/Intersection.cuh:182-193). Arcs (u,v) fall in bin 16*min(15, log2 d+(u)) + min(15, log2 d+(v)). A route table assigns each
 * of the 256 bins one route; counts are exact for every table. */
enum {
  GARY_ROUTE_BITMAP = 0,         /* streamed through the pivot's LDS set */
  GARY_ROUTE_CACHED_BSEARCH = 1, /* edge-centric two-phase cached binary search */
  GARY_ROUTE_MERGE = 2,          /* edge-centric merge path */
  GARY_ROUTE_SKIP = 3,           /* not counted (per-bin timing) */
  GARY_ROUTE_BINS = 256
};
/* Counts with a caller-given table (h_routes[256], optional predicted ns per arc h_costs[256]);
 * pivots whose predicted bitmap cost exceeds cost_big ns get a workgroup (0: degree split).
 * Requires the COO (populate). Accumulates into *d_count. */
int gary_tc_count_routed(const gary_device_graph *g, const uint8_t *h_routes,
                         const float *h_costs, float cost_big, const gary_launch_cfg *cfg,
                         uint64_t *d_count, void *stream);
/* The plain degree-threshold table used by GARY_TC_HEURISTIC (with the current thresholds). */
int gary_selector_heuristic_table(uint8_t *h_routes);
/* Thresholds of that table (one setting for the process, every device), as log2 out-degrees of
 * the arc (u,v): log2 d+(u) >= bitmap_from -> bitmap; else both logs <= small_max -> merge;
 * log2 d+(v) - log2 d+(u) >= skew -> cached binary search; else merge. Library defaults:
 * (6, 5, 3); the autotuner searches them per workload. Waits for each device that already holds
 * a table (no count in flight reads it while it changes). */
int gary_selector_set_heuristic(int32_t bitmap_from, int32_t small_max, int32_t skew);

/* MLP selector: 6 bin features -> 32 -> 32 (ReLU) -> 3 predicted log2(ns per arc) for
 * {bitmap, cached binary search, merge}. Parameters: W1[32][6] b1[32] W2[32][32] b2[32]
 * W3[3][32] b3[3] (GARY_MLP_PARAMS floats). */
#define GARY_MLP_IN 6
#define GARY_MLP_HIDDEN 32
#define GARY_MLP_OUT 3
#define GARY_MLP_PARAMS (GARY_MLP_HIDDEN * GARY_MLP_IN + GARY_MLP_HIDDEN + \
                         GARY_MLP_HIDDEN * GARY_MLP_HIDDEN + GARY_MLP_HIDDEN + \
                         GARY_MLP_OUT * GARY_MLP_HIDDEN + GARY_MLP_OUT)
/* Features of the 256 bins of a graph (h_features[256][6]); h_arcs[256] = arcs per bin
 * (sampled estimate), may be NULL. Requires the COO. */
int gary_selector_bin_features(const gary_device_graph *g, float *h_features, float *h_arcs);
/* Forward pass on the device: h_y[n][3] = MLP(h_x[n][6]). */
int gary_mlp_forward(const float *h_params, const float *h_x, int32_t n, float *h_y);
/* Mean over n samples of 0.5 * Σ_o mask[o] (y_o - t_o)^2 and its gradient (device kernels). */
int gary_mlp_loss_grad(const float *h_params, const float *h_x, const float *h_t,
                       const float *h_mask, int32_t n, float *h_loss, float *h_grad);
/* Full-batch Adam on the device-resident model; seed != 0 re-initialises it first. */
int gary_selector_train(const float *h_x, const float *h_t, const float *h_mask, int32_t n,
                        int32_t epochs, float lr, uint64_t seed, float *h_loss);
/* The same training from a fresh initialisation (seed != 0) into h_params (GARY_MLP_PARAMS
 * floats), leaving every model of the library as it is (trains the 4-cycle routing model). */
int gary_mlp_train(const float *h_x, const float *h_t, const float *h_mask, int32_t n,
                   int32_t epochs, float lr, uint64_t seed, float *h_params, float *h_loss);
/* Model parameters in/out; `trained` enables GARY_TC_AUTO_LEARNED (else it runs the heuristic
 * table); `allowed` is a bit mask of the routes the model may pick; the learned table leaves the
 * all-bitmap routing only when another route subset is predicted to save more than `margin`
 * (a fraction, 0 <= margin < 1) of its time. */
int gary_selector_set_params(const float *h_params, int trained, float cost_big, uint32_t allowed,
                             float margin);
int gary_selector_get_params(float *h_params);
/* Loads a selector file (the JSON object gary_amd/selector.py writes: "params" with
 * GARY_MLP_PARAMS numbers, optional "cost_big", "allowed", "margin", "fixed_nnz" and
 * "fixed_costs") into the current device and enables GARY_TC_AUTO_LEARNED; a missing or malformed
 * file fails with GARY_ERR_INVALID and leaves the model unchanged. */
int gary_selector_load(const char *path);
/* *trained = 1 when GARY_TC_AUTO_LEARNED runs a model on the current device, 0 when it falls
 * back to the heuristic table. */
int gary_selector_trained(int *trained);
/* Fixed cost (ns) of using each route at all, as up to 4 points (h_nnz[k] increasing,
 * h_costs[k][3]) interpolated linearly in log2(1 + nnz); the learned table picks the route subset
 * with the least predicted total. */
int gary_selector_set_fixed_costs(int32_t npoints, const float *h_nnz, const float *h_costs);
/* The table GARY_TC_AUTO_LEARNED would use on this graph (fused features + MLP inference). */
int gary_selector_learned_table(const gary_device_graph *g, uint8_t *h_routes, float *h_costs);

/* ---- batched triangle counting (one persistent launch over many graphs, hipGraph) -------- */
/* No counterpart in the reference; its batched shapes are the batched search test kernels
 * (test/SearchTest.cu:69-136) and the block-scope queue CTAQueue (include/ops/DeviceQueue.cuh:33-182),
 * whose role the persistent kernels' atomic work cursors take here. */
typedef struct gary_batch gary_batch;
/* The batch is the disjoint union of its graphs in one oriented CSR: graph g owns the vertex ids
 * [vbase[g], vbase[g+1]) with vbase the prefix sum of graph_vertices, and every arc stays inside
 * its graph (checked). indPtr has Σ graph_vertices + 1 entries. All device memory is allocated
 * here; counting allocates nothing. */
int gary_batch_create(int device, int32_t ngraphs, const int32_t *graph_vertices, int64_t nnz,
                      const int64_t *indPtr, const int32_t *colIdx, gary_batch **out);
int gary_batch_destroy(gary_batch *b);
/* Bin thresholds / grid of the pivot path; drops the captured graph. */
int gary_batch_set_cfg(gary_batch *b, const gary_launch_cfg *cfg);
int gary_batch_device_graph(gary_batch *b, gary_device_graph *out);
/* Counts every graph, overwriting the batch's per-graph counters. use_graph != 0 replays the
 * hipGraph captured at the first such call (one hipGraphLaunch); 0 enqueues the kernels. */
int gary_batch_tc_launch(gary_batch *b, int use_graph, void *stream);
/* Copies the per-graph counts of the last launch (waits for `stream`). */
int gary_batch_counts(gary_batch *b, uint64_t *h_counts, void *stream);
/* gary_batch_tc_launch(b, 1, NULL) + gary_batch_counts(b, h_counts, NULL). */
int gary_batch_tc_count(gary_batch *b, uint64_t *h_counts);

/* ---- 4-vertex homomorphism counts (graph_query::rectCountHom, declared but never defined at
 *      include/kernels/GraphQueryHom.cuh:48; K4 and diamond have no reference counterpart) --- */
enum { GARY_HOM_K4 = 1, GARY_HOM_DIAMOND = 2, GARY_HOM_C4 = 4, GARY_HOM_ALL = 7 };
typedef struct gary_hom4_counts {
  uint64_t tri;         /* #triangles (with GARY_HOM_K4 or GARY_HOM_DIAMOND) */
  uint64_t k4;          /* #4-cliques */
  uint64_t hom_k4;      /* 24 * k4 */
  uint64_t hom_diamond; /* sum over arcs (a,b) of |N(a) ∩ N(b)|^2 */
  uint64_t hom_c4;      /* tr(A^4) = 8 * c4 + 4 * sum_v C(d(v),2) + 2m */
  uint64_t c4;          /* #4-cycles */
} gary_hom4_counts;
/* Counts on a symmetric, loop-free CSR whose ids are ranked by (degree, id)
 * (GARY_CSR_SYMMETRIC_RANKED); any other input fails with GARY_ERR_INVALID. Synchronous on
 * `stream`; values that exceed 64 bits fail with GARY_ERR_OVERFLOW. */
int gary_hom4_count(const gary_device_graph *g, int patterns, gary_hom4_counts *out,
                    void *stream);

/* 4-cycle vertex routes. The 4-cycles are counted at their highest-ranked vertex u from the
 * wedges u-v-w (v, w < u). Each vertex with nl >= 2 lower neighbours falls in bin
 * 16 * min(15, floor(log2 nl)) + min(15, floor(log2 S) / 2), S = the sum of d(v) over its lower
 * neighbours v (an upper bound on its wedges; above 64 lower neighbours, nl / 64 times the sum over
 * 64 evenly spaced ones), and a table of GARY_C4_BINS route bytes sends each bin to one counting
 * path: */
enum {
  GARY_C4_LIGHT = 0,  /* LDS hash table of one wave, then of one workgroup, then dense counters,
                         as the vertex's wedges overflow each (nl < 256 only; above, the windows) */
  GARY_C4_WINDOW = 1, /* LDS counters over windows of 65,536 endpoint ids */
  GARY_C4_DENSE = 2,  /* one workgroup with a dense counter array in device memory */
  GARY_C4_SKIP = 3    /* not counted (for timing the others) */
};
#define GARY_C4_BINS 256
/* How a call routes its vertices: AUTO is LEARNED when a 4-cycle model is set on the device
 * (gary_hom4_c4_set_model / gary_hom4_c4_load_model), else FIXED; FIXED is the degree split
 * (nl >= 256 to the windows, the rest light); LEARNED takes per bin the route with the least
 * MLP-predicted ns per vertex (fails with GARY_ERR_STATE without a model); TABLE uses the
 * caller's table. */
enum {
  GARY_C4_ROUTES_AUTO = 0,
  GARY_C4_ROUTES_FIXED = 1,
  GARY_C4_ROUTES_LEARNED = 2,
  GARY_C4_ROUTES_TABLE = 3
};
/* gary_hom4_count with the 4-cycle routes chosen by `c4_routes` (h_table: GARY_C4_BINS route
 * bytes for GARY_C4_ROUTES_TABLE, else ignored). No reference counterpart; counts do not depend
 * on the routes, except that GARY_C4_SKIP bins are left out of c4 and hom_c4. */
int gary_hom4_count_routed(const gary_device_graph *g, int patterns, int c4_routes,
                           const uint8_t *h_table, gary_hom4_counts *out, void *stream);
/* Per-bin MLP inputs of the 4-cycle routing of g (h_x: GARY_C4_BINS x GARY_MLP_IN floats) and the
 * vertices in each bin (h_vertices: GARY_C4_BINS floats). Synchronous. */
int gary_hom4_c4_features(const gary_device_graph *g, float *h_x, float *h_vertices);
/* The table a routing mode (AUTO, FIXED or LEARNED) gives g. Synchronous. */
int gary_hom4_c4_table(const gary_device_graph *g, int c4_routes, uint8_t *h_table);
/* Sets the 4-cycle model of the current device: GARY_MLP_PARAMS floats (the layout of the
 * selector's MLP; outputs = log2 ns per vertex of routes LIGHT, WINDOW, DENSE) and the predicted
 * relative gain needed to leave the fixed route (0 <= margin < 1). h_params = NULL removes it. */
int gary_hom4_c4_set_model(const float *h_params, float margin);
/* Reads "params" (GARY_MLP_PARAMS numbers) and optional "margin" from the JSON file that
 * gary_amd/c4_selector.py writes, and sets them as above. */
int gary_hom4_c4_load_model(const char *path);
/* Whether the current device has a 4-cycle model; h_params (GARY_MLP_PARAMS floats) and margin
 * receive it when non-null and set. */
int gary_hom4_c4_model(int *trained, float *h_params, float *margin);

/* Number of k-cliques (1 <= k <= 10) of the simple graph given as a GARY_CSR_SYMMETRIC_RANKED CSR:
 * for each vertex u, the (k-1)-cliques of the DAG induced on its higher-ranked neighbours, counted
 * on the device. k = 3 is the triangle count, k = 4 the K4 count of gary_hom4_count. No reference
 * counterpart (the reference counts triangles only). Fails with GARY_ERR_UNSUPPORTED when a
 * vertex has more than 2,048 higher-ranked neighbours. */
int gary_kclique_count(const gary_device_graph *g, int k, uint64_t *count, void *stream);

/* ---- block-scope queue (include/ops/DeviceQueue.cuh:33-182, CTAQueue) ------------------------ */

/* Runs the device queue of gary_amd/csrc/block_queue.h (the reference's CTAQueue: clear / fill /
 * size / tryEnqX / tryDeqX / tryEnq / tryDeq on a 128-slot ring in LDS) in one workgroup of 16
 * waves, lane 0 of each wave calling it, as the reference's test does
 * (test/CTACommTest.cu:91-119). Asynchronous on `stream`; d_out is zeroed first.
 *   mode 0, exclusive: the queue is filled with d_items[0..n), n <= 128, and every wave dequeues
 *     with tryDeqX until it is empty: d_out[w*128 + k] = the k-th item wave w took,
 *     d_status[w] = how many it took (d_out: 16*128 entries).
 *   mode 1, concurrent: waves 0-7 enqueue their share of d_items[0..n) (values < n) through the
 *     ring while waves 8-15 dequeue until n items are out; d_out[x] counts the times item x came
 *     out (n entries), d_status[w] = the items wave w moved, or -1 if a call of that wave was
 *     stuck or an item was out of range.
 * Every retry loop in the kernel is bounded, so it always ends. */
int gary_block_queue_exercise(int mode, const uint32_t *d_items, uint32_t n, uint32_t *d_out,
                              int32_t *d_status, void *stream);

/* Triangle listing on a degree-oriented CSR (GARY_CSR_ORIENT_DEGREE or ORIENT_ID; needs no COO):
 * every triangle once, as the int32 triple (u, v, w) with v in N+(u) and w in N+(u) ∩ N+(v),
 * written to the device array out[3*cap]. The order of the triples depends on scheduling.
 * *total receives the number of triangles; when it exceeds cap only cap triples are written and
 * the call returns GARY_ERR_OVERFLOW. The intersection-with-output of
 * include/ops/Intersection.cuh:18-50,76-112 applied to the triangle query of
 * include/kernels/GraphQueryHom.cuh:15-43 (which the reference only counts). Synchronous.
 * Scratch: a chunk of 256 to 2,048 triples per wave (about cap / 16384), at most 192 MiB. */
int gary_tri_list(const gary_device_graph *g, int32_t *out, uint64_t cap, uint64_t *total,
                  void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GARY_H_ */
