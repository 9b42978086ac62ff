// tc_internal.h — constants shared by the counting kernels and their dispatchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "gary.h"

namespace gary {

// Intersection methods of the edge-centric kernels.
constexpr int kIsectBsearch = 0;  // plain binary search of the shorter list in the longer
constexpr int kIsectCached = 1;   // two-phase search with a per-group LDS index cache
constexpr int kIsectMerge = 2;    // merge-path merge of both lists

// Per-arc routes of the selector-driven path (one byte per arc).
constexpr uint8_t kRouteBitmap = 0;  // streamed through the pivot's LDS set
constexpr uint8_t kRouteCached = 1;  // edge-centric two-phase cached binary search
constexpr uint8_t kRouteMerge = 2;   // edge-centric merge path
constexpr uint8_t kRouteSkip = 3;    // no work (empty intersection, or excluded for timing)
constexpr int kRouteBins = 256;      // (min(15, log2 d+(u)), min(15, log2 d+(v)))

// Device-resident strategy table: route per bin and predicted cost per arc (ns) per bin.
// Pivot-path launch knobs: pivots with out-degree <= bin_small share a workgroup (one per wave),
// those up to bin_hub get a workgroup of `waves` waves, larger ones the hub kernel; grid = 0
// takes the default (2048 workgroups).
// int32 words of pivot-path counters that follow the three pivot lists (3 * stride words).
constexpr int kPivotCounterWords = 1536;

// Default out-degree split between per-wave and per-workgroup pivots (measured best on
// R-MAT 20 with the heavy-first pivot order).
constexpr int kDefaultBinSmall = 128;

struct PivotCfg {
  int bin_small = kDefaultBinSmall, bin_hub = 2048, grid = 0, waves = 4;
  int direction = GARY_PIVOT_AUTO;  // GARY_PIVOT_*
};

// Reverse rows of the pivot path. An arc (u, v) is streamed from the endpoint whose out-list is
// shorter: N+(v) through the set of N+(u) at pivot u (forward), or, when d+(u) + kRevSlack <
// d+(v), N+(u) through the set of N+(v) at pivot v (reverse). The reversed arcs are grouped by
// v per call (sources of out-degree kRevMaxSourceDeg or more are never reversed) into rrows, one
// (row start, out-degree) of the source u per arc, so a pivot reads
// its reverse rows without a dependent load of indPtr[u] and d+(u); rrange[v] = (begin, end) of
// v's list there.
// rrange == nullptr streams every arc forward. The grouping is a two-level counting sort
// (tc_kernels.hip) over 2^lg buckets and `tiles` equal ranges of arcs; it needs m <= 2^24 and
// nnz < 2^32.
constexpr int32_t kRevSlack = 128;
constexpr int32_t kRevMaxSourceDeg = 1 << 19;  // sources with a longer out-list stay forward
constexpr int kRevMaxBuckets = 2048;    // LDS counters per tile
constexpr int kRevMaxBucketIds = 8192;  // LDS counters per bucket (ids per bucket)
constexpr int kRevChunk = 4096;         // arcs of a tile staged in LDS at a time

struct RevShape {
  int32_t lg = 9;        // log2 of the bucket count
  int32_t k = 9;         // bits of the hashed id: ceil(log2 m), at least lg
  int32_t ids = 1;       // ids per bucket: 2^(k - lg)
  int32_t tiles = 1;     // arc ranges of the two tile passes
  int64_t tile_arcs = 1; // arcs per tile (a multiple of kRevChunk; the last tile may hold fewer)
};

struct RevBufs {
  uint2 *rrange = nullptr;    // [m] (begin, end) in rrows
  uint2 *rrows = nullptr;     // [nnz] (indPtr[u], d+(u)) of each reversed arc's source
  uint32_t *ghist = nullptr;  // [2^lg * tiles + 1]
  uint32_t *part = nullptr;   // scan partials of ghist
  uint64_t *pairs = nullptr;  // [nnz] (v >> lg) << 51 | d+(u) << 32 | indPtr[u], by bucket
  uint8_t *dsat = nullptr;    // [m] min(d+(v), 255)
  RevShape sh;
};

// Whether reverse rows apply to a graph of m vertices and nnz arcs; whether a launch with
// `direction` uses them on graph g (AUTO: from kRevAutoArcs arcs on); the scratch bytes of
// RevBufs (0 when they do not apply) and their carving from `p`.
constexpr int64_t kRevAutoArcs = 1 << 20;
bool rev_supported(int32_t m, int64_t nnz);
bool rev_wanted(int direction, const gary_device_graph &g);
size_t rev_bytes(int32_t m, int64_t nnz);
RevBufs rev_carve(void *p, int32_t m, int64_t nnz);

struct RouteTable {
  uint8_t route[kRouteBins];
  float cost[kRouteBins];
  uint32_t rowmask;  // bit i: some bin (i, *) routes to the bitmap
  uint32_t mixed;    // nonzero: some bin routes elsewhere than the bitmap
  uint32_t pad[2];
};

// Counts with every arc routed by `table` (device pointer). Cost-based pivot split: pivots whose
// predicted bitmap cost exceeds cost_big (ns) get a workgroup; 0 keeps the degree split.
int launch_routed(const gary_device_graph &g, const RouteTable *table, const PivotCfg &pc,
                  float cost_big, unsigned long long *cnt, hipStream_t st);

// Edge-centric passes of the routed path over compacted arc lists (counts on the device).
int launch_edge_lists(const gary_device_graph &g, const uint32_t *cached, const int32_t *n_cached,
                      const uint32_t *merge, const int32_t *n_merge, unsigned long long *cnt,
                      hipStream_t st);

// HEURISTIC / AUTO_LEARNED strategies (selector.hip).
int launch_selected(const gary_device_graph &g, bool learned, const PivotCfg &pc,
                    unsigned long long *cnt, hipStream_t st);

// Pivot path with caller-provided buffers (see tc_kernels.hip); gid_of/gcount select
// per-graph accumulation into gcount[ngraphs] (zeroed by the launch).
int launch_pivot_core(const gary_device_graph &g, const PivotCfg &pc,
                      int32_t *lists, size_t stride, int32_t *counters,
                      const int32_t *gid_of, unsigned long long *gcount, int32_t ngraphs,
                      unsigned long long *cnt, hipStream_t st, bool timed,
                      const uint8_t *route = nullptr, const float *pcost = nullptr,
                      float cost_big = 0.f, const RouteTable *table = nullptr,
                      const RevBufs *rev = nullptr);

// Exclusive scan of a[0..len) in place on `s` (values and total below 2^32); `part` holds
// exclusive_scan_parts(len) values (graph_build.hip).
uint32_t exclusive_scan_parts(uint32_t len);
int exclusive_scan_u32(uint32_t *a, uint32_t len, uint32_t *part, hipStream_t s);

// Triangle listing on the pivot path: every triangle (u, v, w) of a degree-oriented CSR
// written to out[3 * cap], claims counted in *total (zeroed by the caller).
int launch_tri_list(const gary_device_graph &g, unsigned long long *total, int32_t *out,
                    unsigned long long cap, hipStream_t st);

}  // namespace gary
