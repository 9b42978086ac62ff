// tc_kernels.hip — triangle counting over an acyclically oriented CSR/COO on gfx950.
//
// Count = Σ over arcs (u, v) of |N+(u) ∩ N+(v)| (graph_query::triangleCountHom,
// include/kernels/GraphQueryHom.cuh:15-43). Two kernel families:
//
// 1. Edge-centric (the reference's formulation): a group of G lanes (G = 8..64) per COO edge,
//    INTERLEAVED or CHUNKED over the edge range (DeviceIterator.cuh:117-131), intersecting by
//    plain binary search (Intersection.cuh:56-71), two-phase LDS-cached binary search
//    (:118-135, Search.cuh:93-138) or a merge-path merge (the reference's intersectMerge is a
//    stub, :172-180).
//
// 2. Pivot-centric ("bitmap"): a wave (or a workgroup for hubs) owns a pivot vertex u, builds an
//    LDS membership structure of N+(u) — a bitmap over [min N+(u), max N+(u)] when that window
//    fits, else an open-addressing hash — and streams every N+(v), v in N+(u), through it. The
//    (v, w) pairs are flattened across the 64 lanes by a wave prefix scan over deg(v), so lanes
//    stay busy however short the lists are and every N+(v) is read with consecutive addresses.
//    Each N+(u) is read once per pivot rather than once per edge.
//
// Counts are per-lane 32-bit partials, reduced per wave to 64 bits and added with one atomic per
// wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "device_search.h"
#include "device_util.h"
#include "tc_internal.h"
#include "workspace.h"

namespace gary {

// ======================================================================================
// Group-cooperative intersections (count only)
// ======================================================================================

// Merge-path split: number of A elements among the first k of merge(A, B), A taken first on ties.
__device__ __forceinline__ int32_t merge_path_split(const int32_t *__restrict__ A, int32_t na,
                                                    const int32_t *__restrict__ B, int32_t nb,
                                                    int32_t k) {
  int32_t lo = k > nb ? k - nb : 0, hi = k < na ? k : na;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (A[mid] <= B[k - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Count matches in merged positions [k0, k1): a match is counted when B[j] is consumed right
// after an equal A element, so every common value is counted exactly once across lanes.
__device__ __forceinline__ uint32_t merge_count_segment(const int32_t *__restrict__ A, int32_t na,
                                                        const int32_t *__restrict__ B, int32_t nb,
                                                        int32_t k0, int32_t k1) {
  if (k0 >= k1) return 0;
  int32_t i = merge_path_split(A, na, B, nb, k0), j = k0 - i;
  uint32_t c = 0;
  int32_t a = i < na ? A[i] : INT32_MAX;
  int32_t b = j < nb ? B[j] : INT32_MAX;
  int32_t prev_a = i > 0 ? A[i - 1] : INT32_MIN;
  for (int32_t k = k0; k < k1; ++k) {
    if (i < na && (j >= nb || a <= b)) {
      prev_a = a;
      ++i;
      a = i < na ? A[i] : INT32_MAX;
    } else {
      c += (i > 0 && prev_a == b);
      ++j;
      b = j < nb ? B[j] : INT32_MAX;
    }
  }
  return c;
}

template <int G, int METHOD>
__device__ __forceinline__ uint32_t isect_group(const int32_t *__restrict__ A, int32_t na,
                                                const int32_t *__restrict__ B, int32_t nb,
                                                int32_t *cache, int32_t n_cache, int glane) {
  if (na == 0 || nb == 0) return 0;  // Intersection.cuh:190,197
  uint32_t c = 0;
  if constexpr (METHOD == kIsectMerge) {
    const int32_t L = na + nb, per = (L + G - 1) / G;
    const int32_t k0 = min(glane * per, L), k1 = min(k0 + per, L);
    c = merge_count_segment(A, na, B, nb, k0, k1);
  } else {
    if (na > nb) {  // swap so that |A| <= |B|
      const int32_t *t = A; A = B; B = t;
      int32_t tn = na; na = nb; nb = tn;
    }
    if constexpr (METHOD == kIsectBsearch) {
      for (int32_t k = glane; k < na; k += G) c += bsearch(B, nb, A[k]);
    } else {
      build_cache_group(B, nb, cache, n_cache, glane, G);
      wave_sync();
      for (int32_t k = glane; k < na; k += G) c += bsearch_2phase(B, nb, cache, n_cache, A[k]);
      wave_sync();  // the cache is rebuilt for the next edge
    }
  }
  return c;
}

// ======================================================================================
// Edge-centric kernel: one group of G lanes per COO edge.
// ======================================================================================

template <int G, int METHOD, int PART>
__global__ void __launch_bounds__(256) tc_edge_kernel(gary_device_graph g, int32_t n_cache,
                                                      unsigned long long *count) {
  extern __shared__ int32_t smem[];
  const int glane = lane_id() & (G - 1);
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G;
  const int64_t gid = (int64_t)blockIdx.x * gpb + gib;
  const int64_t ngroups = (int64_t)gridDim.x * gpb;
  int32_t *cache = smem + (size_t)gib * n_cache;
  const int64_t E = g.coo.nnz;
  const uint32_t m = (uint32_t)g.csr.m;
  int64_t e, e_end, step;
  if constexpr (PART == GARY_PARTITION_CHUNKED) {
    e = gid * E / ngroups;
    e_end = (gid + 1) * E / ngroups;
    step = 1;
  } else {
    e = gid;
    e_end = E;
    step = ngroups;
  }
  uint32_t c = 0;
  for (; e < e_end; e += step) {
    const int32_t src = g.coo.rowIdx[e], dst = g.coo.colIdx[e];
    if ((uint32_t)src >= m || (uint32_t)dst >= m) continue;
    const int32_t *A = g.csr.colIdx + g.csr.indPtr[src];
    const int32_t *B = g.csr.colIdx + g.csr.indPtr[dst];
    c += isect_group<G, METHOD>(A, g.csr.deg[src], B, g.csr.deg[dst], cache, n_cache, glane);
  }
  wave_atomic_add(count, c);
}

// ======================================================================================
// Pivot-centric kernels
// ======================================================================================

__device__ __forceinline__ uint32_t hash_slot(int32_t w, int32_t log_t) {
  return ((uint32_t)w * 0x9E3779B1u) >> (32 - log_t);
}

// LDS membership structure of one pivot list L[0..d): bitmap over [lo, hi] if the window fits
// `bitmap_words` words and is not much larger than a hash table would be, else a hash table of
// 2^log_t slots (log_t chosen with load factor <= 1/2).
struct PivotSet {
  int32_t lo, hi;
  int32_t log_t;   // hash slots = 1 << log_t; 0 means bitmap mode
  int32_t words;   // bitmap words (bitmap mode)
};

__device__ __forceinline__ PivotSet pivot_plan(int32_t lo, int32_t hi, int32_t d,
                                               int32_t table_words) {
  PivotSet s;
  s.lo = lo;
  s.hi = hi;
  int32_t log_t = 6;
  while ((1 << log_t) < 2 * d) ++log_t;
  const int64_t span = (int64_t)hi - lo + 1;
  const int64_t words = (span + 31) >> 5;
  if (words <= table_words && words <= 4 * (1 << log_t)) {
    s.log_t = 0;
    s.words = (int32_t)words;
  } else {
    s.log_t = log_t;
    s.words = 0;
  }
  return s;
}

// Probe w against the pivot set (one LDS read for the bitmap; linear probing for the hash).
__device__ __forceinline__ uint32_t pivot_probe(const int32_t *table, const PivotSet &s,
                                                int32_t w) {
  if (w < s.lo || w > s.hi) return 0;
  if (s.log_t == 0) {
    const uint32_t off = (uint32_t)(w - s.lo);
    return ((uint32_t)table[off >> 5] >> (off & 31)) & 1u;
  }
  const uint32_t mask = (1u << s.log_t) - 1;
  uint32_t h = hash_slot(w, s.log_t);
  while (true) {
    const int32_t t = table[h];
    if (t == w) return 1;
    if (t == -1) return 0;
    h = (h + 1) & mask;
  }
}

// Wave-per-pivot kernel. Each wave owns `table_words` words of LDS for its pivot set plus the
// 64-entry row tables of the flattened (v, w) loop.
template <int WAVES, int TABLE_WORDS>
__global__ void __launch_bounds__(WAVES * 64)
    tc_pivot_wave_kernel(gary_device_graph g, const int32_t *__restrict__ pivots,
                         const int32_t *__restrict__ d_npivots, unsigned long long *count) {
  const int32_t npivots = *d_npivots;
  __shared__ int32_t s_table[WAVES][TABLE_WORDS];
  __shared__ int32_t s_end[WAVES][64];
  __shared__ int64_t s_ptr[WAVES][64];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  int32_t *table = s_table[wib];
  int32_t *rend = s_end[wib];
  int64_t *rptr = s_ptr[wib];
  const uint32_t m = (uint32_t)g.csr.m;
  const int64_t nw = (int64_t)gridDim.x * WAVES;
  uint32_t c = 0;
  for (int64_t p = (int64_t)blockIdx.x * WAVES + wib; p < npivots; p += nw) {
    const int32_t u = pivots ? pivots[p] : (int32_t)p;
    if ((uint32_t)u >= m) continue;
    const int32_t d = g.csr.deg[u];
    if (d < 2) continue;
    const int32_t *L = g.csr.colIdx + g.csr.indPtr[u];
    const PivotSet s = pivot_plan(L[0], L[d - 1], d, TABLE_WORDS);
    if (s.log_t == 0 || (1 << s.log_t) <= TABLE_WORDS) {
      // ---- build the LDS set
      const int32_t fill_words = s.log_t == 0 ? s.words : (1 << s.log_t);
      const int32_t fill_val = s.log_t == 0 ? 0 : -1;
      for (int32_t i = lane; i < fill_words; i += kWave) table[i] = fill_val;
      wave_sync();
      for (int32_t i = lane; i < d; i += kWave) {
        const int32_t w = L[i];
        if (s.log_t == 0) {
          const uint32_t off = (uint32_t)(w - s.lo);  // < 32*words for a sorted list
          if (off < (uint32_t)s.words * 32u) atomicOr((uint32_t *)&table[off >> 5], 1u << (off & 31));
        } else {
          const uint32_t mask = (1u << s.log_t) - 1;
          uint32_t h = hash_slot(w, s.log_t);
          while (atomicCAS(&table[h], -1, w) != -1) h = (h + 1) & mask;
        }
      }
      wave_sync();
      // ---- stream N+(v) for v in N+(u), 64 rows at a time, flattened over lanes
      for (int32_t c0 = 0; c0 < d; c0 += kWave) {
        const int32_t i = c0 + lane;
        int32_t dv = 0;
        int64_t pv = 0;
        if (i < d) {
          const int32_t v = L[i];
          if ((uint32_t)v < m) {
            dv = g.csr.deg[v];
            pv = g.csr.indPtr[v];
          }
        }
        const uint32_t incl = wave_inclusive_scan_u32((uint32_t)dv);
        const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
        rend[lane] = (int32_t)incl;
        rptr[lane] = pv;
        wave_sync();
        for (uint32_t f = lane; f < total; f += kWave) {
          // row j = first index with rend[j] > f
          int32_t j = 0;
#pragma unroll
          for (int32_t half = 32; half > 0; half >>= 1)
            if ((uint32_t)rend[j + half - 1] <= f) j += half;
          j += ((uint32_t)rend[j] <= f);
          const uint32_t start = j ? (uint32_t)rend[j - 1] : 0u;
          const int32_t w = g.csr.colIdx[rptr[j] + (f - start)];
          c += pivot_probe(table, s, w);
        }
        wave_sync();
      }
    } else {
      // The pivot set does not fit this wave's LDS: intersect edge by edge (plain search of the
      // shorter list in the longer one). The dispatcher sends such pivots to the hub kernel, so
      // this path only guards against a mis-binned call.
      for (int32_t k = 0; k < d; ++k) {
        const int32_t v = L[k];
        if ((uint32_t)v >= m) continue;
        c += isect_group<64, kIsectBsearch>(L, d, g.csr.colIdx + g.csr.indPtr[v], g.csr.deg[v],
                                            nullptr, 0, lane);
      }
    }
  }
  wave_atomic_add(count, c);
}

// Workgroup-per-hub-pivot kernel: the pivot set lives in dynamic LDS of `table_words` words;
// rows of the flattened loop are processed BLOCK at a time.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    tc_pivot_block_kernel(gary_device_graph g, const int32_t *__restrict__ pivots,
                          const int32_t *__restrict__ d_npivots, int32_t table_words,
                          unsigned long long *count) {
  const int32_t npivots = *d_npivots;
  extern __shared__ int32_t dyn[];
  __shared__ int32_t s_end[BLOCK];
  __shared__ int64_t s_ptr[BLOCK];
  __shared__ uint32_t s_wsum[BLOCK / 64];
  __shared__ uint64_t s_red[BLOCK / 64];
  int32_t *table = dyn;
  const int lane = lane_id();
  const int wib = threadIdx.x / kWave;
  const uint32_t m = (uint32_t)g.csr.m;
  uint32_t c = 0;
  for (int64_t p = blockIdx.x; p < npivots; p += gridDim.x) {
    const int32_t u = pivots ? pivots[p] : (int32_t)p;
    if ((uint32_t)u >= m) continue;
    const int32_t d = g.csr.deg[u];
    if (d < 2) continue;
    const int32_t *L = g.csr.colIdx + g.csr.indPtr[u];
    const PivotSet s = pivot_plan(L[0], L[d - 1], d, table_words);
    if (s.log_t != 0 && (1 << s.log_t) > table_words) {
      // set too large for LDS: edge-by-edge binary search, one wave per arc
      for (int32_t k = wib; k < d; k += BLOCK / 64) {
        const int32_t v = L[k];
        if ((uint32_t)v >= m) continue;
        c += isect_group<64, kIsectBsearch>(L, d, g.csr.colIdx + g.csr.indPtr[v], g.csr.deg[v],
                                            nullptr, 0, lane);
      }
      continue;
    }
    const int32_t fill_words = s.log_t == 0 ? s.words : (1 << s.log_t);
    const int32_t fill_val = s.log_t == 0 ? 0 : -1;
    __syncthreads();  // previous pivot's readers are done with the table
    for (int32_t i = threadIdx.x; i < fill_words; i += BLOCK) table[i] = fill_val;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < d; i += BLOCK) {
      const int32_t w = L[i];
      if (s.log_t == 0) {
        const uint32_t off = (uint32_t)(w - s.lo);  // < 32*words for a sorted list
        if (off < (uint32_t)s.words * 32u) atomicOr((uint32_t *)&table[off >> 5], 1u << (off & 31));
      } else {
        const uint32_t mask = (1u << s.log_t) - 1;
        uint32_t h = hash_slot(w, s.log_t);
        while (atomicCAS(&table[h], -1, w) != -1) h = (h + 1) & mask;
      }
    }
    __syncthreads();
    for (int32_t c0 = 0; c0 < d; c0 += BLOCK) {
      const int32_t i = c0 + threadIdx.x;
      int32_t dv = 0;
      int64_t pv = 0;
      if (i < d) {
        const int32_t v = L[i];
        if ((uint32_t)v < m) {
          dv = g.csr.deg[v];
          pv = g.csr.indPtr[v];
        }
      }
      // block-wide inclusive scan of dv
      uint32_t incl = wave_inclusive_scan_u32((uint32_t)dv);
      if (lane == kWave - 1) s_wsum[wib] = incl;
      __syncthreads();
      uint32_t off = 0;
      for (int k = 0; k < wib; ++k) off += s_wsum[k];
      uint32_t total = 0;
      for (int k = 0; k < BLOCK / 64; ++k) total += s_wsum[k];
      incl += off;
      s_end[threadIdx.x] = (int32_t)incl;
      s_ptr[threadIdx.x] = pv;
      __syncthreads();
      for (uint32_t f = threadIdx.x; f < total; f += BLOCK) {
        int32_t j = 0;
        for (int32_t half = BLOCK / 2; half > 0; half >>= 1)
          if ((uint32_t)s_end[j + half - 1] <= f) j += half;
        j += ((uint32_t)s_end[j] <= f);
        const uint32_t start = j ? (uint32_t)s_end[j - 1] : 0u;
        const int32_t w = g.csr.colIdx[s_ptr[j] + (f - start)];
        c += pivot_probe(table, s, w);
      }
      __syncthreads();  // s_end / s_ptr / s_wsum are rewritten by the next chunk
    }
  }
  block_atomic_add<BLOCK>(count, c, s_red);
}

// ======================================================================================
// Degree binning of pivots (device work lists)
// ======================================================================================

// bins: 0 = wave kernel with the small table, 1 = wave kernel with the large table,
// 2 = workgroup (hub) kernel. Vertices with out-degree < 2 are dropped (no triangle has them as
// its lowest vertex). Wave-aggregated atomics: one atomic per wave and bin.
__global__ void bin_pivots_kernel(const int32_t *__restrict__ deg, int32_t m, int32_t t_small,
                                  int32_t t_large, int32_t *__restrict__ lists, int64_t stride,
                                  int32_t *__restrict__ counters) {
  const int lane = lane_id();
  for (int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ll; base < m;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = base + lane;
    int bin = -1;
    if (v < m) {
      const int32_t d = deg[v];
      if (d >= 2) bin = d <= t_small ? 0 : (d <= t_large ? 1 : 2);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const uint64_t mask = __ballot(bin == b);
      if (!mask) continue;
      int32_t pos0 = 0;
      const int leader = __ffsll((unsigned long long)mask) - 1;
      if (lane == leader) pos0 = atomicAdd(&counters[b], __popcll(mask));
      pos0 = __shfl(pos0, leader, kWave);
      if (bin == b) lists[b * stride + pos0 + lanes_below(mask)] = (int32_t)v;
    }
  }
}

}  // namespace gary

// ======================================================================================
// Host dispatch (C ABI)
// ======================================================================================

namespace gary {
namespace {

struct TcCfg {
  int block = 256, grid = 0, group = 64, cache_depth = 64, partition = GARY_PARTITION_INTERLEAVED;
  int bin_small = 128, bin_hub = 1024;
};

constexpr int kSmallWords = 256;    // wave kernel, small pivot sets (<= 128 keys hashed)
constexpr int kLargeWords = 2048;   // wave kernel, large pivot sets (<= 1024 keys hashed)
constexpr int kHubWords = 8192;     // workgroup kernel (32 KiB of dynamic LDS)
constexpr int kHubBlock = 512;
constexpr int kPivotWaves = 4;

int resolve_cfg(const gary_launch_cfg *in, TcCfg &c) {
  if (in) {
    if (in->block) c.block = in->block;
    if (in->grid) c.grid = in->grid;
    if (in->group) c.group = in->group;
    if (in->cache_depth) c.cache_depth = in->cache_depth;
    c.partition = in->partition;
    if (in->bin_small) c.bin_small = in->bin_small;
    if (in->bin_hub) c.bin_hub = in->bin_hub;
  }
  GARY_CHECK_ARG(c.block >= 64 && c.block <= 256 && c.block % 64 == 0,
                 "launch cfg: block=%d must be a multiple of 64 in [64, 256]", c.block);
  GARY_CHECK_ARG(c.group == 8 || c.group == 16 || c.group == 32 || c.group == 64,
                 "launch cfg: group=%d must be 8, 16, 32 or 64", c.group);
  GARY_CHECK_ARG(c.cache_depth >= 1 && c.cache_depth <= 4096 &&
                     (c.cache_depth & (c.cache_depth - 1)) == 0,
                 "launch cfg: cache_depth=%d must be a power of two <= 4096", c.cache_depth);
  GARY_CHECK_ARG(c.partition == GARY_PARTITION_INTERLEAVED || c.partition == GARY_PARTITION_CHUNKED,
                 "launch cfg: bad partition %d", c.partition);
  GARY_CHECK_ARG(c.grid >= 0, "launch cfg: grid=%d", c.grid);
  GARY_CHECK_ARG(c.bin_small >= 2 && c.bin_small <= kSmallWords / 2 && c.bin_hub >= c.bin_small &&
                     c.bin_hub <= kLargeWords / 2,
                 "launch cfg: bins small=%d hub=%d out of range", c.bin_small, c.bin_hub);
  const size_t lds = (size_t)(c.block / c.group) * c.cache_depth * sizeof(int32_t);
  GARY_CHECK_ARG(lds <= 65536, "launch cfg: index cache needs %zu bytes of LDS", lds);
  return GARY_OK;
}

int check_graph(const gary_device_graph *g, bool need_coo) {
  GARY_CHECK_ARG(g, "null DeviceGraph");
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "DeviceGraph: m=%d nnz=%lld", g->csr.m,
                 (long long)g->csr.nnz);
  if (g->csr.m > 0) GARY_CHECK_ARG(g->csr.indPtr && g->csr.deg, "DeviceGraph: null CSR arrays");
  if (g->csr.nnz > 0) GARY_CHECK_ARG(g->csr.colIdx, "DeviceGraph: null colIdx");
  if (need_coo && g->coo.nnz > 0)
    GARY_CHECK_ARG(g->coo.rowIdx && g->coo.colIdx,
                   "DeviceGraph: the COO rowIdx is required (call populate first)");
  return GARY_OK;
}

template <int G, int METHOD>
int launch_edge_g(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                  hipStream_t st) {
  const int grid = c.grid ? c.grid : 4096;
  const size_t lds = METHOD == kIsectCached
                         ? (size_t)(c.block / G) * c.cache_depth * sizeof(int32_t) : 0;
  if (c.partition == GARY_PARTITION_CHUNKED)
    tc_edge_kernel<G, METHOD, GARY_PARTITION_CHUNKED><<<grid, c.block, lds, st>>>(g, c.cache_depth, cnt);
  else
    tc_edge_kernel<G, METHOD, GARY_PARTITION_INTERLEAVED><<<grid, c.block, lds, st>>>(g, c.cache_depth, cnt);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

template <int METHOD>
int launch_edge(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                hipStream_t st) {
  switch (c.group) {
    case 8: return launch_edge_g<8, METHOD>(g, c, cnt, st);
    case 16: return launch_edge_g<16, METHOD>(g, c, cnt, st);
    case 32: return launch_edge_g<32, METHOD>(g, c, cnt, st);
    default: return launch_edge_g<64, METHOD>(g, c, cnt, st);
  }
}

// Pivot bins: workspace = 3 lists of m int32 + 3 counters.
int launch_pivot(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                 hipStream_t st) {
  const int32_t m = g.csr.m;
  if (m == 0 || g.csr.nnz == 0) return GARY_OK;
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t stride = ((size_t)m + 63) & ~(size_t)63;
  GARY_TRY(workspace_get(dev, st, sizeof(int32_t) * (3 * stride + 64), &ws));
  int32_t *lists = (int32_t *)ws;
  int32_t *counters = lists + 3 * stride;
  GARY_HIP(hipMemsetAsync(counters, 0, sizeof(int32_t) * 4, st));
  const int bgrid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 2048);
  bin_pivots_kernel<<<bgrid, 256, 0, st>>>(g.csr.deg, m, c.bin_small, c.bin_hub, lists,
                                           (int64_t)stride, counters);
  GARY_HIP(hipGetLastError());
  const int grid = c.grid ? c.grid : 2048;
  // Pivot counts stay on the device: kernels read them through the counters (no host sync).
  tc_pivot_wave_kernel<kPivotWaves, kSmallWords><<<grid, kPivotWaves * 64, 0, st>>>(
      g, lists, counters + 0, cnt);
  GARY_HIP(hipGetLastError());
  tc_pivot_wave_kernel<kPivotWaves, kLargeWords><<<grid, kPivotWaves * 64, 0, st>>>(
      g, lists + stride, counters + 1, cnt);
  GARY_HIP(hipGetLastError());
  tc_pivot_block_kernel<kHubBlock><<<512, kHubBlock, kHubWords * sizeof(int32_t), st>>>(
      g, lists + 2 * stride, counters + 2, kHubWords, cnt);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

}  // namespace
}  // namespace gary

extern "C" {

int64_t gary_index_cache_smem(int32_t block, int64_t cache_size) {
  if (block <= 0 || cache_size < 0) return 0;
  return (int64_t)(block / gary::kWave) * cache_size;
}

int gary_tc_count(const gary_device_graph *g, int strategy, const gary_launch_cfg *cfg,
                  uint64_t *d_count, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(d_count, "gary_tc_count: null counter");
  GARY_CHECK_ARG(strategy >= 0 && strategy < GARY_TC_NUM_STRATEGIES,
                 "gary_tc_count: unknown strategy %d", strategy);
  const bool edge = strategy == GARY_TC_MERGE || strategy == GARY_TC_CACHED_BSEARCH ||
                    strategy == GARY_TC_BSEARCH;
  GARY_TRY(check_graph(g, edge));
  TcCfg c;
  if (strategy == GARY_TC_MERGE) c.group = 16;
  GARY_TRY(resolve_cfg(cfg, c));
  auto *cnt = (unsigned long long *)d_count;
  hipStream_t st = (hipStream_t)stream;
  if (g->csr.nnz == 0 || g->csr.m == 0) return GARY_OK;
  switch (strategy) {
    case GARY_TC_BSEARCH: return launch_edge<kIsectBsearch>(*g, c, cnt, st);
    case GARY_TC_CACHED_BSEARCH: return launch_edge<kIsectCached>(*g, c, cnt, st);
    case GARY_TC_MERGE: return launch_edge<kIsectMerge>(*g, c, cnt, st);
    case GARY_TC_BITMAP:
    case GARY_TC_HEURISTIC:
    case GARY_TC_AUTO_LEARNED:
    default: return launch_pivot(*g, c, cnt, st);
  }
  GARY_API_END
}

int gary_tc_count_sync(const gary_device_graph *g, int strategy, const gary_launch_cfg *cfg,
                       uint64_t *h_count) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(h_count, "gary_tc_count_sync: null output");
  int dev = 0;
  GARY_TRY(gary::current_device(&dev));
  void *ws = nullptr;
  // the counter lives in its own workspace slot (key: stream handle 1, never a real stream)
  GARY_TRY(gary::workspace_get(dev, (void *)1, sizeof(uint64_t), &ws));
  GARY_HIP(hipMemset(ws, 0, sizeof(uint64_t)));
  GARY_TRY(gary_tc_count(g, strategy, cfg, (uint64_t *)ws, nullptr));
  GARY_HIP(hipMemcpy(h_count, ws, sizeof(uint64_t), hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_tc_count_legacy(const gary_device_graph *g, uint32_t *d_count, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(d_count, "gary_tc_count_legacy: null counter");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0;
  GARY_TRY(gary::current_device(&dev));
  void *ws = nullptr;
  GARY_TRY(gary::workspace_get(dev, (void *)2, sizeof(uint64_t), &ws));
  GARY_HIP(hipMemsetAsync(ws, 0, sizeof(uint64_t), st));
  gary_launch_cfg cfg{};
  cfg.group = 64;
  cfg.cache_depth = 64;
  GARY_TRY(gary_tc_count(g, GARY_TC_CACHED_BSEARCH, &cfg, (uint64_t *)ws, stream));
  uint64_t total = 0;
  uint32_t prev = 0;
  GARY_HIP(hipMemcpyAsync(&total, ws, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(&prev, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  GARY_HIP(hipStreamSynchronize(st));
  const uint64_t sum = total + prev;
  if (sum > UINT32_MAX)
    return gary::set_error(GARY_ERR_OVERFLOW,
                           "triangleCountLegacy: count %llu does not fit 32 bits",
                           (unsigned long long)sum);
  const uint32_t out = (uint32_t)sum;
  GARY_HIP(hipMemcpyAsync(d_count, &out, sizeof(uint32_t), hipMemcpyHostToDevice, st));
  GARY_HIP(hipStreamSynchronize(st));
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
