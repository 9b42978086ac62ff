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
// 2. Pivot-centric ("bitmap"): a wave (small pivots) or a workgroup (larger ones; an 8-wave
//    kernel with a larger set for the largest) owns a pivot vertex u, builds an LDS membership
//    set of N+(u) — a two-level rank bitmap, or an open-addressing hash when that does not fit —
//    and streams every N+(v), v in N+(u), through it. The rows of 64 v are flattened into
//    16-byte chunks read by consecutive lanes, so lanes stay busy however short the lists are
//    and a wave's loads are contiguous. Each N+(u) is read once per pivot rather than once per
//    arc. Pivots are binned on the device by out-degree, heaviest first.
//
// Counts: each lane sums 32-bit counts per edge (at most min(d(u), d(v)) < 2^31) or per pivot
// call of stream_rows / pivot_fallback (bounded below 2^32 there) into a 64-bit partial, so no
// partial wraps however many edges or pivots a persistent lane takes; partials are reduced per
// wave and added with one atomic per workgroup.
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
      for (int32_t k = glane; k < na; k += G) c += bsearch_2phase_seg(B, nb, cache, n_cache, A[k]);
      wave_sync();  // the cache is rebuilt for the next edge
    }
  }
  return c;
}

// ======================================================================================
// Edge-centric kernel: one group of G lanes per COO edge.
// ======================================================================================

// DYNAMIC partition: edges claimed kDynChunk at a time per group from a device-wide cursor.
constexpr int kDynChunk = 16;

template <int G, int METHOD, int PART>
__global__ void __launch_bounds__(256) tc_edge_kernel(gary_device_graph g, int32_t n_cache,
                                                      unsigned long long *count,
                                                      const uint8_t *__restrict__ route,
                                                      uint8_t want,
                                                      unsigned long long *dyn_cursor = nullptr) {
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
  if constexpr (PART == GARY_PARTITION_DYNAMIC) {
    uint64_t c = 0;
    const int leader = lane_id() & ~(G - 1);
    while (true) {
      unsigned long long base = 0;
      if (glane == 0) base = atomicAdd(dyn_cursor, (unsigned long long)kDynChunk);
      base = __shfl(base, leader, kWave);
      if ((int64_t)base >= E) break;
      const int64_t end = min((int64_t)base + kDynChunk, E);
      for (e = (int64_t)base; e < end; ++e) {
        if (route && route[e] != want) continue;
        const int32_t src = g.coo.rowIdx[e], dst = g.coo.colIdx[e];
        if ((uint32_t)src >= m || (uint32_t)dst >= m) continue;
        const int32_t *A = g.csr.colIdx + g.csr.indPtr[src];
        const int32_t *B = g.csr.colIdx + g.csr.indPtr[dst];
        c += isect_group<G, METHOD>(A, g.csr.deg[src], B, g.csr.deg[dst], cache, n_cache, glane);
      }
    }
    __shared__ unsigned long long s_acc_dyn;
    block_atomic_add(count, c, &s_acc_dyn);
    return;
  }
  if constexpr (PART == GARY_PARTITION_CHUNKED) {
    e = gid * E / ngroups;
    e_end = (gid + 1) * E / ngroups;
    step = 1;
  } else {
    e = gid;
    e_end = E;
    step = ngroups;
  }
  uint64_t c = 0;
  for (; e < e_end; e += step) {
    if (route && route[e] != want) continue;
    const int32_t src = g.coo.rowIdx[e], dst = g.coo.colIdx[e];
    if ((uint32_t)src >= m || (uint32_t)dst >= m) continue;
    const int32_t *A = g.csr.colIdx + g.csr.indPtr[src];
    const int32_t *B = g.csr.colIdx + g.csr.indPtr[dst];
    c += isect_group<G, METHOD>(A, g.csr.deg[src], B, g.csr.deg[dst], cache, n_cache, glane);
  }
  __shared__ unsigned long long s_acc;
  block_atomic_add(count, c, &s_acc);
}

// List-driven variant of the routed path: one group of G lanes per arc, the arcs of the cached
// list (two-phase cached binary search, cache depth n_cache) first, then those of the merge list
// (arc = index into the CSR). One launch serves both lists; a group changes method only where
// the lists meet.
template <int G>
__global__ void __launch_bounds__(256) tc_edge_list_kernel(gary_device_graph g, int32_t n_cache,
                                                           const uint32_t *__restrict__ cached,
                                                           const int32_t *__restrict__ n_cached,
                                                           const uint32_t *__restrict__ merge,
                                                           const int32_t *__restrict__ n_merge,
                                                           unsigned long long *count) {
  extern __shared__ int32_t smem[];
  const int glane = lane_id() & (G - 1);
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G;
  const int64_t ngroups = (int64_t)gridDim.x * gpb;
  int32_t *cache = smem + (size_t)gib * n_cache;
  const int64_t nc = *n_cached, E = nc + *n_merge;
  const uint32_t m = (uint32_t)g.csr.m;
  uint64_t c = 0;
  for (int64_t k = (int64_t)blockIdx.x * gpb + gib; k < E; k += ngroups) {
    const uint32_t e = k < nc ? cached[k] : merge[k - nc];
    const int32_t src = g.coo.rowIdx[e], dst = g.csr.colIdx[e];
    if ((uint32_t)src >= m || (uint32_t)dst >= m) continue;
    const int32_t *A = g.csr.colIdx + g.csr.indPtr[src];
    const int32_t *B = g.csr.colIdx + g.csr.indPtr[dst];
    if (k < nc)
      c += isect_group<G, kIsectCached>(A, g.csr.deg[src], B, g.csr.deg[dst], cache, n_cache, glane);
    else
      c += isect_group<G, kIsectMerge>(A, g.csr.deg[src], B, g.csr.deg[dst], nullptr, 0, glane);
  }
  __shared__ unsigned long long s_acc;
  block_atomic_add(count, c, &s_acc);
}

// ======================================================================================
// Pivot-centric kernels
// ======================================================================================
//
// A pivot u owns an LDS membership set of N+(u); every row N+(v), v in N+(u), is streamed
// through it by a whole wave (lanes stride the row, four loads in flight per lane). On R-MAT
// scale 20, 99% of the probes fall in rows with at least 64 entries, so striding rows keeps the
// lanes busy without any per-probe row lookup.

__device__ __forceinline__ uint32_t hash_slot(int32_t w, int32_t log_t) {
  return ((uint32_t)w * 0x9E3779B1u) >> (32 - log_t);
}

// Membership set of one pivot list N+(u) in LDS.
//
// Bitmap mode (log_t == 0), a two-level rank bitmap over x = w - base, x < span:
//   word[cwords] : low 16 bits: bit c set iff some element has (x >> s) == 16 * i + c;
//                  high 16 bits: number of set coarse bits in the words before i
//   word cwords  : zero; a probe clamps its coarse index to it, so ids outside the span (and
//                  below base, which wrap) miss without a range branch
//   exact[...]   : one block of 2^s bits per set coarse bit, in coarse-bit order, from word
//                  cwords + 1
// A probe reads one word (coarse bit and rank together) and, when the coarse bit is set, one
// exact word; elements of one sorted row share coarse words, so neighbouring lanes mostly read
// the same addresses. (Coarse bits and ranks in separate arrays took one more dependent read
// per probe: 3.59 -> 3.50 ms per R-MAT 20 count with them packed.)
// Hash mode (log_t > 0): open addressing over 2^log_t slots (empty = -1, linear probing), used
// when the bitmap does not fit.
struct PivotSet {
  int32_t base;
  uint32_t span;
  int32_t s;       // bitmap: log2 of the ids per coarse bit (>= 5)
  int32_t cwords;  // bitmap: coarse words
  int32_t log_t;   // hash: log2 slots; 0 in bitmap mode
  int32_t words;   // words of the table in use (cleared at build)
  uint32_t ebase;  // bitmap: byte offset of the exact blocks, 4 * (cwords + 1)
  uint32_t emask;  // bitmap: (2^(s-3) - 1) & ~3, the in-block byte offset bits of x >> 3
};

// Words of the bitmap at block size 2^s, and their step words(s + 1) - words(s). The step
// strictly increases with s (the exact blocks' d * 2^(s-5) doubles while the coarse words' drop
// ceil(span / 2^(s+4)) - ceil(span / 2^(s+5)) never grows), so the least words over s in [5, 20]
// are at the first s whose step is >= 0 (or at 20), the same s a scan of every s picks (first
// minimum). The search starts at the continuous optimum, 2^(2s-1) = span / d, and moves at most
// a step or two: scanning all 16 block sizes was 0.16 ms of scalar work per R-MAT 20 launch.
__device__ __forceinline__ int64_t plan_words(uint32_t span, int32_t d, int32_t s) {
  return (((int64_t)span + (1ll << (s + 4)) - 1) >> (s + 4)) + 1 + ((int64_t)d << (s - 5));
}
__device__ __forceinline__ int64_t plan_step(uint32_t span, int32_t d, int32_t s) {
  const int64_t drop = (((int64_t)span + (1ll << (s + 4)) - 1) >> (s + 4)) -
                       (((int64_t)span + (1ll << (s + 5)) - 1) >> (s + 5));
  return ((int64_t)d << (s - 5)) - drop;
}

__device__ __forceinline__ bool pivot_plan(int32_t lo, int32_t hi, int32_t d, int32_t table_words,
                                           PivotSet &ps) {
  ps.base = lo;
  ps.span = (uint32_t)(hi - lo) + 1u;
  // bit lengths instead of a division: log2(span / d) to within one
  int32_t best_s = ((32 - __clz((int)ps.span)) - (32 - __clz(max(d, 1))) + 2) >> 1;
  best_s = min(max(best_s, 5), 20);
  while (best_s > 5 && plan_step(ps.span, d, best_s - 1) >= 0) --best_s;
  while (best_s < 20 && plan_step(ps.span, d, best_s) < 0) ++best_s;
  const int64_t best = plan_words(ps.span, d, best_s);
  if (best <= table_words) {
    ps.s = best_s;
    ps.cwords = (int32_t)(((int64_t)ps.span + (1ll << (best_s + 4)) - 1) >> (best_s + 4));
    ps.log_t = 0;
    ps.words = (int32_t)best;
    ps.ebase = 4u * (uint32_t)(ps.cwords + 1);
    ps.emask = ((1u << (best_s - 3)) - 1u) & ~3u;
    return true;
  }
  int32_t log_t = 5;
  while ((1 << log_t) < 4 * d) ++log_t;
  while ((1 << log_t) > table_words && (1 << (log_t - 1)) >= 2 * d) --log_t;
  if ((1 << log_t) <= table_words) {
    ps.s = 0;
    ps.cwords = 0;
    ps.log_t = log_t;
    ps.words = 1 << log_t;
    return true;
  }
  return false;
}

// Builds the set from the sorted list L[0..d). tid/nthr: this thread among the builders (a wave
// or a workgroup); BLOCK selects workgroup-wide synchronisation (the prefix scan then runs on
// wave 0 of the workgroup).
template <bool BLOCK>
__device__ __forceinline__ void pivot_sync() {
  if constexpr (BLOCK) __syncthreads();
  else wave_sync();
}

template <bool BLOCK>
__device__ __forceinline__ void pivot_build(int32_t *t, const PivotSet &ps,
                                            const int32_t *__restrict__ L, int32_t d, int tid,
                                            int nthr) {
  const int32_t fillv = ps.log_t ? -1 : 0;
  for (int32_t i = tid; i < ps.words; i += nthr) t[i] = fillv;
  pivot_sync<BLOCK>();
  if (ps.log_t) {
    const uint32_t mask = (1u << ps.log_t) - 1;
    for (int32_t i = tid; i < d; i += nthr) {
      const int32_t w = L[i];
      uint32_t h = hash_slot(w, ps.log_t);
      while (atomicCAS(&t[h], -1, w) != -1) h = (h + 1) & mask;
    }
    pivot_sync<BLOCK>();
    return;
  }
  uint32_t *coarse = (uint32_t *)t;
  uint32_t *exact = (uint32_t *)(t + ps.cwords + 1);
  const int32_t s = ps.s;
  for (int32_t i = tid; i < d; i += nthr) {
    const uint32_t x = (uint32_t)(L[i] - ps.base);
    if (x < ps.span) atomicOr(&coarse[x >> (s + 4)], 1u << ((x >> s) & 15));
  }
  pivot_sync<BLOCK>();
  if (tid < kWave) {  // rank prefix of the set coarse bits, into the high half of each word
    uint32_t carry = 0;
    for (int32_t c0 = 0; c0 < ps.cwords; c0 += kWave) {
      const int32_t i = c0 + tid;
      const uint32_t w = i < ps.cwords ? coarse[i] : 0u;
      const uint32_t v = (uint32_t)__popc(w);
      const uint32_t incl = wave_inclusive_scan_u32(v);
      if (i < ps.cwords) coarse[i] = w | ((carry + incl - v) << 16);
      carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
    }
  }
  pivot_sync<BLOCK>();
  for (int32_t i = tid; i < d; i += nthr) {
    const uint32_t x = (uint32_t)(L[i] - ps.base);
    if (x >= ps.span) continue;
    const uint32_t word = coarse[x >> (s + 4)], b = (x >> s) & 15;
    const uint32_t r = (word >> 16) + (uint32_t)__popc(word & ((1u << b) - 1));
    atomicOr(&exact[(r << (s - 5)) + ((x & ((1u << s) - 1)) >> 5)], 1u << (x & 31));
  }
  pivot_sync<BLOCK>();
}

// pivot_build for one wave and a list of at most kWave elements held in registers: e = L[lane],
// -1 past the end of the list.
__device__ __forceinline__ void pivot_build_reg(int32_t *t, const PivotSet &ps, int32_t e,
                                                int lane) {
  const int32_t fillv = ps.log_t ? -1 : 0;
  for (int32_t i = lane; i < ps.words; i += kWave) t[i] = fillv;
  wave_sync();
  if (ps.log_t) {
    if (e >= 0) {
      const uint32_t mask = (1u << ps.log_t) - 1;
      uint32_t h = hash_slot(e, ps.log_t);
      while (atomicCAS(&t[h], -1, e) != -1) h = (h + 1) & mask;
    }
    wave_sync();
    return;
  }
  uint32_t *coarse = (uint32_t *)t;
  uint32_t *exact = (uint32_t *)(t + ps.cwords + 1);
  const int32_t s = ps.s;
  const uint32_t x = (uint32_t)(e - ps.base);
  const bool in = e >= 0 && x < ps.span;
  if (in) atomicOr(&coarse[x >> (s + 4)], 1u << ((x >> s) & 15));
  wave_sync();
  uint32_t carry = 0;  // rank prefix of the set coarse bits, into the high half of each word
  for (int32_t c0 = 0; c0 < ps.cwords; c0 += kWave) {
    const int32_t i = c0 + lane;
    const uint32_t w = i < ps.cwords ? coarse[i] : 0u;
    const uint32_t v = (uint32_t)__popc(w);
    const uint32_t incl = wave_inclusive_scan_u32(v);
    if (i < ps.cwords) coarse[i] = w | ((carry + incl - v) << 16);
    carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
  }
  wave_sync();
  if (in) {
    const uint32_t word = coarse[x >> (s + 4)], b = (x >> s) & 15;
    const uint32_t r = (word >> 16) + (uint32_t)__popc(word & ((1u << b) - 1));
    atomicOr(&exact[(r << (s - 5)) + ((x & ((1u << s) - 1)) >> 5)], 1u << (x & 31));
  }
  wave_sync();
}

__device__ __forceinline__ uint32_t pivot_probe(const int32_t *t, const PivotSet &ps, int32_t w) {
  const uint32_t x = (uint32_t)(w - ps.base);
  if (x >= ps.span) return 0;
  if (ps.log_t == 0) {
    const int32_t s = ps.s;
    const uint32_t word = (uint32_t)t[x >> (s + 4)], b = (x >> s) & 15;
    if (!((word >> b) & 1u)) return 0;
    const uint32_t r = (word >> 16) + (uint32_t)__popc(word & ((1u << b) - 1));
    const uint32_t e = (uint32_t)t[ps.cwords + 1 + (r << (s - 5)) + ((x & ((1u << s) - 1)) >> 5)];
    return (e >> (x & 31)) & 1u;
  }
  const uint32_t mask = (1u << ps.log_t) - 1;
  uint32_t h = hash_slot(w, ps.log_t);
  while (true) {
    const int32_t v = t[h];
    if (v == w) return 1;
    if (v == -1) return 0;
    h = (h + 1) & mask;
  }
}

constexpr int kStreamWindows = 2;  // listing: 16-byte-chunk windows loaded before any is probed
constexpr int kChunk = 8;          // counting: elements per streamed chunk (32 bytes)

// Per-wave LDS scratch of the flattened row streaming: one slot per lane.
struct RowSlots {
  int32_t s[kWave];
};

// The bitmap and hash probes of pivot_probe, apart, so the element loops below are compiled
// once per mode with no per-element mode branch. The bitmap probe is written for its gfx950
// instructions, which bound the kernel: a clamp instead of the range branch, bit-field
// extracts for the coarse bit, its rank mask and the exact bit (whose offset the extract takes
// mod 32), and the exact word's byte address from per-pivot ebase / emask. 8 vector
// instructions reach the coarse bit and 8 more the exact bit (before: 10 and 11; R-MAT 20 count
// 3.11 -> 2.90 ms).
__device__ __forceinline__ uint32_t probe_bitmap(const int32_t *t, const PivotSet &ps, int32_t w) {
  const uint32_t x = (uint32_t)(w - ps.base);
  const int32_t s = ps.s;
  const uint32_t word = (uint32_t)t[min(x >> (s + 4), (uint32_t)ps.cwords)];
  const uint32_t b = __builtin_amdgcn_ubfe(x, s, 4);
  if (!__builtin_amdgcn_ubfe(word, b, 1)) return 0;
  const uint32_t r = (word >> 16) + (uint32_t)__popc(__builtin_amdgcn_ubfe(word, 0, b));
  const uint32_t off = (r << (s - 3)) + ((x >> 3) & ps.emask) + ps.ebase;
  const uint32_t e = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(t) + off);
  return __builtin_amdgcn_ubfe(e, x, 1);  // the bit field offset is x mod 32
}

__device__ __forceinline__ uint32_t probe_hash(const int32_t *t, const PivotSet &ps, int32_t w) {
  const uint32_t mask = (1u << ps.log_t) - 1;
  uint32_t h = hash_slot(w, ps.log_t);
  while (true) {
    const int32_t v = t[h];
    if (v == w) return 1;
    if (v == -1) return 0;
    h = (h + 1) & mask;
  }
}

template <int CH, bool HASH>
__device__ __forceinline__ uint32_t probe_chunk(const int4 (&data)[CH / 4], int32_t jlo, int32_t jhi,
                                                const int32_t *table, const PivotSet &ps) {
  if constexpr (HASH) {  // a hash probe loops: only the chunk's own elements are probed
    uint32_t c = 0;
#pragma unroll
    for (int h = 0; h < CH / 4; ++h) {
      const int32_t el[4] = {data[h].x, data[h].y, data[h].z, data[h].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int pos = 4 * h + q;
        if (jlo <= pos && pos < jhi) c += probe_hash(table, ps, el[q]);
      }
    }
    return c;
  } else {
    // Every slot is probed (a bitmap probe reads inside the set whatever the value) into a hit
    // mask, and the slots of the chunk's row [jlo, jhi) are selected once: a per-slot range
    // check cost two compares, a cleared result and an add per element.
    uint32_t hits = 0;
#pragma unroll
    for (int h = 0; h < CH / 4; ++h) {
      const int32_t el[4] = {data[h].x, data[h].y, data[h].z, data[h].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) hits |= probe_bitmap(table, ps, el[q]) << (4 * h + q);
    }
    const int32_t lo = min(max(jlo, 0), CH), hi = min(max(jhi, 0), CH);
    const uint32_t valid = hi > lo ? (((1u << (hi - lo)) - 1u) << lo) : 0u;
    return (uint32_t)__popc(hits & valid);
  }
}

// One 64-row batch of stream_rows: lane l holds the (start, length) of its row (length 0: no
// row). `tag` carries the slot tags from batch to batch of one pivot (1 << 30 before the first).
__device__ __forceinline__ uint32_t stream_batch(const int32_t *__restrict__ col, uint64_t nnz,
                                                 int64_t pv, int32_t dv, const int32_t *table,
                                                 const PivotSet &ps, bool hash, RowSlots &rs,
                                                 int lane, int32_t &tag) {
  constexpr int CH = kChunk;
  constexpr int LG = 3;
  uint32_t c = 0;
  const int32_t nch = dv > 0 ? (int32_t)(((pv + dv + CH - 1) >> LG) - (pv >> LG)) : 0;
  const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nch);
  const int32_t start = incl - nch;
  const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
  // chunk of batch position x: base_r + x; its valid slots: [A - CH*x, A + dv - CH*x)
  // (unsigned: the terms may wrap, the results for the row's own positions fit int32)
  const uint32_t base_r = (uint32_t)(pv >> LG) - (uint32_t)start;
  const uint32_t A = (uint32_t)(pv & (CH - 1)) + (uint32_t)CH * (uint32_t)start;
  const uint32_t Bv = A + (uint32_t)dv;
  for (int32_t w0 = 0; w0 < total; w0 += kWave) {
    if (tag >= (1 << 30)) {  // stale slots could outrank new tags: clear them
      rs.s[lane] = -1;
      tag = 0;
      wave_sync();
    }
    // slot of the window position where each row with chunks here begins (0 for the row
    // running into the window), tagged so that slots left by earlier windows lose the max
    if (nch > 0 && start < w0 + kWave && incl > w0) rs.s[max(0, start - w0)] = tag | lane;
    wave_sync();
    const int32_t row = wave_inclusive_max_i32(rs.s[lane]) & (kWave - 1);
    wave_sync();
    tag += kWave;
    const int32_t x = w0 + lane;
    const uint32_t ch = (uint32_t)__shfl((int32_t)base_r, row, kWave) + (uint32_t)x;
    const int32_t jlo = (int32_t)((uint32_t)__shfl((int32_t)A, row, kWave) - (uint32_t)CH * (uint32_t)x);
    int32_t jhi = (int32_t)((uint32_t)__shfl((int32_t)Bv, row, kWave) - (uint32_t)CH * (uint32_t)x);
    int4 data[CH / 4];
#pragma unroll
    for (int h = 0; h < CH / 4; ++h) data[h] = make_int4(0, 0, 0, 0);
    if (x < total) {
      const uint64_t e0 = (uint64_t)ch << LG;
      data[0] = *reinterpret_cast<const int4 *>(col + e0);
#pragma unroll
      for (int h = 1; h < CH / 4; ++h)
        if (e0 + 4 * h < nnz) data[h] = *reinterpret_cast<const int4 *>(col + e0 + 4 * h);
    } else {
      jhi = 0;
    }
    c += hash ? probe_chunk<CH, true>(data, jlo, jhi, table, ps)
              : probe_chunk<CH, false>(data, jlo, jhi, table, ps);
  }
  return c;
}

// Row header of forward row r of pivot list L: (row start, length) of N+(L[r]), length 0 when
// the arc is routed elsewhere or reversed (d+(v) > dmax: counted at pivot v).
__device__ __forceinline__ void forward_row(const gary_device_graph &g, int32_t v, int32_t dmax,
                                            bool routed, int64_t &pv, int32_t &dv) {
  pv = 0;
  dv = 0;
  if (routed && (uint32_t)v < (uint32_t)g.csr.m) {
    dv = g.csr.deg[v];
    pv = g.csr.indPtr[v];
    if (dv > dmax) dv = 0;
  }
}

// The rows r0, r0 + stride, ... (< d) of pivot list L, 64 rows per batch. A batch is flattened
// into the aligned chunks of kChunk elements (32 bytes) that overlap its rows; each window of 64
// chunks gives lane l chunk w0 + l, so a wave's loads cover one contiguous stretch of colIdx.
// The row owning chunk w0 + l is the last row starting at or before it: each row with chunks in
// the window writes (tag | its lane) into the LDS slot of its first chunk there (the row running
// into the window writes slot 0), and an inclusive max-scan over the slots hands every lane its
// row. The tag grows by window, so slots left by earlier windows lose the max and need no reset.
// One shuffle each then gives the lane its chunk in colIdx and the range of its valid slots.
// Elements of a chunk outside its row are skipped. A chunk is read as two aligned 16-byte
// halves; the second only when it holds an index below nnz, so no read leaves the 16-byte
// blocks that hold colIdx elements. Chunk indices are 32-bit (colIdx below 2^34 elements).
// (16-byte chunks with two windows in flight: 3.38 ms per R-MAT 20 count; 32-byte chunks with
// a reset slot array and four shuffles: 3.21 ms; this form: 3.10 ms.)
// The rows streamed are N+(v) for the row vertices L[0..d) (the pivot's out-list: forward rows,
// skipped when d+(v) > dmax, since such an arc is reversed and streamed at v) and then the rows
// R2[0..d2) given by (row start, length) (the sources of the arcs reversed onto this pivot), rows
// r0, r0 + stride, ... of the concatenation. The returned per-lane count is below 2^32: a lane
// probes at most one chunk (CH elements) per window, and a call has at most (Σ_v (d(v) / CH +
// 2)) / 64 + (d + d2) / 64 + 1 windows, so the count is at most nnz / 64 + 3(d + d2) / 8 + 8 <
// 2^28 + 2^31 + 8 for colIdx below 2^34 elements and d + d2 < 2^32.
__device__ __forceinline__ uint32_t stream_rows(const gary_device_graph &g,
                                                const int32_t *__restrict__ L, int32_t d,
                                                const uint2 *__restrict__ R2, int32_t d2,
                                                int32_t dmax, int32_t r0, int32_t stride,
                                                const int32_t *table, const PivotSet &ps,
                                                RowSlots &rs, int lane,
                                                const uint8_t *__restrict__ Lroute) {
  const uint64_t nnz = (uint64_t)g.csr.nnz;
  const bool hash = ps.log_t != 0;
  const int32_t dtot = d + d2;
  uint32_t c = 0;
  int32_t tag = 1 << 30;  // forces a slot reset before the first window
  for (int32_t b = r0; b < dtot; b += stride * kWave) {
    const int32_t r = b + lane * stride;
    int32_t dv = 0;
    int64_t pv = 0;
    if (r < d) {
      forward_row(g, L[r], dmax, !Lroute || Lroute[r] == kRouteBitmap, pv, dv);
    } else if (r < dtot) {
      const uint2 rw = R2[r - d];
      pv = rw.x;
      dv = (int32_t)rw.y;
    }
    c += stream_batch(g.csr.colIdx, nnz, pv, dv, table, ps, hash, rs, lane, tag);
  }
  return c;
}

// Edge-by-edge fallback for a pivot whose set does not fit in LDS: rows r0, r0+stride, ... of
// the same row lists as stream_rows. A lane's count per row is at most ceil(min(d, d(v)) / 64)
// + 1, so the call's is below nnz / 64 + d + d2 < 2^32 for d + d2 < 2^31.
__device__ __forceinline__ uint32_t pivot_fallback(const gary_device_graph &g,
                                                   const int32_t *__restrict__ L, int32_t d,
                                                   const uint2 *__restrict__ R2, int32_t d2,
                                                   int32_t dmax, int32_t r0, int32_t stride,
                                                   int lane, const uint8_t *__restrict__ Lroute) {
  const uint32_t m = (uint32_t)g.csr.m;
  uint32_t c = 0;
  for (int32_t k = r0; k < d + d2; k += stride) {
    int64_t pv;
    int32_t dv;
    if (k < d) {
      if (Lroute && Lroute[k] != kRouteBitmap) continue;
      const int32_t v = L[k];
      if ((uint32_t)v >= m) continue;
      dv = g.csr.deg[v];
      if (dv > dmax) continue;  // reversed: counted at pivot v
      pv = g.csr.indPtr[v];
    } else {
      const uint2 rw = R2[k - d];
      pv = rw.x;
      dv = (int32_t)rw.y;
    }
    c += isect_group<64, kIsectBsearch>(L, d, g.csr.colIdx + pv, dv, nullptr, 0, lane);
  }
  return c;
}

// Persistent pivot kernel.
// Phase 1: each workgroup takes the next pivot of `big` from an atomic cursor, builds one
//   shared set of TABLE_WORDS words and its WAVES waves split the rows.
// Phase 2: each wave takes a few pivots of `small` at a time and works alone with a
//   1/WAVES share of the table memory.
// counts[0] = #small, counts[1] = #big; cursors[1] (the big list's) and the small list's striped
// cursors start at zero.
// Per-graph mode (gid_of != nullptr, batched disjoint unions): each wave adds its partial count
// to gcount[gid_of[u]] whenever the pivot's graph changes, instead of to *count.
// The small list is ordered heaviest first; its first split[0] pivots (weight bucket >=
// kLightLog) are claimed H at a time, the rest Lc at a time (below).
// Small-pivot claims are striped over kClaimStripes cursors kStripeSpacing words apart: claim k
// belongs to stripe k % kClaimStripes, whose cursor hands out k / kClaimStripes; a wave starts at
// stripe blockIdx.x % kClaimStripes and moves on when its stripe is exhausted. With one cursor,
// 8,192 waves taking short pivots queue on it: claims alone (no pivot work) took 0.46 ms of an
// R-MAT 20 launch; striped, the launch went 1.99 -> 1.90 ms (Chung-Lu 4M 7.12 -> 6.74 ms). Checking
// an exhausted stripe with a coherent load before its atomic cost more than it saved (2.17 ms).
// A workgroup remembers the stripes its waves found exhausted (an LDS bit each), so at the end
// of the launch a stripe's cursor takes one failed fetch-add per workgroup rather than per wave.
constexpr int kClaimStripes = 8;
constexpr int kStripeSpacing = 64;  // words
// Claim sizes (pivots per atomic on the small list's cursors), forward / with reversed rows. A
// claim costs a fetch-add serialised with every other wave's on one address (about 10 ns each)
// and one chain of header loads; with reversed rows the small pivots stream little (their arcs
// to much longer out-lists are streamed at the other end), so larger claims pay: R-MAT 20
// 2.24 -> 2.03 ms per launch at 8 / 24 against 4 / 16, while forward streaming loses balance
// with them (2.90 -> 3.13 ms) and gains a little with smaller ones (2 / 12 against 4 / 16:
// R-MAT 20 forward 2.84 -> 2.81 ms per count, the 2,000-graph batch 13.79 -> 13.57 ms; 4 / 64
// and 8 / 64: 17.5 ms).
constexpr int kSmallBatchHeavy = 2, kSmallBatchHeavyRev = 8;
constexpr int kSmallBatchLight = 12, kSmallBatchLightRev = 24;
static_assert(kSmallBatchHeavyRev <= kWave && kSmallBatchLightRev <= kWave &&
                  kSmallBatchHeavy <= kWave && kSmallBatchLight <= kWave,
              "one pivot header per lane");
constexpr int kLightLog = 4;
static_assert(kLightLog >= 1 && kLightLog <= 15, "light-pivot threshold");

// Reverse rows (rrange != nullptr): pivot u also streams the rows rrows[rrange[u].x ..
// rrange[u].y), the sources x of the arcs (x, u) with d+(x) + kRevSlack < d+(u), and skips its
// own out-neighbours v with d+(v) > d+(u) + kRevSlack, whose arcs are streamed at v.
__device__ __forceinline__ int32_t rev_rows(const uint2 *__restrict__ rrange,
                                            const uint2 *__restrict__ rrows, int32_t u,
                                            const uint2 *&R2) {
  if (!rrange) {
    R2 = nullptr;
    return 0;
  }
  const uint2 r = rrange[u];
  R2 = rrows + r.x;
  return (int32_t)(r.y - r.x);
}

template <int WAVES, int TABLE_WORDS>
__global__ void __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(8)))
    tc_pivot_kernel(gary_device_graph g, const int32_t *__restrict__ small,
                    const int32_t *__restrict__ big, const int32_t *__restrict__ counts,
                    int32_t *__restrict__ cursors, unsigned long long *count,
                    const int32_t *__restrict__ gid_of, unsigned long long *gcount,
                    const uint8_t *__restrict__ route, const RouteTable *__restrict__ rtable,
                    const int32_t *__restrict__ split, const uint2 *__restrict__ rrange,
                    const uint2 *__restrict__ rrows, int32_t *__restrict__ stripes) {
  __shared__ int32_t table[TABLE_WORDS];
  if (rtable && !rtable->mixed) route = nullptr;  // every bin routes to the bitmap
  __shared__ RowSlots s_rows[WAVES];
  __shared__ uint32_t s_spent;  // bit x: this workgroup has found claim stripe x exhausted
  // No separate LDS word for the pivot index (it is passed through table[0] between pivots,
  // when the set is dead): every word kept out of LDS helps eight workgroups fit a CU.
  const int lane = lane_id();
  const int wib = threadIdx.x / kWave;
  RowSlots &rb = s_rows[wib];
  const int32_t n_small = counts[0], n_big = counts[1];
  if (n_small == 0 && n_big == 0) return;
  const uint32_t m = (uint32_t)g.csr.m;
  const int32_t *__restrict__ col = g.csr.colIdx;
  const uint64_t nnz = (uint64_t)g.csr.nnz;
  uint64_t c = 0;
  if (threadIdx.x == 0) s_spent = 0u;  // published by phase 1's first barrier
  // ---- phase 1: big pivots, one workgroup each
  while (true) {
    if (threadIdx.x == 0) table[0] = atomicAdd(&cursors[1], 1);
    __syncthreads();
    const int32_t p = table[0];
    __syncthreads();  // every thread holds p before the next set overwrites table[0]
    if (p >= n_big) break;
    const int32_t u = big[p];
    const int32_t d = (uint32_t)u < m ? g.csr.deg[u] : 0;
    if (d >= 2) {
      const int32_t *L = g.csr.colIdx + g.csr.indPtr[u];
      const uint2 *L2;
      const int32_t d2 = rev_rows(rrange, rrows, u, L2);
      const int32_t dmax = rrange && d < kRevMaxSourceDeg ? d + kRevSlack : INT32_MAX;
      PivotSet s;
      if (pivot_plan(L[0], L[d - 1], d, TABLE_WORDS, s)) {
        pivot_build<true>(table, s, L, d, threadIdx.x, WAVES * 64);
        c += stream_rows(g, L, d, L2, d2, dmax, wib, WAVES, table, s, rb, lane,
                         route ? route + g.csr.indPtr[u] : nullptr);
      } else {
        c += pivot_fallback(g, L, d, L2, d2, dmax, wib, WAVES, lane,
                            route ? route + g.csr.indPtr[u] : nullptr);
      }
      if (gid_of) {
        wave_atomic_add(&gcount[gid_of[u]], c);
        c = 0;
      }
    }
    __syncthreads();  // the set (and table[0]) are rewritten for the next pivot
  }
  // ---- phase 2: small pivots, one wave each
  int32_t *wt = table + wib * (TABLE_WORDS / WAVES);
  int32_t cur = -1;  // graph of the wave's partial count (per-graph mode)
  // claim k covers pivots [k*H, k*H + H) while k < ceil(n_heavy / H), then light claims of L
  const int32_t H = rrange ? kSmallBatchHeavyRev : kSmallBatchHeavy;
  const int32_t Lc = rrange ? kSmallBatchLightRev : kSmallBatchLight;
  const int32_t n_heavy = min(split[0], n_small);
  const int32_t heavy_claims = (n_heavy + H - 1) / H;
  const int32_t all_claims = heavy_claims + (n_small - n_heavy + Lc - 1) / Lc;
  while (true) {
    int32_t k = all_claims;
    if (lane == 0 && all_claims > 0) {
      for (int t = 0; t < kClaimStripes; ++t) {
        const int x = (int)((blockIdx.x + t) & (kClaimStripes - 1));
        if ((__atomic_load_n(&s_spent, __ATOMIC_RELAXED) >> x) & 1u) continue;
        const int64_t kk = x + (int64_t)kClaimStripes * atomicAdd(&stripes[x * kStripeSpacing], 1);
        if (kk < all_claims) {
          k = (int32_t)kk;
          break;
        }
        atomicOr(&s_spent, 1u << x);
      }
    }
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, kWave));
    const int32_t base = k < heavy_claims ? k * H : n_heavy + (k - heavy_claims) * Lc;
    if (base >= n_small) break;
    const int32_t end = min(base + (k < heavy_claims ? H : Lc), k < heavy_claims ? n_heavy : n_small);
    // The claim's pivot headers, one pivot per lane: id, out-degree, row start, reversed rows,
    // graph and first and last out-neighbour, so the claim pays one chain of dependent loads
    // instead of one per pivot (R-MAT 20: the small pivots' plan and set builds took 0.81 ms
    // of the launch with a chain per pivot).
    const int32_t cnt = end - base;
    int32_t hu = -1, hd = 0, hlo = 0, hhi = 0, hg = -1;
    int64_t hp = 0;
    uint2 hr = make_uint2(0u, 0u);
    if (lane < cnt) {
      hu = small[base + lane];
      if ((uint32_t)hu < m) {
        hd = g.csr.deg[hu];
        hp = g.csr.indPtr[hu];
        if (rrange) hr = rrange[hu];
        if (gid_of) hg = gid_of[hu];
      }
      if (hd >= 2) {
        hlo = g.csr.colIdx[hp];
        hhi = g.csr.colIdx[hp + hd - 1];
      }
    }
    // The claim's set plans, one pivot per lane in vector code, packed as 1 | s << 1 |
    // log_t << 6 | cwords << 10 | piped << 22 (0: the set does not fit; the words follow from
    // them and d; piped: at most kWave out-neighbours and no reverse rows) with the span
    // beside it. Planned per pivot, the plan's scalar code ran on the CU's one
    // scalar unit for every pivot of every wave: 1.8 ms of a 2,000-graph batch launch (20 M
    // pivots).
    static_assert(TABLE_WORDS / WAVES <= 4096, "coarse words fit 12 bits");
    int32_t hplan = 0;
    uint32_t hspan = 0;
    if (lane < cnt && hd >= 2) {
      PivotSet ps;
      if (pivot_plan(hlo, hhi, hd, TABLE_WORDS / WAVES, ps)) {
        const bool pip = hd <= kWave && hr.y == hr.x;
        hplan = 1 | ps.s << 1 | ps.log_t << 6 | ps.cwords << 10 | (pip ? 1 << 22 : 0);
        hspan = ps.span;
      }
    }
    // Pivots without reverse rows and with at most kWave out-neighbours are pipelined: the next
    // such pivot's out-list is loaded (one element per lane) while this one is built and
    // streamed, and the row headers of this one's out-neighbours are loaded before its set is
    // built, so a pivot waits on two dependent global loads (row headers, then chunks) instead
    // of four (out-list, out-list again for the rows, row headers, chunks). (Two deep, the next
    // pivot's row headers in flight as well, the live registers spilled: 2,000-graph batch
    // 13.23 -> 13.80 ms, R-MAT 20 2.024 -> 2.042 ms.)
    auto readlane64 = [&](int64_t x, int32_t jj) {
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(x >> 32), jj);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)x, jj);
      return (int64_t)(((uint64_t)hi << 32) | lo);
    };
    auto piped = [&](int32_t jj) {
      return ((__builtin_amdgcn_readlane(hplan, jj) >> 22) & 1) != 0;  // set only with d >= 2
    };
    auto prefetch = [&](int32_t jj) {
      if (jj >= cnt || !piped(jj)) return -1;
      const int32_t dj = __builtin_amdgcn_readlane(hd, jj);
      return lane < dj ? col[readlane64(hp, jj) + lane] : -1;
    };
    int32_t nx = prefetch(0);
    for (int32_t j = 0; j < cnt; ++j) {
      const int32_t e = nx;
      nx = prefetch(j + 1);
      const int32_t d = __builtin_amdgcn_readlane(hd, j);
      if (d < 2) continue;
      if (gid_of) {
        const int32_t gu = __builtin_amdgcn_readlane(hg, j);
        if (gu != cur) {
          if (cur >= 0) wave_atomic_add(&gcount[cur], c);
          c = 0;
          cur = gu;
        }
      }
      const int64_t pu = readlane64(hp, j);
      const int32_t *L = col + pu;
      const uint32_t rb0 = (uint32_t)__builtin_amdgcn_readlane((int32_t)hr.x, j);
      const int32_t d2 = (int32_t)((uint32_t)__builtin_amdgcn_readlane((int32_t)hr.y, j) - rb0);
      const uint2 *L2 = rrange ? rrows + rb0 : nullptr;
      const int32_t dmax = rrange && d < kRevMaxSourceDeg ? d + kRevSlack : INT32_MAX;
      const uint8_t *Lroute = route ? route + pu : nullptr;
      const int32_t pl = __builtin_amdgcn_readlane(hplan, j);
      if (pl == 0) {
        c += pivot_fallback(g, L, d, L2, d2, dmax, 0, 1, lane, Lroute);
        continue;
      }
      PivotSet s;
      s.base = __builtin_amdgcn_readlane(hlo, j);
      s.span = (uint32_t)__builtin_amdgcn_readlane((int32_t)hspan, j);
      s.s = (pl >> 1) & 31;
      s.log_t = (pl >> 6) & 15;
      s.cwords = (pl >> 10) & 4095;
      s.words = s.log_t ? 1 << s.log_t : s.cwords + 1 + (d << (s.s - 5));
      s.ebase = 4u * (uint32_t)(s.cwords + 1);
      s.emask = s.log_t ? 0u : ((1u << (s.s - 3)) - 1u) & ~3u;
      if ((pl >> 22) & 1) {
        int64_t pv;
        int32_t dv;
        forward_row(g, e, dmax, e >= 0 && (!Lroute || Lroute[lane] == kRouteBitmap), pv, dv);
        pivot_build_reg(wt, s, e, lane);
        int32_t tag = 1 << 30;
        c += stream_batch(col, nnz, pv, dv, wt, s, s.log_t != 0, rb, lane, tag);
      } else {
        pivot_build<false>(wt, s, L, d, lane, kWave);
        c += stream_rows(g, L, d, L2, d2, dmax, 0, 1, wt, s, rb, lane, Lroute);
      }
      wave_sync();  // the set is rebuilt for the next pivot
    }
  }
  if (!gid_of) {
    __shared__ unsigned long long s_acc;
    block_atomic_add(count, c, &s_acc);
  } else if (cur >= 0) {
    wave_atomic_add(&gcount[cur], c);
  }
}

// Zeroes `n` words (graph-capturable stand-in for a memset).
__global__ void fill_zero_kernel(uint32_t *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ======================================================================================
// Degree binning of pivots (device work lists)
// ======================================================================================

// bins: 0 = small (2 <= d <= t_small), 1 = big (t_small < d <= t_big), 2 = hub (d > t_big).
// Vertices with out-degree < 2 are dropped (no triangle has them as its lowest vertex).
// Routed mode (rowmask != 0xffffffff): pivots whose degree row routes no bin to the bitmap are
// dropped; with predicted costs (pcost) the small/big split is by cost (cost_big, ns) instead of
// by out-degree. Hubs always go by degree.
// Within each list, pivots are ordered heaviest first by bucket 15 - min(15, log2 weight)
// (weight = out-degree, or the predicted cost in cost mode), so the persistent kernels take the
// long pivots early and the tail of the launch is made of short ones.
constexpr int kBuckets = 16;
constexpr int32_t kRevRowsBig = 256;  // reverse-row mode: pivots with more rows get a workgroup
constexpr int kKeys = 3 * kBuckets;
// counter words: [0] #small, [1] #big, [2] always 0, [3] #hub, [4..5] main cursors,
// [6..7] hub cursors, [8, 8+kKeys) histogram, [8+kKeys, 8+2*kKeys) scatter cursors
// The histogram and the scatter cursors have kBinCopies copies: workgroup b of both kernels
// (the same vertex range) uses copy b % kBinCopies, so fewer workgroups contend for one word.
constexpr int kBinCopies = 8;
constexpr int kHistOff = 8;
constexpr int kCurOff = kHistOff + kBinCopies * kKeys;
// [kSplitOff] heavy/light split of the small list; [+2] the same for the hub launch (whose
// small list is empty: always 0)
constexpr int kSplitOff = kCurOff + kBinCopies * kKeys;
// [kStripeOff, + kClaimStripes * kStripeSpacing) the small list's striped claim cursors (the hub
// launch, whose small list is empty, makes no claim on them)
constexpr int kStripeOff = 1024;
static_assert(kSplitOff + 4 <= kStripeOff &&
                  kStripeOff + kClaimStripes * kStripeSpacing <= kPivotCounterWords,
              "pivot counter words");

// ordered == false (per-graph counting of batched unions) puts every small pivot in the
// lightest bucket: the small list then keeps vertex order, so consecutive pivots of a wave mostly
// belong to one graph and its per-graph partial count is flushed rarely (big and hub pivots
// flush once each anyway and stay heaviest first).
__device__ __forceinline__ int pivot_key(const int32_t *__restrict__ deg, int64_t v, int64_t end,
                                         int32_t t_small, int32_t t_big, uint32_t rowmask,
                                         const float *__restrict__ pcost, float cost_big,
                                         bool ordered, const uint2 *__restrict__ rrange) {
  if (v >= end) return -1;
  const int32_t d = deg[v];
  if (d < 2) return -1;
  const int lg = min(15, 31 - __clz(d));
  // rows reversed onto v (their bins were checked against the route table when listed)
  int32_t rr = 0;
  if (rrange) {
    const uint2 r = rrange[v];
    rr = (int32_t)(r.y - r.x);
  }
  if (rr == 0 && !((rowmask >> lg) & 1u)) return -1;
  if (d > t_big) return 2 * kBuckets + (15 - lg);
  if (pcost && cost_big > 0.f) {
    const float cst = pcost[v];
    if (!(cst > 0.f) && rr == 0) return -1;
    const int lc = min(15, max(0, (int)log2f(fmaxf(cst, 1.f))));
    return cst > cost_big || (rrange && d + rr > kRevRowsBig) ? kBuckets + (15 - lc)
                                                  : (ordered ? 15 - lc : 15);
  }
  if (rrange) {
    // weight: rows (at most d forward and rr reversed) times the set size, each row being at
    // most d + kRevSlack long; the pivot gets a workgroup for a large set or many rows
    const uint64_t w = (uint64_t)(d + rr) * (uint64_t)d;
    const int lw = min(15, (63 - __clzll(w)) >> 1);
    const bool big = d > t_small || d + rr > kRevRowsBig;
    return big ? kBuckets + (15 - lw) : (ordered ? 15 - lw : 15);
  }
  return d <= t_small ? (ordered ? 15 - lg : 15) : kBuckets + (15 - lg);
}

__global__ void __launch_bounds__(256)
    bin_hist_kernel(const int32_t *__restrict__ deg, int32_t m, int32_t t_small, int32_t t_big,
                    int32_t *__restrict__ counters, const RouteTable *__restrict__ table,
                    const float *__restrict__ pcost, float cost_big, bool ordered,
                    const uint2 *__restrict__ rrange) {
  const uint32_t rowmask = table ? table->rowmask : 0xffffffffu;
  __shared__ int32_t s_hist[kKeys];
  const int lane = lane_id();
  for (int i = threadIdx.x; i < kKeys; i += blockDim.x) s_hist[i] = 0;
  __syncthreads();
  const int64_t chunk = (((int64_t)m + gridDim.x - 1) / gridDim.x + 63) & ~63ll;
  const int64_t beg = (int64_t)blockIdx.x * chunk, end = min((int64_t)m, beg + chunk);
  for (int64_t base = beg + (threadIdx.x & ~63); base < end; base += blockDim.x) {
    const int key =
        pivot_key(deg, base + lane, end, t_small, t_big, rowmask, pcost, cost_big, ordered,
                  rrange);
    for_each_key(key, [&](int k, uint64_t mask, int leader) {
      if (lane == leader) atomicAdd(&s_hist[k], __popcll(mask));
    });
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kKeys; i += blockDim.x)
    if (s_hist[i]) atomicAdd(&counters[kHistOff + (blockIdx.x % kBinCopies) * kKeys + i], s_hist[i]);
}

// Same vertex ranges as bin_hist_kernel; bucket starts come from the finished histogram.
__global__ void __launch_bounds__(256)
    bin_scatter_kernel(const int32_t *__restrict__ deg, int32_t m, int32_t t_small,
                       int32_t t_big, int32_t *__restrict__ lists, int64_t stride,
                       int32_t *__restrict__ counters, const RouteTable *__restrict__ table,
                       const float *__restrict__ pcost, float cost_big, bool ordered,
                       const uint2 *__restrict__ rrange) {
  const uint32_t rowmask = table ? table->rowmask : 0xffffffffu;
  __shared__ int32_t s_cnt[kKeys], s_base[kKeys], s_start[kKeys], s_tot[kKeys], s_pre[kKeys];
  const int lane = lane_id();
  const int copy = blockIdx.x % kBinCopies;
  if (threadIdx.x < kKeys) {  // key totals, and the part counted by lower copies
    int32_t tot = 0, pre = 0;
    for (int c = 0; c < kBinCopies; ++c) {
      const int32_t h = counters[kHistOff + c * kKeys + threadIdx.x];
      pre += c < copy ? h : 0;
      tot += h;
    }
    s_tot[threadIdx.x] = tot;
    s_pre[threadIdx.x] = pre;
  }
  __syncthreads();
  if (threadIdx.x < 3) {  // bucket starts of list threadIdx.x (heaviest bucket first)
    const int b = threadIdx.x;
    int32_t run = 0;
    for (int k = 0; k < kBuckets; ++k) {
      s_start[b * kBuckets + k] = run;
      run += s_tot[b * kBuckets + k];
    }
    if (blockIdx.x == 0) {
      counters[b == 2 ? 3 : b] = run;
      if (b == 0) counters[kSplitOff] = s_start[15 - kLightLog + 1];  // buckets with lg >= kLightLog
    }
  }
  for (int i = threadIdx.x; i < kKeys; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
  const int64_t chunk = (((int64_t)m + gridDim.x - 1) / gridDim.x + 63) & ~63ll;
  const int64_t beg = (int64_t)blockIdx.x * chunk, end = min((int64_t)m, beg + chunk);
  for (int64_t base = beg + (threadIdx.x & ~63); base < end; base += blockDim.x) {
    const int key =
        pivot_key(deg, base + lane, end, t_small, t_big, rowmask, pcost, cost_big, ordered,
                  rrange);
    for_each_key(key, [&](int k, uint64_t mask, int leader) {
      if (lane == leader) atomicAdd(&s_cnt[k], __popcll(mask));
    });
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kKeys; i += blockDim.x) {
    s_base[i] = s_cnt[i]
        ? s_start[i] + s_pre[i] + atomicAdd(&counters[kCurOff + copy * kKeys + i], s_cnt[i]) : 0;
    s_cnt[i] = 0;
  }
  __syncthreads();
  for (int64_t base = beg + (threadIdx.x & ~63); base < end; base += blockDim.x) {
    const int key =
        pivot_key(deg, base + lane, end, t_small, t_big, rowmask, pcost, cost_big, ordered,
                  rrange);
    for_each_key(key, [&](int k, uint64_t mask, int leader) {
      int32_t pos0 = 0;
      if (lane == leader) pos0 = atomicAdd(&s_cnt[k], __popcll(mask));
      pos0 = __shfl(pos0, leader, kWave);
      if (key == k)
        lists[(k / kBuckets) * stride + s_base[k] + pos0 + lanes_below(mask)] = (int32_t)(base + lane);
    });
  }
}

// ======================================================================================
// Reverse rows: the arcs streamed at their target
// ======================================================================================
//
// Arc (u, v) is reversed when d+(u) + kRevSlack < d+(v): its intersection is then taken by
// streaming N+(u) through the set of N+(v) at pivot v, so each arc streams the shorter of the
// two out-lists (on R-MAT 20: 1.97 G probes and 260 M 32-byte chunks instead of 4.43 G and
// 567 M). Only arcs the route table sends to the bitmap are listed (all arcs when unrouted).
//
// The reversed arcs are grouped by target with a two-level counting sort whose counters all live
// in LDS (device-scope atomics run at about 6.5 G/s here: one global fetch-add per arc and pass
// took 3.8 ms for the 11.1 M reversed arcs of R-MAT 20 at slack 8):
//   1. rev_hist_kernel: tile t (a range of the COO's arcs, a multiple of kRevChunk) counts its
//      reversed arcs per bucket into ghist[bucket * tiles + t];
//   2. an exclusive scan of ghist gives every (bucket, tile) its first position;
//   3. rev_place_kernel: the tile reads its arcs again, chunk by chunk, counting-sorts each
//      chunk's reversed arcs by bucket in LDS and writes them, in runs, as 64-bit words
//      h << 51 | d+(u) << 32 | indPtr[u] (h < 2^13 ids per bucket, d+(u) < 2^19);
//   4. rev_group_kernel: one workgroup per bucket counts its words per target in LDS, scans the
//      counts, writes rrange[v] for every v of the bucket and places each source's row (start,
//      length).
// Buckets and ids within them come from a multiplicative hash of the target (rev_hash).
// The tile passes read each arc's source from the COO rowIdx, so the grouping needs the COO
// (without it the pivot path streams every arc forward). Walking 64-row batches of the CSR
// instead, with rows found by a slot write and a max-scan per 64 arcs, took 174 + 238 us for
// the two tile passes on R-MAT 20: two dependent loads per step and one step in flight.

constexpr int kRevBlock = 512;
constexpr int kRevPer = kRevChunk / kRevBlock;  // arcs of a chunk per thread
constexpr int kRevGroupBlock = 1024;
constexpr int kRevUnroll = 4;  // pairs a thread of rev_group_kernel loads before using them
static_assert(kRevChunk % kRevBlock == 0, "whole arcs per thread");
static_assert(kRevMaxBuckets <= 65536 && kRevMaxBucketIds <= (1 << 13) &&
                  kRevMaxSourceDeg <= (1 << 19),
              "a pair word holds h in 13 bits and d+(u) in 19; a bucket id fits 16 bits");

// Hashed id of target v: x = v * kRevMul mod 2^k, a bijection on [0, 2^k); its top lg bits are
// the bucket, its low k - lg bits the dense index h within the bucket, and v = x * kRevMulInv
// mod 2^k. The product mixes all bits of v into the top ones, so R-MAT's hubs (ids with few set
// bits) spread over the buckets: on R-MAT 20 the largest bucket holds 26.5 K of 7.7 M reversed
// arcs (mean 15 K); folding the id's bits (v ^ v >> lg ^ v >> 2 lg) left 84 K in one bucket and
// the grouping pass at 125 us, and buckets by id range 688 K of 11.1 M arcs (slack 8).
constexpr uint32_t kRevMul = 0x9E3779B1u;
constexpr uint32_t kRevMulInv = 0x0E8B2F51u;  // kRevMul * kRevMulInv = 1 mod 2^32
static_assert(kRevMul * kRevMulInv == 1u, "multiplicative inverse mod 2^32");

__device__ __forceinline__ uint32_t rev_hash(uint32_t v, const RevShape &sh) {
  return (v * kRevMul) & ((sh.k == 32 ? 0u : (1u << sh.k)) - 1u);
}
__device__ __forceinline__ uint32_t rev_bucket(uint32_t v, const RevShape &sh) {
  return rev_hash(v, sh) >> (sh.k - sh.lg);
}

// dsat[v] = min(d+(v), 255): the reversal test reads one byte per arc (an array of m bytes stays
// in L2) and the exact out-degree only when both out-degrees are large.
__global__ void __launch_bounds__(256) rev_dsat_kernel(const int32_t *__restrict__ deg, int32_t m,
                                                       uint8_t *__restrict__ dsat) {
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < m; v += (int64_t)gridDim.x * 256)
    dsat[v] = (uint8_t)min(deg[v], 255);
}

// The arcs [e, e + kRevChunk) of the COO (clipped to e_end), kRevPer per thread: v[q] = target
// of a reversed arc (d+(u) + kRevSlack < d+(v) and d+(u) < kRevMaxSourceDeg), else -1; u[q] its
// source, du[q] = d+(u). The test reads the byte dsat[v] and the exact out-degree only when it
// saturates; it is written as selects: with early returns inside the unrolled loop the
// compiler kept v[q] for every arc.
template <bool ROWS = false>
__device__ __forceinline__ void rev_load_chunk(const gary_device_graph &g,
                                               const uint8_t *__restrict__ route,
                                               const uint8_t *__restrict__ dsat, int64_t e,
                                               int64_t e_end, int32_t (&u)[kRevPer],
                                               int32_t (&v)[kRevPer], int32_t (&du)[kRevPer],
                                               uint32_t *pu = nullptr) {
  const uint32_t m = (uint32_t)g.csr.m;
#pragma unroll
  for (int q = 0; q < kRevPer; ++q) {
    const int64_t a = e + q * kRevBlock + threadIdx.x;
    const bool in = a < e_end && (!route || route[a] == kRouteBitmap);
    u[q] = in ? g.coo.rowIdx[a] : -1;
    v[q] = in ? g.csr.colIdx[a] : -1;
  }
  int32_t sv[kRevPer];
#pragma unroll
  for (int q = 0; q < kRevPer; ++q) {
    const bool ok = (uint32_t)u[q] < m && (uint32_t)v[q] < m;
    du[q] = ok ? g.csr.deg[u[q]] : INT32_MAX - kRevSlack;
    sv[q] = ok ? (int32_t)dsat[v[q]] : 0;
    // the source's row start beside its degree, not after the test (one dependent load less)
    if constexpr (ROWS) pu[q] = ok ? (uint32_t)g.csr.indPtr[u[q]] : 0u;
  }
#pragma unroll
  for (int q = 0; q < kRevPer; ++q) {
    const int32_t need = du[q] + kRevSlack;  // reversed iff need < d+(v)
    const bool sat = sv[q] == 255 && need >= 255;
    const int32_t dv = sat ? g.csr.deg[v[q]] : sv[q];
    v[q] = need < dv && du[q] < kRevMaxSourceDeg ? v[q] : -1;
  }
}

__global__ void __launch_bounds__(kRevBlock)
    rev_hist_kernel(gary_device_graph g, const uint8_t *__restrict__ route,
                    const RouteTable *__restrict__ rtable, const uint8_t *__restrict__ dsat,
                    RevShape sh, uint32_t *__restrict__ ghist) {
  __shared__ uint32_t s_hist[kRevMaxBuckets];
  if (rtable && !rtable->mixed) route = nullptr;  // every bin routes to the bitmap
  const int nb = 1 << sh.lg;
  for (int i = threadIdx.x; i < nb; i += kRevBlock) s_hist[i] = 0;
  __syncthreads();
  const int64_t e_beg = (int64_t)blockIdx.x * sh.tile_arcs;
  const int64_t e_end = min(g.csr.nnz, e_beg + sh.tile_arcs);
  // two chunks per trip, their loads in flight together
  for (int64_t e = e_beg; e < e_end; e += 2 * kRevChunk) {
    int32_t u[kRevPer], v[kRevPer], du[kRevPer], u2[kRevPer], v2[kRevPer], du2[kRevPer];
    rev_load_chunk(g, route, dsat, e, e_end, u, v, du);
    rev_load_chunk(g, route, dsat, e + kRevChunk, e_end, u2, v2, du2);
#pragma unroll
    for (int q = 0; q < kRevPer; ++q) {
      if (v[q] >= 0) atomicAdd(&s_hist[rev_bucket((uint32_t)v[q], sh)], 1u);
      if (v2[q] >= 0) atomicAdd(&s_hist[rev_bucket((uint32_t)v2[q], sh)], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += kRevBlock)
    ghist[(int64_t)i * sh.tiles + blockIdx.x] = s_hist[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) ghist[(int64_t)nb * sh.tiles] = 0u;
}

// Each chunk's reversed arcs are counting-sorted by bucket in LDS and written out in runs, so the
// consecutive lanes of a store write consecutive positions of one bucket (written from the
// lanes' registers, each store touched up to 64 lines: 1.26 ms for the 32 M reversed arcs of the
// Chung-Lu 4M graph).
__global__ void __launch_bounds__(kRevBlock)
    rev_place_kernel(gary_device_graph g, const uint8_t *__restrict__ route,
                     const RouteTable *__restrict__ rtable, const uint8_t *__restrict__ dsat,
                     RevShape sh, const uint32_t *__restrict__ ghist,
                     uint64_t *__restrict__ pairs) {
  __shared__ uint64_t s_word[kRevChunk];  // the staged arcs' words, by bucket
  __shared__ uint16_t s_bk[kRevChunk];    // and their buckets
  __shared__ uint32_t s_cur[kRevMaxBuckets];   // the tile's next position per bucket
  __shared__ uint32_t s_off[kRevMaxBuckets];   // chunk: start of each bucket in s_stage
  __shared__ uint32_t s_fill[kRevMaxBuckets];  // chunk: count, then fill cursor
  __shared__ uint32_t s_wave[kRevBlock / kWave];
  if (rtable && !rtable->mixed) route = nullptr;
  const int nb = 1 << sh.lg;
  for (int i = threadIdx.x; i < nb; i += kRevBlock)
    s_cur[i] = ghist[(int64_t)i * sh.tiles + blockIdx.x];
  const int64_t e_beg = (int64_t)blockIdx.x * sh.tile_arcs;
  const int64_t e_end = min(g.csr.nnz, e_beg + sh.tile_arcs);
  for (int64_t e = e_beg; e < e_end; e += kRevChunk) {
    for (int i = threadIdx.x; i < nb; i += kRevBlock) s_fill[i] = 0;
    __syncthreads();
    int32_t u[kRevPer], v[kRevPer], du[kRevPer];
    uint32_t pu[kRevPer];  // row starts of the sources (consecutive arcs mostly share one)
    rev_load_chunk<true>(g, route, dsat, e, e_end, u, v, du, pu);
#pragma unroll
    for (int q = 0; q < kRevPer; ++q)
      if (v[q] >= 0) atomicAdd(&s_fill[rev_bucket((uint32_t)v[q], sh)], 1u);
    __syncthreads();
    // exclusive scan of the nb bucket counts (each thread owns per consecutive ones)
    const int per = (nb + kRevBlock - 1) / kRevBlock;
    const int i0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (int q = 0; q < per; ++q)
      if (i0 + q < nb) sum += s_fill[i0 + q];
    const uint32_t incl = wave_inclusive_scan_u32(sum);
    if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
    __syncthreads();
    uint32_t run = incl - sum, total = 0;
    for (int q = 0; q < kRevBlock / kWave; ++q) {
      if (q < (int)(threadIdx.x / kWave)) run += s_wave[q];
      total += s_wave[q];
    }
    for (int q = 0; q < per; ++q)
      if (i0 + q < nb) {
        const uint32_t c = s_fill[i0 + q];
        s_off[i0 + q] = run;
        s_fill[i0 + q] = run;
        run += c;
      }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRevPer; ++q)
      if (v[q] >= 0) {
        const uint32_t x = rev_hash((uint32_t)v[q], sh);
        const uint32_t b = x >> (sh.k - sh.lg);
        const uint32_t at = atomicAdd(&s_fill[b], 1u);
        s_word[at] = ((uint64_t)(x & (uint32_t)(sh.ids - 1)) << 51) | ((uint64_t)du[q] << 32) |
                     pu[q];
        s_bk[at] = (uint16_t)b;
      }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kRevBlock) {
      const uint32_t b = s_bk[j];
      pairs[s_cur[b] + (j - s_off[b])] = s_word[j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kRevBlock) s_cur[i] += s_fill[i] - s_off[i];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kRevGroupBlock)
    rev_group_kernel(gary_device_graph g, RevShape sh, const uint32_t *__restrict__ ghist,
                     const uint64_t *__restrict__ pairs, uint2 *__restrict__ rrange,
                     uint2 *__restrict__ rrows) {
  const int32_t m = g.csr.m;
  constexpr int kW = kRevGroupBlock / kWave;
  __shared__ uint32_t s_cnt[kRevMaxBucketIds];
  __shared__ uint32_t s_wave[kW];
  const uint32_t bk = blockIdx.x;
  const uint32_t beg = ghist[(int64_t)bk * sh.tiles];
  const uint32_t end = ghist[(int64_t)(bk + 1) * sh.tiles];  // the last bucket reads the total
  for (int i = threadIdx.x; i < sh.ids; i += kRevGroupBlock) s_cnt[i] = 0;
  __syncthreads();
  for (uint32_t i0 = beg + threadIdx.x; i0 < end; i0 += kRevUnroll * kRevGroupBlock) {
    uint32_t h[kRevUnroll];
#pragma unroll
    for (int q = 0; q < kRevUnroll; ++q) {
      const uint32_t i = i0 + q * kRevGroupBlock;
      h[q] = i < end ? (uint32_t)(pairs[i] >> 51) : 0xffffffffu;
    }
#pragma unroll
    for (int q = 0; q < kRevUnroll; ++q)
      if (h[q] != 0xffffffffu) atomicAdd(&s_cnt[h[q]], 1u);
  }
  __syncthreads();
  // exclusive scan of the sh.ids counters: each thread owns `per` consecutive ones
  const int per = (sh.ids + kRevGroupBlock - 1) / kRevGroupBlock;
  const int i0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (int q = 0; q < per; ++q)
    if (i0 + q < sh.ids) sum += s_cnt[i0 + q];
  const uint32_t incl = wave_inclusive_scan_u32(sum);
  if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
  __syncthreads();
  uint32_t run = beg + incl - sum;
  for (int q = 0; q < (int)(threadIdx.x / kWave); ++q) run += s_wave[q];
  const uint32_t kmask = (sh.k == 32 ? 0u : (1u << sh.k)) - 1u;
  for (int q = 0; q < per; ++q) {
    const int hq = i0 + q;
    if (hq >= sh.ids) break;
    const uint32_t c = s_cnt[hq];
    s_cnt[hq] = run;  // the list start of this id, then its fill cursor
    const uint32_t x = (bk << (sh.k - sh.lg)) | (uint32_t)hq;
    const int64_t v = (int64_t)((x * kRevMulInv) & kmask);
    if (v < m) rrange[v] = make_uint2(run, run + c);
    run += c;
  }
  __syncthreads();
  for (uint32_t i0 = beg + threadIdx.x; i0 < end; i0 += kRevUnroll * kRevGroupBlock) {
    uint64_t pr[kRevUnroll];
#pragma unroll
    for (int q = 0; q < kRevUnroll; ++q) {
      const uint32_t i = i0 + q * kRevGroupBlock;
      pr[q] = i < end ? pairs[i] : ~0ull;
    }
#pragma unroll
    for (int q = 0; q < kRevUnroll; ++q)
      if (pr[q] != ~0ull)
        rrows[atomicAdd(&s_cnt[(uint32_t)(pr[q] >> 51)], 1u)] =
            make_uint2((uint32_t)pr[q], (uint32_t)(pr[q] >> 32) & (kRevMaxSourceDeg - 1u));
  }
}

// ======================================================================================
// Triangle listing on the pivot path
// ======================================================================================
//
// The counting kernel's sets, row streaming and schedule, with every hit (v, w) of pivot u
// written out as the triple (u, v, w): hits are appended in ballot order to a per-wave LDS
// stage of kListStage triples; a full stage is copied to the wave's own chunk of global scratch
// (no atomic), and a full chunk claims one output range with a single atomic and is copied out
// with coalesced loads and stores. Claiming per stage, 1.66 M atomics on one address for the
// 424 M triangles of R-MAT 20, serialised at one L2 channel and took two thirds of the time.

constexpr int kListStage = 256;

struct ListOut {
  unsigned long long *total;  // triples found (claimed output positions)
  int32_t *out;               // 3 * cap int32
  unsigned long long cap;
  int32_t *scratch;           // 3 * chunk int32 per wave: blockIdx.x * waves + wave
  int32_t chunk;              // triples per wave chunk, >= kListStage
};

// A wave's listing state: its LDS stage and its scratch chunk (fills are wave-uniform).
struct ListWave {
  int32_t *stage;
  int32_t fill;
  int32_t *chunk;
  int32_t cfill;
};

// Claims an output range for the wave's chunk and copies it out.
__device__ __forceinline__ void list_chunk_flush(ListWave &lw, const ListOut &lo) {
  // the chunk was stored by every lane of this wave: make those stores visible to its loads
  __threadfence_block();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(lo.total, (unsigned long long)lw.cfill);
  base = __shfl(base, 0, kWave);
  for (int32_t i = lane_id(); i < 3 * lw.cfill; i += kWave)
    if (base + (unsigned long long)(i / 3) < lo.cap) lo.out[3 * base + (unsigned long long)i] = lw.chunk[i];
  lw.cfill = 0;
}

// Moves the LDS stage to the scratch chunk (flushing the chunk first when it is too full).
__device__ __forceinline__ void list_flush(ListWave &lw, const ListOut &lo) {
  wave_sync();  // the staged triples are visible to every lane
  if (lw.cfill + lw.fill > lo.chunk) list_chunk_flush(lw, lo);
  int32_t *dst = lw.chunk + 3 * lw.cfill;
  for (int32_t i = lane_id(); i < 3 * lw.fill; i += kWave) dst[i] = lw.stage[i];
  lw.cfill += lw.fill;
  lw.fill = 0;
  wave_sync();  // the stage is read before it is refilled
}

// Stages this lane's triple when `hit`; every lane of the wave calls it.
__device__ __forceinline__ void list_emit(ListWave &lw, const ListOut &lo, bool hit, int32_t u,
                                          int32_t v, int32_t w) {
  const uint64_t hits = __ballot(hit);
  if (!hits) return;
  const int32_t c = __popcll(hits);
  if (lw.fill + c > kListStage) list_flush(lw, lo);
  if (hit) {
    const int32_t at = 3 * (lw.fill + lanes_below(hits));
    lw.stage[at] = u;
    lw.stage[at + 1] = v;
    lw.stage[at + 2] = w;
  }
  lw.fill += c;
}

// stream_rows with output: the owning row's vertex travels with each chunk.
template <int K>
__device__ __forceinline__ void stream_rows_list(const gary_device_graph &g, int32_t u,
                                                 const int32_t *__restrict__ L, int32_t d,
                                                 int32_t r0, int32_t stride, const int32_t *table,
                                                 const PivotSet &ps, RowSlots &rs, int lane,
                                                 ListWave &lw, const ListOut &lo) {
  const uint32_t m = (uint32_t)g.csr.m;
  const int64_t nnz = g.csr.nnz;
  const int32_t *__restrict__ col = g.csr.colIdx;
  for (int32_t b = r0; b < d; b += stride * kWave) {
    const int32_t r = b + lane * stride;
    int32_t dv = 0, v = 0;
    int64_t pv = 0;
    if (r < d) {
      v = L[r];
      if ((uint32_t)v < m) {
        dv = g.csr.deg[v];
        pv = g.csr.indPtr[v];
      }
    }
    const int32_t nch = dv > 0 ? (int32_t)(((pv + dv + 3) >> 2) - (pv >> 2)) : 0;
    const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nch);
    const int32_t start = incl - nch;
    const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
    const int32_t cb = (int32_t)(uint32_t)(pv >> 2);
    const int32_t rlo = (int32_t)(pv & 3);
    const int32_t rhi = rlo + dv;
    for (int32_t q0 = 0; q0 < total; q0 += K * kWave) {
      int4 data[K];
      int32_t jlo[K], jhi[K], vr[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int32_t w0 = q0 + k * kWave;
        rs.s[lane] = -1;
        wave_sync();
        if (nch > 0 && start < w0 + kWave && incl > w0) rs.s[max(0, start - w0)] = lane;
        wave_sync();
        const int32_t row = max(0, wave_inclusive_max_i32(rs.s[lane]));
        const int32_t j = w0 + lane - __shfl(start, row, kWave);
        const uint32_t ch = (uint32_t)__shfl(cb, row, kWave) + (uint32_t)j;
        vr[k] = __shfl(v, row, kWave);
        jlo[k] = max(0, __shfl(rlo, row, kWave) - 4 * j);
        jhi[k] = min(4, __shfl(rhi, row, kWave) - 4 * j);
        data[k] = make_int4(0, 0, 0, 0);
        if (w0 + lane < total) {
          const int64_t e0 = 4 * (int64_t)ch;
          if (e0 + 3 < nnz) {
            data[k] = *reinterpret_cast<const int4 *>(col + e0);
          } else {
            data[k].x = e0 < nnz ? col[e0] : 0;
            data[k].y = e0 + 1 < nnz ? col[e0 + 1] : 0;
            data[k].z = e0 + 2 < nnz ? col[e0 + 2] : 0;
          }
        } else {
          jhi[k] = 0;
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool h0 = (jlo[k] <= 0 && 0 < jhi[k]) ? pivot_probe(table, ps, data[k].x) != 0 : false;
        const bool h1 = (jlo[k] <= 1 && 1 < jhi[k]) ? pivot_probe(table, ps, data[k].y) != 0 : false;
        const bool h2 = (jlo[k] <= 2 && 2 < jhi[k]) ? pivot_probe(table, ps, data[k].z) != 0 : false;
        const bool h3 = (jlo[k] <= 3 && 3 < jhi[k]) ? pivot_probe(table, ps, data[k].w) != 0 : false;
        list_emit(lw, lo, h0, u, vr[k], data[k].x);
        list_emit(lw, lo, h1, u, vr[k], data[k].y);
        list_emit(lw, lo, h2, u, vr[k], data[k].z);
        list_emit(lw, lo, h3, u, vr[k], data[k].w);
      }
    }
  }
}

// Row by row with binary search in HBM, for a pivot whose set does not fit in LDS.
__device__ __forceinline__ void fallback_list(const gary_device_graph &g, int32_t u,
                                              const int32_t *__restrict__ L, int32_t d,
                                              int32_t r0, int32_t stride, int lane,
                                              ListWave &lw, const ListOut &lo) {
  const uint32_t m = (uint32_t)g.csr.m;
  for (int32_t k = r0; k < d; k += stride) {
    const int32_t v = L[k];
    if ((uint32_t)v >= m) continue;
    const int32_t *N = g.csr.colIdx + g.csr.indPtr[v];
    const int32_t dv = g.csr.deg[v];
    for (int32_t p0 = 0; p0 < dv; p0 += kWave) {
      const int32_t p = p0 + lane;
      const int32_t w = p < dv ? N[p] : 0;
      bool hit = false;
      if (p < dv) {
        int32_t lo_i = 0, hi_i = d;
        while (lo_i < hi_i) {
          const int32_t mid = (lo_i + hi_i) >> 1;
          if (L[mid] < w) lo_i = mid + 1;
          else hi_i = mid;
        }
        hit = lo_i < d && L[lo_i] == w;
      }
      list_emit(lw, lo, hit, u, v, w);
    }
  }
}

// The schedule of tc_pivot_kernel (big pivots one workgroup each, then small pivots one wave
// each, heaviest first), listing instead of counting.
template <int WAVES, int TABLE_WORDS>
__global__ void __launch_bounds__(WAVES * 64)
    tri_list_pivot_kernel(gary_device_graph g, const int32_t *__restrict__ small,
                          const int32_t *__restrict__ big, const int32_t *__restrict__ counts,
                          int32_t *__restrict__ cursors, ListOut lo,
                          const int32_t *__restrict__ split) {
  __shared__ int32_t table[TABLE_WORDS];
  __shared__ RowSlots s_rows[WAVES];
  __shared__ int32_t staged[WAVES][3 * kListStage];
  const int lane = lane_id();
  const int wib = threadIdx.x / kWave;
  RowSlots &rb = s_rows[wib];
  ListWave lw{staged[wib], 0,
              lo.scratch + ((int64_t)blockIdx.x * WAVES + wib) * 3 * (int64_t)lo.chunk, 0};
  const int32_t n_small = counts[0], n_big = counts[1];
  if (n_small == 0 && n_big == 0) return;
  const uint32_t m = (uint32_t)g.csr.m;
  while (true) {
    if (threadIdx.x == 0) table[0] = atomicAdd(&cursors[1], 1);
    __syncthreads();
    const int32_t p = table[0];
    __syncthreads();
    if (p >= n_big) break;
    const int32_t u = big[p];
    const int32_t d = (uint32_t)u < m ? g.csr.deg[u] : 0;
    if (d >= 2) {
      const int32_t *L = g.csr.colIdx + g.csr.indPtr[u];
      PivotSet s;
      if (pivot_plan(L[0], L[d - 1], d, TABLE_WORDS, s)) {
        pivot_build<true>(table, s, L, d, threadIdx.x, WAVES * 64);
        stream_rows_list<kStreamWindows>(g, u, L, d, wib, WAVES, table, s, rb, lane, lw, lo);
      } else {
        fallback_list(g, u, L, d, wib, WAVES, lane, lw, lo);
      }
    }
    __syncthreads();
  }
  int32_t *wt = table + wib * (TABLE_WORDS / WAVES);
  const int32_t n_heavy = min(split[0], n_small);
  const int32_t heavy_claims = (n_heavy + kSmallBatchHeavy - 1) / kSmallBatchHeavy;
  while (true) {
    int32_t k = 0;
    if (lane == 0) k = atomicAdd(&cursors[0], 1);
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, kWave));
    const int32_t base = k < heavy_claims
                             ? k * kSmallBatchHeavy
                             : n_heavy + (k - heavy_claims) * kSmallBatchLight;
    if (base >= n_small) break;
    const int32_t end = min(base + (k < heavy_claims ? kSmallBatchHeavy : kSmallBatchLight),
                            k < heavy_claims ? n_heavy : n_small);
    for (int32_t p = base; p < end; ++p) {
      const int32_t u = small[p];
      if ((uint32_t)u >= m) continue;
      const int32_t d = g.csr.deg[u];
      if (d < 2) continue;
      const int32_t *L = g.csr.colIdx + g.csr.indPtr[u];
      PivotSet s;
      if (!pivot_plan(L[0], L[d - 1], d, TABLE_WORDS / WAVES, s)) {
        fallback_list(g, u, L, d, 0, 1, lane, lw, lo);
        continue;
      }
      pivot_build<false>(wt, s, L, d, lane, kWave);
      stream_rows_list<kStreamWindows>(g, u, L, d, 0, 1, wt, s, rb, lane, lw, lo);
      wave_sync();
    }
  }
  if (lw.fill > 0) list_flush(lw, lo);
  if (lw.cfill > 0) list_chunk_flush(lw, lo);
}

}  // namespace gary

// ======================================================================================
// Host dispatch (C ABI)
// ======================================================================================

namespace gary {
namespace {

struct TcCfg {
  int block = 256, grid = 0, group = 64, cache_depth = 64, partition = GARY_PARTITION_INTERLEAVED;
  int bin_small = kDefaultBinSmall, bin_hub = 2048, direction = GARY_PIVOT_AUTO;
};

constexpr int kPivotWords = 4096;  // 16 KiB: a big pivot's set, or one small-pivot set per wave
constexpr int kHubWaves = 8;
constexpr int kHubWords = 8192;    // 32 KiB hub set
constexpr int kHubGrid = 512;
constexpr int kListGrid = 2048;  // main listing launch: four waves per workgroup

// Pivot path: `block` picks the waves per workgroup of the main pivot kernel (128, 256 or 512
// threads; other values take 256).
PivotCfg pivot_cfg(const TcCfg &c) {
  PivotCfg p;
  p.bin_small = c.bin_small;
  p.bin_hub = c.bin_hub;
  p.grid = c.grid;
  p.waves = c.block == 128 ? 2 : c.block == 512 ? 8 : 4;
  p.direction = c.direction;
  return p;
}

int resolve_cfg(const gary_launch_cfg *in, TcCfg &c) {
  if (in) {
    if (in->block) c.block = in->block;
    if (in->grid) c.grid = in->grid;
    if (in->group) c.group = in->group;
    if (in->cache_depth) c.cache_depth = in->cache_depth;
    c.partition = in->partition;
    if (in->bin_small) c.bin_small = in->bin_small;
    if (in->bin_hub) c.bin_hub = in->bin_hub;
    c.direction = in->direction;
  }
  GARY_CHECK_ARG(c.block >= 64 && c.block <= 512 && c.block % 64 == 0,
                 "launch cfg: block=%d must be a multiple of 64 in [64, 512]", c.block);
  GARY_CHECK_ARG(c.group == 8 || c.group == 16 || c.group == 32 || c.group == 64,
                 "launch cfg: group=%d must be 8, 16, 32 or 64", c.group);
  GARY_CHECK_ARG(c.cache_depth >= 1 && c.cache_depth <= 4096 &&
                     (c.cache_depth & (c.cache_depth - 1)) == 0,
                 "launch cfg: cache_depth=%d must be a power of two <= 4096", c.cache_depth);
  GARY_CHECK_ARG(c.partition == GARY_PARTITION_INTERLEAVED || c.partition == GARY_PARTITION_CHUNKED ||
                     c.partition == GARY_PARTITION_DYNAMIC,
                 "launch cfg: bad partition %d", c.partition);
  GARY_CHECK_ARG(c.grid >= 0, "launch cfg: grid=%d", c.grid);
  GARY_CHECK_ARG(c.direction == GARY_PIVOT_AUTO || c.direction == GARY_PIVOT_MIN_DIRECTION ||
                     c.direction == GARY_PIVOT_FORWARD,
                 "launch cfg: bad pivot direction %d", c.direction);
  GARY_CHECK_ARG(c.bin_small >= 2 && c.bin_hub >= c.bin_small && c.bin_hub <= kPivotWords / 2,
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
  unsigned long long *cursor = nullptr;
  if (c.partition == GARY_PARTITION_DYNAMIC) {
    // the cursor lives in the stream's workspace and is cleared by a kernel (capture-safe)
    int dev = 0;
    GARY_TRY(current_device(&dev));
    void *ws = nullptr;
    GARY_TRY(workspace_get(dev, (void *)st, 256, &ws));
    cursor = (unsigned long long *)ws;
    fill_zero_kernel<<<1, 64, 0, st>>>((uint32_t *)cursor, 2);
    GARY_HIP(hipGetLastError());
  }
  GARY_TRY(timing_begin(st));
  if (c.partition == GARY_PARTITION_DYNAMIC)
    tc_edge_kernel<G, METHOD, GARY_PARTITION_DYNAMIC><<<grid, c.block, lds, st>>>(
        g, c.cache_depth, cnt, nullptr, 0, cursor);
  else if (c.partition == GARY_PARTITION_CHUNKED)
    tc_edge_kernel<G, METHOD, GARY_PARTITION_CHUNKED><<<grid, c.block, lds, st>>>(g, c.cache_depth, cnt,
                                                                                nullptr, 0);
  else
    tc_edge_kernel<G, METHOD, GARY_PARTITION_INTERLEAVED><<<grid, c.block, lds, st>>>(g, c.cache_depth, cnt,
                                                                                    nullptr, 0);
  GARY_HIP(hipGetLastError());
  GARY_TRY(timing_end(st));
  return GARY_OK;
}

template <int METHOD>
int launch_edge(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                hipStream_t st) {
  GARY_CHECK_ARG(c.block <= 256, "edge-centric strategies: block=%d exceeds 256", c.block);
  switch (c.group) {
    case 8: return launch_edge_g<8, METHOD>(g, c, cnt, st);
    case 16: return launch_edge_g<16, METHOD>(g, c, cnt, st);
    case 32: return launch_edge_g<32, METHOD>(g, c, cnt, st);
    default: return launch_edge_g<64, METHOD>(g, c, cnt, st);
  }
}

}  // namespace

bool rev_supported(int32_t m, int64_t nnz) {
  return m >= 1 && m <= (1 << 24) && nnz >= 1 && nnz < (1ll << 32);
}

bool rev_wanted(int direction, const gary_device_graph &g) {
  // the grouping lists the reversed arcs from the COO: without its rowIdx every arc streams
  // forward
  if (!rev_supported(g.csr.m, g.csr.nnz) || !g.coo.rowIdx || g.coo.nnz != g.csr.nnz) return false;
  return direction == GARY_PIVOT_MIN_DIRECTION ||
         (direction == GARY_PIVOT_AUTO && g.csr.nnz >= kRevAutoArcs);
}

namespace {

RevShape rev_shape(int32_t m, int64_t nnz) {
  RevShape r;
  int lg = 1;
  while ((1ll << lg) < (int64_t)m) ++lg;
  r.lg = std::min(11, std::max(9, lg - 13));
  r.k = std::max(lg, r.lg);
  r.ids = 1 << (r.k - r.lg);
  const int64_t chunks = (nnz + kRevChunk - 1) / kRevChunk;
  // two chunks per tile (staging 4,096 arcs keeps the placement pass at 3 workgroups per CU;
  // 8,192 arcs at one: R-MAT 20 count 2.187 -> 2.171 ms, Chung-Lu 4M 8.02 -> 7.98 ms)
  r.tiles = (int32_t)std::max<int64_t>(1, std::min<int64_t>(4096, (chunks + 1) / 2));
  r.tile_arcs = (chunks + r.tiles - 1) / r.tiles * kRevChunk;
  r.tiles = (int32_t)((nnz + r.tile_arcs - 1) / r.tile_arcs);
  return r;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t rev_bytes(int32_t m, int64_t nnz) {
  if (!rev_supported(m, nnz)) return 0;
  const RevShape sh = rev_shape(m, nnz);
  const uint32_t hist = (uint32_t)(1u << sh.lg) * (uint32_t)sh.tiles + 1u;
  return al256(sizeof(uint2) * (size_t)m) + al256(sizeof(uint32_t) * hist) +
         al256(sizeof(uint32_t) * exclusive_scan_parts(hist)) +
         al256(sizeof(uint64_t) * (size_t)nnz) + al256(sizeof(uint2) * (size_t)nnz) +
         al256((size_t)m);
}

RevBufs rev_carve(void *p, int32_t m, int64_t nnz) {
  RevBufs r;
  if (!rev_supported(m, nnz)) return r;
  r.sh = rev_shape(m, nnz);
  const uint32_t hist = (uint32_t)(1u << r.sh.lg) * (uint32_t)r.sh.tiles + 1u;
  char *c = (char *)p;
  r.rrange = (uint2 *)c;
  c += al256(sizeof(uint2) * (size_t)m);
  r.ghist = (uint32_t *)c;
  c += al256(sizeof(uint32_t) * hist);
  r.part = (uint32_t *)c;
  c += al256(sizeof(uint32_t) * exclusive_scan_parts(hist));
  r.pairs = (uint64_t *)c;
  c += al256(sizeof(uint64_t) * (size_t)nnz);
  r.rrows = (uint2 *)c;
  c += al256(sizeof(uint2) * (size_t)nnz);
  r.dsat = (uint8_t *)c;
  return r;
}

// Pivot path: bin pivots by out-degree on the device, then the hub launch (d > bin_hub, 8-wave
// workgroups with 32 KiB sets) and the main persistent launch (big pivots per workgroup, small
// pivots per wave). `lists` holds 3 * stride int32, `counters` 8 int32 (zeroed here by a kernel,
// so the sequence can be captured in a graph). With `rev`, the reversed arcs are grouped by
// target first (the counting sort above: two passes over the arcs, a scan, one pass per bucket)
// and pivots are binned by their rows as well. No host round trip, no allocation.
int launch_pivot_core(const gary_device_graph &g, const PivotCfg &pc,
                      int32_t *lists, size_t stride, int32_t *counters,
                      const int32_t *gid_of, unsigned long long *gcount, int32_t ngraphs,
                      unsigned long long *cnt, hipStream_t st, bool timed,
                      const uint8_t *route, const float *pcost, float cost_big,
                      const RouteTable *table, const RevBufs *rev) {
  const int32_t m = g.csr.m;
  if (m == 0 || g.csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(((uintptr_t)g.csr.colIdx & 15) == 0,
                 "pivot strategies read colIdx in 16-byte chunks: colIdx must be 16-byte aligned");
  GARY_CHECK_ARG(g.csr.nnz < (1ll << 34),
                 "pivot strategies index colIdx in 32-bit chunks: nnz must be below 2^34");
  // counters: [0] #small, [1] #big, [2] always 0, [3] #hub; cursors: [4..5] main, [6..7] hub;
  // then the binning histogram and scatter cursors
  fill_zero_kernel<<<1, 256, 0, st>>>((uint32_t *)counters, kPivotCounterWords);
  GARY_HIP(hipGetLastError());
  if (gcount) {
    fill_zero_kernel<<<std::max<int64_t>(1, std::min<int64_t>((ngraphs * 2 + 255) / 256, 1024)), 256,
                       0, st>>>((uint32_t *)gcount, (int64_t)ngraphs * 2);
    GARY_HIP(hipGetLastError());
  }
  const int bgrid = (int)std::min<int64_t>(((int64_t)m + 1023) / 1024, 4096);
  GARY_CHECK_ARG(pc.waves == 2 || pc.waves == 4 || pc.waves == 8,
                 "pivot path: %d waves per workgroup (2, 4 or 8)", pc.waves);
  GARY_CHECK_ARG(pc.bin_small >= 2 && pc.bin_small <= kPivotWords / pc.waves / 2 &&
                     pc.bin_hub >= pc.bin_small && pc.bin_hub <= kPivotWords / 2,
                 "pivot path: bins small=%d hub=%d out of range for %d waves", pc.bin_small,
                 pc.bin_hub, pc.waves);
  const uint2 *rrange = nullptr;
  const uint2 *rrows = nullptr;
  if (rev && rev->rrange) {
    const RevShape &sh = rev->sh;
    GARY_CHECK_ARG(g.coo.rowIdx && g.coo.nnz == g.csr.nnz,
                   "reverse rows need the COO rowIdx (call populate first)");
    GARY_CHECK_ARG(rev_supported(m, g.csr.nnz) && (1 << sh.lg) <= kRevMaxBuckets &&
                       sh.ids <= kRevMaxBucketIds && (int64_t)sh.tiles * sh.tile_arcs >= g.csr.nnz &&
                       sh.tile_arcs % kRevChunk == 0,
                   "reverse rows: m=%d nnz=%lld out of range", m, (long long)g.csr.nnz);
    const uint32_t hist = (uint32_t)(1u << sh.lg) * (uint32_t)sh.tiles + 1u;
    rev_dsat_kernel<<<(int)std::min<int64_t>(((int64_t)m + 255) / 256, 2048), 256, 0, st>>>(
        g.csr.deg, m, rev->dsat);
    GARY_HIP(hipGetLastError());
    rev_hist_kernel<<<sh.tiles, kRevBlock, 0, st>>>(g, route, table, rev->dsat, sh, rev->ghist);
    GARY_HIP(hipGetLastError());
    GARY_TRY(exclusive_scan_u32(rev->ghist, hist, rev->part, st));
    rev_place_kernel<<<sh.tiles, kRevBlock, 0, st>>>(g, route, table, rev->dsat, sh, rev->ghist,
                                                     rev->pairs);
    GARY_HIP(hipGetLastError());
    rev_group_kernel<<<1 << sh.lg, kRevGroupBlock, 0, st>>>(g, sh, rev->ghist, rev->pairs,
                                                            rev->rrange, rev->rrows);
    GARY_HIP(hipGetLastError());
    rrange = rev->rrange;
    rrows = rev->rrows;
  }
  const bool ordered = gid_of == nullptr;
  bin_hist_kernel<<<bgrid, 256, 0, st>>>(g.csr.deg, m, pc.bin_small, pc.bin_hub, counters, table,
                                         pcost, cost_big, ordered, rrange);
  GARY_HIP(hipGetLastError());
  bin_scatter_kernel<<<bgrid, 256, 0, st>>>(g.csr.deg, m, pc.bin_small, pc.bin_hub, lists,
                                            (int64_t)stride, counters, table, pcost, cost_big,
                                            ordered, rrange);
  GARY_HIP(hipGetLastError());
  // hubs: small list empty (counters[2] == 0), big list = bin 2
  tc_pivot_kernel<kHubWaves, kHubWords><<<kHubGrid, kHubWaves * 64, 0, st>>>(
      g, lists, lists + 2 * stride, counters + 2, counters + 6, cnt, gid_of,
      gcount, route, table, counters + kSplitOff + 2, rrange, rrows, counters + kStripeOff);
  GARY_HIP(hipGetLastError());
  if (timed) GARY_TRY(timing_begin(st));
  const int grid = pc.grid ? pc.grid : 2048;
  if (pc.waves == 2)
    tc_pivot_kernel<2, kPivotWords><<<grid, 2 * 64, 0, st>>>(
        g, lists, lists + stride, counters, counters + 4, cnt, gid_of, gcount, route,
        table, counters + kSplitOff, rrange, rrows, counters + kStripeOff);
  else if (pc.waves == 8)
    tc_pivot_kernel<8, kPivotWords><<<grid, 8 * 64, 0, st>>>(
        g, lists, lists + stride, counters, counters + 4, cnt, gid_of, gcount, route,
        table, counters + kSplitOff, rrange, rrows, counters + kStripeOff);
  else
    tc_pivot_kernel<4, kPivotWords><<<grid, 4 * 64, 0, st>>>(
        g, lists, lists + stride, counters, counters + 4, cnt, gid_of, gcount, route,
        table, counters + kSplitOff, rrange, rrows, counters + kStripeOff);
  GARY_HIP(hipGetLastError());
  if (timed) GARY_TRY(timing_end(st));
  return GARY_OK;
}

namespace {

int launch_pivot(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                 hipStream_t st);

}  // namespace

int launch_edge_lists(const gary_device_graph &g, const uint32_t *cached, const int32_t *n_cached,
                      const uint32_t *merge, const int32_t *n_merge, unsigned long long *cnt,
                      hipStream_t st) {
  // 2048 workgroups (eight per CU): enough groups in flight for long lists, and a short launch
  // when the table sends no arc to the edge routes
  tc_edge_list_kernel<16><<<2048, 256, (256 / 16) * 32 * sizeof(int32_t), st>>>(
      g, 32, cached, n_cached, merge, n_merge, cnt);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

namespace {

int launch_pivot(const gary_device_graph &g, const TcCfg &c, unsigned long long *cnt,
                 hipStream_t st) {
  const int32_t m = g.csr.m;
  if (m == 0 || g.csr.nnz == 0) return GARY_OK;
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t stride = ((size_t)m + 63) & ~(size_t)63;
  const size_t core = sizeof(int32_t) * (3 * stride + kPivotCounterWords);
  const bool rev = rev_wanted(c.direction, g);
  GARY_TRY(workspace_get(dev, st, core + (rev ? rev_bytes(m, g.csr.nnz) : 0), &ws));
  int32_t *lists = (int32_t *)ws;
  RevBufs rb;
  if (rev) rb = rev_carve((char *)ws + core, m, g.csr.nnz);
  return launch_pivot_core(g, pivot_cfg(c), lists, stride, lists + 3 * stride, nullptr, nullptr,
                           0, cnt, st, true, nullptr, nullptr, 0.f, nullptr, rev ? &rb : nullptr);
}

}  // namespace

int launch_tri_list(const gary_device_graph &g, unsigned long long *total, int32_t *out,
                    unsigned long long cap, hipStream_t st) {
  const int32_t m = g.csr.m;
  if (m == 0 || g.csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(((uintptr_t)g.csr.colIdx & 15) == 0,
                 "listing reads colIdx in 16-byte chunks: colIdx must be 16-byte aligned");
  GARY_CHECK_ARG(g.csr.nnz < (1ll << 34),
                 "listing indexes colIdx in 32-bit chunks: nnz must be below 2^34");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t stride = ((size_t)m + 63) & ~(size_t)63;
  GARY_TRY(workspace_get(dev, st, sizeof(int32_t) * (3 * stride + kPivotCounterWords), &ws));
  int32_t *lists = (int32_t *)ws;
  int32_t *counters = lists + 3 * stride;
  const PivotCfg pc;
  fill_zero_kernel<<<1, 256, 0, st>>>((uint32_t *)counters, kPivotCounterWords);
  GARY_HIP(hipGetLastError());
  const int bgrid = (int)std::min<int64_t>(((int64_t)m + 1023) / 1024, 4096);
  bin_hist_kernel<<<bgrid, 256, 0, st>>>(g.csr.deg, m, pc.bin_small, pc.bin_hub, counters,
                                         nullptr, nullptr, 0.f, true, nullptr);
  GARY_HIP(hipGetLastError());
  bin_scatter_kernel<<<bgrid, 256, 0, st>>>(g.csr.deg, m, pc.bin_small, pc.bin_hub, lists,
                                            (int64_t)stride, counters, nullptr, nullptr, 0.f, true,
                                            nullptr);
  GARY_HIP(hipGetLastError());
  // per-wave scratch chunks (the hub launch's waves reuse the main launch's first chunks; the
  // launches run in stream order). A chunk holds about 1/16384 of the output, 256 to 2,048
  // triples: at most 192 MiB of scratch, for one claim per chunk. On R-MAT 20 (424 M triangles)
  // a 2,048-triple cap took 8.1 ms, 8,192 10.1 ms (the chunk is gone from L2 before it is read
  // back), 1,024 8.9 ms and 512 12.9 ms (claims on one address again).
  constexpr int64_t kListWaves = std::max(kHubGrid * kHubWaves, kListGrid * 4);
  int32_t chunk = kListStage;
  while (chunk < 2048 && (unsigned long long)chunk * 16384ull < cap) chunk *= 2;
  void *scratch = nullptr;
  GARY_TRY(workspace_get(dev, (void *)11, sizeof(int32_t) * 3 * (size_t)chunk * kListWaves,
                         &scratch));
  const ListOut lo{total, out, cap, (int32_t *)scratch, chunk};
  tri_list_pivot_kernel<kHubWaves, kHubWords><<<kHubGrid, kHubWaves * 64, 0, st>>>(
      g, lists, lists + 2 * stride, counters + 2, counters + 6, lo,
      counters + kSplitOff + 2);
  GARY_HIP(hipGetLastError());
  tri_list_pivot_kernel<4, kPivotWords><<<kListGrid, 4 * 64, 0, st>>>(
      g, lists, lists + stride, counters, counters + 4, lo, counters + kSplitOff);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

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
                    strategy == GARY_TC_BSEARCH || strategy == GARY_TC_HEURISTIC ||
                    strategy == GARY_TC_AUTO_LEARNED;
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
    case GARY_TC_HEURISTIC:
    case GARY_TC_AUTO_LEARNED:
      return launch_selected(*g, strategy == GARY_TC_AUTO_LEARNED, pivot_cfg(c), cnt, st);
    case GARY_TC_BITMAP:
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
