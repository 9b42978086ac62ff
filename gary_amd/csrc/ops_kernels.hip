// ops_kernels.hip — search and intersection operators as standalone device operations.
//
// Reference: include/ops/Search.cuh (binary_search, build_cache, binary_search_2phase) as driven
// by the test kernels of test/SearchTest.cu:14-183, and include/ops/Intersection.cuh
// (intersectBinarySearch count / emit variants, :18-135, 239-300) as driven by
// test/SetOpTest.cu:15-43. The reference's kernels run on one warp or a handful of blocks; these
// spread a single large operation over the whole GPU (grid-stride over keys, one LDS index
// cache per wave) and keep the reference's per-key semantics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "device_search.h"
#include "device_util.h"
#include "workspace.h"

namespace gary {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Keys a thread searches in lock-step: four once there are more keys than the grid has threads
// (4 Mi keys: 72.6 -> 60.0 us plain), one below that (32 Ki keys: 6.4 against 7.5 us).
constexpr int kSearchKeys = 4;

// ---- search ------------------------------------------------------------------------------

// binarySearchKernel / binarySearch2PhaseKernel / the count variants: every workgroup builds one
// index cache of the whole list (Search.cuh:103-109) for its 256 threads, then threads
// grid-stride the keys. Phase 1 is a branchless lower bound in the cache for all K keys of a
// thread at once; phase 2 (and the plain search) halve in HBM down to kSegment candidates and
// finish with one segment read per key (bsearch_multi_seg).
template <bool TWO_PHASE, bool COUNT, int K>
__global__ void __launch_bounds__(kBlock)
    search_kernel(const int32_t *__restrict__ list, int32_t len, const int32_t *__restrict__ keys,
                  int64_t nkeys, uint8_t *__restrict__ results, int32_t n_cache,
                  unsigned long long *count) {
  extern __shared__ int32_t cache[];
  if constexpr (TWO_PHASE) {
    build_cache_group(list, len, cache, n_cache, threadIdx.x, kBlock);
    __syncthreads();
  }
  // K keys per thread and trip, searched in lock-step: key r of a trip is i0 + r * stride, so
  // the key loads and result stores of each r stay coalesced
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  uint32_t c = 0;
  for (int64_t i0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; i0 < nkeys; i0 += K * stride) {
    int32_t key[K], n[K];
    const int32_t *p[K];
    bool hit[K], f[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const int64_t i = i0 + r * stride;
      key[r] = i < nkeys ? keys[i] : 0;
      p[r] = list;
      n[r] = i < nkeys ? len : 0;
      hit[r] = false;
    }
    if constexpr (TWO_PHASE) {  // phase 1 in the LDS cache (Search.cuh:120-138)
      int32_t lb[K];
      lower_bound_pow2_multi<K>(cache, n_cache, key, lb);
#pragma unroll
      for (int r = 0; r < K; ++r) {
        // a cache hit, or pos = lb - 1: the last cache entry below the key (none: a miss)
        hit[r] = n[r] > 0 && lb[r] < n_cache && cache[lb[r]] == key[r];
        const int32_t pos = lb[r] - 1;
        if (hit[r] || pos < 0) {
          n[r] = 0;
        } else {
          const int32_t bottom = (int32_t)((int64_t)pos * len / n_cache);
          const int32_t top = (int32_t)((int64_t)(pos + 1) * len / n_cache);
          p[r] = list + bottom;
          n[r] = n[r] > 0 ? top - bottom : 0;
        }
      }
    }
    bsearch_multi_seg<K>(p, n, key, f);
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const int64_t i = i0 + r * stride;
      if (i >= nkeys) continue;
      const bool found = hit[r] || f[r];
      if constexpr (COUNT) c += found;
      else results[i] = found;
    }
  }
  if constexpr (COUNT) {
    __shared__ unsigned long long s_acc;
    block_atomic_add(count, c, &s_acc);
  }
}

// batchedBinarySearch{,2Phase}Kernel: one wave per item, INTERLEAVED over items.
template <bool TWO_PHASE>
__global__ void __launch_bounds__(kBlock)
    search_batched_kernel(int32_t batch, const int32_t *const *__restrict__ lists,
                          const int64_t *__restrict__ list_sizes,
                          const int32_t *const *__restrict__ keys,
                          const int64_t *__restrict__ key_sizes,
                          uint8_t *const *__restrict__ results, int32_t n_cache) {
  extern __shared__ int32_t smem[];
  int32_t *cache = smem + (threadIdx.x / kWave) * n_cache;
  const int lane = lane_id();
  for (int64_t it = global_wave_id(); it < batch; it += num_global_waves()) {
    const int32_t *list = lists[it];
    const int32_t len = (int32_t)list_sizes[it];
    const int32_t *k = keys[it];
    const int64_t nk = key_sizes[it];
    uint8_t *res = results[it];
    if (len <= 0) {
      for (int64_t i = lane; i < nk; i += kWave) res[i] = 0;
      continue;
    }
    if constexpr (TWO_PHASE) {
      wave_sync();
      build_cache_group(list, len, cache, n_cache, lane, kWave);
      wave_sync();
    }
    for (int64_t i = lane; i < nk; i += kWave)
      res[i] = TWO_PHASE ? bsearch_2phase(list, len, cache, n_cache, k[i]) : bsearch(list, len, k[i]);
  }
}

// ---- intersection --------------------------------------------------------------------------

enum { kIsBsearch = GARY_ISECT_BSEARCH, kIsCached = GARY_ISECT_CACHED_BSEARCH,
       kIsMerge = GARY_ISECT_MERGE, kIsBitmap = GARY_ISECT_BITMAP };

// Membership flags of the keys of A in B (A is the shorter list). For EMIT, flags[i] = 1 when
// A[i] is in B and each block writes its found-count to block_counts[blockIdx]; otherwise the
// count is accumulated into *count.
template <int METHOD, bool EMIT>
__global__ void __launch_bounds__(kBlock)
    isect_probe_kernel(const int32_t *__restrict__ A, int32_t na, const int32_t *__restrict__ B,
                       int32_t nb, int32_t n_cache, uint8_t *__restrict__ flags,
                       int32_t *__restrict__ block_counts, unsigned long long *count) {
  extern __shared__ int32_t smem[];
  __shared__ uint32_t s_wave[kWavesPerBlock];
  int32_t *cache = smem + (threadIdx.x / kWave) * n_cache;
  if constexpr (METHOD == kIsCached) {
    build_cache_group(B, nb, cache, n_cache, lane_id(), kWave);
    wave_sync();
  }
  uint32_t c = 0;
  // EMIT: contiguous per-block chunks so that block order = key order for the compaction
  const int64_t per = EMIT ? ((int64_t)na + gridDim.x - 1) / gridDim.x : 0;
  const int64_t beg = EMIT ? (int64_t)blockIdx.x * per : (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t end = EMIT ? min((int64_t)na, beg + per) : (int64_t)na;
  const int64_t step = EMIT ? kBlock : (int64_t)gridDim.x * kBlock;
  for (int64_t i = EMIT ? beg + threadIdx.x : beg; i < end; i += step) {
    const int32_t key = A[i];
    bool hit = false;
    const int32_t *p[1] = {B};
    int32_t n[1] = {nb};
    if constexpr (METHOD == kIsCached) {  // phase 1 in the LDS cache
      int32_t pos = 0;
      if (bsearch_pos(cache, n_cache, key, pos)) {
        hit = true;
        n[0] = 0;
      } else if (pos < 0) {
        n[0] = 0;
      } else {
        const int32_t bottom = (int32_t)((int64_t)pos * nb / n_cache);
        p[0] = B + bottom;
        n[0] = (int32_t)((int64_t)(pos + 1) * nb / n_cache) - bottom;
      }
    }
    const int32_t k[1] = {key};
    bool f[1];
    bsearch_multi<1>(p, n, k, f);  // branchless: no early exit, fewer divergent trips
    const bool found = hit || f[0];
    c += found;
    if constexpr (EMIT) flags[i] = found;
  }
  if constexpr (EMIT) {
    const uint32_t s = wave_sum_u32(c);
    if (lane_id() == 0) s_wave[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < kWavesPerBlock; ++w) t += s_wave[w];
      block_counts[blockIdx.x] = (int32_t)t;
    }
  } else {
    __shared__ unsigned long long s_acc;
    block_atomic_add(count, c, &s_acc);
  }
}

// Merge-path split: the number of A elements among the first k of merge(A, B) (ties take A
// first), i.e. the first i with A[i] > B[k - 1 - i]. One wave finds it in about log64 steps: the
// lanes test 64 evenly spaced candidates at once, and the search narrows to the gap after the
// last candidate that holds.
__device__ int32_t mp_split_wave(const int32_t *__restrict__ A, int32_t na,
                                 const int32_t *__restrict__ B, int32_t nb, int64_t k) {
  const int lane = lane_id();
  int64_t lo = k > nb ? k - nb : 0, hi = k < na ? k : na;  // answer in [lo, hi]
  while (hi - lo > kWave) {
    const int64_t step = (hi - lo + kWave - 1) / kWave;
    const int64_t i = lo + lane * step;
    const bool holds = i < hi && A[i] <= B[k - 1 - i];
    const uint64_t yes = __ballot(holds);
    if (!yes) {
      hi = lo;
    } else {
      const int t = 63 - __clzll((long long)yes);  // last candidate that holds
      const int64_t at = lo + t * step;
      hi = min(hi, at + step);
      lo = at + 1;
    }
  }
  const int64_t i = lo + lane;
  const bool holds = i < hi && A[i] <= B[k - 1 - i];
  return (int32_t)(lo + __popcll(__ballot(holds)));
}

constexpr int kMergePer = 16;                   // merge-path steps per thread
constexpr int kMergeTile = kBlock * kMergePer;  // per workgroup
constexpr int kMergeGridMax = 8192;            // workgroups of one merge launch

// Merge-path count in tiles: a workgroup takes kMergeTile consecutive merge-path steps, finds
// where they start and end in A and B (waves 0 and 1, one split each), loads both ranges into
// LDS with coalesced reads, and each thread merges kMergePer steps of the tile from LDS. A pair
// is counted when B[j] is taken right after an equal A element (the element before the tile's
// first A is kept in LDS for that test). Each workgroup stores its count in partial[blockIdx.x]
// (sum_partials_kernel adds them up): one device-scope fetch-add per workgroup on the caller's
// counter serialised about 9 ns each (2,048 workgroups: 9 us more than 1,024).
__global__ void __launch_bounds__(kBlock)
    isect_merge_kernel(const int32_t *__restrict__ A, int32_t na, const int32_t *__restrict__ B,
                       int32_t nb, unsigned long long *__restrict__ partial) {
  __shared__ int32_t sA[kMergeTile + 1];  // sA[0] = A[i0 - 1]
  __shared__ int32_t sB[kMergeTile];
  __shared__ int32_t split[2];
  const int64_t L = (int64_t)na + nb;
  const int w = threadIdx.x / kWave;
  uint32_t c = 0;
  for (int64_t d0 = (int64_t)blockIdx.x * kMergeTile; d0 < L;
       d0 += (int64_t)gridDim.x * kMergeTile) {
    const int64_t d1 = min(d0 + (int64_t)kMergeTile, L);
    if (w < 2) {
      const int32_t s = mp_split_wave(A, na, B, nb, w == 0 ? d0 : d1);
      if (lane_id() == 0) split[w] = s;
    }
    __syncthreads();
    const int32_t i0 = split[0], i1 = split[1];
    const int32_t j0 = (int32_t)(d0 - i0), j1 = (int32_t)(d1 - i1);
    const int32_t la = i1 - i0, lb = j1 - j0;
    // the tile's A and B ranges (la, lb <= kMergeTile): every load of the thread is issued
    // before any is stored, so the fill waits on one load latency (a loop storing each load in
    // turn waited on up to 32: 41.6 us per 4 Mi x 4 Mi count)
    int32_t va[kMergePer], vb[kMergePer];
#pragma unroll
    for (int q = 0; q < kMergePer; ++q) {
      const int32_t t = (int32_t)threadIdx.x + q * kBlock;
      va[q] = t < la ? A[i0 + t] : 0;
      vb[q] = t < lb ? B[j0 + t] : 0;
    }
#pragma unroll
    for (int q = 0; q < kMergePer; ++q) {
      const int32_t t = (int32_t)threadIdx.x + q * kBlock;
      if (t < la) sA[1 + t] = va[q];
      if (t < lb) sB[t] = vb[q];
    }
    if (threadIdx.x == 0) sA[0] = i0 > 0 ? A[i0 - 1] : 0;
    __syncthreads();
    const int32_t ll = la + lb;
    const int32_t k0 = min((int32_t)threadIdx.x * kMergePer, ll);
    const int32_t k1 = min(k0 + kMergePer, ll);
    if (k0 < k1) {
      int32_t lo = max(0, k0 - lb), hi = min(k0, la);
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sA[1 + mid] <= sB[k0 - 1 - mid]) lo = mid + 1;
        else hi = mid;
      }
      int32_t i = lo, j = k0 - lo;
      for (int32_t k = k0; k < k1; ++k) {
        if (i < la && (j >= lb || sA[1 + i] <= sB[j])) {
          ++i;
        } else {
          c += (i0 + i > 0 && sA[i] == sB[j]);
          ++j;
        }
      }
    }
    __syncthreads();  // the tile's LDS is reloaded by the next iteration
  }
  __shared__ uint64_t s_w[kBlock / kWave];
  const uint64_t s = wave_sum_u64(c);
  if (lane_id() == 0) s_w[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) tot += s_w[i];
    partial[blockIdx.x] = tot;
  }
}

// *count += the n partial counts (one workgroup).
__global__ void __launch_bounds__(kBlock)
    sum_partials_kernel(const unsigned long long *__restrict__ partial, int32_t n,
                        unsigned long long *count) {
  uint64_t v = 0;
  for (int32_t i = threadIdx.x; i < n; i += kBlock) v += partial[i];
  __shared__ unsigned long long s_acc;
  block_atomic_add(count, v, &s_acc);
}

// Bitmap count: each block owns windows of WORDS*32 consecutive ids; it marks the elements of B
// inside its window in an LDS bitmap, then tests the elements of A inside the window.
template <int WORDS>
__global__ void __launch_bounds__(kBlock)
    isect_bitmap_kernel(const int32_t *__restrict__ A, int32_t na, const int32_t *__restrict__ B,
                        int32_t nb, unsigned long long *count) {
  __shared__ uint32_t bits[WORDS];
  __shared__ int32_t s_range[4];
  const int64_t lo_id = min(A[0], B[0]);
  const int64_t hi_id = max(A[na - 1], B[nb - 1]);
  const int64_t span = hi_id - lo_id + 1;
  const int64_t nwin = (span + (int64_t)WORDS * 32 - 1) / ((int64_t)WORDS * 32);
  uint32_t c = 0;
  for (int64_t w = blockIdx.x; w < nwin; w += gridDim.x) {
    const int64_t w0 = lo_id + w * (int64_t)WORDS * 32;
    const int64_t w1 = w0 + (int64_t)WORDS * 32;  // exclusive
    for (int i = threadIdx.x; i < WORDS; i += kBlock) bits[i] = 0;
    if (threadIdx.x == 0) {
      // index ranges of B and A inside [w0, w1): lower bounds by binary search
      int32_t l = 0, r = nb;
      while (l < r) { int32_t md = (l + r) >> 1; if (B[md] < w0) l = md + 1; else r = md; }
      s_range[0] = l;
      r = nb;
      while (l < r) { int32_t md = (l + r) >> 1; if (B[md] < w1) l = md + 1; else r = md; }
      s_range[1] = l;
      l = 0; r = na;
      while (l < r) { int32_t md = (l + r) >> 1; if (A[md] < w0) l = md + 1; else r = md; }
      s_range[2] = l;
      r = na;
      while (l < r) { int32_t md = (l + r) >> 1; if (A[md] < w1) l = md + 1; else r = md; }
      s_range[3] = l;
    }
    __syncthreads();
    for (int32_t i = s_range[0] + threadIdx.x; i < s_range[1]; i += kBlock) {
      const uint32_t off = (uint32_t)(B[i] - w0);
      atomicOr(&bits[off >> 5], 1u << (off & 31));
    }
    __syncthreads();
    for (int32_t i = s_range[2] + threadIdx.x; i < s_range[3]; i += kBlock) {
      const uint32_t off = (uint32_t)(A[i] - w0);
      c += (bits[off >> 5] >> (off & 31)) & 1u;
    }
    __syncthreads();
  }
  __shared__ unsigned long long s_acc;
  block_atomic_add(count, c, &s_acc);
}

// Exclusive scan of block counts in one workgroup (nblocks <= 4096) and the total.
__global__ void __launch_bounds__(kBlock)
    scan_counts_kernel(int32_t *__restrict__ counts, int32_t n, unsigned long long *total) {
  __shared__ uint32_t s_w[kWavesPerBlock];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int32_t base = 0; base < n; base += kBlock) {
    const int32_t i = base + threadIdx.x;
    const uint32_t v = i < n ? (uint32_t)counts[i] : 0u;
    uint32_t incl = wave_inclusive_scan_u32(v);
    if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = incl;
    __syncthreads();
    uint32_t off = s_carry;
    for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) off += s_w[w];
    if (i < n) counts[i] = (int32_t)(off + incl - v);
    __syncthreads();
    if (threadIdx.x == kBlock - 1) s_carry = off + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

// Ordered compaction: block b writes its found keys starting at offsets[b], in key order.
__global__ void __launch_bounds__(kBlock)
    compact_kernel(const int32_t *__restrict__ A, int32_t na, const uint8_t *__restrict__ flags,
                   const int32_t *__restrict__ offsets, int32_t *__restrict__ out) {
  __shared__ uint32_t s_w[kWavesPerBlock];
  const int64_t per = ((int64_t)na + gridDim.x - 1) / gridDim.x;
  const int64_t beg = (int64_t)blockIdx.x * per, end = min((int64_t)na, beg + per);
  uint32_t base = (uint32_t)offsets[blockIdx.x];
  for (int64_t c0 = beg; c0 < end; c0 += kBlock) {
    const int64_t i = c0 + threadIdx.x;
    const bool f = i < end && flags[i];
    const uint64_t mask = __ballot(f);
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t off = base;
    for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) off += s_w[w];
    if (f) out[off + lanes_below(mask)] = A[i];
    uint32_t tot = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) tot += s_w[w];
    __syncthreads();
    base += tot;
  }
}

int check_search_args(int64_t len, int64_t nkeys, int mode, int cache_depth) {
  GARY_CHECK_ARG(len >= 0 && len <= INT32_MAX && nkeys >= 0, "search: len=%lld nkeys=%lld",
                 (long long)len, (long long)nkeys);
  GARY_CHECK_ARG(mode == GARY_SEARCH_PLAIN || mode == GARY_SEARCH_2PHASE, "search: bad mode %d",
                 mode);
  if (mode == GARY_SEARCH_2PHASE)
    GARY_CHECK_ARG(cache_depth >= 1 && cache_depth <= 1024 && (cache_depth & (cache_depth - 1)) == 0,
                   "search: cache_depth=%d must be a power of two in [1, 1024]", cache_depth);
  return GARY_OK;
}

int grid_for(int64_t items, int cap = 4096) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((items + kBlock - 1) / kBlock, cap));
}

// search_kernel with kSearchKeys keys per thread when the keys outnumber the grid's threads.
template <bool COUNT>
void launch_search(int grid, int mode, const int32_t *list, int64_t len, const int32_t *keys,
                   int64_t nkeys, uint8_t *results, int cache_depth, unsigned long long *cnt,
                   hipStream_t st) {
  const bool multi = nkeys > (int64_t)grid * kBlock;
  const size_t smem = (size_t)cache_depth * sizeof(int32_t);
  if (mode == GARY_SEARCH_2PHASE) {
    if (multi)
      search_kernel<true, COUNT, kSearchKeys><<<grid, kBlock, smem, st>>>(
          list, (int32_t)len, keys, nkeys, results, cache_depth, cnt);
    else
      search_kernel<true, COUNT, 1><<<grid, kBlock, smem, st>>>(list, (int32_t)len, keys, nkeys,
                                                                results, cache_depth, cnt);
  } else {
    if (multi)
      search_kernel<false, COUNT, kSearchKeys><<<grid, kBlock, 0, st>>>(
          list, (int32_t)len, keys, nkeys, results, 0, cnt);
    else
      search_kernel<false, COUNT, 1><<<grid, kBlock, 0, st>>>(list, (int32_t)len, keys, nkeys,
                                                              results, 0, cnt);
  }
}

}  // namespace
}  // namespace gary

namespace gary {
namespace {
// Zeroes n bytes: a kernel rather than a memset node, so callers may capture the operators in a
// graph.
__global__ void fill_bytes_kernel(uint8_t *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
}  // namespace
}  // namespace gary

extern "C" {

int gary_search(const int32_t *d_list, int64_t len, const int32_t *d_keys, int64_t nkeys,
                uint8_t *d_results, int mode, int cache_depth, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_TRY(check_search_args(len, nkeys, mode, cache_depth));
  if (nkeys == 0) return GARY_OK;
  GARY_CHECK_ARG(d_keys && d_results, "gary_search: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (len == 0) {
    fill_bytes_kernel<<<(int)std::min<int64_t>((nkeys + 255) / 256, 4096), 256, 0, st>>>(
        d_results, nkeys);
    GARY_HIP(hipGetLastError());
    return GARY_OK;
  }
  GARY_CHECK_ARG(d_list, "gary_search: null list");
  launch_search<false>(grid_for(nkeys), mode, d_list, len, d_keys, nkeys, d_results,
                       cache_depth, nullptr, st);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

int gary_search_count(const int32_t *d_list, int64_t len, const int32_t *d_keys, int64_t nkeys,
                      uint64_t *d_count, int mode, int cache_depth, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_TRY(check_search_args(len, nkeys, mode, cache_depth));
  GARY_CHECK_ARG(d_count, "gary_search_count: null counter");
  if (nkeys == 0 || len == 0) return GARY_OK;
  GARY_CHECK_ARG(d_list && d_keys, "gary_search_count: null pointer");
  hipStream_t st = (hipStream_t)stream;
  launch_search<true>(grid_for(nkeys), mode, d_list, len, d_keys, nkeys, nullptr, cache_depth,
                      (unsigned long long *)d_count, st);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

int gary_search_batched(int32_t batch, const int32_t *const *d_lists, const int64_t *d_list_sizes,
                        const int32_t *const *d_keys, const int64_t *d_key_sizes,
                        uint8_t *const *d_results, int mode, int cache_depth, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_TRY(check_search_args(0, 0, mode, cache_depth));
  GARY_CHECK_ARG(batch >= 0, "gary_search_batched: batch=%d", batch);
  if (batch == 0) return GARY_OK;
  GARY_CHECK_ARG(d_lists && d_list_sizes && d_keys && d_key_sizes && d_results,
                 "gary_search_batched: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int grid = (int)std::min<int64_t>((batch + kWavesPerBlock - 1) / kWavesPerBlock, 4096);
  if (mode == GARY_SEARCH_2PHASE)
    search_batched_kernel<true><<<grid, kBlock, kWavesPerBlock * cache_depth * sizeof(int32_t), st>>>(
        batch, d_lists, d_list_sizes, d_keys, d_key_sizes, d_results, cache_depth);
  else
    search_batched_kernel<false><<<grid, kBlock, 0, st>>>(batch, d_lists, d_list_sizes, d_keys,
                                                          d_key_sizes, d_results, 0);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

int gary_intersect_count(const int32_t *d_a, int64_t na, const int32_t *d_b, int64_t nb,
                         int method, int cache_depth, uint64_t *d_count, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(d_count, "gary_intersect_count: null counter");
  GARY_CHECK_ARG(na >= 0 && nb >= 0 && na <= INT32_MAX && nb <= INT32_MAX,
                 "gary_intersect_count: sizes %lld %lld", (long long)na, (long long)nb);
  GARY_CHECK_ARG(method >= GARY_ISECT_BSEARCH && method <= GARY_ISECT_BITMAP,
                 "gary_intersect_count: bad method %d", method);
  if (method == GARY_ISECT_CACHED_BSEARCH)
    GARY_TRY(check_search_args(0, 0, GARY_SEARCH_2PHASE, cache_depth));
  if (na == 0 || nb == 0) return GARY_OK;
  GARY_CHECK_ARG(d_a && d_b, "gary_intersect_count: null list");
  hipStream_t st = (hipStream_t)stream;
  auto *cnt = (unsigned long long *)d_count;
  if (method == GARY_ISECT_MERGE) {
    // one workgroup per tile (grid-striding beyond kMergeGridMax), partial counts in the
    // stream's workspace, summed by one workgroup
    int dev = 0;
    GARY_TRY(current_device(&dev));
    const int64_t tiles = (na + nb + kMergeTile - 1) / kMergeTile;
    const int grid = (int)std::min<int64_t>(tiles, kMergeGridMax);
    void *ws = nullptr;
    GARY_TRY(workspace_get(dev, st, sizeof(unsigned long long) * (size_t)grid, &ws));
    auto *partial = (unsigned long long *)ws;
    isect_merge_kernel<<<grid, kBlock, 0, st>>>(d_a, (int32_t)na, d_b, (int32_t)nb, partial);
    GARY_HIP(hipGetLastError());
    sum_partials_kernel<<<1, kBlock, 0, st>>>(partial, grid, cnt);
  } else if (method == GARY_ISECT_BITMAP) {
    isect_bitmap_kernel<4096><<<2048, kBlock, 0, st>>>(d_a, (int32_t)na, d_b, (int32_t)nb, cnt);
  } else {
    if (na > nb) { std::swap(d_a, d_b); std::swap(na, nb); }
    const int grid = grid_for(na);
    if (method == GARY_ISECT_CACHED_BSEARCH)
      isect_probe_kernel<kIsCached, false>
          <<<grid, kBlock, kWavesPerBlock * cache_depth * sizeof(int32_t), st>>>(
              d_a, (int32_t)na, d_b, (int32_t)nb, cache_depth, nullptr, nullptr, cnt);
    else
      isect_probe_kernel<kIsBsearch, false><<<grid, kBlock, 0, st>>>(
          d_a, (int32_t)na, d_b, (int32_t)nb, 0, nullptr, nullptr, cnt);
  }
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

int gary_intersect_emit(const int32_t *d_a, int64_t na, const int32_t *d_b, int64_t nb,
                        int method, int cache_depth, int32_t *d_out, uint64_t *d_count,
                        void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(d_count, "gary_intersect_emit: null counter");
  GARY_CHECK_ARG(na >= 0 && nb >= 0 && na <= INT32_MAX && nb <= INT32_MAX,
                 "gary_intersect_emit: sizes %lld %lld", (long long)na, (long long)nb);
  GARY_CHECK_ARG(method == GARY_ISECT_BSEARCH || method == GARY_ISECT_CACHED_BSEARCH,
                 "gary_intersect_emit: method %d unsupported (bsearch or cached_bsearch)", method);
  if (method == GARY_ISECT_CACHED_BSEARCH)
    GARY_TRY(check_search_args(0, 0, GARY_SEARCH_2PHASE, cache_depth));
  hipStream_t st = (hipStream_t)stream;
  if (na == 0 || nb == 0) {
    fill_bytes_kernel<<<1, 64, 0, st>>>((uint8_t *)d_count, (int64_t)sizeof(uint64_t));
    GARY_HIP(hipGetLastError());
    return GARY_OK;
  }
  GARY_CHECK_ARG(d_a && d_b && d_out, "gary_intersect_emit: null pointer");
  if (na > nb) { std::swap(d_a, d_b); std::swap(na, nb); }
  const int grid = grid_for(na, 2048);
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t flag_bytes = ((size_t)na + 255) & ~(size_t)255;
  GARY_TRY(workspace_get(dev, st, flag_bytes + sizeof(int32_t) * (size_t)grid, &ws));
  uint8_t *flags = (uint8_t *)ws;
  int32_t *bcounts = (int32_t *)((char *)ws + flag_bytes);
  const size_t lds = method == GARY_ISECT_CACHED_BSEARCH
                         ? kWavesPerBlock * cache_depth * sizeof(int32_t) : 0;
  if (method == GARY_ISECT_CACHED_BSEARCH)
    isect_probe_kernel<kIsCached, true><<<grid, kBlock, lds, st>>>(
        d_a, (int32_t)na, d_b, (int32_t)nb, cache_depth, flags, bcounts, nullptr);
  else
    isect_probe_kernel<kIsBsearch, true><<<grid, kBlock, 0, st>>>(
        d_a, (int32_t)na, d_b, (int32_t)nb, 0, flags, bcounts, nullptr);
  GARY_HIP(hipGetLastError());
  scan_counts_kernel<<<1, kBlock, 0, st>>>(bcounts, grid, (unsigned long long *)d_count);
  GARY_HIP(hipGetLastError());
  compact_kernel<<<grid, kBlock, 0, st>>>(d_a, (int32_t)na, flags, bcounts, d_out);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
