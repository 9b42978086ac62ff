// graph_build.hip — graph construction on the device: R-MAT tuples, then symmetrize, drop
// self-loops, dedup and orient into the int32 CSR the counting kernels read.
//
// The reference builds its CSR on the host only: GraphCOO::symmetrize (src/Graph.cpp:19-32),
// GraphCOO::orientation (:34-61), toCSR_ (:63-91) with the OpenMP radix sort of
// common/ParRadixSort.h:218-306, and the degree array of the GraphCSR constructor
// (include/Graph.h:242-248). The device build produces exactly what gary_csr_from_edges produces
// on the host (same modes, same sorted lists, same indPtr), in a few milliseconds instead of
// seconds.
//
// Pipeline (every step a pass over HBM; no host round trip except reading the edge count):
//   1. canonical keys lo<<vb | hi of non-loop tuples (loops become a sentinel that sorts last);
//   2. stable LSD radix sort of the keys on their 2*vb significant bits;
//   3. ordered compaction of the distinct keys (dedup);
//   4. degrees: run lengths of lo within a wave plus atomics for hi;
//   5. arc keys src<<vb | dst for the requested mode (one per edge oriented by (degree, id) or by
//      id, two per edge symmetric; ranked ids come from a sort of (degree, id) keys);
//   6. radix sort of the arc keys; colIdx = low bits, indPtr[v] = first arc with src >= v.
//
// Radix sort (reduce-then-scan, 8-bit digits): per tile of 4096 keys, every wave ranks its keys
// by digit with ballots (the lanes holding a digit are found with 8 ballots, the lowest of them
// updates the wave's LDS counter), so equal digits keep input order within the tile; a scan of
// the digit-major per-tile histograms gives each (digit, tile) its output offset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "prng.h"

namespace {

using gary::kWave;
using gary::lane_id;
using gary::lanes_below;

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kIpt = 16;                 // keys per lane per tile
constexpr int kTile = kBlock * kIpt;     // 4096 keys per workgroup
constexpr int kWaveSpan = kWave * kIpt;  // 1024 consecutive keys per wave
constexpr int kDigitBits = 8;
constexpr int kBins = 1 << kDigitBits;

inline uint32_t tiles_of(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }

// Lanes of the wave holding the same digit as this lane (invalid lanes belong to no group).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t p = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kDigitBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    p &= bit ? m : ~m;
  }
  return p;
}

// ---- R-MAT tuples ------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) gen_rmat_kernel(uint64_t seed, int scale, uint64_t n,
                                                          uint64_t tA, uint64_t tAB, uint64_t tABC,
                                                          int32_t *__restrict__ src,
                                                          int32_t *__restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    uint32_t u, v;
    gary::rmat_tuple(seed, i, scale, tA, tAB, tABC, &u, &v);
    src[i] = (int32_t)u;
    dst[i] = (int32_t)v;
  }
}

// ---- radix sort --------------------------------------------------------------------------------

// bh[d * tiles + t] = number of keys of tile t whose digit at `shift` is d.
__global__ void __launch_bounds__(kBlock) rs_hist_kernel(const uint64_t *__restrict__ keys,
                                                         uint32_t n, int shift,
                                                         uint32_t *__restrict__ bh,
                                                         uint32_t tiles) {
  __shared__ uint32_t h[kWaves][kBins];
  for (int i = threadIdx.x; i < kWaves * kBins; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint32_t dig[kIpt];
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    dig[k] = i < n ? (uint32_t)(keys[i] >> shift) & (kBins - 1) : 0u;
  }
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const bool valid = base + (uint32_t)k * kWave + (uint32_t)lane < n;
    const uint64_t p = digit_peers(dig[k], valid);
    if (valid && lanes_below(p) == 0) h[w][dig[k]] += (uint32_t)__popcll(p);
    gary::wave_sync();
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kBins; d += kBlock) {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += h[q][d];
    bh[(size_t)d * tiles + blockIdx.x] = s;
  }
}

// Stable scatter of one tile: out[offset(digit, tile) + rank within tile] = key.
__global__ void __launch_bounds__(kBlock) rs_scatter_kernel(const uint64_t *__restrict__ in,
                                                            uint64_t *__restrict__ out,
                                                            uint32_t n, int shift,
                                                            const uint32_t *__restrict__ off,
                                                            uint32_t tiles) {
  __shared__ uint32_t h[kWaves][kBins];
  for (int i = threadIdx.x; i < kWaves * kBins; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint64_t key[kIpt];
  uint32_t rank[kIpt];
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    key[k] = i < n ? in[i] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const bool valid = base + (uint32_t)k * kWave + (uint32_t)lane < n;
    const uint32_t d = (uint32_t)(key[k] >> shift) & (kBins - 1);
    const uint64_t p = digit_peers(d, valid);
    uint32_t r = 0;
    if (valid) {
      r = h[w][d];
      if (lanes_below(p) == 0) h[w][d] = r + (uint32_t)__popcll(p);
      r += (uint32_t)lanes_below(p);
    }
    rank[k] = r;
    gary::wave_sync();
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kBins; d += kBlock) {
    uint32_t run = off[(size_t)d * tiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const uint32_t c = h[q][d];
      h[q][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(key[k] >> shift) & (kBins - 1);
      out[h[w][d] + rank[k]] = key[k];
    }
  }
}

// ---- exclusive scan of uint32 arrays (values and total below 2^32) ------------------------------

// Exclusive scan of the kTile values at a[0..len) held by one workgroup, in place, plus `carry`;
// returns the tile's total to every thread. Coalesced loads into LDS, 16 consecutive values per
// thread, a wave scan of the thread sums, then the wave totals.
__device__ uint32_t block_scan_tile(uint32_t *__restrict__ a, uint32_t len, uint32_t carry,
                                    uint32_t *lds /* kTile + kWaves */) {
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = (uint32_t)k * kBlock + threadIdx.x;
    lds[i] = i < len ? a[i] : 0u;
  }
  __syncthreads();
  uint32_t v[kIpt], s = 0;
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    v[k] = lds[threadIdx.x * kIpt + k];
    s += v[k];
  }
  const uint32_t incl = gary::wave_inclusive_scan_u32(s);
  const int w = threadIdx.x / kWave;
  uint32_t *wt = lds + kTile;
  __syncthreads();
  if (lane_id() == kWave - 1) wt[w] = incl;
  __syncthreads();
  uint32_t before = carry, total = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) {
    if (q < w) before += wt[q];
    total += wt[q];
  }
  uint32_t run = before + incl - s;
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    lds[threadIdx.x * kIpt + k] = run;
    run += v[k];
  }
  __syncthreads();
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = (uint32_t)k * kBlock + threadIdx.x;
    if (i < len) a[i] = lds[i];
  }
  __syncthreads();
  return total;
}

__global__ void __launch_bounds__(kBlock) scan_reduce_kernel(const uint32_t *__restrict__ a,
                                                             uint32_t len,
                                                             uint32_t *__restrict__ part) {
  __shared__ uint32_t ws[kWaves];
  const uint32_t base = blockIdx.x * (uint32_t)kTile;
  uint32_t s = 0;
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kBlock + threadIdx.x;
    if (i < len) s += a[i];
  }
  s = gary::wave_sum_u32(s);
  if (lane_id() == 0) ws[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kWaves; ++q) t += ws[q];
    part[blockIdx.x] = t;
  }
}

// One workgroup: exclusive scan of `len` values in place, tile after tile.
__global__ void __launch_bounds__(kBlock) scan_single_kernel(uint32_t *__restrict__ a,
                                                             uint32_t len) {
  __shared__ uint32_t lds[kTile + kWaves];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < len; b += kTile)
    carry += block_scan_tile(a + b, std::min<uint32_t>(kTile, len - b), carry, lds);
}

__global__ void __launch_bounds__(kBlock) scan_down_kernel(uint32_t *__restrict__ a, uint32_t len,
                                                           const uint32_t *__restrict__ part) {
  __shared__ uint32_t lds[kTile + kWaves];
  const uint32_t base = blockIdx.x * (uint32_t)kTile;
  block_scan_tile(a + base, std::min<uint32_t>(kTile, len - base), part[blockIdx.x], lds);
}

// ---- keys, dedup, degrees, arcs, CSR -------------------------------------------------------------

// key = lo << vb | hi of each tuple; loops -> sentinel (all 2*vb bits set, never a valid edge);
// ids outside [0, n) raise *bad.
__global__ void __launch_bounds__(kBlock) canon_kernel(const int32_t *__restrict__ src,
                                                       const int32_t *__restrict__ dst,
                                                       uint64_t nt, int32_t n, int vb,
                                                       uint64_t sentinel,
                                                       uint64_t *__restrict__ keys,
                                                       int32_t *__restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nt;
       i += (uint64_t)gridDim.x * kBlock) {
    const int32_t u = src[i], v = dst[i];
    uint64_t k = sentinel;
    if (u < 0 || v < 0 || u >= n || v >= n) {
      *bad = 1;
    } else if (u != v) {
      const uint32_t lo = (uint32_t)min(u, v), hi = (uint32_t)max(u, v);
      k = ((uint64_t)lo << vb) | hi;
    }
    keys[i] = k;
  }
}

// Keys of a sorted array that start a run of equal keys and are not the sentinel.
__device__ __forceinline__ bool is_first(const uint64_t *keys, uint32_t i, uint64_t k,
                                         uint64_t sentinel) {
  return k != sentinel && (i == 0 || keys[i - 1] != k);
}

// cnt[t] = distinct non-sentinel keys of tile t.
__global__ void __launch_bounds__(kBlock) uniq_count_kernel(const uint64_t *__restrict__ keys,
                                                            uint32_t n, uint64_t sentinel,
                                                            uint32_t *__restrict__ cnt) {
  __shared__ uint32_t ws[kWaves];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint32_t c = 0;
#pragma unroll 4
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    if (i < n) c += is_first(keys, i, keys[i], sentinel);
  }
  c = gary::wave_sum_u32(c);
  if (lane == 0) ws[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kWaves; ++q) t += ws[q];
    cnt[blockIdx.x] = t;
  }
}

// Ordered compaction of the distinct keys: tile t writes from off[t].
__global__ void __launch_bounds__(kBlock) uniq_write_kernel(const uint64_t *__restrict__ keys,
                                                            uint32_t n, uint64_t sentinel,
                                                            const uint32_t *__restrict__ off,
                                                            uint64_t *__restrict__ out) {
  __shared__ uint32_t ws[kWaves];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint64_t key[kIpt];
  uint64_t keep[kIpt];
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    key[k] = i < n ? keys[i] : sentinel;
  }
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    keep[k] = __ballot(i < n && is_first(keys, i, key[k], sentinel));
    c += (uint32_t)__popcll(keep[k]);
  }
  if (lane == 0) ws[w] = c;
  __syncthreads();
  uint32_t pos = off[blockIdx.x];
  for (int q = 0; q < w; ++q) pos += ws[q];
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    if ((keep[k] >> lane) & 1ull) out[pos + (uint32_t)lanes_below(keep[k])] = key[k];
    pos += (uint32_t)__popcll(keep[k]);
  }
}

// deg[v] += number of distinct edges at v. Keys are sorted by lo, so the lanes of a wave that
// share lo are consecutive: the first lane of each run adds the run length.
__global__ void __launch_bounds__(kBlock) degree_kernel(const uint64_t *__restrict__ keys,
                                                        uint32_t m, int vb,
                                                        int32_t *__restrict__ deg) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  const uint32_t stride = gridDim.x * (uint32_t)kBlock;
  for (uint32_t b = blockIdx.x * (uint32_t)kBlock + (threadIdx.x & ~(kWave - 1)); b < m;
       b += stride) {
    const uint32_t i = b + (uint32_t)lane_id();
    const bool valid = i < m;
    const uint64_t k = valid ? keys[i] : 0ull;
    const int32_t lo = valid ? (int32_t)(k >> vb) : -1;
    const int32_t prev = __shfl_up(lo, 1, kWave);
    const uint64_t heads = __ballot(valid && (lane_id() == 0 || prev != lo));
    const int nvalid = (int)__popcll(__ballot(valid));  // valid lanes are a prefix of the wave
    if (valid) {
      atomicAdd(&deg[(uint32_t)(k & mask)], 1);
      if ((heads >> lane_id()) & 1ull) {
        const uint64_t later = heads & ~((2ull << lane_id()) - 1);
        const int next = later ? (__ffsll((unsigned long long)later) - 1) : nvalid;
        atomicAdd(&deg[lo], next - lane_id());
      }
    }
  }
}

enum ArcMode { kArcDegree = 0, kArcSym = 1, kArcId = 2, kArcRanked = 3 };

// Arc keys src << vb | dst of the distinct edges for the mode (two per edge when symmetric).
__global__ void __launch_bounds__(kBlock) arcs_kernel(const uint64_t *__restrict__ ukeys,
                                                      uint32_t m, int vb, int mode,
                                                      const int32_t *__restrict__ deg,
                                                      const int32_t *__restrict__ newid,
                                                      uint64_t *__restrict__ arcs) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x; i < m;
       i += gridDim.x * (uint32_t)kBlock) {
    const uint64_t k = ukeys[i];
    uint32_t lo = (uint32_t)(k >> vb), hi = (uint32_t)(k & mask);
    if (mode == kArcDegree) {
      const int32_t dl = deg[lo], dh = deg[hi];
      const bool lo_first = dl < dh || (dl == dh && lo < hi);
      arcs[i] = lo_first ? ((uint64_t)lo << vb | hi) : ((uint64_t)hi << vb | lo);
    } else if (mode == kArcId) {  // src > dst
      arcs[i] = (uint64_t)hi << vb | lo;
    } else {
      if (mode == kArcRanked) {
        lo = (uint32_t)newid[lo];
        hi = (uint32_t)newid[hi];
      }
      arcs[2 * (size_t)i] = (uint64_t)lo << vb | hi;
      arcs[2 * (size_t)i + 1] = (uint64_t)hi << vb | lo;
    }
  }
}

// Ranking keys (deg v) << vb | v of every vertex.
__global__ void __launch_bounds__(kBlock) rank_keys_kernel(const int32_t *__restrict__ deg,
                                                           int32_t n, int vb,
                                                           uint64_t *__restrict__ keys) {
  for (int32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
    keys[v] = (uint64_t)(uint32_t)deg[v] << vb | (uint32_t)v;
}

// newid[v] = position of (deg v, v) in ascending order.
__global__ void __launch_bounds__(kBlock) rank_ids_kernel(const uint64_t *__restrict__ sorted,
                                                          int32_t n, int vb,
                                                          int32_t *__restrict__ newid) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    newid[(uint32_t)(sorted[i] & mask)] = i;
}

// colIdx[i] = dst of arc i; indPtr[v] = first arc with src >= v (indPtr[n] = narcs).
__global__ void __launch_bounds__(kBlock) csr_kernel(const uint64_t *__restrict__ arcs,
                                                     uint32_t narcs, int32_t n, int vb,
                                                     int64_t *__restrict__ indPtr,
                                                     int32_t *__restrict__ colIdx) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x; i <= narcs;
       i += gridDim.x * (uint32_t)kBlock) {
    const int32_t lo_v = i == 0 ? 0 : (int32_t)(arcs[i - 1] >> vb) + 1;
    const int32_t hi_v = i == narcs ? n : (int32_t)(arcs[i] >> vb);
    for (int32_t v = lo_v; v <= hi_v; ++v) indPtr[v] = (int64_t)i;
    if (i < narcs) colIdx[i] = (int32_t)(arcs[i] & mask);
  }
}

// ---- host side ----------------------------------------------------------------------------------

int grid_for(uint64_t items, int cap = 8192) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((items + kBlock - 1) / kBlock, (uint64_t)cap));
}

int bits_for(int64_t n) {
  int b = 0;
  while (b < 63 && (int64_t(1) << b) < n) ++b;
  return b;
}

// Device buffers released at scope exit.
struct DevBufs {
  std::vector<void *> ptrs;
  ~DevBufs() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T **p, size_t count) {
    void *q = nullptr;
    GARY_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    ptrs.push_back(q);
    *p = (T *)q;
    return GARY_OK;
  }
};

// Exclusive scan of a[0..len) on `s`; part must hold tiles_of(len) values.
int scan_u32(uint32_t *a, uint32_t len, uint32_t *part, hipStream_t s) {
  if (len == 0) return GARY_OK;
  const uint32_t nt = tiles_of(len);
  if (nt == 1) {
    scan_single_kernel<<<1, kBlock, 0, s>>>(a, len);
  } else {
    scan_reduce_kernel<<<nt, kBlock, 0, s>>>(a, len, part);
    scan_single_kernel<<<1, kBlock, 0, s>>>(part, nt);
    scan_down_kernel<<<nt, kBlock, 0, s>>>(a, len, part);
  }
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

// Stable sort of keys[0..n) on bits [0, bits); tmp holds n keys, bh 256*tiles_of(n) counters and
// part tiles_of(256*tiles_of(n)) counters. The sorted keys end in *result (keys or tmp).
int radix_sort(uint64_t *keys, uint64_t *tmp, uint32_t n, int bits, uint32_t *bh, uint32_t *part,
               hipStream_t s, uint64_t **result) {
  uint64_t *in = keys, *out = tmp;
  const uint32_t tiles = tiles_of(n);
  if (n > 1) {
    for (int shift = 0; shift < bits; shift += kDigitBits) {
      rs_hist_kernel<<<tiles, kBlock, 0, s>>>(in, n, shift, bh, tiles);
      GARY_HIP(hipGetLastError());
      GARY_TRY(scan_u32(bh, (uint32_t)kBins * tiles, part, s));
      rs_scatter_kernel<<<tiles, kBlock, 0, s>>>(in, out, n, shift, bh, tiles);
      GARY_HIP(hipGetLastError());
      std::swap(in, out);
    }
  }
  *result = in;
  return GARY_OK;
}

// Scratch sizes for radix_sort over up to n keys.
size_t sort_bh_count(uint64_t n) { return (size_t)kBins * tiles_of(n); }
size_t sort_part_count(uint64_t n) { return tiles_of(sort_bh_count(n)) + 1; }

// The build: the CSR of `mode` into indPtr[n+1] / colIdx (capacity cap), stored count in
// *nnz_out. When colIdx is null, *colIdx_alloc receives a hipMalloc'ed array of exactly the
// stored count (at least one element), owned by the caller.
int build_csr(int32_t n, int64_t nt, const int32_t *src, const int32_t *dst, int mode,
              int64_t *indPtr, int32_t *colIdx, int64_t cap, int64_t *nnz_out,
              int32_t **colIdx_alloc, hipStream_t s) {
  const bool sym = mode == GARY_CSR_SYMMETRIC || mode == GARY_CSR_SYMMETRIC_RANKED;
  GARY_CHECK_ARG((sym ? 2 * nt : nt) < ((int64_t)1 << 31),
                 "device CSR build: %lld tuples (limit 2^31 - 1 keys)", (long long)nt);
  const int vb = bits_for(n);
  GARY_CHECK_ARG(2 * vb <= 62, "device CSR build: n=%d too large", n);
  const uint64_t sentinel = (vb == 0) ? 0 : ((1ull << (2 * vb)) - 1);
  DevBufs bufs;
  // 1-3: canonical keys, sort, dedup
  int64_t m = 0;
  uint64_t *ukeys = nullptr;
  const uint64_t keycap = std::max<uint64_t>((uint64_t)nt, sym ? 2 * (uint64_t)nt : 0);
  uint64_t *ka = nullptr, *kb = nullptr;
  uint32_t *bh = nullptr, *part = nullptr;
  int32_t *flag = nullptr;
  GARY_TRY(bufs.alloc(&ka, std::max<uint64_t>(keycap, (uint64_t)n)));
  GARY_TRY(bufs.alloc(&kb, std::max<uint64_t>(keycap, (uint64_t)n)));
  const uint64_t sortmax = std::max<uint64_t>(keycap, (uint64_t)n);
  GARY_TRY(bufs.alloc(&bh, std::max(sort_bh_count(sortmax), (size_t)tiles_of(sortmax) + 1)));
  GARY_TRY(bufs.alloc(&part, sort_part_count(sortmax)));
  GARY_TRY(bufs.alloc(&flag, 2));
  GARY_HIP(hipMemsetAsync(flag, 0, 2 * sizeof(int32_t), s));
  if (nt > 0) {
    canon_kernel<<<grid_for((uint64_t)nt), kBlock, 0, s>>>(src, dst, (uint64_t)nt, n, vb,
                                                            sentinel, ka, flag);
    GARY_HIP(hipGetLastError());
    uint64_t *sorted = nullptr;
    GARY_TRY(radix_sort(ka, kb, (uint32_t)nt, 2 * vb, bh, part, s, &sorted));
    uint64_t *other = sorted == ka ? kb : ka;
    const uint32_t tiles = tiles_of((uint64_t)nt);
    uint32_t *cnt = bh;  // reuse: tiles + 1 counters
    uniq_count_kernel<<<tiles, kBlock, 0, s>>>(sorted, (uint32_t)nt, sentinel, cnt);
    GARY_HIP(hipGetLastError());
    GARY_HIP(hipMemsetAsync(cnt + tiles, 0, sizeof(uint32_t), s));
    GARY_TRY(scan_u32(cnt, tiles + 1, part, s));
    uniq_write_kernel<<<tiles, kBlock, 0, s>>>(sorted, (uint32_t)nt, sentinel, cnt, other);
    GARY_HIP(hipGetLastError());
    uint32_t hm = 0;
    int32_t hbad = 0;
    GARY_HIP(hipMemcpyAsync(&hm, cnt + tiles, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GARY_HIP(hipMemcpyAsync(&hbad, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GARY_HIP(hipStreamSynchronize(s));
    GARY_CHECK_ARG(!hbad, "device CSR build: vertex id out of [0, n)");
    m = hm;
    ukeys = other;
  }
  const int64_t need = sym ? 2 * m : m;
  if (!colIdx_alloc && (need > cap || (need > 0 && !colIdx)))
    return gary::set_error(GARY_ERR_INVALID, "device CSR build: colIdx capacity %lld < %lld",
                           (long long)cap, (long long)need);
  if (colIdx_alloc) {
    void *q = nullptr;
    GARY_HIP(hipMalloc(&q, sizeof(int32_t) * (size_t)std::max<int64_t>(need, 1)));
    *colIdx_alloc = (int32_t *)q;
    colIdx = (int32_t *)q;
  }
  // 4: degrees (the ranked and degree modes need them)
  int32_t *deg = nullptr, *newid = nullptr;
  if (mode == GARY_CSR_ORIENT_DEGREE || mode == GARY_CSR_SYMMETRIC_RANKED) {
    GARY_TRY(bufs.alloc(&deg, (size_t)n));
    GARY_HIP(hipMemsetAsync(deg, 0, sizeof(int32_t) * (size_t)std::max(n, 1), s));
    if (m > 0) {
      degree_kernel<<<grid_for((uint64_t)m), kBlock, 0, s>>>(ukeys, (uint32_t)m, vb, deg);
      GARY_HIP(hipGetLastError());
    }
  }
  uint64_t *arcbuf = ukeys == ka ? kb : ka;  // free half of the key buffers
  if (mode == GARY_CSR_SYMMETRIC_RANKED && n > 0) {
    GARY_TRY(bufs.alloc(&newid, (size_t)n));
    rank_keys_kernel<<<grid_for((uint64_t)n), kBlock, 0, s>>>(deg, n, vb, arcbuf);
    GARY_HIP(hipGetLastError());
    // ukeys is still live, so the sort's ping-pong partner is a buffer of its own; degrees are
    // at most n - 1 < 2^vb, so the keys have 2*vb significant bits.
    uint64_t *rk = nullptr;
    GARY_TRY(bufs.alloc(&rk, (size_t)n));
    uint64_t *sorted = nullptr;
    GARY_TRY(radix_sort(arcbuf, rk, (uint32_t)n, 2 * vb, bh, part, s, &sorted));
    rank_ids_kernel<<<grid_for((uint64_t)n), kBlock, 0, s>>>(sorted, n, vb, newid);
    GARY_HIP(hipGetLastError());
  }
  // 5-6: arcs, sort, CSR. Symmetric modes write 2m arcs: ukeys must not alias arcbuf's range,
  // and the sort's ping-pong partner is a fresh buffer when the arc count exceeds nt.
  const int am = mode == GARY_CSR_ORIENT_DEGREE ? kArcDegree
               : mode == GARY_CSR_ORIENT_ID     ? kArcId
               : mode == GARY_CSR_SYMMETRIC     ? kArcSym
                                                : kArcRanked;
  if (m > 0) {
    arcs_kernel<<<grid_for((uint64_t)m), kBlock, 0, s>>>(ukeys, (uint32_t)m, vb, am, deg, newid,
                                                         arcbuf);
    GARY_HIP(hipGetLastError());
  }
  uint64_t *sorted = arcbuf;
  if (need > 1) {
    // ukeys is dead once the arcs are written: its buffer is the ping-pong partner
    GARY_TRY(radix_sort(arcbuf, ukeys, (uint32_t)need, 2 * vb, bh, part, s, &sorted));
  }
  if (n > 0 || need > 0) {
    csr_kernel<<<grid_for((uint64_t)need + 1), kBlock, 0, s>>>(sorted, (uint32_t)need, n, vb,
                                                              indPtr, colIdx);
    GARY_HIP(hipGetLastError());
  } else {
    GARY_HIP(hipMemsetAsync(indPtr, 0, sizeof(int64_t), s));
  }
  GARY_HIP(hipStreamSynchronize(s));
  *nnz_out = need;
  return GARY_OK;
}

}  // namespace

struct gary_ctx;
namespace gary {
// Installs device-built CSR arrays into a context (defined with the context).
int ctx_adopt(gary_ctx *ctx, int32_t m, int64_t nnz, int64_t *indPtr, int32_t *colIdx);
int ctx_device(gary_ctx *ctx, int *device);
}  // namespace gary

extern "C" {

int gary_gen_rmat_device(int32_t scale, int64_t ntuples, double a, double b, double c,
                         uint64_t seed, int32_t *src, int32_t *dst, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(scale >= 1 && scale <= 30 && ntuples >= 0, "gary_gen_rmat_device: scale=%d",
                 scale);
  GARY_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0,
                 "gary_gen_rmat_device: bad a,b,c");
  GARY_CHECK_ARG(ntuples == 0 || (src && dst), "gary_gen_rmat_device: null output");
  if (ntuples == 0) return GARY_OK;
  const gary::RmatThresholds t = gary::rmat_thresholds(a, b, c);
  gen_rmat_kernel<<<grid_for((uint64_t)ntuples, 16384), kBlock, 0, (hipStream_t)stream>>>(
      seed, scale, (uint64_t)ntuples, t.tA, t.tAB, t.tABC, src, dst);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

int gary_csr_from_edges_device(int32_t n, int64_t ntuples, const int32_t *src,
                               const int32_t *dst, int mode, int64_t *indPtr, int32_t *colIdx,
                               int64_t cap, int64_t *nnz_out, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(n >= 0 && ntuples >= 0, "gary_csr_from_edges_device: n=%d ntuples=%lld", n,
                 (long long)ntuples);
  GARY_CHECK_ARG(mode >= GARY_CSR_ORIENT_DEGREE && mode <= GARY_CSR_SYMMETRIC_RANKED,
                 "gary_csr_from_edges_device: bad mode %d", mode);
  GARY_CHECK_ARG(indPtr && nnz_out && (ntuples == 0 || (src && dst)),
                 "gary_csr_from_edges_device: null pointer");
  return build_csr(n, ntuples, src, dst, mode, indPtr, colIdx, cap, nnz_out, nullptr,
                   (hipStream_t)stream);
  GARY_API_END
}

int gary_ctx_build_from_edges(gary_ctx *ctx, int32_t n, int64_t ntuples, const int32_t *src,
                              const int32_t *dst, int mode, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx, "gary_ctx_build_from_edges: null ctx");
  GARY_CHECK_ARG(n >= 0 && ntuples >= 0, "gary_ctx_build_from_edges: n=%d ntuples=%lld", n,
                 (long long)ntuples);
  GARY_CHECK_ARG(mode >= GARY_CSR_ORIENT_DEGREE && mode <= GARY_CSR_SYMMETRIC_RANKED,
                 "gary_ctx_build_from_edges: bad mode %d", mode);
  GARY_CHECK_ARG(ntuples == 0 || (src && dst), "gary_ctx_build_from_edges: null pointer");
  int device = 0, cur = 0;
  GARY_TRY(gary::ctx_device(ctx, &device));
  GARY_HIP(hipGetDevice(&cur));
  if (cur != device) GARY_HIP(hipSetDevice(device));
  int64_t *indPtr = nullptr;
  GARY_HIP(hipMalloc(&indPtr, sizeof(int64_t) * ((size_t)n + 1)));
  int32_t *colIdx = nullptr;
  int64_t nnz = 0;
  const int st = build_csr(n, ntuples, src, dst, mode, indPtr, nullptr, 0, &nnz, &colIdx,
                           (hipStream_t)stream);
  if (st != GARY_OK) {
    (void)hipFree(indPtr);
    if (colIdx) (void)hipFree(colIdx);
    return st;
  }
  return gary::ctx_adopt(ctx, n, nnz, indPtr, colIdx);
  GARY_API_END
}

}  // extern "C"
