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
// Pipeline (every step a pass over HBM; the host reads back only edge counts):
//   1. both arcs u<<vb|v and v<<vb|u of every non-loop tuple (loops become a sentinel that sorts
//      last);
//   2. stable LSD radix sort of the arcs on their 2*vb significant bits;
//   3. ordered compaction of the first arc of each run (dedup): the symmetric arc list in
//      (src, dst) order, whose row lengths are the degrees;
//   4. the mode's arcs: all of them (symmetric), an order-preserving compaction of one arc per
//      edge (oriented by (degree, id) or by id; it stays sorted), or relabeled by the rank of
//      (degree, id) and sorted again (ranked);
//   5. colIdx = dst of each arc; indPtr = exclusive scan of the row lengths, which come from the
//      first and last arc of every row.
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
#include "workspace.h"

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

// ---- Chung-Lu tuples ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) gen_chung_lu_kernel(uint64_t seed, int32_t n, uint64_t nt,
                                                              const uint64_t *__restrict__ cum,
                                                              int32_t *__restrict__ src,
                                                              int32_t *__restrict__ dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nt;
       i += (uint64_t)gridDim.x * kBlock) {
    uint64_t s = gary::stream_state(seed, i);
    src[i] = gary::chung_lu_draw(cum, n, gary::next64(s));
    dst[i] = gary::chung_lu_draw(cum, n, gary::next64(s));
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
  // counts only (no ranks): one LDS atomic add per key into the wave's histogram
#pragma unroll
  for (int k = 0; k < kIpt; ++k)
    if (base + (uint32_t)k * kWave + (uint32_t)lane < n) atomicAdd(&h[w][dig[k]], 1u);
  __syncthreads();
  for (int d = threadIdx.x; d < kBins; d += kBlock) {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) s += h[q][d];
    bh[(size_t)d * tiles + blockIdx.x] = s;
  }
}

// Stable scatter of one tile: out[offset(digit, tile) + rank within tile] = key. The tile is
// first reordered by digit in LDS, so the stores of consecutive threads go to consecutive
// addresses of a digit's run (whole cache lines) instead of 64 lanes spread over many runs.
__global__ void __launch_bounds__(kBlock) rs_scatter_kernel(const uint64_t *__restrict__ in,
                                                            uint64_t *__restrict__ out,
                                                            uint32_t n, int shift,
                                                            const uint32_t *__restrict__ off,
                                                            uint32_t tiles) {
  __shared__ uint32_t h[kWaves][kBins];
  __shared__ uint32_t dstart[kBins];  // first tile-local position of each digit
  __shared__ uint32_t gbase[kBins];   // output offset of each digit's run from this tile
  __shared__ uint32_t wtot[kWaves];
  __shared__ uint64_t sorted[kTile];
  for (int i = threadIdx.x; i < kWaves * kBins; i += kBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t tile0 = blockIdx.x * (uint32_t)kTile;
  const uint32_t base = tile0 + (uint32_t)w * kWaveSpan;
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
  // digit d (one per thread: kBlock == kBins): the waves' counts become their prefixes within
  // the digit, and a block scan of the digit totals gives the digit's first tile position
  static_assert(kBlock == kBins, "one digit per thread");
  const int d0 = threadIdx.x;
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) {
    const uint32_t c = h[q][d0];
    h[q][d0] = cnt;
    cnt += c;
  }
  const uint32_t incl = gary::wave_inclusive_scan_u32(cnt);
  if (lane == kWave - 1) wtot[w] = incl;
  gbase[d0] = off[(size_t)d0 * tiles + blockIdx.x];
  __syncthreads();
  uint32_t before = 0;
  for (int q = 0; q < w; ++q) before += wtot[q];
  dstart[d0] = before + incl - cnt;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    if (base + (uint32_t)k * kWave + (uint32_t)lane < n) {
      const uint32_t d = (uint32_t)(key[k] >> shift) & (kBins - 1);
      sorted[dstart[d] + h[w][d] + rank[k]] = key[k];
    }
  }
  __syncthreads();
  const uint32_t valid = min((uint32_t)kTile, n - tile0);
  for (uint32_t j = threadIdx.x; j < valid; j += kBlock) {
    const uint64_t k = sorted[j];
    const uint32_t d = (uint32_t)(k >> shift) & (kBins - 1);
    out[gbase[d] + (j - dstart[d])] = k;
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

// ---- keys, dedup, orientation, CSR --------------------------------------------------------------

// Both arcs u<<vb|v and v<<vb|u of each tuple at 2i, 2i+1; loops -> sentinel (all 2*vb bits set,
// never a valid arc); ids outside [0, n) raise *bad.
__global__ void __launch_bounds__(kBlock) canon_kernel(const int32_t *__restrict__ src,
                                                       const int32_t *__restrict__ dst,
                                                       uint64_t nt, int32_t n, int vb,
                                                       uint64_t sentinel,
                                                       uint64_t *__restrict__ keys,
                                                       int32_t *__restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < nt;
       i += (uint64_t)gridDim.x * kBlock) {
    const int32_t u = src[i], v = dst[i];
    uint64_t a = sentinel, b = sentinel;
    if (u < 0 || v < 0 || u >= n || v >= n) {
      *bad = 1;
    } else if (u != v) {
      a = ((uint64_t)(uint32_t)u << vb) | (uint32_t)v;
      b = ((uint64_t)(uint32_t)v << vb) | (uint32_t)u;
    }
    keys[2 * i] = a;
    keys[2 * i + 1] = b;
  }
}

// Keep rules of the ordered compactions. Each sees the sorted array, the index and the key.
struct KeepFirst {  // first key of each run of equal keys, sentinel dropped (dedup)
  uint64_t sentinel;
  __device__ bool operator()(const uint64_t *keys, uint32_t i, uint64_t k) const {
    return k != sentinel && (i == 0 || keys[i - 1] != k);
  }
};
struct KeepDegreeOrder {  // arc s -> d with (deg s, s) < (deg d, d)
  int vb;
  uint64_t mask;
  const uint32_t *deg;
  __device__ bool operator()(const uint64_t *, uint32_t, uint64_t k) const {
    const uint32_t s = (uint32_t)(k >> vb), d = (uint32_t)(k & mask);
    const uint32_t ds = deg[s], dd = deg[d];
    return ds < dd || (ds == dd && s < d);
  }
};
struct KeepIdOrder {  // arc s -> d with s > d (the reference's GraphCOO::orientation)
  int vb;
  uint64_t mask;
  __device__ bool operator()(const uint64_t *, uint32_t, uint64_t k) const {
    return (uint32_t)(k >> vb) > (uint32_t)(k & mask);
  }
};

// cnt[t] = kept keys of tile t.
template <class Keep>
__global__ void __launch_bounds__(kBlock) keep_count_kernel(const uint64_t *__restrict__ keys,
                                                            uint32_t n, Keep keep,
                                                            uint32_t *__restrict__ cnt) {
  __shared__ uint32_t ws[kWaves];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint32_t c = 0;
#pragma unroll 4
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    if (i < n) c += keep(keys, i, keys[i]);
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

// Ordered compaction of the kept keys: tile t writes from off[t].
template <class Keep>
__global__ void __launch_bounds__(kBlock) keep_write_kernel(const uint64_t *__restrict__ keys,
                                                            uint32_t n, Keep keep,
                                                            const uint32_t *__restrict__ off,
                                                            uint64_t *__restrict__ out) {
  __shared__ uint32_t ws[kWaves];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile + (uint32_t)w * kWaveSpan;
  uint64_t key[kIpt];
  uint64_t kept[kIpt];
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    key[k] = i < n ? keys[i] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    const uint32_t i = base + (uint32_t)k * kWave + (uint32_t)lane;
    kept[k] = __ballot(i < n && keep(keys, i, key[k]));
    c += (uint32_t)__popcll(kept[k]);
  }
  if (lane == 0) ws[w] = c;
  __syncthreads();
  uint32_t pos = off[blockIdx.x];
  for (int q = 0; q < w; ++q) pos += ws[q];
#pragma unroll
  for (int k = 0; k < kIpt; ++k) {
    if ((kept[k] >> lane) & 1ull) out[pos + (uint32_t)lanes_below(kept[k])] = key[k];
    pos += (uint32_t)__popcll(kept[k]);
  }
}

// Rows of arcs sorted by (src, dst): start[s] = first arc of row s, end[s] = one past its last
// (both zero for rows without arcs); colIdx[i] = dst of arc i when colIdx is given.
__global__ void __launch_bounds__(kBlock) row_bounds_kernel(const uint64_t *__restrict__ arcs,
                                                            uint32_t narcs, int vb,
                                                            uint32_t *__restrict__ start,
                                                            uint32_t *__restrict__ end,
                                                            int32_t *__restrict__ colIdx) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x; i < narcs;
       i += gridDim.x * (uint32_t)kBlock) {
    const uint64_t k = arcs[i];
    const uint32_t s = (uint32_t)(k >> vb);
    if (i == 0 || (uint32_t)(arcs[i - 1] >> vb) != s) start[s] = i;
    if (i + 1 == narcs || (uint32_t)(arcs[i + 1] >> vb) != s) end[s] = i + 1;
    if (colIdx) colIdx[i] = (int32_t)(k & mask);
  }
}

// len[v] = end[v] - start[v] in place of end, for v < n, and len[n] = 0 (the scan's total slot).
__global__ void __launch_bounds__(kBlock) row_len_kernel(const uint32_t *__restrict__ start,
                                                         uint32_t *__restrict__ end, int32_t n) {
  for (int32_t v = blockIdx.x * kBlock + threadIdx.x; v <= n; v += gridDim.x * kBlock)
    end[v] = v < n ? end[v] - start[v] : 0u;
}

// indPtr[v] = the exclusive row-length prefix, widened to 64 bits.
__global__ void __launch_bounds__(kBlock) widen_kernel(const uint32_t *__restrict__ pre,
                                                       int32_t n, int64_t *__restrict__ indPtr) {
  for (int32_t v = blockIdx.x * kBlock + threadIdx.x; v <= n; v += gridDim.x * kBlock)
    indPtr[v] = (int64_t)pre[v];
}

// Ranking keys (deg v) << vb | v of every vertex.
__global__ void __launch_bounds__(kBlock) rank_keys_kernel(const uint32_t *__restrict__ deg,
                                                           int32_t n, int vb,
                                                           uint64_t *__restrict__ keys) {
  for (int32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
    keys[v] = (uint64_t)deg[v] << vb | (uint32_t)v;
}

// newid[v] = position of (deg v, v) in ascending order.
__global__ void __launch_bounds__(kBlock) rank_ids_kernel(const uint64_t *__restrict__ sorted,
                                                          int32_t n, int vb,
                                                          int32_t *__restrict__ newid) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    newid[(uint32_t)(sorted[i] & mask)] = i;
}

// Arc s -> d becomes newid[s] -> newid[d].
__global__ void __launch_bounds__(kBlock) relabel_kernel(const uint64_t *__restrict__ arcs,
                                                         uint32_t narcs, int vb,
                                                         const int32_t *__restrict__ newid,
                                                         uint64_t *__restrict__ out) {
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  for (uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x; i < narcs;
       i += gridDim.x * (uint32_t)kBlock) {
    const uint64_t k = arcs[i];
    out[i] = (uint64_t)(uint32_t)newid[(uint32_t)(k >> vb)] << vb |
             (uint32_t)newid[(uint32_t)(k & mask)];
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

// Byte offsets of 256-byte aligned sub-buffers of one scratch allocation.
struct Carve {
  size_t bytes = 0;
  template <class T>
  size_t take(size_t count) {
    const size_t at = bytes;
    bytes += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    return at;
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

// Ordered compaction of keys[0..n) by `keep` into out; *kept receives the count (read back to the
// host, so this synchronizes the stream). cnt holds tiles_of(n) + 1 counters.
template <class Keep>
int compact(const uint64_t *keys, uint32_t n, Keep keep, uint64_t *out, uint32_t *cnt,
            uint32_t *part, hipStream_t s, uint32_t *kept) {
  *kept = 0;
  if (n == 0) return GARY_OK;
  const uint32_t tiles = tiles_of(n);
  keep_count_kernel<Keep><<<tiles, kBlock, 0, s>>>(keys, n, keep, cnt);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipMemsetAsync(cnt + tiles, 0, sizeof(uint32_t), s));
  GARY_TRY(scan_u32(cnt, tiles + 1, part, s));
  keep_write_kernel<Keep><<<tiles, kBlock, 0, s>>>(keys, n, keep, cnt, out);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipMemcpyAsync(kept, cnt + tiles, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  GARY_HIP(hipStreamSynchronize(s));
  return GARY_OK;
}

// Row lengths of arcs sorted by (src, dst) into len[0..n) (len[n] = 0); start and len hold n + 1
// counters; colIdx (optional) receives the dst of every arc.
int row_lengths(const uint64_t *arcs, uint32_t narcs, int32_t n, int vb, uint32_t *start,
                uint32_t *len, int32_t *colIdx, hipStream_t s) {
  GARY_HIP(hipMemsetAsync(start, 0, sizeof(uint32_t) * ((size_t)n + 1), s));
  GARY_HIP(hipMemsetAsync(len, 0, sizeof(uint32_t) * ((size_t)n + 1), s));
  if (narcs > 0) {
    row_bounds_kernel<<<grid_for(narcs), kBlock, 0, s>>>(arcs, narcs, vb, start, len, colIdx);
    GARY_HIP(hipGetLastError());
  }
  row_len_kernel<<<grid_for((uint64_t)n + 1), kBlock, 0, s>>>(start, len, n);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

// The CSR of arcs sorted by (src, dst): colIdx = dst, indPtr = exclusive prefix of row lengths.
int csr_from_sorted(const uint64_t *arcs, uint32_t narcs, int32_t n, int vb, uint32_t *start,
                    uint32_t *len, uint32_t *part, int64_t *indPtr, int32_t *colIdx,
                    hipStream_t s) {
  GARY_TRY(row_lengths(arcs, narcs, n, vb, start, len, colIdx, s));
  GARY_TRY(scan_u32(len, (uint32_t)n + 1, part, s));
  widen_kernel<<<grid_for((uint64_t)n + 1), kBlock, 0, s>>>(len, n, indPtr);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

// The build: the CSR of `mode` into indPtr[n+1] / colIdx (capacity cap), stored count in
// *nnz_out. When colIdx_alloc is given, it receives a hipMalloc'ed colIdx of exactly the stored
// count (at least one element), owned by the caller, and colIdx / cap are ignored.
//
// Both arcs of every tuple are sorted together, so the deduplicated list is the symmetric CSR in
// (src, dst) order and the degrees are its row lengths. The oriented modes keep one arc per edge
// by an order-preserving compaction, which stays sorted; only the ranked mode sorts again, after
// relabeling.
int build_csr(int32_t n, int64_t nt, const int32_t *src, const int32_t *dst, int mode,
              int64_t *indPtr, int32_t *colIdx, int64_t cap, int64_t *nnz_out,
              int32_t **colIdx_alloc, hipStream_t s) {
  GARY_CHECK_ARG(2 * nt < ((int64_t)1 << 31),
                 "device CSR build: %lld tuples (limit 2^30 - 1)", (long long)nt);
  const int vb = bits_for(n);
  const uint64_t mask = (vb == 0) ? 0 : ((1ull << vb) - 1);
  const uint64_t sentinel = (vb == 0) ? 0 : ((1ull << (2 * vb)) - 1);
  const uint32_t nkeys = (uint32_t)(2 * nt);
  const uint64_t cap_keys = std::max<uint64_t>(nkeys, (uint64_t)n);
  // Scratch: one library workspace slot per (device, stream), carved into the sort's two key
  // buffers and counters, so repeated builds allocate nothing.
  int device = 0;
  GARY_TRY(gary::current_device(&device));
  Carve plan;
  const size_t o_ka = plan.take<uint64_t>(cap_keys), o_kb = plan.take<uint64_t>(cap_keys);
  const size_t o_bh = plan.take<uint32_t>(std::max(sort_bh_count(cap_keys),
                                                   (size_t)tiles_of(cap_keys) + 1));
  const size_t o_part = plan.take<uint32_t>(
      std::max(sort_part_count(cap_keys), (size_t)tiles_of((uint64_t)n + 1) + 1));
  const size_t o_start = plan.take<uint32_t>((size_t)n + 1), o_len = plan.take<uint32_t>((size_t)n + 1);
  const size_t o_deg = plan.take<uint32_t>((size_t)n + 1), o_newid = plan.take<int32_t>((size_t)n);
  const size_t o_rk = plan.take<uint64_t>((size_t)n), o_flag = plan.take<int32_t>(1);
  char *ws = nullptr;
  // A scratch slot of its own per stream (the stream handle plus an odd offset: never a stream
  // handle, never one of the library's small fixed keys), so a build never reallocates the
  // workspace that the counting calls of the stream use (and that a captured graph of them
  // points into).
  GARY_TRY(gary::workspace_get(device, (char *)s + 4097, plan.bytes, (void **)&ws));
  uint64_t *ka = (uint64_t *)(ws + o_ka), *kb = (uint64_t *)(ws + o_kb);
  uint32_t *bh = (uint32_t *)(ws + o_bh), *part = (uint32_t *)(ws + o_part);
  uint32_t *start = (uint32_t *)(ws + o_start), *len = (uint32_t *)(ws + o_len);
  int32_t *flag = (int32_t *)(ws + o_flag);
  GARY_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
  // 1-3: both arcs of every tuple, sorted, deduplicated -> the symmetric arc list `sym`
  uint64_t *sym = ka, *other = kb;
  uint32_t m2 = 0;
  if (nt > 0) {
    canon_kernel<<<grid_for((uint64_t)nt), kBlock, 0, s>>>(src, dst, (uint64_t)nt, n, vb,
                                                            sentinel, ka, flag);
    GARY_HIP(hipGetLastError());
    uint64_t *sorted = nullptr;
    GARY_TRY(radix_sort(ka, kb, nkeys, 2 * vb, bh, part, s, &sorted));
    sym = sorted == ka ? kb : ka;
    GARY_TRY(compact(sorted, nkeys, KeepFirst{sentinel}, sym, bh, part, s, &m2));
    other = sorted;
    int32_t hbad = 0;
    GARY_HIP(hipMemcpyAsync(&hbad, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GARY_HIP(hipStreamSynchronize(s));
    GARY_CHECK_ARG(!hbad, "device CSR build: vertex id out of [0, n)");
  }
  const bool symmetric = mode == GARY_CSR_SYMMETRIC || mode == GARY_CSR_SYMMETRIC_RANKED;
  const int64_t need = symmetric ? (int64_t)m2 : (int64_t)m2 / 2;
  if (colIdx_alloc) {
    void *q = nullptr;
    GARY_HIP(hipMalloc(&q, sizeof(int32_t) * (size_t)std::max<int64_t>(need, 1)));
    *colIdx_alloc = (int32_t *)q;
    colIdx = (int32_t *)q;
  } else if (need > cap || (need > 0 && !colIdx)) {
    return gary::set_error(GARY_ERR_INVALID, "device CSR build: colIdx capacity %lld < %lld",
                           (long long)cap, (long long)need);
  }
  if (mode == GARY_CSR_SYMMETRIC) {
    GARY_TRY(csr_from_sorted(sym, m2, n, vb, start, len, part, indPtr, colIdx, s));
  } else if (mode == GARY_CSR_ORIENT_ID) {
    uint32_t kept = 0;
    GARY_TRY(compact(sym, m2, KeepIdOrder{vb, mask}, other, bh, part, s, &kept));
    GARY_TRY(csr_from_sorted(other, kept, n, vb, start, len, part, indPtr, colIdx, s));
  } else {
    // degrees = row lengths of the symmetric list
    uint32_t *deg = (uint32_t *)(ws + o_deg);
    GARY_TRY(row_lengths(sym, m2, n, vb, start, deg, nullptr, s));
    if (mode == GARY_CSR_ORIENT_DEGREE) {
      uint32_t kept = 0;
      GARY_TRY(compact(sym, m2, KeepDegreeOrder{vb, mask, deg}, other, bh, part, s, &kept));
      GARY_TRY(csr_from_sorted(other, kept, n, vb, start, len, part, indPtr, colIdx, s));
    } else {  // ranked: new id = rank of (deg v, v); relabel and sort again
      int32_t *newid = (int32_t *)(ws + o_newid);
      uint64_t *rk = (uint64_t *)(ws + o_rk);
      uint64_t *arcs = sym;
      if (n > 0) {
        // degrees are at most n - 1 < 2^vb, so the ranking keys have 2*vb significant bits
        rank_keys_kernel<<<grid_for((uint64_t)n), kBlock, 0, s>>>(deg, n, vb, other);
        GARY_HIP(hipGetLastError());
        uint64_t *sorted = nullptr;
        GARY_TRY(radix_sort(other, rk, (uint32_t)n, 2 * vb, bh, part, s, &sorted));
        rank_ids_kernel<<<grid_for((uint64_t)n), kBlock, 0, s>>>(sorted, n, vb, newid);
        GARY_HIP(hipGetLastError());
      }
      if (m2 > 0) {
        relabel_kernel<<<grid_for(m2), kBlock, 0, s>>>(sym, m2, vb, newid, other);
        GARY_HIP(hipGetLastError());
        GARY_TRY(radix_sort(other, sym, m2, 2 * vb, bh, part, s, &arcs));
      }
      GARY_TRY(csr_from_sorted(arcs, m2, n, vb, start, len, part, indPtr, colIdx, s));
    }
  }
  GARY_HIP(hipStreamSynchronize(s));
  *nnz_out = need;
  return GARY_OK;
}

}  // namespace

struct gary_ctx;
namespace gary {
// Exclusive scan of a[0..len) in place on `s` (values and total below 2^32); `part` holds
// exclusive_scan_parts(len) values.
uint32_t exclusive_scan_parts(uint32_t len) { return tiles_of(len) + 1; }
int exclusive_scan_u32(uint32_t *a, uint32_t len, uint32_t *part, hipStream_t s) {
  return scan_u32(a, len, part, s);
}
// The Chung-Lu cumulative weight table (host_graph.cpp).
void chung_lu_cum(int32_t n, int64_t ntuples, double gamma, double max_w,
                  std::vector<uint64_t> &cum);
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

int gary_gen_chung_lu_device(int32_t n, int64_t ntuples, double gamma, double max_w,
                             uint64_t seed, int32_t *src, int32_t *dst, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(n >= 2 && ntuples >= 0 && gamma > 1.0 && max_w >= 1.0,
                 "gary_gen_chung_lu_device: bad arguments");
  GARY_CHECK_ARG(ntuples == 0 || (src && dst), "gary_gen_chung_lu_device: null output");
  if (ntuples == 0) return GARY_OK;
  std::vector<uint64_t> cum;
  gary::chung_lu_cum(n, ntuples, gamma, max_w, cum);
  hipStream_t s = (hipStream_t)stream;
  int device = 0;
  GARY_TRY(gary::current_device(&device));
  void *ws = nullptr;  // the table, in a workspace slot of its own
  GARY_TRY(gary::workspace_get(device, (void *)10, sizeof(uint64_t) * cum.size(), &ws));
  GARY_HIP(hipMemcpyAsync(ws, cum.data(), sizeof(uint64_t) * cum.size(), hipMemcpyHostToDevice, s));
  gen_chung_lu_kernel<<<grid_for((uint64_t)ntuples, 16384), kBlock, 0, s>>>(
      seed, n, (uint64_t)ntuples, (const uint64_t *)ws, src, dst);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipStreamSynchronize(s));  // the host table is released on return
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
