// hom4_kernels.hip — 4-vertex homomorphism counts (K4, diamond, 4-cycle) on gfx950.
//
// The reference declares graph_query::rectCountHom but never defines it
// (include/kernels/GraphQueryHom.cuh:48), and counts neither 4-cliques nor diamonds. Definitions
// (G simple, A its adjacency matrix):
//   hom(K4)      = 24 * #K4
//   hom(diamond) = Σ over arcs (a,b) of t(a,b)^2,  t(a,b) = |N(a) ∩ N(b)|
//   hom(C4)      = tr(A^4) = 8 * #C4 + 4 * Σ_v C(d(v),2) + 2m
//
// Input: a symmetric CSR whose vertex ids are ranked by (degree, id) (GARY_CSR_SYMMETRIC_RANKED),
// so the upper part of each row is the degree orientation N+(u) and the lower part N-(u).
//
// K4, triangles and per-edge triangle counts come from one pivot pass: pivot u builds the local
// graph G_u on N+(u) as bitset rows (row i = local out-neighbours of the i-th vertex of N+(u)),
// found by streaming every N+(v) through an LDS set of N+(u) that also returns element
// positions. Then T_u = #local arcs, #K4_u = Σ over local arcs (i,j) of popc(row_i & row_j), and
// every triangle (u,v,w) adds one to t(u,v), t(u,w) and t(v,w): the first two are row and column
// counts of G_u (one atomic per arc per pivot), the third one atomic per triangle.
// C4 counts wedges u-v-w with u the highest-ranked vertex (v, w < u) in per-vertex counter maps:
// #C4 = Σ_u Σ_w C(c_u(w), 2) (Chiba-Nishizeki style vertex ordering; every 4-cycle is counted
// once, at its highest vertex and the opposite vertex).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "mlp.h"
#include "tc_internal.h"
#include "workspace.h"

namespace gary {
namespace {

constexpr int kHomWaves = 4;  // c4_kernel
constexpr int kHomBlock = kHomWaves * 64;
constexpr int kSetWords = 4096;   // LDS words for the position-returning set of N+(u)
constexpr int kMaxHomDeg = 2048;  // largest out-degree handled
// Pivot passes (local graphs of N+(u)) run twice, on two lists of pivots split by out-degree:
// pivots whose local graph fits a 40 KiB LDS pool take four 512-thread workgroups per CU; the
// rest take one 1,024-thread workgroup per CU holding the whole 160 KiB (rows up to d+ = 1,053
// stay in LDS; beyond, they go to a per-workgroup slice of global scratch, where every hit is a
// global atomic). Pool sizes are in 32-bit words and leave room for the other LDS arrays of
// each kernel.
constexpr int kSmallWaves = 8;
constexpr int kBigWaves = 16;
constexpr int kLdsWords = 40960;  // 160 KiB per CU
constexpr int kHomSmallPool = kLdsWords / 4 - 1;                         // + s_item
constexpr int kHomBigPool = kLdsWords - 1;                               // + s_item
constexpr int kKcSmallPool = kLdsWords / 4 - kSmallWaves * 64 - 1;      // + stage, s_item
constexpr int kKcBigPool = kLdsWords - kBigWaves * 64 - 1;              // + stage, s_item
// N+(v) loads a lane issues before resolving them (R-MAT 22 K4 + diamond: 2 -> 78.1 ms,
// 4 -> 72.3 ms, 8 -> 70.6 ms, 16 -> 71.6 ms)
constexpr int kStreamUnroll = 8;

// LDS words beyond the set for a pivot of out-degree d: bitset rows (2 * ceil(d / 64) words
// each) and, for the 4-vertex pass, d column counters.
__host__ __device__ constexpr int64_t local_words(int32_t d, bool counters) {
  return (int64_t)d * (((d + 63) >> 6) * 2) + (counters ? d : 0);
}

// Smallest out-degree whose local graph does not fit `pool` (the first pivot of the big list).
constexpr int32_t first_big(int32_t pool, bool counters) {
  int32_t d = 1;
  while (d <= kMaxHomDeg && kSetWords + local_words(d, counters) <= pool) ++d;
  return d;
}

__device__ __forceinline__ uint32_t hslot(int32_t w, int32_t log_t) {
  return ((uint32_t)w * 0x9E3779B1u) >> (32 - log_t);
}

// Set of a sorted list L[0..d) that returns element positions.
// Bitmap mode (log_t == 0): coarse[cw] | cprefix[cw] | exact[E] | eprefix[E], E = nblocks<<(s-5):
//   position(w) = eprefix[word] + popc(exact[word] & below(w)).
// Hash mode: keys[T] | idx[T] (linear probing, empty = -1).
struct PosSet {
  int32_t base;
  uint32_t span;
  int32_t s, cwords, ewords;  // bitmap
  int32_t log_t;              // hash
  int32_t words;
};

// Words of the position bitmap at block size 2^s: 2 ceil(span / 2^(s+5)) + 2 d 2^(s-5). As for
// the triangle kernel's set, the step words(s + 1) - words(s) strictly increases with s, so the
// least words over s in [5, 20] are at the first s whose step is >= 0 (the first minimum a scan
// of every s picks), searched from the continuous optimum 2^(2s) = span / d.
__device__ __forceinline__ int64_t posset_step(uint32_t span, int32_t d, int32_t s) {
  const int64_t drop = (((int64_t)span + (1ll << (s + 5)) - 1) >> (s + 5)) -
                       (((int64_t)span + (1ll << (s + 6)) - 1) >> (s + 6));
  return 2 * ((int64_t)d << (s - 5)) - 2 * drop;
}

__device__ bool posset_plan(int32_t lo, int32_t hi, int32_t d, int32_t budget, PosSet &ps) {
  ps.base = lo;
  ps.span = (uint32_t)(hi - lo) + 1u;
  int32_t bs = ((32 - __clz((int)ps.span)) - (32 - __clz(max(d, 1))) + 1) >> 1;
  bs = min(max(bs, 5), 20);
  while (bs > 5 && posset_step(ps.span, d, bs - 1) >= 0) --bs;
  while (bs < 20 && posset_step(ps.span, d, bs) < 0) ++bs;
  const int64_t best = 2 * ((((int64_t)ps.span + (1ll << (bs + 5)) - 1) >> (bs + 5)) +
                            ((int64_t)d << (bs - 5)));
  if (best <= budget) {
    ps.s = bs;
    ps.cwords = (int32_t)(((int64_t)ps.span + (1ll << (bs + 5)) - 1) >> (bs + 5));
    ps.ewords = d << (bs - 5);
    ps.log_t = 0;
    ps.words = (int32_t)best;
    return true;
  }
  int32_t log_t = 5;
  while ((1 << log_t) < 2 * d) ++log_t;
  if (2 * (1 << log_t) <= budget) {
    ps.log_t = log_t;
    ps.s = ps.cwords = ps.ewords = 0;
    ps.words = 2 << log_t;
    return true;
  }
  return false;
}

// Workgroup-wide build (all threads of the block call it).
__device__ void posset_build(int32_t *t, const PosSet &ps, const int32_t *__restrict__ L,
                             int32_t d) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int32_t i = tid; i < ps.words; i += nthr) t[i] = ps.log_t ? -1 : 0;
  __syncthreads();
  if (ps.log_t) {
    const uint32_t mask = (1u << ps.log_t) - 1;
    int32_t *keys = t, *idx = t + (1 << ps.log_t);
    for (int32_t i = tid; i < d; i += nthr) {
      uint32_t h = hslot(L[i], ps.log_t);
      while (atomicCAS(&keys[h], -1, L[i]) != -1) h = (h + 1) & mask;
      idx[h] = i;
    }
    __syncthreads();
    return;
  }
  // coarse word k and its rank prefix at t[2k], t[2k + 1]; exact word k and its prefix at
  // t[2 cwords + 2k], t[2 cwords + 2k + 1]: a lookup level is one 8-byte LDS read
  const int32_t s = ps.s;
  uint32_t *coarse = (uint32_t *)t;
  uint32_t *exact = (uint32_t *)(t + 2 * ps.cwords);
  for (int32_t i = tid; i < d; i += nthr) {
    const uint32_t x = (uint32_t)(L[i] - ps.base);
    if (x < ps.span) atomicOr(&coarse[2 * (x >> (s + 5))], 1u << ((x >> s) & 31));
  }
  __syncthreads();
  if (tid < kWave) {
    uint32_t carry = 0;
    for (int32_t c0 = 0; c0 < ps.cwords; c0 += kWave) {
      const int32_t i = c0 + tid;
      const uint32_t v = i < ps.cwords ? (uint32_t)__popc(coarse[2 * i]) : 0u;
      const uint32_t inc = wave_inclusive_scan_u32(v);
      if (i < ps.cwords) coarse[2 * i + 1] = carry + inc - v;
      carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
    }
  }
  __syncthreads();
  for (int32_t i = tid; i < d; i += nthr) {
    const uint32_t x = (uint32_t)(L[i] - ps.base);
    if (x >= ps.span) continue;
    const uint32_t cw = x >> (s + 5), b = (x >> s) & 31;
    const uint32_t r = coarse[2 * cw + 1] + (uint32_t)__popc(coarse[2 * cw] & ((1u << b) - 1));
    atomicOr(&exact[2 * ((r << (s - 5)) + ((x & ((1u << s) - 1)) >> 5))], 1u << (x & 31));
  }
  __syncthreads();
  if (tid < kWave) {
    uint32_t carry = 0;
    for (int32_t c0 = 0; c0 < ps.ewords; c0 += kWave) {
      const int32_t i = c0 + tid;
      const uint32_t v = i < ps.ewords ? (uint32_t)__popc(exact[2 * i]) : 0u;
      const uint32_t inc = wave_inclusive_scan_u32(v);
      if (i < ps.ewords) exact[2 * i + 1] = carry + inc - v;
      carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
    }
  }
  __syncthreads();
}

// Position of w in L, or -1.
__device__ __forceinline__ int32_t posset_find(const int32_t *t, const PosSet &ps, int32_t w) {
  const uint32_t x = (uint32_t)(w - ps.base);
  if (x >= ps.span) return -1;
  if (ps.log_t == 0) {
    const int32_t s = ps.s;
    const uint32_t cw = x >> (s + 5), b = (x >> s) & 31;
    const uint2 c = reinterpret_cast<const uint2 *>(t)[cw];  // {word, rank prefix}
    if (!((c.x >> b) & 1u)) return -1;
    const uint32_t r = c.y + (uint32_t)__popc(c.x & ((1u << b) - 1));
    const uint32_t ei = (r << (s - 5)) + ((x & ((1u << s) - 1)) >> 5);
    const uint2 e = reinterpret_cast<const uint2 *>(t + 2 * ps.cwords)[ei];
    if (!((e.x >> (x & 31)) & 1u)) return -1;
    return (int32_t)(e.y + (uint32_t)__popc(e.x & ((1u << (x & 31)) - 1)));
  }
  const uint32_t mask = (1u << ps.log_t) - 1;
  uint32_t h = hslot(w, ps.log_t);
  while (true) {
    const int32_t k = t[h];
    if (k == w) return t[(1 << ps.log_t) + h];
    if (k == -1) return -1;
    h = (h + 1) & mask;
  }
}

// ---- upper / lower split of the ranked symmetric CSR -------------------------------------------

// cut[u] = |{w in N(u) : w < u}|; up_start[u] = indPtr[u] + cut[u]; up_len[u] = deg[u] - cut[u].
// Flags bad[0] if ids are not in non-decreasing degree order or a row holds u itself.
// Also sums up_len - lo_len into *asym: zero for a symmetric graph.
__global__ void hom_split_kernel(gary_device_graph g, int64_t *__restrict__ up_start,
                                 int32_t *__restrict__ up_len, int32_t *__restrict__ lo_len,
                                 int32_t *__restrict__ bad, unsigned long long *asym) {
  const int32_t m = g.csr.m;
  int64_t diff = 0;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = g.csr.indPtr[u];
    const int32_t d = g.csr.deg[u];
    const int32_t *N = g.csr.colIdx + p;
    int32_t l = 0, r = d;
    while (l < r) {
      const int32_t mid = (l + r) >> 1;
      if (N[mid] < (int32_t)u) l = mid + 1;
      else r = mid;
    }
    if (l < d && N[l] == (int32_t)u) *bad = 1;
    if (u + 1 < m && g.csr.deg[u + 1] < d) *bad = 1;
    up_start[u] = p + l;
    up_len[u] = d - l;
    lo_len[u] = l;
    diff += (int64_t)(d - l) - l;
  }
  const uint64_t s = wave_sum_u64((uint64_t)diff);
  if (lane_id() == 0 && s) atomicAdd(asym, (unsigned long long)s);
}

// list[0 .. *count) = the vertices u with lo <= val[u] < hi, in any order (per-wave runs
// ascend); *max_val (when given) = the largest such val[u]. Builds the pivot lists of the local
// graph passes (val = out-degree) and the vertex classes of the 4-cycle pass (val = lower
// degree).
__global__ void range_list_kernel(const int32_t *__restrict__ val, int32_t m, int32_t lo,
                                  int32_t hi, int32_t *__restrict__ list,
                                  int32_t *__restrict__ count, int32_t *__restrict__ max_val) {
  const int lane = lane_id();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < m;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = base + lane;
    const int32_t x = u < m ? val[u] : 0;
    const bool in = u < m && x >= lo && x < hi;
    const uint64_t mask = __ballot(in);
    if (!mask) continue;
    int32_t pos = 0;
    if (lane == 0) pos = atomicAdd(count, __popcll(mask));
    pos = __shfl(pos, 0, kWave);
    if (in) {
      list[pos + lanes_below(mask)] = (int32_t)u;
      if (max_val) atomicMax(max_val, x);
    }
  }
}

// ---- pivot pass: triangles, K4, per-edge triangle counts ---------------------------------------

struct HomTotals {
  unsigned long long tri, k4;
};

// One workgroup per pivot, claimed from the list `items` (count *nitems; the big list from its
// end, where the highest-ranked and densest pivots are). LDS pool: the position set, then d
// column counters, then the local bitset rows when they fit, else this block's slice of
// `row_scratch` (global memory).
template <int WAVES, int POOL, bool BIG>
__global__ void __launch_bounds__(WAVES * kWave)
    hom_local_kernel(gary_device_graph g, const int64_t *__restrict__ up_start,
                     const int32_t *__restrict__ up_len, const int32_t *__restrict__ items,
                     const int32_t *__restrict__ nitems, uint32_t *__restrict__ tcount,
                     int32_t *__restrict__ cursor, uint32_t *__restrict__ row_scratch,
                     int64_t scratch_words, HomTotals *tot, int32_t *__restrict__ unsupported) {
  constexpr int kBlock = WAVES * kWave;
  __shared__ __align__(8) int32_t pool[POOL];
  __shared__ int32_t s_item;
  int32_t *const set = pool;
  int32_t *const colcnt = pool + kSetWords;
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  const int32_t m = g.csr.m;
  const int32_t *__restrict__ col = g.csr.colIdx;
  const int32_t n_items = *nitems;
  uint64_t tri = 0, k4 = 0;
  // thread 0 claims kClaim consecutive items per atomic and hands them out one at a time
  // through s_item: four small pivots per claim on the one cursor (R-MAT 22 K4 + diamond
  // 83 -> 81 ms); big pivots one at a time, heaviest first.
  constexpr int32_t kClaim = BIG ? 1 : 4;
  int32_t next = 0, end = 0;  // thread 0's claimed range
  while (true) {
    if (threadIdx.x == 0) {
      if (next >= end) {
        next = atomicAdd(cursor, kClaim);
        end = next + kClaim;
      }
      s_item = next++;
    }
    __syncthreads();
    const int32_t it = s_item;
    if (it >= n_items) break;
    const int32_t u = items[BIG ? n_items - 1 - it : it];
    const int32_t d = ((uint32_t)u < (uint32_t)m) ? up_len[u] : 0;
    if (d < 2) {
      __syncthreads();
      continue;
    }
    const int64_t ubase = up_start[u];
    const int32_t *L = col + ubase;
    PosSet ps;
    const int32_t W2 = ((d + 63) >> 6) * 2;  // 32-bit words per local row (<= 64)
    const int64_t row_words = (int64_t)d * W2;
    const int64_t lds_rows = (int64_t)POOL - kSetWords - d;  // LDS words left for the rows
    if (d > kMaxHomDeg || lds_rows < 0 || !posset_plan(L[0], L[d - 1], d, kSetWords, ps) ||
        (row_words > lds_rows && row_words > scratch_words)) {
      if (threadIdx.x == 0) *unsupported = d;
      __syncthreads();
      continue;
    }
    uint32_t *rows = row_words <= lds_rows ? (uint32_t *)(pool + kSetWords + d)
                                           : row_scratch + (int64_t)blockIdx.x * scratch_words;
    for (int64_t i = threadIdx.x; i < row_words; i += kBlock) rows[i] = 0;
    for (int32_t i = threadIdx.x; i < d; i += kBlock) colcnt[i] = 0;
    posset_build(set, ps, L, d);  // begins and ends with a workgroup barrier
    // stream the rows: wave wib takes local rows i = wib, wib + WAVES, ...; a lane issues
    // kStreamUnroll loads of N+(v) before looking them up. (Streaming the rows as flattened
    // 32-byte chunks, as the triangle kernel does, cut this pass to 67.9 ms per R-MAT 22 K4 +
    // diamond call without the t(v, w) atomics, but a chunk per lane spread each atomic
    // instruction over many rows' lines: 113.8 ms with them, 90.2 ms transposed to 8 chunks per
    // instruction, against 81.3 ms here, where a wave's lanes hit one row. Two rows streamed by
    // one wave as a single run, which fills the lanes a short row leaves idle, changed nothing:
    // 78.9 against 77.5 ms on R-MAT 22, 14.9 against 15.0 ms on R-MAT 20.)
    // The headers (start, length of N+(v)) of a wave's next 64 rows are loaded lane-parallel,
    // one row a lane, and handed out by shuffles, so a row's loads wait on one round trip to
    // memory instead of three dependent ones (R-MAT 22 K4 + diamond 77.7 -> 72.6 ms).
    for (int32_t r0 = wib; r0 < d; r0 += WAVES * kWave) {
      const int32_t ih = r0 + lane * WAVES;
      int64_t hb = 0;
      int32_t hl = 0;
      if (ih < d) {
        const int32_t v = L[ih];
        if ((uint32_t)v < (uint32_t)m) {
          hb = up_start[v];
          hl = up_len[v];
        }
      }
      const int32_t nr = min(kWave, (d - r0 + WAVES - 1) / WAVES);
      for (int32_t q = 0; q < nr; ++q) {
        const int32_t i = r0 + q * WAVES;
        const int32_t dv = __shfl(hl, q, kWave);
        if (dv == 0) continue;
        const uint32_t vb_hi = (uint32_t)__shfl((int32_t)(hb >> 32), q, kWave);
        const uint32_t vb_lo = (uint32_t)__shfl((int32_t)hb, q, kWave);
        const int64_t vb = (int64_t)(((uint64_t)vb_hi << 32) | vb_lo);
        for (int32_t p0 = lane; p0 < dv; p0 += kStreamUnroll * kWave) {
          int32_t w[kStreamUnroll];
#pragma unroll
          for (int k = 0; k < kStreamUnroll; ++k) {
            const int32_t p = p0 + k * kWave;
            w[k] = p < dv ? col[vb + p] : -1;  // -1 lies outside every set
          }
#pragma unroll
          for (int k = 0; k < kStreamUnroll; ++k) {
            const int32_t j = posset_find(set, ps, w[k]);
            if (j >= 0) {
              atomicOr(&rows[(int64_t)i * W2 + (j >> 5)], 1u << (j & 31));
              atomicAdd(&colcnt[j], 1);
              atomicAdd(&tcount[vb + p0 + k * kWave], 1u);  // t(v, w) from pivot u
            }
          }
        }
      }
    }
    __syncthreads();
    // row counts (t(u, v_i)) + column counts (t(u, v_j)); triangles of u = Σ row counts
    for (int32_t i = threadIdx.x; i < d; i += kBlock) {
      uint32_t rc = 0;
      for (int32_t w = 0; w < W2; ++w) rc += (uint32_t)__popc(rows[(int64_t)i * W2 + w]);
      tri += rc;
      const uint32_t add = rc + (uint32_t)colcnt[i];
      if (add) atomicAdd(&tcount[ubase + i], add);
    }
    // K4: for every local arc (i, j): popc(row_i & row_j). A wave takes row i; lane l holds its
    // 32-bit word l, and the row's set bits are enumerated 64 at a time: arc a belongs to the
    // last word whose exclusive bit prefix is <= a (a search over lanes by shuffles), and is the
    // (a - prefix)-th set bit of that word. Row j only has bits above j, so the AND starts at the
    // word holding bit j + 1.
    for (int32_t i = wib; i < d; i += WAVES) {
      const uint32_t *ri = rows + (int64_t)i * W2;
      const uint32_t word = lane < W2 ? ri[lane] : 0u;
      const int32_t nb = __popc(word);
      const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nb);
      const int32_t excl = incl - nb;
      const int32_t arcs = __builtin_amdgcn_readlane(incl, kWave - 1);
      // words of row i past its last nonzero one add nothing to popc(row_i & row_j)
      const int32_t wend = 64 - (int32_t)__clzll((unsigned long long)__ballot(word != 0u));
      for (int32_t a0 = 0; a0 < arcs; a0 += kWave) {
        const int32_t a = a0 + lane;
        int32_t o = 0;  // last word with excl <= a
#pragma unroll
        for (int32_t step = kWave / 2; step > 0; step >>= 1)
          if (__shfl(excl, o + step, kWave) <= a) o += step;
        uint32_t wv = (uint32_t)__shfl((int32_t)word, o, kWave);
        int32_t k = a - __shfl(excl, o, kWave);
        if (a >= arcs) continue;
        int32_t pos = 0;  // position of the k-th set bit of wv
#pragma unroll
        for (int32_t width = 16; width > 0; width >>= 1) {
          const int32_t c = __popc(wv & ((1u << width) - 1u));
          if (k >= c) {
            k -= c;
            wv >>= width;
            pos += width;
          }
        }
        const int32_t j = o * 32 + pos;
        const uint32_t *rj = rows + (int64_t)j * W2;
        uint32_t c = 0;
        for (int32_t w = (j + 1) >> 5; w < wend; ++w) c += (uint32_t)__popc(ri[w] & rj[w]);
        k4 += c;
      }
    }
    __syncthreads();
  }
  const uint64_t st = wave_sum_u64(tri), sk = wave_sum_u64(k4);
  if (lane == 0) {
    if (st) atomicAdd(&tot->tri, (unsigned long long)st);
    if (sk) atomicAdd(&tot->k4, (unsigned long long)sk);
  }
}

// 128-bit accumulation into {lo, hi} from 64-bit wave partials.
__device__ __forceinline__ void add_u128(unsigned long long *acc, uint64_t v) {
  if (!v) return;
  const unsigned long long old = atomicAdd(&acc[0], (unsigned long long)v);
  if (old + v < old) atomicAdd(&acc[1], 1ull);
}

// Σ over upper arcs of t^2 (128-bit), and Σ_v C(d(v), 2) (128-bit).
__global__ void hom_sums_kernel(gary_device_graph g, const int64_t *__restrict__ up_start,
                                const int32_t *__restrict__ up_len,
                                const uint32_t *__restrict__ tcount,
                                unsigned long long *sum_t2, unsigned long long *sum_c2) {
  const int32_t m = g.csr.m;
  const int lane = lane_id();
  uint64_t t2 = 0, c2 = 0;
  for (int64_t u = global_wave_id(); u < m; u += num_global_waves()) {
    const int64_t b = up_start[u];
    const int32_t d = up_len[u];
    for (int32_t i = lane; i < d; i += kWave) {
      const uint64_t t = tcount[b + i];
      t2 += t * t;
    }
    if (lane == 0) {
      const uint64_t dd = (uint64_t)g.csr.deg[u];
      c2 += dd * (dd - (dd > 0)) / 2;
    }
    // flush before a lane partial could wrap (t < 2^32, so one row adds < 2^64 / 64 per lane)
    if ((u & 1023) == 1023) {
      add_u128(sum_t2, wave_sum_u64(t2) * (lane == 0));
      add_u128(sum_c2, wave_sum_u64(c2) * (lane == 0));
      t2 = c2 = 0;
    }
  }
  const uint64_t a = wave_sum_u64(t2), b = wave_sum_u64(c2);
  if (lane == 0) {
    add_u128(sum_t2, a);
    add_u128(sum_c2, b);
  }
}

// ---- k-cliques ----------------------------------------------------------------------------------
//
// #K_k = Σ over pivots u of the (k-1)-cliques of the local DAG G_u on N+(u) (rows as in
// hom_local_kernel: bit j of row i set iff the j-th vertex of N+(u) is an out-neighbour of the
// i-th; only bits above i). A wave takes top-level rows i = wave, wave + 4, ... and walks the
// candidate sets depth first: S_1 = row_i, S_{t+1} = S_t & row_x for each x in S_t; lane l holds
// 32-bit word l of the current set (2W <= 64 words since d <= 2048), so an intersection is one
// AND per lane and the set bits of S_t are taken in order with a ballot and a readlane. The last
// level counts the arcs inside a set lane-parallel, as hom_local_kernel does for K4: the set is
// staged in LDS, its set bits are spread over the lanes, and lane a adds popc(S & row_y) for its
// bit y.

// Arcs of the local DAG inside the set whose 32-bit word `lane` is S (per-lane partial).
__device__ uint64_t kc_arcs_in(const uint32_t *rows, int32_t w2, uint32_t S, uint32_t *stage) {
  const int lane = lane_id();
  stage[lane] = S;
  wave_sync();
  const int32_t nb = __popc(S);
  const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nb);
  const int32_t excl = incl - nb;
  const int32_t bits = __builtin_amdgcn_readlane(incl, kWave - 1);
  // words of S past its last nonzero one add nothing to popc(S & row_y)
  const int32_t wend = 64 - (int32_t)__clzll((unsigned long long)__ballot(S != 0u));
  uint64_t acc = 0;
  for (int32_t a0 = 0; a0 < bits; a0 += kWave) {
    const int32_t a = a0 + lane;
    int32_t o = 0;  // last word with excl <= a
#pragma unroll
    for (int32_t step = kWave / 2; step > 0; step >>= 1)
      if (__shfl(excl, o + step, kWave) <= a) o += step;
    uint32_t wv = (uint32_t)__shfl((int32_t)S, o, kWave);
    int32_t k = a - __shfl(excl, o, kWave);
    if (a >= bits) continue;
    int32_t pos = 0;  // position of the k-th set bit of wv
#pragma unroll
    for (int32_t width = 16; width > 0; width >>= 1) {
      const int32_t c = __popc(wv & ((1u << width) - 1u));
      if (k >= c) {
        k -= c;
        wv >>= width;
        pos += width;
      }
    }
    const int32_t y = o * 32 + pos;
    const uint32_t *ry = rows + (int64_t)y * w2;
    uint32_t c = 0;
    for (int32_t w = (y + 1) >> 5; w < wend; ++w) c += (uint32_t)__popc(stage[w] & ry[w]);
    acc += c;
  }
  wave_sync();
  return acc;
}

// Depth-first levels above the last: for each x in S, recurse on S & row_x.
template <int D>
__device__ uint64_t kc_dfs(const uint32_t *rows, int32_t w2, uint32_t S, uint32_t *stage) {
  const int lane = lane_id();
  uint32_t rem = S;
  uint64_t acc = 0;
  while (true) {
    const uint64_t nz = __ballot(rem != 0u);
    if (!nz) break;
    const int l = __ffsll((unsigned long long)nz) - 1;
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readlane((int)rem, l);
    const int32_t x = l * 32 + (__ffs(wv) - 1);
    if (lane == l) rem &= rem - 1u;
    const uint32_t sx = lane < w2 ? (S & rows[(int64_t)x * w2 + lane]) : 0u;
    if constexpr (D == 1) {
      acc += kc_arcs_in(rows, w2, sx, stage);
    } else {
      acc += kc_dfs<D - 1>(rows, w2, sx, stage);
    }
  }
  return acc;
}

// K5 on a local graph whose rows fit 16 words (d <= 512): the wave's four 16-lane groups take
// the set bits x of row i in turn (group g the bits g, g + 4, ...), each with its own set
// S2 = row_i & row_x, whose inner arcs it counts lane-parallel as kc_arcs_in does over a whole
// wave. S holds word (lane mod 16) of row i in every group. Per arc (i, x) of the local graph
// the full-wave walk paid one ballot, one staging round and one arc enumeration over 64 lanes
// for a set that mostly fits a few words: 120 ms for K5 on R-MAT 20.
constexpr int kKcGroup = 16;

__device__ __forceinline__ int32_t group_incl_scan(int32_t v, int lg) {
#pragma unroll
  for (int o = 1; o < kKcGroup; o <<= 1) {
    const int32_t t = __shfl_up(v, o, kKcGroup);
    if (lg >= o) v += t;
  }
  return v;
}

// Position of the k-th set bit (k < popc) of the group's words, given their exclusive popcount
// prefix: the last word o with excl <= k, then the bit inside it by halving popcounts.
__device__ __forceinline__ int32_t group_kth_bit(uint32_t word, int32_t excl, int32_t k) {
  int32_t o = 0;
#pragma unroll
  for (int32_t step = kKcGroup / 2; step > 0; step >>= 1)
    if (__shfl(excl, o + step, kKcGroup) <= k) o += step;
  uint32_t wv = (uint32_t)__shfl((int32_t)word, o, kKcGroup);
  int32_t kk = k - __shfl(excl, o, kKcGroup);
  int32_t pos = 0;
#pragma unroll
  for (int32_t width = 16; width > 0; width >>= 1) {
    const int32_t c = __popc(wv & ((1u << width) - 1u));
    if (kk >= c) {
      kk -= c;
      wv >>= width;
      pos += width;
    }
  }
  return o * 32 + pos;
}

__device__ uint64_t kc5_row_grouped(const uint32_t *rows, int32_t w2, uint32_t S, uint32_t *stage) {
  const int lane = lane_id(), lg = lane & (kKcGroup - 1), base = lane & ~(kKcGroup - 1);
  const int grp = lane / kKcGroup;
  constexpr int kGroups = kWave / kKcGroup;
  const int32_t nb = __popc(S);
  const int32_t incl = group_incl_scan(nb, lg);
  const int32_t excl = incl - nb;
  const int32_t nbits = __builtin_amdgcn_readlane(incl, kKcGroup - 1);  // the same in every group
  uint64_t acc = 0;
  for (int32_t k0 = 0; k0 < nbits; k0 += kGroups) {
    const int32_t k = k0 + grp;
    const bool act = k < nbits;
    const int32_t x = act ? group_kth_bit(S, excl, k) : 0;
    const uint32_t S2 = act && lg < w2 ? (S & rows[(int64_t)x * w2 + lg]) : 0u;
    stage[lane] = S2;
    wave_sync();
    const int32_t nb2 = __popc(S2);
    const int32_t incl2 = group_incl_scan(nb2, lg);
    const int32_t excl2 = incl2 - nb2;
    const int32_t bits2 = __shfl(incl2, kKcGroup - 1, kKcGroup);
    // words of S2 past its last nonzero one add nothing
    const uint32_t nzw = (uint32_t)(__ballot(S2 != 0u) >> base) & ((1u << kKcGroup) - 1u);
    const int32_t wend = 32 - __clz((int)nzw);
    int32_t maxbits = 0;
#pragma unroll
    for (int q = 0; q < kGroups; ++q)
      maxbits = max(maxbits, __builtin_amdgcn_readlane(incl2, q * kKcGroup + kKcGroup - 1));
    for (int32_t a0 = 0; a0 < maxbits; a0 += kKcGroup) {
      const int32_t a = a0 + lg;
      const int32_t y = group_kth_bit(S2, excl2, min(a, max(bits2 - 1, 0)));
      if (a >= bits2) continue;
      const uint32_t *ry = rows + (int64_t)y * w2;
      uint32_t c = 0;
      for (int32_t w = (y + 1) >> 5; w < wend; ++w) c += (uint32_t)__popc(stage[base + w] & ry[w]);
      acc += c;
    }
    wave_sync();  // the stage is rewritten for the next bits
  }
  return acc;
}

struct U128Acc {
  unsigned long long lo, hi;
};

// One workgroup per pivot from a pivot list, the local graph built as in hom_local_kernel (LDS
// pool: the position set, then the rows when they fit). K = clique size, 3 <= K <= 10.
template <int K, int WAVES, int POOL, bool BIG>
__global__ void __launch_bounds__(WAVES * kWave)
    kclique_kernel(gary_device_graph g, const int64_t *__restrict__ up_start,
                   const int32_t *__restrict__ up_len, const int32_t *__restrict__ items,
                   const int32_t *__restrict__ nitems, int32_t *__restrict__ cursor,
                   uint32_t *__restrict__ row_scratch, int64_t scratch_words, U128Acc *total,
                   int32_t *__restrict__ unsupported) {
  constexpr int kBlock = WAVES * kWave;
  __shared__ __align__(8) int32_t pool[POOL];
  __shared__ uint32_t stage[WAVES][kWave];
  __shared__ int32_t s_item;
  int32_t *const set = pool;
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  const int32_t m = g.csr.m;
  const int32_t *__restrict__ col = g.csr.colIdx;
  const int32_t n_items = *nitems;
  uint64_t acc = 0;
  // thread 0 claims kClaim consecutive items per atomic and hands them out one at a time
  // through s_item. One at a time here: the clique search of a pivot varies far more than the
  // local-graph pass does, and claims of four made K5 on R-MAT 20 18% slower.
  constexpr int32_t kClaim = 1;
  int32_t next = 0, end = 0;  // thread 0's claimed range
  while (true) {
    if (threadIdx.x == 0) {
      if (next >= end) {
        next = atomicAdd(cursor, kClaim);
        end = next + kClaim;
      }
      s_item = next++;
    }
    __syncthreads();
    const int32_t it = s_item;
    if (it >= n_items) break;
    const int32_t u = items[BIG ? n_items - 1 - it : it];
    const int32_t d = ((uint32_t)u < (uint32_t)m) ? up_len[u] : 0;
    if (d < K - 1) {  // fewer than K - 1 out-neighbours: no K-clique has u as its lowest vertex
      __syncthreads();
      continue;
    }
    const int32_t *L = col + up_start[u];
    PosSet ps;
    const int32_t w2 = ((d + 63) >> 6) * 2;  // 32-bit words per local row (<= 64)
    const int64_t row_words = (int64_t)d * w2;
    const int64_t lds_rows = (int64_t)POOL - kSetWords;
    if (d > kMaxHomDeg || !posset_plan(L[0], L[d - 1], d, kSetWords, ps) ||
        (row_words > lds_rows && row_words > scratch_words)) {
      if (threadIdx.x == 0) *unsupported = d;
      __syncthreads();
      continue;
    }
    uint32_t *rows = row_words <= lds_rows ? (uint32_t *)(pool + kSetWords)
                                           : row_scratch + (int64_t)blockIdx.x * scratch_words;
    for (int64_t i = threadIdx.x; i < row_words; i += kBlock) rows[i] = 0;
    posset_build(set, ps, L, d);  // begins and ends with a workgroup barrier
    if (threadIdx.x == 0) s_item = 0;  // row counter of the count phase (u is read by now)
    // Stream the rows N+(v_i): wave wib takes the local rows i = b0 + wib + WAVES * l of each
    // batch b0 (lane l's row), flattened into the aligned 8-element (32-byte) chunks that overlap
    // them, 64 chunks per window, lane l taking chunk w0 + l: the chunk's row comes from a tagged
    // slot (the wave's stage words, which the count phase rewrites, so every pivot starts with a
    // clear) and a max-scan, its colIdx position and valid slots from shuffles, as in the
    // triangle kernel's row streaming. (A wave per row with lanes striding it left most lanes
    // idle on the short rows.)
    int32_t *const slots = reinterpret_cast<int32_t *>(stage[wib]);
    int32_t tag = 1 << 30;
    for (int32_t b0 = 0; b0 < d; b0 += WAVES * kWave) {
      const int32_t i = b0 + wib + WAVES * lane;
      int64_t pv = 0;
      int32_t dv = 0;
      if (i < d) {
        const int32_t v = L[i];
        if ((uint32_t)v < (uint32_t)m) {
          pv = up_start[v];
          dv = up_len[v];
        }
      }
      const int32_t nch = dv > 0 ? (int32_t)(((pv + dv + 7) >> 3) - (pv >> 3)) : 0;
      const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nch);
      const int32_t start = incl - nch;
      const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
      const uint32_t base_r = (uint32_t)(pv >> 3) - (uint32_t)start;
      const uint32_t A = (uint32_t)(pv & 7) + 8u * (uint32_t)start;
      const uint32_t Bv = A + (uint32_t)dv;
      for (int32_t w0 = 0; w0 < total; w0 += kWave) {
        if (tag >= (1 << 30)) {  // stale slots could outrank new tags: clear them
          slots[lane] = -1;
          tag = 0;
          wave_sync();
        }
        if (nch > 0 && start < w0 + kWave && incl > w0) slots[max(0, start - w0)] = tag | lane;
        wave_sync();
        const int32_t row = wave_inclusive_max_i32(slots[lane]) & (kWave - 1);
        wave_sync();
        tag += kWave;
        const int32_t xw = w0 + lane;
        const uint32_t ch = (uint32_t)__shfl((int32_t)base_r, row, kWave) + (uint32_t)xw;
        const int32_t jlo = (int32_t)((uint32_t)__shfl((int32_t)A, row, kWave) - 8u * (uint32_t)xw);
        int32_t jhi = (int32_t)((uint32_t)__shfl((int32_t)Bv, row, kWave) - 8u * (uint32_t)xw);
        const int32_t ri = b0 + wib + WAVES * row;  // local row of the chunk
        int4 lo4 = make_int4(-1, -1, -1, -1), hi4 = make_int4(-1, -1, -1, -1);
        if (xw < total) {
          const uint64_t e0 = (uint64_t)ch << 3;
          lo4 = *reinterpret_cast<const int4 *>(col + e0);
          if (e0 + 4 < (uint64_t)g.csr.nnz) hi4 = *reinterpret_cast<const int4 *>(col + e0 + 4);
        } else {
          jhi = 0;
        }
        const int32_t el[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (q < jlo || q >= jhi) continue;
          const int32_t j = posset_find(set, ps, el[q]);
          if (j >= 0) atomicOr(&rows[(int64_t)ri * w2 + (j >> 5)], 1u << (j & 31));
        }
      }
    }
    __syncthreads();
    // Rows to waves: for K >= 5 one at a time from s_item (a row's clique search varies widely:
    // K5 on R-MAT 20 139 -> 131 ms), else statically (claims cost K4 about 1%).
    for (int32_t r = wib;; r += WAVES) {
      int32_t i = r;
      if constexpr (K >= 5) {
        if (lane == 0) i = atomicAdd(&s_item, 1);
        i = __shfl(i, 0, kWave);
      }
      if (i >= d) break;
      if constexpr (K == 5) {
        if (w2 <= kKcGroup) {  // four 16-lane groups per wave (row i in each)
          const int lg = lane & (kKcGroup - 1);
          const uint32_t S16 = lg < w2 ? rows[(int64_t)i * w2 + lg] : 0u;
          acc += kc5_row_grouped(rows, w2, S16, stage[wib]);
          continue;
        }
      }
      const uint32_t S = lane < w2 ? rows[(int64_t)i * w2 + lane] : 0u;
      if constexpr (K == 3) {
        acc += (uint32_t)__popc(S);
      } else if constexpr (K == 4) {
        acc += kc_arcs_in(rows, w2, S, stage[wib]);
      } else {
        acc += kc_dfs<K - 4>(rows, w2, S, stage[wib]);
      }
    }
    __syncthreads();
  }
  const uint64_t s = wave_sum_u64(acc);
  if (lane == 0 && s) {
    const unsigned long long old = atomicAdd(&total->lo, (unsigned long long)s);
    if (old + s < old) atomicAdd(&total->hi, 1ull);
  }
}

// ---- 4-cycles ------------------------------------------------------------------------------------

constexpr int kC4HeavyLo = 256;  // lower neighbours from which a vertex takes the window path

// First index of N(v) with value >= key.
__device__ __forceinline__ int32_t lower_bound(const int32_t *__restrict__ N, int32_t d,
                                               int32_t key) {
  int32_t l = 0, r = d;
  while (l < r) {
    const int32_t mid = (l + r) >> 1;
    if (N[mid] < key) l = mid + 1;
    else r = mid;
  }
  return l;
}

// Light vertices with few wedges: one wave per vertex u (static round robin over the list),
// counting the endpoints of its wedges in the wave's own LDS hash table (kLightSlots keys, 8-bit
// counts: a count never exceeds the lower degree, which is below 256). A vertex with more than
// kLightMaxWedges wedges goes to `overflow` for c4_kernel. The runs N(v) ∩ [0, u) of the (at most
// 255) lower neighbours are flattened over the wave as in the window kernel. On R-MAT 22 this
// takes 85% of the 499K light vertices (8% of their wedges); one workgroup per vertex spent most
// of its time on per-vertex barriers and latency.
constexpr int kLightWaves = 4;
constexpr int kLightSlots = 1024;
constexpr int kLightMaxWedges = 512;  // load factor <= 1/2: probing always ends
static_assert(kLightSlots == 1 << 10, "hslot(e, 10) indexes the table");
static_assert(kC4HeavyLo <= 4 * kWave && kC4HeavyLo <= 256, "neighbours per lane; 8-bit counts");

__global__ void __launch_bounds__(kLightWaves * kWave)
    c4_light_wave_kernel(gary_device_graph g, const int32_t *__restrict__ lo_len,
                         const int32_t *__restrict__ light, const int32_t *__restrict__ nlight,
                         int32_t *__restrict__ overflow, int32_t *__restrict__ noverflow,
                         unsigned long long *c4_acc) {
  __shared__ int32_t keys_all[kLightWaves][kLightSlots];
  __shared__ uint32_t cnt_all[kLightWaves][kLightSlots / 4];
  __shared__ int32_t slots_all[kLightWaves][kWave];
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  int32_t *keys = keys_all[wib];
  uint32_t *cnt = cnt_all[wib];
  int32_t *slots = slots_all[wib];
  const int32_t *__restrict__ col = g.csr.colIdx;
  for (int32_t x = lane; x < kLightSlots; x += kWave) keys[x] = -1;
  for (int32_t x = lane; x < kLightSlots / 4; x += kWave) cnt[x] = 0u;
  wave_sync();
  uint64_t acc = 0;
  const int64_t n_items = *nlight;
  for (int64_t it = global_wave_id(); it < n_items; it += num_global_waves()) {
    const int32_t u = light[it];
    const int32_t nl = lo_len[u];  // 2 <= nl < kC4HeavyLo = 4 * kWave
    const int32_t *Nu = col + g.csr.indPtr[u];
    int64_t pv[kC4HeavyLo / kWave];
    int32_t cut[kC4HeavyLo / kWave];
    int32_t mine = 0;
#pragma unroll
    for (int r = 0; r < kC4HeavyLo / kWave; ++r) {
      const int32_t k = r * kWave + lane;
      pv[r] = 0;
      cut[r] = 0;
      if (k < nl) {
        const int32_t v = Nu[k];
        pv[r] = g.csr.indPtr[v];
        cut[r] = lower_bound(col + pv[r], g.csr.deg[v], u);  // N(v) ∩ [0, u)
        mine += cut[r];
      }
    }
    if (wave_sum_u64((uint64_t)mine) > (uint64_t)kLightMaxWedges) {
      if (lane == 0) overflow[atomicAdd(noverflow, 1)] = u;
      continue;
    }
#pragma unroll
    for (int r = 0; r < kC4HeavyLo / kWave; ++r) {
      if (r * kWave >= nl) break;
      const int32_t len = cut[r];
      const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)len);
      const int32_t start = incl - len;
      const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
      for (int32_t w0 = 0; w0 < total; w0 += kWave) {
        slots[lane] = -1;
        wave_sync();
        if (len > 0 && start < w0 + kWave && incl > w0) slots[max(0, start - w0)] = lane;
        wave_sync();
        const int32_t row = max(0, wave_inclusive_max_i32(slots[lane]));
        const int64_t rp = __shfl(pv[r], row, kWave);
        const int32_t rs = __shfl(start, row, kWave);
        const int32_t e = w0 + lane < total ? col[rp + (w0 + lane - rs)] : -1;
        wave_sync();  // slots are rewritten by the next step
        if (e >= 0) {
          uint32_t h = hslot(e, 10);
          while (true) {
            const int32_t old = atomicCAS(&keys[h], -1, e);
            if (old == -1 || old == e) {
              atomicAdd(&cnt[h >> 2], 1u << (8 * (h & 3)));
              break;
            }
            h = (h + 1) & (kLightSlots - 1);
          }
        }
      }
    }
    wave_sync();
    for (int32_t x = lane; x < kLightSlots; x += kWave) {
      if (keys[x] != -1) {
        const uint64_t c = (cnt[x >> 2] >> (8 * (x & 3))) & 0xffu;
        acc += c * (c - 1) / 2;
        keys[x] = -1;
      }
    }
    wave_sync();
    for (int32_t x = lane; x < kLightSlots / 4; x += kWave) cnt[x] = 0u;
    wave_sync();
  }
  const uint64_t sacc = wave_sum_u64(acc);
  if (lane == 0) add_u128(c4_acc, sacc);
}

// Light vertices with more wedges (up to kBlockMaxWedges): one workgroup per vertex, counting in
// an LDS hash table of packed words (vertex id << 8 | count; ids below 2^24 - 1, counts below 256
// as above) sized per vertex to the next power of two >= 2 * wedges. The wedges of u form one
// flat range [0, W): thread t takes endpoints t, t + block, ..., finding its lower neighbour by
// a binary search over the run offsets in LDS. A vertex with more wedges goes to `overflow`.
constexpr int kBlockThreads = 1024;
constexpr int kBlockSlots = 32768;  // 128 KiB
constexpr int kBlockMaxWedges = kBlockSlots / 2;
constexpr uint32_t kBlockEmpty = 0xffffffffu;
constexpr int32_t kBlockMaxId = (1 << 24) - 2;  // largest vertex id a packed key holds

__global__ void __launch_bounds__(kBlockThreads)
    c4_light_block_kernel(gary_device_graph g, const int32_t *__restrict__ lo_len,
                          const int32_t *__restrict__ items, const int32_t *__restrict__ nitems,
                          int32_t *__restrict__ overflow, int32_t *__restrict__ noverflow,
                          int32_t *__restrict__ cursor, unsigned long long *c4_acc) {
  __shared__ uint32_t table[kBlockSlots];
  __shared__ int64_t pv_s[kC4HeavyLo];
  __shared__ int32_t off_s[kC4HeavyLo + 1];  // exclusive prefix of the run lengths
  __shared__ int32_t s_item;
  const int lane = lane_id();
  const int32_t *__restrict__ col = g.csr.colIdx;
  for (int32_t x = threadIdx.x; x < kBlockSlots; x += kBlockThreads) table[x] = kBlockEmpty;
  uint64_t acc = 0;
  const int32_t n_items = *nitems;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(cursor, 1);
    __syncthreads();
    const int32_t it = s_item;
    if (it >= n_items) break;
    const int32_t u = items[it];
    const int32_t nl = lo_len[u];  // 2 <= nl < kC4HeavyLo
    const int32_t *Nu = col + g.csr.indPtr[u];
    if (threadIdx.x < kC4HeavyLo) {
      const int32_t k = threadIdx.x;
      int32_t cut = 0;
      if (k < nl) {
        const int32_t v = Nu[k];
        pv_s[k] = g.csr.indPtr[v];
        cut = lower_bound(col + pv_s[k], g.csr.deg[v], u);  // N(v) ∩ [0, u)
      }
      off_s[k + 1] = cut;
    }
    if (threadIdx.x == 0) off_s[0] = 0;
    __syncthreads();
    if (threadIdx.x < kWave) {  // inclusive scan of off_s[1 .. kC4HeavyLo]
      uint32_t carry = 0;
      for (int32_t c0 = 0; c0 < kC4HeavyLo; c0 += kWave) {
        const uint32_t v = (uint32_t)off_s[c0 + lane + 1];
        const uint32_t inc = wave_inclusive_scan_u32(v);
        off_s[c0 + lane + 1] = (int32_t)(carry + inc);
        carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
      }
    }
    __syncthreads();
    const int32_t W = off_s[nl];
    if (W > kBlockMaxWedges) {
      if (threadIdx.x == 0) overflow[atomicAdd(noverflow, 1)] = u;
      __syncthreads();
      continue;
    }
    int32_t log_t = 10;
    while ((1 << log_t) < 2 * W) ++log_t;
    const uint32_t mask = (1u << log_t) - 1u;
    for (int32_t e = threadIdx.x; e < W; e += kBlockThreads) {
      int32_t lo = 0, hi = nl - 1;  // last k with off_s[k] <= e
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (off_s[mid] <= e) lo = mid;
        else hi = mid - 1;
      }
      const int32_t w = col[pv_s[lo] + (e - off_s[lo])];
      const uint32_t key = (uint32_t)w << 8;
      uint32_t h = hslot(w, log_t);
      while (true) {
        uint32_t cur = table[h];
        if (cur == kBlockEmpty) {
          cur = atomicCAS(&table[h], kBlockEmpty, key | 1u);
          if (cur == kBlockEmpty) break;
        }
        if ((cur & ~0xffu) == key) {
          atomicAdd(&table[h], 1u);
          break;
        }
        h = (h + 1) & mask;
      }
    }
    __syncthreads();
    for (int32_t x = threadIdx.x; x <= (int32_t)mask; x += kBlockThreads) {
      const uint32_t t = table[x];
      if (t != kBlockEmpty) {
        const uint64_t c = t & 0xffu;
        acc += c * (c - 1) / 2;
        table[x] = kBlockEmpty;
      }
    }
    __syncthreads();
  }
  const uint64_t sacc = wave_sum_u64(acc);
  if (lane == 0) add_u128(c4_acc, sacc);
}

// One workgroup per light vertex u (2 <= lower degree < kC4HeavyLo) that has too many wedges
// for c4_light_wave_kernel, from the list that kernel writes: counts wedges u-v-w (v, w < u)
// into this block's dense counter slice (m entries, all zero between vertices), listing each
// endpoint w the first time its counter leaves zero; then reads and clears the listed counters,
// adding C(c, 2). The second pass touches the distinct endpoints only. (Claiming every vertex id
// from the cursor instead, most of them with nothing to count, cost a global atomic and two
// barriers per id: 105 ms of the C4 count on R-MAT 22.)
__global__ void __launch_bounds__(kHomBlock)
    c4_kernel(gary_device_graph g, const int32_t *__restrict__ lo_len,
              const int32_t *__restrict__ light, const int32_t *__restrict__ nlight,
              uint32_t *__restrict__ counters, int32_t *__restrict__ touched_all,
              int32_t *__restrict__ cursor, unsigned long long *c4_acc) {
  __shared__ int32_t s_item, s_nt;
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  const int32_t m = g.csr.m;
  const int32_t *__restrict__ col = g.csr.colIdx;
  uint32_t *cnt = counters + (int64_t)blockIdx.x * m;
  int32_t *touched = touched_all + (int64_t)blockIdx.x * m;
  uint64_t acc = 0;
  const int32_t n_items = *nlight;
  while (true) {
    if (threadIdx.x == 0) {
      s_item = atomicAdd(cursor, 1);
      s_nt = 0;
    }
    __syncthreads();
    const int32_t item = s_item;
    if (item >= n_items) break;
    const int32_t u = light[item];
    const int32_t nl = lo_len[u];
    const int32_t *Nu = col + g.csr.indPtr[u];
    {
      for (int32_t k = wib; k < nl; k += kHomWaves) {
        const int32_t v = Nu[k];
        const int32_t *Nv = col + g.csr.indPtr[v];
        const int32_t cut = lower_bound(Nv, g.csr.deg[v], u);
        for (int32_t p0 = 0; p0 < cut; p0 += kWave) {
          const int32_t p = p0 + lane;
          int32_t w = -1;
          bool first = false;
          if (p < cut) {
            w = Nv[p];
            first = atomicAdd(&cnt[w], 1u) == 0u;
          }
          const uint64_t fm = __ballot(first);
          if (fm) {  // wave-uniform
            int32_t base = 0;
            if (lane == 0) base = atomicAdd(&s_nt, __popcll(fm));
            base = __shfl(base, 0, kWave);
            if (first) touched[base + lanes_below(fm)] = w;
          }
        }
      }
      __syncthreads();
      const int32_t nt = s_nt;
      for (int32_t k = threadIdx.x; k < nt; k += kHomBlock) {
        // read-and-clear at L2, where the counting atomics landed (a plain load could hit a
        // line this CU cached before those atomics)
        const uint64_t c = atomicExch(&cnt[touched[k]], 0u);
        acc += c * (c - 1) / 2;
      }
    }
    __syncthreads();
  }
  const uint64_t s = wave_sum_u64(acc);
  if (lane == 0) add_u128(c4_acc, s);
}

// ---- 4-cycles at heavy vertices: endpoint windows in LDS -----------------------------------------
//
// A vertex u with many lower neighbours has most of the wedges and endpoints spread over all of
// [0, u), so its dense counters would live in HBM. Here the endpoint range is cut into windows of
// kC4Win ids whose counters fit LDS (16-bit counters, two per word: a count never exceeds the
// lower degree of u; a vertex with 65536 or more lower neighbours counts each window in two
// halves with 32-bit counters); a work item is (u, kC4Group consecutive windows). Items are
// claimed window group by window group, so the workgroups read the same id range of the
// neighbour lists together (R-MAT 22: 281 ms, against 334 ms vertex by vertex), and within a
// group from the highest-ranked vertex (most wedges) down, so a group ends with small items. The
// endpoints of N(v) in a window form one run of N(v) (sorted), bounded by the run index below.
// The runs of 64 neighbours v are flattened over a wave (as in the triangle kernel's row
// streaming) so that every lane counts one endpoint per step. (Loading the cursor with nx[i] and
// the three index entries together, two dependent loads per batch instead of three, made no
// difference: 153.1 -> 153.7 ms per R-MAT 22 C4 call.)
//
// Run index. Each neighbour's run in a window is read from a per-call index of the CSR: for
// every row v, one 64-bit entry per window of kC4Win ids that N(v) touches (window id << 40 |
// position of its first element of N(v) there) and a last entry at the row's end. A visit with
// a run reads two consecutive entries for the run's bounds and keeps the next entry's window id
// as the neighbour's next window, so no run end is searched for in N(v). (Searching each run's
// end in N(v) cost log2 d(v) dependent reads, in as many cache lines, per run: 1.0e9 runs of 36
// endpoints on average on R-MAT 22.)
constexpr int kRunPosBits = 40;
constexpr uint64_t kRunPosMask = (1ull << kRunPosBits) - 1;
constexpr int32_t kRunEndWin = (1 << 23) - 1;  // window id of a row's last entry
constexpr int kC4Words = 32768; // 128 KiB of LDS counters: one workgroup per CU
constexpr int kC4Win = 65536;   // ids per window: 16-bit counters, two per word
constexpr int kC4Group = 16;    // windows per work item (each item starts its cursors with a
                                // binary search per neighbour; 4 windows: 327 ms on R-MAT 22)
constexpr int kC4Block = 1024;
constexpr int kC4Waves = kC4Block / kWave;
constexpr int kC4Windows = 2;   // 64-chunk windows whose loads are issued together

// Entries per row of the run index: the windows N(v) touches, plus the row's last entry. One wave
// per row; cnt[m] = 0 (the scan's total slot).
__global__ void __launch_bounds__(256) run_count_kernel(gary_device_graph g, int32_t win_shift,
                                                        uint32_t *__restrict__ cnt) {
  const int lane = lane_id();
  const int32_t *__restrict__ col = g.csr.colIdx;
  for (int64_t v = global_wave_id(); v <= g.csr.m; v += num_global_waves()) {
    if (v == g.csr.m) {
      if (lane == 0) cnt[v] = 0;
      continue;
    }
    const int64_t p0 = g.csr.indPtr[v];
    const int32_t d = g.csr.deg[v];
    uint32_t c = 0;
    for (int32_t e0 = 0; e0 < d; e0 += kWave) {
      const int32_t e = e0 + lane;
      const bool first = e < d && (e == 0 || (col[p0 + e] >> win_shift) != (col[p0 + e - 1] >> win_shift));
      c += (uint32_t)__popcll(__ballot(first));
    }
    if (lane == 0) cnt[v] = c + 1;
  }
}

// The entries themselves, at ptr[v] (the scanned counts), in window order.
__global__ void __launch_bounds__(256) run_write_kernel(gary_device_graph g, int32_t win_shift,
                                                        const uint32_t *__restrict__ ptr,
                                                        uint64_t *__restrict__ runs) {
  const int lane = lane_id();
  const int32_t *__restrict__ col = g.csr.colIdx;
  for (int64_t v = global_wave_id(); v < g.csr.m; v += num_global_waves()) {
    const int64_t p0 = g.csr.indPtr[v];
    const int32_t d = g.csr.deg[v];
    uint64_t at = ptr[v];
    for (int32_t e0 = 0; e0 < d; e0 += kWave) {
      const int32_t e = e0 + lane;
      const int32_t w = e < d ? col[p0 + e] : 0;
      const bool first = e < d && (e == 0 || (w >> win_shift) != (col[p0 + e - 1] >> win_shift));
      const uint64_t mask = __ballot(first);
      if (first)
        runs[at + lanes_below(mask)] =
            ((uint64_t)(w >> win_shift) << kRunPosBits) | (uint64_t)(p0 + e);
      at += (uint64_t)__popcll(mask);
    }
    if (lane == 0) runs[at] = ((uint64_t)kRunEndWin << kRunPosBits) | (uint64_t)(p0 + d);
  }
}

__global__ void __launch_bounds__(kC4Block)
    c4_window_kernel(gary_device_graph g, const int32_t *__restrict__ lo_len,
                     const int32_t *__restrict__ heavy, const int32_t *__restrict__ nheavy_p,
                     int32_t jmax, const uint32_t *__restrict__ run_ptr,
                     const uint64_t *__restrict__ runs, int64_t *__restrict__ cursors,
                     int32_t *__restrict__ next_win, int64_t stride,
                     int32_t *__restrict__ item_cursor, unsigned long long *c4_acc) {
  __shared__ __align__(16) uint32_t cnt[kC4Words];
  __shared__ int32_t slots[kC4Waves][kWave];
  __shared__ int32_t s_item, s_kskip, s_batch, s_any;
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  const int32_t *__restrict__ col = g.csr.colIdx;
  const int64_t nheavy = *nheavy_p;
  const int64_t nitems = nheavy * jmax;
  int64_t *cur = cursors + (int64_t)blockIdx.x * stride;     // run-index entry of the next run
  int32_t *nx = next_win + (int64_t)blockIdx.x * stride;     // window id of that run
  uint64_t acc = 0;
  int32_t tag = 1 << 30;  // slot tags of this wave's chunk hand-off (first use clears the slots)
  for (int32_t x = threadIdx.x; x < kC4Words; x += kC4Block) cnt[x] = 0u;  // zero between windows
  if (threadIdx.x == 0) {
    s_batch = 0;
    s_any = 0;
    // An endpoint w of degree <= 1 closes no 4-cycle (c_u(w) <= d(w) <= 1), and ids are ranked
    // by degree, so the windows below the first id of degree >= 2 are skipped whole (on R-MAT 22,
    // 35 of its 64 windows: 43% of the ids are isolated and 56% have degree below 2).
    int32_t lo = 0, hi = g.csr.m;
    while (lo < hi) {
      const int32_t mid = (int32_t)(((int64_t)lo + hi) >> 1);
      if (g.csr.deg[mid] < 2) lo = mid + 1;
      else hi = mid;
    }
    s_kskip = lo / kC4Win;
  }
  __syncthreads();
  const int32_t kskip = s_kskip;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(item_cursor, 1);
    __syncthreads();
    const int64_t it = s_item;
    __syncthreads();  // everyone holds the item before s_item is written again
    if (it >= nitems) break;
    const int32_t u = heavy[nheavy - 1 - it % nheavy];  // the list ascends by id (rank)
    const int32_t j = (int32_t)(it / nheavy);
    const int32_t K = (u + kC4Win - 1) / kC4Win;
    const int32_t k0 = max(j * kC4Group, kskip), k1 = min(K, j * kC4Group + kC4Group);
    if (k0 >= k1) continue;
    const int32_t nl = lo_len[u];
    const int32_t *Nu = col + g.csr.indPtr[u];
    for (int32_t i = threadIdx.x; i < nl; i += kC4Block) {  // cursors at the first window
      const int32_t v = Nu[i];
      int64_t lo = run_ptr[v], hi = (int64_t)run_ptr[v + 1] - 1;  // the last entry ends the row
      while (lo < hi) {  // first entry of a window >= k0
        const int64_t mid = (lo + hi) >> 1;
        if ((int32_t)(runs[mid] >> kRunPosBits) < k0) lo = mid + 1;
        else hi = mid;
      }
      cur[i] = lo;
      nx[i] = (int32_t)(runs[lo] >> kRunPosBits);
    }
    __syncthreads();
    // Neighbours per wave. N(u) ascends by rank, i.e. by degree, and a neighbour's runs grow with
    // its degree, so contiguous shares would leave the last wave the longest runs: below kC4Block
    // neighbours each wave takes every kC4Waves-th one (an even share of every degree range, so
    // every wave of the one-per-CU workgroup has loads in flight); from kC4Block on, the waves take
    // batches of 64 consecutive neighbours, the first kC4Waves by wave index and the rest from an
    // LDS counter, so a wave that drew short runs takes more batches.
    const bool few = nl < kC4Block;
    const int32_t per = few ? (nl + kC4Waves - 1) / kC4Waves : kWave;
    // Passes over the index windows k0 .. k1 - 1 (kC4Win ids each). A count c_u(w) is at most
    // min(nl, d(w)), and ids are ranked by degree, so d(w) <= deg[whi - 1] in a pass over
    // [wlo, whi). The counter width of a pass is the narrowest that cannot overflow:
    //   8 bits, 4 per word: 2 * kC4Win ids, two index windows (their runs joined);
    //   16 bits, 2 per word: one index window;
    //   32 bits: half an index window (the run is read in both halves; the cursor moves after
    //   the second).
    int32_t k = k0, half = 0;
    while (k < k1 && k * kC4Win < u) {
      const int32_t wbase = k * kC4Win;
      int32_t bits, wlo, whi, nidx = k + 1 < k1 ? 2 : 1;
      const int32_t whi8 = min(wbase + nidx * kC4Win, u), whi16 = min(wbase + kC4Win, u);
      if (nl < 256 || g.csr.deg[whi8 - 1] < 256) {
        bits = 8;
        wlo = wbase;
        whi = whi8;
      } else if (nl < 65536 || g.csr.deg[whi16 - 1] < 65536) {
        bits = 16;
        nidx = 1;
        wlo = wbase;
        whi = whi16;
      } else {
        nidx = 1;
        bits = 32;
        wlo = wbase + half * (kC4Win / 2);
        whi = min(wlo + kC4Win / 2, u);
      }
      // the cursors move at the end of the index windows of this pass
      const bool advance = bits != 32 || whi == min(wbase + kC4Win, u);
      const int lgc = bits == 8 ? 2 : bits == 16 ? 1 : 0;  // log2 counters per word
      const uint32_t cmask = bits == 32 ? 0xffffffffu : (1u << bits) - 1;
      for (int32_t bi = wib; few ? bi == wib : bi * kWave < nl;) {
        const int32_t i = few ? wib + kC4Waves * lane : bi * kWave + lane;
        int64_t p = 0;
        int32_t len = 0;
        // the run of N(v) in index windows [k, k + nidx): nx[i] caches the window of the
        // neighbour's next run, so a neighbour with none here costs one 4-byte load; two
        // consecutive entries bound a run, and a third joins the next window's run
        if (lane < per && i < nl && nx[i] < k + nidx) {
          const int64_t jc = cur[i];
          const uint64_t a = runs[jc];
          uint64_t b = runs[jc + 1];
          int64_t nxt = jc + 1;
          if (nidx == 2 && (int32_t)(a >> kRunPosBits) == k && (int32_t)(b >> kRunPosBits) == k + 1) {
            b = runs[jc + 2];  // entry jc + 1 is not the row's last, so jc + 2 exists
            nxt = jc + 2;
          }
          p = (int64_t)(a & kRunPosMask);
          len = (int32_t)((int64_t)(b & kRunPosMask) - p);
          // the pass that reaches u: the run stops at u itself (u is in N(v), the first
          // element >= u)
          if (whi == u) len = lower_bound(col + p, len, u);
          if (advance) {
            cur[i] = nxt;
            nx[i] = (int32_t)(b >> kRunPosBits);
          }
        }
        // The batch's runs are flattened into the aligned 8-element (32-byte) chunks that overlap
        // them, 64 per window, lane l taking chunk w0 + l, as in the triangle kernel's row
        // streaming: the lane's run comes from a tagged slot and a max-scan, its chunk and valid
        // slots from three shuffles, so that hand-off is paid once per 8 endpoints (once per
        // endpoint, with 4 endpoints per lane in flight: 153 ms per R-MAT 22 C4 call).
        // Chunk indices are 32-bit (the window path needs nnz < 2^32).
        const int32_t nch = len > 0 ? (int32_t)(((p + len + 7) >> 3) - (p >> 3)) : 0;
        const int32_t incl = (int32_t)wave_inclusive_scan_u32((uint32_t)nch);
        const int32_t start = incl - nch;
        const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
        const uint32_t base_r = (uint32_t)(p >> 3) - (uint32_t)start;
        const uint32_t A = (uint32_t)(p & 7) + 8u * (uint32_t)start;
        const uint32_t Bv = A + (uint32_t)len;
        const int32_t span = whi - wlo;
        if (total > 0 && lane == 0) s_any = 1;  // the pass has counters to sweep
        // kC4Windows windows per trip: each window's hand-off, then its loads, so the loads of
        // the trip's windows are in flight together before any endpoint is counted
        for (int32_t w0 = 0; w0 < total; w0 += kC4Windows * kWave) {
          int4 data[kC4Windows][2];
          int32_t jlo[kC4Windows], jhi[kC4Windows];
#pragma unroll
          for (int r = 0; r < kC4Windows; ++r) {
            const int32_t ws = w0 + r * kWave;
            data[r][0] = data[r][1] = make_int4(0, 0, 0, 0);
            jlo[r] = 0;
            jhi[r] = 0;
            if (ws >= total) continue;  // wave-uniform
            if (tag >= (1 << 30)) {  // stale slots could outrank new tags: clear them
              slots[wib][lane] = -1;
              tag = 0;
              wave_sync();
            }
            if (nch > 0 && start < ws + kWave && incl > ws) slots[wib][max(0, start - ws)] = tag | lane;
            wave_sync();
            const int32_t row = wave_inclusive_max_i32(slots[wib][lane]) & (kWave - 1);
            wave_sync();
            tag += kWave;
            const int32_t xw = ws + lane;
            const uint32_t ch = (uint32_t)__shfl((int32_t)base_r, row, kWave) + (uint32_t)xw;
            jlo[r] = (int32_t)((uint32_t)__shfl((int32_t)A, row, kWave) - 8u * (uint32_t)xw);
            jhi[r] = (int32_t)((uint32_t)__shfl((int32_t)Bv, row, kWave) - 8u * (uint32_t)xw);
            if (xw < total) {
              const uint64_t e0 = (uint64_t)ch << 3;
              data[r][0] = *reinterpret_cast<const int4 *>(col + e0);
              if (e0 + 4 < (uint64_t)g.csr.nnz)
                data[r][1] = *reinterpret_cast<const int4 *>(col + e0 + 4);
            } else {
              jhi[r] = 0;
            }
          }
#pragma unroll
          for (int r = 0; r < kC4Windows; ++r) {
            const int32_t el[8] = {data[r][0].x, data[r][0].y, data[r][0].z, data[r][0].w,
                                   data[r][1].x, data[r][1].y, data[r][1].z, data[r][1].w};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const int32_t x = el[q] - wlo;
              // slots of this lane's run, endpoints in [wlo, whi) (others: not this pass)
              if (q < jlo[r] || q >= jhi[r] || x < 0 || x >= span) continue;
              const uint32_t sh = (uint32_t)(x & ((1 << lgc) - 1)) * (uint32_t)bits;
              atomicAdd(&cnt[x >> lgc], 1u << sh);
            }
          }
        }
        if (few) break;
        int32_t nb = 0;
        if (lane == 0) nb = atomicAdd(&s_batch, 1);
        bi = kC4Waves + __builtin_amdgcn_readfirstlane(__shfl(nb, 0, kWave));
      }
      __syncthreads();
      // read and clear every counter of the pass (counting with fetch-adds that return nothing
      // and listing first touches, up to 2,048 a pass, cost more: 91.6 -> 84.7 ms per R-MAT 22
      // C4 call with the sweep), four words per 16-byte LDS access. C(c, 2) per word: 8-bit
      // counters (c <= 255) by two byte dot products, (sum c^2 - sum c) / 2, exact in 32 bits;
      // wider ones in 64 bits. (Counter by counter in 64-bit arithmetic, the sweeps cost about
      // 60 vector instructions per word: some 30 ms of an R-MAT 22 C4 call.)
      static_assert(kC4Words % 4 == 0, "whole 16-byte groups of counter words");
      // (a pass in which no neighbour had a run has nothing to sweep)
      const int32_t words4 = s_any ? ((((whi - wlo) + (1 << lgc) - 1) >> lgc) + 3) >> 2 : 0;
      uint4 *cnt4 = reinterpret_cast<uint4 *>(cnt);
      for (int32_t x = threadIdx.x; x < words4; x += kC4Block) {
        const uint4 v4 = cnt4[x];
        cnt4[x] = make_uint4(0u, 0u, 0u, 0u);
        const uint32_t w[4] = {v4.x, v4.y, v4.z, v4.w};
        if (bits == 8) {
          uint32_t s = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            s += (__builtin_amdgcn_udot4(w[q], w[q], 0u, false) -
                  __builtin_amdgcn_udot4(w[q], 0x01010101u, 0u, false)) >> 1;
          acc += s;
        } else if (bits == 16) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint64_t lo = w[q] & 0xffffu, hi = w[q] >> 16;
            acc += (lo * (lo - (lo > 0)) + hi * (hi - (hi > 0))) >> 1;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint64_t c = w[q];
            acc += (c * (c - (c > 0))) >> 1;
          }
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        s_batch = 0;
        s_any = 0;
      }
      __syncthreads();
      if (bits == 32 && !advance) {
        half = 1;
      } else {
        half = 0;
        k += nidx;
      }
    }
  }
  const uint64_t sacc = wave_sum_u64(acc);
  if (lane == 0) add_u128(c4_acc, sacc);
}


// ---- 4-cycle vertex routing ----------------------------------------------------------------------
//
// Each vertex u with two or more lower neighbours is binned by (min(15, log2 nl), min(15,
// log2 S / 2)), nl = |N(u) ∩ [0, u)| and S = Σ_{v ∈ N(u), v < u} d(v): S bounds the wedges
// u-v-w (v, w < u) from above (estimated from 64 lower neighbours when nl > 64). A table of kRouteBins
// route bytes sends each bin to the light kernels (LDS hash tables of one wave, then of one
// workgroup, then the dense counters, as a vertex's wedges overflow each), to the LDS endpoint
// windows, or to the dense counters directly. The fixed split is the degree threshold (nl >=
// kC4HeavyLo to the windows); a learned table takes, per bin, the route with the least
// predicted ns per vertex from the MLP, and leaves the fixed route only for a predicted gain
// above a margin. The light kernels hold at most kC4HeavyLo - 1 lower neighbours, so a vertex
// with more that a table sends there goes to the windows.
constexpr uint8_t kC4Light = GARY_C4_LIGHT, kC4Window = GARY_C4_WINDOW, kC4Dense = GARY_C4_DENSE,
                  kC4Skip = GARY_C4_SKIP;
static_assert(GARY_C4_BINS == kRouteBins, "one route byte per bin");

// A learned table sends a bin to the dense counters only below this S bin (S < 2^24 per vertex:
// one device-memory atomic per wedge), the range its timings cover.
constexpr int kC4DenseMaxSBin = 11;

struct C4BinSums {  // per bin: vertices, and Σ of round(256 * log2 x) for x = nl, S, K
  unsigned long long n, lnl, ls, lk;
};

__device__ __forceinline__ int c4_bin(int32_t nl, unsigned long long S) {
  const int a = min(15, 31 - __clz(max(nl, 1)));
  const int b = min(15, (63 - __clzll((long long)max(S, 1ull))) >> 1);
  return a * 16 + b;
}

__host__ __device__ __forceinline__ uint8_t c4_fixed_route(int bin) {
  return (1 << (bin >> 4)) >= kC4HeavyLo ? kC4Window : kC4Light;  // nl >= 2^(bin >> 4)
}

__device__ __forceinline__ unsigned long long log2_fix8(unsigned long long x) {
  return (unsigned long long)(log2f((float)x) * 256.f + 0.5f);
}

// MLP inputs of one bin: means of log2 nl, log2 S and log2 K (K = ceil(u / kC4Win), the
// endpoint windows of u), log2 m, log2(1 + vertices in the bin), mean log2(S / nl).
__device__ __forceinline__ void c4_feature_vector(const C4BinSums &s, float log_m, float *x) {
  const float n = (float)s.n;
  const float lnl = (float)s.lnl / (256.f * n), ls = (float)s.ls / (256.f * n);
  const float lk = (float)s.lk / (256.f * n);
  x[0] = lnl / 8.f;
  x[1] = ls / 16.f;
  x[2] = lk / 8.f;
  x[3] = log_m / 16.f;
  x[4] = log2f(1.f + n) / 16.f;
  x[5] = (ls - lnl) / 16.f;
}

// cls[u] = bin of u (0 for nl < 2, which has no 4-cycle at its top) and the per-bin sums. S is
// exact for nl <= kWave, summed by the vertex's own lane; above, it is estimated as
// floor(nl * Σ_{i < 64} d(N[floor(i * nl / 64)]) / 64) from 64 evenly spaced lower neighbours,
// one wave step per vertex (a bin spans a factor of 4 in S; summing every lower arc of the hubs
// cost 3 ms on R-MAT 22).
__global__ void __launch_bounds__(256)
    c4_class_kernel(gary_device_graph g, const int32_t *__restrict__ lo_len,
                    uint8_t *__restrict__ cls, C4BinSums *__restrict__ sums) {
  __shared__ unsigned long long s_sum[kRouteBins * 4];
  for (int i = threadIdx.x; i < kRouteBins * 4; i += blockDim.x) s_sum[i] = 0ull;
  __syncthreads();
  const int lane = lane_id();
  const int32_t m = g.csr.m;
  const int32_t *__restrict__ col = g.csr.colIdx;
  const int32_t *__restrict__ deg = g.csr.deg;
  for (int64_t base = global_wave_id() * kWave; base < m; base += num_global_waves() * kWave) {
    const int64_t u = base + lane;
    const int32_t nl = u < m ? lo_len[u] : 0;
    const int64_t p = u < m ? g.csr.indPtr[u] : 0;
    unsigned long long S = 0;
    if (nl <= kWave)
      for (int32_t k = 0; k < nl; ++k) S += (unsigned)deg[col[p + k]];
    uint64_t big = __ballot(nl > kWave);
    while (big) {  // wave-uniform: one sampled step per vertex
      const int j = __builtin_ctzll(big);
      big &= big - 1;
      const int64_t pj = __shfl(p, j, kWave);
      const int32_t nj = __shfl(nl, j, kWave);
      const int32_t k = (int32_t)(((int64_t)lane * nj) / kWave);
      const uint64_t part = wave_sum_u64((unsigned)deg[col[pj + k]]);
      if (lane == j) S = (unsigned long long)(((unsigned __int128)part * (unsigned)nj) / kWave);
    }
    if (u < m) {
      if (nl >= 2) {
        const int b = c4_bin(nl, S);
        cls[u] = (uint8_t)b;
        const unsigned long long K = ((unsigned long long)u + kC4Win - 1) / kC4Win;
        atomicAdd(&s_sum[b * 4 + 0], 1ull);
        atomicAdd(&s_sum[b * 4 + 1], log2_fix8((unsigned long long)nl));
        atomicAdd(&s_sum[b * 4 + 2], log2_fix8(S));
        atomicAdd(&s_sum[b * 4 + 3], log2_fix8(max(K, 1ull)));
      } else {
        cls[u] = 0;
      }
    }
  }
  __syncthreads();
  unsigned long long *out = &sums[0].n;
  for (int i = threadIdx.x; i < kRouteBins * 4; i += blockDim.x)
    if (s_sum[i]) atomicAdd(&out[i], s_sum[i]);
}

// Features of every bin for the host (training rows): X[kRouteBins][kIn], vertices per bin.
__global__ void __launch_bounds__(kRouteBins)
    c4_features_kernel(const C4BinSums *__restrict__ sums, float log_m, float *__restrict__ X,
                       float *__restrict__ nv) {
  const int b = threadIdx.x;
  const C4BinSums s = sums[b];
  nv[b] = (float)s.n;
  float x[kIn] = {};
  if (s.n) c4_feature_vector(s, log_m, x);
  for (int i = 0; i < kIn; ++i) X[b * kIn + i] = x[i];
}

// Learned table: output o of the MLP is log2(predicted ns per vertex) of route o (light,
// window, dense). Empty bins keep the fixed route.
__global__ void __launch_bounds__(kRouteBins)
    c4_mlp_table_kernel(const float *__restrict__ P, const C4BinSums *__restrict__ sums,
                        float log_m, float margin, uint8_t *__restrict__ table) {
  __shared__ float sP[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
  const int b = threadIdx.x;
  const C4BinSums s = sums[b];
  const uint8_t fixed = c4_fixed_route(b);
  if (s.n == 0) {
    table[b] = fixed;
    return;
  }
  float x[kIn], h1[kHid], h2[kHid], y[kOut];
  c4_feature_vector(s, log_m, x);
  mlp_forward(sP, x, h1, h2, y);
  const bool light_ok = (1 << (b >> 4)) < kC4HeavyLo;
  const bool dense_ok = (b & 15) <= kC4DenseMaxSBin;
  int best = fixed;
  float by = y[fixed] + log2f(fmaxf(1.f - margin, 1e-6f));  // an alternative must beat this
  for (int o = 0; o < kOut; ++o) {
    if (o == fixed || (o == kC4Light && !light_ok) || (o == kC4Dense && !dense_ok)) continue;
    if (y[o] < by) {
      by = y[o];
      best = o;
    }
  }
  table[b] = (uint8_t)best;
}

// The light and dense vertex lists from the table (per-wave runs ascend, as in
// range_list_kernel); the window list is built by window_flag_kernel / window_list_kernel.
__global__ void __launch_bounds__(256)
    c4_route_kernel(const int32_t *__restrict__ lo_len, const uint8_t *__restrict__ cls,
                    int32_t m, const uint8_t *__restrict__ table, int32_t *__restrict__ light,
                    int32_t *__restrict__ nlight, int32_t *__restrict__ dense,
                    int32_t *__restrict__ ndense) {
  __shared__ uint8_t s_table[kRouteBins];
  for (int i = threadIdx.x; i < kRouteBins; i += blockDim.x) s_table[i] = table[i];
  __syncthreads();
  const int lane = lane_id();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < m;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = base + lane;
    const int32_t nl = u < m ? lo_len[u] : 0;
    const int r = nl >= 2 ? s_table[cls[u]] : kC4Skip;
    const bool to_light = r == kC4Light && nl < kC4HeavyLo, to_dense = r == kC4Dense;
    const uint64_t ml = __ballot(to_light), md = __ballot(to_dense);
    int32_t pl = 0, pd = 0;
    if (lane == 0) {
      if (ml) pl = atomicAdd(nlight, __popcll(ml));
      if (md) pd = atomicAdd(ndense, __popcll(md));
    }
    pl = __shfl(pl, 0, kWave);
    pd = __shfl(pd, 0, kWave);
    if (to_light) light[pl + lanes_below(ml)] = (int32_t)u;
    if (to_dense) dense[pd + lanes_below(md)] = (int32_t)u;
  }
}

// The window list in id order. f[u] = 1 when u takes the windows (table == nullptr: nl >=
// kC4HeavyLo; else its bin's route, light with nl >= kC4HeavyLo counting as windows), f[m] = 0,
// and *max_nl = the largest nl among them; an exclusive scan of f then gives each its slot
// (window_list_kernel). The list order sets which work items run together: left to the order
// of atomics, the R-MAT 22 count took 175 to 236 ms from call to call, in id order 173 ms.
__global__ void __launch_bounds__(256)
    window_flag_kernel(const int32_t *__restrict__ lo_len, const uint8_t *__restrict__ cls,
                       const uint8_t *__restrict__ table, int32_t m, uint32_t *__restrict__ f,
                       int32_t *__restrict__ max_nl) {
  int32_t mx = 0;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u <= m;
       u += (int64_t)gridDim.x * blockDim.x) {
    bool w = false;
    if (u < m) {
      const int32_t nl = lo_len[u];
      if (!table) {
        w = nl >= kC4HeavyLo;
      } else if (nl >= 2) {
        const int r = table[cls[u]];
        w = r == kC4Window || (r == kC4Light && nl >= kC4HeavyLo);
      }
      if (w) mx = max(mx, nl);
    }
    f[u] = w ? 1u : 0u;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o, kWave));
  if (lane_id() == 0 && mx > 0) atomicMax(max_nl, mx);
}

__global__ void __launch_bounds__(256)
    window_list_kernel(const uint32_t *__restrict__ off, int32_t m, int32_t *__restrict__ list,
                       int32_t *__restrict__ count) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m;
       u += (int64_t)gridDim.x * blockDim.x)
    if (off[u + 1] != off[u]) list[off[u]] = (int32_t)u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = (int32_t)off[m];
}

// ---- 4-cycle routing state (host) ------------------------------------------------------------------

struct C4Model {
  float *P = nullptr;  // device, kParams floats
  bool trained = false;
  float margin = 0.f;
};

std::mutex g_c4_mu;
std::map<int, C4Model> *g_c4_models = new std::map<int, C4Model>();

int c4_model_of(int dev, C4Model *out) {
  std::lock_guard<std::mutex> lock(g_c4_mu);
  *out = (*g_c4_models)[dev];
  return GARY_OK;
}

int c4_set_model(int dev, const float *h_params, float margin) {
  std::lock_guard<std::mutex> lock(g_c4_mu);
  C4Model &md = (*g_c4_models)[dev];
  if (!h_params) {
    md.trained = false;
    return GARY_OK;
  }
  if (!md.P) GARY_HIP(hipMalloc(&md.P, sizeof(float) * kParams));
  GARY_HIP(hipMemcpy(md.P, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice));
  md.trained = true;
  md.margin = margin;
  return GARY_OK;
}

// Zeroes the bin sums, then bins every vertex (cls) and sums the bins' features.
int c4_classify(const gary_device_graph &g, const int32_t *lo_len, uint8_t *cls, C4BinSums *csum,
                hipStream_t st) {
  GARY_HIP(hipMemsetAsync(csum, 0, sizeof(C4BinSums) * kRouteBins, st));
  const int grid = (int)std::min<int64_t>(((int64_t)g.csr.m + 255) / 256, 1024);
  c4_class_kernel<<<std::max(grid, 1), 256, 0, st>>>(g, lo_len, cls, csum);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

// Features (X, nv) or the table of a routing mode (table) for gary_hom4_c4_features /
// gary_hom4_c4_table: checks the input as gary_hom4_count does, bins the vertices, and copies
// the result back (synchronous on the null stream).
int c4_inspect(const gary_device_graph &g, int c4_routes, float *h_x, float *h_nv,
               uint8_t *h_table) {
  const int32_t m = g.csr.m;
  if (h_table)
    for (int b = 0; b < kRouteBins; ++b) h_table[b] = c4_fixed_route(b);
  if (h_x) std::fill(h_x, h_x + (size_t)kRouteBins * kIn, 0.f);
  if (h_nv) std::fill(h_nv, h_nv + kRouteBins, 0.f);
  if (m == 0 || g.csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(g.csr.indPtr && g.csr.colIdx && g.csr.deg, "gary_hom4_c4: null arrays");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  C4Model model;
  GARY_TRY(c4_model_of(dev, &model));
  if (h_table && !h_x) {
    if (c4_routes == GARY_C4_ROUTES_FIXED ||
        (c4_routes == GARY_C4_ROUTES_AUTO && !model.trained))
      return GARY_OK;
    if (!model.trained)
      return set_error(GARY_ERR_STATE, "gary_hom4_c4_table: no 4-cycle routing model on this device");
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_us = al(sizeof(int64_t) * (size_t)m), b_len = al(sizeof(int32_t) * (size_t)m);
  const size_t b_misc = 256, b_cls = al((size_t)m);
  const size_t b_csum = al(sizeof(C4BinSums) * kRouteBins);
  const size_t b_x = al(sizeof(float) * kRouteBins * (kIn + 1)), b_tab = al(kRouteBins);
  void *ws = nullptr;
  GARY_TRY(workspace_get(dev, (void *)12,
                         b_us + 2 * b_len + b_misc + b_cls + b_csum + b_x + b_tab, &ws));
  char *p = (char *)ws;
  int64_t *up_start = (int64_t *)p; p += b_us;
  int32_t *up_len = (int32_t *)p; p += b_len;
  int32_t *lo_len = (int32_t *)p; p += b_len;
  int32_t *bad = (int32_t *)p; p += b_misc;
  uint8_t *cls = (uint8_t *)p; p += b_cls;
  C4BinSums *csum = (C4BinSums *)p; p += b_csum;
  float *X = (float *)p; p += b_x;
  uint8_t *table = (uint8_t *)p;
  GARY_HIP(hipMemsetAsync(bad, 0, b_misc, nullptr));
  hom_split_kernel<<<std::min<int64_t>(((int64_t)m + 255) / 256, 4096), 256>>>(
      g, up_start, up_len, lo_len, bad, (unsigned long long *)((char *)bad + 192));
  GARY_HIP(hipGetLastError());
  GARY_TRY(c4_classify(g, lo_len, cls, csum, nullptr));
  if (h_x) {
    c4_features_kernel<<<1, kRouteBins>>>(csum, log2f((float)m), X, X + kRouteBins * kIn);
  } else {
    c4_mlp_table_kernel<<<1, kRouteBins>>>(model.P, csum, log2f((float)m), model.margin, table);
  }
  GARY_HIP(hipGetLastError());
  int32_t flag = 0;
  unsigned long long asym = 0;
  GARY_HIP(hipMemcpy(&flag, bad, sizeof flag, hipMemcpyDeviceToHost));
  GARY_HIP(hipMemcpy(&asym, (char *)bad + 192, sizeof asym, hipMemcpyDeviceToHost));
  if (flag || asym)
    return set_error(GARY_ERR_INVALID,
                     "gary_hom4_c4: the graph must be symmetric, loop-free and ranked "
                     "(ids in non-decreasing degree order: GARY_CSR_SYMMETRIC_RANKED)");
  if (h_x) {
    GARY_HIP(hipMemcpy(h_x, X, sizeof(float) * kRouteBins * kIn, hipMemcpyDeviceToHost));
    GARY_HIP(hipMemcpy(h_nv, X + kRouteBins * kIn, sizeof(float) * kRouteBins,
                       hipMemcpyDeviceToHost));
  } else {
    GARY_HIP(hipMemcpy(h_table, table, kRouteBins, hipMemcpyDeviceToHost));
  }
  return GARY_OK;
}

}  // namespace
}  // namespace gary

extern "C" {

int gary_hom4_count_routed(const gary_device_graph *g, int patterns, int c4_routes,
                           const uint8_t *h_table, gary_hom4_counts *out, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && out, "gary_hom4_count: null pointer");
  GARY_CHECK_ARG(c4_routes >= GARY_C4_ROUTES_AUTO && c4_routes <= GARY_C4_ROUTES_TABLE,
                 "gary_hom4_count: bad 4-cycle routing mode %d", c4_routes);
  if (c4_routes == GARY_C4_ROUTES_TABLE) {
    GARY_CHECK_ARG(h_table, "gary_hom4_count: GARY_C4_ROUTES_TABLE without a table");
    for (int b = 0; b < kRouteBins; ++b)
      GARY_CHECK_ARG(h_table[b] <= kC4Skip, "gary_hom4_count: bad route %d in bin %d", h_table[b], b);
  }
  GARY_CHECK_ARG((patterns & ~GARY_HOM_ALL) == 0, "gary_hom4_count: bad pattern mask %d", patterns);
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_hom4_count: bad sizes");
  *out = gary_hom4_counts{};
  const int32_t m = g->csr.m;
  if (m == 0 || g->csr.nnz == 0 || patterns == 0) return GARY_OK;
  GARY_CHECK_ARG(g->csr.indPtr && g->csr.colIdx && g->csr.deg, "gary_hom4_count: null arrays");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0;
  GARY_TRY(current_device(&dev));
  int ncu = 0;
  GARY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // hom_local_kernel: four small-pivot workgroups per CU (40 KiB of LDS each), one big-pivot
  // workgroup per CU (160 KiB). c4_kernel: one per CU; its wedge atomics saturate the memory
  // system there (half as many: 2.8 s on R-MAT 22, twice as many: 2.1 s, against 2.06 s).
  const int grid_small = std::max(1, ncu * 4);
  const int grid_big = std::max(1, ncu);
  const int grid_c4 = std::max(1, ncu);
  const int grid_win = std::max(1, ncu);  // 140 KiB of LDS each: one per CU
  const int lgrid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 4096);
  const int64_t scratch_words = local_words(kMaxHomDeg, false);
  // workspace layout (bytes, 256-aligned pieces)
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_up_start = al(sizeof(int64_t) * (size_t)m);
  const size_t b_up_len = al(sizeof(int32_t) * (size_t)m);
  const size_t b_lo_len = al(sizeof(int32_t) * (size_t)m);
  const size_t b_t = al(sizeof(uint32_t) * (size_t)g->csr.nnz);
  const size_t b_misc = 256;  // bad flag, cursors, totals, 128-bit sums
  const bool need_local = patterns & (GARY_HOM_K4 | GARY_HOM_DIAMOND);
  const bool need_c4 = patterns & GARY_HOM_C4;
  const size_t b_rows = need_local ? al(sizeof(uint32_t) * (size_t)scratch_words * grid_big) : 0;
  const size_t b_piv = need_local ? al(sizeof(int32_t) * 2 * (size_t)m) : 0;  // pivot lists
  const size_t b_cnt = need_c4 ? al(sizeof(uint32_t) * (size_t)m * grid_c4) : 0;
  const size_t b_touched = need_c4 ? al(sizeof(int32_t) * (size_t)m * grid_c4) : 0;
  // heavy and light vertex lists, and the light vertices each light kernel passes on
  const size_t b_heavy = need_c4 ? al(sizeof(int32_t) * 4 * (size_t)m) : 0;
  // 4-cycle routing by table: the model (LEARNED, or AUTO with a model), bins, sums, table
  C4Model model;
  if (need_c4 && (c4_routes == GARY_C4_ROUTES_AUTO || c4_routes == GARY_C4_ROUTES_LEARNED)) {
    GARY_TRY(c4_model_of(dev, &model));
    if (c4_routes == GARY_C4_ROUTES_LEARNED && !model.trained)
      return set_error(GARY_ERR_STATE, "gary_hom4_count: no 4-cycle routing model on this device");
  }
  const bool by_table =
      need_c4 && (c4_routes == GARY_C4_ROUTES_TABLE ||
                  (c4_routes != GARY_C4_ROUTES_FIXED && model.trained));
  const size_t b_cls = by_table ? al((size_t)m) : 0;
  const size_t b_csum = by_table ? al(sizeof(C4BinSums) * kRouteBins) : 0;
  const size_t b_tab = by_table ? al(kRouteBins) : 0;
  // window-list flags / slots (m + 1) and their scan's partial sums
  const uint32_t nflag = (uint32_t)m + 1;
  const size_t b_flag = need_c4 ? al(sizeof(uint32_t) * nflag) : 0;
  const size_t b_fpart = need_c4 ? al(sizeof(uint32_t) * exclusive_scan_parts(nflag)) : 0;
  void *ws = nullptr;
  GARY_TRY(workspace_get(dev, (void *)3,
                         b_up_start + b_up_len + b_lo_len + b_t + b_misc + b_rows + b_piv +
                             b_cnt + b_touched + b_heavy + b_cls + b_csum + b_tab + b_flag +
                             b_fpart, &ws));
  char *p = (char *)ws;
  int64_t *up_start = (int64_t *)p; p += b_up_start;
  int32_t *up_len = (int32_t *)p; p += b_up_len;
  int32_t *lo_len = (int32_t *)p; p += b_lo_len;
  uint32_t *tcount = (uint32_t *)p; p += b_t;
  char *misc = p; p += b_misc;
  uint32_t *rows = (uint32_t *)p; p += b_rows;
  int32_t *piv_small = (int32_t *)p; p += b_piv;
  int32_t *piv_big = piv_small + m;
  uint32_t *cnts = (uint32_t *)p; p += b_cnt;
  int32_t *touched = (int32_t *)p; p += b_touched;
  int32_t *heavy = (int32_t *)p;
  int32_t *light = heavy + m;
  int32_t *light_over = light + m;
  int32_t *light_over2 = light_over + m;
  p += b_heavy;
  uint8_t *cls = (uint8_t *)p; p += b_cls;
  C4BinSums *csum = (C4BinSums *)p; p += b_csum;
  uint8_t *d_table = (uint8_t *)p; p += b_tab;
  uint32_t *wflag = (uint32_t *)p; p += b_flag;
  uint32_t *wpart = (uint32_t *)p; p += b_fpart;
  // [0] bad, [1] unsupported, [2..3] cursors, [4] #heavy, [5] max lower degree of a heavy
  // vertex, [6] window-item cursor, [7] #light, [8] #small pivots, [9] #big pivots, [10] cursor,
  // [11] #light vertices with too many wedges for the wave path, [12] #those with too many for
  // the workgroup hash path, [13] cursor
  int32_t *bad = (int32_t *)misc;
  HomTotals *tot = (HomTotals *)(misc + 64);    // tri, k4
  unsigned long long *sums = (unsigned long long *)(misc + 128);  // t2 lo/hi, c2 lo/hi, c4 lo/hi
  GARY_HIP(hipMemsetAsync(misc, 0, b_misc, st));
  GARY_HIP(hipMemsetAsync(tcount, 0, b_t, st));
  if (need_c4) GARY_HIP(hipMemsetAsync(cnts, 0, b_cnt, st));
  hom_split_kernel<<<std::min<int64_t>(((int64_t)m + 255) / 256, 4096), 256, 0, st>>>(
      *g, up_start, up_len, lo_len, bad, (unsigned long long *)(misc + 192));
  GARY_HIP(hipGetLastError());
  if (need_local) {  // big pivots first, then the small ones
    constexpr int32_t dbig = first_big(kHomSmallPool, true);
    range_list_kernel<<<lgrid, 256, 0, st>>>(up_len, m, 2, dbig, piv_small, bad + 8, nullptr);
    GARY_HIP(hipGetLastError());
    range_list_kernel<<<lgrid, 256, 0, st>>>(up_len, m, dbig, INT32_MAX, piv_big, bad + 9, nullptr);
    GARY_HIP(hipGetLastError());
    hom_local_kernel<kBigWaves, kHomBigPool, true><<<grid_big, kBigWaves * kWave, 0, st>>>(
        *g, up_start, up_len, piv_big, bad + 9, tcount, bad + 10, rows, scratch_words, tot, bad + 1);
    GARY_HIP(hipGetLastError());
    hom_local_kernel<kSmallWaves, kHomSmallPool, false><<<grid_small, kSmallWaves * kWave, 0, st>>>(
        *g, up_start, up_len, piv_small, bad + 8, tcount, bad + 2, nullptr, 0, tot, bad + 1);
    GARY_HIP(hipGetLastError());
  }
  hom_sums_kernel<<<1024, 256, 0, st>>>(*g, up_start, up_len, tcount, sums, sums + 2);
  GARY_HIP(hipGetLastError());
  if (need_c4) {
    const bool use_block = m - 1 <= kBlockMaxId;  // packed keys hold every id
    if (!by_table) {
      range_list_kernel<<<lgrid, 256, 0, st>>>(lo_len, m, 2, kC4HeavyLo, light, bad + 7, nullptr);
      GARY_HIP(hipGetLastError());
    } else {
      if (c4_routes == GARY_C4_ROUTES_TABLE)  // pageable copy: h_table outlives this call's sync
        GARY_HIP(hipMemcpyAsync(d_table, h_table, kRouteBins, hipMemcpyHostToDevice, st));
      GARY_TRY(c4_classify(*g, lo_len, cls, csum, st));
      if (c4_routes != GARY_C4_ROUTES_TABLE) {
        c4_mlp_table_kernel<<<1, kRouteBins, 0, st>>>(model.P, csum, log2f((float)m),
                                                      model.margin, d_table);
        GARY_HIP(hipGetLastError());
      }
      // dense vertices join the list that the last light kernel passes to c4_kernel
      c4_route_kernel<<<lgrid, 256, 0, st>>>(lo_len, cls, m, d_table, light, bad + 7,
                                             use_block ? light_over2 : light_over,
                                             use_block ? bad + 12 : bad + 11);
      GARY_HIP(hipGetLastError());
    }
    window_flag_kernel<<<lgrid, 256, 0, st>>>(lo_len, by_table ? cls : nullptr,
                                              by_table ? d_table : nullptr, m, wflag, bad + 5);
    GARY_HIP(hipGetLastError());
    GARY_TRY(exclusive_scan_u32(wflag, nflag, wpart, st));
    window_list_kernel<<<lgrid, 256, 0, st>>>(wflag, m, heavy, bad + 4);
    GARY_HIP(hipGetLastError());
    c4_light_wave_kernel<<<std::max(1, ncu * 7), kLightWaves * kWave, 0, st>>>(
        *g, lo_len, light, bad + 7, light_over, bad + 11, sums + 4);  // 21 KiB of LDS: 7 per CU
    GARY_HIP(hipGetLastError());
    const int32_t *rest = light_over, *nrest = bad + 11;
    if (use_block) {
      c4_light_block_kernel<<<std::max(1, ncu), kBlockThreads, 0, st>>>(
          *g, lo_len, light_over, bad + 11, light_over2, bad + 12, bad + 13, sums + 4);
      GARY_HIP(hipGetLastError());
      rest = light_over2;
      nrest = bad + 12;
    }
    c4_kernel<<<grid_c4, kHomBlock, 0, st>>>(*g, lo_len, rest, nrest, cnts, touched, bad + 3,
                                              sums + 4);
    GARY_HIP(hipGetLastError());
    int32_t hv[2];  // #heavy, max lower degree: sizes the per-workgroup cursor slices
    GARY_HIP(hipMemcpyAsync(hv, bad + 4, sizeof hv, hipMemcpyDeviceToHost, st));
    GARY_HIP(hipStreamSynchronize(st));
    if (hv[0] > 0) {
      const int64_t stride = std::max<int32_t>(hv[1], 1);
      const int32_t windows = (m + kC4Win - 1) / kC4Win;
      const int32_t jmax = (windows + kC4Group - 1) / kC4Group;
      GARY_CHECK_ARG((int64_t)hv[0] * jmax < INT32_MAX, "gary_hom4_count: too many window items");
      // run index: at most one entry per element plus one per row
      GARY_CHECK_ARG((uint64_t)g->csr.nnz + (uint64_t)m + 1 < (1ull << 32) &&
                         (uint64_t)g->csr.nnz < (1ull << kRunPosBits) && windows < kRunEndWin,
                     "gary_hom4_count: 4-cycles need nnz + m below 2^32");
      GARY_CHECK_ARG(((uintptr_t)g->csr.colIdx & 15) == 0,
                     "gary_hom4_count: the 4-cycle windows read colIdx in 16-byte chunks: colIdx "
                     "must be 16-byte aligned");
      const uint32_t nptr = (uint32_t)m + 1;
      const size_t b_ptr = al(sizeof(uint32_t) * nptr);
      const size_t b_part = al(sizeof(uint32_t) * exclusive_scan_parts(nptr));
      const size_t b_runs = al(sizeof(uint64_t) * ((size_t)g->csr.nnz + (size_t)m));
      const size_t b_cur = al(sizeof(int64_t) * (size_t)stride * grid_win);
      const size_t b_nx = al(sizeof(int32_t) * (size_t)stride * grid_win);
      void *cws = nullptr;
      GARY_TRY(workspace_get(dev, (void *)8, b_ptr + b_part + b_runs + b_cur + b_nx, &cws));
      char *q = (char *)cws;
      uint32_t *run_ptr = (uint32_t *)q; q += b_ptr;
      uint32_t *part = (uint32_t *)q; q += b_part;
      uint64_t *runs = (uint64_t *)q; q += b_runs;
      int64_t *cur = (int64_t *)q; q += b_cur;
      int32_t *nxw = (int32_t *)q;
      const int rgrid = (int)std::min<int64_t>(((int64_t)m + 4) / 4, 8192);  // 4 waves per block
      run_count_kernel<<<rgrid, 256, 0, st>>>(*g, 16, run_ptr);
      GARY_HIP(hipGetLastError());
      GARY_TRY(exclusive_scan_u32(run_ptr, nptr, part, st));
      run_write_kernel<<<rgrid, 256, 0, st>>>(*g, 16, run_ptr, runs);
      GARY_HIP(hipGetLastError());
      c4_window_kernel<<<grid_win, kC4Block, 0, st>>>(*g, lo_len, heavy, bad + 4, jmax, run_ptr,
                                                      runs, cur, nxw, stride, bad + 6, sums + 4);
      GARY_HIP(hipGetLastError());
    }
  }
  int32_t flags[2];
  unsigned long long asym = 0;
  HomTotals ht;
  unsigned long long sh[6];
  GARY_HIP(hipMemcpyAsync(flags, bad, sizeof flags, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(&asym, misc + 192, sizeof asym, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(&ht, tot, sizeof ht, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(sh, sums, sizeof sh, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipStreamSynchronize(st));
  if (flags[0] || asym)
    return set_error(GARY_ERR_INVALID,
                     "gary_hom4_count: the graph must be symmetric, loop-free and ranked "
                     "(ids in non-decreasing degree order: GARY_CSR_SYMMETRIC_RANKED)");
  if (flags[1])
    return set_error(GARY_ERR_UNSUPPORTED,
                     "gary_hom4_count: a pivot with out-degree %d exceeds the local-graph limit %d",
                     flags[1], kMaxHomDeg);
  using u128 = unsigned __int128;
  const u128 lim = (u128)UINT64_MAX;
  const u128 t2 = ((u128)sh[1] << 64) | sh[0];
  const u128 c2 = ((u128)sh[3] << 64) | sh[2];
  const u128 c4 = ((u128)sh[5] << 64) | sh[4];
  const u128 diam = 2 * t2;
  const u128 hk4 = (u128)ht.k4 * 24;
  const u128 hc4 = 8 * c4 + 4 * c2 + (u128)g->csr.nnz;
  out->tri = ht.tri;
  out->k4 = ht.k4;
  if (need_local) {
    if (hk4 > lim || diam > lim)
      return set_error(GARY_ERR_OVERFLOW, "gary_hom4_count: hom(K4) or hom(diamond) exceeds 64 bits");
    out->hom_k4 = (uint64_t)hk4;
    out->hom_diamond = (uint64_t)diam;
  }
  if (need_c4) {
    if (hc4 > lim || c4 > lim)
      return set_error(GARY_ERR_OVERFLOW, "gary_hom4_count: hom(C4) exceeds 64 bits");
    out->hom_c4 = (uint64_t)hc4;
    out->c4 = (uint64_t)c4;
  }
  return GARY_OK;
  GARY_API_END
}

int gary_hom4_count(const gary_device_graph *g, int patterns, gary_hom4_counts *out,
                    void *stream) {
  return gary_hom4_count_routed(g, patterns, GARY_C4_ROUTES_AUTO, nullptr, out, stream);
}

int gary_hom4_c4_features(const gary_device_graph *g, float *h_x, float *h_vertices) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && h_x && h_vertices, "gary_hom4_c4_features: null pointer");
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_hom4_c4_features: bad sizes");
  return c4_inspect(*g, GARY_C4_ROUTES_AUTO, h_x, h_vertices, nullptr);
  GARY_API_END
}

int gary_hom4_c4_table(const gary_device_graph *g, int c4_routes, uint8_t *h_table) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && h_table, "gary_hom4_c4_table: null pointer");
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_hom4_c4_table: bad sizes");
  GARY_CHECK_ARG(c4_routes >= GARY_C4_ROUTES_AUTO && c4_routes <= GARY_C4_ROUTES_LEARNED,
                 "gary_hom4_c4_table: mode %d is not AUTO, FIXED or LEARNED", c4_routes);
  return c4_inspect(*g, c4_routes, nullptr, nullptr, h_table);
  GARY_API_END
}

int gary_hom4_c4_set_model(const float *h_params, float margin) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(margin >= 0.f && margin < 1.f, "gary_hom4_c4_set_model: margin %g outside [0, 1)",
                 (double)margin);
  if (h_params)
    for (int i = 0; i < kParams; ++i)
      GARY_CHECK_ARG(std::isfinite(h_params[i]), "gary_hom4_c4_set_model: parameter %d is not finite", i);
  int dev = 0;
  GARY_TRY(current_device(&dev));
  return c4_set_model(dev, h_params, margin);
  GARY_API_END
}

int gary_hom4_c4_load_model(const char *path) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(path, "gary_hom4_c4_load_model: null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return set_error(GARY_ERR_INVALID, "gary_hom4_c4_load_model: cannot open %s", path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string doc = ss.str();
  std::vector<double> params, margin;
  if (!selector_file_numbers(doc, "params", params) || params.size() != (size_t)kParams)
    return set_error(GARY_ERR_INVALID, "gary_hom4_c4_load_model: %s has no \"params\" array of "
                     "%d numbers", path, kParams);
  selector_file_numbers(doc, "margin", margin);
  std::vector<float> p(params.begin(), params.end());
  return gary_hom4_c4_set_model(p.data(), margin.empty() ? 0.f : (float)margin[0]);
  GARY_API_END
}

int gary_hom4_c4_model(int *trained, float *h_params, float *margin) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(trained, "gary_hom4_c4_model: null pointer");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  C4Model md;
  GARY_TRY(c4_model_of(dev, &md));
  *trained = md.trained ? 1 : 0;
  if (md.trained && h_params)
    GARY_HIP(hipMemcpy(h_params, md.P, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  if (md.trained && margin) *margin = md.margin;
  return GARY_OK;
  GARY_API_END
}

int gary_kclique_count(const gary_device_graph *g, int k, uint64_t *count, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && count, "gary_kclique_count: null pointer");
  GARY_CHECK_ARG(k >= 1 && k <= 10, "gary_kclique_count: k=%d outside [1, 10]", k);
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_kclique_count: bad sizes");
  *count = 0;
  const int32_t m = g->csr.m;
  if (k == 1) {
    *count = (uint64_t)m;
    return GARY_OK;
  }
  if (m == 0 || g->csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(g->csr.indPtr && g->csr.colIdx && g->csr.deg, "gary_kclique_count: null arrays");
  GARY_CHECK_ARG(((uintptr_t)g->csr.colIdx & 15) == 0,
                 "gary_kclique_count: the local-graph pass reads colIdx in 16-byte chunks: colIdx "
                 "must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0;
  GARY_TRY(current_device(&dev));
  int ncu = 0;
  GARY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid_small = std::max(1, ncu * 4), grid_big = std::max(1, ncu);
  const int lgrid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 4096);
  const int64_t scratch_words = local_words(kMaxHomDeg, false);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_up_start = al(sizeof(int64_t) * (size_t)m);
  const size_t b_len = al(sizeof(int32_t) * (size_t)m);
  // [0] bad, [1] unsupported, [2..3] cursors, [4] #small, [5] #big; +64: total; +192: asym
  const size_t b_misc = 256;
  const size_t b_rows = k >= 3 ? al(sizeof(uint32_t) * (size_t)scratch_words * grid_big) : 0;
  const size_t b_piv = k >= 3 ? al(sizeof(int32_t) * 2 * (size_t)m) : 0;  // pivot lists
  void *ws = nullptr;
  GARY_TRY(workspace_get(dev, (void *)3, b_up_start + 2 * b_len + b_misc + b_rows + b_piv, &ws));
  char *p = (char *)ws;
  int64_t *up_start = (int64_t *)p; p += b_up_start;
  int32_t *up_len = (int32_t *)p; p += b_len;
  int32_t *lo_len = (int32_t *)p; p += b_len;
  char *misc = p; p += b_misc;
  uint32_t *rows = (uint32_t *)p; p += b_rows;
  int32_t *piv_small = (int32_t *)p;
  int32_t *piv_big = piv_small + m;
  int32_t *bad = (int32_t *)misc;
  U128Acc *total = (U128Acc *)(misc + 64);
  GARY_HIP(hipMemsetAsync(misc, 0, b_misc, st));
  hom_split_kernel<<<std::min<int64_t>(((int64_t)m + 255) / 256, 4096), 256, 0, st>>>(
      *g, up_start, up_len, lo_len, bad, (unsigned long long *)(misc + 192));
  GARY_HIP(hipGetLastError());
  if (k >= 3) {  // big pivots first, then the small ones
    constexpr int32_t dbig = first_big(kKcSmallPool, false);
    range_list_kernel<<<lgrid, 256, 0, st>>>(up_len, m, k - 1, dbig, piv_small, bad + 4, nullptr);
    GARY_HIP(hipGetLastError());
    range_list_kernel<<<lgrid, 256, 0, st>>>(up_len, m, dbig, INT32_MAX, piv_big, bad + 5, nullptr);
    GARY_HIP(hipGetLastError());
    switch (k) {
#define GARY_KC(K)                                                                               \
  case K:                                                                                        \
    kclique_kernel<K, kBigWaves, kKcBigPool, true><<<grid_big, kBigWaves * kWave, 0, st>>>(      \
        *g, up_start, up_len, piv_big, bad + 5, bad + 3, rows, scratch_words, total, bad + 1);   \
    GARY_HIP(hipGetLastError());                                                                 \
    kclique_kernel<K, kSmallWaves, kKcSmallPool, false>                                          \
        <<<grid_small, kSmallWaves * kWave, 0, st>>>(*g, up_start, up_len, piv_small, bad + 4,   \
                                                     bad + 2, nullptr, 0, total, bad + 1);       \
    break;
      GARY_KC(3) GARY_KC(4) GARY_KC(5) GARY_KC(6) GARY_KC(7) GARY_KC(8) GARY_KC(9) GARY_KC(10)
#undef GARY_KC
    }
    GARY_HIP(hipGetLastError());
  }
  int32_t flags[2];
  unsigned long long asym = 0;
  U128Acc tot;
  GARY_HIP(hipMemcpyAsync(flags, bad, sizeof flags, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(&asym, misc + 192, sizeof asym, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipMemcpyAsync(&tot, total, sizeof tot, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipStreamSynchronize(st));
  if (flags[0] || asym)
    return set_error(GARY_ERR_INVALID,
                     "gary_kclique_count: the graph must be symmetric, loop-free and ranked "
                     "(ids in non-decreasing degree order: GARY_CSR_SYMMETRIC_RANKED)");
  if (k == 2) {
    *count = (uint64_t)g->csr.nnz / 2;
    return GARY_OK;
  }
  if (flags[1])
    return set_error(GARY_ERR_UNSUPPORTED,
                     "gary_kclique_count: a pivot with out-degree %d exceeds the local-graph limit %d",
                     flags[1], kMaxHomDeg);
  if (tot.hi) return set_error(GARY_ERR_OVERFLOW, "gary_kclique_count: the count exceeds 64 bits");
  *count = tot.lo;
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
