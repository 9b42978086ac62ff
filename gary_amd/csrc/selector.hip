// selector.hip — strategy selection for triangle counting: routed execution, the plain
// degree-threshold heuristic, and a small FP32 MLP (trained and evaluated on the device) that
// predicts the per-arc cost of each strategy from cheap features.
//
// Arcs (u,v) are binned by (min(15, log2 d+(u)), min(15, log2 d+(v))). A route table gives each
// bin one strategy (pivot bitmap, two-phase cached binary search, merge) and a predicted cost per
// arc. Counts do not depend on the table: every route computes |N+(u) ∩ N+(v)| exactly.
//
// The heuristic restates the case analysis of the reference's dispatcher comment
// (include/ops/Intersection.cuh:182-187: both lists small; one small, one large; both large but
// skewed; both large and even) with degree thresholds, and sends pivots with out-degree >= 64 to
// the bitmap (their LDS set is amortised over many arcs).
//
// The MLP (6 -> 32 -> 32 -> 3, ReLU) maps bin features to log2(ns per arc) for each strategy. It
// is trained with the kernels below (forward, per-sample backward, ordered gradient reduction,
// Adam) on per-bin timings that the library measures itself; inference is one fused kernel that
// picks the route subset with the least predicted total (leaving the all-bitmap routing only for
// a gain above a margin) and writes the route table and the predicted costs, which can also
// drive the pivot split (workgroup vs wave) by predicted cost instead of by degree.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <mutex>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "mlp.h"
#include "tc_internal.h"
#include "workspace.h"

namespace gary {

namespace {

__device__ __forceinline__ int ilog2c(int32_t x) { return min(15, 31 - __clz(max(x, 1))); }

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ---- routing --------------------------------------------------------------------------------

// Routes every arc (COO order) and compacts the cached / merge arcs into lists. Each workgroup
// owns kRouteChunk consecutive arcs: pass 1 routes them into LDS and counts the edge-route arcs,
// one global atomic per list reserves the workgroup's slots, pass 2 writes route bytes and list
// entries from LDS. With pcost != nullptr it also adds the predicted cost of bitmap arcs to their
// pivot: arcs of a vertex are contiguous, so a segmented wave scan leaves one atomic per vertex
// segment. lists: cached / merge arc ids; counts[0..1] (zeroed beforehand).
constexpr int kRouteChunk = 4096;

__global__ void __launch_bounds__(256)
    route_kernel(gary_device_graph g, const RouteTable *__restrict__ table,
                 uint8_t *__restrict__ route, float *__restrict__ pcost,
                 uint32_t *__restrict__ cached, uint32_t *__restrict__ merge,
                 int32_t *__restrict__ counts) {
  __shared__ uint8_t s_route[kRouteBins];
  __shared__ float s_cost[kRouteBins];
  __shared__ uint8_t s_r[kRouteChunk];
  __shared__ int32_t s_cnt[2], s_base[2];
  // all-bitmap tables: the pivot kernels ignore `route`, so only the predicted costs are needed
  const bool mixed = table->mixed != 0;
  if (!mixed && !pcost) return;
  for (int i = threadIdx.x; i < kRouteBins; i += blockDim.x) {
    s_route[i] = table->route[i];
    s_cost[i] = table->cost[i] > 0.f ? table->cost[i] : 1.f;
  }
  if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int32_t m = g.csr.m;
  const int lane = lane_id();
  const int64_t e0 = (int64_t)blockIdx.x * kRouteChunk, e1 = min(g.csr.nnz, e0 + kRouteChunk);
  for (int64_t base = e0 + (threadIdx.x & ~63); base < e1; base += blockDim.x) {
    const int64_t e = base + lane;
    uint8_t r = kRouteSkip;
    int32_t u = -1;
    float cst = 0.f;
    if (e < e1) {
      u = g.coo.rowIdx[e];
      const int32_t v = g.csr.colIdx[e];
      if ((uint32_t)u < (uint32_t)m && (uint32_t)v < (uint32_t)m) {
        const int32_t a = g.csr.deg[u], b = g.csr.deg[v];
        if (a >= 2 && b >= 1) {
          const int bin = ilog2c(a) * 16 + ilog2c(b);
          r = s_route[bin];
          if (r == kRouteBitmap) cst = s_cost[bin];
        }
      }
      s_r[e - e0] = r;
    }
    const uint64_t mc = __ballot(r == kRouteCached), mm = __ballot(r == kRouteMerge);
    if (lane == 0) {
      if (mc) atomicAdd(&s_cnt[0], __popcll(mc));
      if (mm) atomicAdd(&s_cnt[1], __popcll(mm));
    }
    if (pcost) {  // segmented inclusive scan over equal pivots (non-decreasing u in a wave)
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const float c2 = __shfl_up(cst, o, kWave);
        const int32_t u2 = __shfl_up(u, o, kWave);
        if (lane >= o && u2 == u) cst += c2;
      }
      const int32_t un = __shfl_down(u, 1, kWave);
      if (u >= 0 && (lane == kWave - 1 || un != u) && cst > 0.f) atomicAdd(&pcost[u], cst);
    }
  }
  if (!mixed) return;
  __syncthreads();
  if (threadIdx.x < 2) {
    s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&counts[threadIdx.x], s_cnt[threadIdx.x]) : 0;
    s_cnt[threadIdx.x] = 0;
  }
  __syncthreads();
  for (int64_t base = e0 + (threadIdx.x & ~63); base < e1; base += blockDim.x) {
    const int64_t e = base + lane;
    const uint8_t r = e < e1 ? s_r[e - e0] : kRouteSkip;
    const uint64_t mc = __ballot(r == kRouteCached), mm = __ballot(r == kRouteMerge);
    int32_t oc = 0, om = 0;
    if (lane == 0) {
      if (mc) oc = atomicAdd(&s_cnt[0], __popcll(mc));
      if (mm) om = atomicAdd(&s_cnt[1], __popcll(mm));
    }
    oc = __shfl(oc, 0, kWave);
    om = __shfl(om, 0, kWave);
    if (r == kRouteCached) cached[s_base[0] + oc + lanes_below(mc)] = (uint32_t)e;
    if (r == kRouteMerge) merge[s_base[1] + om + lanes_below(mm)] = (uint32_t)e;
    if (e < e1) route[e] = r;
  }
}

// ---- bin features (sampled) -------------------------------------------------------------------
// sums[bin] = {arcs seen, overlap samples, Σ |N+(u) ∩ N+(v)| / min(d+(u), d+(v))}. Arcs are
// sampled in runs of 64 consecutive ones (one wave, coalesced COO reads): every `hist_every`-th
// run is counted (scaled back by hist_every). Every arc whose index is a multiple of `ovl_every`
// also measures its overlap, one wave per sampled arc: up to 64 evenly spaced keys of the shorter
// list, one per lane, are binary-searched in the longer one (the overlap is the hit fraction).
struct BinSums {
  float arcs, samples, ovl;
};
// Workgroups add their partial sums into one of kSumCopies copies of the 256 bins (by workgroup
// index), so fewer of them contend for one address; readers add the copies.
constexpr int kSumCopies = 8;

__device__ __forceinline__ BinSums read_sums(const BinSums *__restrict__ sums, int b) {
  BinSums t{0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < kSumCopies; ++c) {
    const BinSums x = sums[c * kRouteBins + b];
    t.arcs += x.arcs;
    t.samples += x.samples;
    t.ovl += x.ovl;
  }
  return t;
}

__global__ void bin_features_kernel(gary_device_graph g, int32_t hist_every, int32_t ovl_every,
                                    BinSums *__restrict__ sums) {
  __shared__ float s_arcs[kRouteBins], s_samp[kRouteBins], s_ovl[kRouteBins];
  for (int i = threadIdx.x; i < kRouteBins; i += blockDim.x) s_arcs[i] = s_samp[i] = s_ovl[i] = 0.f;
  __syncthreads();
  const int32_t m = g.csr.m;
  const int64_t nnz = g.csr.nnz;
  const int lane = lane_id();
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  const int64_t wave0 = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int64_t nruns = (nnz + kWave - 1) / kWave;
  for (int64_t k = wave0 * hist_every; k < nruns; k += nwaves * hist_every) {
    const int64_t e = k * kWave + lane;
    int bin = -1;
    if (e < nnz) {
      const int32_t u = g.coo.rowIdx[e], v = g.coo.colIdx[e];
      if ((uint32_t)u < (uint32_t)m && (uint32_t)v < (uint32_t)m) {
        const int32_t a = g.csr.deg[u], b = g.csr.deg[v];
        if (a >= 2 && b >= 1) bin = ilog2c(a) * 16 + ilog2c(b);
      }
    }
    for_each_key(bin, [&](int key, uint64_t mask, int leader) {  // one LDS atomic per bin
      if (lane == leader) atomicAdd(&s_arcs[key], (float)(__popcll(mask) * hist_every));
    });
  }
  const int64_t novl = (nnz + ovl_every - 1) / ovl_every;
  for (int64_t k = wave0; k < novl; k += nwaves) {
    const int64_t e = k * ovl_every;
    const int32_t u = g.coo.rowIdx[e], v = g.coo.colIdx[e];
    if ((uint32_t)u >= (uint32_t)m || (uint32_t)v >= (uint32_t)m) continue;  // wave-uniform
    const int32_t a = g.csr.deg[u], b = g.csr.deg[v];
    if (a < 2 || b < 1) continue;
    const int bin = ilog2c(a) * 16 + ilog2c(b);
    const int32_t *A = g.csr.colIdx + g.csr.indPtr[u], *B = g.csr.colIdx + g.csr.indPtr[v];
    int32_t na = a, nb = b;
    if (na > nb) { const int32_t *t = A; A = B; B = t; int32_t tn = na; na = nb; nb = tn; }
    // up to 64 keys of the shorter list, evenly spaced, one per lane (a bounded chain of
    // dependent loads per sample)
    // The search starts from 64 splitters of the longer list held one per lane
    // (B[k*nb/64]): a search over lanes by shuffles, then a binary search of at most
    // ceil(nb/64) elements in memory.
    const int32_t ns = min(na, kWave);
    const int32_t key = lane < ns ? A[(int32_t)(((int64_t)lane * na) / ns)] : INT32_MAX;
    const int32_t sp = B[(int32_t)(((int64_t)lane * nb) / kWave)];
    int32_t ks = 0;  // last splitter <= key (0 also when key < B[0])
#pragma unroll
    for (int32_t step = kWave / 2; step > 0; step >>= 1)
      if (__shfl(sp, ks + step, kWave) <= key) ks += step;
    uint32_t hit = 0;
    if (lane < ns) {
      int32_t l = (int32_t)(((int64_t)ks * nb) / kWave);
      int32_t r = (int32_t)(((int64_t)(ks + 1) * nb) / kWave) - 1;
      if (ks == kWave - 1) r = nb - 1;
      while (l <= r) {
        const int32_t mid = (l + r) >> 1;
        const int32_t x = B[mid];
        if (x == key) { ++hit; break; }
        if (x < key) l = mid + 1;
        else r = mid - 1;
      }
    }
    hit = wave_sum_u32(hit);
    if (lane == 0) {
      atomicAdd(&s_samp[bin], 1.f);
      atomicAdd(&s_ovl[bin], (float)hit / (float)ns);
    }
  }
  __syncthreads();
  BinSums *mine = sums + (blockIdx.x % kSumCopies) * kRouteBins;
  for (int i = threadIdx.x; i < kRouteBins; i += blockDim.x) {
    if (s_arcs[i] > 0.f) atomicAdd(&mine[i].arcs, s_arcs[i]);
    if (s_samp[i] > 0.f) {
      atomicAdd(&mine[i].samples, s_samp[i]);
      atomicAdd(&mine[i].ovl, s_ovl[i]);
    }
  }
}

__device__ __forceinline__ void bin_feature_vector(int bin, const BinSums &s, float log_nnz,
                                                   float *x) {
  const float i = (float)(bin >> 4), j = (float)(bin & 15);
  x[0] = i / 16.f;
  x[1] = j / 16.f;
  x[2] = (i - j) / 16.f;
  x[3] = s.samples > 0.f ? s.ovl / s.samples : 0.f;
  x[4] = log2f(1.f + s.arcs) / 32.f;
  x[5] = log_nnz / 32.f;
}

__global__ void bin_feature_rows_kernel(const BinSums *__restrict__ sums, float log_nnz,
                                        float *__restrict__ X, float *__restrict__ arcs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= kRouteBins) return;
  BinSums s = read_sums(sums, b);
  bin_feature_vector(b, s, log_nnz, X + (int64_t)b * kIn);
  arcs[b] = s.arcs;
}

// ---- MLP -------------------------------------------------------------------------------------------

__global__ void mlp_forward_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                   int32_t n, float *__restrict__ Y) {
  __shared__ float sP[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn], h1[kHid], h2[kHid], y[kOut];
  for (int i = 0; i < kIn; ++i) x[i] = X[(int64_t)s * kIn + i];
  mlp_forward(sP, x, h1, h2, y);
  for (int o = 0; o < kOut; ++o) Y[(int64_t)s * kOut + o] = y[o];
}

// Per-sample gradient of the masked squared error 0.5 * Σ_o mask[o] (y_o - t_o)^2, written to
// G[s][kParams]; the loss terms go to L[s].
__global__ void mlp_backward_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                    const float *__restrict__ T, const float *__restrict__ M,
                                    int32_t n, float *__restrict__ G, float *__restrict__ L) {
  __shared__ float sP[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn], h1[kHid], h2[kHid], y[kOut], dy[kOut], d2[kHid], d1[kHid];
  for (int i = 0; i < kIn; ++i) x[i] = X[(int64_t)s * kIn + i];
  mlp_forward(sP, x, h1, h2, y);
  float loss = 0.f;
  for (int o = 0; o < kOut; ++o) {
    const float r = (y[o] - T[(int64_t)s * kOut + o]) * M[(int64_t)s * kOut + o];
    dy[o] = r;
    loss += 0.5f * r * r;
  }
  L[s] = loss;
  float *g = G + (int64_t)s * kParams;
  for (int o = 0; o < kOut; ++o) {
    g[oB3 + o] = dy[o];
    for (int i = 0; i < kHid; ++i) g[oW3 + o * kHid + i] = dy[o] * h2[i];
  }
  for (int i = 0; i < kHid; ++i) {
    float a = 0.f;
    for (int o = 0; o < kOut; ++o) a += sP[oW3 + o * kHid + i] * dy[o];
    d2[i] = h2[i] > 0.f ? a : 0.f;
  }
  for (int k = 0; k < kHid; ++k) {
    g[oB2 + k] = d2[k];
    for (int i = 0; i < kHid; ++i) g[oW2 + k * kHid + i] = d2[k] * h1[i];
  }
  for (int i = 0; i < kHid; ++i) {
    float a = 0.f;
    for (int k = 0; k < kHid; ++k) a += sP[oW2 + k * kHid + i] * d2[k];
    d1[i] = h1[i] > 0.f ? a : 0.f;
  }
  for (int k = 0; k < kHid; ++k) {
    g[oB1 + k] = d1[k];
    for (int i = 0; i < kIn; ++i) g[oW1 + k * kIn + i] = d1[k] * x[i];
  }
}

// grad[p] = (1/n) Σ_s G[s][p], summed in sample order (deterministic); loss = (1/n) Σ_s L[s].
__global__ void mlp_reduce_kernel(const float *__restrict__ G, const float *__restrict__ L,
                                  int32_t n, float *__restrict__ grad, float *__restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kParams) {
    float a = 0.f;
    for (int32_t s = 0; s < n; ++s) a += G[(int64_t)s * kParams + p];
    grad[p] = a / (float)n;
  }
  if (p == kParams) {
    float a = 0.f;
    for (int32_t s = 0; s < n; ++s) a += L[s];
    *loss = a / (float)n;
  }
}

// Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr m^/(sqrt(v^)+eps).
__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ grad,
                            float *__restrict__ M1, float *__restrict__ M2, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = grad[p];
  const float m = b1 * M1[p] + (1.f - b1) * g;
  const float v = b2 * M2[p] + (1.f - b2) * g * g;
  M1[p] = m;
  M2[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Fused inference: features of every bin -> MLP -> predicted ns per arc of each route; then the
// route subset R (non-empty, within `allowed`) with the least predicted total
//   Σ_{r in R} (fixed_a[r] + fixed_b[r] * nnz) + Σ_bins arcs * min_{r in R} cost(bin, r)
// is chosen, and every bin takes its cheapest route in R. One workgroup of 256 threads.
// Fixed cost (ns) of using each route, piecewise linear in log2(nnz) through up to 4 points
// (clamped at the ends).
struct FixedCosts {
  float lx[4];
  float f[4][kOut];
  int n;
};

__device__ __forceinline__ float fixed_cost(const FixedCosts &fc, float lx, int o) {
  if (fc.n <= 0) return 0.f;
  if (fc.n == 1 || lx <= fc.lx[0]) return fc.f[0][o];
  for (int k = 1; k < fc.n; ++k)
    if (lx <= fc.lx[k]) {
      const float t = (lx - fc.lx[k - 1]) / fmaxf(fc.lx[k] - fc.lx[k - 1], 1e-6f);
      return fc.f[k - 1][o] + t * (fc.f[k][o] - fc.f[k - 1][o]);
    }
  return fc.f[fc.n - 1][o];
}

__global__ void __launch_bounds__(kRouteBins)
    mlp_table_kernel(const float *__restrict__ P, const BinSums *__restrict__ sums, float log_nnz,
                     float nnz, uint32_t allowed, FixedCosts fixed, float margin,
                     RouteTable *__restrict__ table) {
  __shared__ float sP[kParams];
  __shared__ float s_tot[8];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) sP[i] = P[i];
  if (threadIdx.x < 8) s_tot[threadIdx.x] = 0.f;
  if (threadIdx.x == 0) {
    table->rowmask = 0;
    table->mixed = 0;
  }
  __syncthreads();
  const int b = threadIdx.x;
  float x[kIn], h1[kHid], h2[kHid], y[kOut], c[kOut];
  const BinSums s = read_sums(sums, b);
  bin_feature_vector(b, s, log_nnz, x);
  mlp_forward(sP, x, h1, h2, y);
  for (int o = 0; o < kOut; ++o) c[o] = exp2f(fminf(y[o], 30.f));
  for (uint32_t R = 1; R < 8; ++R) {
    if ((R & allowed) != R) continue;
    float best = INFINITY;
    for (int o = 0; o < kOut; ++o)
      if ((R >> o) & 1u) best = fminf(best, c[o]);
    float contrib = s.arcs * best;
    contrib = wave_sum_f32(contrib);
    if (lane_id() == 0) atomicAdd(&s_tot[R], contrib);
  }
  __syncthreads();
  uint32_t bestR = 0;
  float bestT = INFINITY;
  for (uint32_t R = 1; R < 8; ++R) {
    if ((R & allowed) != R) continue;
    float t = s_tot[R];
    for (int o = 0; o < kOut; ++o)
      if ((R >> o) & 1u) t += fmaxf(0.f, fixed_cost(fixed, log_nnz, o));
    if (t < bestT) {
      bestT = t;
      bestR = R;
    }
  }
  if ((allowed & 1u) && bestR != 1u) {  // leave the bitmap only for a predicted gain > margin
    const float tb = s_tot[1] + fmaxf(0.f, fixed_cost(fixed, log_nnz, 0));
    if (bestT > (1.f - margin) * tb) bestR = 1u;
  }
  int best = 0;
  float by = INFINITY;
  for (int o = 0; o < kOut; ++o)
    if (((bestR >> o) & 1u) && c[o] < by) {
      by = c[o];
      best = o;
    }
  table->route[b] = (uint8_t)best;  // output o == route code o (bitmap, cached, merge)
  table->cost[b] = by;
  const uint64_t bm = __ballot(best == 0);  // bins 64w .. 64w+63 of this wave
  if (lane_id() == 0) {
    uint32_t rows = 0;
    for (int k = 0; k < 64; ++k)
      if ((bm >> k) & 1ull) rows |= 1u << ((b + k) >> 4);
    atomicOr(&table->rowmask, rows);
    if (bm != ~0ull) atomicOr(&table->mixed, 1u);
  }
}

// ---- selector state per device -------------------------------------------------------------------

// Thresholds of the plain heuristic, as log2 out-degrees: pivots with log2 d+(u) >= bitmap_from
// go to the bitmap (the pivot's set serves >= 2^bitmap_from arcs); otherwise both lists small
// (both logs <= small_max) -> merge; one list 2^skew times longer than the other -> cached
// binary search; else merge. The defaults (6, 5, 3) are the library's plain heuristic; the last
// three cases are the reference dispatcher's (include/ops/Intersection.cuh:182-187).
struct HeuristicThresholds {
  int32_t bitmap_from = 6, small_max = 5, skew = 3;
};

struct Model : MlpTrainState {
  float fixed_lx[4] = {0.f, 0.f, 0.f, 0.f};  // fixed route costs: points (log2(1 + nnz), ns)
  float fixed_f[4][3] = {};
  int fixed_n = 0;
  bool trained = false;
  float cost_big = 0.f;  // ns; pivots above get a workgroup (0 = degree split)
  uint32_t allowed = 7;
  float margin = 0.f;  // predicted relative gain needed to route arcs away from the bitmap
};

struct DeviceState {
  RouteTable *heuristic = nullptr;  // device
  RouteTable *sync_table = nullptr; // device: output of the synchronous learned-table call
  BinSums *sums = nullptr;          // device
  Model model;
};

std::mutex g_mu;
std::map<int, DeviceState> *g_state = new std::map<int, DeviceState>();
HeuristicThresholds g_thresholds;  // of every device's heuristic table (guarded by g_mu)

void set_rowmask(RouteTable &t) {
  t.rowmask = 0;
  for (int b = 0; b < kRouteBins; ++b)
    if (t.route[b] == kRouteBitmap) t.rowmask |= 1u << (b >> 4);
  t.mixed = 0;
  for (int b = 0; b < kRouteBins; ++b)
    if (t.route[b] != kRouteBitmap) t.mixed = 1;
  t.pad[0] = t.pad[1] = 0;
}

void heuristic_table_host(RouteTable &t, const HeuristicThresholds &h = HeuristicThresholds()) {
  for (int b = 0; b < kRouteBins; ++b) {
    const int i = b >> 4, j = b & 15;  // log2 d+(u), log2 d+(v)
    uint8_t r;
    if (i >= h.bitmap_from) r = kRouteBitmap;
    else if (i <= h.small_max && j <= h.small_max) r = kRouteMerge;
    else if (j - i >= h.skew) r = kRouteCached;
    else r = kRouteMerge;
    t.route[b] = r;
    t.cost[b] = 1.f;
  }
  set_rowmask(t);
}

int state_for(int dev, DeviceState **out) {
  DeviceState &s = (*g_state)[dev];
  if (!s.heuristic) {
    RouteTable h;
    heuristic_table_host(h, g_thresholds);
    GARY_HIP(hipMalloc(&s.heuristic, sizeof(RouteTable)));
    GARY_HIP(hipMalloc(&s.sync_table, sizeof(RouteTable)));
    GARY_HIP(hipMalloc(&s.sums, sizeof(BinSums) * kRouteBins * kSumCopies));
    GARY_HIP(hipMemcpy(s.heuristic, &h, sizeof h, hipMemcpyHostToDevice));
  }
  *out = &s;
  return GARY_OK;
}

FixedCosts fixed_of(const Model &m) {
  FixedCosts f{};
  f.n = m.fixed_n;
  for (int k = 0; k < 4; ++k) {
    f.lx[k] = m.fixed_lx[k];
    for (int o = 0; o < kOut; ++o) f.f[k][o] = m.fixed_f[k][o];
  }
  return f;
}

int ensure_model(DeviceState &s) { return mlp_alloc(s.model); }

__global__ void fill_zero_kernel_sel(uint32_t *__restrict__ p, int32_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

int compute_bin_sums(const gary_device_graph &g, BinSums *sums, hipStream_t st) {
  // a fill kernel, not a memset node, so a caller may capture the learned strategy in a graph
  fill_zero_kernel_sel<<<(kRouteBins * kSumCopies * 3 + 255) / 256, 256, 0, st>>>(
      (uint32_t *)sums, kRouteBins * kSumCopies * 3);
  GARY_HIP(hipGetLastError());
  if (g.csr.nnz == 0) return GARY_OK;
  // a bounded sample whatever the graph size: at most ~16K runs of 64 arcs for the histogram
  // and ~4K arcs for the overlap (every 16th run and every 4096th arc on smaller graphs), so the
  // overlap samples take one round of the 4,096 waves (32K runs and 8K samples: 26 us on
  // R-MAT 20)
  const int64_t nruns = (g.csr.nnz + kWave - 1) / kWave;
  const int32_t hist_every = (int32_t)std::max<int64_t>(16, (nruns + 16383) / 16384);
  const int32_t ovl_every = (int32_t)std::max<int64_t>(4096, (g.csr.nnz + 4095) / 4096);
  bin_features_kernel<<<1024, 256, 0, st>>>(g, hist_every, ovl_every, sums);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
}

}  // namespace

namespace {

// Per-(device, stream) buffers of one routed count: per-arc route bytes, per-pivot predicted
// costs, the pivot lists and counters, the compacted edge-route arc lists and their counts, and
// the route table and bin sums of the call. Keeping the table and the sums here (not per device)
// lets counts on different streams of one device run at the same time.
struct RoutedWs {
  uint8_t *route = nullptr;
  float *pcost = nullptr;
  int32_t *lists = nullptr;
  uint32_t *cached = nullptr, *merge = nullptr;
  int32_t *counts = nullptr;
  RouteTable *table = nullptr;
  BinSums *sums = nullptr;
  size_t stride = 0;
  RevBufs rev;  // reversed arcs of the bitmap route (pivot path)
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int routed_ws(const gary_device_graph &g, hipStream_t st, RoutedWs &w) {
  int dev = 0;
  GARY_TRY(current_device(&dev));
  const int64_t m = std::max<int64_t>(g.csr.m, 0), nnz = std::max<int64_t>(g.csr.nnz, 0);
  w.stride = ((size_t)m + 63) & ~(size_t)63;
  const size_t b_route = align256((size_t)nnz);
  const size_t b_pcost = align256(sizeof(float) * w.stride);
  const size_t b_lists = align256(sizeof(int32_t) * (3 * w.stride + kPivotCounterWords));
  const size_t b_arcs = align256(sizeof(uint32_t) * (size_t)nnz);
  const size_t b_counts = 256;
  const size_t b_table = align256(sizeof(RouteTable));
  const size_t b_sums = align256(sizeof(BinSums) * kRouteBins * kSumCopies);
  const size_t b_rev = rev_bytes((int32_t)m, nnz);
  void *ws = nullptr;
  GARY_TRY(workspace_get(dev, (void *)((uintptr_t)st ^ 0x5E1Eull),
                         b_route + b_pcost + b_lists + 2 * b_arcs + b_counts + b_table + b_sums +
                             b_rev,
                         &ws));
  char *p = (char *)ws;
  w.route = (uint8_t *)p; p += b_route;
  w.pcost = (float *)p; p += b_pcost;
  w.lists = (int32_t *)p; p += b_lists;
  w.cached = (uint32_t *)p; p += b_arcs;
  w.merge = (uint32_t *)p; p += b_arcs;
  w.counts = (int32_t *)p; p += b_counts;
  w.table = (RouteTable *)p; p += b_table;
  w.sums = (BinSums *)p; p += b_sums;
  w.rev = rev_carve(p, (int32_t)m, nnz);
  return GARY_OK;
}

// What the host knows of a route table: which routes it uses. A table computed on the device
// (the learned one) is unknown, and every launch is made.
struct TableShape {
  bool any_bitmap = true;  // some bin goes to the pivot kernels
  bool any_edge = true;    // some bin goes to the edge-centric kernels
  bool mixed = true;       // some bin goes elsewhere than the bitmap (per-arc route bytes needed)
};

TableShape shape_of(const RouteTable &t) {
  TableShape sh;
  sh.any_bitmap = sh.any_edge = false;
  for (int b = 0; b < kRouteBins; ++b) {
    sh.any_bitmap |= t.route[b] == kRouteBitmap;
    sh.any_edge |= t.route[b] == kRouteCached || t.route[b] == kRouteMerge;
  }
  sh.mixed = t.mixed != 0;
  return sh;
}

// Counts with every arc routed by `table` (device pointer, read by the kernels of this call).
// Launches the table cannot need are left out (shape from the host copy of the table).
int launch_routed_ws(const gary_device_graph &g, const RoutedWs &w, const RouteTable *table,
                     const PivotCfg &pc, float cost_big, unsigned long long *cnt, hipStream_t st,
                     const TableShape &sh = TableShape()) {
  const int32_t m = g.csr.m;
  const bool by_cost = cost_big > 0.f && sh.any_bitmap;
  if (sh.mixed || by_cost) {
    fill_zero_kernel_sel<<<1, 64, 0, st>>>((uint32_t *)w.counts, 2);
    if (by_cost)
      fill_zero_kernel_sel<<<(int)std::min<int64_t>(((int64_t)m + 255) / 256, 4096), 256, 0,
                             st>>>((uint32_t *)w.pcost, m);
    route_kernel<<<(int)((g.csr.nnz + kRouteChunk - 1) / kRouteChunk), 256, 0, st>>>(
        g, table, w.route, by_cost ? w.pcost : nullptr, w.cached, w.merge, w.counts);
    GARY_HIP(hipGetLastError());
  }
  if (sh.any_bitmap)
    GARY_TRY(launch_pivot_core(g, pc, w.lists, w.stride, w.lists + 3 * w.stride, nullptr, nullptr,
                               0, cnt, st, true, w.route, by_cost ? w.pcost : nullptr, cost_big,
                               table,
                               rev_wanted(pc.direction, g) && w.rev.rrange ? &w.rev : nullptr));
  if (!sh.any_edge) return GARY_OK;
  return launch_edge_lists(g, w.cached, w.counts, w.merge, w.counts + 1, cnt, st);
}

int check_routed(const gary_device_graph &g) {
  GARY_CHECK_ARG(g.coo.rowIdx, "routed counting needs the COO rowIdx (call populate first)");
  GARY_CHECK_ARG(g.csr.nnz <= (int64_t)UINT32_MAX, "routed counting: more than 2^32-1 arcs");
  return GARY_OK;
}

}  // namespace

// The numbers of the value of top-level "key" in a model file (a number, or an array of
// numbers, nested arrays flattened in order). The files are the JSON objects that
// gary_amd/selector.py and gary_amd/c4_selector.py write: their numeric keys ("params",
// "margin", ...) come before any nested report, so the first occurrence of a key is the
// top-level one. Returns false when the key is absent or its value holds no number.
bool selector_file_numbers(const std::string &doc, const char *key, std::vector<double> &out) {
  out.clear();
  const std::string k = std::string("\"") + key + "\"";
  size_t at = doc.find(k);
  if (at == std::string::npos) return false;
  at = doc.find_first_not_of(" \t\r\n", at + k.size());
  if (at == std::string::npos || doc[at] != ':') return false;
  at = doc.find_first_not_of(" \t\r\n", at + 1);
  if (at == std::string::npos) return false;
  int depth = 0;
  do {
    const char c = doc[at];
    if (c == '[') {
      ++depth;
      ++at;
    } else if (c == ']') {
      --depth;
      ++at;
    } else if (c == '-' || c == '+' || c == '.' || (c >= '0' && c <= '9')) {
      char *end = nullptr;
      const double v = std::strtod(doc.c_str() + at, &end);
      if (end == doc.c_str() + at) return false;
      out.push_back(v);
      at = (size_t)(end - doc.c_str());
    } else if (c == ',' || c == ' ' || c == '\t' || c == '\r' || c == '\n') {
      ++at;
    } else {
      return false;  // not a number or an array of numbers
    }
  } while (depth > 0 && at < doc.size());
  return depth == 0 && !out.empty();
}


int mlp_alloc(MlpTrainState &m) {
  if (m.P) return GARY_OK;
  GARY_HIP(hipMalloc(&m.P, sizeof(float) * kParams));
  GARY_HIP(hipMalloc(&m.M1, sizeof(float) * kParams));
  GARY_HIP(hipMalloc(&m.M2, sizeof(float) * kParams));
  GARY_HIP(hipMalloc(&m.grad, sizeof(float) * kParams));
  GARY_HIP(hipMalloc(&m.loss, sizeof(float)));
  GARY_HIP(hipMemset(m.M1, 0, sizeof(float) * kParams));
  GARY_HIP(hipMemset(m.M2, 0, sizeof(float) * kParams));
  GARY_HIP(hipMemset(m.P, 0, sizeof(float) * kParams));
  return GARY_OK;
}

int mlp_init(MlpTrainState &s, uint64_t seed) {
  std::vector<float> p(kParams, 0.f);
  uint64_t st = seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int cnt, int fan_in) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < cnt; ++i) p[(size_t)off + i] = (float)((2.0 * next() - 1.0) * lim);
  };
  fill(oW1, kHid * kIn, kIn);
  fill(oW2, kHid * kHid, kHid);
  fill(oW3, kOut * kHid, kHid);
  GARY_HIP(hipMemcpy(s.P, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  GARY_HIP(hipMemset(s.M1, 0, sizeof(float) * kParams));
  GARY_HIP(hipMemset(s.M2, 0, sizeof(float) * kParams));
  s.steps = 0;
  return GARY_OK;
}

int mlp_train(MlpTrainState &s, const float *h_x, const float *h_t, const float *h_mask, int32_t n,
              int32_t epochs, float lr, float *h_loss) {
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t fl = (size_t)n * (kIn + 2 * kOut + kParams + 1);
  GARY_TRY(workspace_get(dev, (void *)7, sizeof(float) * fl, &ws));
  float *X = (float *)ws, *T = X + (size_t)n * kIn, *M = T + (size_t)n * kOut;
  float *G = M + (size_t)n * kOut, *L = G + (size_t)n * kParams;
  GARY_HIP(hipMemcpy(X, h_x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(T, h_t, sizeof(float) * (size_t)n * kOut, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(M, h_mask, sizeof(float) * (size_t)n * kOut, hipMemcpyHostToDevice));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int32_t ep = 0; ep < epochs; ++ep) {
    ++s.steps;
    const float c1 = 1.f - std::pow(b1, (float)s.steps), c2 = 1.f - std::pow(b2, (float)s.steps);
    mlp_backward_kernel<<<(n + 127) / 128, 128>>>(s.P, X, T, M, n, G, L);
    mlp_reduce_kernel<<<(kParams + 1 + 127) / 128, 128>>>(G, L, n, s.grad, s.loss);
    adam_kernel<<<(kParams + 127) / 128, 128>>>(s.P, s.grad, s.M1, s.M2, lr, b1, b2, eps, c1, c2);
  }
  GARY_HIP(hipGetLastError());
  mlp_backward_kernel<<<(n + 127) / 128, 128>>>(s.P, X, T, M, n, G, L);
  mlp_reduce_kernel<<<(kParams + 1 + 127) / 128, 128>>>(G, L, n, s.grad, s.loss);
  GARY_HIP(hipGetLastError());
  float loss = 0.f;
  GARY_HIP(hipMemcpy(&loss, s.loss, sizeof(float), hipMemcpyDeviceToHost));
  if (h_loss) *h_loss = loss;
  return GARY_OK;
}

int launch_routed(const gary_device_graph &g, const RouteTable *table, const PivotCfg &pc,
                  float cost_big, unsigned long long *cnt, hipStream_t st) {
  if (g.csr.m == 0 || g.csr.nnz == 0) return GARY_OK;
  GARY_TRY(check_routed(g));
  RoutedWs w;
  GARY_TRY(routed_ws(g, st, w));
  return launch_routed_ws(g, w, table, pc, cost_big, cnt, st);
}

// HEURISTIC and AUTO_LEARNED strategies of gary_tc_count. The learned table and its bin sums
// live in the stream's workspace; only the model parameters and the heuristic table are per
// device (both change only through the selector's set-up calls).
int launch_selected(const gary_device_graph &g, bool learned, const PivotCfg &pc,
                    unsigned long long *cnt, hipStream_t st) {
  if (g.csr.m == 0 || g.csr.nnz == 0) return GARY_OK;
  GARY_TRY(check_routed(g));
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  RoutedWs w;
  GARY_TRY(routed_ws(g, st, w));
  if (!learned || !s->model.trained) {
    RouteTable h;
    heuristic_table_host(h, g_thresholds);
    return launch_routed_ws(g, w, s->heuristic, pc, 0.f, cnt, st, shape_of(h));
  }
  GARY_TRY(compute_bin_sums(g, w.sums, st));
  mlp_table_kernel<<<1, kRouteBins, 0, st>>>(s->model.P, w.sums,
                                             log2f(1.f + (float)g.csr.nnz), (float)g.csr.nnz,
                                             s->model.allowed, fixed_of(s->model),
                                             s->model.margin, w.table);
  GARY_HIP(hipGetLastError());
  return launch_routed_ws(g, w, w.table, pc, s->model.cost_big, cnt, st);
}

}  // namespace gary

extern "C" {

int gary_tc_count_routed(const gary_device_graph *g, const uint8_t *h_routes,
                         const float *h_costs, float cost_big, const gary_launch_cfg *cfg,
                         uint64_t *d_count, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && h_routes && d_count, "gary_tc_count_routed: null pointer");
  for (int b = 0; b < kRouteBins; ++b)
    GARY_CHECK_ARG(h_routes[b] <= kRouteSkip, "gary_tc_count_routed: bad route %d in bin %d",
                   h_routes[b], b);
  RouteTable t;
  std::memcpy(t.route, h_routes, kRouteBins);
  for (int b = 0; b < kRouteBins; ++b) t.cost[b] = h_costs ? h_costs[b] : 1.f;
  set_rowmask(t);
  const int bs = cfg && cfg->bin_small ? cfg->bin_small : kDefaultBinSmall;
  const int bh = cfg && cfg->bin_hub ? cfg->bin_hub : 2048;
  const int grid = cfg ? cfg->grid : 0;
  hipStream_t st = (hipStream_t)stream;
  if (g->csr.m == 0 || g->csr.nnz == 0) return GARY_OK;
  GARY_TRY(check_routed(*g));
  RoutedWs w;
  GARY_TRY(routed_ws(*g, st, w));
  GARY_HIP(hipMemcpyAsync(w.table, &t, sizeof t, hipMemcpyHostToDevice, st));
  GARY_HIP(hipStreamSynchronize(st));  // `t` lives on this stack frame
  PivotCfg pc;
  pc.bin_small = bs;
  pc.bin_hub = bh;
  pc.grid = grid;
  if (cfg && cfg->block) pc.waves = cfg->block == 128 ? 2 : cfg->block == 512 ? 8 : 4;
  return launch_routed_ws(*g, w, w.table, pc, cost_big, (unsigned long long *)d_count, st,
                          shape_of(t));
  GARY_API_END
}

int gary_selector_heuristic_table(uint8_t *h_routes) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_routes, "gary_selector_heuristic_table: null pointer");
  RouteTable t;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    heuristic_table_host(t, g_thresholds);
  }
  std::memcpy(h_routes, t.route, kRouteBins);
  return GARY_OK;
  GARY_API_END
}

int gary_selector_set_heuristic(int32_t bitmap_from, int32_t small_max, int32_t skew) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(bitmap_from >= 0 && bitmap_from <= 16 && small_max >= 0 && small_max <= 16 &&
                     skew >= 1 && skew <= 16,
                 "gary_selector_set_heuristic: thresholds %d %d %d out of range", bitmap_from,
                 small_max, skew);
  std::lock_guard<std::mutex> lock(g_mu);
  g_thresholds.bitmap_from = bitmap_from;
  g_thresholds.small_max = small_max;
  g_thresholds.skew = skew;
  RouteTable t;
  heuristic_table_host(t, g_thresholds);
  if (g_state->empty()) return GARY_OK;  // device tables are built from g_thresholds later
  int cur = 0;
  GARY_TRY(current_device(&cur));
  for (auto &kv : *g_state) {
    if (!kv.second.heuristic) continue;
    GARY_HIP(hipSetDevice(kv.first));
    GARY_HIP(hipDeviceSynchronize());  // no count in flight reads the old table
    GARY_HIP(hipMemcpy(kv.second.heuristic, &t, sizeof t, hipMemcpyHostToDevice));
  }
  GARY_HIP(hipSetDevice(cur));
  return GARY_OK;
  GARY_API_END
}

int gary_selector_bin_features(const gary_device_graph *g, float *h_features, float *h_arcs) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && h_features, "gary_selector_bin_features: null pointer");
  GARY_CHECK_ARG(g->csr.nnz == 0 || g->coo.rowIdx, "bin features need the COO rowIdx (populate)");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  GARY_TRY(compute_bin_sums(*g, s->sums, nullptr));
  void *ws = nullptr;
  GARY_TRY(workspace_get(dev, (void *)4, sizeof(float) * kRouteBins * (kIn + 1), &ws));
  float *X = (float *)ws, *A = X + kRouteBins * kIn;
  bin_feature_rows_kernel<<<1, kRouteBins>>>(s->sums, log2f(1.f + (float)g->csr.nnz), X, A);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipMemcpy(h_features, X, sizeof(float) * kRouteBins * kIn, hipMemcpyDeviceToHost));
  if (h_arcs) GARY_HIP(hipMemcpy(h_arcs, A, sizeof(float) * kRouteBins, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_selector_set_params(const float *h_params, int trained, float cost_big,
                             uint32_t allowed, float margin) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_params, "gary_selector_set_params: null pointer");
  GARY_CHECK_ARG(allowed >= 1 && allowed <= 7, "gary_selector_set_params: allowed mask %u", allowed);
  GARY_CHECK_ARG(margin >= 0.f && margin < 1.f, "gary_selector_set_params: margin %g", margin);
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  GARY_TRY(ensure_model(*s));
  GARY_HIP(hipMemcpy(s->model.P, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice));
  GARY_HIP(hipMemset(s->model.M1, 0, sizeof(float) * kParams));
  GARY_HIP(hipMemset(s->model.M2, 0, sizeof(float) * kParams));
  s->model.steps = 0;
  s->model.trained = trained != 0;
  s->model.cost_big = cost_big;
  s->model.allowed = allowed;
  s->model.margin = margin;
  return GARY_OK;
  GARY_API_END
}

int gary_selector_set_fixed_costs(int32_t npoints, const float *h_nnz, const float *h_costs) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(npoints >= 0 && npoints <= 4 && (npoints == 0 || (h_nnz && h_costs)),
                 "gary_selector_set_fixed_costs: 0..4 points");
  for (int k = 1; k < npoints; ++k)
    GARY_CHECK_ARG(h_nnz[k] > h_nnz[k - 1], "gary_selector_set_fixed_costs: nnz not increasing");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  s->model.fixed_n = npoints;
  for (int k = 0; k < npoints; ++k) {
    s->model.fixed_lx[k] = log2f(1.f + h_nnz[k]);
    for (int o = 0; o < kOut; ++o) s->model.fixed_f[k][o] = h_costs[k * kOut + o];
  }
  return GARY_OK;
  GARY_API_END
}

int gary_selector_trained(int *trained) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(trained, "gary_selector_trained: null pointer");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  *trained = s->model.trained ? 1 : 0;
  return GARY_OK;
  GARY_API_END
}

int gary_selector_load(const char *path) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(path, "gary_selector_load: null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return set_error(GARY_ERR_INVALID, "gary_selector_load: cannot open %s", path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string doc = ss.str();
  std::vector<double> params, cost_big, allowed, margin, fixed_nnz, fixed_costs;
  if (!selector_file_numbers(doc, "params", params) || params.size() != (size_t)kParams)
    return set_error(GARY_ERR_INVALID, "gary_selector_load: %s has no \"params\" array of %d "
                     "numbers", path, kParams);
  selector_file_numbers(doc, "cost_big", cost_big);
  selector_file_numbers(doc, "allowed", allowed);
  selector_file_numbers(doc, "margin", margin);
  selector_file_numbers(doc, "fixed_nnz", fixed_nnz);
  selector_file_numbers(doc, "fixed_costs", fixed_costs);
  GARY_CHECK_ARG(fixed_nnz.size() <= 4 && fixed_costs.size() == fixed_nnz.size() * kOut,
                 "gary_selector_load: %s: %zu fixed-cost points with %zu costs", path,
                 fixed_nnz.size(), fixed_costs.size());
  std::vector<float> p(params.begin(), params.end());
  std::vector<float> fx(fixed_nnz.begin(), fixed_nnz.end()), fc(fixed_costs.begin(), fixed_costs.end());
  GARY_TRY(gary_selector_set_params(p.data(), 1, cost_big.empty() ? 0.f : (float)cost_big[0],
                                    allowed.empty() ? 7u : (uint32_t)allowed[0],
                                    margin.empty() ? 0.f : (float)margin[0]));
  return gary_selector_set_fixed_costs((int32_t)fx.size(), fx.empty() ? nullptr : fx.data(),
                                       fc.empty() ? nullptr : fc.data());
  GARY_API_END
}

int gary_selector_get_params(float *h_params) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_params, "gary_selector_get_params: null pointer");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  GARY_TRY(ensure_model(*s));
  GARY_HIP(hipMemcpy(h_params, s->model.P, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_mlp_forward(const float *h_params, const float *h_x, int32_t n, float *h_y) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_params && h_x && h_y && n >= 0 && n <= (1 << 20), "gary_mlp_forward: bad args");
  if (n == 0) return GARY_OK;
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t fl = (size_t)kParams + (size_t)n * (kIn + kOut);
  GARY_TRY(workspace_get(dev, (void *)5, sizeof(float) * fl, &ws));
  float *P = (float *)ws, *X = P + kParams, *Y = X + (size_t)n * kIn;
  GARY_HIP(hipMemcpy(P, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(X, h_x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice));
  mlp_forward_kernel<<<(n + 127) / 128, 128>>>(P, X, n, Y);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipMemcpy(h_y, Y, sizeof(float) * (size_t)n * kOut, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_mlp_loss_grad(const float *h_params, const float *h_x, const float *h_t,
                       const float *h_mask, int32_t n, float *h_loss, float *h_grad) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_params && h_x && h_t && h_mask && h_loss && h_grad && n >= 1 && n <= (1 << 16),
                 "gary_mlp_loss_grad: bad args");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  void *ws = nullptr;
  const size_t fl = 2 * (size_t)kParams + 1 + (size_t)n * (kIn + 2 * kOut + kParams + 1);
  GARY_TRY(workspace_get(dev, (void *)6, sizeof(float) * fl, &ws));
  float *P = (float *)ws, *grad = P + kParams, *loss = grad + kParams, *X = loss + 1;
  float *T = X + (size_t)n * kIn, *M = T + (size_t)n * kOut, *G = M + (size_t)n * kOut;
  float *L = G + (size_t)n * kParams;
  GARY_HIP(hipMemcpy(P, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(X, h_x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(T, h_t, sizeof(float) * (size_t)n * kOut, hipMemcpyHostToDevice));
  GARY_HIP(hipMemcpy(M, h_mask, sizeof(float) * (size_t)n * kOut, hipMemcpyHostToDevice));
  mlp_backward_kernel<<<(n + 127) / 128, 128>>>(P, X, T, M, n, G, L);
  mlp_reduce_kernel<<<(kParams + 1 + 127) / 128, 128>>>(G, L, n, grad, loss);
  GARY_HIP(hipGetLastError());
  GARY_HIP(hipMemcpy(h_grad, grad, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  GARY_HIP(hipMemcpy(h_loss, loss, sizeof(float), hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_selector_train(const float *h_x, const float *h_t, const float *h_mask, int32_t n,
                        int32_t epochs, float lr, uint64_t seed, float *h_loss) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_x && h_t && h_mask && n >= 1 && n <= (1 << 16) && epochs >= 0 && lr > 0.f,
                 "gary_selector_train: bad args");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  GARY_TRY(ensure_model(*s));
  if (seed) GARY_TRY(mlp_init(s->model, seed));  // else continue from the current parameters
  return mlp_train(s->model, h_x, h_t, h_mask, n, epochs, lr, h_loss);
  GARY_API_END
}

int gary_mlp_train(const float *h_x, const float *h_t, const float *h_mask, int32_t n,
                   int32_t epochs, float lr, uint64_t seed, float *h_params, float *h_loss) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(h_x && h_t && h_mask && h_params && n >= 1 && n <= (1 << 16) && epochs >= 0 &&
                     lr > 0.f && seed != 0,
                 "gary_mlp_train: bad args");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  static std::map<int, MlpTrainState> *trainers = new std::map<int, MlpTrainState>();
  MlpTrainState &t = (*trainers)[dev];
  GARY_TRY(mlp_alloc(t));
  GARY_TRY(mlp_init(t, seed));
  GARY_TRY(mlp_train(t, h_x, h_t, h_mask, n, epochs, lr, h_loss));
  GARY_HIP(hipMemcpy(h_params, t.P, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

int gary_selector_learned_table(const gary_device_graph *g, uint8_t *h_routes, float *h_costs) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && h_routes, "gary_selector_learned_table: null pointer");
  GARY_CHECK_ARG(g->csr.nnz == 0 || g->coo.rowIdx, "learned table needs the COO rowIdx (populate)");
  int dev = 0;
  GARY_TRY(current_device(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceState *s = nullptr;
  GARY_TRY(state_for(dev, &s));
  if (!s->model.trained) return set_error(GARY_ERR_STATE, "the selector has no trained model");
  GARY_TRY(compute_bin_sums(*g, s->sums, nullptr));
  mlp_table_kernel<<<1, kRouteBins>>>(s->model.P, s->sums, log2f(1.f + (float)g->csr.nnz),
                                      (float)g->csr.nnz, s->model.allowed, fixed_of(s->model),
                                      s->model.margin, s->sync_table);
  GARY_HIP(hipGetLastError());
  RouteTable t;
  GARY_HIP(hipMemcpy(&t, s->sync_table, sizeof t, hipMemcpyDeviceToHost));
  std::memcpy(h_routes, t.route, kRouteBins);
  if (h_costs) std::memcpy(h_costs, t.cost, sizeof(float) * kRouteBins);
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
