// host_graph.cpp — seeded graph generators and the COO -> CSR build (host side of the library).
//
// The reference builds its CSR on the host with OpenMP (symmetrize src/Graph.cpp:19-32,
// orientation :34-61, toCSR_ :63-91 with the radix sort of common/ParRadixSort.h:218-306).
// This file keeps those semantics and adds what the reference leaves out: self-loop removal,
// dedup, and orientation by (degree, id). Parallelism is std::thread (no OpenMP runtime is
// linked into the HIP library).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "common.h"
#include "prng.h"

namespace {

using gary::bounded32;
using gary::mix64;
using gary::next64;
using gary::stream_state;

// ---- parallel helpers -----------------------------------------------------------------------

int num_workers() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return (int)std::min<unsigned>(h, 16);
}

// Number of chunks parallel_chunks splits [0, n) into (small ranges run inline as one chunk).
int plan_workers(int64_t n, int workers) { return (workers <= 1 || n < 4096) ? 1 : workers; }

// Runs f(worker, begin, end) on plan_workers(n, workers) contiguous chunks of [0, n).
void parallel_chunks(int64_t n, int workers, const std::function<void(int, int64_t, int64_t)> &f) {
  workers = plan_workers(n, workers);
  if (workers == 1) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> ts;
  for (int w = 0; w < workers; ++w) {
    int64_t b = n * w / workers, e = n * (w + 1) / workers;
    ts.emplace_back([=, &f] { f(w, b, e); });
  }
  for (auto &t : ts) t.join();
}

// ---- LSD radix sort of 64-bit keys (only the low `bits` bits are significant) ----------------

void radix_sort_u64(std::vector<uint64_t> &keys, int bits, int workers) {
  const int64_t n = (int64_t)keys.size();
  if (n < 2 || bits <= 0) return;
  constexpr int D = 11, R = 1 << D;
  std::vector<uint64_t> tmp(n);
  uint64_t *in = keys.data(), *out = tmp.data();
  std::vector<int64_t> hist((size_t)workers * R);
  int passes = (bits + D - 1) / D;
  for (int p = 0; p < passes; ++p) {
    int shift = p * D;
    std::fill(hist.begin(), hist.end(), 0);
    parallel_chunks(n, workers, [&](int w, int64_t b, int64_t e) {
      int64_t *h = hist.data() + (size_t)w * R;
      for (int64_t i = b; i < e; ++i) h[(in[i] >> shift) & (R - 1)]++;
    });
    int64_t run = 0;  // digit-major, worker-minor exclusive scan keeps the sort stable
    for (int d = 0; d < R; ++d)
      for (int w = 0; w < workers; ++w) {
        int64_t c = hist[(size_t)w * R + d];
        hist[(size_t)w * R + d] = run;
        run += c;
      }
    parallel_chunks(n, workers, [&](int w, int64_t b, int64_t e) {
      int64_t *h = hist.data() + (size_t)w * R;
      for (int64_t i = b; i < e; ++i) out[h[(in[i] >> shift) & (R - 1)]++] = in[i];
    });
    std::swap(in, out);
  }
  if (in != keys.data()) std::memcpy(keys.data(), in, sizeof(uint64_t) * (size_t)n);
}

int bits_for(int64_t n) {
  int b = 0;
  while (b < 63 && (int64_t(1) << b) < n) ++b;
  return b;
}

}  // namespace

namespace gary {

// Integer cumulative Chung-Lu weights cum[0..n] (cum[0] = 0, every weight at least 1 unit):
// w_i = C (i + i0)^(-1/(gamma-1)) with i0 chosen so that w_0 = max_w while the weights sum to
// 2 * ntuples, scaled to a 2^62 total. Shared by the host and device generators.
void chung_lu_cum(int32_t n, int64_t ntuples, double gamma, double max_w,
                  std::vector<uint64_t> &cum) {
  // w_i = C (i + i0)^(-beta), beta = 1/(gamma-1), with i0 chosen so that w_0 = max_w while the
  // weights sum to 2*ntuples (bisection on i0; the sum is monotone in i0 at fixed w_0).
  const double beta = 1.0 / (gamma - 1.0), target = 2.0 * (double)ntuples;
  auto total_for = [&](double i0) {
    double cst = max_w * std::pow(i0, beta), t = 0;
    for (int32_t i = 0; i < n; ++i) t += cst * std::pow((double)i + i0, -beta);
    return t;
  };
  double lo = 1e-6, hi = (double)n * 16;
  for (int it = 0; it < 200 && hi - lo > 1e-9 * hi; ++it) {
    double mid = 0.5 * (lo + hi);
    if (total_for(mid) < target) lo = mid;
    else hi = mid;
  }
  const double i0 = 0.5 * (lo + hi), cst = max_w * std::pow(i0, beta);
  // Integer cumulative weights: sampling is integer-only from here on.
  cum.assign((size_t)n + 1, 0);
  double wsum = 0;
  for (int32_t i = 0; i < n; ++i) wsum += cst * std::pow((double)i + i0, -beta);
  const double scale = 4611686018427387904.0 / wsum;  // 2^62 / total
  double run = 0;
  for (int32_t i = 0; i < n; ++i) {
    run += cst * std::pow((double)i + i0, -beta);
    cum[(size_t)i + 1] = (uint64_t)(run * scale);
  }
  for (int32_t i = 0; i < n; ++i)  // keep every weight >= 1 unit so every vertex is drawable
    if (cum[(size_t)i + 1] <= cum[i]) cum[(size_t)i + 1] = cum[i] + 1;
}

}  // namespace gary

extern "C" {

int gary_gen_er(int32_t n, int64_t m, uint64_t seed, int32_t *src, int32_t *dst) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(n >= 0 && m >= 0, "gary_gen_er: n=%d m=%lld", n, (long long)m);
  GARY_CHECK_ARG((m == 0) || (src && dst), "gary_gen_er: null output");
  const int64_t pairs = (int64_t)n * (n - 1) / 2;
  GARY_CHECK_ARG(m <= pairs, "gary_gen_er: m=%lld exceeds C(n,2)=%lld", (long long)m,
                 (long long)pairs);
  // Rejection sampling of distinct unordered pairs in draw order (one sequential stream),
  // with an open-addressing set of canonical keys.
  size_t cap = 16;
  while (cap < (size_t)(2 * m + 2)) cap <<= 1;
  std::vector<uint64_t> table(cap, ~0ull);
  uint64_t s = mix64(seed);
  int64_t got = 0;
  while (got < m) {
    uint64_t x = next64(s);
    uint32_t u = bounded32((uint32_t)(x >> 32), (uint32_t)n);
    uint32_t v = bounded32((uint32_t)x, (uint32_t)n);
    if (u == v) continue;
    uint64_t key = u < v ? ((uint64_t)u << 32 | v) : ((uint64_t)v << 32 | u);
    size_t h = (size_t)(mix64(key) & (cap - 1));
    bool fresh = true;
    while (table[h] != ~0ull) {
      if (table[h] == key) { fresh = false; break; }
      h = (h + 1) & (cap - 1);
    }
    if (!fresh) continue;
    table[h] = key;
    src[got] = (int32_t)u;
    dst[got] = (int32_t)v;
    ++got;
  }
  return GARY_OK;
  GARY_API_END
}

int gary_gen_rmat(int32_t scale, int64_t ntuples, double a, double b, double c, uint64_t seed,
                  int32_t *src, int32_t *dst) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(scale >= 1 && scale <= 30 && ntuples >= 0, "gary_gen_rmat: scale=%d", scale);
  GARY_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "gary_gen_rmat: bad a,b,c");
  GARY_CHECK_ARG(ntuples == 0 || (src && dst), "gary_gen_rmat: null output");
  const gary::RmatThresholds t = gary::rmat_thresholds(a, b, c);
  parallel_chunks(ntuples, num_workers(), [&](int, int64_t beg, int64_t end) {
    for (int64_t i = beg; i < end; ++i) {
      uint32_t u, v;
      gary::rmat_tuple(seed, (uint64_t)i, scale, t.tA, t.tAB, t.tABC, &u, &v);
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
    }
  });
  return GARY_OK;
  GARY_API_END
}

int gary_gen_chung_lu(int32_t n, int64_t ntuples, double gamma, double max_w, uint64_t seed,
                      int32_t *src, int32_t *dst) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(n >= 2 && ntuples >= 0 && gamma > 1.0 && max_w >= 1.0,
                 "gary_gen_chung_lu: bad arguments");
  GARY_CHECK_ARG(ntuples == 0 || (src && dst), "gary_gen_chung_lu: null output");
  std::vector<uint64_t> cum;
  gary::chung_lu_cum(n, ntuples, gamma, max_w, cum);
  parallel_chunks(ntuples, num_workers(), [&](int, int64_t beg, int64_t end) {
    for (int64_t i = beg; i < end; ++i) {
      uint64_t s = stream_state(seed, (uint64_t)i);
      src[i] = gary::chung_lu_draw(cum.data(), n, next64(s));
      dst[i] = gary::chung_lu_draw(cum.data(), n, next64(s));
    }
  });
  return GARY_OK;
  GARY_API_END
}

int gary_csr_from_edges(int32_t n, int64_t ntuples, const int32_t *src, const int32_t *dst,
                        int mode, int64_t *indPtr, int32_t *colIdx, int64_t cap,
                        int64_t *nnz_out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(n >= 0 && ntuples >= 0, "gary_csr_from_edges: n=%d ntuples=%lld", n,
                 (long long)ntuples);
  GARY_CHECK_ARG(mode >= GARY_CSR_ORIENT_DEGREE && mode <= GARY_CSR_SYMMETRIC_RANKED,
                 "gary_csr_from_edges: bad mode %d", mode);
  const bool ranked = mode == GARY_CSR_SYMMETRIC_RANKED;
  if (ranked) mode = GARY_CSR_SYMMETRIC;
  GARY_CHECK_ARG(indPtr && nnz_out && (ntuples == 0 || (src && dst)),
                 "gary_csr_from_edges: null pointer");
  const int W = num_workers();
  // 1. canonical (lo, hi) keys of non-loop tuples; the key order sorts by lo then hi.
  std::vector<int64_t> kept((size_t)W + 1, 0);
  std::atomic<int> bad{0};
  parallel_chunks(ntuples, W, [&](int w, int64_t b, int64_t e) {
    int64_t c = 0;
    for (int64_t i = b; i < e; ++i) {
      int32_t u = src[i], v = dst[i];
      if (u < 0 || v < 0 || u >= n || v >= n) { bad = 1; return; }
      c += (u != v);
    }
    kept[(size_t)w + 1] = c;
  });
  GARY_CHECK_ARG(!bad, "gary_csr_from_edges: vertex id out of [0, n)");
  const int Wused = plan_workers(ntuples, W);
  for (int w = 0; w < Wused; ++w) kept[(size_t)w + 1] += kept[(size_t)w];
  std::vector<uint64_t> keys((size_t)kept[(size_t)Wused]);
  parallel_chunks(ntuples, W, [&](int w, int64_t b, int64_t e) {
    int64_t o = kept[(size_t)w];
    for (int64_t i = b; i < e; ++i) {
      uint32_t u = (uint32_t)src[i], v = (uint32_t)dst[i];
      if (u == v) continue;
      keys[(size_t)o++] = u < v ? ((uint64_t)u << 32 | v) : ((uint64_t)v << 32 | u);
    }
  });
  // 2. sort + dedup
  const int vb = bits_for(n);
  for (auto &k : keys) k = (k >> 32) << vb | (k & 0xffffffffu);  // pack to 2*vb bits
  radix_sort_u64(keys, 2 * vb, W);
  const uint64_t lomask = (vb == 0) ? 0 : ((uint64_t(1) << vb) - 1);
  int64_t m = 0;
  for (size_t i = 0; i < keys.size(); ++i)
    if (m == 0 || keys[(size_t)m - 1] != keys[i]) keys[(size_t)m++] = keys[i];
  keys.resize((size_t)m);
  const int64_t need = mode == GARY_CSR_SYMMETRIC ? 2 * m : m;
  if (need > cap || (need > 0 && !colIdx))
    return gary::set_error(GARY_ERR_INVALID, "gary_csr_from_edges: colIdx capacity %lld < %lld",
                           (long long)cap, (long long)need);
  // 3. degrees of the simple undirected graph
  std::vector<std::atomic<int64_t>> deg((size_t)n);
  for (auto &d : deg) d.store(0, std::memory_order_relaxed);
  parallel_chunks(m, W, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      deg[(size_t)(keys[(size_t)i] >> vb)].fetch_add(1, std::memory_order_relaxed);
      deg[(size_t)(keys[(size_t)i] & lomask)].fetch_add(1, std::memory_order_relaxed);
    }
  });
  // 4. arc direction and out-degrees
  auto lo_first = [&](uint32_t lo, uint32_t hi) {
    if (mode == GARY_CSR_ORIENT_ID) return false;  // keep src > dst: hi -> lo
    int64_t dl = deg[lo].load(std::memory_order_relaxed), dh = deg[hi].load(std::memory_order_relaxed);
    return dl < dh || (dl == dh && lo < hi);
  };
  std::vector<std::atomic<int64_t>> fill((size_t)n + 1);
  for (auto &f : fill) f.store(0, std::memory_order_relaxed);
  parallel_chunks(m, W, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint32_t lo = (uint32_t)(keys[(size_t)i] >> vb), hi = (uint32_t)(keys[(size_t)i] & lomask);
      if (mode == GARY_CSR_SYMMETRIC) {
        fill[lo].fetch_add(1, std::memory_order_relaxed);
        fill[hi].fetch_add(1, std::memory_order_relaxed);
      } else {
        fill[lo_first(lo, hi) ? lo : hi].fetch_add(1, std::memory_order_relaxed);
      }
    }
  });
  // ranked: new id of v = position of (deg v, v) in ascending order (counting sort by degree)
  std::vector<int32_t> newid;
  if (ranked) {
    int64_t maxd = 0;
    for (int32_t v = 0; v < n; ++v) maxd = std::max(maxd, deg[(size_t)v].load(std::memory_order_relaxed));
    std::vector<int64_t> start((size_t)maxd + 2, 0);
    for (int32_t v = 0; v < n; ++v) start[(size_t)deg[(size_t)v].load(std::memory_order_relaxed) + 1]++;
    for (int64_t k = 0; k <= maxd; ++k) start[(size_t)k + 1] += start[(size_t)k];
    newid.resize((size_t)n);
    for (int32_t v = 0; v < n; ++v)
      newid[(size_t)v] = (int32_t)start[(size_t)deg[(size_t)v].load(std::memory_order_relaxed)]++;
    // out-degree counts move with the vertices
    std::vector<int64_t> cnt((size_t)n);
    for (int32_t v = 0; v < n; ++v) cnt[(size_t)newid[(size_t)v]] = fill[(size_t)v].load(std::memory_order_relaxed);
    for (int32_t v = 0; v < n; ++v) fill[(size_t)v].store(cnt[(size_t)v], std::memory_order_relaxed);
  }
  auto relabel = [&](uint32_t x) -> uint32_t { return ranked ? (uint32_t)newid[x] : x; };
  indPtr[0] = 0;
  for (int32_t v = 0; v < n; ++v) {
    indPtr[v + 1] = indPtr[v] + fill[(size_t)v].load(std::memory_order_relaxed);
    fill[(size_t)v].store(indPtr[v], std::memory_order_relaxed);
  }
  // 5. scatter, then sort each list
  parallel_chunks(m, W, [&](int, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint32_t lo = (uint32_t)(keys[(size_t)i] >> vb), hi = (uint32_t)(keys[(size_t)i] & lomask);
      if (mode == GARY_CSR_SYMMETRIC) {
        const uint32_t a = relabel(lo), b = relabel(hi);
        colIdx[fill[a].fetch_add(1, std::memory_order_relaxed)] = (int32_t)b;
        colIdx[fill[b].fetch_add(1, std::memory_order_relaxed)] = (int32_t)a;
      } else if (lo_first(lo, hi)) {
        colIdx[fill[lo].fetch_add(1, std::memory_order_relaxed)] = (int32_t)hi;
      } else {
        colIdx[fill[hi].fetch_add(1, std::memory_order_relaxed)] = (int32_t)lo;
      }
    }
  });
  parallel_chunks(n, W, [&](int, int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      if (indPtr[v + 1] - indPtr[v] > 1) std::sort(colIdx + indPtr[v], colIdx + indPtr[v + 1]);
  });
  *nnz_out = need;
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
