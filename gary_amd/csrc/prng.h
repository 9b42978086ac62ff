// prng.h — the integer PRNG and the R-MAT quadrant walk shared by the host and device graph
// generators, so a tuple drawn on the GPU equals the one drawn on the host bit for bit.
//
// splitmix64 (Steele, Lea, Flood 2014) streams: item i of a generator seeded with `seed` starts
// from mix64(seed ^ mix64(i + 1)). Draws onto [0, n) use Lemire's multiply-shift.
#pragma once

#include <cstdint>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define GARY_HD __host__ __device__ __forceinline__
#else
#define GARY_HD inline
#endif

namespace gary {

GARY_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

GARY_HD uint64_t next64(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ull;
  return mix64(s);
}

// Independent stream for item i of a generator seeded with `seed`.
GARY_HD uint64_t stream_state(uint64_t seed, uint64_t i) { return mix64(seed ^ mix64(i + 1)); }

// Lemire's multiply-shift map of a 32-bit draw onto [0, n).
GARY_HD uint32_t bounded32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

// R-MAT quadrant thresholds on 32-bit draws: [0,tA) -> (0,0), [tA,tAB) -> (0,1),
// [tAB,tABC) -> (1,0), the rest -> (1,1).
struct RmatThresholds {
  uint64_t tA, tAB, tABC;
};

inline RmatThresholds rmat_thresholds(double a, double b, double c) {
  return {(uint64_t)(a * 4294967296.0), (uint64_t)((a + b) * 4294967296.0),
          (uint64_t)((a + b + c) * 4294967296.0)};
}

// Tuple i of an R-MAT graph of 2^scale vertices: one bit of u and v per level, two levels per
// 64-bit draw (high half first). No noise, no vertex permutation.
GARY_HD void rmat_tuple(uint64_t seed, uint64_t i, int scale, uint64_t tA, uint64_t tAB,
                        uint64_t tABC, uint32_t *u_out, uint32_t *v_out) {
  uint64_t s = stream_state(seed, i);
  uint32_t u = 0, v = 0;
  uint64_t x = 0;
  for (int lvl = 0; lvl < scale; ++lvl) {
    if ((lvl & 1) == 0) x = next64(s);
    const uint64_t r = (lvl & 1) ? (x & 0xffffffffu) : (x >> 32);
    const uint32_t ub = r >= tAB, vb = (r >= tA && r < tAB) || r >= tABC;
    u = (u << 1) | ub;
    v = (v << 1) | vb;
  }
  *u_out = u;
  *v_out = v;
}

// Chung-Lu endpoint: t = floor(x * total / 2^64) with total = cum[n], then the largest i with
// cum[i] <= t (cum: n + 1 non-decreasing integer cumulative weights, cum[0] = 0).
GARY_HD int32_t chung_lu_draw(const uint64_t *cum, int32_t n, uint64_t x) {
  const uint64_t total = cum[n];
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t t = __umul64hi(x, total);
#else
  const uint64_t t = (uint64_t)(((unsigned __int128)x * total) >> 64);
#endif
  int32_t lo = 0, hi = n + 1;  // first index with cum > t is in [lo, hi]
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (cum[mid] <= t) lo = mid + 1;
    else hi = mid;
  }
  const int32_t i = lo - 1;
  return i < n - 1 ? i : n - 1;
}

}  // namespace gary
