// device_search.h — sorted-list search primitives for one lane (gfx950).
//
// Restates include/ops/Search.cuh for wave64 kernels:
//   bsearch            <- binary_search(Indexable, key)           Search.cuh:48-61
//   bsearch_pos        <- binary_search(Indexable, key, int &pos) Search.cuh:67-81
//   build_cache_group  <- build_cache(list, cache, group)         Search.cuh:103-109
//   bsearch_2phase     <- binary_search_2phase(list, cache, key)  Search.cuh:120-138
//   bsearch_2phase_seg <- binary_search_2phase, phase 2 ending in a segment read
// The cache depth N is a runtime power of two here (the reference fixes N=32 at compile time,
// test/TriQueryTest.cu:42). Phase 2 halves its slice only down to at most kSegment candidates
// and then reads them in one round: the (at most three) aligned 16-byte blocks that cover them,
// issued together, instead of the last log2(kSegment) dependent loads of a binary search.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gary {

// Membership of K keys, key[k] searched in the sorted range p[k][0 .. n[k]), resolved in
// lock-step. The branchless lower bound (P.-V. Khuong and P. Morin, "Array layouts for
// comparison-based searching", 2017) takes ceil(log2 n) steps whatever the key, so the K loads of
// a step are independent and in flight together; the early exit of `bsearch` on an equal element
// gains nothing when most keys are absent. found[k] is the same answer as bsearch gives.
template <int K>
__device__ __forceinline__ void bsearch_multi(const int32_t *const (&p)[K], const int32_t (&n)[K],
                                              const int32_t (&key)[K], bool (&found)[K]) {
  const int32_t *base[K];
  int32_t len[K], most = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    base[k] = p[k];
    len[k] = n[k];
    most = max(most, len[k]);
  }
  while (most > 1) {  // invariant: the lower bound of key[k] lies in [base[k], base[k] + len[k]]
    most = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (len[k] > 1) {
        const int32_t h = len[k] >> 1;
        base[k] = base[k][h] < key[k] ? base[k] + h : base[k];
        len[k] -= h;
      }
      most = max(most, len[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    found[k] = false;
    if (n[k] > 0) {  // lower bound = base, or base + 1 when *base < key
      const int32_t v0 = base[k][0];
      const int32_t v1 = base[k] + 1 < p[k] + n[k] ? base[k][1] : v0;
      found[k] = v0 == key[k] || (v0 < key[k] && v1 == key[k]);
    }
  }
}

// Phase 2 stops halving at this many candidates (the lower bound lies among the first
// kSegment + 1 elements from `base`, which three aligned 16-byte blocks always cover).
constexpr int32_t kSegment = 8;

// Whether key is one of the elements of [base, base + kSegment] that lie inside [lo, hi): reads
// the aligned 16-byte blocks that hold base[0 .. kSegment]. Each block read holds at least one
// element of [lo, hi) (the first holds base[0], the others are read only when they start below
// hi), so no read leaves the 16-byte blocks of the list. Any element equal to key is a member:
// comparing more elements of the list than the candidates cannot give a false answer.
__device__ __forceinline__ bool segment_has(const int32_t *lo, const int32_t *hi,
                                            const int32_t *base, int32_t key) {
  const int32_t *a = reinterpret_cast<const int32_t *>((uintptr_t)base & ~(uintptr_t)15);
  int4 blk[3];
  blk[0] = *reinterpret_cast<const int4 *>(a);
  blk[1] = a + 4 < hi ? *reinterpret_cast<const int4 *>(a + 4) : make_int4(0, 0, 0, 0);
  blk[2] = a + 8 < hi ? *reinterpret_cast<const int4 *>(a + 8) : make_int4(0, 0, 0, 0);
  bool f = false;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const int32_t v[4] = {blk[b].x, blk[b].y, blk[b].z, blk[b].w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t *at = a + 4 * b + q;
      f |= at >= lo && at < hi && v[q] == key;
    }
  }
  return f;
}

// bsearch_multi with the segment finish: lock-step halving while more than kSegment candidates
// remain, then one segment read per key.
template <int K>
__device__ __forceinline__ void bsearch_multi_seg(const int32_t *const (&p)[K],
                                                  const int32_t (&n)[K], const int32_t (&key)[K],
                                                  bool (&found)[K]) {
  const int32_t *base[K];
  int32_t len[K], most = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    base[k] = p[k];
    len[k] = n[k];
    most = max(most, len[k]);
  }
  while (most > kSegment) {  // the lower bound of key[k] lies in [base[k], base[k] + len[k]]
    most = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (len[k] > kSegment) {
        const int32_t h = len[k] >> 1;
        base[k] = base[k][h] < key[k] ? base[k] + h : base[k];
        len[k] -= h;
      }
      most = max(most, len[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    found[k] = n[k] > 0 && segment_has(p[k], p[k] + n[k], base[k], key[k]);
}

// Lower bounds (number of entries < key) of K keys in an LDS array of power-of-two length n,
// resolved in lock-step.
template <int K>
__device__ __forceinline__ void lower_bound_pow2_multi(const int32_t *a, int32_t n,
                                                       const int32_t (&key)[K], int32_t (&lb)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) lb[k] = 0;
  for (int32_t half = n >> 1; half > 0; half >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) lb[k] += a[lb[k] + half - 1] < key[k] ? half : 0;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) lb[k] += a[lb[k]] < key[k];
}

// Membership of key in list[0 .. size) (Search.cuh:48-61). The early exit is kept here: the
// edge-centric strategies search short degree-oriented lists, where the branchless form
// (bsearch_multi) measured slower (heuristic strategy on R-MAT 20: 4.77 -> 4.96 ms).
__device__ __forceinline__ bool bsearch(const int32_t *__restrict__ list, int32_t size, int32_t key) {
  int32_t l = 0, r = size - 1;
  while (r >= l) {
    int32_t mid = l + ((r - l) >> 1);
    int32_t v = list[mid];
    if (v == key) return true;
    if (v < key) l = mid + 1;
    else r = mid - 1;
  }
  return false;
}

// On a miss `pos` = index of the last element < key (-1 if key < list[0]).
__device__ __forceinline__ bool bsearch_pos(const int32_t *__restrict__ list, int32_t size,
                                            int32_t key, int32_t &pos) {
  int32_t l = 0, r = size - 1;
  while (r >= l) {
    int32_t p = l + ((r - l) >> 1);
    int32_t v = list[p];
    if (v == key) {
      pos = p;
      return true;
    }
    if (v < key) l = p + 1;
    else r = p - 1;
  }
  pos = r;
  return false;
}

// Branch-light lower bound over an LDS array of power-of-two length n (returns the number of
// entries < key, i.e. the insertion index).
__device__ __forceinline__ int32_t lower_bound_pow2(const int32_t *a, int32_t n, int32_t key) {
  int32_t base = 0;
  for (int32_t half = n >> 1; half > 0; half >>= 1)
    if (a[base + half - 1] < key) base += half;
  return base + (a[base] < key);
}

// cache[i] = list[i*size/N] for i in [0, N), lanes of a group of `gsize` threads cooperating.
__device__ __forceinline__ void build_cache_group(const int32_t *__restrict__ list, int64_t size,
                                                  int32_t *cache, int32_t n_cache, int glane,
                                                  int gsize) {
  for (int32_t i = glane; i < n_cache; i += gsize) cache[i] = list[(int64_t)i * size / n_cache];
}

// bsearch_2phase with the segment finish in phase 2 (one key).
__device__ __forceinline__ bool bsearch_2phase_seg(const int32_t *__restrict__ list, int32_t size,
                                                   const int32_t *cache, int32_t n_cache,
                                                   int32_t key) {
  int32_t pos = 0;
  if (bsearch_pos(cache, n_cache, key, pos)) return true;
  if (pos < 0) return false;
  const int32_t bottom = (int32_t)((int64_t)pos * size / n_cache);
  const int32_t top = (int32_t)((int64_t)(pos + 1) * size / n_cache);
  const int32_t *p[1] = {list + bottom};
  const int32_t n[1] = {top - bottom}, k[1] = {key};
  bool f[1];
  bsearch_multi_seg<1>(p, n, k, f);
  return f[0];
}

// Phase 1 in the LDS cache, phase 2 in list[pos*size/N, (pos+1)*size/N) (Search.cuh:120-138).
__device__ __forceinline__ bool bsearch_2phase(const int32_t *__restrict__ list, int32_t size,
                                               const int32_t *cache, int32_t n_cache,
                                               int32_t key) {
  int32_t pos = 0;
  if (bsearch_pos(cache, n_cache, key, pos)) return true;
  if (pos < 0) return false;
  int32_t bottom = (int32_t)((int64_t)pos * size / n_cache);
  int32_t top = (int32_t)((int64_t)(pos + 1) * size / n_cache);
  return bsearch(list + bottom, top - bottom, key);
}

}  // namespace gary
