// ops/Intersection.h — device sorted-set intersection (include/ops/Intersection.cuh:18-359 of
// the reference): count-only (binary search, cached two-phase search, merge, bitmap) and
// intersection with sorted output.
#pragma once

#include "../DeviceGraph.h"

namespace ops {

enum class IntersectMethod : int {
  BinarySearch = GARY_ISECT_BSEARCH,
  CachedBinarySearch = GARY_ISECT_CACHED_BSEARCH,
  Merge = GARY_ISECT_MERGE,
  Bitmap = GARY_ISECT_BITMAP,
};

// *d_count += |A ∩ B|.
inline void intersect_count(const vidT *d_a, int64_t na, const vidT *d_b, int64_t nb,
                            uint64_t *d_count,
                            IntersectMethod m = IntersectMethod::CachedBinarySearch, int N = 32,
                            void *stream = nullptr) {
  gary::check(gary_intersect_count(d_a, na, d_b, nb, (int)m, N, d_count, stream));
}

// d_out[0 .. *d_count) = A ∩ B in ascending order; *d_count is overwritten.
inline void intersect(const vidT *d_a, int64_t na, const vidT *d_b, int64_t nb, vidT *d_out,
                      uint64_t *d_count,
                      IntersectMethod m = IntersectMethod::CachedBinarySearch, int N = 32,
                      void *stream = nullptr) {
  gary::check(gary_intersect_emit(d_a, na, d_b, nb, (int)m, N, d_out, d_count, stream));
}

}  // namespace ops
