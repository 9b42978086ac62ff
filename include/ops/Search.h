// ops/Search.h — device search operators (include/ops/Search.cuh:33-170 of the reference).
// binary_search over one sorted device list for many keys, plain or two-phase with an index
// cache of N entries per wave (build_cache + binary_search_2phase), and the batched form.
#pragma once

#include "../DeviceGraph.h"

namespace ops {

// results[i] = keys[i] in list (binarySearchKernel / binarySearch2PhaseKernel).
inline void binary_search(const vidT *d_list, int64_t len, const vidT *d_keys, int64_t nkeys,
                          uint8_t *d_results, bool twoPhase = false, int N = 32,
                          void *stream = nullptr) {
  gary::check(gary_search(d_list, len, d_keys, nkeys, d_results,
                          twoPhase ? GARY_SEARCH_2PHASE : GARY_SEARCH_PLAIN, N, stream));
}

// *d_count += #keys found (binarySearchCountKernel / binarySearch2PhaseCountKernel).
inline void binary_search_count(const vidT *d_list, int64_t len, const vidT *d_keys,
                                int64_t nkeys, uint64_t *d_count, bool twoPhase = false,
                                int N = 32, void *stream = nullptr) {
  gary::check(gary_search_count(d_list, len, d_keys, nkeys, d_count,
                                twoPhase ? GARY_SEARCH_2PHASE : GARY_SEARCH_PLAIN, N, stream));
}

}  // namespace ops
