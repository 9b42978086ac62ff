// tc_internal.h — constants shared by the counting kernels and their dispatchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "gary.h"

namespace gary {

// Intersection methods of the edge-centric kernels.
constexpr int kIsectBsearch = 0;  // plain binary search of the shorter list in the longer
constexpr int kIsectCached = 1;   // two-phase search with a per-group LDS index cache
constexpr int kIsectMerge = 2;    // merge-path merge of both lists

// Pivot path with caller-provided buffers (see tc_kernels.hip); gid_of/gcount select
// per-graph accumulation into gcount[ngraphs] (zeroed by the launch).
int launch_pivot_core(const gary_device_graph &g, int bin_small, int bin_hub, int grid,
                      int32_t *lists, size_t stride, int32_t *counters,
                      const int32_t *gid_of, unsigned long long *gcount, int32_t ngraphs,
                      unsigned long long *cnt, hipStream_t st, bool timed);

}  // namespace gary
