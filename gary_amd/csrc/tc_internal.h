// tc_internal.h — constants shared by the counting kernels and their dispatchers.
#pragma once

namespace gary {

// Intersection methods of the edge-centric kernels.
constexpr int kIsectBsearch = 0;  // plain binary search of the shorter list in the longer
constexpr int kIsectCached = 1;   // two-phase search with a per-group LDS index cache
constexpr int kIsectMerge = 2;    // merge-path merge of both lists

}  // namespace gary
