// kernels/GraphQueryHom.h — host launchers with the names of the reference's query kernels
// (include/kernels/GraphQueryHom.cuh:9-49, src/GraphQueryHom.cu:11-38).
//
// The reference launches `triangleCountHom<useCache, N, Policy><<<grid, block, smem, stream>>>(
// dataGraph, count)`; here the same template arguments pick the edge-centric strategy (one
// 64-lane wave per edge, LDS index cache of N entries, INTERLEAVED / CHUNKED edge ranges) and
// the launch goes through the C ABI. triangleCount() exposes every strategy.
#pragma once

#include <utility>

#include "../DeviceGraph.h"

namespace graph_query {

enum class Strategy : int {
  Learned = GARY_TC_AUTO_LEARNED,
  Heuristic = GARY_TC_HEURISTIC,
  Merge = GARY_TC_MERGE,
  CachedBinarySearch = GARY_TC_CACHED_BSEARCH,
  BinarySearch = GARY_TC_BSEARCH,
  Bitmap = GARY_TC_BITMAP,
};

// Elements of index cache per workgroup (getIndexCacheSMem); multiply by sizeof(vidT).
inline size_t getIndexCacheSMem(int blockDim, size_t cacheSize) {
  return (size_t)gary_index_cache_smem(blockDim, (int64_t)cacheSize);
}

// getIndexCacheSMem(dim3 blockDim, size_t cacheSize) (src/GraphQueryHom.cu:11-13): any block
// shape with x, y, z members (HIP's dim3), whose threads x * y * z make the block.
template <class Dim3, class = decltype(std::declval<const Dim3 &>().x *
                                       std::declval<const Dim3 &>().y *
                                       std::declval<const Dim3 &>().z)>
inline size_t getIndexCacheSMem(const Dim3 &blockDim, size_t cacheSize) {
  return getIndexCacheSMem((int)(blockDim.x * blockDim.y * blockDim.z), cacheSize);
}

// count += Σ over arcs (u,v) of |N+(u) ∩ N+(v)| (caller-zeroed device counter).
template <bool useCache = true, size_t N = 32, WorkPartition Policy = interleavedPolicy>
void triangleCountHom(DeviceGraph dataGraph, uint64_t *count, int gridDim = 0, int blockDim = 256,
                      void *stream = nullptr) {
  static_assert(Policy == INTERLEAVED || Policy == CHUNKED || Policy == DYNAMIC,
                "Unsupported work partitioning policy");
  gary_launch_cfg cfg{};
  cfg.block = blockDim;
  cfg.grid = gridDim;
  cfg.group = WAVE_SIZE;
  cfg.cache_depth = (int32_t)N;
  cfg.partition = Policy == CHUNKED   ? GARY_PARTITION_CHUNKED
                  : Policy == DYNAMIC ? GARY_PARTITION_DYNAMIC
                                      : GARY_PARTITION_INTERLEAVED;
  gary::check(gary_tc_count(&dataGraph, useCache ? GARY_TC_CACHED_BSEARCH : GARY_TC_BSEARCH, &cfg,
                            count, stream));
}

// Any strategy; cfg may be null (library defaults / autotuned values).
inline void triangleCount(DeviceGraph dataGraph, Strategy s, uint64_t *count,
                          const gary_launch_cfg *cfg = nullptr, void *stream = nullptr) {
  gary::check(gary_tc_count(&dataGraph, (int)s, cfg, count, stream));
}

// Loads a trained selector file (gary_amd/selector_model.json) so Strategy::Learned runs its
// model on the current device; without one, Learned runs the heuristic table.
inline void loadSelector(const std::string &path) { gary::check(gary_selector_load(path.c_str())); }

inline bool selectorTrained() {
  int t = 0;
  gary::check(gary_selector_trained(&t));
  return t != 0;
}

// triangleCountLegacy(DeviceGraph, unsigned*): 32-bit counter; throws gary::Error with
// GARY_ERR_OVERFLOW instead of wrapping past 2^32.
inline void triangleCountLegacy(DeviceGraph dataGraph, unsigned *count, void *stream = nullptr) {
  gary::check(gary_tc_count_legacy(&dataGraph, (uint32_t *)count, stream));
}

}  // namespace graph_query
