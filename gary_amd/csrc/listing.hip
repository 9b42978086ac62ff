// listing.hip — triangle listing: every triangle of a degree-oriented CSR written out as one
// (u, v, w) triple with v in N+(u) and w in N+(u) ∩ N+(v).
//
// This is the reference's intersection-with-output (intersectBinarySearch's emit variants,
// include/ops/Intersection.cuh:18-50, 76-112: lanes stride one list, search the other, and hits
// claim output slots) applied to every arc of the triangle query
// (include/kernels/GraphQueryHom.cuh:15-43), which the reference only counts.
//
// One wave per pivot u (atomic cursor). N+(u) is copied into the wave's LDS slice when it fits
// (up to kListLds ids), else searched in place in HBM. For each v in N+(u), lanes stride N+(v)
// and binary-search each w in N+(u). Hits are appended (ballot order) to a per-wave staging
// buffer of kStage triples in LDS; a full buffer claims one contiguous output range with a single
// atomic on the global counter and is copied out with coalesced stores. Triples are complete and
// unique; their order depends on scheduling. (Claiming output for every ballot instead, about
// 10^8 atomics on one address for R-MAT 20, took 658 ms.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "device_util.h"
#include "workspace.h"

namespace gary {
namespace {

constexpr int kListWaves = 4;
constexpr int kListBlock = kListWaves * kWave;
constexpr int kListLds = 1024;  // ids of N+(u) per wave held in LDS (4 KiB)
constexpr int kStage = 256;     // staged triples per wave (3 KiB)

__device__ __forceinline__ bool in_sorted(const int32_t *L, int32_t n, int32_t x) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (L[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && L[lo] == x;
}

// Copies the wave's `fill` staged triples to out[3*base ...) with base claimed from *total;
// triples at or beyond cap are dropped. Every lane of the wave calls it.
__device__ __forceinline__ void flush_stage(const int32_t *stage, int32_t fill,
                                            unsigned long long *total, int32_t *out,
                                            unsigned long long cap) {
  wave_sync();  // the staged triples are visible to every lane
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(total, (unsigned long long)fill);
  base = __shfl(base, 0, kWave);
  for (int32_t i = lane_id(); i < 3 * fill; i += kWave)
    if (base + (unsigned long long)(i / 3) < cap) out[3 * base + (unsigned long long)i] = stage[i];
  wave_sync();  // the stage is read before it is refilled
}

__global__ void __launch_bounds__(kListBlock)
    tri_list_kernel(gary_device_graph g, int32_t *__restrict__ cursor,
                    unsigned long long *__restrict__ total, int32_t *__restrict__ out,
                    unsigned long long cap) {
  __shared__ int32_t cache[kListWaves][kListLds];
  __shared__ int32_t staged[kListWaves][3 * kStage];
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  const int32_t m = g.csr.m;
  const int64_t *__restrict__ ptr = g.csr.indPtr;
  const int32_t *__restrict__ col = g.csr.colIdx;
  int32_t *mine = cache[wib];
  int32_t *stage = staged[wib];
  int32_t fill = 0;  // staged triples (the same in every lane)
  while (true) {
    int32_t u = 0;
    if (lane == 0) u = atomicAdd(cursor, 1);
    u = __shfl(u, 0, kWave);
    if (u >= m) break;
    const int64_t ub = ptr[u];
    const int32_t du = (int32_t)(ptr[u + 1] - ub);
    if (du < 2) continue;
    const bool lds = du <= kListLds;
    if (lds) {
      for (int32_t i = lane; i < du; i += kWave) mine[i] = col[ub + i];
      wave_sync();
    }
    const int32_t *U = lds ? mine : col + ub;
    for (int32_t i = 0; i < du; ++i) {
      const int32_t v = U[i];
      const int64_t vb = ptr[v];
      const int32_t dv = (int32_t)(ptr[v + 1] - vb);
      for (int32_t p0 = 0; p0 < dv; p0 += kWave) {
        const int32_t p = p0 + lane;
        const int32_t w = p < dv ? col[vb + p] : 0;
        const bool hit = p < dv && in_sorted(U, du, w);
        const uint64_t hits = __ballot(hit);
        if (!hits) continue;
        const int32_t c = __popcll(hits);
        if (fill + c > kStage) {
          flush_stage(stage, fill, total, out, cap);
          fill = 0;
        }
        if (hit) {
          const int32_t at = 3 * (fill + lanes_below(hits));
          stage[at] = u;
          stage[at + 1] = v;
          stage[at + 2] = w;
        }
        fill += c;
      }
    }
    wave_sync();  // the LDS copy of N+(u) is rewritten by the next pivot
  }
  if (fill > 0) flush_stage(stage, fill, total, out, cap);
}

}  // namespace
}  // namespace gary

extern "C" {

int gary_tri_list(const gary_device_graph *g, int32_t *out, uint64_t cap, uint64_t *total,
                  void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && total, "gary_tri_list: null pointer");
  GARY_CHECK_ARG(cap == 0 || out, "gary_tri_list: null output with capacity %llu",
                 (unsigned long long)cap);
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_tri_list: bad sizes");
  *total = 0;
  if (g->csr.m == 0 || g->csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(g->csr.indPtr && g->csr.colIdx, "gary_tri_list: null arrays");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, ncu = 0;
  GARY_TRY(current_device(&dev));
  GARY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  void *ws = nullptr;  // [0] pivot cursor (int32), [8] total (u64)
  GARY_TRY(workspace_get(dev, (void *)9, 256, &ws));
  GARY_HIP(hipMemsetAsync(ws, 0, 16, st));
  int32_t *cursor = (int32_t *)ws;
  auto *tot = (unsigned long long *)((char *)ws + 8);
  tri_list_kernel<<<std::max(1, ncu * 8), kListBlock, 0, st>>>(*g, cursor, tot, out,
                                                               (unsigned long long)cap);
  GARY_HIP(hipGetLastError());
  unsigned long long h = 0;
  GARY_HIP(hipMemcpyAsync(&h, tot, sizeof h, hipMemcpyDeviceToHost, st));
  GARY_HIP(hipStreamSynchronize(st));
  *total = h;
  if (h > cap)
    return set_error(GARY_ERR_OVERFLOW,
                     "gary_tri_list: %llu triangles, output capacity %llu (the first %llu written)",
                     h, (unsigned long long)cap, (unsigned long long)cap);
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
