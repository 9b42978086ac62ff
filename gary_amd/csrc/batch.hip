// batch.hip — thousands of graphs counted in one persistent launch, captured in a hipGraph.
//
// The reference's only batching is one warp per item over pointer arrays
// (batchedBinarySearch{,2Phase}Kernel, test/SearchTest.cu:69-136), with a persistent kernel and
// a block-scope work queue explored separately (test/CTACommTest.cu, include/ops/DeviceQueue.cuh).
// Here a batch is the disjoint union of its graphs: one CSR with graph g owning the vertex range
// [vbase[g], vbase[g+1]). The degree orientation of a union is the union of the orientations, so
// the pivot path runs unchanged over the whole batch — its atomic cursors are the device work
// queue — and adds each wave's partial count to its pivot's graph. The whole sequence (counter
// reset by a kernel, binning, hub launch, main launch) is captured once into a hipGraph and then
// replayed with a single hipGraphLaunch per count; nothing is allocated after creation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "device_util.h"
#include "tc_internal.h"

struct gary_batch {
  int device = 0;
  int32_t ngraphs = 0, nv = 0;
  int64_t nnz = 0;
  int64_t *indPtr = nullptr;
  int32_t *colIdx = nullptr, *deg = nullptr, *gid_of = nullptr;
  int32_t *lists = nullptr;  // 3 * stride + 64
  size_t stride = 0;
  unsigned long long *gcount = nullptr;
  gary_launch_cfg cfg{};
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int use_device(int device) {
  int cur = -1;
  GARY_HIP(hipGetDevice(&cur));
  if (cur != device) GARY_HIP(hipSetDevice(device));
  return GARY_OK;
}

__global__ void batch_deg_kernel(const int64_t *__restrict__ indPtr, int32_t nv,
                                 int32_t *__restrict__ deg, int32_t *__restrict__ bad) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = indPtr[v + 1] - indPtr[v];
    if (d < 0 || d > INT32_MAX) *bad = 1;
    deg[v] = (int32_t)(d < 0 ? 0 : d);
  }
}

// gid_of[v] = g for v in [vbase[g], vbase[g+1]): one wave per graph.
__global__ void batch_gid_kernel(const int32_t *__restrict__ vbase, int32_t ngraphs,
                                 int32_t *__restrict__ gid_of) {
  for (int64_t gi = gary::global_wave_id(); gi < ngraphs; gi += gary::num_global_waves())
    for (int32_t v = vbase[gi] + gary::lane_id(); v < vbase[gi + 1]; v += gary::kWave)
      gid_of[v] = (int32_t)gi;
}

void release(gary_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (void *p : {(void *)b->indPtr, (void *)b->colIdx, (void *)b->deg, (void *)b->gid_of,
                  (void *)b->lists, (void *)b->gcount})
    if (p) (void)hipFree(p);
  *b = gary_batch{};
}

gary_device_graph view(const gary_batch *b) {
  gary_device_graph g{};
  g.csr = gary_device_csr{b->nv, b->nnz, b->indPtr, b->colIdx, b->deg};
  g.coo = gary_device_coo{b->nv, b->nnz, nullptr, b->colIdx, b->deg};
  return g;
}

int enqueue(gary_batch *b, hipStream_t st) {
  const gary_device_graph g = view(b);
  gary::PivotCfg pc;
  if (b->cfg.bin_small) pc.bin_small = b->cfg.bin_small;
  if (b->cfg.bin_hub) pc.bin_hub = b->cfg.bin_hub;
  pc.grid = b->cfg.grid;
  if (b->cfg.block) pc.waves = b->cfg.block == 128 ? 2 : b->cfg.block == 512 ? 8 : 4;
  return gary::launch_pivot_core(g, pc, b->lists, b->stride, b->lists + 3 * b->stride, b->gid_of,
                                 b->gcount, b->ngraphs, nullptr, st, false);
}

int capture(gary_batch *b) {
  if (b->exec) return GARY_OK;
  if (!b->cap) GARY_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  GARY_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue(b, b->cap);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap, &graph);
  if (rc != GARY_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return gary::set_error(GARY_ERR_HIP, "batch graph capture failed: %s", hipGetErrorString(e));
  }
  b->graph = graph;
  GARY_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return GARY_OK;
}

}  // namespace

extern "C" {

int gary_batch_create(int device, int32_t ngraphs, const int32_t *graph_vertices, int64_t nnz,
                      const int64_t *indPtr, const int32_t *colIdx, gary_batch **out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(out && graph_vertices && indPtr && (nnz == 0 || colIdx),
                 "gary_batch_create: null pointer");
  GARY_CHECK_ARG(ngraphs >= 1 && nnz >= 0, "gary_batch_create: ngraphs=%d nnz=%lld", ngraphs,
                 (long long)nnz);
  std::vector<int32_t> vbase((size_t)ngraphs + 1, 0);
  int64_t nv = 0;
  for (int32_t i = 0; i < ngraphs; ++i) {
    GARY_CHECK_ARG(graph_vertices[i] >= 0, "gary_batch_create: graph %d has %d vertices", i,
                   graph_vertices[i]);
    nv += graph_vertices[i];
    GARY_CHECK_ARG(nv <= INT32_MAX, "gary_batch_create: more than 2^31-1 vertices in total");
    vbase[(size_t)i + 1] = (int32_t)nv;
  }
  GARY_CHECK_ARG(indPtr[0] == 0 && indPtr[nv] == nnz, "gary_batch_create: indPtr[0]=%lld "
                 "indPtr[nv]=%lld, nnz=%lld", (long long)indPtr[0], (long long)indPtr[nv],
                 (long long)nnz);
  // every arc stays inside its graph's vertex range (the union is disjoint)
  for (int32_t gi = 0; gi < ngraphs; ++gi)
    for (int64_t e = indPtr[vbase[gi]]; e < indPtr[vbase[(size_t)gi + 1]]; ++e)
      GARY_CHECK_ARG(colIdx[e] >= vbase[gi] && colIdx[e] < vbase[(size_t)gi + 1],
                     "gary_batch_create: arc %lld of graph %d leaves the graph", (long long)e, gi);
  GARY_TRY(use_device(device));
  auto *b = new gary_batch();
  b->device = device;
  b->ngraphs = ngraphs;
  b->nv = (int32_t)nv;
  b->nnz = nnz;
  b->stride = ((size_t)nv + 63) & ~(size_t)63;
  int32_t *d_vbase = nullptr, *bad = nullptr;
  auto fail = [&](int rc) {
    if (d_vbase) (void)hipFree(d_vbase);
    if (bad) (void)hipFree(bad);
    release(b);
    delete b;
    return rc;
  };
#define BATCH_HIP(call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (void)hipGetLastError();                                                              \
      return fail(gary::set_error(e_ == hipErrorOutOfMemory ? GARY_ERR_OOM : GARY_ERR_HIP,   \
                                  "%s: %s", #call, hipGetErrorString(e_)));                  \
    }                                                                                       \
  } while (0)
  BATCH_HIP(hipMalloc(&b->indPtr, sizeof(int64_t) * ((size_t)nv + 1)));
  BATCH_HIP(hipMalloc(&b->colIdx, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)));
  BATCH_HIP(hipMalloc(&b->deg, sizeof(int32_t) * (size_t)std::max<int64_t>(nv, 1)));
  BATCH_HIP(hipMalloc(&b->gid_of, sizeof(int32_t) * (size_t)std::max<int64_t>(nv, 1)));
  BATCH_HIP(hipMalloc(&b->lists, sizeof(int32_t) * (3 * b->stride + gary::kPivotCounterWords)));
  BATCH_HIP(hipMalloc(&b->gcount, sizeof(unsigned long long) * (size_t)ngraphs));
  BATCH_HIP(hipMalloc(&d_vbase, sizeof(int32_t) * ((size_t)ngraphs + 1)));
  BATCH_HIP(hipMalloc(&bad, sizeof(int32_t)));
  BATCH_HIP(hipMemcpy(b->indPtr, indPtr, sizeof(int64_t) * ((size_t)nv + 1), hipMemcpyHostToDevice));
  if (nnz)
    BATCH_HIP(hipMemcpy(b->colIdx, colIdx, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice));
  BATCH_HIP(hipMemcpy(d_vbase, vbase.data(), sizeof(int32_t) * vbase.size(), hipMemcpyHostToDevice));
  BATCH_HIP(hipMemset(bad, 0, sizeof(int32_t)));
  BATCH_HIP(hipMemset(b->gcount, 0, sizeof(unsigned long long) * (size_t)ngraphs));
  if (nv > 0) {
    batch_deg_kernel<<<(int)std::min<int64_t>((nv + 255) / 256, 4096), 256>>>(b->indPtr, (int32_t)nv,
                                                                            b->deg, bad);
    batch_gid_kernel<<<(int)std::min<int64_t>((ngraphs + 3) / 4, 4096), 256>>>(d_vbase, ngraphs,
                                                                             b->gid_of);
    BATCH_HIP(hipGetLastError());
  }
  int32_t hbad = 0;
  BATCH_HIP(hipMemcpy(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost));
#undef BATCH_HIP
  (void)hipFree(d_vbase);
  (void)hipFree(bad);
  d_vbase = bad = nullptr;
  if (hbad) return fail(gary::set_error(GARY_ERR_INVALID, "gary_batch_create: indPtr not monotone"));
  *out = b;
  return GARY_OK;
  GARY_API_END
}

int gary_batch_destroy(gary_batch *b) {
  GARY_API_BEGIN
  if (!b) return GARY_OK;
  GARY_TRY(use_device(b->device));
  release(b);
  delete b;
  return GARY_OK;
  GARY_API_END
}

int gary_batch_set_cfg(gary_batch *b, const gary_launch_cfg *cfg) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(b, "gary_batch_set_cfg: null batch");
  gary_launch_cfg c = cfg ? *cfg : gary_launch_cfg{};
  const int bs = c.bin_small ? c.bin_small : gary::kDefaultBinSmall, bh = c.bin_hub ? c.bin_hub : 2048;
  const int waves = c.block == 128 ? 2 : c.block == 512 ? 8 : 4;
  GARY_CHECK_ARG(c.block == 0 || c.block == 128 || c.block == 256 || c.block == 512,
                 "gary_batch_set_cfg: block=%d (0, 128, 256 or 512)", c.block);
  GARY_CHECK_ARG(bs >= 2 && bs <= 2048 / waves && bh >= bs && bh <= 2048 && c.grid >= 0,
                 "gary_batch_set_cfg: bins small=%d hub=%d grid=%d out of range", bs, bh, c.grid);
  GARY_TRY(use_device(b->device));
  if (b->exec) {  // the captured graph bakes in the old configuration
    GARY_HIP(hipDeviceSynchronize());
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  b->cfg = c;
  return GARY_OK;
  GARY_API_END
}

int gary_batch_device_graph(gary_batch *b, gary_device_graph *out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(b && out, "gary_batch_device_graph: null pointer");
  *out = view(b);
  return GARY_OK;
  GARY_API_END
}

int gary_batch_tc_launch(gary_batch *b, int use_graph, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(b, "gary_batch_tc_launch: null batch");
  GARY_TRY(use_device(b->device));
  if (!use_graph) return enqueue(b, (hipStream_t)stream);
  GARY_TRY(capture(b));
  GARY_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return GARY_OK;
  GARY_API_END
}

int gary_batch_counts(gary_batch *b, uint64_t *h_counts, void *stream) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(b && h_counts, "gary_batch_counts: null pointer");
  GARY_TRY(use_device(b->device));
  GARY_HIP(hipMemcpyAsync(h_counts, b->gcount, sizeof(uint64_t) * (size_t)b->ngraphs,
                          hipMemcpyDeviceToHost, (hipStream_t)stream));
  GARY_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GARY_OK;
  GARY_API_END
}

int gary_batch_tc_count(gary_batch *b, uint64_t *h_counts) {
  GARY_API_BEGIN
  GARY_TRY(gary_batch_tc_launch(b, 1, nullptr));
  return gary_batch_counts(b, h_counts, nullptr);
  GARY_API_END
}

}  // extern "C"
