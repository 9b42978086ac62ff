// device_graph.hip — DeviceGraphCtx: owner of one graph's device buffers.
//
// Reference: include/DeviceGraph.cuh:75-133, src/DeviceGraph.cu:14-96. The reference keeps
// thrust device_vectors and builds the COO row array on the host (GraphCSR::toCOO,
// src/Graph.cpp:141-155) before copying it; here the degree array and the COO rows are derived
// on the device from the uploaded row pointers, so only indPtr and colIdx cross PCIe.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "device_util.h"

struct gary_ctx {
  int device = 0;
  int32_t m = 0;
  int64_t nnz = 0;
  int64_t *indPtr = nullptr;
  int32_t *colIdx = nullptr;
  int32_t *deg = nullptr;
  int32_t *rowIdx = nullptr;
  bool loaded = false;
};

namespace {

int use_device(int device) {
  int cur = -1;
  GARY_HIP(hipGetDevice(&cur));
  if (cur != device) GARY_HIP(hipSetDevice(device));
  return GARY_OK;
}

void release(gary_ctx *c) {
  if (c->indPtr) (void)hipFree(c->indPtr);
  if (c->colIdx) (void)hipFree(c->colIdx);
  if (c->deg) (void)hipFree(c->deg);
  if (c->rowIdx) (void)hipFree(c->rowIdx);
  c->indPtr = nullptr;
  c->colIdx = nullptr;
  c->deg = nullptr;
  c->rowIdx = nullptr;
  c->loaded = false;
  c->m = 0;
  c->nnz = 0;
}

// deg[v] = indPtr[v+1] - indPtr[v]; also validates monotonicity (flag set on violation).
__global__ void derive_deg_kernel(const int64_t *__restrict__ indPtr, int32_t m,
                                  int32_t *__restrict__ deg, int32_t *__restrict__ bad) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < m;
       v += (int64_t)gridDim.x * blockDim.x) {
    int64_t d = indPtr[v + 1] - indPtr[v];
    if (d < 0 || d > INT32_MAX) {
      *bad = 1;
      d = 0;
    }
    deg[v] = (int32_t)d;
  }
}

// rowIdx[e] = v for e in [indPtr[v], indPtr[v+1]): one wave per vertex row, lanes stride the row.
__global__ void derive_rows_kernel(const int64_t *__restrict__ indPtr, int32_t m,
                                   int32_t *__restrict__ rowIdx) {
  const int64_t nw = gary::num_global_waves();
  const int lane = gary::lane_id();
  for (int64_t v = gary::global_wave_id(); v < m; v += nw) {
    int64_t b = indPtr[v], e = indPtr[v + 1];
    for (int64_t i = b + lane; i < e; i += gary::kWave) rowIdx[i] = (int32_t)v;
  }
}

}  // namespace

namespace gary {

int ctx_device(gary_ctx *ctx, int *device) {
  GARY_CHECK_ARG(ctx && device, "null ctx");
  *device = ctx->device;
  return GARY_OK;
}

// Takes ownership of device arrays built on the ctx's device (indPtr[m+1], colIdx[nnz]) and
// derives deg, as gary_ctx_load_csr does after its copies.
int ctx_adopt(gary_ctx *ctx, int32_t m, int64_t nnz, int64_t *indPtr, int32_t *colIdx) {
  GARY_TRY(use_device(ctx->device));
  release(ctx);
  ctx->indPtr = indPtr;
  ctx->colIdx = colIdx;
  GARY_HIP(hipMalloc(&ctx->deg, sizeof(int32_t) * (size_t)(m > 0 ? m : 1)));
  int32_t *bad = nullptr;
  GARY_HIP(hipMalloc(&bad, sizeof(int32_t)));
  GARY_HIP(hipMemset(bad, 0, sizeof(int32_t)));
  if (m > 0) {
    int grid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 4096);
    derive_deg_kernel<<<grid, 256>>>(ctx->indPtr, m, ctx->deg, bad);
    GARY_HIP(hipGetLastError());
  }
  GARY_HIP(hipDeviceSynchronize());
  GARY_HIP(hipFree(bad));
  ctx->m = m;
  ctx->nnz = nnz;
  ctx->loaded = true;
  return GARY_OK;
}

}  // namespace gary

extern "C" {

int gary_ctx_create(int device, gary_ctx **out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(out, "gary_ctx_create: null out");
  int n = 0;
  GARY_HIP(hipGetDeviceCount(&n));
  GARY_CHECK_ARG(device >= 0 && device < n, "gary_ctx_create: device %d of %d", device, n);
  auto *c = new gary_ctx();
  c->device = device;
  *out = c;
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_destroy(gary_ctx *ctx) {
  GARY_API_BEGIN
  if (!ctx) return GARY_OK;
  GARY_TRY(use_device(ctx->device));
  release(ctx);
  delete ctx;
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_clear(gary_ctx *ctx) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx, "gary_ctx_clear: null ctx");
  GARY_TRY(use_device(ctx->device));
  release(ctx);
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_load_csr(gary_ctx *ctx, int32_t m, int64_t nnz, const int64_t *indPtr,
                      const int32_t *colIdx) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx, "gary_ctx_load_csr: null ctx");
  GARY_CHECK_ARG(m >= 0 && nnz >= 0, "gary_ctx_load_csr: m=%d nnz=%lld", m, (long long)nnz);
  GARY_CHECK_ARG(indPtr && (nnz == 0 || colIdx), "gary_ctx_load_csr: null array");
  GARY_CHECK_ARG(indPtr[0] == 0 && indPtr[m] == nnz,
                 "gary_ctx_load_csr: indPtr[0]=%lld indPtr[m]=%lld nnz=%lld",
                 (long long)indPtr[0], (long long)indPtr[m], (long long)nnz);
  GARY_TRY(use_device(ctx->device));
  release(ctx);
  GARY_HIP(hipMalloc(&ctx->indPtr, sizeof(int64_t) * ((size_t)m + 1)));
  GARY_HIP(hipMalloc(&ctx->colIdx, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1)));
  GARY_HIP(hipMalloc(&ctx->deg, sizeof(int32_t) * (size_t)(m > 0 ? m : 1)));
  GARY_HIP(hipMemcpy(ctx->indPtr, indPtr, sizeof(int64_t) * ((size_t)m + 1),
                     hipMemcpyHostToDevice));
  if (nnz > 0)
    GARY_HIP(hipMemcpy(ctx->colIdx, colIdx, sizeof(int32_t) * (size_t)nnz,
                       hipMemcpyHostToDevice));
  int32_t *bad = nullptr;
  GARY_HIP(hipMalloc(&bad, sizeof(int32_t)));
  GARY_HIP(hipMemset(bad, 0, sizeof(int32_t)));
  if (m > 0) {
    int grid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 4096);
    derive_deg_kernel<<<grid, 256>>>(ctx->indPtr, m, ctx->deg, bad);
    GARY_HIP(hipGetLastError());
  }
  int32_t hbad = 0;
  GARY_HIP(hipMemcpy(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost));
  GARY_HIP(hipFree(bad));
  if (hbad) {
    release(ctx);
    return gary::set_error(GARY_ERR_INVALID, "gary_ctx_load_csr: indPtr is not monotone");
  }
  ctx->m = m;
  ctx->nnz = nnz;
  ctx->loaded = true;
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_populate(gary_ctx *ctx) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx, "gary_ctx_populate: null ctx");
  if (!ctx->loaded) return gary::set_error(GARY_ERR_STATE, "DeviceGraphCtx is not initialized");
  GARY_TRY(use_device(ctx->device));
  if (ctx->rowIdx) return GARY_OK;
  GARY_HIP(hipMalloc(&ctx->rowIdx, sizeof(int32_t) * (size_t)(ctx->nnz > 0 ? ctx->nnz : 1)));
  if (ctx->m > 0 && ctx->nnz > 0) {
    int grid = (int)std::min<int64_t>(((int64_t)ctx->m + 3) / 4, 8192);
    derive_rows_kernel<<<grid, 256>>>(ctx->indPtr, ctx->m, ctx->rowIdx);
    GARY_HIP(hipGetLastError());
    GARY_HIP(hipDeviceSynchronize());
  }
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_device_graph(gary_ctx *ctx, gary_device_graph *out) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx && out, "gary_ctx_device_graph: null pointer");
  if (!ctx->loaded) return gary::set_error(GARY_ERR_STATE, "DeviceGraphCtx is not initialized");
  out->csr = gary_device_csr{ctx->m, ctx->nnz, ctx->indPtr, ctx->colIdx, ctx->deg};
  out->coo = gary_device_coo{ctx->m, ctx->nnz, ctx->rowIdx, ctx->colIdx, ctx->deg};
  return GARY_OK;
  GARY_API_END
}

int gary_ctx_copy_to_host(gary_ctx *ctx, int64_t *indPtr, int32_t *colIdx, int32_t *deg) {
  GARY_API_BEGIN
  GARY_CHECK_ARG(ctx, "gary_ctx_copy_to_host: null ctx");
  if (!ctx->loaded) return gary::set_error(GARY_ERR_STATE, "DeviceGraphCtx is not initialized");
  GARY_TRY(use_device(ctx->device));
  if (indPtr)  // all m+1 row pointers (the reference copies m, src/DeviceGraph.cu:71-77)
    GARY_HIP(hipMemcpy(indPtr, ctx->indPtr, sizeof(int64_t) * ((size_t)ctx->m + 1),
                       hipMemcpyDeviceToHost));
  if (colIdx && ctx->nnz > 0)
    GARY_HIP(hipMemcpy(colIdx, ctx->colIdx, sizeof(int32_t) * (size_t)ctx->nnz,
                       hipMemcpyDeviceToHost));
  if (deg && ctx->m > 0)
    GARY_HIP(hipMemcpy(deg, ctx->deg, sizeof(int32_t) * (size_t)ctx->m, hipMemcpyDeviceToHost));
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
