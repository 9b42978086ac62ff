// noalloc.hip — triangle counting without any allocation: the pivot-bitmap path of gary_tc_count
// (strategy GARY_TC_BITMAP) with its pivot lists and counters in a fixed device array.
//
// For callers that cannot allocate device memory (code under stream capture, tight serving
// loops, hosts that forbid hipMalloc). The array `scratch` is a file-scope device variable: each
// GPU holds its own copy for the life of the process. Its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that address is handed to the
// kernels; the host-side name of the array is not a device address.
//
// Every call on one GPU uses the same array and leaves its contents behind: the counters are
// cleared by a kernel at the start of each call, and the pivot lists are rewritten before they
// are read. Calls that overlap on one GPU (two streams, two host threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"
#include "tc_internal.h"
#include "workspace.h"

namespace {

// 16 MiB: three pivot lists of round_up(m, 64) int32 entries each, then the pivot counters.
constexpr size_t kScratchWords = (size_t)4 << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) int32_t scratch[kScratchWords];

namespace {

std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch` (looked up on the first call on the device).
int scratch_ptr(int dev, int32_t **out) {
  GARY_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "gary_tc_count_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    GARY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GARY_CHECK_ARG(p != nullptr, "gary_tc_count_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (int32_t *)p;
  }
  *out = g_scratch_ptr[dev];
  return GARY_OK;
}

}  // namespace

// Largest vertex count: 3 * round_up(m, 64) + kPivotCounterWords <= kScratchWords.
constexpr int64_t kMaxVertices =
    (int64_t)(((kScratchWords - gary::kPivotCounterWords) / 3) & ~(size_t)63);
static_assert(kMaxVertices == 1397568, "the limit stated in gary.h");

extern "C" {

int gary_tc_count_noalloc(const gary_device_graph *g, const gary_launch_cfg *cfg,
                          uint64_t *d_count, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(g && d_count, "gary_tc_count_noalloc: null pointer");
  GARY_CHECK_ARG(g->csr.m >= 0 && g->csr.nnz >= 0, "gary_tc_count_noalloc: m=%d nnz=%lld",
                 g->csr.m, (long long)g->csr.nnz);
  GARY_CHECK_ARG(g->csr.m <= kMaxVertices,
                 "gary_tc_count_noalloc: %d vertices exceed the %lld its %zu-byte scratch holds",
                 g->csr.m, (long long)kMaxVertices, kScratchWords * sizeof(int32_t));
  if (g->csr.m == 0 || g->csr.nnz == 0) return GARY_OK;
  GARY_CHECK_ARG(g->csr.indPtr && g->csr.deg && g->csr.colIdx,
                 "gary_tc_count_noalloc: null CSR arrays");
  PivotCfg pc;
  if (cfg) {
    if (cfg->block) pc.waves = cfg->block == 128 ? 2 : cfg->block == 512 ? 8 : 4;
    if (cfg->grid) pc.grid = cfg->grid;
    if (cfg->bin_small) pc.bin_small = cfg->bin_small;
    if (cfg->bin_hub) pc.bin_hub = cfg->bin_hub;
  }
  GARY_CHECK_ARG(pc.grid >= 0, "gary_tc_count_noalloc: grid=%d", pc.grid);
  int dev = 0;
  GARY_TRY(current_device(&dev));
  int32_t *lists = nullptr;
  GARY_TRY(scratch_ptr(dev, &lists));
  const size_t stride = ((size_t)g->csr.m + 63) & ~(size_t)63;
  return launch_pivot_core(*g, pc, lists, stride, lists + 3 * stride, nullptr, nullptr, 0,
                           (unsigned long long *)d_count, (hipStream_t)stream, false);
  GARY_API_END
}

}  // extern "C"
