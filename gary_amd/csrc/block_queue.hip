// block_queue.hip — drives the block-scope queue (block_queue.h) the way the reference's queue test
// does (test/CTACommTest.cu:91-119, FairSleepKernel), and with producers and consumers at once.
// The reference uses CTAQueue in that test only; gary_block_queue_exercise is its counterpart
// behind the C ABI, so the queue's contract is checked on the device.
//
// One workgroup of kQWaves waves, a kQSlots-slot queue in LDS, lane 0 of each wave calling it.
//   mode 0 (exclusive): fill from items[0 .. n), n <= kQSlots; every wave takes items with tryDeqX
//     until the queue is empty; the k-th item wave w takes goes to out[w * kQSlots + k] and
//     status[w] = the number it took.
//   mode 1 (concurrent): waves 0 .. kQWaves/2 - 1 enqueue their share of items[0 .. n) with tryEnq
//     while the other waves dequeue with tryDeq until n items are out; an item x (< n) adds one
//     to out[x]; status[w] = the items wave w moved, or -1 if one of its calls was stuck or an
//     item was out of range.
// Every retry loop is bounded (kQTries failed calls in a row), so the kernel always ends.
#include <hip/hip_runtime.h>

#include "block_queue.h"
#include "common.h"
#include "device_util.h"

namespace gary {
namespace {

constexpr unsigned kQSlots = 128;
constexpr int kQWaves = 16;
constexpr int kQTries = 1 << 16;  // failed calls in a row before a wave gives up

__global__ void __launch_bounds__(kQWaves * kWave)
    queue_exercise_kernel(int mode, const uint32_t *__restrict__ items, uint32_t n,
                          uint32_t *__restrict__ out, int32_t *__restrict__ status) {
  __shared__ uint32_t data[kQSlots], seq[kQSlots], pos[2];
  __shared__ uint32_t consumed;
  BlockQueue<kQSlots> q{data, seq, pos};
  const int lane = lane_id(), wib = threadIdx.x / kWave;
  if (mode == 0) q.fill(items, n);
  else q.clear();
  if (threadIdx.x == 0) consumed = 0;
  __syncthreads();
  if (lane != 0) return;
  if (mode == 0) {
    uint32_t k = 0, item = 0;
    while (q.tryDeqX(item) == kQueueOk) out[wib * kQSlots + k++] = item;
    status[wib] = (int32_t)k;
    return;
  }
  constexpr int half = kQWaves / 2;
  int32_t moved = 0;
  bool bad = false;
  if (wib < half) {  // producer
    const uint32_t per = (n + half - 1) / half;
    const uint32_t lo = min(n, (uint32_t)wib * per), hi = min(n, lo + per);
    for (uint32_t i = lo; i < hi && !bad; ++i) {
      int r = kQueueEmptyOrFull;
      for (int t = 0; t < kQTries && r == kQueueEmptyOrFull; ++t) {
        r = q.tryEnq(items[i]);
        if (r == kQueueEmptyOrFull) __builtin_amdgcn_s_sleep(2);  // full: let consumers drain
      }
      if (r == kQueueOk) ++moved;
      else bad = true;
    }
  } else {  // consumer
    int fails = 0;
    while (fails < kQTries &&
           __hip_atomic_load(&consumed, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < n) {
      uint32_t item = 0;
      const int r = q.tryDeq(item);
      if (r == kQueueOk) {
        fails = 0;
        if (item < n) atomicAdd(&out[item], 1u);
        else bad = true;
        atomicAdd(&consumed, 1u);
        ++moved;
      } else if (r == kQueueStuck) {
        bad = true;
        break;
      } else {
        ++fails;
        __builtin_amdgcn_s_sleep(2);  // empty: let producers fill
      }
    }
  }
  status[wib] = bad ? -1 : moved;
}

}  // namespace
}  // namespace gary

extern "C" {

int gary_block_queue_exercise(int mode, const uint32_t *d_items, uint32_t n, uint32_t *d_out,
                              int32_t *d_status, void *stream) {
  GARY_API_BEGIN
  using namespace gary;
  GARY_CHECK_ARG(mode == 0 || mode == 1, "gary_block_queue_exercise: mode %d is not 0 or 1", mode);
  GARY_CHECK_ARG(d_status && d_out && (n == 0 || d_items),
                 "gary_block_queue_exercise: null pointer");
  GARY_CHECK_ARG(mode == 1 || n <= kQSlots,
                 "gary_block_queue_exercise: exclusive mode holds at most %u items", kQSlots);
  GARY_CHECK_ARG(n <= (1u << 24), "gary_block_queue_exercise: at most 2^24 items");
  hipStream_t st = (hipStream_t)stream;
  const size_t out_bytes = sizeof(uint32_t) * (mode == 0 ? (size_t)kQWaves * kQSlots : (size_t)n);
  if (out_bytes) GARY_HIP(hipMemsetAsync(d_out, 0, out_bytes, st));
  queue_exercise_kernel<<<1, kQWaves * kWave, 0, st>>>(mode, d_items, n, d_out, d_status);
  GARY_HIP(hipGetLastError());
  return GARY_OK;
  GARY_API_END
}

}  // extern "C"
