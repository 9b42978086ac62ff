// block_queue.h — the reference's block-scope queue (CTAQueue, include/ops/DeviceQueue.cuh:33-182)
// for gfx950: a ring of N slots, normally in LDS, with absolute head / tail positions.
//
// Same interface and contract as the reference:
//   clear()            block-cooperative; empties the queue
//   fill(src, n)       block-cooperative; the queue holds src[0 .. min(n, N)) in order
//   size()             head - tail, never below 0
//   tryEnqX / tryDeqX  exclusive mode (only enqueues, or only dequeues, at a time): reserve a
//                      position with one atomic, undo it on overshoot, then plain slot access
//   tryEnq / tryDeq    concurrent mode: producers and consumers at the same time
//
// Two differences, on purpose:
//   * Concurrent mode gives every slot a sequence number (the bounded MPMC ring published by
//     D. Vyukov): a producer may take position p only while seq[p % N] == p, a consumer only while
//     seq[p % N] == p + 1, each by one CAS on head / tail. Neither side waits for the other, and
//     an item can never be stranded. The reference instead undoes an overshooting fetch_add;
//     under a race of three producers that can give two of them the same position, and the item
//     written last then sits in a slot no consumer position covers.
//   * Concurrent calls are made by one lane per wave (each wave is one producer or consumer), as
//     the reference's own test does per warp (test/CTACommTest.cu:91-119). The lanes of a wave64
//     share one program counter, so per-lane retry loops on one queue would serialise a wave
//     against itself.
// A concurrent call retries its CAS at most kQueueSpins times and then reports kQueueStuck.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace gary {

enum QueueResult : int { kQueueEmptyOrFull = 0, kQueueOk = 1, kQueueStuck = -1 };

constexpr int kQueueSpins = 1 << 16;

template <unsigned N>
struct BlockQueue {
  uint32_t *data;  // N items
  uint32_t *seq;   // N sequence numbers (concurrent mode)
  uint32_t *pos;   // pos[0] = head (exclusive), pos[1] = tail; absolute, taken mod N on access

  __device__ static uint32_t load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ static void store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ static bool cas(uint32_t *p, uint32_t expected, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }

  // Block-cooperative: every thread of the block calls it; a barrier must follow.
  __device__ void clear() { fill(nullptr, 0); }

  // Block-cooperative: every thread of the block calls it; a barrier must follow.
  __device__ void fill(const uint32_t *src, unsigned n) {
    n = n < N ? n : N;
    for (unsigned i = threadIdx.x; i < N; i += blockDim.x) {
      if (i < n) data[i] = src[i];
      seq[i] = i < n ? i + 1 : i;  // full slots wait for consumer position i, empty ones for
                                   // producer position i
    }
    if (threadIdx.x == 0) {
      pos[0] = n;
      pos[1] = 0;
    }
  }

  __device__ unsigned size() const {
    const int h = (int)load(&pos[0]), t = (int)load(&pos[1]);
    return h > t ? (unsigned)(h - t) : 0u;
  }

  // Exclusive mode.
  __device__ int tryEnqX(uint32_t item) {
    const unsigned at = atomicAdd(&pos[0], 1u);
    if (at >= load(&pos[1]) + N) {
      atomicSub(&pos[0], 1u);  // overshoot: the queue is full
      return kQueueEmptyOrFull;
    }
    data[at % N] = item;
    return kQueueOk;
  }

  __device__ int tryDeqX(uint32_t &item) {
    const unsigned at = atomicAdd(&pos[1], 1u);
    if (at >= load(&pos[0])) {
      atomicSub(&pos[1], 1u);  // overshoot: the queue is empty
      return kQueueEmptyOrFull;
    }
    item = data[at % N];
    return kQueueOk;
  }

  // Concurrent mode, one lane per wave.
  __device__ int tryEnq(uint32_t item) {
    for (int spin = 0; spin < kQueueSpins; ++spin) {
      const uint32_t p = load(&pos[0]);
      const int diff = (int)(load(&seq[p % N]) - p);
      if (diff < 0) return kQueueEmptyOrFull;  // slot p - N not consumed yet: full
      if (diff == 0 && cas(&pos[0], p, p + 1)) {
        data[p % N] = item;
        store(&seq[p % N], p + 1);  // publishes the item to consumer position p
        return kQueueOk;
      }
      // another producer took p: retry at the new head
    }
    return kQueueStuck;
  }

  __device__ int tryDeq(uint32_t &item) {
    for (int spin = 0; spin < kQueueSpins; ++spin) {
      const uint32_t p = load(&pos[1]);
      const int diff = (int)(load(&seq[p % N]) - (p + 1));
      if (diff < 0) return kQueueEmptyOrFull;  // item p not published yet: empty
      if (diff == 0 && cas(&pos[1], p, p + 1)) {
        item = data[p % N];
        store(&seq[p % N], p + N);  // frees the slot for producer position p + N
        return kQueueOk;
      }
      // another consumer took p: retry at the new tail
    }
    return kQueueStuck;
  }
};

}  // namespace gary
