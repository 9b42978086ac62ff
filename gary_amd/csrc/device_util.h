// device_util.h — wave64 building blocks for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gary {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Orders LDS/global accesses of the lanes of one wave (waves execute in lock-step; this stops
// the compiler from moving memory operations across the point and waits for issued LDS ops).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of `mask` strictly below this lane (v_mbcnt).
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Calls f(key, mask, leader) once per distinct key among the wave's lanes with key >= 0 (every
// lane of the wave must call it): mask = the lanes holding that key, leader = the lowest of them.
template <class F>
__device__ __forceinline__ void for_each_key(int key, F &&f) {
  uint64_t todo = __ballot(key >= 0);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int k = __shfl(key, leader, kWave);
    const uint64_t mask = __ballot(key == k);
    f(k, mask, leader);
    todo &= ~mask;
  }
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, kWave);
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o, kWave);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o, kWave);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Sum over groups of `G` consecutive lanes (G a power of two <= 64); every lane of a group
// receives its group's sum.
template <int G>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, kWave);
  return v;
}

// Inclusive prefix sum across the 64 lanes.
__device__ __forceinline__ uint32_t wave_inclusive_scan_u32(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

// Inclusive max-scan over the wave with DPP: row shifts by 1, 2, 4, 8 inside each 16-lane row,
// then the last lane of rows 0 and 2 into rows 1 and 3 (row_bcast:15), then lane 31 into rows
// 2 and 3 (row_bcast:31). Lanes without a source keep INT32_MIN (the identity).
__device__ __forceinline__ int32_t wave_inclusive_max_i32(int32_t v) {
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp((int)INT32_MIN, v, 0x143, 0xc, 0xf, false));
  return v;
}

// Inclusive prefix sum within groups of G consecutive lanes.
template <int G>
__device__ __forceinline__ uint32_t group_inclusive_scan_u32(uint32_t v) {
  const int l = lane_id() & (G - 1);
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

// Accumulate a per-lane partial into a 64-bit device counter with one atomic per wave.
__device__ __forceinline__ void wave_atomic_add(unsigned long long *dst, uint64_t v) {
  uint64_t s = wave_sum_u64(v);
  if (lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

// Adds the block's sum of v to *dst with one global atomic (through the LDS word `acc`). Every
// thread of the block calls it, once, at the end of the kernel. Waves of a persistent kernel
// finish together, and one global atomic per wave on a single address then serialises at one L2
// channel (about 10 ns each: 8,192 waves add up to about 80 us).
__device__ __forceinline__ void block_atomic_add(unsigned long long *dst, uint64_t v,
                                                 unsigned long long *acc) {
  __syncthreads();
  if (threadIdx.x == 0) *acc = 0;
  __syncthreads();
  const uint64_t s = wave_sum_u64(v);
  if (lane_id() == 0 && s) atomicAdd(acc, (unsigned long long)s);
  __syncthreads();
  if (threadIdx.x == 0 && *acc) atomicAdd(dst, *acc);
}

// Block-wide u64 sum into a counter: wave sums -> LDS -> one atomic per workgroup.
template <int BLOCK>
__device__ __forceinline__ void block_atomic_add(unsigned long long *dst, uint64_t v,
                                                 uint64_t *lds /* BLOCK/64 entries */) {
  uint64_t s = wave_sum_u64(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) t += lds[i];
    if (t) atomicAdd(dst, (unsigned long long)t);
  }
}

// Grid-stride helpers.
__device__ __forceinline__ int64_t global_wave_id() {
  return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
}
__device__ __forceinline__ int64_t num_global_waves() {
  return ((int64_t)gridDim.x * blockDim.x) / kWave;
}

}  // namespace gary
