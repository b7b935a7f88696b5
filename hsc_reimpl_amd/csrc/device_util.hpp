// Device helpers: wave/block reductions, bounded grid barrier.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace hsc {

#define HSC_ERR_DEVICE_TIMEOUT_CODE 13

__device__ __forceinline__ long long rt_ticks() {
  // constant 100 MHz counter (s_memrealtime)
  return (long long)__builtin_amdgcn_s_memrealtime();
}

// Butterfly sum over the 64 lanes; every lane gets the bitwise-same value
// (each step adds two commutative partial sums).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum over groups of G consecutive lanes.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
  return v;
}

// Deterministic block sum (fixed order); result valid in thread 0.
// `red` must hold kBlock/64 doubles.  Contains __syncthreads.
__device__ __forceinline__ double block_sum_t0(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) s += red[i];
  }
  __syncthreads();
  return s;
}

// Grid barrier for a co-resident (cooperative) grid.  `bar` is a monotonic
// counter zeroed before the launch; barrier number `epoch` completes when it
// reaches epoch*nwg.  The wait is bounded: after `timeout` ticks (or when
// another workgroup has raised `err`) it sets `err` and returns false so
// every workgroup leaves the kernel instead of hanging.
__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned& epoch, unsigned nwg,
                                          int* err, long long timeout) {
  __shared__ int s_ok;
  __syncthreads();
  epoch += 1;
  if (threadIdx.x == 0) {
    const unsigned target = epoch * nwg;
    // release: one agent-scope fence before arriving; the spin itself uses
    // relaxed loads (an acquire load per spin would invalidate the L2 each
    // time) and a single agent-scope fence after the wait is the acquire.
    __threadfence();
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long t0 = rt_ticks();
    int ok = 1;
    for (;;) {
      unsigned v = __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((int)(v - target) >= 0) break;
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
      if (rt_ticks() - t0 > timeout) {
        __hip_atomic_store(err, HSC_ERR_DEVICE_TIMEOUT_CODE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__device__ __forceinline__ int tile_of_chunk(int c, int nC, int T) {
  return (int)(((long long)c * T) / nC);
}

}  // namespace hsc
