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

// v of lane (lane ^ M) for M in {1, 2, 4, 8}, through DPP lane moves inside
// the VALU (no LDS permute round trip): quad_perm for 1 and 2; 4 = half-row
// mirror (lane ^ 7) then the quad reversal (lane ^ 3); 8 = row mirror
// (lane ^ 15) then the half-row mirror (lane ^ 7).
template <int M>
__device__ __forceinline__ int dpp_xor_i32(int v) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8, "DPP xor distance");
  if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  if constexpr (M == 4) {
    const int t = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);             // row_half_mirror
    return __builtin_amdgcn_mov_dpp(t, 0x1B, 0xF, 0xF, true);                      // quad_perm [3,2,1,0]
  }
  if constexpr (M == 8) {
    const int t = __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true);             // row_mirror
    return __builtin_amdgcn_mov_dpp(t, 0x141, 0xF, 0xF, true);                     // row_half_mirror
  }
  return v;
}
template <int M>
__device__ __forceinline__ float dpp_xor(float v) {
  return __int_as_float(dpp_xor_i32<M>(__float_as_int(v)));
}
template <int M>
__device__ __forceinline__ double dpp_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_xor_i32<M>((int)(unsigned)(b & 0xffffffffll));
  const int hi = dpp_xor_i32<M>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over groups of G consecutive lanes: v += v(lane ^ m) for m = G/2 .. 1,
// so every lane of a group gets the bitwise-same value (each step adds two
// commutative partial sums).  Distances up to 8 move through DPP.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "group size");
  if constexpr (G == 64) v += __shfl_xor(v, 32, 64);
  if constexpr (G >= 32) v += __shfl_xor(v, 16, G);
  if constexpr (G >= 16) v += dpp_xor<8>(v);
  if constexpr (G >= 8) v += dpp_xor<4>(v);
  if constexpr (G >= 4) v += dpp_xor<2>(v);
  if constexpr (G >= 2) v += dpp_xor<1>(v);
  return v;
}

// Butterfly sum over the 64 lanes; every lane gets the bitwise-same value
// (each step adds two commutative partial sums).  (LDS permutes: DPP moves
// measured slower here in FP64, as the reductions' critical path.)
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Deterministic block sum (fixed order); result valid in thread 0.
// `red` must hold blockDim.x/64 doubles.  Contains __syncthreads.
__device__ __forceinline__ double block_sum_t0(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    for (int i = 0; i < nw; ++i) s += red[i];
  }
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- grid barrier
// Hierarchical barrier for a co-resident (cooperative) grid on the 8-XCD
// MI355X.  Each XCD has its own L2; making one workgroup's stores visible to
// another XCD needs an L2 writeback (release) and an L2 invalidate (acquire).
// Doing the writeback once per XCD instead of once per workgroup is what
// makes the barrier cheap: workgroups wait for their own stores, arrive on a
// per-XCD counter, the last arrival of each XCD writes the XCD's L2 back and
// joins an 8-way device counter, then releases its XCD through a flag; every
// workgroup then does the agent-scope acquire.  All counters are monotonic
// (zeroed before the launch); every wait is bounded by `timeout` ticks and
// ends early when another workgroup raised `err`.
struct GridBarrier {
  unsigned* ctr;          // [kBarrierWords] zeroed before the launch
  int* err;
  long long timeout;
  unsigned nwg;
  unsigned flat_epoch, epoch, xcc, nx, nact;
};
// counter layout (kBarrierWords words): [0] flat, [32] XCD leaders, [64+32x] XCD arrivals,
// [320+32x] XCD release flags, [576+32x] XCD workgroup counts

__device__ __forceinline__ unsigned hw_xcc_id() {
  // HW_REG_XCC_ID (hardware register 20), bits [3:0]
  return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7u;
}

__device__ __forceinline__ void wait_vm_zero() {
  // s_waitcnt vmcnt(0): this wave's stores have reached the L2
  __builtin_amdgcn_s_waitcnt(0x0F70);
}

__device__ __forceinline__ bool gb_spin_until(unsigned* addr, unsigned target, int* err, long long t0,
                                              long long timeout) {
  for (;;) {
    const unsigned v = __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int)(v - target) >= 0) return true;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (rt_ticks() - t0 > timeout) {
      __hip_atomic_store(err, HSC_ERR_DEVICE_TIMEOUT_CODE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// flat barrier (every workgroup fences): used once per launch to count the
// workgroups of each XCD
__device__ bool gb_flat(GridBarrier& B) {
  __shared__ int s_ok;
  wait_vm_zero();  // every wave's stores reach L2 (s_barrier alone does not wait for them)
  __syncthreads();
  B.flat_epoch += 1;
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(B.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool ok = gb_spin_until(B.ctr, B.flat_epoch * B.nwg, B.err, rt_ticks(), B.timeout);
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__device__ bool gb_init(GridBarrier& B, unsigned* ctr, int* err, long long timeout, unsigned nwg) {
  __shared__ unsigned s_x, s_nx, s_na;
  B.ctr = ctr;
  B.err = err;
  B.timeout = timeout;
  B.nwg = nwg;
  B.flat_epoch = 0;
  B.epoch = 0;
  if (threadIdx.x == 0) {
    s_x = hw_xcc_id();
    __hip_atomic_fetch_add(ctr + 576 + 32 * s_x, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!gb_flat(B)) return false;
  if (threadIdx.x == 0) {
    s_nx = __hip_atomic_load(ctr + 576 + 32 * s_x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned na = 0;
    for (int x = 0; x < 8; ++x)
      na += __hip_atomic_load(ctr + 576 + 32 * x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0u;
    s_na = na;
  }
  __syncthreads();
  B.xcc = s_x;
  B.nx = s_nx;
  B.nact = s_na;
  return true;
}

__device__ bool gb_sync(GridBarrier& B) {
  __shared__ int s_ok;
  wait_vm_zero();  // every wave's stores reach L2 before the XCD leader writes the L2 back
  __syncthreads();
  B.epoch += 1;
  if (threadIdx.x == 0) {
    const long long t0 = rt_ticks();
    bool ok;
    const unsigned prev =
        __hip_atomic_fetch_add(B.ctr + 64 + 32 * B.xcc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev + 1u == B.epoch * B.nx) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // write this XCD's L2 back
      __hip_atomic_fetch_add(B.ctr + 32, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = gb_spin_until(B.ctr + 32, B.epoch * B.nact, B.err, t0, B.timeout);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(B.ctr + 320 + 32 * B.xcc, B.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      ok = gb_spin_until(B.ctr + 320 + 32 * B.xcc, B.epoch, B.err, t0, B.timeout);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__device__ __forceinline__ int tile_of_chunk(int c, int nC, int T) {
  const long long prod = (long long)c * T;
  if (prod < (1LL << 31)) return (int)((unsigned)prod / (unsigned)nC);  // 32-bit division when it fits
  return (int)(prod / nC);
}

}  // namespace hsc
