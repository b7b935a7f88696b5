// Assembly before the HSC path: general 3x3-block triplets (duplicates
// allowed, one or both triangles) -> the upper-triangle BCOO of unique blocks
// that the plugin receives (libuipc's MatrixConverter::ge2sym + convert,
// called at global_linear_system.cu:232-234; MatrixConverter itself is not
// part of the reference sources).
//
// Keys (row, col) with row <= col; a lower block (row > col) is folded onto
// (col, row) transposed, or dropped when the input holds both triangles of a
// symmetric matrix.  A stable radix sort of the keys keeps equal keys in input
// order, and each unique block is summed by one thread in that order, so the
// result does not depend on the launch and matches a sequential sum.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "kernels.hpp"

namespace hsc {

namespace {

using u64 = unsigned long long;
constexpr u64 kDropKey = ~0ull;  // sorts last

__global__ void k_asm_keys(const int* __restrict__ rows, const int* __restrict__ cols, long long count, int fold,
                           u64* __restrict__ keys, int* __restrict__ idx) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = rows[i], c = cols[i];
    u64 key = kDropKey;
    if (r <= c) key = ((u64)(unsigned)r << 32) | (u64)(unsigned)c;
    else if (fold) key = ((u64)(unsigned)c << 32) | (u64)(unsigned)r;
    keys[i] = key;
    idx[i] = (int)i;
  }
}

// flag[j] = 1 when sorted entry j starts a new kept block
__global__ void k_asm_heads(const u64* __restrict__ keys, long long count, int* __restrict__ flag) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (long long)gridDim.x * blockDim.x)
    flag[j] = keys[j] != kDropKey && (j == 0 || keys[j] != keys[j - 1]) ? 1 : 0;
}

// each head sums its run of equal keys in input order (a folded lower block
// transposed: column-major v[r + 3c] lands on (c, r))
__global__ void k_asm_reduce(const u64* __restrict__ keys, const int* __restrict__ idx,
                             const int* __restrict__ flag, const int* __restrict__ pos, long long count,
                             const int* __restrict__ rows, const int* __restrict__ cols,
                             const double* __restrict__ vals, int* __restrict__ orow, int* __restrict__ ocol,
                             double* __restrict__ ovals) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (long long)gridDim.x * blockDim.x) {
    if (!flag[j]) continue;
    const u64 key = keys[j];
    double s[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) s[q] = 0.0;
    for (long long e = j; e < count && keys[e] == key; ++e) {
      const int i = idx[e];
      const double* v = vals + 9 * (size_t)i;
      if (rows[i] > cols[i]) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) s[c + 3 * r] += v[r + 3 * c];
      } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) s[q] += v[q];
      }
    }
    const int p = pos[j];
    orow[p] = (int)(key >> 32);
    ocol[p] = (int)(key & 0xffffffffull);
#pragma unroll
    for (int q = 0; q < 9; ++q) ovals[9 * (size_t)p + q] = s[q];
  }
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

}  // namespace

hipError_t assemble_bcoo(const int* rows, const int* cols, const double* vals, long long count, int fold,
                         int* orow, int* ocol, double* ovals, long long* out_count, hipStream_t st) {
  *out_count = 0;
  if (count <= 0) return hipSuccess;
  if (count >= (1LL << 31)) return hipErrorInvalidValue;
  const int n = (int)count;
  u64 *k0 = nullptr, *k1 = nullptr;
  int *i0 = nullptr, *i1 = nullptr, *flag = nullptr, *pos = nullptr;
  void* tmp = nullptr;
  size_t sort_bytes = 0, scan_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, i0, i1, n, 0, 64, st);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, pos, n, st);
  const size_t tmp_bytes = std::max(sort_bytes, scan_bytes);
  if (e == hipSuccess) e = hipMallocAsync((void**)&k0, sizeof(u64) * n, st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&k1, sizeof(u64) * n, st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&i0, sizeof(int) * n, st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&i1, sizeof(int) * n, st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&flag, sizeof(int) * n, st);
  if (e == hipSuccess) e = hipMallocAsync((void**)&pos, sizeof(int) * n, st);
  if (e == hipSuccess) e = hipMallocAsync(&tmp, std::max<size_t>(tmp_bytes, 16), st);
  int last[2] = {0, 0};
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_asm_keys, dim3(grid_for(n)), dim3(256), 0, st, rows, cols, count, fold, k0, i0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, k0, k1, i0, i1, n, 0, 64, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_asm_heads, dim3(grid_for(n)), dim3(256), 0, st, k1, count, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, flag, pos, n, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_asm_reduce, dim3(grid_for(n)), dim3(256), 0, st, k1, i1, flag, pos, count, rows, cols,
                       vals, orow, ocol, ovals);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&last[0], pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&last[1], flag + n - 1, sizeof(int), hipMemcpyDeviceToHost, st);
  for (void* p : {(void*)k0, (void*)k1, (void*)i0, (void*)i1, (void*)flag, (void*)pos, tmp})
    if (p) (void)hipFreeAsync(p, st);
  const hipError_t se = hipStreamSynchronize(st);
  if (e == hipSuccess) e = se;
  if (e == hipSuccess) *out_count = (long long)last[0] + last[1];
  return e;
}

}  // namespace hsc
