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

#include <cstdio>
#include <map>
#include <mutex>

#include "../../include/hsc.h"
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

// the sorted positions of the kept blocks' first entries, in output order
// (heads[pos[j]] = j), and their count (the exclusive-scan total)
__global__ void k_asm_compact(const int* __restrict__ flag, const int* __restrict__ pos, long long count,
                              int* __restrict__ heads, int* __restrict__ nheads) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (long long)gridDim.x * blockDim.x) {
    if (flag[j]) heads[pos[j]] = (int)j;
    if (j == count - 1) nheads[0] = pos[j] + flag[j];
  }
}

// each kept block sums its run of equal keys in input order, one lane per
// value component (9 lanes per block, kRedHeads blocks per workgroup), so
// every sum keeps the sequential order and the 9 lanes read each 72-byte
// input block contiguously (a folded lower block is transposed: component
// o = c + 3r of the output takes r + 3c of the input)
constexpr int kRedHeads = 64;

__global__ __launch_bounds__(9 * kRedHeads) void k_asm_reduce(
    const int* __restrict__ heads, const int* __restrict__ nheads, const u64* __restrict__ keys,
    const int* __restrict__ idx, long long count, const int* __restrict__ rows, const int* __restrict__ cols,
    const double* __restrict__ vals, int* __restrict__ orow, int* __restrict__ ocol, double* __restrict__ ovals) {
  const int h = threadIdx.x / 9, o = threadIdx.x - 9 * h;
  const int ot = o / 3 + 3 * (o % 3);
  const long long nh = nheads[0];
  for (long long t = (long long)blockIdx.x * kRedHeads + h; t < nh; t += (long long)gridDim.x * kRedHeads) {
    const long long j = heads[t];
    const u64 key = keys[j];
    double s = 0.0;
    for (long long e = j; e < count && keys[e] == key; ++e) {
      const int i = idx[e];
      s += vals[9 * (size_t)i + (rows[i] > cols[i] ? ot : o)];
    }
    ovals[9 * (size_t)t + o] = s;
    if (o == 0) {
      orow[t] = (int)(key >> 32);
      ocol[t] = (int)(key & 0xffffffffull);
    }
  }
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

// ABD-FEM coupling (abd_fem_linear_subsystem.cu:18-61): contact block H (3x3,
// ABD vertex a, FEM vertex f) -> J(xbar_a)^T H (12x3), written as the 4
// blocks (4 body + k, f), zero when the body or the vertex is fixed.  With
// x = p + A xbar and q = (p, A's rows), x_i = p_i + A_i . xbar, so block 0 is
// H and block k >= 1 (DOFs of row k-1 of A) is xbar outer H's row k-1: entry
// (m, l) = xbar[m] H(k-1, l) -- the layout of J^T g in the gravity setup at
// affine_body_dynamics.cu:609-613 (segment 3+3k = g[k] xbar).
//
// A workgroup takes tiles of kCplTile contacts: their H (contiguous), xbar,
// block row / column and flags are staged in LDS, then the tile's 36 output
// values and 4 index pairs per contact are written as contiguous runs (one
// element per lane), so every store is coalesced.
constexpr int kCplTile = 64;
constexpr int kCplThreads = 256;

__global__ __launch_bounds__(kCplThreads) void k_abd_fem_coupling(AbdFemCouplingParams P) {
  __shared__ double sH[kCplTile * 9];
  __shared__ double sX[kCplTile * 3];
  __shared__ int sRow[kCplTile], sCol[kCplTile], sZero[kCplTile];
  const int tid = threadIdx.x;
  for (long long t = blockIdx.x; t * kCplTile < P.count; t += gridDim.x) {
    const long long c0 = t * kCplTile;
    const int nc = (int)(P.count - c0 < kCplTile ? P.count - c0 : kCplTile);
    for (int i = tid; i < 9 * nc; i += kCplThreads) sH[i] = P.hess[9 * c0 + i];
    if (tid < nc) {
      const int av = P.abd_vertex[c0 + tid], fv = P.fem_vertex[c0 + tid];
      const int body = av >= 0 && av < P.n_abd_vertices ? P.v2b[av] : -1;
      const bool ok = body >= 0 && body < P.n_bodies && fv >= 0 && fv < P.n_fem_vertices;
      if (!ok) atomicOr(P.err, 1);
      sRow[tid] = ok ? P.abd_block_offset + 4 * body : -1;
      sCol[tid] = ok ? P.fem_block_offset + fv : -1;
      sZero[tid] = !ok || P.body_fixed[body] != 0 || P.vertex_fixed[fv] != 0;
#pragma unroll
      for (int m = 0; m < 3; ++m) sX[3 * tid + m] = ok ? P.xbar[3 * (long long)av + m] : 0.0;
    }
    __syncthreads();
    double* ov = P.oval + 36 * c0;
    for (int e = tid; e < 36 * nc; e += kCplThreads) {
      const int cl = e / 36, r = e - 36 * cl, k = r / 9, q = r - 9 * k;
      const int m = q % 3, l = q / 3;
      double v = 0.0;
      if (!sZero[cl]) {
        const double h = sH[9 * cl + (k == 0 ? q : (k - 1) + 3 * l)];
        v = k == 0 ? h : sX[3 * cl + m] * h;
      }
      ov[e] = v;
    }
    for (int e = tid; e < 4 * nc; e += kCplThreads) {
      const int cl = e >> 2;
      P.orow[4 * c0 + e] = sRow[cl] < 0 ? -1 : sRow[cl] + (e & 3);
      P.ocol[4 * c0 + e] = sCol[cl];
    }
    __syncthreads();  // the next tile overwrites the staging
  }
}

}  // namespace

hipError_t launch_abd_fem_coupling(const AbdFemCouplingParams& P, hipStream_t st) {
  if (P.count <= 0) return hipSuccess;
  const long long tiles = (P.count + kCplTile - 1) / kCplTile;
  const int grid = (int)std::min<long long>(tiles, 8192);
  hipLaunchKernelGGL(k_abd_fem_coupling, dim3(grid), dim3(kCplThreads), 0, st, P);
  return hipGetLastError();
}

// Workspace of one assembly of n triplets: two key and two index arrays
// (sort input and output), head flags, their scan, the compacted heads,
// hipcub's temporary storage and the head count, each 256-byte aligned.
namespace {
size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

hipError_t asm_layout(int n, hipStream_t st, AsmLayout* L) {
  u64* nk = nullptr;
  int* ni = nullptr;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, L->sort_bytes, nk, nk, ni, ni, n, 0, 64, st);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, L->scan_bytes, ni, ni, n, st);
  L->tmp = up256(std::max<size_t>(std::max(L->sort_bytes, L->scan_bytes), 16));
  L->kb = up256(sizeof(u64) * (size_t)n);
  L->ib = up256(sizeof(int) * (size_t)n);
  L->total = 2 * L->kb + 5 * L->ib + L->tmp + 256;
  return e;
}

// The assembly in a caller-provided workspace of at least L.total bytes
// (nothing in it needs to start at zero: every array is written before it is
// read).  Synchronises st; *out_count = number of unique kept blocks.
hipError_t assemble_in(const int* rows, const int* cols, const double* vals, long long count, int fold, int* orow,
                       int* ocol, double* ovals, long long* out_count, char* arena, const AsmLayout& L,
                       hipStream_t st) {
  const int n = (int)count;
  u64* k0 = (u64*)arena;
  u64* k1 = (u64*)(arena + L.kb);
  int* i0 = (int*)(arena + 2 * L.kb);
  int* i1 = (int*)(arena + 2 * L.kb + L.ib);
  int* flag = (int*)(arena + 2 * L.kb + 2 * L.ib);
  int* pos = (int*)(arena + 2 * L.kb + 3 * L.ib);
  void* tmp = arena + 2 * L.kb + 4 * L.ib;
  int* heads = (int*)(arena + 2 * L.kb + 4 * L.ib + L.tmp);
  int* dcount = (int*)(arena + 2 * L.kb + 5 * L.ib + L.tmp);
  size_t sort_bytes = L.sort_bytes, scan_bytes = L.scan_bytes;
  int hcount = 0;
  hipLaunchKernelGGL(k_asm_keys, dim3(grid_for(n)), dim3(256), 0, st, rows, cols, count, fold, k0, i0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, k0, k1, i0, i1, n, 0, 64, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_asm_heads, dim3(grid_for(n)), dim3(256), 0, st, k1, count, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, flag, pos, n, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_asm_compact, dim3(grid_for(n)), dim3(256), 0, st, flag, pos, count, heads, dcount);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const int grid = (int)std::min<long long>((count + kRedHeads - 1) / kRedHeads, 1024);
    hipLaunchKernelGGL(k_asm_reduce, dim3(grid), dim3(9 * kRedHeads), 0, st, heads, dcount, k1, i1, count, rows,
                       cols, vals, orow, ocol, ovals);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hcount, dcount, sizeof(int), hipMemcpyDeviceToHost, st);
  const hipError_t se = hipStreamSynchronize(st);
  if (e == hipSuccess) e = se;
  if (e == hipSuccess) *out_count = hcount;
  return e;
}

namespace {

// Library-held workspace of hsc_assemble_bcoo, one per device, grown on
// demand and kept for the process; calls are serialised by its lock.
struct AsmArena {
  char* p = nullptr;
  size_t cap = 0;
};
std::mutex g_asm_mu;
std::map<int, AsmArena> g_asm_arena;

}  // namespace

hipError_t assemble_bcoo(const int* rows, const int* cols, const double* vals, long long count, int fold,
                         int* orow, int* ocol, double* ovals, long long* out_count, hipStream_t st) {
  *out_count = 0;
  if (count <= 0) return hipSuccess;
  if (count >= (1LL << 31)) return hipErrorInvalidValue;
  AsmLayout L;
  hipError_t e = asm_layout((int)count, st, &L);
  if (e != hipSuccess) return e;
  int dev = 0;
  e = hipStreamGetDevice(st, &dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_asm_mu);
  AsmArena& a = g_asm_arena[dev];
  if (a.cap < L.total) {
    if (a.p) {
      e = hipStreamSynchronize(st);
      if (e == hipSuccess) e = hipFree(a.p);
      a.p = nullptr;
      a.cap = 0;
      if (e != hipSuccess) return e;
    }
    e = hipMalloc((void**)&a.p, L.total);
    if (e != hipSuccess) {
      a.p = nullptr;
      return e;
    }
    note_alloc();
    a.cap = L.total;
  }
  return assemble_in(rows, cols, vals, count, fold, orow, ocol, ovals, out_count, a.p, L, st);
}

}  // namespace hsc

