// Persistent Schur-complement PCG for the HSC solve on gfx950 (MI355X).
//
// One cooperative launch runs the whole PCG of every system in a batch
// (linear_hsc.cu:359-393 of the reference, one system; core.py:95-138 for the
// true-residual schedule).  Per iteration and system:
//   A  p <- z + beta p (on the fly), q = D p, <p,q>, t = W^T p       [SpMV + low-rank]
//   B  r -= alpha (q - W t), x += alpha p, z0 = w Dinv r, <r,r>     [elementwise]
//   C1..C(r-1)  z <- z + w Dinv (r - D z) (Horner form of the Neumann
//        series sum_k (I - w Dinv D)^k w Dinv r), last one with <r,z>
// S = D - V C V^T is applied as D - W W^T with W = V L, C = L L^T.
// Scalars live in LDS; every workgroup reduces the per-chunk partials in the
// same fixed order, so all of them take the same decisions and the result is
// bitwise reproducible.  Grid barriers have bounded waits.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace hsc {

enum { ST_RUN = 0, ST_DONE = 1 };
enum { CHK_NONE = 0, CHK_CONVERGE = 1, CHK_REPLACE = 2, CHK_NORM = 3 };
constexpr int kBreakdown = 7, kNotConverged = 8;

__device__ __forceinline__ double combine(double z, double p, double beta) { return fma(beta, p, z); }
__device__ __forceinline__ float combine(float z, float p, float beta) { return fmaf(beta, p, z); }

// The same pointer in the global address space: loads through it compile to
// global_load (vector memory counter only) instead of flat_load, which the
// compiler emits when a pointer's origin (a select between buffers, a struct
// field) hides that it is global, and which also holds LDS waits back.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

// gbl() when the kernel's values are FP64 (MT = double), the pointer as it is
// otherwise: FP64 solves measured 5% faster with global loads in the row
// SpMVs, passes and stage A; the FP32 persistent kernel gained nothing (1-5%
// slower in the runs measured), so its passes keep the compiler's choice
template <typename MT, typename T>
__device__ __forceinline__ auto gsel(const T* p) {
  if constexpr (sizeof(MT) == 8) return gbl(p);
  else return p;
}

template <typename MT> __device__ __forceinline__ const MT* mat_vals(const SysDesc& S);
template <> __device__ __forceinline__ const double* mat_vals<double>(const SysDesc& S) { return S.val64; }
template <> __device__ __forceinline__ const float* mat_vals<float>(const SysDesc& S) { return S.val32; }
template <typename MT> __device__ __forceinline__ const MT* inv_vals(const SysDesc& S);
template <> __device__ __forceinline__ const double* inv_vals<double>(const SysDesc& S) { return S.invD64; }
template <> __device__ __forceinline__ const float* inv_vals<float>(const SysDesc& S) { return S.invD32; }
template <typename MT> __device__ __forceinline__ const MT* w_vals(const SysDesc& S);
template <> __device__ __forceinline__ const double* w_vals<double>(const SysDesc& S) { return S.W64; }
template <> __device__ __forceinline__ const float* w_vals<float>(const SysDesc& S) { return S.W32; }

// One block of a lane's row sum: a += B x when `in` (a select).  The fused
// multiply-adds are written out, so every code path that computes a row
// (registers, LDS or memory operands) rounds the same way: the sums are
// bitwise the same whatever the compiler's contraction choices.
template <typename MT>
__device__ __forceinline__ void blk_step(bool in, const MT (&v)[9], MT x0, MT x1, MT x2, MT& a0, MT& a1, MT& a2) {
  const MT b0 = a0 + fma(v[6], x2, fma(v[3], x1, v[0] * x0));
  const MT b1 = a1 + fma(v[7], x2, fma(v[4], x1, v[1] * x0));
  const MT b2 = a2 + fma(v[8], x2, fma(v[5], x1, v[2] * x0));
  a0 = in ? b0 : a0;
  a1 = in ? b1 : a1;
  a2 = in ? b2 : a2;
}

// Output component of a preconditioner pass: z + w d.(r - D z)_row, with d
// the component's row of the inverse diagonal block and a = (D z)_row
// (fused multiply-adds written out, as in blk_step).
template <typename VT>
__device__ __forceinline__ VT pass_out(VT zo, VT w, VT d0, VT d1, VT d2, double r0, double r1, double r2, VT a0,
                                       VT a1, VT a2) {
  const VT e0 = (VT)r0 - a0, e1 = (VT)r1 - a1, e2 = (VT)r2 - a2;
  return fma(w, fma(d2, e2, fma(d1, e1, d0 * e0)), zo);
}

// y_row = sum_j D_ij src(j) over one block row with G lanes (lane strides the
// row's 3x3 blocks); every lane of the group gets the full 3-vector.  With
// U > 1 each round puts U blocks per lane in flight at once (column indices
// first, then the gathers and value streams, then the FMAs), so a row costs
// two dependent latencies per round instead of two per block; every lane
// still adds its blocks p0+lane, p0+lane+G, ... in increasing order.
template <int G, typename MT, int U = 1, typename Src>
__device__ __forceinline__ void row_spmv(const SysDesc& S, const MT* __restrict__ val, int row,
                                         int lane, Src src, MT& a0, MT& a1, MT& a2) {
  a0 = a1 = a2 = MT(0);
  if (row < S.nb) {
    const auto rowptr = gsel<MT>(S.rowptr);
    const auto colidx = gsel<MT>(S.colidx);
    const int p0 = rowptr[row], p1 = rowptr[row + 1];
    if constexpr (U == 1) {
      for (int p = p0 + lane; p < p1; p += G) {
        const int c = colidx[p];
        MT x0, x1, x2;
        src(c, x0, x1, x2);
        const auto v = gsel<MT>(val) + val_index(p, 0);
        const MT vb[9] = {v[0],             v[kValChunk],     v[2 * kValChunk], v[3 * kValChunk], v[4 * kValChunk],
                          v[5 * kValChunk], v[6 * kValChunk], v[7 * kValChunk], v[8 * kValChunk]};
        blk_step(true, vb, x0, x1, x2, a0, a1, a2);
      }
    } else {
      for (int pb = p0 + lane; pb < p1; pb += G * U) {
        int pp[U], cc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          pp[u] = min(pb + u * G, p1 - 1);  // rows are never empty (diagonal block)
          cc[u] = colidx[pp[u]];
        }
        MT x[U][3], v[U][9];
        const auto vg = gsel<MT>(val);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          src(cc[u], x[u][0], x[u][1], x[u][2]);
#pragma unroll
          for (int e = 0; e < 9; ++e) v[u][e] = vg[val_index(pp[u], e)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) blk_step(pb + u * G < p1, v[u], x[u][0], x[u][1], x[u][2], a0, a1, a2);
      }
    }
  }
  a0 = group_sum<G>(a0);
  a1 = group_sum<G>(a1);
  a2 = group_sum<G>(a2);
}

// blocks per lane in flight per round of the streaming SpMV: 1 in the
// latency schedule (its rows are register-resident), more in the throughput
// schedule, where each workgroup streams many row tiles
template <int G, bool RES>
struct StreamU {
  static constexpr int value = RES ? 1 : (G >= 16 ? 1 : (G == 8 ? 2 : 3));
};

template <typename T>
__device__ __forceinline__ T pick3(int d, T a0, T a1, T a2) { return d == 0 ? a0 : (d == 1 ? a1 : a2); }

// ---------------------------------------------------------------- stages
// Stage A on one row tile: p_new = z + beta p_old, q = D p_new, dot += p_new.q
template <int G, typename VT, int U = 1>
__device__ __forceinline__ void tile_A(const SysDesc& S, int tl, const VT* __restrict__ z,
                                       const VT* __restrict__ pold, VT* __restrict__ pnew,
                                       VT* __restrict__ q, VT beta, double& dot) {
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  const int row = tile_row(S, tl, kBlock / G, g);
  VT a0, a1, a2;
  row_spmv<G, VT, U>(S, mat_vals<VT>(S), row, lane,
                  [&](int c, VT& x0, VT& x1, VT& x2) {
                    x0 = combine(z[3 * c], pold[3 * c], beta);
                    x1 = combine(z[3 * c + 1], pold[3 * c + 1], beta);
                    x2 = combine(z[3 * c + 2], pold[3 * c + 2], beta);
                  },
                  a0, a1, a2);
  if (row < S.nb && lane < 3) {
    const int idx = 3 * row + lane;
    const VT pn = combine(z[idx], pold[idx], beta);
    const VT qv = pick3(lane, a0, a1, a2);
    pnew[idx] = pn;
    q[idx] = qv;
    dot += (double)pn * (double)qv;
  }
}

// Neumann (Horner) step on one row tile: zout = zin + w Dinv (r - D zin)
template <int G, typename VT, int U = 1>
__device__ __forceinline__ void tile_N(const SysDesc& S, int tl, const VT* __restrict__ zin,
                                       VT* __restrict__ zout, VT omega, bool last, double& rz) {
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  const int row = tile_row(S, tl, kBlock / G, g);
  VT a0, a1, a2;
  row_spmv<G, VT, U>(S, mat_vals<VT>(S), row, lane,
                  [&](int c, VT& x0, VT& x1, VT& x2) {
                    x0 = zin[3 * c];
                    x1 = zin[3 * c + 1];
                    x2 = zin[3 * c + 2];
                  },
                  a0, a1, a2);
  if (row < S.nb && lane < 3) {
    const int i3 = 3 * row;
    const double r0 = S.r[i3], r1 = S.r[i3 + 1], r2 = S.r[i3 + 2];
    const VT* inv = inv_vals<VT>(S) + 9 * row + 3 * lane;
    const VT out = pass_out(zin[i3 + lane], omega, inv[0], inv[1], inv[2], r0, r1, r2, a0, a1, a2);
    zout[i3 + lane] = out;
    if (last) rz += pick3(lane, r0, r1, r2) * (double)out;
  }
}

// ---------------------------------------------------------------- resident rows
// When a workgroup owns exactly one row tile in the Neumann chunking (the
// latency schedule), the first U blocks of each lane's share of its row --
// column indices and values -- and the row's inverse diagonal block are
// loaded once per launch and kept in registers; every Neumann pass then
// costs one gather latency instead of the row-pointer -> index -> gather
// chain.  Blocks beyond the first U per lane are streamed as before, and a
// lane holding such a tail adds a dependent load to the whole pass: FP32 keeps
// 5 blocks per lane at 4 lanes per row (20 per row, the 78k-tet system's
// longest row is 19; 251 VGPRs, no spills), which took its passes' streamed
// tails from 4.5% of the blocks to none and the PCG launch from 3.91 to 3.60 ms.
// FP64 keeps 3 (12 per row, 249 VGPRs; streamed tails 35% -> 17% of the
// blocks, plain FP64 PCG 14.5 -> 13.6 ms); 4 would pass 256 VGPRs.
template <typename MT>
struct ResidentU {
  static constexpr int value = 5;
};
template <>
struct ResidentU<double> {
  static constexpr int value = 3;
};

template <int G, typename MT>
struct RowRes {
  static constexpr int U = ResidentU<MT>::value * 4 / G > 0 ? ResidentU<MT>::value * 4 / G : 1;
  int p0, p1;          // block range of the row (p1 == p0: no row)
  int c[U];
  MT v[U][9];          // zero beyond the row's end (the column then is the row itself)
  MT d0, d1, d2;       // row `lane` (< 3) of the inverse diagonal block
};

template <int G, typename MT>
__device__ __forceinline__ void row_preload(const SysDesc& S, const MT* __restrict__ val, const MT* __restrict__ inv,
                                            int row, int lane, RowRes<G, MT>& rr) {
  constexpr int U = RowRes<G, MT>::U;
  rr.p0 = rr.p1 = 0;
  rr.d0 = rr.d1 = rr.d2 = MT(0);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    rr.c[u] = 0;
#pragma unroll
    for (int e = 0; e < 9; ++e) rr.v[u][e] = MT(0);
  }
  if (row >= S.nb) return;
  rr.p0 = S.rowptr[row];
  rr.p1 = S.rowptr[row + 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int p = rr.p0 + lane + u * G;
    rr.c[u] = row;
    if (p < rr.p1) {
      rr.c[u] = S.colidx[p];
#pragma unroll
      for (int e = 0; e < 9; ++e) rr.v[u][e] = val[val_index(p, e)];
    }
  }
  if (lane < 3) {
    rr.d0 = inv[9 * row + 3 * lane];
    rr.d1 = inv[9 * row + 3 * lane + 1];
    rr.d2 = inv[9 * row + 3 * lane + 2];
  }
}

// row_spmv with the first U blocks of each lane from registers.  The gathers
// of the resident blocks are issued unconditionally (absent blocks gather
// the row itself, rows past the system column 0, with zero values), then
// `pre` issues the caller's own loads of the row, and only then is anything
// accumulated: every load of a pass is in flight together, one memory
// latency per pass.  Blocks beyond the row's end are skipped by a select,
// so each lane adds exactly its blocks p0+lane, p0+lane+G, ... in order.
template <int G, typename MT, typename Src, typename Pre>
__device__ __forceinline__ void row_spmv_res(const SysDesc& S, const MT* __restrict__ val, const RowRes<G, MT>& rr,
                                             int lane, Src src, Pre pre, MT& a0, MT& a1, MT& a2) {
  constexpr int U = RowRes<G, MT>::U;
  a0 = a1 = a2 = MT(0);
  MT x[U][3];
#pragma unroll
  for (int u = 0; u < U; ++u) src(rr.c[u], x[u][0], x[u][1], x[u][2]);
  pre();
#pragma unroll
  for (int u = 0; u < U; ++u) blk_step(rr.p0 + lane + u * G < rr.p1, rr.v[u], x[u][0], x[u][1], x[u][2], a0, a1, a2);
  if (rr.p0 + lane + U * G < rr.p1) {
    const auto colidx = gsel<MT>(S.colidx);
    for (int p = rr.p0 + lane + U * G; p < rr.p1; p += G) {
      const int c = colidx[p];
      MT x0, x1, x2;
      src(c, x0, x1, x2);
      const auto v = gsel<MT>(val) + val_index(p, 0);
      const MT vb[9] = {v[0],             v[kValChunk],     v[2 * kValChunk], v[3 * kValChunk], v[4 * kValChunk],
                        v[5 * kValChunk], v[6 * kValChunk], v[7 * kValChunk], v[8 * kValChunk]};
      blk_step(true, vb, x0, x1, x2, a0, a1, a2);
    }
  }
  a0 = group_sum<G>(a0);
  a1 = group_sum<G>(a1);
  a2 = group_sum<G>(a2);
}

// z store of a preconditioner pass: plain, or (WT) a relaxed agent-scope
// atomic store, written through this XCD's L2 to the coherence point
template <bool WT, typename VT>
__device__ __forceinline__ void zst(VT* p, VT v) {
  if constexpr (WT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// Neumann step on the resident row tile (same arithmetic as tile_N).  The
// row's r and z are loaded after the gathers are issued, by every lane (a
// lane that owns no output reads row 0), so no branch holds them back.
template <int G, typename VT, bool WT = false>
__device__ __forceinline__ void tile_N_res(const SysDesc& S, int row, const RowRes<G, VT>& rr,
                                           const VT* __restrict__ zin, VT* __restrict__ zout, VT omega, bool last,
                                           double& rz) {
  const int lane = threadIdx.x % G;
  const bool own = row < S.nb && lane < 3;
  const int ro = own ? row : 0, lo = own ? lane : 0;
  double r0, r1, r2;
  VT zo;
  const auto zg = gbl(zin);
  const auto rg = gbl(S.r);
  VT a0, a1, a2;
  row_spmv_res<G, VT>(S, mat_vals<VT>(S), rr, lane,
                      [&](int c, VT& x0, VT& x1, VT& x2) {
                        x0 = zg[3 * c];
                        x1 = zg[3 * c + 1];
                        x2 = zg[3 * c + 2];
                      },
                      [&]() {
                        r0 = rg[3 * ro];
                        r1 = rg[3 * ro + 1];
                        r2 = rg[3 * ro + 2];
                        zo = zg[3 * ro + lo];
                      },
                      a0, a1, a2);
  if (own) {
    const VT out = pass_out(zo, omega, rr.d0, rr.d1, rr.d2, r0, r1, r2, a0, a1, a2);
    zst<WT>(zout + 3 * row + lane, out);
    if (last) rz += pick3(lane, r0, r1, r2) * (double)out;
  }
}

// stage A on the resident row tile (same arithmetic as tile_A; loads as in
// tile_N_res)
template <int G, typename VT>
__device__ __forceinline__ void tile_A_res(const SysDesc& S, int row, const RowRes<G, VT>& rr,
                                           const VT* __restrict__ z, const VT* __restrict__ pold,
                                           VT* __restrict__ pnew, VT* __restrict__ q, VT beta, double& dot) {
  const int lane = threadIdx.x % G;
  const bool own = row < S.nb && lane < 3;
  const int io = own ? 3 * row + lane : 0;
  VT zo, po;
  const auto zg = gbl(z);
  const auto pg = gbl(pold);
  VT a0, a1, a2;
  row_spmv_res<G, VT>(S, mat_vals<VT>(S), rr, lane,
                      [&](int c, VT& x0, VT& x1, VT& x2) {
                        x0 = combine(zg[3 * c], pg[3 * c], beta);
                        x1 = combine(zg[3 * c + 1], pg[3 * c + 1], beta);
                        x2 = combine(zg[3 * c + 2], pg[3 * c + 2], beta);
                      },
                      [&]() {
                        zo = zg[io];
                        po = pg[io];
                      },
                      a0, a1, a2);
  if (own) {
    const int idx = 3 * row + lane;
    const VT pn = combine(zo, po, beta);
    const VT qv = pick3(lane, a0, a1, a2);
    pnew[idx] = pn;
    q[idx] = qv;
    dot += (double)pn * (double)qv;
  }
}

// D x in FP64 on one row tile (true residual / operator tests)
template <int G, int U = 1>
__device__ __forceinline__ void tile_Dx64(const SysDesc& S, int tl, const double* __restrict__ x,
                                          double* __restrict__ y) {
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  const int row = tile_row(S, tl, kBlock / G, g);
  double a0, a1, a2;
  row_spmv<G, double, U>(S, S.val64, row, lane,
                      [&](int c, double& x0, double& x1, double& x2) {
                        x0 = x[3 * c];
                        x1 = x[3 * c + 1];
                        x2 = x[3 * c + 2];
                      },
                      a0, a1, a2);
  if (row < S.nb && lane < 3) y[3 * row + lane] = pick3(lane, a0, a1, a2);
}

// low-rank partial t[l] += sum_{j in tile} W[j,l] src(act_dof[j]); acc[c]
// holds column l = lane + 64c.  Each wave owns RW consecutive rows of the
// tile: lanes 0..RW-1 fetch the source values at once, then every lane
// issues its W loads for RB rows at a time before the multiply-adds; with
// FP32 W both batches of a wave's rows are in flight together.  Loads are
// unconditional (indices clamped into the tile) and masked by a select.
template <typename WT, typename Src>
__device__ __forceinline__ void tile_Wt(const SysDesc& S, int tl, const WT* __restrict__ W, Src src,
                                        double acc[4]) {
  constexpr int RW = kLrTile / (kBlock / 64);  // rows per wave
  constexpr int RB = 8;                        // rows per batch of W loads
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j0 = tl * kLrTile;
  const int jw = j0 + wave * RW;
  const int j1 = min(j0 + kLrTile, S.a);
  const int k = S.k;
  double my_src = 0.0;
  if (lane < RW && jw + lane < j1) my_src = src(S.act_dof[jw + lane]);
  auto batch = [&](int jb, int c) {
    const int l = lane + 64 * c;
    const int lc = min(l, k - 1);
    WT w[RB];
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) {
      const int j = min(jw + jb + jj, j1 - 1);
      w[jj] = W[(size_t)j * k + lc];
    }
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) {
      const bool ok = l < k && jw + jb + jj < j1;
      acc[c] += (ok ? (double)w[jj] : 0.0) * __shfl(my_src, jb + jj, 64);
    }
  };
  if (jw >= j1) return;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    if (64 * c >= k) break;
    if constexpr (sizeof(WT) == 4) {
#pragma unroll
      for (int jb = 0; jb < RW; jb += RB) batch(jb, c);
    } else {
#pragma unroll 1
      for (int jb = 0; jb < RW; jb += RB) batch(jb, c);
    }
  }
}

// List the coupled nodes of one node tile in LDS (an = the node's first W
// row, -1 if uncoupled): s_list[slot] = an, s_slot[tid] = slot.  Uniform
// over the workgroup.
__device__ __forceinline__ void lowrank_list(int an, int* s_list, int* s_slot, int* s_cnt) {
  const int tid = threadIdx.x;
  if (tid == 0) *s_cnt = 0;
  __syncthreads();
  if (an >= 0) {
    const int slot = atomicAdd(s_cnt, 1);
    s_list[slot] = an;
    s_slot[tid] = slot;
  }
  __syncthreads();
}

// g_d = sum_l W[an+d, l] t[l] for the listed coupled nodes: groups of 4
// lanes take one node each (128 nodes per pass of the workgroup), every lane
// strides the rank by 4 and has its W loads for the three rows in flight
// before the group sum.  Uniform over the workgroup.
template <typename WT>
__device__ __forceinline__ void lowrank_apply(int k, int an, const WT* __restrict__ W, const double* __restrict__ t,
                                              double g[3], const int* s_list, const int* s_slot,
                                              double (*s_g)[3], const int* s_cnt, const WT* wc = nullptr) {
  constexpr int GL = 4;               // lanes per node
  constexpr int NPR = kBlock / GL;    // nodes per pass
  const int tid = threadIdx.x;
  const int m = *s_cnt;
  const int sub = tid % GL, grp = tid / GL;
  for (int base = 0; base < m; base += NPR) {
    const int it = base + grp;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    if (it < m) {
      // W rows of the listed node: from the LDS cache (list order) or global
      const WT* w0 = wc ? wc + (size_t)it * 3 * k : W + (size_t)s_list[it] * k;
      const WT* w1 = w0 + k;
      const WT* w2 = w1 + k;
#pragma unroll 8
      for (int l = sub; l < k; l += GL) {
        const double tl_ = t[l];
        p0 += (double)w0[l] * tl_;
        p1 += (double)w1[l] * tl_;
        p2 += (double)w2[l] * tl_;
      }
    }
    p0 = group_sum<GL>(p0);
    p1 = group_sum<GL>(p1);
    p2 = group_sum<GL>(p2);
    if (it < m && sub == 0) {
      s_g[it][0] = p0;
      s_g[it][1] = p1;
      s_g[it][2] = p2;
    }
  }
  __syncthreads();
  if (an >= 0) {
    const int slot = s_slot[tid];
    g[0] = s_g[slot][0];
    g[1] = s_g[slot][1];
    g[2] = s_g[slot][2];
  }
}

template <typename WT>
__device__ __forceinline__ void tile_lowrank_g(const SysDesc& S, int an, const WT* __restrict__ W,
                                               const double* __restrict__ t, double g[3], int* s_list,
                                               int* s_slot, double (*s_g)[3], int* s_cnt) {
  lowrank_list(an, s_list, s_slot, s_cnt);
  lowrank_apply(S.k, an, W, t, g, s_list, s_slot, s_g, s_cnt);
}

// sum over the four waves' acc, written to dst[l] for l < k (thread-ordered)
__device__ __forceinline__ void flush_Wt(double acc[4], int k, double* dst, double (*lr)[64]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = 0; c < 4; ++c) {
    if (64 * c >= k) break;
    lr[wave][lane] = acc[c];
    __syncthreads();
    if (threadIdx.x < 64) {
      const int l = lane + 64 * c;
      double s = lr[0][lane];
      for (int w = 1; w < kBlock / 64; ++w) s += lr[w][lane];
      if (l < k) dst[l] = s;
    }
    __syncthreads();
    acc[c] = 0.0;
  }
  for (int c = 0; c < 4; ++c) acc[c] = 0.0;
}

// ---------------------------------------------------------------- helpers
struct TileRange {
  int t0, t1;
};

__device__ __forceinline__ int sys_of_tile_begin(const SysDesc* sys, int nsys, int t, int which) {
  int s = 0;
  for (; s < nsys; ++s) {
    const int e = which == 0 ? sys[s].rt_end : (which == 1 ? sys[s].et_end : sys[s].lt_end);
    if (t < e) break;
  }
  return s;
}

__device__ __forceinline__ int tiles_begin(const SysDesc& S, int which) {
  return which == 0 ? S.rt_begin : (which == 1 ? S.et_begin : S.lt_begin);
}
__device__ __forceinline__ int tiles_end(const SysDesc& S, int which) {
  return which == 0 ? S.rt_end : (which == 1 ? S.et_end : S.lt_end);
}

// Iterate the tiles [t0,t1) of a tiling grouped by system: body(s, local_tile)
// for each tile, flush(s) after the last tile of each system.  Uniform over
// the workgroup.
template <typename Body, typename Flush>
__device__ __forceinline__ void for_tiles(const SysDesc* sys, int nsys, TileRange tr, int which,
                                          Body body, Flush flush) {
  if (tr.t0 >= tr.t1) return;
  int s = sys_of_tile_begin(sys, nsys, tr.t0, which);
  int t = tr.t0;
  while (t < tr.t1 && s < nsys) {
    const int b = tiles_begin(sys[s], which);
    const int e = min(tr.t1, tiles_end(sys[s], which));
    for (; t < e; ++t) body(s, t - b);
    flush(s);
    ++s;
  }
}

// Chunks that own no tile of a tiling (T tiles over nC chunks) never write a
// partial; their slots are skipped by the reductions.
__device__ __forceinline__ bool chunk_empty(int c, int nC, int T) {
  if (nC <= T) return false;  // every chunk owns at least one tile (no division needed)
  return tile_of_chunk(c, nC, T) == tile_of_chunk(c + 1, nC, T);
}

// reduce_chunks_all in two halves: the first kPreN chunk loads of every lane
// issued early (so that they overlap other work), the sums later in the same
// order as reduce_chunks_all adds them
constexpr int kPreN = 3;
struct ChunkPre {
  double x[kPreN];
  int c0, c1, s, l, L;
  bool act;
};
template <typename Rng, typename Want>
__device__ __forceinline__ void chunks_prefetch(const double* __restrict__ part, int nsys, int nC, Rng rng,
                                                Want want, ChunkPre& pf) {
  int L = 64;
  while (L > 1 && L * nsys > kBlock) L >>= 1;
  pf.L = L;
  pf.s = threadIdx.x / L;
  pf.l = threadIdx.x % L;
  pf.act = pf.s < nsys && want(pf.s);
  pf.c0 = 0;
  pf.c1 = -1;
  if (pf.act) rng(pf.s, pf.c0, pf.c1);
#pragma unroll
  for (int i = 0; i < kPreN; ++i) {
    const int c = pf.c0 + pf.l + i * L;
    pf.x[i] = pf.act && c <= pf.c1 ? part[(size_t)pf.s * nC + c] : 0.0;
  }
}
__device__ __forceinline__ void chunks_finish(const double* __restrict__ part, int nC, int T, const ChunkPre& pf,
                                              double* out) {
  double v = 0.0;
  if (pf.act) {
#pragma unroll
    for (int i = 0; i < kPreN; ++i) {
      const int c = pf.c0 + pf.l + i * pf.L;
      if (c <= pf.c1) v += chunk_empty(c, nC, T) ? 0.0 : pf.x[i];
    }
    for (int c = pf.c0 + pf.l + kPreN * pf.L; c <= pf.c1; c += pf.L) {
      const double x = part[(size_t)pf.s * nC + c];
      v += chunk_empty(c, nC, T) ? 0.0 : x;
    }
  }
  for (int m = pf.L >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if (pf.act && pf.l == 0) out[pf.s] = v;
}

// Per-system sums of chunk partials for every system at once: groups of L
// lanes (L = 64 for up to 8 systems, halved until all systems fit in the
// workgroup) take one system each, every lane strides the system's chunks
// by L with all loads issued before the sums, and the group adds its lanes
// by butterfly (for up to 8 systems: wave w sums system w over its 64
// lanes).  out[s] is valid after the next __syncthreads.
template <typename Rng, typename Want>
__device__ __forceinline__ void reduce_chunks_all(const double* __restrict__ part, int nsys, int nC, int T, Rng rng,
                                                  Want want, double* out) {
  int L = 64;
  while (L > 1 && L * nsys > kBlock) L >>= 1;
  const int tid = threadIdx.x;
  const int s = tid / L, l = tid % L;
  double v = 0.0;
  const bool act = s < nsys && want(s);
  if (act) {
    int c0, c1;
    rng(s, c0, c1);
#pragma unroll 4
    for (int c = c0 + l; c <= c1; c += L) {
      const double x = part[(size_t)s * nC + c];
      v += chunk_empty(c, nC, T) ? 0.0 : x;
    }
  }
  for (int m = L >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if (act && l == 0) out[s] = v;
}

// Four reductions of reduce_chunks_all's kind over one chunk range per
// system (arrays a, b with nA chunks per system; f, e with nF), their loads
// in flight together; each sum is added in reduce_chunks_all's order.
template <typename Rng>
__device__ __forceinline__ void reduce4_chunks(const double* __restrict__ f, int nF, const double* __restrict__ a,
                                               const double* __restrict__ b2, int nA, const double* __restrict__ e,
                                               int nsys, int T, Rng rng, double* of, double* oa, double* ob,
                                               double* oe) {
  int L = 64;
  while (L > 1 && L * nsys > kBlock) L >>= 1;
  const int tid = threadIdx.x;
  const int s = tid / L, l = tid % L;
  double vf = 0.0, va = 0.0, vb = 0.0, ve = 0.0;
  const bool act = s < nsys;
  if (act) {
    int c0, c1;
    rng(s, c0, c1);
#pragma unroll 2
    for (int c = c0 + l; c <= c1; c += L) {
      const double xf = f[(size_t)s * nF + c], xa = a[(size_t)s * nA + c], xb = b2[(size_t)s * nA + c];
      const double xe = e[(size_t)s * nF + c];
      const bool emp = chunk_empty(c, nF, T);
      vf += emp ? 0.0 : xf;
      va += emp ? 0.0 : xa;
      vb += emp ? 0.0 : xb;
      ve += emp ? 0.0 : xe;
    }
  }
  for (int m = L >> 1; m >= 1; m >>= 1) {
    vf += __shfl_xor(vf, m, 64);
    va += __shfl_xor(va, m, 64);
    vb += __shfl_xor(vb, m, 64);
    ve += __shfl_xor(ve, m, 64);
  }
  if (act && l == 0) {
    of[s] = vf;
    oa[s] = va;
    ob[s] = vb;
    oe[s] = ve;
  }
}

// |t_s|^2 for every system at once (same grouping as reduce_chunks_all)
template <typename Want>
__device__ __forceinline__ void sum_tt_all(const SysDesc* sys, int nsys, const double* t, Want want, double* out) {
  int L = 64;
  while (L > 1 && L * nsys > kBlock) L >>= 1;
  const int tid = threadIdx.x;
  const int s = tid / L, l = tid % L;
  double v = 0.0;
  const bool in = s < nsys;
  const bool act = in && want(s);
  if (act)
    for (int j = l; j < sys[s].k; j += L) v += t[sys[s].t_off + j] * t[sys[s].t_off + j];
  for (int m = L >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if (in && l == 0) out[s] = act ? v : 0.0;
}

// ---------------------------------------------------------------- kernel
template <int G, typename VT, bool RES, bool PIPE = false>
__global__ __launch_bounds__(kBlock) void pcg_persistent(PcgParams P) {
  constexpr int SU = StreamU<G, RES>::value;
  // pass pairs: FP32 resident rows at 4 lanes per row, standard loop
  constexpr bool kPairs = G == 4 && RES && !PIPE && sizeof(VT) == 4;
  // dynamic LDS: the system descriptors, the low-rank sums of the node
  // tile's coupled nodes, then the t vectors (sum of ranks), the coupled W
  // rows of the resident node tile and the halo region
  extern __shared__ __align__(16) double s_dyn[];
  SysDesc* const s_sys = reinterpret_cast<SysDesc*>(s_dyn);
  double(*const s_g)[3] = reinterpret_cast<double(*)[3]>(reinterpret_cast<char*>(s_dyn) + P.sys_bytes);
  double* const s_t = reinterpret_cast<double*>(reinterpret_cast<char*>(s_dyn) + P.sys_bytes + P.sg_bytes);
  __shared__ double s_rz[kMaxSys], s_alpha[kMaxSys], s_beta[kMaxSys], s_normb[kMaxSys];
  __shared__ double s_rel[kMaxSys], s_v0[kMaxSys], s_v1[kMaxSys], s_tt[kMaxSys];
  __shared__ int s_st[kMaxSys], s_it[kMaxSys], s_chk[kMaxSys], s_rs[kMaxSys];
  __shared__ int s_nchk[kMaxSys], s_nrs[kMaxSys];
  __shared__ double s_red[kBlock / 64];
  __shared__ double s_lr[kBlock / 64][64];
  __shared__ int s_any, s_pfail;
  __shared__ int s_list[kNodeTileMax], s_slot[kNodeTileMax], s_cnt;
  __shared__ double s_tred[kBlock];
  __shared__ int s_pairs;

  // the descriptors live in LDS for the launch: after every grid barrier's
  // acquire they are read without a trip to memory
  {
    const int words = (int)(P.nsys * sizeof(SysDesc) / sizeof(int));
    const int* src = reinterpret_cast<const int*>(P.sys);
    int* dst = reinterpret_cast<int*>(s_sys);
    for (int e = threadIdx.x; e < words; e += kBlock) dst[e] = src[e];
    __syncthreads();
  }
  const SysDesc* sys = s_sys;
  const int nsys = P.nsys;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const bool is_lr = b < P.n_lr_wg;
  GridBarrier gbar;
  // step weights (all equal to omega for the Neumann series)
  auto step_w = [&](int j) { return (VT)P.omegas[j < kMaxPolySteps ? j : kMaxPolySteps - 1]; };
  const VT omega = step_w(0);

  TileRange trA{0, 0}, trF{0, 0}, trE{0, 0}, trL{0, 0};
  const int cA = P.chunkA[b], cF = P.chunkF[b], cL = P.chunkL[b];
  if (!is_lr) trA = {tile_of_chunk(cA, P.nA, P.rt_total), tile_of_chunk(cA + 1, P.nA, P.rt_total)};
  if (cF >= 0) {  // (no row or node tiles for the low-rank workgroups of the aligned split)
    trF = {tile_of_chunk(cF, P.nF, P.rt_total), tile_of_chunk(cF + 1, P.nF, P.rt_total)};
    trE = {tile_of_chunk(cF, P.nF, P.et_total), tile_of_chunk(cF + 1, P.nF, P.et_total)};
  }
  if (is_lr) trL = {tile_of_chunk(cL, P.nL, P.lt_total), tile_of_chunk(cL + 1, P.nL, P.lt_total)};

  for (int s = tid; s < nsys; s += kBlock) {
    s_st[s] = ST_RUN;
    s_it[s] = 0;
    s_chk[s] = CHK_NONE;
    s_rs[s] = 0;
    s_nchk[s] = 0;
    s_nrs[s] = 0;
    s_rel[s] = 1.0;
    s_beta[s] = 0.0;
    s_rz[s] = 0.0;
  }
  __syncthreads();

  // optional stage timers (workgroup-private slots, thread 0 only)
  long long t_last = 0;
  auto sync_at = [&](int si) -> bool {
    long long t_before = 0;
    if (P.prof && tid == 0) t_before = rt_ticks();
    const bool ok = gb_sync(gbar);
    if (P.prof && tid == 0) {
      const long long t_after = rt_ticks();
      long long* slot = P.prof + (size_t)b * 32;
      slot[2 * si] += t_before - t_last;
      slot[2 * si + 1] += t_after - t_before;
      slot[20 + si] += 1;
      t_last = t_after;
    }
    return ok;
  };

  // resident Neumann row tile (exactly one row tile in the Neumann chunking)
  const bool resF = RES && trF.t1 - trF.t0 == 1;
  int sF = 0, rowF = 0;
  RowRes<G, VT> rrF;
  if (resF) {
    sF = sys_of_tile_begin(sys, nsys, trF.t0, 0);
    rowF = tile_row(sys[sF], trF.t0 - sys[sF].rt_begin, kBlock / G, tid / G);
    row_preload<G, VT>(sys[sF], mat_vals<VT>(sys[sF]), inv_vals<VT>(sys[sF]), rowF, tid % G, rrF);
  }
  // stage A on the same resident tile (aligned split)
  const bool resA = resF && !is_lr && P.nA == P.nF && cA == cF;
  // resident node tile (exactly one node tile): its node and coupled-node
  // list are fixed for the launch, and stage B's vector loads are issued
  // before the stage-A reductions
  const bool resE = RES && trE.t1 - trE.t0 == 1;
  int sE = 0, iE = 0, anE = -1;
  if (resE) {
    sE = sys_of_tile_begin(sys, nsys, trE.t0, 1);
    iE = tile_row(sys[sE], trE.t0 - sys[sE].et_begin, P.node_tile, tid);
    if (tid < P.node_tile && iE < sys[sE].nb && sys[sE].k > 0) anE = sys[sE].act_node[iE];
  }

  const bool started = gb_init(gbar, P.barrier, P.error, P.timeout_ticks, (unsigned)P.nwg);
  if (P.prof && tid == 0) t_last = rt_ticks();
  if (started) {
  if (resE) lowrank_list(anE, s_list, s_slot, &s_cnt);  // kept for the whole launch
  // W rows of the resident tile's coupled nodes, cached in LDS for stage B
  VT* s_wc = nullptr;
  if (resE && P.wc_rows > 0 && sys[sE].k > 0 && s_cnt <= P.wc_rows) {
    s_wc = reinterpret_cast<VT*>(reinterpret_cast<char*>(s_t) + P.wc_off);
    const SysDesc& S = sys[sE];
    const int k = S.k, m = s_cnt, n3k = 3 * k;
    const VT* W = w_vals<VT>(S);
    for (int e = tid; e < m * n3k; e += kBlock) {
      const int slot = e / n3k, rem = e - slot * n3k;
      s_wc[e] = W[(size_t)s_list[slot] * k + rem];
    }
    __syncthreads();
  }
  // pass pairs: on when every system's halo packs were built (the same
  // decision in every workgroup); each row workgroup copies its tile's pack
  // into LDS for the launch
  unsigned char* const s_halo = reinterpret_cast<unsigned char*>(s_t) + P.halo_off;
  if constexpr (kPairs) {
    if (tid == 0) {
      int ok = P.halo;
      for (int s = 0; s < nsys && ok; ++s)
        ok = sys[s].halo != nullptr && sys[s].zt != nullptr && *reinterpret_cast<const int*>(sys[s].halo) == 0;
      s_pairs = ok;
    }
    __syncthreads();
    if (s_pairs && resF) {
      const SysDesc& S = sys[sF];
      const unsigned char* pk = S.halo + kHaloBaseBytes + (size_t)(trF.t0 - S.rt_begin) * P.halo_stride;
      const int nblk = reinterpret_cast<const int*>(pk)[2];
      for (int e = tid; e < kHaloOffBlk / 16; e += kBlock)
        reinterpret_cast<int4*>(s_halo)[e] = reinterpret_cast<const int4*>(pk)[e];
      // the halo blocks' column slots (ushort pairs), then their 9 value planes
      for (int e = tid; e < (nblk + 1) / 2; e += kBlock)
        reinterpret_cast<int*>(s_halo + kHaloOffBlk)[e] = reinterpret_cast<const int*>(pk + kHaloOffBlk)[e];
      const size_t voff = halo_val_off(P.hbl);
      for (int e = tid; e < 9 * nblk; e += kBlock)
        reinterpret_cast<float*>(s_halo + voff)[e] = reinterpret_cast<const float*>(pk + voff)[e];
      __syncthreads();
    }
  }

  int zc = 0, pc = 0;
  const int order = P.order;
  double* partE_rr = P.partE;
  double* partE_rz = P.partE + (size_t)nsys * P.nF;

  // ---- node-tile stage: z0 = w Dinv r (and residual norms) ------------
  // kind 0: init (r == rhs), 1: stage B, 2: T2
  // stage-B vectors of the resident node tile, loaded before the reductions
  VT preq[3] = {VT(0), VT(0), VT(0)}, prep[3] = {VT(0), VT(0), VT(0)};
  double prer[3] = {0.0, 0.0, 0.0}, prex[3] = {0.0, 0.0, 0.0};
  VT pi0 = VT(0), pi1 = VT(0), pi2 = VT(0), pi3 = VT(0), pi4 = VT(0), pi5 = VT(0), pi6 = VT(0), pi7 = VT(0),
     pi8 = VT(0);
  bool pre_ok = false;
  auto prefetch_B = [&]() {
    pre_ok = false;
    if (!resE || s_st[sE] != ST_RUN) return;
    const SysDesc& S = sys[sE];
    if (tid >= P.node_tile || iE >= S.nb) return;
    const int i3 = 3 * iE;
    const VT* q = (const VT*)S.q;
    const VT* pn = (const VT*)S.p[pc];
    const VT* inv = inv_vals<VT>(S) + 9 * iE;
    for (int d = 0; d < 3; ++d) {
      preq[d] = q[i3 + d];
      prep[d] = pn[i3 + d];
      prer[d] = S.r[i3 + d];
      prex[d] = S.x[i3 + d];
    }
    pi0 = inv[0];
    pi1 = inv[1];
    pi2 = inv[2];
    pi3 = inv[3];
    pi4 = inv[4];
    pi5 = inv[5];
    pi6 = inv[6];
    pi7 = inv[7];
    pi8 = inv[8];
    pre_ok = true;
  };

  // tsel: the t vector of the true residual's low-rank term (0: s_t; 2: the
  // pipelined loop's W^T x copy), chosen by index so its loads stay LDS loads
  auto node_stage = [&](int kind, int tsel) {
    const double* tx = s_t + (tsel ? 2 * P.t_total : 0);
    double rr = 0.0, rz = 0.0;
    for_tiles(sys, nsys, trE, 1,
              [&](int s, int tl) {
                const SysDesc& S = sys[s];
                const int st = s_st[s];
                if (st != ST_RUN) return;
                if (kind == 2 && s_chk[s] == CHK_NONE) return;
                const int i = tile_row(S, tl, P.node_tile, tid);
                const bool valid = tid < P.node_tile && i < S.nb;
                double g[3] = {0.0, 0.0, 0.0};
                if (kind != 0 && S.k > 0) {
                  const int an = resE ? anE : (valid ? S.act_node[i] : -1);
                  if (!resE) lowrank_list(an, s_list, s_slot, &s_cnt);
                  if (kind == 1)
                    lowrank_apply(S.k, an, w_vals<VT>(S), s_t + S.t_off, g, s_list, s_slot, s_g, &s_cnt,
                                  resE ? s_wc : (const VT*)nullptr);
                  else
                    lowrank_apply(S.k, an, S.W64, tx + S.t_off, g, s_list, s_slot, s_g, &s_cnt);
                }
                if (!valid) return;
                const int i3 = 3 * i;
                double rv[3];
                bool write_r = false, write_z = true;
                if (kind == 0) {
                  rv[0] = S.r[i3];
                  rv[1] = S.r[i3 + 1];
                  rv[2] = S.r[i3 + 2];
                } else {
                  if (kind == 1) {
                    const double al = s_alpha[s];
                    if (pre_ok) {
                      for (int d = 0; d < 3; ++d) {
                        const double qd = (double)preq[d] - g[d];
                        rv[d] = prer[d] - al * qd;
                        S.x[i3 + d] = prex[d] + al * (double)prep[d];
                      }
                    } else {
                      const VT* q = (const VT*)S.q;
                      const VT* pn = (const VT*)S.p[pc];
                      for (int d = 0; d < 3; ++d) {
                        const double qd = (double)q[i3 + d] - g[d];
                        rv[d] = S.r[i3 + d] - al * qd;
                        S.x[i3 + d] += al * (double)pn[i3 + d];
                      }
                    }
                    write_r = true;
                  } else {
                    for (int d = 0; d < 3; ++d) rv[d] = S.rhs[i3 + d] - S.q64[i3 + d] + g[d];
                    const int ck = s_chk[s];
                    write_r = (ck == CHK_CONVERGE || ck == CHK_REPLACE);
                    write_z = write_r;
                  }
                }
                if (write_r)
                  for (int d = 0; d < 3; ++d) S.r[i3 + d] = rv[d];
                rr += rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                if (write_z) {
                  VT i0 = pi0, i1 = pi1, i2 = pi2, i3v = pi3, i4 = pi4, i5 = pi5, i6 = pi6, i7 = pi7, i8 = pi8;
                  if (!pre_ok) {
                    const VT* ig = inv_vals<VT>(S) + 9 * i;
                    i0 = ig[0];
                    i1 = ig[1];
                    i2 = ig[2];
                    i3v = ig[3];
                    i4 = ig[4];
                    i5 = ig[5];
                    i6 = ig[6];
                    i7 = ig[7];
                    i8 = ig[8];
                  }
                  VT* z = (VT*)S.z[zc];
                  const VT r0 = (VT)rv[0], r1 = (VT)rv[1], r2 = (VT)rv[2];
                  const VT z0 = omega * (i0 * r0 + i1 * r1 + i2 * r2);
                  const VT z1 = omega * (i3v * r0 + i4 * r1 + i5 * r2);
                  const VT z2 = omega * (i6 * r0 + i7 * r1 + i8 * r2);
                  z[i3] = z0;
                  z[i3 + 1] = z1;
                  z[i3 + 2] = z2;
                  if (order == 1) {
                    rz += rv[0] * (double)z0;
                    rz += rv[1] * (double)z1;
                    rz += rv[2] * (double)z2;
                  }
                }
              },
              [&](int s) {
                const bool active = s_st[s] == ST_RUN && (kind != 2 || s_chk[s] != CHK_NONE);
                // T2 in norm-only mode keeps stage B's z and <r,z> partials
                const bool new_z = active && (kind != 2 || s_chk[s] == CHK_CONVERGE || s_chk[s] == CHK_REPLACE);
                // <r, z0> is the preconditioned product only for order 1
                // (higher orders take it from the last Neumann pass)
                const double a = block_sum_t0(rr, s_red);
                const double c = order == 1 ? block_sum_t0(rz, s_red) : 0.0;
                if (tid == 0) {
                  if (active) partE_rr[(size_t)s * P.nF + cF] = a;
                  if (new_z && order == 1) partE_rz[(size_t)s * P.nF + cF] = c;
                }
                rr = 0.0;
                rz = 0.0;
              });
  };

  // ---- point-to-point sync between preconditioner passes --------------
  // A pass reads z only at the columns of its tile's rows, so pass j+1 of a
  // row tile may start once the tiles its rows reference (and, for a
  // symmetric pattern, the tiles that reference it) have finished pass j.
  // With two z buffers this also covers the reuse of a buffer: a neighbour
  // that has finished pass j has finished reading the buffer pass j+1
  // overwrites.  A pass followed by such a wait stores z with agent-scope
  // atomic stores (written through to the coherence point), waits for them
  // to complete and then publishes the tile's pass count: for stores that
  // are themselves agent-scope atomics that wait is the release, and no
  // writeback of this XCD's L2 is needed (measured: a writeback per
  // workgroup and pass costs more than the grid barrier it replaces).  The
  // waiter polls its neighbours' counts and acquires (invalidating stale
  // lines in its L1 and L2), then reads z with plain loads, so the
  // arithmetic is the grid-barrier path's, bit for bit.  Only the resident
  // row workgroups of the aligned split take part; the low-rank workgroups
  // idle until the next grid barrier.
  unsigned pseq = 0;
  auto pass_wait = [&]() -> bool {
    pseq += 1;
    if (!resF) return true;
    long long t_before = 0;
    if (P.prof && tid == 0) t_before = rt_ticks();
    wait_vm_zero();
    __syncthreads();
    const SysDesc& S = sys[sF];
    const int tg = trF.t0, tl = tg - S.rt_begin;
    if (tid == 0) {
      __hip_atomic_store(P.pass_flags + tg, pseq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_pfail = 0;
    }
    __syncthreads();
    if (s_st[sF] == ST_RUN) {
      const int* L = S.nbr + (size_t)tl * kNbrStride;
      const int cnt = L[0];
      const int n = cnt >= 0 ? cnt : S.rt_end - S.rt_begin;
      const long long t0 = rt_ticks();
      for (int i = tid; i < n; i += kBlock) {
        const int u = S.rt_begin + (cnt >= 0 ? L[1 + i] : i);
        if (!gb_spin_until(P.pass_flags + u, pseq, P.error, t0, P.timeout_ticks)) s_pfail = 1;
      }
    }
    __syncthreads();
    const bool ok = s_pfail == 0;
    if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    if (P.prof && tid == 0) {
      const long long t_after = rt_ticks();
      long long* slot = P.prof + (size_t)b * 32;
      slot[14] += t_before - t_last;
      slot[15] += t_after - t_before;
      slot[27] += 1;
      t_last = t_after;
    }
    return ok;
  };

  // ---- pass pairs: passes j and j+1 on the resident row tile --------------
  // The z of the tile's node set (own rows, halo rows, their other columns)
  // is gathered once into LDS, with the halo rows' r; pass j then runs on
  // the own rows and the halo rows from LDS, and pass j+1 on the own rows
  // from pass j's values in LDS.  A halo row is computed exactly as its
  // owner computes it (same 4-lane split of its blocks, same order, same
  // helpers), so pass j+1 sees the bits the owner's pass j produced.
  //
  // Between pairs the tiles exchange tagged values instead of waiting on
  // each other: a pass value is published as one 64-bit atomic word (its
  // FP32 bits, the round number above them), stored write-through without
  // waiting, and a pair's gather re-reads each word it needs until the tag
  // is the round it expects.  The word is single-copy atomic, so a matching
  // tag proves its value.  Rounds alternate between two buffers, and a tile
  // cannot overwrite a buffer before every reader is done with it: a tile
  // that reads this tile's rows has a row in this tile's node set (the set
  // reaches two steps, and the pattern is symmetric), so this tile's next
  // gather waits for that reader's output of the round that followed its read.
  // zin: plain input (the first step of an iteration, after a grid barrier),
  // else the tagged input zti of round rin; zout: plain output (the last
  // pair), else the tagged output zto of round rout.  Returns false when a
  // bounded wait expired (error flag set).
  // (r is the same in every pass of an iteration: the own rows' r stays in
  // registers and the halo rows' r in LDS from the first pair on)
  double pr0 = 0.0, pr1 = 0.0, pr2 = 0.0;
  auto pair_passes = [&](const SysDesc& S, const VT* zin, const unsigned long long* zti, unsigned rin, VT* zout,
                         unsigned long long* zto, unsigned rout, VT w1, VT w2, bool last2, bool first,
                         double& rz) -> bool {
    if constexpr (kPairs) {
      constexpr int R = kHaloTileRows;
      const int lane = tid & 3, g = tid >> 2;
      const bool own = rowF < S.nb && lane < 3;
      const int ro = own ? rowF : 0, lo = own ? lane : 0;
      const int tlF = trF.t0 - S.rt_begin;
      const int r0t = tile_first(S, tlF, R), r1t = tile_end(S, tlF, R);  // the tile's rows
      const int* hdr = reinterpret_cast<const int*>(s_halo);
      const int nh = hdr[0], nset = hdr[1];
      const int2* hrows = reinterpret_cast<const int2*>(s_halo + kHaloOffRows);
      const float* hinv = reinterpret_cast<const float*>(s_halo + kHaloOffInv);
      const unsigned short* oslot = reinterpret_cast<const unsigned short*>(s_halo + kHaloOffSlot);
      const int* setn = reinterpret_cast<const int*>(s_halo + kHaloOffSet);
      const unsigned short* bslot = reinterpret_cast<const unsigned short*>(s_halo + kHaloOffBlk);
      const float* hval = reinterpret_cast<const float*>(s_halo + halo_val_off(P.hbl));
      float4* zs = reinterpret_cast<float4*>(s_halo + halo_pack_bytes(P.hbl));
      float4* rs = zs + kHaloSet;
      float4* zl = rs + kHaloRows;
      const auto rg = gbl(S.r);
      // gather: set slots tid and tid + 512 (and, in the first pair of an
      // iteration, r of the own row and of the halo row at slot tid)
      {
        const int k0 = tid, k1 = tid + kBlock;
        const bool v0 = k0 < nset && (k0 >= R || r0t + k0 < r1t);
        const bool v1 = k1 < nset;
        const int n0 = v0 ? (k0 < R ? r0t + k0 : setn[k0 - R]) : 0;
        const int n1 = v1 ? setn[k1 - R] : 0;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c = a;
        if (zti == nullptr) {
          const auto zg = gbl(zin);
          a = make_float4(zg[3 * n0], zg[3 * n0 + 1], zg[3 * n0 + 2], 0.0f);
          c = make_float4(zg[3 * n1], zg[3 * n1 + 1], zg[3 * n1 + 2], 0.0f);
        } else {
          bool need0 = v0, need1 = v1, fail = false;
          const long long t0 = rt_ticks();
          for (;;) {
            unsigned long long w[6];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
              w[e] = __hip_atomic_load(zti + 3 * n0 + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              w[3 + e] = __hip_atomic_load(zti + 3 * n1 + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (need0 && (unsigned)(w[0] >> 32) == rin && (unsigned)(w[1] >> 32) == rin &&
                (unsigned)(w[2] >> 32) == rin) {
              a = make_float4(__uint_as_float((unsigned)w[0]), __uint_as_float((unsigned)w[1]),
                              __uint_as_float((unsigned)w[2]), 0.0f);
              need0 = false;
            }
            if (need1 && (unsigned)(w[3] >> 32) == rin && (unsigned)(w[4] >> 32) == rin &&
                (unsigned)(w[5] >> 32) == rin) {
              c = make_float4(__uint_as_float((unsigned)w[3]), __uint_as_float((unsigned)w[4]),
                              __uint_as_float((unsigned)w[5]), 0.0f);
              need1 = false;
            }
            if (!need0 && !need1) break;
            if (__hip_atomic_load(P.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
              fail = true;
              break;
            }
            if (rt_ticks() - t0 > P.timeout_ticks) {
              __hip_atomic_store(P.error, HSC_ERR_DEVICE_TIMEOUT_CODE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              fail = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          if (fail) s_pfail = 1;
        }
        if (first) {
          const bool hr0 = k0 >= R && k0 < R + nh;
          const double h0 = rg[3 * n0], h1 = rg[3 * n0 + 1], h2 = rg[3 * n0 + 2];
          pr0 = rg[3 * ro];
          pr1 = rg[3 * ro + 1];
          pr2 = rg[3 * ro + 2];
          if (hr0) rs[k0 - R] = make_float4((VT)h0, (VT)h1, (VT)h2, 0.0f);
        }
        if (v0) zs[k0] = a;
        if (v1) zs[k1] = c;
      }
      __syncthreads();
      if (s_pfail) return false;
      const double r0 = pr0, r1 = pr1, r2 = pr2;
      // pass j on the own rows
      {
        float4 x[kHaloU];
#pragma unroll
        for (int u = 0; u < kHaloU; ++u) x[u] = zs[oslot[u * kBlock + tid]];
        VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
#pragma unroll
        for (int u = 0; u < kHaloU; ++u)
          blk_step(rrF.p0 + lane + u * G < rrF.p1, rrF.v[u], x[u].x, x[u].y, x[u].z, a0, a1, a2);
        a0 = group_sum<G>(a0);
        a1 = group_sum<G>(a1);
        a2 = group_sum<G>(a2);
        const float4 zo = zs[g];
        if (own)
          reinterpret_cast<float*>(zl + g)[lane] =
              pass_out(lane == 0 ? zo.x : (lane == 1 ? zo.y : zo.z), w1, rrF.d0, rrF.d1, rrF.d2, r0, r1, r2, a0, a1,
                       a2);
      }
      // pass j on the halo rows.  Up to 128 of them: 4 lanes per row, one
      // round.  More (the tiles where two bodies meet): 2 lanes per row in
      // one round instead of a second round, lane m forming the partial
      // sums of lanes m and m + 2 of the 4-lane split in turn and adding
      // them, then the pair's sum: a_m + a_(m+2), then the two -- the
      // operands and order of the 4-lane butterfly, so every row is summed
      // bit for bit as its owner sums it.
      if (nh > R) {
        const int i = tid >> 1, m = tid & 1;
        const bool hv = i < nh;
        const int2 rw = hv ? hrows[i] : make_int2(0, 0);
        VT s0 = VT(0), s1 = VT(0), s2 = VT(0);
#pragma unroll 1
        for (int h2 = 0; h2 < 2; ++h2) {
          const int vl = m + 2 * h2;
          int pb[kHaloU], sl[kHaloU];
#pragma unroll
          for (int u = 0; u < kHaloU; ++u) {
            const int bi = vl + 4 * u;
            pb[u] = bi < rw.x ? rw.y + bi : 0;
          }
#pragma unroll
          for (int u = 0; u < kHaloU; ++u) sl[u] = bslot[pb[u]];
          float4 xb[kHaloU];
#pragma unroll
          for (int u = 0; u < kHaloU; ++u) xb[u] = zs[sl[u]];
          VT b0 = VT(0), b1 = VT(0), b2 = VT(0);
#pragma unroll
          for (int u = 0; u < kHaloU; ++u) {
            const float* vp = hval + 9 * (size_t)pb[u];
            VT v[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) v[e] = vp[e];
            blk_step(vl + 4 * u < rw.x, v, xb[u].x, xb[u].y, xb[u].z, b0, b1, b2);
          }
          if (h2 == 0) {
            s0 = b0;
            s1 = b1;
            s2 = b2;
          } else {
            s0 = s0 + b0;
            s1 = s1 + b1;
            s2 = s2 + b2;
          }
        }
        s0 = group_sum<2>(s0);
        s1 = group_sum<2>(s1);
        s2 = group_sum<2>(s2);
        if (hv) {
          const float* d = hinv + 9 * i;
          const float4 rr = rs[i];
          const float4 zo = zs[R + i];
          float* zi = reinterpret_cast<float*>(zl + R + i);
          zi[m] = pass_out(m == 0 ? zo.x : zo.y, w1, d[3 * m], d[3 * m + 1], d[3 * m + 2], (double)rr.x, (double)rr.y,
                           (double)rr.z, s0, s1, s2);
          if (m == 0)
            zi[2] = pass_out(zo.z, w1, d[6], d[7], d[8], (double)rr.x, (double)rr.y, (double)rr.z, s0, s1, s2);
        }
      } else
#pragma unroll
      for (int q = 0; q < kHaloRows / R; ++q) {
        if (q * R >= nh) break;  // (the same for the whole workgroup)
        const int i = q * R + g;
        const bool hv = i < nh;
        const int2 rw = hv ? hrows[i] : make_int2(0, 0);
        // the blocks' slots, then their z, then their values: the LDS reads
        // of a level are issued together
        int pb[kHaloU], sl[kHaloU];
#pragma unroll
        for (int u = 0; u < kHaloU; ++u) {
          const int bi = lane + 4 * u;
          pb[u] = bi < rw.x ? rw.y + bi : 0;
        }
#pragma unroll
        for (int u = 0; u < kHaloU; ++u) sl[u] = bslot[pb[u]];
        float4 xb[kHaloU];
#pragma unroll
        for (int u = 0; u < kHaloU; ++u) xb[u] = zs[sl[u]];
        VT b0 = VT(0), b1 = VT(0), b2 = VT(0);
#pragma unroll
        for (int u = 0; u < kHaloU; ++u) {
          const float* vp = hval + 9 * (size_t)pb[u];
          VT v[9];
#pragma unroll
          for (int e = 0; e < 9; ++e) v[e] = vp[e];
          blk_step(lane + 4 * u < rw.x, v, xb[u].x, xb[u].y, xb[u].z, b0, b1, b2);
        }
        b0 = group_sum<G>(b0);
        b1 = group_sum<G>(b1);
        b2 = group_sum<G>(b2);
        if (hv && lane < 3) {
          const float* d = hinv + 9 * i + 3 * lane;
          const float4 rr = rs[i];
          const float4 zo = zs[R + i];
          reinterpret_cast<float*>(zl + R + i)[lane] =
              pass_out(lane == 0 ? zo.x : (lane == 1 ? zo.y : zo.z), w1, d[0], d[1], d[2], (double)rr.x,
                       (double)rr.y, (double)rr.z, b0, b1, b2);
        }
      }
      __syncthreads();
      // pass j+1 on the own rows, from pass j's values
      float4 x[kHaloU];
#pragma unroll
      for (int u = 0; u < kHaloU; ++u) x[u] = zl[oslot[u * kBlock + tid]];
      const float4 zo2 = zl[g];
      VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
#pragma unroll
      for (int u = 0; u < kHaloU; ++u)
        blk_step(rrF.p0 + lane + u * G < rrF.p1, rrF.v[u], x[u].x, x[u].y, x[u].z, a0, a1, a2);
      a0 = group_sum<G>(a0);
      a1 = group_sum<G>(a1);
      a2 = group_sum<G>(a2);
      if (own) {
        const VT out = pass_out(lane == 0 ? zo2.x : (lane == 1 ? zo2.y : zo2.z), w2, rrF.d0, rrF.d1, rrF.d2, r0, r1,
                                r2, a0, a1, a2);
        if (zto)
          __hip_atomic_store(zto + 3 * rowF + lane, ((unsigned long long)rout << 32) | __float_as_uint(out),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
          zout[3 * rowF + lane] = out;
        if (last2) rz += pick3(lane, r0, r1, r2) * (double)out;
      }
    }
    return true;
  };

  // the first pass alone (an odd pass count in pair mode): on the own rows
  // from the plain input, published as tagged values of round rout
  auto single_tagged = [&](const SysDesc& S, const VT* zin, unsigned long long* zto, unsigned rout, VT w) {
    if constexpr (kPairs) {
      const int lane = tid & 3;
      const bool own = rowF < S.nb && lane < 3;
      const int ro = own ? rowF : 0, lo = own ? lane : 0;
      double r0, r1, r2;
      VT zo;
      const auto zg = gbl(zin);
      const auto rg = gbl(S.r);
      VT a0, a1, a2;
      row_spmv_res<G, VT>(S, mat_vals<VT>(S), rrF, lane,
                          [&](int c, VT& x0, VT& x1, VT& x2) {
                            x0 = zg[3 * c];
                            x1 = zg[3 * c + 1];
                            x2 = zg[3 * c + 2];
                          },
                          [&]() {
                            r0 = rg[3 * ro];
                            r1 = rg[3 * ro + 1];
                            r2 = rg[3 * ro + 2];
                            zo = zg[3 * ro + lo];
                          },
                          a0, a1, a2);
      if (own) {
        const VT out = pass_out(zo, w, rrF.d0, rrF.d1, rrF.d2, r0, r1, r2, a0, a1, a2);
        __hip_atomic_store(zto + 3 * rowF + lane, ((unsigned long long)rout << 32) | __float_as_uint(out),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  };

  // ---- Neumann steps C1..C(order-1) -----------------------------------
  // (in pair mode: pairs (1, 2), (3, 4), ..., then a single last pass when
  // order - 1 is odd)
  // pass-pair rounds: every tagged write round of the launch has its own
  // number (1, 2, ...), and round r writes buffer r & 1
  unsigned round = 0;
  auto neumann_steps = [&]() -> bool {
    if (kPairs && s_pairs && order >= 3) {
      // pairs (1, 2), ... or, for an odd pass count, pass 1 alone and then
      // pairs (2, 3), ...: the first step reads the plain z, the last pair
      // writes the plain z, the steps between exchange tagged values.  A
      // failed bounded wait of any thread sets s_pfail (cleared here).
      if (tid == 0) s_pfail = 0;
      __syncthreads();
      const int n = order - 1;
      int j = 1;
      const unsigned long long* zti = nullptr;
      unsigned rin = 0;
      const size_t zt_stride = 3 * (size_t)(resF ? sys[sF].nb : 0);
      if (n & 1) {
        round += 1;
        if (resF && s_st[sF] == ST_RUN) {
          const SysDesc& S = sys[sF];
          single_tagged(S, (const VT*)S.z[zc], S.zt + (round & 1) * zt_stride, round, step_w(1));
        }
        zti = resF ? sys[sF].zt + (round & 1) * zt_stride : nullptr;
        rin = round;
        j = 2;
      }
      for (; j < order; j += 2) {
        const bool last2 = j + 1 == order - 1;
        const bool first = j <= 2;
        unsigned long long* zto = nullptr;
        if (!last2) round += 1;
        double rz = 0.0;
        bool ok = true;
        if (resF) {
          const SysDesc& S = sys[sF];
          if (!last2) zto = S.zt + (round & 1) * zt_stride;
          if (s_st[sF] == ST_RUN)
            ok = pair_passes(S, (const VT*)S.z[zc], zti, rin, (VT*)S.z[zc ^ 1], zto, round, step_w(j), step_w(j + 1),
                             last2, first, rz);
          if (last2) {
            const double c = block_sum_t0(rz, s_red);
            if (s_st[sF] == ST_RUN && tid == 0) P.partF[(size_t)sF * P.nF + cF] = c;
          }
        }
        if (!ok) return false;
        zti = zto;
        rin = round;
      }
      zc ^= 1;
      return sync_at(1);
    }
    for (int j = 1; j < order;) {
      const bool last = (j == order - 1);
      double rz = 0.0;
      auto flushF = [&](int s) {
        if (last) {
          const double c = block_sum_t0(rz, s_red);
          if (s_st[s] == ST_RUN && tid == 0) P.partF[(size_t)s * P.nF + cF] = c;
          rz = 0.0;
        }
      };
      if (resF) {
        const SysDesc& S = sys[sF];
        if (s_st[sF] == ST_RUN) {
          // write-through stores before a point-to-point wait
          const VT* zi = (const VT*)S.z[zc];
          VT* zn = (VT*)S.z[zc ^ 1];
          const VT w = step_w(j);
          if (P.pass_sync && !last)
            tile_N_res<G, VT, true>(S, rowF, rrF, zi, zn, w, last, rz);
          else
            tile_N_res<G, VT>(S, rowF, rrF, zi, zn, w, last, rz);
        }
        flushF(sF);
      } else {
        for_tiles(sys, nsys, trF, 0,
                  [&](int s, int tl) {
                    if (s_st[s] != ST_RUN) return;
                    const SysDesc& S = sys[s];
                    tile_N<G, VT, SU>(S, tl, (const VT*)S.z[zc], (VT*)S.z[zc ^ 1], step_w(j), last, rz);
                  },
                  flushF);
      }
      zc ^= 1;
      if (P.pass_sync && !last) {
        if (!pass_wait()) return false;
      } else {
        if (!sync_at(1)) return false;
      }
      ++j;
    }
    return true;
  };

  // reduce rz (after the last preconditioner stage) into s_v0
  auto all_sys = [](int) { return true; };
  auto rngF = [&](int s, int& c0, int& c1) { c0 = sys[s].cF0; c1 = sys[s].cF1; };
  auto rngE = [&](int s, int& c0, int& c1) { c0 = sys[s].cE0; c1 = sys[s].cE1; };
  auto rngA = [&](int s, int& c0, int& c1) { c0 = sys[s].cA0; c1 = sys[s].cA1; };
  auto reduce_rz = [&]() {
    if (order > 1)
      reduce_chunks_all(P.partF, nsys, P.nF, P.rt_total, rngF, all_sys, s_v0);
    else
      reduce_chunks_all(partE_rz, nsys, P.nF, P.et_total, rngE, all_sys, s_v0);
    __syncthreads();
  };

  // Reduce the low-rank partials into s_t (and |t|^2 into s_tt) for the
  // systems selected by `want`; with with_pA also the stage-A partials
  // <p, D p> into s_v1 (every system).  Entry q of the concatenated t vectors
  // is summed over its system's low-rank chunks: when all ranks fit in one
  // workgroup (sum k <= kBlock) by ng groups of threads (every ng-th chunk,
  // groups added in order), else one thread per entry.  The stage-A loads are
  // issued before the first wait, so all latencies overlap.
  // tsel: destination s_t + tsel * t_total (0: t; 1, 2: the pipelined loop's copies)
  auto reduce_t = [&](auto want, bool with_pA, int tsel) {
    double* tdst = s_t + tsel * P.t_total;
    ChunkPre preA;
    if (with_pA) chunks_prefetch(P.partA, nsys, P.nA, rngA, all_sys, preA);
    const int T = P.t_total;
    const int ng = T > 0 && T <= kBlock ? kBlock / T : 1;
    auto sum_entry = [&](int q, int g0, int gs) {
      int s = 0, hi = nsys - 1;  // last system with t_off <= q (binary search)
      while (s < hi) {
        const int mid = (s + hi + 1) >> 1;
        if (sys[mid].t_off <= q) s = mid;
        else hi = mid - 1;
      }
      const SysDesc& S = sys[s];
      double v = 0.0;
      if (want(s) && S.k > 0) {
        const int l = q - S.t_off;
#pragma unroll 8
        for (int c = S.cL0 + g0; c <= S.cL1; c += gs) {
          const double x = P.partL[S.lrp_off + (size_t)(c - S.cL0) * S.k + l];
          v += chunk_empty(c, P.nL, P.lt_total) ? 0.0 : x;
        }
      }
      return v;
    };
    if (T > 0 && T <= kBlock) {
      const int q = tid % T, grp = tid / T;
      if (grp < ng) s_tred[grp * T + q] = sum_entry(q, grp, ng);
    } else {
      for (int q = tid; q < T; q += kBlock) tdst[q] = sum_entry(q, 0, 1);
    }
    if (with_pA) chunks_finish(P.partA, P.nA, P.rt_total, preA, s_v1);
    __syncthreads();
    if (T > 0 && T <= kBlock) {
      if (tid < T) {
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) acc += s_tred[g * T + tid];
        tdst[tid] = acc;
      }
      __syncthreads();
    }
    sum_tt_all(sys, nsys, tdst, want, s_tt);
    __syncthreads();
  };

  // low-rank stage for LR workgroups; mode 0: src = z + beta p_old, 1: src = x
  auto lr_stage = [&](int mode) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for_tiles(sys, nsys, trL, 2,
              [&](int s, int tl) {
                const SysDesc& S = sys[s];
                if (s_st[s] != ST_RUN) return;
                if (mode == 0) {
                  const VT* z = (const VT*)S.z[zc];
                  const VT* po = (const VT*)S.p[pc];
                  const VT be = (VT)s_beta[s];
                  tile_Wt(S, tl, w_vals<VT>(S),
                          [&](int dof) { return (double)combine(z[dof], po[dof], be); }, acc);
                } else {
                  if (s_chk[s] == CHK_NONE) return;
                  const double* x = S.x;
                  tile_Wt(S, tl, S.W64, [&](int dof) { return x[dof]; }, acc);
                }
              },
              [&](int s) {
                const SysDesc& S = sys[s];
                const bool active = s_st[s] == ST_RUN && (mode == 0 || s_chk[s] != CHK_NONE);
                double* dst = P.partL + S.lrp_off + (size_t)(cL - S.cL0) * S.k;
                if (active && S.k > 0) {
                  flush_Wt(acc, S.k, dst, s_lr);
                } else {
                  for (int c = 0; c < 4; ++c) acc[c] = 0.0;
                }
              });
  };

  // ================================================================ init
  node_stage(0, 0);
  if (!sync_at(0)) goto out;
  reduce_chunks_all(partE_rr, nsys, P.nF, P.et_total, rngE, all_sys, s_v1);
  __syncthreads();
  for (int s = tid; s < nsys; s += kBlock) {
    const double nb = sqrt(s_v1[s]);
    s_normb[s] = nb;
    if (!(nb > 0.0)) {
      s_st[s] = ST_DONE;
      s_rel[s] = 0.0;
    }
  }
  __syncthreads();
  if constexpr (PIPE) {
    // ============================================================ pipelined loop
    // One grid barrier per iteration.  The passes end with a neighbour-tile
    // wait, so D z (the D part of S z) is formed by each row workgroup for
    // its own rows without a grid barrier; with q_old = D p_old kept in FP64,
    //   q = D p = D z + beta q_old,
    //   p.D p = z.D z + 2 beta z.q_old + beta^2 p_old.D p_old,
    //   t = W^T p = W^T z + beta W^T p_old,
    // so the dots and W^T z of one iteration are reduced at one barrier
    // together with r.z (the last pass) and |r|^2 (the previous update).
    // The residual check is on the r the preconditioner just used, so
    // convergence is seen after one more preconditioner application than in
    // the standard loop, with the same iterate.  A true-residual check
    // (convergence or replacement) that keeps a system running recomputes
    // its z from the new r before its next update.
    __shared__ double s_dz[kPipeRows * 3];  // D z of the resident rows (row layout -> node layout)
    __shared__ double s_zdz[kMaxSys], s_zq[kMaxSys], s_rr[kMaxSys], s_pdp[kMaxSys];
    __shared__ int s_first[kMaxSys], s_redo[kMaxSys], s_skip[kMaxSys];
    double* s_tz = s_t + P.t_total;      // W^T z
    double* partA2 = P.partA + (size_t)nsys * P.nA;
    const int lrow = tid / G;
    for (int s = tid; s < nsys; s += kBlock) {
      s_first[s] = 1;
      s_redo[s] = 0;
      s_skip[s] = 0;
      s_pdp[s] = 0.0;
      s_rz[s] = 0.0;
    }
    __syncthreads();
    auto pipe_wait = [&]() -> bool { return P.pass_sync ? pass_wait() : sync_at(1); };
    // W^T v over the coupled nodes of this row workgroup's node tile (v = the
    // last pass's z, or x for a true residual), written to partR for the
    // reduction at the next barrier.  Each wave sums its consecutive nodes in
    // order and the waves' sums are added in order, so the result does not
    // depend on scheduling.  (The low-rank workgroups idle in this loop.)
    __shared__ double s_vc[kPipeRows * 3];
    __shared__ int s_an[kPipeRows];
    if (tid < kPipeRows) s_an[tid] = (resE && tid < P.node_tile) ? anE : -1;
    __syncthreads();
    auto rows_Wt = [&](bool use_x) {
      if (!resF) return;
      const SysDesc& S = sys[sF];
      const int k = S.k;
      if (k == 0 || s_st[sF] != ST_RUN || (use_x && s_chk[sF] == CHK_NONE)) return;
      if (tid < kPipeRows) {
        const int i = iE;
        const bool cpl = s_an[tid] >= 0 && i < S.nb;
        for (int d = 0; d < 3; ++d)
          s_vc[3 * tid + d] = !cpl ? 0.0 : (use_x ? S.x[3 * i + d] : (double)((const VT*)S.z[zc])[3 * i + d]);
      }
      __syncthreads();
      constexpr int NW = kBlock / 64;
      constexpr int per = kPipeRows / NW;
      const int wv = tid >> 6, lane = tid & 63;
      double* dst = P.partR + (size_t)S.t_off * P.nF + (size_t)cF * k;
      for (int l0 = 0; l0 < k; l0 += 64) {
        const int l = l0 + lane;
        double acc = 0.0;
        if (l < k) {
          for (int u = 0; u < per; ++u) {
            const int node = wv * per + u;
            const int an = s_an[node];
            if (an < 0) continue;
            double w0, w1, w2;
            if (use_x) {
              w0 = S.W64[(size_t)an * k + l];
              w1 = S.W64[(size_t)(an + 1) * k + l];
              w2 = S.W64[(size_t)(an + 2) * k + l];
            } else if (s_wc) {
              const VT* wr = s_wc + (size_t)s_slot[node] * 3 * k;
              w0 = (double)wr[l];
              w1 = (double)wr[k + l];
              w2 = (double)wr[2 * k + l];
            } else {
              const VT* W = w_vals<VT>(S);
              w0 = (double)W[(size_t)an * k + l];
              w1 = (double)W[(size_t)(an + 1) * k + l];
              w2 = (double)W[(size_t)(an + 2) * k + l];
            }
            acc += w0 * s_vc[3 * node] + w1 * s_vc[3 * node + 1] + w2 * s_vc[3 * node + 2];
          }
        }
        s_lr[wv][lane] = acc;
        __syncthreads();
        if (tid < 64 && l < k) {
          double t = 0.0;
          for (int w = 0; w < NW; ++w) t += s_lr[w][tid];
          dst[l] = t;
        }
        __syncthreads();
      }
    };
    // t entries of the systems selected by `want` from the row workgroups'
    // partials: entry q of system s sums its chunks cF0..cF1 in order (groups
    // of threads take every ng-th chunk; the groups are added in order)
    auto reduce_rows_t = [&](auto want, int tsel) {
      double* tdst = s_t + tsel * P.t_total;
      const int T = P.t_total;
      const int ng = T > 0 && T <= kBlock ? kBlock / T : 1;
      auto sum_entry = [&](int q, int g0, int gs) {
        int s = 0, hi = nsys - 1;
        while (s < hi) {
          const int mid = (s + hi + 1) >> 1;
          if (sys[mid].t_off <= q) s = mid;
          else hi = mid - 1;
        }
        const SysDesc& S = sys[s];
        double v = 0.0;
        if (want(s) && S.k > 0) {
          const int l = q - S.t_off;
          const double* src = P.partR + (size_t)S.t_off * P.nF;
#pragma unroll 8
          for (int c = S.cF0 + g0; c <= S.cF1; c += gs) v += src[(size_t)c * S.k + l];
        }
        return v;
      };
      if (T > 0 && T <= kBlock) {
        const int q = tid % T, grp = tid / T;
        if (grp < ng) s_tred[grp * T + q] = sum_entry(q, grp, ng);
      } else {
        for (int q = tid; q < T; q += kBlock) tdst[q] = sum_entry(q, 0, 1);
      }
      __syncthreads();
      if (T > 0 && T <= kBlock) {
        if (tid < T) {
          double acc = 0.0;
          for (int g = 0; g < ng; ++g) acc += s_tred[g * T + tid];
          tdst[tid] = acc;
        }
      }
      __syncthreads();
    };
    for (;;) {
      if (tid == 0) {
        int any = 0;
        for (int s = 0; s < nsys; ++s) any |= (s_st[s] == ST_RUN);
        s_any = any;
      }
      __syncthreads();
      if (!s_any) break;

      // ---------------- passes 1 .. order-1, each after the neighbours' previous one
      if (!pipe_wait()) goto out;
      for (int j = 1; j < order; ++j) {
        const bool last = j == order - 1;
        double rz = 0.0;
        if (resF) {
          const SysDesc& S = sys[sF];
          if (s_st[sF] == ST_RUN) {
            const VT* zi = (const VT*)S.z[zc];
            VT* zn = (VT*)S.z[zc ^ 1];
            if (P.pass_sync)
              tile_N_res<G, VT, true>(S, rowF, rrF, zi, zn, step_w(j), last, rz);
            else
              tile_N_res<G, VT>(S, rowF, rrF, zi, zn, step_w(j), last, rz);
          }
          if (last) {
            const double c = block_sum_t0(rz, s_red);
            if (s_st[sF] == ST_RUN && tid == 0) P.partF[(size_t)sF * P.nF + cF] = c;
          }
        }
        zc ^= 1;
        if (!pipe_wait()) goto out;
      }

      // ---------------- D z and W^T z on the resident rows
      if (resF) {
        const SysDesc& S = sys[sF];
        double zdz = 0.0, zq = 0.0;
        if (s_st[sF] == ST_RUN) {
          const auto zg = gbl((const VT*)S.z[zc]);
          const auto qg = gbl((const double*)S.q);
          const int lane = tid % G;
          const bool own = rowF < S.nb && lane < 3;
          const int io = own ? 3 * rowF + lane : 0;
          VT zr;
          double qo;
          VT a0, a1, a2;
          row_spmv_res<G, VT>(S, mat_vals<VT>(S), rrF, lane,
                              [&](int c, VT& x0, VT& x1, VT& x2) {
                                x0 = zg[3 * c];
                                x1 = zg[3 * c + 1];
                                x2 = zg[3 * c + 2];
                              },
                              [&]() {
                                zr = zg[io];
                                qo = qg[io];
                              },
                              a0, a1, a2);
          if (own) {
            const VT dz = pick3(lane, a0, a1, a2);
            const double zo = (double)zr;
            zdz = zo * (double)dz;
            if (!s_first[sF]) zq = zo * qo;
            s_dz[3 * lrow + lane] = (double)dz;
          }
        }
        const double d0 = block_sum_t0(zdz, s_red);
        const double d1 = block_sum_t0(zq, s_red);
        if (s_st[sF] == ST_RUN && tid == 0) {
          P.partA[(size_t)sF * P.nA + cA] = d0;
          partA2[(size_t)sF * P.nA + cA] = d1;
        }
        rows_Wt(false);
      }
      if (!sync_at(2)) goto out;

      // ---------------- reductions (every workgroup, fixed order)
      // r.z, z.D z, z.q_old and |r|^2 (the aligned split chunks rows and
      // nodes alike: one chunk range per system), then W^T z
      reduce4_chunks(P.partF, P.nF, P.partA, partA2, P.nA, partE_rr, nsys, P.rt_total, rngF, s_v0, s_zdz, s_zq,
                     s_rr);
      reduce_rows_t([&](int s) { return s_st[s] == ST_RUN; }, 1);
      if (tid == 0) {
        int any = 0;
        for (int s = 0; s < nsys; ++s) {
          s_chk[s] = CHK_NONE;
          s_redo[s] = 0;
          if (s_st[s] != ST_RUN || s_skip[s]) continue;
          const SysDesc& S = sys[s];
          const double rel = sqrt(s_rr[s]) / s_normb[s];
          s_rel[s] = rel;
          int ck = CHK_NONE;
          if (rel <= S.tol) ck = CHK_CONVERGE;
          else if (P.replace_every > 0 && s_it[s] > 0 && (s_it[s] % P.replace_every) == 0) ck = CHK_REPLACE;
          s_chk[s] = ck;
          any |= (ck != CHK_NONE);
        }
        s_any = any;
      }
      __syncthreads();

      // ---------------- true residual of the checked systems (T1, T2)
      if (s_any) {
        if (resF) {
          const SysDesc& S = sys[sF];
          if (s_st[sF] == ST_RUN && s_chk[sF] != CHK_NONE) tile_Dx64<G, SU>(S, trF.t0 - S.rt_begin, S.x, S.q64);
          rows_Wt(true);
        }
        if (!sync_at(4)) goto out;
        reduce_rows_t([&](int s) { return s_st[s] == ST_RUN && s_chk[s] != CHK_NONE; }, 2);
        node_stage(2, 2);
        if (!sync_at(5)) goto out;
        reduce_chunks_all(partE_rr, nsys, P.nF, P.et_total, rngE,
                          [&](int s) { return s_st[s] == ST_RUN && s_chk[s] != CHK_NONE; }, s_v1);
        __syncthreads();
        if (tid == 0) {
          for (int s = 0; s < nsys; ++s) {
            if (s_st[s] != ST_RUN || s_chk[s] == CHK_NONE) continue;
            const SysDesc& S = sys[s];
            const double rel = sqrt(s_v1[s]) / s_normb[s];
            s_nchk[s] += 1;
            s_rel[s] = rel;
            if (rel <= S.tol) {
              s_st[s] = ST_DONE;
            } else {
              // r replaced by the true residual and z0 recomputed from it:
              // the passes run again before the next update (after a failed
              // convergence check the iteration restarts with beta = 0)
              if (s_chk[s] == CHK_CONVERGE) {
                s_first[s] = 1;
                s_nrs[s] += 1;
              }
              s_redo[s] = 1;
              s_skip[s] = 1;
            }
          }
        }
        __syncthreads();
      }
      if (tid == 0) {
        for (int s = 0; s < nsys; ++s)
          if (s_st[s] == ST_RUN && !s_redo[s] && s_it[s] >= sys[s].max_iter) s_st[s] = kNotConverged;
      }
      __syncthreads();

      // ---------------- beta, t = W^T p, alpha
      for (int s = tid; s < nsys; s += kBlock)
        if (s_st[s] == ST_RUN && !s_redo[s]) s_beta[s] = s_first[s] ? 0.0 : s_v0[s] / s_rz[s];
      __syncthreads();
      for (int s = 0; s < nsys; ++s) {
        if (s_st[s] != ST_RUN || s_redo[s]) continue;
        const int off = sys[s].t_off, k = sys[s].k;
        const double be = s_beta[s];
        const bool first = s_first[s];
        for (int l = tid; l < k; l += kBlock) s_t[off + l] = first ? s_tz[off + l] : s_tz[off + l] + be * s_t[off + l];
      }
      __syncthreads();
      sum_tt_all(sys, nsys, s_t, [&](int s) { return s_st[s] == ST_RUN && !s_redo[s]; }, s_tt);
      __syncthreads();
      for (int s = tid; s < nsys; s += kBlock) {
        if (s_st[s] != ST_RUN || s_redo[s]) continue;
        const double be = s_beta[s], rz = s_v0[s];
        const double pdp = s_first[s] ? s_zdz[s] : s_zdz[s] + be * (2.0 * s_zq[s] + be * s_pdp[s]);
        const double pSp = pdp - s_tt[s];
        if (!(isfinite(pSp) && pSp > 0.0 && isfinite(rz) && rz > 0.0)) {
          s_st[s] = kBreakdown;
        } else {
          s_alpha[s] = rz / pSp;
          s_rz[s] = rz;
          s_pdp[s] = pdp;
        }
      }
      __syncthreads();

      // ---------------- update on the resident node tile; z0 = w Dinv r
      {
        double rr = 0.0;
        for_tiles(sys, nsys, trE, 1,
                  [&](int s, int tl) {
                    const SysDesc& S = sys[s];
                    if (s_st[s] != ST_RUN || s_redo[s]) return;
                    const int i = tile_row(S, tl, P.node_tile, tid);
                    const bool valid = tid < P.node_tile && i < S.nb;
                    double g[3] = {0.0, 0.0, 0.0};
                    if (S.k > 0)
                      lowrank_apply(S.k, anE, w_vals<VT>(S), s_t + S.t_off, g, s_list, s_slot, s_g, &s_cnt, s_wc);
                    if (!valid) return;
                    const int i3 = 3 * i;
                    VT* z = (VT*)S.z[zc];
                    VT* pv = (VT*)S.p[0];
                    double* qv = (double*)S.q;
                    const double al = s_alpha[s], bed = s_beta[s];
                    const VT be = (VT)bed;
                    const bool first = s_first[s];
                    double rv[3];
                    for (int d = 0; d < 3; ++d) {
                      const VT zd = z[i3 + d];
                      const VT pn = first ? zd : combine(zd, pv[i3 + d], be);
                      const double dz = s_dz[3 * tid + d];
                      const double qn = first ? dz : fma(bed, qv[i3 + d], dz);
                      pv[i3 + d] = pn;
                      qv[i3 + d] = qn;
                      rv[d] = S.r[i3 + d] - al * (qn - g[d]);
                      S.x[i3 + d] += al * (double)pn;
                    }
                    for (int d = 0; d < 3; ++d) S.r[i3 + d] = rv[d];
                    rr += rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                    const VT* ig = inv_vals<VT>(S) + 9 * i;
                    const VT r0 = (VT)rv[0], r1 = (VT)rv[1], r2 = (VT)rv[2];
                    const VT z0 = omega * (ig[0] * r0 + ig[1] * r1 + ig[2] * r2);
                    const VT z1 = omega * (ig[3] * r0 + ig[4] * r1 + ig[5] * r2);
                    const VT z2 = omega * (ig[6] * r0 + ig[7] * r1 + ig[8] * r2);
                    if (P.pass_sync) {
                      zst<true>(z + i3, z0);
                      zst<true>(z + i3 + 1, z1);
                      zst<true>(z + i3 + 2, z2);
                    } else {
                      z[i3] = z0;
                      z[i3 + 1] = z1;
                      z[i3 + 2] = z2;
                    }
                  },
                  [&](int s) {
                    const double a = block_sum_t0(rr, s_red);
                    if (s_st[s] == ST_RUN && !s_redo[s] && tid == 0) partE_rr[(size_t)s * P.nF + cF] = a;
                    rr = 0.0;
                  });
      }
      if (tid == 0) {
        for (int s = 0; s < nsys; ++s) {
          if (s_st[s] != ST_RUN || s_redo[s]) continue;
          s_it[s] += 1;
          s_first[s] = 0;
          s_skip[s] = 0;
        }
      }
      __syncthreads();
    }
  } else {
  if (!neumann_steps()) goto out;
  reduce_rz();
  for (int s = tid; s < nsys; s += kBlock) {
    s_rz[s] = s_v0[s];
    s_beta[s] = 0.0;
  }
  __syncthreads();

  // ================================================================ loop
  for (;;) {
    if (tid == 0) {
      int any = 0;
      for (int s = 0; s < nsys; ++s) any |= (s_st[s] == ST_RUN);
      s_any = any;
    }
    __syncthreads();
    if (!s_any) break;

    // ---------------- stage A
    if (is_lr) {
      lr_stage(0);
    } else if (resA) {
      double dot = 0.0;
      const SysDesc& S = sys[sF];
      if (s_st[sF] == ST_RUN)
        tile_A_res<G, VT>(S, rowF, rrF, (const VT*)S.z[zc], (const VT*)S.p[pc], (VT*)S.p[pc ^ 1], (VT*)S.q,
                          (VT)s_beta[sF], dot);
      const double d = block_sum_t0(dot, s_red);
      if (s_st[sF] == ST_RUN && tid == 0) P.partA[(size_t)sF * P.nA + cA] = d;
    } else {
      double dot = 0.0;
      for_tiles(sys, nsys, trA, 0,
                [&](int s, int tl) {
                  if (s_st[s] != ST_RUN) return;
                  const SysDesc& S = sys[s];
                  tile_A<G, VT, SU>(S, tl, (const VT*)S.z[zc], (const VT*)S.p[pc], (VT*)S.p[pc ^ 1],
                                (VT*)S.q, (VT)s_beta[s], dot);
                },
                [&](int s) {
                  const double d = block_sum_t0(dot, s_red);
                  if (s_st[s] == ST_RUN && tid == 0) P.partA[(size_t)s * P.nA + cA] = d;
                  dot = 0.0;
                });
    }
    pc ^= 1;
    if (!sync_at(2)) goto out;
    prefetch_B();
    reduce_t([&](int s) { return s_st[s] == ST_RUN; }, true, 0);
    for (int s = tid; s < nsys; s += kBlock) {
      if (s_st[s] != ST_RUN) continue;
      const double pAp = s_v1[s] - s_tt[s];
      const double rz = s_rz[s];
      if (!(isfinite(pAp) && pAp > 0.0 && isfinite(rz) && rz > 0.0)) {
        s_st[s] = kBreakdown;
      } else {
        s_alpha[s] = rz / pAp;
      }
    }
    __syncthreads();

    // ---------------- stage B
    if (P.prof && tid == 0) P.prof[(size_t)b * 32 + 12] += rt_ticks() - t_last;  // B: reductions part
    node_stage(1, 0);
    pre_ok = false;
    if (!sync_at(3)) goto out;
    reduce_chunks_all(partE_rr, nsys, P.nF, P.et_total, rngE, all_sys, s_v1);
    __syncthreads();
    if (tid == 0) {
      int any = 0;
      for (int s = 0; s < nsys; ++s) {
        s_chk[s] = CHK_NONE;
        if (s_st[s] != ST_RUN) continue;
        const SysDesc& S = sys[s];
        s_it[s] += 1;
        const double rel = sqrt(s_v1[s]) / s_normb[s];
        s_rel[s] = rel;
        int ck = CHK_NONE;
        if (rel <= S.tol) ck = CHK_CONVERGE;
        else if (P.mode_true_every) ck = CHK_NORM;
        else if (P.replace_every > 0 && (s_it[s] % P.replace_every) == 0) ck = CHK_REPLACE;
        s_chk[s] = ck;
        any |= (ck != CHK_NONE);
      }
      s_any = any;
    }
    __syncthreads();

    // ---------------- true residual (T1, T2) when any system asks for it
    if (s_any) {
      if (is_lr) {
        lr_stage(1);
      } else {
        for_tiles(sys, nsys, trA, 0,
                  [&](int s, int tl) {
                    if (s_st[s] != ST_RUN || s_chk[s] == CHK_NONE) return;
                    const SysDesc& S = sys[s];
                    tile_Dx64<G, SU>(S, tl, S.x, S.q64);
                  },
                  [&](int) {});
      }
      if (!sync_at(4)) goto out;
      reduce_t([&](int s) { return s_st[s] == ST_RUN && s_chk[s] != CHK_NONE; }, false, 0);
      node_stage(2, 0);
      if (!sync_at(5)) goto out;
      reduce_chunks_all(partE_rr, nsys, P.nF, P.et_total, rngE,
                        [&](int s) { return s_st[s] == ST_RUN && s_chk[s] != CHK_NONE; }, s_v1);
      __syncthreads();
      if (tid == 0) {
        for (int s = 0; s < nsys; ++s) {
          if (s_st[s] != ST_RUN || s_chk[s] == CHK_NONE) continue;
          const SysDesc& S = sys[s];
          const double rel = sqrt(s_v1[s]) / s_normb[s];
          s_nchk[s] += 1;
          s_rel[s] = rel;
          if (rel <= S.tol) {
            s_st[s] = ST_DONE;
          } else if (s_chk[s] == CHK_CONVERGE) {
            s_rs[s] = 1;
            s_nrs[s] += 1;
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      for (int s = 0; s < nsys; ++s)
        if (s_st[s] == ST_RUN && s_it[s] >= sys[s].max_iter) s_st[s] = kNotConverged;
    }
    __syncthreads();

    // ---------------- preconditioner stages
    if (!neumann_steps()) goto out;
    reduce_rz();
    for (int s = tid; s < nsys; s += kBlock) {
      if (s_st[s] != ST_RUN) continue;
      const double rzn = s_v0[s];
      if (s_rs[s]) {
        s_beta[s] = 0.0;
        s_rs[s] = 0;
      } else {
        s_beta[s] = rzn / s_rz[s];
      }
      s_rz[s] = rzn;
    }
    __syncthreads();
  }
  }  // standard loop

  }  // started
out:
  __syncthreads();
  if (b == 0) {
    for (int s = tid; s < nsys; s += kBlock) {
      SysOut o;
      o.iterations = s_it[s];
      o.status = s_st[s] == ST_DONE ? 0 : (s_st[s] == ST_RUN ? HSC_ERR_DEVICE_TIMEOUT_CODE : s_st[s]);
      o.checks = s_nchk[s];
      o.restarts = s_nrs[s];
      o.pairs = kPairs ? s_pairs : 0;
      o.pad_ = 0;
      o.rel = s_rel[s];
      o.normb = s_normb[s];
      P.out[s] = o;
    }
  }
}

// ---------------------------------------------------------------- operator kernels
// (plain launches: one tile per workgroup; used by the parity entry points)
template <int G>
__global__ __launch_bounds__(kBlock) void op_Dx64(SysDesc S, const double* x, double* y) {
  tile_Dx64<G, StreamU<G, false>::value>(S, blockIdx.x, x, y);
}

// One power-iteration step for lambda_max(Dinv D), fused: z = Dinv D (x s)
// with s = 1 / |x| (|x|^2 the fixed-order sum of the previous step's np_in
// partials, which every workgroup forms the same way), and the tile's
// partial |z|^2.
template <int G>
__global__ __launch_bounds__(kBlock) void op_power_step(SysDesc S, const double* x, const double* part_in,
                                                        int np_in, double* z, double* part) {
  __shared__ double red[kBlock / 64];
  __shared__ double s_sc;
  {
    double v = 0.0;
    for (int i = threadIdx.x; i < np_in; i += kBlock) v += part_in[i];
    const double t = block_sum_t0(v, red);
    if (threadIdx.x == 0) s_sc = 1.0 / sqrt(t);
    __syncthreads();
  }
  const double sc = s_sc;
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  const int row = blockIdx.x * (kBlock / G) + g;
  double a0, a1, a2;
  row_spmv<G, double>(S, S.val64, row, lane,
                      [&](int c, double& x0, double& x1, double& x2) {
                        x0 = x[3 * c] * sc;
                        x1 = x[3 * c + 1] * sc;
                        x2 = x[3 * c + 2] * sc;
                      },
                      a0, a1, a2);
  double zz = 0.0;
  if (row < S.nb && lane < 3) {
    const double* inv = S.invD64 + 9 * row + 3 * lane;
    const double zv = inv[0] * a0 + inv[1] * a1 + inv[2] * a2;
    z[3 * row + lane] = zv;
    zz = zv * zv;
  }
  const double t = block_sum_t0(zz, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <int G, typename VT>
__global__ __launch_bounds__(kBlock) void op_neumann_step(SysDesc S, const VT* zin, VT* zout, double omega) {
  double rz = 0.0;
  tile_N<G, VT, StreamU<G, false>::value>(S, blockIdx.x, zin, zout, (VT)omega, false, rz);
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void op_jacobi(SysDesc S, const double* r, VT* z, double omega) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nb) return;
  const VT* inv = inv_vals<VT>(S) + 9 * i;
  const VT r0 = (VT)r[3 * i], r1 = (VT)r[3 * i + 1], r2 = (VT)r[3 * i + 2];
  for (int d = 0; d < 3; ++d)
    z[3 * i + d] = (VT)omega * (inv[3 * d] * r0 + inv[3 * d + 1] * r1 + inv[3 * d + 2] * r2);
}

// t = W^T x (single workgroup, fixed order) then y[act] -= W t
__global__ __launch_bounds__(kBlock) void op_lowrank64(SysDesc S, const double* x, double* y, double* t) {
  __shared__ double s_tv[kMaxRankLds];
  __shared__ double lr[kBlock / 64][64];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int ntile = (S.a + kLrTile - 1) / kLrTile;
  for (int tl = 0; tl < ntile; ++tl) tile_Wt(S, tl, S.W64, [&](int dof) { return x[dof]; }, acc);
  flush_Wt(acc, S.k, s_tv, lr);
  __syncthreads();
  for (int l = threadIdx.x; l < S.k; l += kBlock) t[l] = s_tv[l];
  for (int j = threadIdx.x; j < S.a; j += kBlock) {
    double g = 0.0;
    for (int l = 0; l < S.k; ++l) g += S.W64[(size_t)j * S.k + l] * s_tv[l];
    y[S.act_dof[j]] -= g;
  }
}

// ---------------------------------------------------------------- chunk-strip operator kernels
// The standalone D x and preconditioner pass stream the matrix once per
// launch: their rate is set by how the loads map onto memory.  Here a wave
// owns the rows whose first block lies in its `cpw` chunks of 64 blocks and
// walks their blocks chunk by chunk, lane j on block 64k + j, so the column
// index and each of the 9 value components of a chunk is one contiguous,
// aligned 64-element load (the component-major chunk layout), no lane idles
// on a short row, and a lane's x gather is one 12-byte (FP32) or 16 + 8-byte
// (FP64) load.  Row sums come out of a head-segmented inclusive scan across
// the wave (DPP row shifts by 1, 2, 4, 8, then the row-15 and row-31
// broadcasts), and the running sum of a row that continues past the chunk's
// end is carried into lane 0 of the next chunk, so rows of any length are
// summed in a fixed order.  Two per-chunk tables built at setup give the
// structure: the chunk's head bits (bit j: block 64k + j is the first block
// of a row) and the number of rows that start before the chunk.  Lanes
// outside the wave's blocks load a clamped in-range address (a cache line a
// lane in range loads anyway) and contribute zero.

// v of another lane through a DPP move (CTRL, row mask RM); 0 where the
// source lane is out of range or the row is masked off
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xF, false));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL, int RM, typename T>
__device__ __forceinline__ void seg_step(bool take, T& a0, T& a1, T& a2) {
  const T t0 = dpp_mov<CTRL, RM>(a0), t1 = dpp_mov<CTRL, RM>(a1), t2 = dpp_mov<CTRL, RM>(a2);
  a0 = take ? a0 + t0 : a0;
  a1 = take ? a1 + t1 : a1;
  a2 = take ? a2 + t2 : a2;
}

// Inclusive scan of (a0, a1, a2) over the lanes of each segment; s0 is the
// first lane of this lane's segment.  Every lane's result is a fixed tree of
// its segment's values.
template <typename T>
__device__ __forceinline__ void seg_scan3(int lane, int s0, T& a0, T& a1, T& a2) {
  const int rl = lane & 15;
  seg_step<0x111, 0xF>(rl >= 1 && lane - 1 >= s0, a0, a1, a2);  // row_shr:1
  seg_step<0x112, 0xF>(rl >= 2 && lane - 2 >= s0, a0, a1, a2);  // row_shr:2
  seg_step<0x114, 0xF>(rl >= 4 && lane - 4 >= s0, a0, a1, a2);  // row_shr:4
  seg_step<0x118, 0xF>(rl >= 8 && lane - 8 >= s0, a0, a1, a2);  // row_shr:8
  // row_bcast:15 into rows 1 and 3 (from lanes 15 and 47)
  seg_step<0x142, 0xA>(((lane >> 4) & 1) != 0 && s0 <= (lane | 15) - 16, a0, a1, a2);
  // row_bcast:31 into rows 2 and 3 (from lane 31)
  seg_step<0x143, 0xC>(lane >= 32 && s0 <= 31, a0, a1, a2);
}

template <typename T>
__device__ __forceinline__ T lane63(T v) {
  if constexpr (sizeof(T) == 4) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  }
}

// x[3c .. 3c+2] as one load of 12 bytes (FP32) or 16 + 8 bytes (FP64)
template <typename VT>
struct __attribute__((aligned(sizeof(VT)))) Vec3 {
  VT v0, v1, v2;
};
template <typename VT>
__device__ __forceinline__ Vec3<VT> gather3(const VT* __restrict__ x, int c) {
  Vec3<VT> t;
  __builtin_memcpy(&t, x + 3 * (size_t)c, sizeof(t));
  return t;
}

// MODE 0: y = D x (FP64).  MODE 1: y = x + w Dinv (r - D x) (VT values and
// vectors, FP64 r), the arithmetic of tile_N per row.
//
// The lane that ends a row writes the row's sum into a 128-row ring in LDS;
// each time 64 rows are finished the wave runs their epilogue with one lane
// per row (consecutive rows: r, z and the inverse blocks are read, and y
// written, as contiguous streams instead of by the few row-ending lanes of
// each chunk).  Workgroups take contiguous runs of chunks per XCD (workgroup
// b runs on XCD b % 8), so the x lines that neighbouring rows share are
// fetched into one L2; the value and index streams, read once per launch,
// are loaded non-temporally so they do not evict x.  At most 72 VGPRs (7
// waves per SIMD): the FP64 pass at 80 ran 6% slower.
constexpr int kStripBlock = 256;  // one wave per SIMD per workgroup: occupancy in steps of one wave

template <int MODE, typename VT>
__global__ __launch_bounds__(kStripBlock) __attribute__((amdgpu_waves_per_eu(7))) void op_strip(
    SysDesc S, StripDesc T, const VT* __restrict__ x, VT* __restrict__ y, VT w) {
  __shared__ VT s_sum[kStripBlock / 64][128][3];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int nwg = (int)gridDim.x, xc = (int)blockIdx.x & 7, q = nwg >> 3, rem = nwg & 7;
  const int b = xc * q + min(xc, rem) + ((int)blockIdx.x >> 3);
  const int wv = b * (kStripBlock / 64) + wl;
  if (wv >= T.nwave) return;
  const int c0 = (int)((long long)wv * T.nch / T.nwave), c1 = (int)((long long)(wv + 1) * T.nch / T.nwave);
  const int rb = T.hrow[c0], re = T.hrow[c1];
  if (rb >= re) return;  // a row longer than the wave's chunks: its owner sums it
  const int hb = S.rowptr[rb], he = S.rowptr[re];
  const VT* val = mat_vals<VT>(S);
  const auto inv = gbl(inv_vals<VT>(S));
  const auto rg = gbl(S.r);
  int rdone = rb;  // rows [rb, rdone) have had their epilogue
  auto epilogue = [&](int cnt) {
    if (lane < cnt) {
      const int row = rdone + lane;
      const int sl = (row - rb) & 127;
      const VT a0 = s_sum[wl][sl][0], a1 = s_sum[wl][sl][1], a2 = s_sum[wl][sl][2];
      if constexpr (MODE == 0) {
        y[3 * row] = a0;
        y[3 * row + 1] = a1;
        y[3 * row + 2] = a2;
      } else {
        const double r0 = rg[3 * row], r1 = rg[3 * row + 1], r2 = rg[3 * row + 2];
        const Vec3<VT> zr = gather3(x, row);
        VT d[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) d[e] = inv[9 * (size_t)row + e];
        y[3 * row] = pass_out(zr.v0, w, d[0], d[1], d[2], r0, r1, r2, a0, a1, a2);
        y[3 * row + 1] = pass_out(zr.v1, w, d[3], d[4], d[5], r0, r1, r2, a0, a1, a2);
        y[3 * row + 2] = pass_out(zr.v2, w, d[6], d[7], d[8], r0, r1, r2, a0, a1, a2);
      }
    }
    rdone += cnt;
  };
  VT cy0 = VT(0), cy1 = VT(0), cy2 = VT(0);
  const int kb = hb >> 6, ke = (he - 1) >> 6;
  for (int k = kb; k <= ke; ++k) {
    const unsigned long long H = T.hbits[k];
    const unsigned long long Hn = k + 1 < T.nch ? T.hbits[k + 1] : 1ull;
    const int p = (k << 6) + lane;
    const bool valid = p >= hb && p < he;
    const int pc = min(max(p, hb), he - 1);
    const unsigned long long m = H & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int s0 = m ? 63 - __clzll((long long)m) : 0;  // first lane of this lane's row in the chunk
    const int row = min(max(T.hrow[k] - 1 + __popcll(m), 0), S.nb - 1);
    const bool next_head = lane < 63 ? ((H >> (lane + 1)) & 1ull) != 0 : (Hn & 1ull) != 0;
    const bool end = valid && (next_head || p + 1 == he);
    const int c = __builtin_nontemporal_load(S.colidx + pc);
    VT v[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) v[e] = __builtin_nontemporal_load(val + val_index(pc, e));
    const Vec3<VT> xv = gather3(x, c);
    const bool carry_in = lane == 0 && (H & 1ull) == 0;
    VT a0 = carry_in ? cy0 : VT(0), a1 = carry_in ? cy1 : VT(0), a2 = carry_in ? cy2 : VT(0);
    blk_step(valid, v, xv.v0, xv.v1, xv.v2, a0, a1, a2);
    seg_scan3(lane, s0, a0, a1, a2);
    cy0 = lane63(a0);
    cy1 = lane63(a1);
    cy2 = lane63(a2);
    if (end) {
      const int sl = (row - rb) & 127;
      s_sum[wl][sl][0] = a0;
      s_sum[wl][sl][1] = a1;
      s_sum[wl][sl][2] = a2;
    }
    // the ring's writes before any lane's read of them (LDS operations of
    // one wave complete in order)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // rows finished through this chunk: those that start in it or before,
    // less the last one unless the next chunk starts with a row
    const int fin = k == ke ? re : min(re, (Hn & 1ull) ? T.hrow[k + 1] : T.hrow[k + 1] - 1);
    while (fin - rdone >= 64) epilogue(64);
    if (k == ke && fin > rdone) epilogue(fin - rdone);
    __builtin_amdgcn_wave_barrier();
  }
}

// head bits and rows-before of each chunk: one thread per chunk k <= nch
// (hrow[nch] = nb closes the table)
__global__ __launch_bounds__(kBlock) void k_chunk_heads(const int* __restrict__ rowptr, int nb, int nch,
                                                        unsigned long long* __restrict__ hbits,
                                                        int* __restrict__ hrow) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k > nch) return;
  const long long lo = 64ll * k;
  int a = 0, b = nb;  // first row whose first block is >= lo
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (rowptr[mid] < lo) a = mid + 1;
    else b = mid;
  }
  hrow[k] = a;
  if (k == nch) return;
  unsigned long long bits = 0;
  for (int r = a; r < nb && rowptr[r] < lo + 64; ++r) bits |= 1ull << (rowptr[r] - lo);
  hbits[k] = bits;
}

// Waves of a strip launch when the caller leaves it open: as many as fit on
// the device at once (occupancy x CUs), at most one per chunk, so the grid
// runs in one round with equal shares of chunks and no tail of late
// workgroups (on 8.3M blocks: 5.1-5.2 TB/s, against 4.9-5.2 for fixed runs of
// 8 to 16 chunks per wave, whose best length differed by kernel).
static int strip_resident_waves(int kernel) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_pair(dev, kernel);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, blocks = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn = kernel == 0   ? (const void*)op_strip<0, double>
                   : kernel == 1 ? (const void*)op_strip<1, double>
                                 : (const void*)op_strip<1, float>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, kStripBlock, 0) != hipSuccess) return 0;
  const int waves = std::max(1, blocks) * cus * (kStripBlock / 64);
  cache[key] = waves;
  return waves;
}

hipError_t launch_chunk_heads(const int* rowptr, int nb, int nnzb, unsigned long long* hbits, int* hrow,
                              hipStream_t st) {
  const int nch = (nnzb + 63) / 64;
  hipLaunchKernelGGL(k_chunk_heads, dim3((nch + kBlock) / kBlock), dim3(kBlock), 0, st, rowptr, nb, nch, hbits,
                     hrow);
  return hipGetLastError();
}

hipError_t launch_strip_op(int kernel, const SysDesc& S, const StripDesc& T, const void* x, void* y, double w,
                           hipStream_t st) {
  if (T.nch == 0) return hipSuccess;
  if (kernel < 0 || kernel > 2) return hipErrorInvalidValue;
  StripDesc D = T;
  if (D.nwave <= 0) {
    const int res = strip_resident_waves(kernel);
    if (res <= 0) return hipErrorInvalidDevice;
    D.nwave = std::min(D.nch, res);
  }
  D.nwave = std::min(D.nwave, D.nch);
  const dim3 grid((D.nwave + kStripBlock / 64 - 1) / (kStripBlock / 64));
  if (kernel == 0)
    hipLaunchKernelGGL((op_strip<0, double>), grid, dim3(kStripBlock), 0, st, S, D, (const double*)x, (double*)y,
                       0.0);
  else if (kernel == 1)
    hipLaunchKernelGGL((op_strip<1, double>), grid, dim3(kStripBlock), 0, st, S, D, (const double*)x, (double*)y,
                       w);
  else
    hipLaunchKernelGGL((op_strip<1, float>), grid, dim3(kStripBlock), 0, st, S, D, (const float*)x, (float*)y,
                       (float)w);
  return hipGetLastError();
}

// ---------------------------------------------------------------- vector kernels
// (host-driven flexible PCG of the learned preconditioner)
__global__ __launch_bounds__(kBlock) void k_dot_partial(const double* a, const double* b, int n, double* part) {
  __shared__ double red[kBlock / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += a[i] * b[i];
  const double t = block_sum_t0(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// fixed-order sum of np partials by one wave: lane l adds l, l+64, ... in
// order, then a butterfly (launch with 64 threads)
__global__ void k_dot_final(const double* part, int np, double* out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 64) s += part[i];
  s = wave_sum(s);
  if (threadIdx.x == 0) *out = s;
}
// x += alpha p ; r -= alpha Ap

// One thread per block row i of D: the neighbour CSR of the BSR pattern
// without the diagonal block (nbr_ptr[i] = rowptr[i] - i, the row's other
// columns in order) and the row's diagonal features Dfeat_i = D_ii / s_i
// (row-major 3x3), sdiag_i = s_i = tr(D_ii) / 3 (1 when not positive).
__global__ void k_node_graph(const int* rowptr, const int* colidx, const double* val64, int nb, int* nbr_ptr,
                             int* nbr, float* Dfeat, float* sdiag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nb) return;
  nbr_ptr[i] = rowptr[i] - i;
  if (i == nb) return;
  int o = rowptr[i] - i, pd = -1;
  for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
    const int c = colidx[p];
    if (c != i) nbr[o++] = c;
    else pd = p;
  }
  double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (pd >= 0)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[3 * r + c] = val64[val_index(pd, r + 3 * c)];
  double s = (B[0] + B[4] + B[8]) / 3.0;
  if (!(s > 0.0)) s = 1.0;
  for (int e = 0; e < 9; ++e) Dfeat[9 * i + e] = (float)(B[e] / s);
  sdiag[i] = (float)s;
}

hipError_t launch_dot_partial(const double* a, const double* b, int n, double* part, int* np_out, hipStream_t st) {
  const int np = min(256, max(1, (n + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_dot_partial, dim3(np), dim3(kBlock), 0, st, a, b, n, part);
  *np_out = np;
  return hipGetLastError();
}
hipError_t launch_node_graph(const int* rowptr, const int* colidx, const double* val64, int nb, int* nbr_ptr,
                             int* nbr, float* Dfeat, float* sdiag, hipStream_t st) {
  hipLaunchKernelGGL(k_node_graph, dim3((nb + 1 + 255) / 256), dim3(256), 0, st, rowptr, colidx, val64, nb, nbr_ptr,
                     nbr, Dfeat, sdiag);
  return hipGetLastError();
}

// explicit instantiations via host launchers
template <int G, typename VT, bool RES, bool PIPE = false>
static hipError_t launch_pcg_t(const PcgParams& P, size_t shmem, hipStream_t st) {
  void* args[] = {(void*)&P};
  return hipLaunchCooperativeKernel((const void*)pcg_persistent<G, VT, RES, PIPE>, dim3(P.nwg), dim3(kBlock), args,
                                    shmem, st);
}

// the same kernel through an ordinary launch (used inside a captured graph,
// whose kernel node is then marked cooperative); the grid never exceeds the
// co-resident capacity computed from the occupancy, and every grid wait is
// bounded
template <int G, typename VT, bool RES, bool PIPE = false>
static hipError_t launch_pcg_plain_t(const PcgParams& P, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((pcg_persistent<G, VT, RES, PIPE>), dim3(P.nwg), dim3(kBlock), shmem, st, P);
  return hipGetLastError();
}

template <int G, typename VT, bool RES, bool PIPE = false>
static hipError_t occ_t(int* blocks, size_t shmem) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, pcg_persistent<G, VT, RES, PIPE>, kBlock, shmem);
}

__global__ void k_zero_u32(unsigned* p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// several buffers zeroed by one launch: segment s covers 32-bit words
// [start[s], start[s+1]) of the concatenation
__global__ void k_zero_multi(ZeroList L) {
  const long long total = L.start[L.n];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    int s = 0;
    while (i >= L.start[s + 1]) ++s;
    static_cast<unsigned*>(L.p[s])[i - L.start[s]] = 0u;
  }
}

hipError_t launch_zero_multi(const void* const* ptrs, const size_t* bytes, int n, hipStream_t st) {
  if (n < 0 || n > kZeroListMax) return hipErrorInvalidValue;
  ZeroList L{};
  L.n = n;
  L.start[0] = 0;
  for (int s = 0; s < n; ++s) {
    if (bytes[s] % 4 != 0) return hipErrorInvalidValue;
    L.p[s] = const_cast<void*>(ptrs[s]);
    L.start[s + 1] = L.start[s] + (long long)(bytes[s] / 4);
  }
  const long long total = L.start[n];
  if (total == 0) return hipSuccess;
  const int grid = (int)std::min<long long>((total + 255) / 256, 2048);
  hipLaunchKernelGGL(k_zero_multi, dim3(grid), dim3(256), 0, st, L);
  return hipGetLastError();
}

// small device results gathered into one staging area by one launch, so
// that a single copy brings them to the host: entry e copies n[e] 32-bit
// words from src[e] (zeros when src[e] is null) to dst[e]
__global__ __launch_bounds__(256) void k_gather_words(GatherList L) {
  for (int e = 0; e < L.count; ++e) {
    const unsigned* src = L.src[e];
    unsigned* dst = L.dst[e];
    for (int i = threadIdx.x; i < L.n[e]; i += blockDim.x) dst[i] = src ? src[i] : 0u;
  }
}

hipError_t launch_gather_words(const GatherList& L, hipStream_t st) {
  if (L.count < 0 || L.count > kGatherMax) return hipErrorInvalidValue;
  for (int e = 0; e < L.count; ++e)
    if (L.n[e] < 0 || (L.n[e] > 0 && !L.dst[e])) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gather_words, dim3(1), dim3(256), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
  if (bytes % 4 != 0) return hipErrorInvalidValue;
  const long long n = (long long)(bytes / 4);
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<long long>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_zero_u32, dim3(grid), dim3(256), 0, st, (unsigned*)p, n);
  return hipGetLastError();
}

// dispatch over lanes per row (4/8/16), precision and schedule
#define HSC_PCG_DISPATCH(FN, ...)                                                       \
  do {                                                                                  \
    if (pipe) {                                                                         \
      if (G == 4 && res) {                                                              \
        if (mixed) return FN<4, float, true, true>(__VA_ARGS__);                        \
        return FN<4, double, true, true>(__VA_ARGS__);                                  \
      }                                                                                 \
    } else if (mixed) {                                                                        \
      if (res) {                                                                        \
        if (G == 4) return FN<4, float, true>(__VA_ARGS__);                             \
        if (G == 8) return FN<8, float, true>(__VA_ARGS__);                             \
        if (G == 16) return FN<16, float, true>(__VA_ARGS__);                           \
      } else {                                                                          \
        if (G == 4) return FN<4, float, false>(__VA_ARGS__);                            \
        if (G == 8) return FN<8, float, false>(__VA_ARGS__);                            \
        if (G == 16) return FN<16, float, false>(__VA_ARGS__);                          \
      }                                                                                 \
    } else {                                                                            \
      if (res) {                                                                        \
        if (G == 4) return FN<4, double, true>(__VA_ARGS__);                            \
        if (G == 8) return FN<8, double, true>(__VA_ARGS__);                            \
        if (G == 16) return FN<16, double, true>(__VA_ARGS__);                          \
      } else {                                                                          \
        if (G == 4) return FN<4, double, false>(__VA_ARGS__);                           \
        if (G == 8) return FN<8, double, false>(__VA_ARGS__);                           \
        if (G == 16) return FN<16, double, false>(__VA_ARGS__);                         \
      }                                                                                 \
    }                                                                                   \
  } while (0)

hipError_t launch_pcg(int G, bool mixed, bool res, bool pipe, const PcgParams& P, size_t shmem, hipStream_t st) {
  HSC_PCG_DISPATCH(launch_pcg_t, P, shmem, st);
  return hipErrorInvalidValue;
}

hipError_t launch_pcg_plain(int G, bool mixed, bool res, bool pipe, const PcgParams& P, size_t shmem,
                            hipStream_t st) {
  HSC_PCG_DISPATCH(launch_pcg_plain_t, P, shmem, st);
  return hipErrorInvalidValue;
}

template <int G, typename VT, bool RES, bool PIPE = false>
static hipError_t lds_t(size_t* bytes) {
  hipFuncAttributes a;
  const hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(pcg_persistent<G, VT, RES, PIPE>));
  if (e == hipSuccess) *bytes = a.sharedSizeBytes;
  return e;
}

hipError_t pcg_static_lds(int G, bool mixed, bool res, bool pipe, size_t* bytes) {
  HSC_PCG_DISPATCH(lds_t, bytes);
  return hipErrorInvalidValue;
}

static hipError_t pcg_occupancy_query(int G, bool mixed, bool res, bool pipe, size_t shmem, int* blocks) {
  HSC_PCG_DISPATCH(occ_t, blocks, shmem);
  return hipErrorInvalidValue;
}

// the runtime's occupancy query costs tens of microseconds of host time; its
// answer depends only on the device, the kernel and the LDS size, so it is
// kept per process
hipError_t pcg_occupancy(int G, bool mixed, bool res, bool pipe, size_t shmem, int* blocks) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int, size_t>, int> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const auto key = std::make_tuple(dev, G, (int)mixed, (int)res, (int)pipe, shmem);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *blocks = it->second;
      return hipSuccess;
    }
  }
  e = pcg_occupancy_query(G, mixed, res, pipe, shmem, blocks);
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> g(mu);
    cache[key] = *blocks;
  }
  return e;
}

template <int G, typename VT, bool RES, bool PIPE = false>
static const void* kernel_fn_t() {
  return (const void*)pcg_persistent<G, VT, RES, PIPE>;
}

const void* pcg_kernel_fn(int G, bool mixed, bool res, bool pipe) {
  HSC_PCG_DISPATCH(kernel_fn_t);
  return nullptr;
}

// one launch of a row-mapped operator kernel (G lanes per block row): 0 =
// FP64 y = D x, 1 = FP64 pass y = x + w Dinv (r - D x), 2 = the pass in FP32
hipError_t launch_row_op(int kernel, int G, const SysDesc& S, const void* x, void* y, double w, hipStream_t st) {
  const int tiles = (S.nb + (kBlock / G) - 1) / (kBlock / G);
  if (tiles == 0) return hipSuccess;
#define HSC_TIMED(GG)                                                                                          \
  if (kernel == 0)                                                                                             \
    hipLaunchKernelGGL(op_Dx64<GG>, dim3(tiles), dim3(kBlock), 0, st, S, (const double*)x, (double*)y);        \
  else if (kernel == 1)                                                                                        \
    hipLaunchKernelGGL((op_neumann_step<GG, double>), dim3(tiles), dim3(kBlock), 0, st, S, (const double*)x,   \
                       (double*)y, w);                                                                         \
  else                                                                                                         \
    hipLaunchKernelGGL((op_neumann_step<GG, float>), dim3(tiles), dim3(kBlock), 0, st, S, (const float*)x,     \
                       (float*)y, w);
  if (kernel < 0 || kernel > 2) return hipErrorInvalidValue;
  switch (G) {
    case 4: HSC_TIMED(4) break;
    case 8: HSC_TIMED(8) break;
    case 16: HSC_TIMED(16) break;
    default: return hipErrorInvalidValue;
  }
#undef HSC_TIMED
  return hipGetLastError();
}

hipError_t launch_op_precond64(const StripDesc* T, int G, const SysDesc& S, int order, const double* omegas,
                               const double* r, double* z, double* tmp, hipStream_t st) {
  const int ntile = (S.nb + kBlock - 1) / kBlock;
  if (S.nb == 0) return hipSuccess;
  auto w = [&](int j) { return omegas[j < kMaxPolySteps ? j : kMaxPolySteps - 1]; };
  // z0 into (order odd ? z : tmp) so that the last step lands in z
  double* bufs[2] = {z, tmp};
  int cur = ((order - 1) % 2 == 0) ? 0 : 1;
  hipLaunchKernelGGL(op_jacobi<double>, dim3(ntile), dim3(kBlock), 0, st, S, r, bufs[cur], w(0));
  SysDesc S2 = S;
  S2.r = const_cast<double*>(r);
  for (int j = 1; j < order; ++j) {
    const hipError_t e = T ? launch_strip_op(1, S2, *T, bufs[cur], bufs[cur ^ 1], w(j), st)
                           : launch_row_op(1, G, S2, bufs[cur], bufs[cur ^ 1], w(j), st);
    if (e != hipSuccess) return e;
    cur ^= 1;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- row-tile neighbours
// bits[t][u] = 1 when a row of tile t (R rows per tile) has a block in a row
// of tile u != t: one workgroup per tile, the tile's blocks are contiguous
__global__ __launch_bounds__(kBlock) void k_tile_nbr_bits(const int* __restrict__ rowptr,
                                                          const int* __restrict__ colidx, int nb, int R,
                                                          const int* __restrict__ tstart, int rtl, int words,
                                                          unsigned* __restrict__ bits) {
  __shared__ unsigned s_bits[kNbrMaxTiles / 32];
  __shared__ int s_ts[kNbrMaxTiles + 1];
  const int t = blockIdx.x;
  for (int w = threadIdx.x; w < words; w += kBlock) s_bits[w] = 0u;
  if (tstart)
    for (int i = threadIdx.x; i <= rtl; i += kBlock) s_ts[i] = tstart[i];
  __syncthreads();
  const int r0 = tstart ? s_ts[t] : t * R, r1 = tstart ? s_ts[t + 1] : min(nb, r0 + R);
  const int p0 = rowptr[r0], p1 = rowptr[r1];
  for (int p = p0 + threadIdx.x; p < p1; p += kBlock) {
    const int c = colidx[p];
    int u = c / R;
    if (tstart) {  // last tile starting at or before c
      int lo = 0, hi = rtl - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_ts[mid] <= c) lo = mid;
        else hi = mid - 1;
      }
      u = lo;
    }
    if (u != t) atomicOr(&s_bits[u >> 5], 1u << (u & 31));
  }
  __syncthreads();
  for (int w = threadIdx.x; w < words; w += kBlock) bits[(size_t)t * words + w] = s_bits[w];
}

// list[t] = {count, u...} over u with bits[t][u] | bits[u][t], ascending;
// count -1 when more than kNbrStride - 1 (the tile then waits on every tile)
__global__ __launch_bounds__(kBlock) void k_tile_nbr_list(const unsigned* __restrict__ bits, int rtl, int words,
                                                          int* __restrict__ list) {
  __shared__ int s_wc[kBlock / 64];
  const int t = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int base = 0;
  for (int u0 = 0; u0 < rtl; u0 += kBlock) {
    const int u = u0 + threadIdx.x;
    bool f = false;
    if (u < rtl && u != t)
      f = (((bits[(size_t)t * words + (u >> 5)] >> (u & 31)) | (bits[(size_t)u * words + (t >> 5)] >> (t & 31))) &
           1u) != 0u;
    const unsigned long long m = __ballot(f);
    if (lane == 0) s_wc[w] = __popcll(m);
    __syncthreads();
    int off = base, tot = base;
    for (int i = 0; i < kBlock / 64; ++i) {
      if (i < w) off += s_wc[i];
      tot += s_wc[i];
    }
    if (f) {
      const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < kNbrStride - 1) list[(size_t)t * kNbrStride + 1 + pos] = u;
    }
    __syncthreads();
    base = tot;
  }
  if (threadIdx.x == 0) list[(size_t)t * kNbrStride] = base <= kNbrStride - 1 ? base : -1;
}

hipError_t launch_tile_neighbors(const int* rowptr, const int* colidx, int nb, int R, const int* tstart, int ntiles,
                                 unsigned* bits, int* list, hipStream_t st) {
  const int rtl = tstart ? ntiles : (nb + R - 1) / R;
  if (rtl <= 0 || rtl > kNbrMaxTiles) return hipErrorInvalidValue;
  const int words = (rtl + 31) / 32;
  hipLaunchKernelGGL(k_tile_nbr_bits, dim3(rtl), dim3(kBlock), 0, st, rowptr, colidx, nb, R, tstart, rtl, words, bits);
  hipLaunchKernelGGL(k_tile_nbr_list, dim3(rtl), dim3(kBlock), 0, st, bits, rtl, words, list);
  return hipGetLastError();
}

// ---------------------------------------------------------------- body-aligned row tiles
// Row tiles of the latency schedule at 4 lanes per row, cut along D's
// connected components: a component of at least kTileRows rows (a body's
// mesh) starts a new tile, and tiles are cut every kTileRows rows from the
// last such start; smaller components share tiles.  A tile that straddles
// two bodies pays for both bodies' halos in every preconditioner pass, and
// with fixed 128-row tiles the five tiles where the 78k-tet system's tori
// meet set the pace of every pass pair.  Row r ends a component when no row
// r' <= r references a column beyond r: the running maximum of each row's
// largest column (rows keep their diagonal first, then columns ascending)
// equals r.  Each row's largest column is read by one thread per row first;
// then one workgroup: each thread takes 16 consecutive rows at a time, and
// workgroup scans carry the running maximum and the last component end.
constexpr int kTileRows = 128;
constexpr int kTileStage = 16;                   // rows per thread per pass
constexpr int kTiledRowsMax = kTileRows * kNbrMaxTiles;

__device__ int block_scan_excl(int v, int* wsum, int* total);

// Exclusive maximum of v over the workgroup (thread order, -1 for thread
// 0); *total gets the maximum over all threads.  Contains __syncthreads.
__device__ int block_scan_max_excl(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    x = lane >= d ? max(x, y) : x;
  }
  int ex = __shfl_up(x, 1, 64);
  ex = lane >= 1 ? ex : -1;
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int off = -1, tot = -1;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const int t = wsum[w];
    off = w < wv ? max(off, t) : off;
    tot = max(tot, t);
  }
  __syncthreads();
  *total = tot;
  return max(off, ex);
}

// reach[r] = max(r, the largest column of row r), one thread per row
__global__ __launch_bounds__(kBlock) void k_row_reach(const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                                      int nb, int* __restrict__ reach) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < nb) reach[r] = max(r, colidx[rowptr[r + 1] - 1]);
}

__global__ __launch_bounds__(kBlock) void k_row_tiles(const int* __restrict__ reach_g,
                                                      const int* __restrict__ act_node, int nb,
                                                      int* __restrict__ tstart, int* __restrict__ out) {
  __shared__ int s_big[kNbrMaxTiles + 1];
  __shared__ int s_tile[kNbrMaxTiles + 1];
  __shared__ int s_w[kBlock / 64];
  const int tid = threadIdx.x;
  if (nb <= 0 || nb > kTiledRowsMax) {
    if (tid == 0) out[0] = out[1] = 0;
    return;
  }
  int carry_max = -1, carry_end = -1, nbig = 0;
  for (int base = 0; base < nb; base += kBlock * kTileStage) {
    // this thread's kTileStage consecutive rows: their ends, then their last
    // columns, all in flight together (indices clamped into the system;
    // rows past it are masked)
    const int r0 = base + tid * kTileStage;
    int reach[kTileStage];
#pragma unroll
    for (int j = 0; j < kTileStage; ++j) reach[j] = reach_g[min(r0 + j, nb - 1)];
    int lm = -1;
#pragma unroll
    for (int j = 0; j < kTileStage; ++j) {
      reach[j] = r0 + j < nb ? reach[j] : 0x7fffffff;
      lm = r0 + j < nb ? max(lm, reach[j]) : lm;
    }
    int tmax = 0;
    const int before = max(carry_max, block_scan_max_excl(lm, s_w, &tmax));
    int run = before, last_end = -1;
#pragma unroll
    for (int j = 0; j < kTileStage; ++j) {
      run = max(run, reach[j]);
      last_end = run == r0 + j ? r0 + j : last_end;  // (rows past nb never match)
    }
    int tend = 0;
    const int prev_end = max(carry_end, block_scan_max_excl(last_end, s_w, &tend));
    // starts of the components of at least kTileRows rows that end in this
    // run, in order
    int prev = prev_end, cnt = 0;
    run = before;
#pragma unroll
    for (int j = 0; j < kTileStage; ++j) {
      run = max(run, reach[j]);
      const bool end = run == r0 + j;
      cnt += end && r0 + j - prev >= kTileRows ? 1 : 0;
      prev = end ? r0 + j : prev;
    }
    int total = 0;
    int pos = nbig + block_scan_excl(cnt, s_w, &total);
    prev = prev_end;
    run = before;
#pragma unroll
    for (int j = 0; j < kTileStage; ++j) {
      run = max(run, reach[j]);
      if (run == r0 + j) {
        if (r0 + j - prev >= kTileRows && pos < kNbrMaxTiles) s_big[pos++] = prev + 1;
        prev = r0 + j;
      }
    }
    nbig += total;
    carry_max = max(carry_max, tmax);
    carry_end = max(carry_end, tend);
  }
  __syncthreads();
  if (nbig > kNbrMaxTiles) {
    if (tid == 0) out[0] = out[1] = 0;
    return;
  }
  // segments: [0 or a big start, the next big start or nb); tiles every
  // kTileRows rows from a segment's start
  const bool lead = nbig == 0 || s_big[0] != 0;
  const int nseg = nbig + (lead ? 1 : 0);
  auto seg_start = [&](int j) { return lead ? (j == 0 ? 0 : s_big[j - 1]) : s_big[j]; };
  auto seg_end = [&](int j) { return j + 1 < nseg ? seg_start(j + 1) : nb; };
  const int per = (nseg + kBlock - 1) / kBlock;
  const int j0 = min(nseg, tid * per), j1 = min(nseg, j0 + per);
  int mine = 0;
  for (int j = j0; j < j1; ++j) mine += (seg_end(j) - seg_start(j) + kTileRows - 1) / kTileRows;
  int ntiles = 0;
  int off = block_scan_excl(mine, s_w, &ntiles);
  if (ntiles > kNbrMaxTiles) {
    if (tid == 0) out[0] = out[1] = 0;
    return;
  }
  for (int j = j0; j < j1; ++j)
    for (int a = seg_start(j), e = seg_end(j); a < e; a += kTileRows) s_tile[off++] = a;
  if (tid == 0) s_tile[ntiles] = nb;
  __syncthreads();
  for (int t = tid; t <= ntiles; t += kBlock) tstart[t] = s_tile[t];
  // most coupled nodes in one tile (sizes the LDS cache of their W rows):
  // each thread counts a run of consecutive rows into their tiles
  int* s_cnt = s_big;
  for (int t = tid; t < ntiles; t += kBlock) s_cnt[t] = 0;
  __syncthreads();
  const int per2 = (nb + kBlock - 1) / kBlock;
  const int a0 = min(nb, tid * per2), a1 = min(nb, a0 + per2);
  if (a0 < a1) {
    int lo = 0, hi = ntiles - 1;  // the tile of row a0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_tile[mid] <= a0) lo = mid;
      else hi = mid - 1;
    }
    int t = lo, tend2 = s_tile[lo + 1], c = 0;
    for (int r = a0; r < a1; r += kTileStage) {
      int an[kTileStage];
#pragma unroll
      for (int j = 0; j < kTileStage; ++j) an[j] = act_node[min(r + j, a1 - 1)];
#pragma unroll
      for (int j = 0; j < kTileStage; ++j) {
        if (r + j >= a1) break;
        if (r + j >= tend2) {
          if (c) atomicAdd(&s_cnt[t], c);
          c = 0;
          while (s_tile[t + 1] <= r + j) ++t;
          tend2 = s_tile[t + 1];
        }
        c += an[j] >= 0 ? 1 : 0;
      }
    }
    if (c) atomicAdd(&s_cnt[t], c);
  }
  __syncthreads();
  int mc = 0;
  for (int t = tid; t < ntiles; t += kBlock) mc = max(mc, s_cnt[t]);
  int mct = 0;
  (void)block_scan_max_excl(mc, s_w, &mct);
  if (tid == 0) {
    out[0] = ntiles;
    out[1] = max(mct, 0);
  }
}

hipError_t launch_row_tiles(const int* rowptr, const int* colidx, const int* act_node, int nb, int* reach,
                            int* tstart, int* out, hipStream_t st) {
  if (nb > 0 && nb <= kTiledRowsMax)
    hipLaunchKernelGGL(k_row_reach, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, rowptr, colidx, nb, reach);
  hipLaunchKernelGGL(k_row_tiles, dim3(1), dim3(kBlock), 0, st, reach, act_node, nb, tstart, out);
  return hipGetLastError();
}
int row_tiles_capacity() { return kNbrMaxTiles + 1; }

// ---------------------------------------------------------------- pass-pair tables
// Exclusive prefix sum of v over the workgroup (thread order); *total gets
// the sum.  wsum holds kBlock / 64 ints.  Contains __syncthreads.
__device__ int block_scan_excl(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    x += lane >= d ? y : 0;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const int t = wsum[w];
    off += w < wv ? t : 0;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// Sorted, distinct node lists through a bitmap over [lo, lo + 32 words): the
// exclusive prefix of the words' popcounts, so a member's rank is pre[w] +
// popc(bits below it).  pre[w] = set bits of bits[0 .. w); returns the
// total.  Each thread counts a run of consecutive words, a workgroup scan
// places the runs.
__device__ int bitmap_prefix(const unsigned* bits, int* pre, int words, int* wsum) {
  const int tid = threadIdx.x;
  const int per = (words + kBlock - 1) / kBlock;
  const int w0 = min(words, tid * per), w1 = min(words, w0 + per);
  int mine = 0;
  for (int w = w0; w < w1; ++w) mine += __popc(bits[w]);
  int total = 0;
  int acc = block_scan_excl(mine, wsum, &total);
  for (int w = w0; w < w1; ++w) {
    pre[w] = acc;
    acc += __popc(bits[w]);
  }
  __syncthreads();
  return total;
}
__device__ __forceinline__ bool bitmap_has(const unsigned* bits, int lo, int words, int c) {
  const int o = c - lo;
  return o >= 0 && (o >> 5) < words && ((bits[o >> 5] >> (o & 31)) & 1u);
}
__device__ __forceinline__ int bitmap_rank(const unsigned* bits, const int* pre, int lo, int c) {
  const int o = c - lo, w = o >> 5;
  return pre[w] + __popc(bits[w] & ((1u << (o & 31)) - 1u));
}
// block min / max / or over the workgroup (every thread gets the result)
__device__ void block_minmaxor(int& lo, int& hi, int& f, int* red) {
  for (int m = 32; m >= 1; m >>= 1) {
    lo = min(lo, __shfl_xor(lo, m, 64));
    hi = max(hi, __shfl_xor(hi, m, 64));
    f |= __shfl_xor(f, m, 64);
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[3 * wv] = lo;
    red[3 * wv + 1] = hi;
    red[3 * wv + 2] = f;
  }
  __syncthreads();
  lo = red[0];
  hi = red[1];
  f = red[2];
  for (int w = 1; w < kBlock / 64; ++w) {
    lo = min(lo, red[3 * w]);
    hi = max(hi, red[3 * w + 1]);
    f |= red[3 * w + 2];
  }
  __syncthreads();
}

// Halo pack of row tile t (layout in common.hpp); one workgroup per tile.
// H1 (the halo rows) = the distinct columns outside the tile that its rows
// reference, ascending; H2 = the distinct columns of the halo rows outside
// the tile and H1, ascending.  A tile that does not fit (more than kHaloRows
// halo rows or kHaloSet set slots or hbl halo blocks, a row of more than
// 4 kHaloU blocks, a column range past a bitmap) sets the system's fail
// word, and the solve keeps one synchronisation per pass.
constexpr int kHaloBitmapWords = 4096;  // column range of 131,072 nodes
constexpr int kHaloBlkMax = 4 * kHaloU * kHaloRows;  // the largest pack capacity (hbl)
// Every phase is spread over the workgroup: the halo rows' blocks are
// numbered 0 .. nblk-1 in halo-row order (block offsets by a workgroup scan)
// and their global block indices and columns staged in LDS, so the column
// loads, the H2 bitmap, the slots and the value copies run one block (or one
// block component) per thread.
__global__ __launch_bounds__(kBlock) void k_tile_halo(const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                                      const float* __restrict__ val32,
                                                      const float* __restrict__ invD32, int nb,
                                                      const int* __restrict__ tstart, int hbl,
                                                      unsigned char* __restrict__ base, long long stride) {
  __shared__ unsigned s_b1[kHaloBitmapWords], s_b2[kHaloBitmapWords];
  __shared__ int s_p1[kHaloBitmapWords], s_p2[kHaloBitmapWords];
  __shared__ int s_qf[kHaloBlkMax], s_col[kHaloBlkMax];
  __shared__ int s_h1[kHaloRows], s_len[kHaloRows], s_lp0[kHaloRows], s_q0[kHaloRows];
  __shared__ int s_red[3 * (kBlock / 64)], s_wsum[kBlock / 64];
  constexpr int R = kHaloTileRows;
  const int t = blockIdx.x, tid = threadIdx.x;
  const int r0 = tstart ? tstart[t] : t * R, r1 = tstart ? tstart[t + 1] : min(nb, r0 + R);
  const int p0 = rowptr[r0], p1 = rowptr[r1];
  int* fail = reinterpret_cast<int*>(base);
  unsigned char* pk = base + kHaloBaseBytes + (size_t)t * stride;
  auto give_up = [&]() {
    if (tid == 0) atomicOr(fail, 1);
  };

  // H1: columns of the tile's rows outside the tile
  int lo1 = 0x7fffffff, hi1 = -1, bad = 0;
  for (int p = p0 + tid; p < p1; p += kBlock) {
    const int c = colidx[p];
    if (c < r0 || c >= r1) {
      lo1 = min(lo1, c);
      hi1 = max(hi1, c);
    }
  }
  for (int r = r0 + tid; r < r1; r += kBlock) bad |= (rowptr[r + 1] - rowptr[r] > 4 * kHaloU) ? 1 : 0;
  block_minmaxor(lo1, hi1, bad, s_red);
  const int words1 = hi1 >= lo1 ? (hi1 - lo1 + 1 + 31) / 32 : 0;
  if (bad || words1 > kHaloBitmapWords) return give_up();
  for (int w = tid; w < words1; w += kBlock) s_b1[w] = 0u;
  __syncthreads();
  for (int p = p0 + tid; p < p1; p += kBlock) {
    const int c = colidx[p];
    if (c < r0 || c >= r1) atomicOr(&s_b1[(c - lo1) >> 5], 1u << ((c - lo1) & 31));
  }
  __syncthreads();
  const int nh = bitmap_prefix(s_b1, s_p1, words1, s_wsum);
  if (nh > kHaloRows) return give_up();
  for (int w = tid; w < words1; w += kBlock) {
    unsigned bits = s_b1[w];
    int idx = s_p1[w];
    while (bits) {
      const int bi = __ffs(bits) - 1;
      s_h1[idx++] = lo1 + 32 * w + bi;
      bits &= bits - 1u;
    }
  }
  __syncthreads();
  // halo rows' lengths and first blocks; their blocks' offsets in the pack
  int len = 0;
  if (tid < nh) {
    const int h = s_h1[tid];
    const int q0 = rowptr[h];
    len = rowptr[h + 1] - q0;
    s_q0[tid] = q0;
    s_len[tid] = len;
  }
  int nblk = 0;
  const int lp0 = block_scan_excl(len, s_wsum, &nblk);
  if (nblk > hbl) return give_up();
  if (tid < nh) {
    s_lp0[tid] = lp0;
    for (int b = 0; b < len; ++b) s_qf[lp0 + b] = s_q0[tid] + b;
  }
  __syncthreads();
  // the halo blocks' columns, and H2's range
  int lo2 = 0x7fffffff, hi2 = -1;
  bad = len > 4 * kHaloU ? 1 : 0;
  for (int f = tid; f < nblk; f += kBlock) {
    const int c = colidx[s_qf[f]];
    s_col[f] = c;
    if ((c < r0 || c >= r1) && !bitmap_has(s_b1, lo1, words1, c)) {
      lo2 = min(lo2, c);
      hi2 = max(hi2, c);
    }
  }
  block_minmaxor(lo2, hi2, bad, s_red);
  const int words2 = hi2 >= lo2 ? (hi2 - lo2 + 1 + 31) / 32 : 0;
  if (bad || words2 > kHaloBitmapWords) return give_up();
  for (int w = tid; w < words2; w += kBlock) s_b2[w] = 0u;
  __syncthreads();
  for (int f = tid; f < nblk; f += kBlock) {
    const int c = s_col[f];
    if ((c < r0 || c >= r1) && !bitmap_has(s_b1, lo1, words1, c))
      atomicOr(&s_b2[(c - lo2) >> 5], 1u << ((c - lo2) & 31));
  }
  __syncthreads();
  const int n2 = bitmap_prefix(s_b2, s_p2, words2, s_wsum);
  const int nset = R + nh + n2;
  if (nset > kHaloSet) return give_up();
  auto slot_of = [&](int c) -> int {
    if (c >= r0 && c < r1) return c - r0;
    if (bitmap_has(s_b1, lo1, words1, c)) return R + bitmap_rank(s_b1, s_p1, lo1, c);
    return R + nh + bitmap_rank(s_b2, s_p2, lo2, c);
  };
  // the pack
  if (tid == 0) {
    int* hdr = reinterpret_cast<int*>(pk);
    hdr[0] = nh;
    hdr[1] = nset;
    hdr[2] = nblk;
    hdr[3] = 0;
  }
  int* setn = reinterpret_cast<int*>(pk + kHaloOffSet);
  for (int w = tid; w < words2; w += kBlock) {
    unsigned bits = s_b2[w];
    int idx = nh + s_p2[w];
    while (bits) {
      const int bi = __ffs(bits) - 1;
      setn[idx++] = lo2 + 32 * w + bi;
      bits &= bits - 1u;
    }
  }
  if (tid < nh) {
    const int h = s_h1[tid];
    setn[tid] = h;
    reinterpret_cast<int2*>(pk + kHaloOffRows)[tid] = make_int2(s_len[tid], s_lp0[tid]);
  }
  float* inv = reinterpret_cast<float*>(pk + kHaloOffInv);
  for (int j = tid; j < 9 * nh; j += kBlock) inv[j] = invD32[9 * (size_t)s_h1[j / 9] + j % 9];
  unsigned short* bslot = reinterpret_cast<unsigned short*>(pk + kHaloOffBlk);
  for (int f = tid; f < nblk; f += kBlock) bslot[f] = (unsigned short)slot_of(s_col[f]);
  float* val = reinterpret_cast<float*>(pk + halo_val_off(hbl));
#pragma unroll 4
  for (int j = tid; j < 9 * nblk; j += kBlock) val[j] = val32[val_index(s_qf[j / 9], j % 9)];
  // set slot of the column of block u of each own-row lane (4 lanes per row);
  // a lane past its row's end gathers its own row (value 0 in the sum)
  {
    const int g = tid >> 2, l = tid & 3, row = r0 + g;
    unsigned short* oslot = reinterpret_cast<unsigned short*>(pk + kHaloOffSlot);
    const int q0 = row < r1 ? rowptr[row] : 0, q1 = row < r1 ? rowptr[row + 1] : 0;
    int cu[kHaloU];
#pragma unroll
    for (int u = 0; u < kHaloU; ++u) {
      const int p = q0 + l + 4 * u;
      cu[u] = p < q1 ? colidx[p] : -1;
    }
#pragma unroll
    for (int u = 0; u < kHaloU; ++u) oslot[u * kBlock + tid] = (unsigned short)(cu[u] >= 0 ? slot_of(cu[u]) : g);
  }
}

hipError_t launch_tile_pairs(const int* rowptr, const int* colidx, const float* val32, const float* invD32, int nb,
                             const int* tstart, int ntiles, int hbl, unsigned char* halo, hipStream_t st) {
  const int rtl = tstart ? ntiles : (nb + kHaloTileRows - 1) / kHaloTileRows;
  if (rtl <= 0 || rtl > kNbrMaxTiles || hbl < 0 || hbl > kHaloBlkMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned*>(halo),
                     (long long)(kHaloBaseBytes / 4));
  hipLaunchKernelGGL(k_tile_halo, dim3(rtl), dim3(kBlock), 0, st, rowptr, colidx, val32, invD32, nb, tstart, hbl,
                     halo, halo_pack_bytes(hbl));
  return hipGetLastError();
}

// power iteration step (partials of |z|^2 into part, one per tile)
int power_step_tiles(int G, int nb) { return (nb + (kBlock / G) - 1) / (kBlock / G); }
hipError_t launch_power_step(int G, const SysDesc& S, const double* x, const double* part_in, int np_in, double* z,
                             double* part, hipStream_t st) {
  const int tiles = power_step_tiles(G, S.nb);
  if (tiles == 0) return hipSuccess;
  switch (G) {
    case 4: hipLaunchKernelGGL(op_power_step<4>, dim3(tiles), dim3(kBlock), 0, st, S, x, part_in, np_in, z, part); break;
    case 8: hipLaunchKernelGGL(op_power_step<8>, dim3(tiles), dim3(kBlock), 0, st, S, x, part_in, np_in, z, part); break;
    case 16: hipLaunchKernelGGL(op_power_step<16>, dim3(tiles), dim3(kBlock), 0, st, S, x, part_in, np_in, z, part); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_sum_partials(const double* part, int np, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(64), 0, st, part, np, out);
  return hipGetLastError();
}

// deterministic pseudo-random start vector in [-1, 1) (power iteration)
__global__ void k_hash_fill(double* x, int n, unsigned seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 2654435761u ^ seed;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  x[i] = (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}

// power-iteration step on the device: out[it] = |z| (from nrm2 = |z|^2),
// x = z / |z|; a zero or non-finite norm leaves x untouched

hipError_t launch_hash_fill(double* x, int n, unsigned seed, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_hash_fill, dim3((n + 255) / 256), dim3(256), 0, st, x, n, seed);
  return hipGetLastError();
}

hipError_t launch_op_lowrank64(const SysDesc& S, const double* x, double* y, double* t, hipStream_t st) {
  if (S.k == 0 || S.a == 0) return hipSuccess;
  hipLaunchKernelGGL(op_lowrank64, dim3(1), dim3(kBlock), 0, st, S, x, y, t);
  return hipGetLastError();
}

}  // namespace hsc
