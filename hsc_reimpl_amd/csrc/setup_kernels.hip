// Setup kernels of the HSC solve on gfx950: split of the BCOO Hessian into
// D (full-storage BSR 3x3, SoA values), the dense ABD block A and the dense
// coupling E over the active FEM DOFs; FEM 3x3 diagonal inverses; ACA of E;
// A^-1; C = U^T A^-1 U = L L^T and W = V L; reduced right-hand side,
// back-substitution and acceptance checks.
// Reference semantics: linear_hsc.cu:61-91,195-304,349-435 and
// core.py:22-50,53-92,141-168 (reproduction/ of the reference).
#include <map>
#include <mutex>
#include "common.hpp"
#include "device_util.hpp"
#include "kernels.hpp"
#include <algorithm>
#include <cfloat>

namespace hsc {

enum SetupErr { E_FEM_DIAG = 0, E_UNCOVERED = 1, E_INTERBODY = 2, E_COUPLING_DOF = 3, E_NONFINITE = 4 };


__device__ __forceinline__ bool in_range(int v, int o, int n) { return v >= o && v < o + n; }

// Classify each stored block (linear_hsc.cu:61-81, :218-233): count the
// full-storage D entries per block row, invert FEM diagonal blocks, scatter
// ABD-ABD blocks into dense A (mirrored), flag active FEM nodes.
__global__ void k_classify(SplitParams S) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.count) return;
  const int R = S.rows[t] * 3, C = S.cols[t] * 3;
  const double* v = S.vals + 9 * t;
  const bool iq = in_range(R, S.qo, S.qm), jq = in_range(C, S.qo, S.qm);
  const bool iff = in_range(R, S.fo, S.fn), jf = in_range(C, S.fo, S.fn);
  if (!(iq || iff) || !(jq || jf)) {
    atomicAdd(&S.err[E_UNCOVERED], 1);
    return;
  }
  double sq = 0.0, B[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      B[r][c] = v[r + 3 * c];
      sq += B[r][c] * B[r][c];
    }
  if (!isfinite(sq)) atomicAdd(&S.err[E_NONFINITE], 1);
  if (iff && jf) {
    const int i = (R - S.fo) / 3, j = (C - S.fo) / 3;
    atomicAdd(&S.row_count[i], 1);
    if (i != j) {
      atomicAdd(&S.row_count[j], 1);
    } else {
      // closed-form 3x3 inverse (adjugate / determinant)
      const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
      const double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
      const double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
      const double det = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
      double inv[9];
      if (!(det > 0.0) || !isfinite(det)) {
        atomicAdd(&S.err[E_FEM_DIAG], 1);
        for (int e = 0; e < 9; ++e) inv[e] = 0.0;
      } else {
        const double id = 1.0 / det;
        inv[0] = c00 * id;
        inv[1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * id;
        inv[2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * id;
        inv[3] = c01 * id;
        inv[4] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id;
        inv[5] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id;
        inv[6] = c02 * id;
        inv[7] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * id;
        inv[8] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id;
      }
      for (int e = 0; e < 9; ++e) {
        S.invD64[9 * i + e] = inv[e];
        S.invD32[9 * i + e] = (float)inv[e];
      }
    }
  } else if (iq && jq) {
    const int a = R - S.qo, b = C - S.qo;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        atomicAdd(&S.A[(size_t)(a + r) * S.qm + b + c], B[r][c]);
        if (a != b) atomicAdd(&S.A[(size_t)(b + c) * S.qm + a + r], B[r][c]);
      }
    if (a / 12 != b / 12) atomicOr(&S.err[E_INTERBODY], 1);
  } else if (sq != 0.0) {
    const int f = iq ? C : R;
    if (!in_range(f, S.fo, S.fn)) {
      atomicAdd(&S.err[E_COUPLING_DOF], 1);
      return;
    }
    S.active_flag[(f - S.fo) / 3] = 1;
  }
}

// Single-workgroup exclusive scan: out[i] = sum_{j<i} in[i], out[n] = total.
// If mul != 1 the output is scaled (used to turn a node count into W rows).
__device__ void wg_scan_strided(const int* in, int* out, int n) {
  constexpr int kPer = 32;  // entries per thread held in registers (n <= 32768)
  __shared__ int s[1024];
  const int tid = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b0 = min(n, tid * per), b1 = min(n, b0 + per);
  const bool reg = n > 0 && per <= kPer;
  int v[kPer];
  int sum = 0;
  if (reg) {
    // all loads issued together (indices clamped, masked by a select)
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int x = in[min(b0 + j, n - 1)];
      v[j] = b0 + j < b1 ? x : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) sum += v[j];
  } else {
    for (int i = b0; i < b1; ++i) sum += in[i];
  }
  s[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  int run = tid ? s[tid - 1] : 0;
  if (reg) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (b0 + j < b1) out[b0 + j] = run;
      run += v[j];
    }
  } else {
    for (int i = b0; i < b1; ++i) {
      out[i] = run;
      run += in[i];
    }
  }
  if (tid == 1023) out[n] = s[1023];
}


// Exclusive scan of n ints by one workgroup of 1024 threads (out[n] = the
// total).  Each of the 16 waves owns a contiguous run of 64-entry chunks,
// loads them coalesced and all at once, scans each chunk across its lanes
// and carries the chunk totals; the waves' totals are then scanned in LDS.
// Runs longer than 32 chunks per wave (n > 32768) take the strided path.
__device__ void wg_scan(const int* in, int* out, int n) {
  constexpr int kWaves = 1024 / 64, kCh = 32;
  __shared__ int s_tot[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int seg = (((n + kWaves - 1) / kWaves) + 63) & ~63;
  const int nch = seg / 64;
  if (n <= 0 || nch > kCh) {
    wg_scan_strided(in, out, n);
    return;
  }
  const int base = w * seg;
  int v[kCh];
#pragma unroll
  for (int j = 0; j < kCh; ++j) {
    const int idx = base + 64 * j + lane;
    const int x = in[min(idx, n - 1)];
    v[j] = (j < nch && idx < n) ? x : 0;
  }
  int carry = 0;
#pragma unroll
  for (int j = 0; j < kCh; ++j) {
    if (j < nch) {
      int x = v[j];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
      }
      v[j] = carry + x - v[j];  // exclusive within the run
      carry += __shfl(x, 63, 64);
    }
  }
  if (lane == 0) s_tot[w] = carry;
  __syncthreads();
  int off = 0;
  for (int q = 0; q < w; ++q) off += s_tot[q];
#pragma unroll
  for (int j = 0; j < kCh; ++j) {
    const int idx = base + 64 * j + lane;
    if (j < nch && idx < n) out[idx] = off + v[j];
  }
  if (threadIdx.x == 0) {
    int t = 0;
    for (int q = 0; q < kWaves; ++q) t += s_tot[q];
    out[n] = t;
  }
}

// blockIdx.x scans (in0 -> out0) or (in1 -> out1), n entries each
__global__ __launch_bounds__(1024) void k_scan2(const int* in0, int* out0, const int* in1, int* out1, int n) {
  wg_scan(blockIdx.x ? in1 : in0, blockIdx.x ? out1 : out0, n);
}

// act_node[i] = 3*prefix (first W row of the node) or -1; act_dof[W row] = FEM DOF
__global__ void k_active(const int* flag, const int* prefix, int nb, int* act_node, int* act_dof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  if (flag[i]) {
    const int w = 3 * prefix[i];
    act_node[i] = w;
    act_dof[w] = 3 * i;
    act_dof[w + 1] = 3 * i + 1;
    act_dof[w + 2] = 3 * i + 2;
  } else {
    act_node[i] = -1;
  }
}


// Place FEM-FEM blocks into their BSR rows (slot order fixed later by the
// per-row sort) and scatter coupling blocks into dense E (linear_hsc.cu:237-247).
__global__ void k_place(PlaceParams S) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.count) return;
  const int R = S.rows[t] * 3, C = S.cols[t] * 3;
  const bool iq = in_range(R, S.qo, S.qm), jq = in_range(C, S.qo, S.qm);
  const bool iff = in_range(R, S.fo, S.fn), jf = in_range(C, S.fo, S.fn);
  if (iff && jf) {
    const int i = (R - S.fo) / 3, j = (C - S.fo) / 3;
    int pos = S.rowptr[i] + atomicAdd(&S.cursor[i], 1);
    S.key_col[pos] = j;
    S.key_src[pos] = (int)(2 * t);
    if (i != j) {
      pos = S.rowptr[j] + atomicAdd(&S.cursor[j], 1);
      S.key_col[pos] = i;
      S.key_src[pos] = (int)(2 * t + 1);
    }
  } else if (iq != jq && (iq ? jf : iff)) {
    const double* v = S.vals + 9 * t;
    const int f = iq ? C : R, q = (iq ? R : C) - S.qo;
    const int col = S.act_node[(f - S.fo) / 3];
    if (col < 0) return;  // zero block
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        // iq: block = E[q+r, col+c]; else block = E^T: E[q+c, col+r]
        const double bv = v[r + 3 * c];
        if (iq) atomicAdd(&S.E[(size_t)(q + r) * S.a + col + c], bv);
        else atomicAdd(&S.E[(size_t)(q + c) * S.a + col + r], bv);
      }
  }
}

// Sort each BSR row by (column, source) -- a deterministic slot order -- and
// gather its blocks into SoA (component c = r + 3*col of the 3x3), FP64 and
// FP32.  One wave per row: each lane ranks its entry against every entry of
// the row (keys exchanged by shuffles; (column, source) is unique within a
// row) and writes it to its sorted slot.
__global__ __launch_bounds__(256) void k_sort_gather(const int* __restrict__ rowptr, int nb,
                                                     const int* __restrict__ key_col, const int* __restrict__ key_src,
                                                     const double* __restrict__ vals, int nnzb, int* colidx,
                                                     double* v64, float* v32) {
  const int row = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
  if (row >= nb) return;  // (uniform over the wave)
  const int p0 = rowptr[row], len = rowptr[row + 1] - p0;
  auto key = [&](int e) { return ((long long)key_col[p0 + e] << 32) | (long long)(unsigned)key_src[p0 + e]; };
  for (int b0 = 0; b0 < len; b0 += 64) {
    const int e = b0 + lane;
    const long long mine = e < len ? key(e) : 0x7fffffffffffffffLL;
    int rank = 0;
    for (int c0 = 0; c0 < len; c0 += 64) {
      const long long kj = c0 + lane < len ? key(c0 + lane) : 0x7fffffffffffffffLL;
      const int cn = min(64, len - c0);
      for (int j = 0; j < cn; ++j) rank += __shfl(kj, j, 64) < mine;
    }
    if (e < len) {
      const int dst = p0 + rank;
      const int src = (int)(mine & 0xffffffffLL);
      const double* v = vals + 9 * (size_t)(src >> 1);
      const bool tr = src & 1;
      colidx[dst] = (int)(mine >> 32);
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          const double bv = tr ? v[c + 3 * r] : v[r + 3 * c];
          v64[val_index(dst, r + 3 * c)] = bv;
          v32[val_index(dst, r + 3 * c)] = (float)bv;
        }
    }
  }
}

// ---------------------------------------------------------------- ACA
// Full-pivot ACA (core.py:22-50; linear_hsc.cu:248-262) as one cooperative
// launch.  R is double-buffered so u = R[:,j], v = R[i,:]/piv are read from
// the previous iterate while the next is written; products and differences
// are rounded separately (no FMA) exactly like NumPy's R -= outer(u, v), so
// pivots agree with the NumPy method.  Pivot ties go to the lowest row-major
// index (np.argmax).
__device__ __forceinline__ void argmax_combine(double& m, long long& i, double m2, long long i2) {
  if (m2 > m || (m2 == m && i2 < i)) {
    m = m2;
    i = i2;
  }
}

__global__ __launch_bounds__(kBlock) void k_aca(AcaParams P) {
  __shared__ double s_m[kBlock / 64], s_q[kBlock / 64];
  __shared__ long long s_i[kBlock / 64];
  __shared__ double s_scale, s_res, s_piv;
  __shared__ long long s_idx;
  __shared__ int s_go;
  const int tid = threadIdx.x, b = blockIdx.x;
  const long long total = (long long)P.qm * P.a;
  const long long gthreads = (long long)P.nwg * kBlock;
  const long long g = (long long)b * kBlock + tid;
  if (P.skip && *P.skip) return;  // uniform: the factors came from the block-by-body ACA
  GridBarrier gbar;
  if (!gb_init(gbar, P.barrier, P.error, P.timeout_ticks, (unsigned)P.nwg)) return;
  // partials alternate between two halves of P.part: a workgroup rewrites a
  // half only two grid barriers after it was written, so every workgroup has
  // finished reading it (a single buffer lets a fast workgroup overwrite a
  // partial that a slower one is still reducing)
  int phase = 0;

  auto wg_reduce = [&](double m, long long idx, double q) {
    double* part = P.part + (size_t)(phase & 1) * 4 * P.nwg;
    for (int off = 32; off >= 1; off >>= 1) {
      double m2 = __shfl_xor(m, off, 64);
      long long i2 = __shfl_xor(idx, off, 64);
      argmax_combine(m, idx, m2, i2);
      q += __shfl_xor(q, off, 64);
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
      s_m[w] = m;
      s_i[w] = idx;
      s_q[w] = q;
    }
    __syncthreads();
    if (tid == 0) {
      double M = s_m[0], Q = s_q[0];
      long long I = s_i[0];
      for (int w2 = 1; w2 < kBlock / 64; ++w2) {
        argmax_combine(M, I, s_m[w2], s_i[w2]);
        Q += s_q[w2];
      }
      part[4 * b + 0] = M;
      part[4 * b + 1] = (double)I;
      part[4 * b + 2] = Q;
    }
  };
  // every workgroup combines all partials the same way (wave 0: lanes strided
  // over workgroups, then a butterfly), so all take the same pivot
  auto all_reduce = [&]() {
    const double* part = P.part + (size_t)(phase & 1) * 4 * P.nwg;
    ++phase;
    if (tid < 64) {
      double M = -1.0, Q = 0.0;
      long long I = 0x7fffffffffffffffLL;
      for (int w = tid; w < P.nwg; w += 64) {
        argmax_combine(M, I, part[4 * w + 0], (long long)part[4 * w + 1]);
        Q += part[4 * w + 2];
      }
      for (int off = 32; off >= 1; off >>= 1) {
        const double m2 = __shfl_xor(M, off, 64);
        const long long i2 = __shfl_xor(I, off, 64);
        argmax_combine(M, I, m2, i2);
        Q += __shfl_xor(Q, off, 64);
      }
      if (tid == 0) {
        s_idx = I;
        s_q[0] = Q;
      }
    }
    __syncthreads();
  };

  double m = -1.0, q = 0.0;
  long long mi = 0x7fffffffffffffffLL;
  for (long long e = g; e < total; e += gthreads) {
    const double v = P.E[e];
    P.R0[e] = v;
    argmax_combine(m, mi, fabs(v), e);
    q += v * v;
  }
  wg_reduce(m, mi, q);
  if (!gb_sync(gbar)) return;
  all_reduce();
  if (tid == 0) {
    s_scale = sqrt(s_q[0]);
    s_res = s_scale > 0.0 ? 1.0 : 0.0;
    if (b == 0) {
      P.hist[0] = s_res;
      *P.scale_out = s_scale;
    }
  }
  __syncthreads();
  const double scale = s_scale;
  const double guard = DBL_EPSILON * fmax(scale, DBL_MIN);
  double* cur = P.R0;
  double* nxt = P.R1;
  int kk = 0;
  for (;;) {
    if (tid == 0) {
      s_go = (kk < P.limit && s_res > P.stop && isfinite(scale)) ? 1 : 0;
      if (s_go) {
        s_piv = cur[s_idx];
        if (!(fabs(s_piv) > guard)) s_go = 0;
      }
    }
    __syncthreads();
    if (!s_go) break;
    const long long pi = s_idx / P.a, pj = s_idx % P.a;
    const double piv = s_piv;
    // factor columns
    for (long long r = g; r < P.qm; r += gthreads) P.U[r + (long long)P.qm * kk] = cur[r * P.a + pj];
    for (long long c = g; c < P.a; c += gthreads) P.V[c + (long long)P.a * kk] = cur[pi * P.a + c] / piv;
    if (b == 0 && tid == 0) {
      P.pivots[2 * kk] = (int)pi;
      P.pivots[2 * kk + 1] = (int)pj;
    }
    m = -1.0;
    q = 0.0;
    mi = 0x7fffffffffffffffLL;
    for (long long e = g; e < total; e += gthreads) {
      const long long r = e / P.a, c = e - r * P.a;
      const double u = cur[r * P.a + pj];
      const double v = cur[pi * P.a + c] / piv;
      const double val = __dsub_rn(cur[e], __dmul_rn(u, v));
      nxt[e] = val;
      argmax_combine(m, mi, fabs(val), e);
      q += val * val;
    }
    wg_reduce(m, mi, q);
    if (!gb_sync(gbar)) return;
    all_reduce();
    ++kk;
    if (tid == 0) {
      s_res = sqrt(s_q[0]) / scale;
      if (b == 0) P.hist[kk] = s_res;
    }
    __syncthreads();
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (b == 0 && tid == 0) *P.rank_out = kk;
}

// ---------------------------------------------------------------- ACA by ABD body
// When every active FEM column of E couples to one ABD body only (each tire
// rim to its own hub, as in the vehicle systems), E is block diagonal once
// its columns are grouped by body, and a rank-1 step of full-pivot ACA
// changes only the block that holds its pivot.  The global pivot sequence is
// then an interleaving of the blocks' own full-pivot sequences (the largest
// entry of a block is the same element whether it is sought in the block or
// in E, ties broken by the same row-major index), with the same pivot rows
// and columns and the same arithmetic per entry.  Each block is therefore
// factored by its own workgroup, with its residual in registers and no grid
// barrier, and the factors are concatenated body by body (one column order;
// V C V^T is the same).  The stop guard is the global eps x ||E||_F.  The
// path is taken when the ACA has no relative-residual stop, no column
// touches two bodies, the blocks cannot reach the rank limit together
// (sum over bodies of min(12, columns) <= limit) and every block fits in
// registers (columns <= kAcaBlkCpt x kBlock); otherwise k_aca or
// k_aca_reg factor the whole E (they exit at once when this path ran).
constexpr int kAcaBlkCpt = 3;                   // block columns per thread (12 rows each)
constexpr int kAcaBlkCols = kAcaBlkCpt * kBlock;  // columns per body

// body of column c (-1: no nonzero, -2: nonzeros in two bodies) and |E[:, c]|^2;
// the 12 rows of a body are loaded together
__global__ void k_col_bodies(AcaBlkParams P) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.a) return;
  int body = -1;
  double q = 0.0;
  for (int bb = 0; bb < P.qm / 12; ++bb) {
    const double* e = P.E + (size_t)(12 * bb) * P.a + c;
    double v[12];
#pragma unroll
    for (int r = 0; r < 12; ++r) v[r] = e[(size_t)r * P.a];
    bool nz = false;
#pragma unroll
    for (int r = 0; r < 12; ++r) {
      q += v[r] * v[r];
      nz = nz || v[r] != 0.0;
    }
    if (nz) body = (body == -1 || body == bb) ? bb : -2;
  }
  P.colbody[c] = body;
  P.colsq[c] = q;
}

// one workgroup: ||E||_F (fixed order), columns per body, and the decision
// plan[0] = 1 when the body path applies
__global__ __launch_bounds__(kBlock) void k_aca_plan(AcaBlkParams P) {
  __shared__ int s_cnt[kAcaBlkMaxBodies];
  __shared__ int s_multi;
  __shared__ double s_red[kBlock / 64];
  const int tid = threadIdx.x, nbody = P.qm / 12;
  for (int bb = tid; bb < kAcaBlkMaxBodies; bb += kBlock) s_cnt[bb] = 0;
  if (tid == 0) s_multi = 0;
  __syncthreads();
  double q = 0.0;
  for (int c = tid; c < P.a; c += kBlock) {
    q += P.colsq[c];
    const int bb = P.colbody[c];
    if (bb == -2) s_multi = 1;
    else if (bb >= 0 && bb < kAcaBlkMaxBodies) atomicAdd(&s_cnt[bb], 1);
  }
  const double tot = block_sum_t0(q, s_red);
  __syncthreads();
  if (tid == 0) {
    P.plan[1] = 0;
    int ok = P.stop == 0.0 && !s_multi && nbody <= kAcaBlkMaxBodies;
    long long kmax = 0;
    for (int bb = 0; bb < nbody && ok; ++bb) {
      if (s_cnt[bb] > kAcaBlkCols) ok = 0;
      kmax += s_cnt[bb] < 12 ? s_cnt[bb] : 12;
    }
    if (kmax > P.limit) ok = 0;
    P.plan[0] = ok;
    *P.scale = sqrt(tot);
  }
  for (int bb = tid; bb < nbody && bb < kAcaBlkMaxBodies; bb += kBlock) P.plan[4 + bb] = s_cnt[bb];
}

// one workgroup per body: full-pivot ACA of the body's 12 x n block.  Thread
// t owns the block's columns t, t + kBlock, ... (all 12 rows of each) in
// registers, so the pivot row's entries v_c = R[pil][c] / piv stay with
// their owners and only the pivot column (12 values) goes through LDS.  The
// block's entries keep the whole-E order: a column's rank among the body's
// columns follows its global index, so (row, local column) orders entries as
// the global row-major index does (the argmax tie-break).
__global__ __launch_bounds__(kBlock) void k_aca_blocks(AcaBlkParams P) {
  __shared__ int s_cols[kAcaBlkCols];
  __shared__ int s_wcnt[kBlock / 64];
  __shared__ double s_m[kBlock / 64], s_q[kBlock / 64];
  __shared__ int s_i[kBlock / 64];
  __shared__ double s_u[12];
  __shared__ double s_piv;
  __shared__ int s_n, s_best;
  __shared__ double s_bestv, s_sq;
  if (!P.plan[0]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int body = blockIdx.x, r0 = 12 * body, a = P.a;
  // the body's columns in ascending order
  if (tid == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < a; base += kBlock) {
    const int c = base + tid;
    const bool mine = c < a && P.colbody[c] == body;
    const unsigned long long m = __ballot(mine);
    if (lane == 0) s_wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = s_n;
    for (int w = 0; w < wave; ++w) off += s_wcnt[w];
    if (mine) s_cols[off + __popcll(m & ((1ull << lane) - 1ull))] = c;
    __syncthreads();
    if (tid == 0) {
      int t = s_n;
      for (int w = 0; w < kBlock / 64; ++w) t += s_wcnt[w];
      s_n = t;
    }
    __syncthreads();
  }
  const int n = s_n;
  constexpr int kNone = 0x7fffffff;
  double R[kAcaBlkCpt][12];
  int gcol[kAcaBlkCpt];
#pragma unroll
  for (int j = 0; j < kAcaBlkCpt; ++j) {
    const int c = tid + j * kBlock;
    gcol[j] = c < n ? s_cols[c] : -1;
#pragma unroll
    for (int r = 0; r < 12; ++r) R[j][r] = c < n ? P.E[(size_t)(r0 + r) * a + gcol[j]] : 0.0;
  }
  const double scale = *P.scale;
  const double guard = DBL_EPSILON * fmax(scale, DBL_MIN);
  // argmax |R| (lowest (row, column) on ties; key = row * 2048 + local column)
  // and |R|^2 of the block, left in s_bestv / s_best / s_sq
  auto reduce = [&]() {
    double m = -1.0, q = 0.0;
    int mi = kNone;
#pragma unroll
    for (int j = 0; j < kAcaBlkCpt; ++j) {
      const int c = tid + j * kBlock;
      if (c < n) {
#pragma unroll
        for (int r = 0; r < 12; ++r) {
          const double v = fabs(R[j][r]);
          const int key = (r << 11) | c;
          if (v > m || (v == m && key < mi)) {
            m = v;
            mi = key;
          }
          q += R[j][r] * R[j][r];
        }
      }
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const double m2 = __shfl_xor(m, off, 64);
      const int i2 = __shfl_xor(mi, off, 64);
      if (m2 > m || (m2 == m && i2 < mi)) {
        m = m2;
        mi = i2;
      }
      q += __shfl_xor(q, off, 64);
    }
    if (lane == 0) {
      s_m[wave] = m;
      s_i[wave] = mi;
      s_q[wave] = q;
    }
    __syncthreads();
    if (tid == 0) {
      double M = s_m[0], Q = s_q[0];
      int I = s_i[0];
      for (int w = 1; w < kBlock / 64; ++w) {
        if (s_m[w] > M || (s_m[w] == M && s_i[w] < I)) {
          M = s_m[w];
          I = s_i[w];
        }
        Q += s_q[w];
      }
      s_bestv = M;
      s_best = I;
      s_sq = Q;
    }
    __syncthreads();
  };
  reduce();
  if (tid == 0) P.hblk[body * 13] = s_sq;
  int k = 0;
  while (n > 0 && k < 12 && s_best != kNone) {
    const int bi = s_best, pil = bi >> 11, pjl = bi & 2047;
    // the pivot column's owner publishes it (and the pivot)
    if (pjl % kBlock == tid) {
      const int j = pjl / kBlock;
#pragma unroll
      for (int jj = 0; jj < kAcaBlkCpt; ++jj)
        if (jj == j) {
#pragma unroll
          for (int r = 0; r < 12; ++r) s_u[r] = R[jj][r];
        }
      s_piv = s_u[pil];
    }
    __syncthreads();
    const double piv = s_piv;
    if (!(fabs(piv) > guard)) break;  // uniform
    if (tid < 12) P.ublk[(size_t)body * 144 + k * 12 + tid] = s_u[tid];
    if (tid == 0) {
      P.pblk[body * 24 + 2 * k] = r0 + pil;
      P.pblk[body * 24 + 2 * k + 1] = s_cols[pjl];
    }
    double u[12];
#pragma unroll
    for (int r = 0; r < 12; ++r) u[r] = s_u[r];
#pragma unroll
    for (int j = 0; j < kAcaBlkCpt; ++j) {
      if (tid + j * kBlock < n) {
        double pr = R[j][0];
#pragma unroll
        for (int r = 1; r < 12; ++r) pr = r == pil ? R[j][r] : pr;
        const double v = pr / piv;
        P.vtmp[(size_t)gcol[j] * 12 + k] = v;
#pragma unroll
        for (int r = 0; r < 12; ++r) R[j][r] = __dsub_rn(R[j][r], __dmul_rn(u[r], v));
      }
    }
    ++k;
    __syncthreads();  // s_u / s_piv are rewritten by the next step
    reduce();
    if (tid == 0) P.hblk[body * 13 + k] = s_sq;
  }
  if (tid == 0) {
    P.rank_b[body] = k;
    // a 13th pivot above the guard (rounding left in a 12-row block): the
    // finish kernel hands the factorisation back to the whole-E ACA
    if (k == 12 && s_best != kNone && s_bestv > guard) P.plan[1] = 1;
  }
}

// the factors in k_aca's format: U, V concatenated body by body, pivots,
// the residual history in that order, rank and scale.  Per-body ranks and
// residuals are staged in LDS; one thread then forms the running sums over
// the bodies (the history's fixed order) and every pivot is written by its
// own thread.
__global__ __launch_bounds__(256) void k_aca_blocks_finish(AcaBlkParams P) {
  __shared__ int s_off[kAcaBlkMaxBodies + 1], s_rk[kAcaBlkMaxBodies];
  __shared__ double s_h0[kAcaBlkMaxBodies], s_hl[kAcaBlkMaxBodies];
  __shared__ double s_done[kAcaBlkMaxBodies], s_rest[kAcaBlkMaxBodies];
  if (!P.plan[0]) return;
  const int tid = threadIdx.x, nbody = P.qm / 12;
  if (P.plan[1]) {  // a block needed more than 12 steps: the whole-E ACA runs instead
    if (blockIdx.x == 0 && tid == 0) P.plan[0] = 0;
    return;
  }
  for (int bb = tid; bb < nbody; bb += blockDim.x) {
    const int rk = P.rank_b[bb];
    s_rk[bb] = rk;
    s_h0[bb] = P.hblk[bb * 13];
    s_hl[bb] = P.hblk[bb * 13 + rk];
  }
  __syncthreads();
  if (tid == 0) {
    int o = 0;
    double rest = 0.0, done = 0.0;
    for (int bb = 0; bb < nbody; ++bb) rest += s_h0[bb];
    for (int bb = 0; bb < nbody; ++bb) {
      s_off[bb] = o;
      o += s_rk[bb];
      rest -= s_h0[bb];
      s_done[bb] = done;
      s_rest[bb] = rest;
      done += s_hl[bb];
    }
    s_off[nbody] = o;
  }
  __syncthreads();
  const int K = s_off[nbody], qm = P.qm, a = P.a;
  const long long nu = (long long)qm * K, nv = (long long)a * K;
  for (long long e = (long long)blockIdx.x * blockDim.x + tid; e < nu + nv; e += (long long)gridDim.x * blockDim.x) {
    if (e < nu) {
      const int g = (int)(e / qm), r = (int)(e % qm), bb = r / 12;
      const int kk = g - s_off[bb];
      P.U[e] = (kk >= 0 && kk < s_rk[bb]) ? P.ublk[(size_t)bb * 144 + kk * 12 + r % 12] : 0.0;
    } else {
      const long long f = e - nu;
      const int g = (int)(f / a), c = (int)(f % a), bb = P.colbody[c];
      const int kk = bb >= 0 ? g - s_off[bb] : -1;
      P.V[f] = (bb >= 0 && kk >= 0 && kk < s_rk[bb]) ? P.vtmp[(size_t)c * 12 + kk] : 0.0;
    }
  }
  if (blockIdx.x == 0) {
    const double scale = *P.scale;
    for (int g = tid; g < K; g += blockDim.x) {
      int bb = 0;
      while (s_off[bb + 1] <= g) ++bb;
      const int kk = g - s_off[bb];
      P.pivots[2 * g] = P.pblk[bb * 24 + 2 * kk];
      P.pivots[2 * g + 1] = P.pblk[bb * 24 + 2 * kk + 1];
      const double tot = s_done[bb] + P.hblk[bb * 13 + kk + 1] + fmax(s_rest[bb], 0.0);
      P.hist[g + 1] = scale > 0.0 ? sqrt(fmax(tot, 0.0)) / scale : 0.0;
    }
    if (tid == 0) {
      P.hist[0] = scale > 0.0 ? 1.0 : 0.0;
      *P.rank_out = K;
      *P.scale_out = scale;
    }
  }
}

// ---------------------------------------------------------------- ACA, register-resident residual
// The same steps and arithmetic as k_aca with the residual held in
// registers: workgroup b owns the columns [b*cpw, (b+1)*cpw) of R (all qm
// rows), thread t the entries t, t + kBlock, ... of that block in (row,
// column) order.  With its argmax partial each workgroup publishes the column
// of its best entry, so after the one grid barrier of a step every workgroup
// reads the pivot column from the winner's slot and holds its own share of
// the pivot row: R is never re-read from memory.  Published columns and
// partials alternate between two slots, as in k_aca.
constexpr int kAcaRegPer = 16;     // residual entries per thread
constexpr int kAcaRegQm = 512;     // ABD DOFs (pivot column staged in LDS)
constexpr int kAcaRegCols = 1024;  // columns per workgroup (pivot-row share staged in LDS)

int aca_reg_grid(int qm, int a, int max_wg) {
  if (qm <= 0 || a <= 0 || qm > kAcaRegQm || max_wg <= 0) return 0;
  const long long total = (long long)qm * a;
  // about 12 entries per thread (measured: 12, 16 and 20 per thread take the
  // same time at 144 x 6,300; the step is bound by the grid barrier)
  int nwg = (int)std::min<long long>(max_wg, std::max<long long>(1, (total + 12LL * kBlock - 1) / (12LL * kBlock)));
  for (; nwg <= max_wg; ++nwg) {
    const int cpw = (a + nwg - 1) / nwg;
    if ((long long)qm * cpw <= (long long)kAcaRegPer * kBlock && cpw <= kAcaRegCols) return nwg;
  }
  return 0;
}

__global__ __launch_bounds__(kBlock) void k_aca_reg(AcaParams P) {
  __shared__ double s_m[kBlock / 64], s_q[kBlock / 64];
  __shared__ long long s_i[kBlock / 64];
  __shared__ double s_u[kAcaRegQm];
  __shared__ double s_v[kAcaRegCols];
  __shared__ double s_scale, s_res, s_piv;
  __shared__ long long s_idx, s_best;
  __shared__ int s_go;
  const int tid = threadIdx.x, b = blockIdx.x;
  const int qm = P.qm, a = P.a, nwg = P.nwg;
  const int cpw = (a + nwg - 1) / nwg;
  const int c0 = min(a, b * cpw), nc = min(a, c0 + cpw) - c0;
  const int cnt = qm * nc;
  if (P.skip && *P.skip) return;  // uniform: the factors came from the block-by-body ACA
  GridBarrier gbar;
  if (!gb_init(gbar, P.barrier, P.error, P.timeout_ticks, (unsigned)nwg)) return;
  int phase = 0;
  double R[kAcaRegPer];
  int rc[kAcaRegPer];  // (row << 10) | local column of each entry, -1: none
#pragma unroll
  for (int j = 0; j < kAcaRegPer; ++j) {
    const int le = tid + j * kBlock;
    rc[j] = -1;
    R[j] = 0.0;
    if (le < cnt) {
      const int r = le / nc, c = le - r * nc;
      rc[j] = (r << 10) | c;
      R[j] = P.E[(size_t)r * a + c0 + c];
    }
  }
  auto rr = [&](int j) { return rc[j] >> 10; };  // (-1 for no entry)
  auto cc = [&](int j) { return rc[j] & 1023; };

  // argmax |R| (lowest row-major index on ties) and sum of squares of the
  // owned entries; the partial into slot `phase`, then the column of the
  // workgroup's best entry into the same slot of the published columns
  auto reduce_publish = [&]() {
    double m = -1.0, q = 0.0;
    long long mi = 0x7fffffffffffffffLL;
#pragma unroll
    for (int j = 0; j < kAcaRegPer; ++j) {
      if (rc[j] >= 0) {
        argmax_combine(m, mi, fabs(R[j]), (long long)rr(j) * a + c0 + cc(j));
        q += R[j] * R[j];
      }
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const double m2 = __shfl_xor(m, off, 64);
      const long long i2 = __shfl_xor(mi, off, 64);
      argmax_combine(m, mi, m2, i2);
      q += __shfl_xor(q, off, 64);
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
      s_m[w] = m;
      s_i[w] = mi;
      s_q[w] = q;
    }
    __syncthreads();
    double* part = P.part + (size_t)(phase & 1) * 4 * nwg;
    if (tid == 0) {
      double M = s_m[0], Q = s_q[0];
      long long I = s_i[0];
      for (int w2 = 1; w2 < kBlock / 64; ++w2) {
        argmax_combine(M, I, s_m[w2], s_i[w2]);
        Q += s_q[w2];
      }
      part[4 * b + 0] = M;
      part[4 * b + 1] = (double)I;
      part[4 * b + 2] = Q;
      s_best = I;
    }
    __syncthreads();
    const long long I = s_best;
    if (nc > 0 && I != 0x7fffffffffffffffLL) {
      const int bc = (int)(I % a) - c0;
      double* cand = P.cand + ((size_t)(phase & 1) * nwg + b) * qm;
#pragma unroll
      for (int j = 0; j < kAcaRegPer; ++j)
        if (rc[j] >= 0 && cc(j) == bc) cand[rr(j)] = R[j];
    }
  };
  // every workgroup combines all partials the same way (as in k_aca)
  auto all_reduce = [&]() {
    const double* part = P.part + (size_t)(phase & 1) * 4 * nwg;
    ++phase;
    if (tid < 64) {
      double M = -1.0, Q = 0.0;
      long long I = 0x7fffffffffffffffLL;
      for (int w = tid; w < nwg; w += 64) {
        argmax_combine(M, I, part[4 * w + 0], (long long)part[4 * w + 1]);
        Q += part[4 * w + 2];
      }
      for (int off = 32; off >= 1; off >>= 1) {
        const double m2 = __shfl_xor(M, off, 64);
        const long long i2 = __shfl_xor(I, off, 64);
        argmax_combine(M, I, m2, i2);
        Q += __shfl_xor(Q, off, 64);
      }
      if (tid == 0) {
        s_idx = I;
        s_q[0] = Q;
      }
    }
    __syncthreads();
  };

  reduce_publish();
  if (!gb_sync(gbar)) return;
  int slot = phase & 1;  // slot of the columns just published
  all_reduce();
  if (tid == 0) {
    s_scale = sqrt(s_q[0]);
    s_res = s_scale > 0.0 ? 1.0 : 0.0;
    if (b == 0) {
      P.hist[0] = s_res;
      *P.scale_out = s_scale;
    }
  }
  __syncthreads();
  const double scale = s_scale;
  const double guard = DBL_EPSILON * fmax(scale, DBL_MIN);
  int kk = 0;
  for (;;) {
    const long long I = s_idx;
    const bool have = I >= 0 && I < (long long)qm * a;
    const int pi = have ? (int)(I / a) : 0, pj = have ? (int)(I % a) : 0;
    if (have) {
      const double* cand = P.cand + ((size_t)slot * nwg + pj / cpw) * qm;  // the pivot column's owner
      for (int r = tid; r < qm; r += kBlock) s_u[r] = cand[r];
    }
    __syncthreads();
    if (tid == 0) {
      s_go = (have && kk < P.limit && s_res > P.stop && isfinite(scale)) ? 1 : 0;
      if (s_go) {
        s_piv = s_u[pi];
        if (!(fabs(s_piv) > guard)) s_go = 0;
      }
    }
    __syncthreads();
    if (!s_go) break;
    const double piv = s_piv;
    // own share of the pivot row, v = R[pi, c] / piv
#pragma unroll
    for (int j = 0; j < kAcaRegPer; ++j)
      if (rr(j) == pi) s_v[cc(j)] = R[j] / piv;
    __syncthreads();
    // factor columns
    if (b == 0)
      for (int r = tid; r < qm; r += kBlock) P.U[r + (size_t)qm * kk] = s_u[r];
    for (int cl = tid; cl < nc; cl += kBlock) P.V[c0 + cl + (size_t)a * kk] = s_v[cl];
    if (b == 0 && tid == 0) {
      P.pivots[2 * kk] = pi;
      P.pivots[2 * kk + 1] = pj;
    }
#pragma unroll
    for (int j = 0; j < kAcaRegPer; ++j)
      if (rc[j] >= 0) R[j] = __dsub_rn(R[j], __dmul_rn(s_u[rr(j)], s_v[cc(j)]));
    reduce_publish();
    if (!gb_sync(gbar)) return;
    slot = phase & 1;
    all_reduce();
    ++kk;
    if (tid == 0) {
      s_res = sqrt(s_q[0]) / scale;
      if (b == 0) P.hist[kk] = s_res;
    }
    __syncthreads();
  }
  if (b == 0 && tid == 0) *P.rank_out = kk;
}

// ---------------------------------------------------------------- small dense LA
// Per-body 12x12 Cholesky + inverse of the diagonal ABD blocks
// (core.py:53-71 BlockInverse; linear_hsc.cu:264-271).  One wave per body.
__global__ __launch_bounds__(64) void k_body_inverse(const double* A, int qm, double* Binv, int* err) {
  __shared__ double L[12][13];
  __shared__ int ok;
  const int body = blockIdx.x, tid = threadIdx.x, o = 12 * body;
  for (int e = tid; e < 144; e += blockDim.x) L[e / 12][e % 12] = A[(size_t)(o + e / 12) * qm + o + e % 12];
  __syncthreads();
  if (tid == 0) {
    ok = 1;
    for (int j = 0; j < 12 && ok; ++j) {
      double d = L[j][j];
      for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
      if (!(d > 0.0) || !isfinite(d)) {
        ok = 0;
        break;
      }
      d = sqrt(d);
      L[j][j] = d;
      for (int i = j + 1; i < 12; ++i) {
        double s = L[i][j];
        for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
        L[i][j] = s / d;
      }
    }
    if (!ok) atomicAdd(err, 1);
  }
  __syncthreads();
  if (!ok) return;
  if (tid < 12) {  // column tid of (L L^T)^-1 (loops unrolled: y stays in registers)
    double y[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      double s = (i == tid) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
      y[i] = s / L[i][i];
    }
#pragma unroll
    for (int i = 11; i >= 0; --i) {
      double s = y[i];
#pragma unroll
      for (int k = i + 1; k < 12; ++k) s -= L[k][i] * y[k];
      y[i] = s / L[i][i];
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) Binv[(size_t)body * 144 + i * 12 + tid] = y[i];
  }
}

// In-place Cholesky of an n x n row-major SPD matrix by one workgroup
// (lower triangle = L; blockDim a multiple of 32).  Returns false
// (uniformly) when not SPD / not finite.  Right-looking, one barrier per
// column: step j reads the pivot and column j as step j - 1 left them, every
// thread forms l_ij = a_ij / d_j itself for the trailing update
// a_ic -= l_ij l_cj (j < c <= i, threads on a 32-wide grid of (i, c)), and
// column j is scaled in place during step j + 1, when nothing reads it.
// The values and operations are those of scaling the column first, so L is
// the same bit for bit.
__device__ bool wg_cholesky(double* M, int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tx = tid & 31, ty = tid >> 5, ny = nt >> 5;
  double dprev = 0.0;
  for (int j = 0; j < n; ++j) {
    const double a = M[(size_t)j * n + j];  // every thread reads the same value
    if (!(a > 0.0) || !isfinite(a)) return false;
    const double d = sqrt(a);
    if (j > 0) {  // column j - 1 is final
      for (int i = j + tid; i < n; i += nt) M[(size_t)i * n + j - 1] /= dprev;
      if (tid == 0) M[(size_t)(j - 1) * n + j - 1] = dprev;
    }
    const int m = n - j - 1;
    for (int r = ty; r < m; r += ny) {
      const int i = j + 1 + r;
      const double li = M[(size_t)i * n + j] / d;
      for (int s = tx; s <= r; s += 32) {
        const int c = j + 1 + s;
        M[(size_t)i * n + c] -= li * (M[(size_t)c * n + j] / d);
      }
    }
    dprev = d;
    __syncthreads();
  }
  if (n > 0 && tid == 0) M[(size_t)(n - 1) * n + n - 1] = dprev;
  __syncthreads();
  return true;
}

// Inv = (L L^T)^-1 from the lower factor in M (n x n); column per thread.
__device__ void wg_chol_inverse(const double* M, int n, double* Inv, double* scratch) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int col = tid; col < n; col += nt) {
    double* y = scratch + (size_t)col * n;
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= M[(size_t)i * n + k] * y[k];
      y[i] = s / M[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= M[(size_t)k * n + i] * y[k];
      y[i] = s / M[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) Inv[(size_t)i * n + col] = y[i];
  }
  __syncthreads();
}

// out = X * Y (n x n row-major), one workgroup
__device__ void wg_gemm(const double* X, const double* Y, double* out, int n) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int m = 0; m < n; ++m) s += X[(size_t)i * n + m] * Y[(size_t)m * n + j];
    out[e] = s;
  }
  __syncthreads();
}

// block-diagonal B^-1 (from Binv) times X, n x n
__device__ void wg_binv_mul(const double* Binv, const double* X, double* out, int n, double scale) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n, body = i / 12, o = 12 * body;
    double s = 0.0;
    for (int m = 0; m < 12; ++m) s += Binv[(size_t)body * 144 + (i - o) * 12 + m] * X[(size_t)(o + m) * n + j];
    out[e] = scale * s;
  }
  __syncthreads();
}

// residual ||I - A P||_F / sqrt(n); Rm = I - A P
__device__ double wg_inverse_residual(const double* A, const double* P, double* Rm, int n, double* red) {
  double acc = 0.0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    double s = (i == j) ? 1.0 : 0.0;
    for (int m = 0; m < n; ++m) s -= A[(size_t)i * n + m] * P[(size_t)m * n + j];
    Rm[e] = s;
    acc += s * s;
  }
  __syncthreads();
  const double tot = block_sum_t0(acc, red);
  __shared__ double s_tot;
  if (threadIdx.x == 0) s_tot = tot;
  __syncthreads();
  return sqrt(s_tot) / sqrt((double)n);
}

__device__ void wg_symmetrize(double* P, int n) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    if (j > i) {
      const double s = 0.5 * (P[(size_t)i * n + j] + P[(size_t)j * n + i]);
      P[(size_t)i * n + j] = s;
      P[(size_t)j * n + i] = s;
    }
  }
  __syncthreads();
}


// The block Neumann polynomial and its doubling when A has no inter-body
// blocks: every matrix of the recurrence (B^-1 A, the terms, P, I - A P) is
// then block diagonal, so each product is one 12 x 12 product per body.  The
// blocks are kept compact, entry (i, jj) of block row i at i * 12 + jj; the
// sums run over the block's 12 columns in the order the dense product takes
// them (its other terms are exact zeros).  The doubling stops on the norm of
// the whole I - A P, as in the dense form.  Returns the residual.
__device__ double wg_abd_neumann_blockdiag(const AbdParams& S, int& doublings, double* red) {
  const int n = S.qm, ne = n * 12;
  double* term = S.s1;  // compact
  double* tmp = S.s2;   // compact
  double* Pc = S.s3;    // compact
  double* R = S.P;      // compact I - A P while iterating; P written dense at the end
  auto blk = [&](const double* M, int i, int jj) {  // dense M's entry (i, o + jj), o = block start
    return M[(size_t)i * n + 12 * (i / 12) + jj];
  };
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const double v = S.omega * S.Binv[e];  // Binv is [body][12][12] = compact
    term[e] = v;
    Pc[e] = v;
  }
  __syncthreads();
  for (int k = 1; k < S.order; ++k) {
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {  // tmp = A term
      const int i = e / 12, jj = e % 12, o = 12 * (i / 12);
      double s = 0.0;
      for (int m = 0; m < 12; ++m) s += blk(S.A, i, m) * term[(size_t)(o + m) * 12 + jj];
      tmp[e] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {  // term -= w B^-1 tmp; P += term
      const int i = e / 12, jj = e % 12, o = 12 * (i / 12);
      double s = 0.0;
      for (int m = 0; m < 12; ++m) s += S.Binv[(size_t)i * 12 + m] * tmp[(size_t)(o + m) * 12 + jj];
      const double t = term[e] - S.omega * s;
      term[e] = t;
      Pc[e] += t;
    }
    __syncthreads();
  }
  auto symmetrize = [&]() {
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const int i = e / 12, jj = e % 12, o = 12 * (i / 12), ii = i - o;
      if (jj > ii) {
        const size_t t = (size_t)(o + jj) * 12 + ii;
        const double s = 0.5 * (Pc[e] + Pc[t]);
        Pc[e] = s;
        Pc[t] = s;
      }
    }
    __syncthreads();
  };
  __shared__ double s_res;
  auto residual = [&]() {  // R = I - A P, its norm / sqrt(n)
    double acc = 0.0;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const int i = e / 12, jj = e % 12, o = 12 * (i / 12);
      double s = (i - o == jj) ? 1.0 : 0.0;
      for (int m = 0; m < 12; ++m) s -= blk(S.A, i, m) * Pc[(size_t)(o + m) * 12 + jj];
      R[e] = s;
      acc += s * s;
    }
    __syncthreads();
    const double tot = block_sum_t0(acc, red);
    if (threadIdx.x == 0) s_res = sqrt(tot) / sqrt((double)n);
    __syncthreads();
    return s_res;
  };
  symmetrize();
  double res = residual();
  if (S.mode == 2) {
    while (res > S.tol && doublings < 48) {
      for (int e = threadIdx.x; e < ne; e += blockDim.x) {  // tmp = P (I - A P)
        const int i = e / 12, jj = e % 12, o = 12 * (i / 12);
        double s = 0.0;
        for (int m = 0; m < 12; ++m) s += Pc[(size_t)i * 12 + m] * R[(size_t)(o + m) * 12 + jj];
        tmp[e] = s;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < ne; e += blockDim.x) Pc[e] += tmp[e];
      __syncthreads();
      symmetrize();
      res = residual();
      ++doublings;
      if (!isfinite(res)) break;
    }
  }
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    S.P[e] = (i / 12 == j / 12) ? Pc[(size_t)i * 12 + (j % 12)] : 0.0;
  }
  __syncthreads();
  return res;
}

// A^-1 approximation P (one workgroup): exact Cholesky inverse, or the block
// Neumann polynomial P_r = sum_{k<r} (I - w B^-1 A)^k w B^-1 symmetrised, then
// optional doubling P <- P + P (I - A P) until ||I-AP||_F/sqrt(qm) <= tol
// (linear_hsc.cu:263-297; core.py:74-92,149-150).
__global__ __launch_bounds__(1024) void k_abd_inverse(AbdParams S) {
  __shared__ double red[1024 / 64];
  const int n = S.qm, nn = n * n;
  int status = 0, doublings = 0;
  if (S.mode != 0 && S.blockdiag) {
    const double res = wg_abd_neumann_blockdiag(S, doublings, red);
    if (S.mode == 2 && !(res <= S.tol)) status = 5;
    if (!isfinite(res)) status = status ? status : 6;
    if (threadIdx.x == 0) {
      S.out[0] = res;
      S.out[1] = (double)doublings;
      S.out[2] = (double)status;
    }
    return;
  }
  if (S.mode == 0) {
    if (S.blockdiag) {
      for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        S.P[e] = (i / 12 == j / 12) ? S.Binv[(size_t)(i / 12) * 144 + (i % 12) * 12 + (j % 12)] : 0.0;
      }
      __syncthreads();
    } else {
      for (int e = threadIdx.x; e < nn; e += blockDim.x) S.s1[e] = S.A[e];
      __syncthreads();
      if (!wg_cholesky(S.s1, n)) status = 4;
      else wg_chol_inverse(S.s1, n, S.P, S.s2);
    }
  } else {
    // term = w B^-1 ; P = term
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
      const int i = e / n, j = e % n;
      const double v = (i / 12 == j / 12) ? S.omega * S.Binv[(size_t)(i / 12) * 144 + (i % 12) * 12 + (j % 12)] : 0.0;
      S.s1[e] = v;
      S.P[e] = v;
    }
    __syncthreads();
    for (int k = 1; k < S.order; ++k) {
      wg_gemm(S.A, S.s1, S.s2, n);                  // A term
      wg_binv_mul(S.Binv, S.s2, S.s3, n, S.omega);  // w B^-1 A term
      for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        S.s1[e] -= S.s3[e];
        S.P[e] += S.s1[e];
      }
      __syncthreads();
    }
    wg_symmetrize(S.P, n);
  }
  double res = 0.0;
  if (status == 0 && S.mode == 0 && S.blockdiag) {
    // both A and P are block diagonal: only the 12x12 diagonal blocks of I - A P are nonzero
    double acc = 0.0;
    for (int e = threadIdx.x; e < n * 12; e += blockDim.x) {
      const int i = e / 12, jj = e % 12, o = 12 * (i / 12), j = o + jj;
      double v = (i == j) ? 1.0 : 0.0;
      for (int m = 0; m < 12; ++m) v -= S.A[(size_t)i * n + o + m] * S.P[(size_t)(o + m) * n + j];
      acc += v * v;
    }
    const double tot = block_sum_t0(acc, red);
    __shared__ double s_tot2;
    if (threadIdx.x == 0) s_tot2 = tot;
    __syncthreads();
    res = sqrt(s_tot2) / sqrt((double)n);
    if (!isfinite(res)) status = 6;
  } else if (status == 0) {
    res = wg_inverse_residual(S.A, S.P, S.s3, n, red);
    if (S.mode == 2) {
      while (res > S.tol && doublings < 48) {
        wg_gemm(S.P, S.s3, S.s2, n);  // P (I - A P)
        for (int e = threadIdx.x; e < nn; e += blockDim.x) S.P[e] += S.s2[e];
        __syncthreads();
        wg_symmetrize(S.P, n);
        res = wg_inverse_residual(S.A, S.P, S.s3, n, red);
        ++doublings;
        if (!isfinite(res)) break;
      }
      if (!isfinite(res) || res > S.tol) status = 5;
    }
    if (!isfinite(res)) status = status ? status : 6;
  }
  if (threadIdx.x == 0) {
    S.out[0] = res;
    S.out[1] = (double)doublings;
    S.out[2] = (double)status;
  }
}


// C[M x N] = op(A) B for small dense FP64 blocks: 16x16 output tiles per
// workgroup staged through LDS.  op(A) = A (row-major, lda) or A^T when
// transA (A stored column-major with leading dim lda, i.e. A^T row-major).
// B is addressed as B[k*ldb_k + n*ldb_n] so both layouts are covered.
__device__ __forceinline__ void gemm_small_tile(const double* A, int lda, int transA, const double* B, int ldb_k,
                                                int ldb_n, double* C, int ldc, int M, int N, int K, double* Ct) {
  __shared__ double As[16][17], Bs[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  double acc = 0.0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int ka = k0 + tx, kb = k0 + ty;
    const int ar = blockIdx.y * 16 + ty;
    As[ty][tx] = (ar < M && ka < K) ? (transA ? A[(size_t)ka * lda + ar] : A[(size_t)ar * lda + ka]) : 0.0;
    Bs[ty][tx] = (kb < K && col < N) ? B[(size_t)kb * ldb_k + (size_t)col * ldb_n] : 0.0;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc += As[ty][kk] * Bs[kk][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    C[(size_t)row * ldc + col] = acc;
    if (Ct) Ct[(size_t)col * M + row] = acc;  // C^T (N x M row-major) too
  }
}

__global__ __launch_bounds__(256) void k_gemm_small(const double* A, int lda, int transA, const double* B,
                                                    int ldb_k, int ldb_n, double* C, int ldc, int M, int N, int K,
                                                    double* Ct) {
  gemm_small_tile(A, lda, transA, B, ldb_k, ldb_n, C, ldc, M, N, K, Ct);
}

// the rank ACA found, read on the device (the coupling kernels are queued
// before the host has read it back)
__device__ __forceinline__ int factor_rank(const FactorParams& S) { return min(max(*S.kdev, 0), S.k); }

// W1 = P U (qm x k row-major) and W1T = W1^T (k x qm); tiles past the rank exit
__global__ __launch_bounds__(256) void k_coupling_PU(FactorParams S) {
  const int k = factor_rank(S), n = S.qm;
  if ((int)blockIdx.x * 16 >= k) return;  // uniform
  gemm_small_tile(S.P, n, 0, S.U, 1, n, S.W1, k, n, k, n, S.W1T);
}
// C = U^T W1 (k x k row-major; U column-major read as row-major U^T)
__global__ __launch_bounds__(256) void k_coupling_UtW(FactorParams S) {
  const int k = factor_rank(S), n = S.qm;
  if ((int)blockIdx.x * 16 >= k || (int)blockIdx.y * 16 >= k) return;  // uniform
  gemm_small_tile(S.U, n, 0, S.W1, k, 1, S.C, k, k, k, n, nullptr);
}

constexpr int kLdsChol = 80;  // C of rank <= 80 is factored in LDS (50 KB)
// C = U^T P U symmetrised (linear_hsc.cu:298-300), copied to Lm; 6 when an
// entry is not finite, else 0.
__device__ int coupling_prelude(const FactorParams& S) {
  const int k = factor_rank(S);
  __syncthreads();
  wg_symmetrize(S.C, k);
  int bad = 0;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    S.Lm[e] = S.C[e];
    if (!isfinite(S.C[e])) bad = 1;
  }
  __shared__ int s_bad;
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  if (bad) atomicOr(&s_bad, 1);
  __syncthreads();
  return s_bad ? 6 : 0;
}

// C = U^T P U symmetrised and its Cholesky C = L L^T.  (The products
// W1 = P U and C = U^T W1 are done by k_gemm_small before this.)
__global__ __launch_bounds__(1024) void k_coupling_core(FactorParams S) {
  const int k = factor_rank(S);
  int status = coupling_prelude(S);
  if (!status) {
    if (k <= kLdsChol) {
      __shared__ double Ls[kLdsChol * kLdsChol];
      for (int e = threadIdx.x; e < k * k; e += blockDim.x) Ls[e] = S.Lm[e];
      __syncthreads();
      if (!wg_cholesky(Ls, k)) status = 6;
      for (int e = threadIdx.x; e < k * k; e += blockDim.x) S.Lm[e] = Ls[e];
      __syncthreads();
    } else if (!wg_cholesky(S.Lm, k)) {
      status = 6;
    }
  }
  if (threadIdx.x == 0) S.out[0] = (double)status;
}

// W = V L (a x k row-major, FP64 and FP32): V C V^T = W W^T.  Each entry
// W[j][l] = sum over m = l .. k-1, in increasing order, of V[j][m] L[m][l].
// k <= 64: a workgroup takes 16 rows of V (column-major, loaded coalesced)
// and L into LDS and writes its 16 x k block of W row-major (coalesced), so
// that a few thousand coupled DOFs spread over a few hundred workgroups;
// larger ranks: one thread per entry.
constexpr int kMakeWRows = 16, kMakeWMaxK = 64;
__global__ __launch_bounds__(256) void k_make_W(const double* V, const double* Lm, int a, const int* kdev, int kmax,
                                                double* W64, float* W32) {
  __shared__ double s_V[kMakeWMaxK][kMakeWRows + 1];
  __shared__ double s_L[kMakeWMaxK * kMakeWMaxK];
  const int k = min(max(*kdev, 0), kmax);
  const int j0 = blockIdx.x * kMakeWRows, rows = min(kMakeWRows, a - j0);
  if (k == 0) return;
  if (k <= kMakeWMaxK) {
    for (int e = threadIdx.x; e < k * kMakeWRows; e += blockDim.x) {
      const int m = e / kMakeWRows, jj = e % kMakeWRows;
      s_V[m][jj] = jj < rows ? V[j0 + jj + (size_t)a * m] : 0.0;
    }
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) s_L[e] = Lm[e];
    __syncthreads();
    for (int o = threadIdx.x; o < rows * k; o += blockDim.x) {
      const int jj = o / k, l = o % k;
      double s = 0.0;
#pragma unroll 4
      for (int m = l; m < k; ++m) s += s_V[m][jj] * s_L[m * k + l];
      W64[(size_t)j0 * k + o] = s;
      W32[(size_t)j0 * k + o] = (float)s;
    }
  } else {
    for (int o = threadIdx.x; o < rows * k; o += blockDim.x) {
      const int j = j0 + o / k, l = o % k;
      double s = 0.0;
#pragma unroll 8
      for (int m = l; m < k; ++m) s += V[j + (size_t)a * m] * Lm[(size_t)m * k + l];
      W64[(size_t)j0 * k + o] = s;
      W32[(size_t)j0 * k + o] = (float)s;
    }
  }
}

// y[0:rows) = M x with M[i][m] at M[i * ld + m] (a row-major matrix, or the
// transpose of a column-major one): one wave per row, lanes over the inner
// index (coalesced), the loads of RB rows in flight together; lane-strided
// partial sums and a butterfly, so the order is fixed.  Called by a whole
// workgroup.
template <int RB>
__device__ __forceinline__ void wave_matvec(const double* __restrict__ M, int ld, int rows, int inner,
                                            const double* x, double* y) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int i0 = wave * RB; i0 < rows; i0 += nw * RB) {
    double acc[RB];
    const double* Mr[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      acc[u] = 0.0;
      Mr[u] = M + (size_t)min(i0 + u, rows - 1) * ld;  // (rows past the end: loaded, not stored)
    }
    for (int m = lane; m < inner; m += 64) {
      const double xm = x[m];
#pragma unroll
      for (int u = 0; u < RB; ++u) acc[u] += Mr[u][m] * xm;
    }
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const double t = wave_sum(acc[u]);
      if (lane == 0 && i0 + u < rows) y[i0 + u] = t;
    }
  }
}

// c = U^T P b_q = (P U)^T b_q  (the coupling part of the reduced
// right-hand side, core.py:163-164; linear_hsc.cu:351-357), with (P U)^T
// kept from setup (W1T, k x qm row-major).  One workgroup: b_q in LDS, one
// wave per output entry.
__global__ __launch_bounds__(512) void k_rhs_small(const double* W1T, const double* b, int qo, int qm, int k,
                                                   double* y) {
  constexpr int kMaxQm = 1536;  // ABD DOFs staged in LDS (128 bodies)
  __shared__ double s_b[kMaxQm];
  const bool lds_b = qm <= kMaxQm;
  if (lds_b)
    for (int i = threadIdx.x; i < qm; i += blockDim.x) s_b[i] = b[qo + i];
  __syncthreads();
  wave_matvec<4>(W1T, qm, k, qm, lds_b ? s_b : b + qo, y);
}

// rhs = b_x - V c (V: a x k column-major, row j = the j-th coupled FEM DOF);
// r = rhs; x_fem = 0
__global__ void k_rhs_vec(const double* b, int fo, int n, const int* act_node, const double* V, int a, int k,
                          const double* c, double* rhs, double* r, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = b[fo + i];
  const int an = act_node[i / 3];
  if (an >= 0 && k > 0) {
    const double* vr = V + an + i % 3;
    double g = 0.0;
#pragma unroll 8
    for (int l = 0; l < k; ++l) g += vr[(size_t)a * l] * c[l];
    v -= g;
  }
  rhs[i] = v;
  r[i] = v;
  x[fo + i] = 0.0;
}

// dq = P (b_q - U V^T dx) written to x[qo:qo+qm]  (core.py:167; linear_hsc.cu:395-399)
// k_backsub_vt: one workgroup per rank column, vt[l] = V[:,l] . dx[active]
__global__ __launch_bounds__(256) void k_backsub_vt(const double* V, const int* act_dof, const double* x, int fo,
                                                    int a, double* vt) {
  __shared__ double red[256 / 64];
  const int l = blockIdx.x;
  double s = 0.0;
#pragma unroll 4
  for (int j = threadIdx.x; j < a; j += blockDim.x) s += V[j + (size_t)a * l] * x[fo + act_dof[j]];
  const double tot = block_sum_t0(s, red);
  if (threadIdx.x == 0) vt[l] = tot;
}
__global__ __launch_bounds__(512) void k_backsub(const double* P, const double* U, const double* b, double* x,
                                                 int qo, int qm, int k, double* tmp) {
  __shared__ double s_vt[kMaxRankLds];
  for (int l = threadIdx.x; l < k; l += blockDim.x) s_vt[l] = tmp[l];
  __syncthreads();
  for (int i = threadIdx.x; i < qm; i += blockDim.x) {
    double s = b[qo + i];
#pragma unroll 8
    for (int l = 0; l < k; ++l) s -= U[i + (size_t)qm * l] * s_vt[l];
    tmp[k + i] = s;
  }
  __syncthreads();
  wave_matvec<6>(P, qm, qm, qm, tmp + k, x + qo);  // P (b_q - U V^T dx)
}

// full H x from the split parts: y_fem = D x_f (precomputed in yf) + E^T x_q,
// y_abd = E x_f[act] + A x_q; then partial sums of (b - Hx)^2, b^2, b.x
__global__ void k_full_residual_fem(const double* yf, const double* E, const int* act_node, int a, int qm,
                                    const double* b, const double* x, int fo, int qo, int n, double* part) {
  __shared__ double red[kBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, bb = 0.0, bx = 0.0;
  if (i < n) {
    double hx = yf[i];
    const int an = act_node[i / 3];
    if (an >= 0) {
      const int col = an + i % 3;
#pragma unroll 8
      for (int q = 0; q < qm; ++q) hx += E[(size_t)q * a + col] * x[qo + q];
    }
    const double bi = b[fo + i];
    rr = (bi - hx) * (bi - hx);
    bb = bi * bi;
    bx = bi * x[fo + i];
  }
  const double s0 = block_sum_t0(rr, red);
  const double s1 = block_sum_t0(bb, red);
  const double s2 = block_sum_t0(bx, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s0;
    part[3 * blockIdx.x + 1] = s1;
    part[3 * blockIdx.x + 2] = s2;
  }
}

// ABD rows of b - H x, one workgroup per row: partial (r^2, b^2, b.x)
__global__ __launch_bounds__(256) void k_full_residual_abd(const double* E, const double* A, const int* act_dof,
                                                           int a, int qm, const double* b, const double* x, int fo,
                                                           int qo, double* part) {
  __shared__ double red[256 / 64];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < a; j += blockDim.x) s += E[(size_t)i * a + j] * x[fo + act_dof[j]];
  for (int m = threadIdx.x; m < qm; m += blockDim.x) s += A[(size_t)i * qm + m] * x[qo + m];
  const double hx = block_sum_t0(s, red);
  if (threadIdx.x == 0) {
    const double bi = b[qo + i];
    part[3 * i] = (bi - hx) * (bi - hx);
    part[3 * i + 1] = bi * bi;
    part[3 * i + 2] = bi * x[qo + i];
  }
}
// fixed-order sums of the partial triples by one wave (lane l adds
// triples l, l+64, ... in order, then a butterfly)
__global__ void k_full_residual_final(const double* part, int nparts, double* out) {
  double R = 0.0, B = 0.0, X = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 64) {
    R += part[3 * p];
    B += part[3 * p + 1];
    X += part[3 * p + 2];
  }
  R = wave_sum(R);
  B = wave_sum(B);
  X = wave_sum(X);
  if (threadIdx.x == 0) {
    out[0] = R;
    out[1] = B;
    out[2] = X;
  }
}

// most active (coupled) nodes in any tile of `tile` consecutive nodes
// (tile <= blockDim.x); *out must start at zero
__global__ void k_max_tile_active(const int* act_node, int nb, int tile, int* out) {
  const int i = blockIdx.x * tile + threadIdx.x;
  const int c = __syncthreads_count(threadIdx.x < tile && i < nb && act_node[i] >= 0);
  if (threadIdx.x == 0) atomicMax(out, c);
}

// ---------------------------------------------------------------- launchers
static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

hipError_t launch_split(const SplitParams& S, hipStream_t st) {
  if (S.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_classify, dim3(cdiv(S.count, 256)), dim3(256), 0, st, S);
  return hipGetLastError();
}
hipError_t launch_scan2(const int* in0, int* out0, const int* in1, int* out1, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_scan2, dim3(2), dim3(1024), 0, st, in0, out0, in1, out1, n);
  return hipGetLastError();
}
hipError_t launch_active(const int* flag, const int* prefix, int nb, int* act_node, int* act_dof, hipStream_t st) {
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_active, dim3(cdiv(nb, 256)), dim3(256), 0, st, flag, prefix, nb, act_node, act_dof);
  return hipGetLastError();
}
hipError_t launch_max_tile_active(const int* act_node, int nb, int tile, int* out, hipStream_t st) {
  if (nb == 0) return hipSuccess;
  if (tile <= 0 || tile > 1024) return hipErrorInvalidValue;
  const int threads = tile <= 64 ? 64 : (tile + 63) / 64 * 64;
  hipLaunchKernelGGL(k_max_tile_active, dim3(cdiv(nb, tile)), dim3(threads), 0, st, act_node, nb, tile, out);
  return hipGetLastError();
}
hipError_t launch_place(const PlaceParams& S, hipStream_t st) {
  if (S.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_place, dim3(cdiv(S.count, 256)), dim3(256), 0, st, S);
  return hipGetLastError();
}
hipError_t launch_sort_gather(const int* rowptr, int nb, int* key_col, int* key_src, const double* vals, int nnzb,
                              int* colidx, double* v64, float* v32, hipStream_t st) {
  if (nb == 0 || nnzb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_gather, dim3(cdiv((long long)nb * 64, 256)), dim3(256), 0, st, rowptr, nb, key_col,
                     key_src, vals, nnzb, colidx, v64, v32);
  return hipGetLastError();
}
// occupancy of a kernel on the current device, queried once per process
// (the runtime's query costs tens of microseconds of host time)
static hipError_t cached_occupancy(int which, const void* fn, int* blocks) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, int> cache;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find({dev, which});
  if (it != cache.end()) {
    *blocks = it->second;
    return hipSuccess;
  }
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, fn, kBlock, 0);
  if (e == hipSuccess) cache[{dev, which}] = *blocks;
  return e;
}
hipError_t aca_occupancy(int* blocks) { return cached_occupancy(0, (const void*)k_aca, blocks); }
hipError_t aca_reg_occupancy(int* blocks) { return cached_occupancy(1, (const void*)k_aca_reg, blocks); }
// Launched as an ordinary kernel: the caller keeps the grid within the
// co-resident bound (one workgroup per CU at most), so every workgroup gets a
// CU once whatever else runs has drained, and every grid wait is bounded.  A
// cooperative launch would first wait for the setup's D-side work on the
// other stream, which this launch overlaps.
hipError_t launch_aca_reg(const AcaParams& P, hipStream_t st) {
  if (P.qm > kAcaRegQm) return hipErrorInvalidValue;
  const int cpw = (P.a + P.nwg - 1) / P.nwg;
  if ((long long)P.qm * cpw > (long long)kAcaRegPer * kBlock || cpw > kAcaRegCols) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_aca_reg, dim3(P.nwg), dim3(kBlock), 0, st, P);
  return hipGetLastError();
}
// the body path: column bodies, plan, one workgroup per body, finish (every
// kernel after the plan does nothing when the plan declines)
hipError_t launch_aca_blocks(const AcaBlkParams& P, hipStream_t st) {
  if (P.a <= 0 || P.qm <= 0) return hipSuccess;
  const int nbody = P.qm / 12;
  hipLaunchKernelGGL(k_col_bodies, dim3(cdiv(P.a, 256)), dim3(256), 0, st, P);
  hipLaunchKernelGGL(k_aca_plan, dim3(1), dim3(kBlock), 0, st, P);
  if (nbody <= kAcaBlkMaxBodies) {
    hipLaunchKernelGGL(k_aca_blocks, dim3(nbody), dim3(kBlock), 0, st, P);
    hipLaunchKernelGGL(k_aca_blocks_finish, dim3(256), dim3(256), 0, st, P);
  }
  return hipGetLastError();
}
hipError_t launch_aca(const AcaParams& P, hipStream_t st) {
  void* args[] = {(void*)&P};
  return hipLaunchCooperativeKernel((const void*)k_aca, dim3(P.nwg), dim3(kBlock), args, 0, st);
}
hipError_t launch_body_inverse(const double* A, int qm, double* Binv, int* err, hipStream_t st) {
  if (qm == 0) return hipSuccess;
  hipLaunchKernelGGL(k_body_inverse, dim3(qm / 12), dim3(64), 0, st, A, qm, Binv, err);
  return hipGetLastError();
}
hipError_t launch_abd_inverse(const AbdParams& S, hipStream_t st) {
  hipLaunchKernelGGL(k_abd_inverse, dim3(1), dim3(1024), 0, st, S);
  return hipGetLastError();
}
hipError_t launch_coupling_core(const FactorParams& S, hipStream_t st) {
  const int n = S.qm, k = S.k;  // grids for the largest rank; the kernels read the rank
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_coupling_PU, dim3((k + 15) / 16, (n + 15) / 16), dim3(256), 0, st, S);
  hipLaunchKernelGGL(k_coupling_UtW, dim3((k + 15) / 16, (k + 15) / 16), dim3(256), 0, st, S);
  hipLaunchKernelGGL(k_coupling_core, dim3(1), dim3(1024), 0, st, S);
  return hipGetLastError();
}
hipError_t launch_make_W(const double* V, const double* Lm, int a, const int* kdev, int kmax, double* W64, float* W32,
                         hipStream_t st) {
  if (a <= 0 || kmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_make_W, dim3(cdiv(a, kMakeWRows)), dim3(256), 0, st, V, Lm, a, kdev, kmax, W64, W32);
  return hipGetLastError();
}
hipError_t launch_rhs(const double* W1T, const double* V, const double* b, int qo, int qm, int fo, int n, int k,
                      int a, const int* act_node, double* y, double* rhs, double* r, double* x, hipStream_t st) {
  if (k > 0) hipLaunchKernelGGL(k_rhs_small, dim3(1), dim3(512), 0, st, W1T, b, qo, qm, k, y);
  if (n > 0)
    hipLaunchKernelGGL(k_rhs_vec, dim3(cdiv(n, 256)), dim3(256), 0, st, b, fo, n, act_node, V, a, k, y, rhs, r, x);
  return hipGetLastError();
}
hipError_t launch_backsub(const double* P, const double* U, const double* V, const int* act_dof, const double* b,
                          double* x, int qo, int qm, int fo, int a, int k, double* tmp, hipStream_t st) {
  if (qm == 0) return hipSuccess;
  if (k > 0) hipLaunchKernelGGL(k_backsub_vt, dim3(k), dim3(256), 0, st, V, act_dof, x, fo, a, tmp);
  hipLaunchKernelGGL(k_backsub, dim3(1), dim3(512), 0, st, P, U, b, x, qo, qm, k, tmp);
  return hipGetLastError();
}
hipError_t launch_full_residual(const double* yf, const double* E, const double* A, const int* act_node,
                                const int* act_dof, int a, int qm, const double* b, const double* x, int fo, int qo,
                                int n, double* part, double* out, hipStream_t st) {
  const int nparts = cdiv(n, kBlock);
  if (nparts > 0)
    hipLaunchKernelGGL(k_full_residual_fem, dim3(nparts), dim3(kBlock), 0, st, yf, E, act_node, a, qm, b, x, fo,
                       qo, n, part);
  if (qm > 0)
    hipLaunchKernelGGL(k_full_residual_abd, dim3(qm), dim3(256), 0, st, E, A, act_dof, a, qm, b, x, fo, qo,
                       part + 3 * nparts);
  hipLaunchKernelGGL(k_full_residual_final, dim3(1), dim3(64), 0, st, part, nparts + qm, out);
  return hipGetLastError();
}

}  // namespace hsc
