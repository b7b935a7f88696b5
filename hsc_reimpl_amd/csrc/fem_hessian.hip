// FEM Hessian generator on the device: per-tet log-Neo-Hookean Hessian with
// the SPD projection by cyclic Jacobi, pulled back to 3x3 node blocks
// (log_neo_hookean_function.h:93-150 and the gradient/Hessian kernel of
// log_neo_hookean_3d.cu:62-135 in the reference's overlays).  F = Ds Dm^-1
// with Ds = [x1-x0, x2-x0, x3-x0]; vec(F) column-major (F(i,j) at i + 3j);
// the shape gradient of node a > 0 is row a-1 of Dm^-1 and node 0's is minus
// their sum.  Each tet emits its 10 blocks (a <= b) as general triplets
// (row = the smaller node index), the input of the assembly.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace hsc {

namespace {

constexpr int kFemBlock = 64;  // tets per workgroup (one wave)

__device__ __forceinline__ double det3(const double F[9]) {  // column-major F[i + 3j]
  return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) + F[6] * (F[1] * F[5] - F[4] * F[2]);
}

// packed upper-triangle index of the symmetric 9x9 entry (r, c)
__host__ __device__ constexpr int pk(int r, int c) {
  return r <= c ? r * 9 - r * (r - 1) / 2 + (c - r) : c * 9 - c * (c - 1) / 2 + (r - c);
}

// The 9x9 Hessian lives in LDS as its 45 upper-triangle entries (one column
// of sH per tet); the Jacobi eigenvectors Q live in registers, every loop
// over them unrolled so that all indices are compile-time constants.  23 KB
// of LDS per 64-tet workgroup and at most 256 VGPRs (two waves per SIMD, a
// few values spilled) put 512 tets in flight per CU, which is what hides the
// latency of the rotations' dependent LDS updates (measured on 78,400 tets:
// 1.09 ms with both matrices in LDS and 96 tets per CU, 0.32 ms at one wave
// per SIMD, 0.25 ms at two).
__global__ __launch_bounds__(kFemBlock, 2) void k_fem_hessian(FemHessianParams P) {
  __shared__ double sH[45][kFemBlock];
  const int t = threadIdx.x;
  const long long e = (long long)blockIdx.x * kFemBlock + t;
  if (e >= P.ntet) return;  // (no workgroup barrier below)
  auto H = [&](int r, int c) -> double& { return sH[pk(r, c)][t]; };
  int v[4];
  for (int a = 0; a < 4; ++a) v[a] = P.tets[4 * e + a];
  const double* Di = P.Dm_inv + 9 * e;  // column-major Dm^-1
  double Ds[9], F[9];
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) Ds[i + 3 * a] = P.x[3 * v[a + 1] + i] - P.x[3 * v[0] + i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 3; ++a) s += Ds[i + 3 * a] * Di[a + 3 * j];
      F[i + 3 * j] = s;
    }
  // raw Hessian: mu I + lambda vec(G) vec(G)^T - q (swap), G = cof(F) / J
  const double mu = P.mu[e], lambda = P.lambda[e];
  const double J = det3(F);
  double G[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      G[i + 3 * j] = (F[i1 + 3 * j1] * F[i2 + 3 * j2] - F[i1 + 3 * j2] * F[i2 + 3 * j1]) / J;
    }
  const double q = lambda * log(J) - mu;
#pragma unroll
  for (int r = 0; r < 9; ++r)
#pragma unroll
    for (int c = r; c < 9; ++c) {
      const int i = r % 3, a = r / 3, j = c % 3, b = c / 3;
      H(r, c) = mu * (i == j && a == b ? 1.0 : 0.0) + lambda * G[i + 3 * a] * G[j + 3 * b] -
                q * G[i + 3 * b] * G[j + 3 * a];
    }
  if (P.project) {
    // bounded cyclic Jacobi eigendecomposition, eigenvalues clamped at 0
    double Q[9][9];
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
#pragma unroll
      for (int c = 0; c < 9; ++c) Q[r][c] = r == c ? 1.0 : 0.0;
#pragma unroll
      for (int c = r; c < 9; ++c) {
        const double h = H(r, c);
        nrm += (r == c ? 1.0 : 2.0) * h * h;
      }
    }
    const double threshold = 1e-13 * sqrt(nrm);
    for (int sweep = 0; sweep < 32; ++sweep) {
      double maximum = 0.0;
#pragma unroll
      for (int p = 0; p < 8; ++p)
#pragma unroll
        for (int qq = p + 1; qq < 9; ++qq) {
          const double apq = H(p, qq);
          maximum = fmax(maximum, fabs(apq));
          if (fabs(apq) <= threshold) continue;
          const double tau = (H(qq, qq) - H(p, p)) / (2.0 * apq);
          const double tt = copysign(1.0 / (fabs(tau) + sqrt(1.0 + tau * tau)), tau);
          const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
          H(p, p) -= tt * apq;
          H(qq, qq) += tt * apq;
          H(p, qq) = 0.0;
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            if (k != p && k != qq) {
              const double kp = H(k, p), kq = H(k, qq);
              H(k, p) = c * kp - s * kq;
              H(k, qq) = s * kp + c * kq;
            }
            const double qp = Q[k][p], qk = Q[k][qq];
            Q[k][p] = c * qp - s * qk;
            Q[k][qq] = s * qp + c * qk;
          }
        }
      if (maximum <= threshold) break;
    }
    double w[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) w[i] = fmax(H(i, i), 0.0);
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = i; j < 9; ++j) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) s += Q[i][k] * w[k] * Q[j][k];
        H(i, j) = s;  // (only Q and w are read here)
      }
  }
  // pullback: block(a, b)(k, l) = sum_{p, m} gradN_a[p] gradN_b[m] H(k + 3p, l + 3m)
  const double scale = P.vol[e] * P.dt2;
  double gN[4][3];
  for (int p = 0; p < 3; ++p) {
    gN[1][p] = Di[0 + 3 * p];
    gN[2][p] = Di[1 + 3 * p];
    gN[3][p] = Di[2 + 3 * p];
    gN[0][p] = -(gN[1][p] + gN[2][p] + gN[3][p]);
  }
  long long o = 10 * e;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) {
      const bool keep = v[i] <= v[j];
      P.orow[o] = keep ? v[i] : v[j];
      P.ocol[o] = keep ? v[j] : v[i];
      // block (l, r) with l the smaller node: (gN_l, gN_r) = keep ? (gN_i, gN_j) : (gN_j, gN_i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          double s = 0.0;
#pragma unroll
          for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) {
              const double gl = keep ? gN[i][p] : gN[j][p];
              const double gr = keep ? gN[j][qq] : gN[i][qq];
              s += gl * gr * H(k + 3 * p, m + 3 * qq);
            }
          P.oval[9 * o + k + 3 * m] = scale * s;  // column-major block
        }
      ++o;
    }
}

}  // namespace

hipError_t launch_fem_hessian(const FemHessianParams& P, hipStream_t st) {
  if (P.ntet <= 0) return hipSuccess;
  const long long grid = (P.ntet + kFemBlock - 1) / kFemBlock;
  if (grid >= (1LL << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fem_hessian, dim3((unsigned)grid), dim3(kFemBlock), 0, st, P);
  return hipGetLastError();
}

}  // namespace hsc
