// Parameter blocks and launchers of the BF16 graph-MLP kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace hsc {

typedef unsigned short bf16_t;

struct MlpGemm {
  const bf16_t* A[2];   // [M x K_s] row-major
  int K[2];             // multiples of 8
  const bf16_t* Bt[2];  // [N x K_s] row-major (row n = column n of B)
  int nsrc;
  const float* bias;    // [N] or null
  int M, N;             // N in {32, 64}
  int ldy;              // stored columns (<= N)
  int relu;
  bf16_t* Y;            // [M x ldy] bf16 or null
  float* Yf;            // [M x ldy] fp32 or null
};

struct MlpWgrad {
  const bf16_t* A;      // [M x lda], uses columns < K
  const bf16_t* dY;     // [M x ldd], uses columns < N
  int lda, ldd, K, N, M;  // K, N <= 64
  int rows_per_chunk;   // multiple of 8
  float* partial;       // [chunks x K x N]
};

hipError_t launch_mlp_gemm(const MlpGemm& g, hipStream_t st);
hipError_t launch_mlp_wgrad(const MlpWgrad& g, float* out, hipStream_t st);
hipError_t launch_mlp_colsum(const bf16_t* dY, int M, int N, int ldd, int rows_per_chunk, float* partial, float* out,
                             hipStream_t st);
hipError_t launch_mlp_aggregate(const bf16_t* H, const int* ptr, const int* nbr, int nb, int M, bf16_t* Y,
                                hipStream_t st);
hipError_t launch_mlp_aggT_mask(const float* G1, const float* G2, const bf16_t* Hprev, const int* ptr, const int* nbr,
                                int nb, int M, bf16_t* dA, hipStream_t st);
hipError_t launch_mlp_loss(const float* O, const float* T, const float* den, int nb, int M, int B, const bf16_t* Wo,
                           const bf16_t* H2, bf16_t* dO, bf16_t* dA2, float* loss_part, hipStream_t st);
hipError_t launch_mlp_adam(float* theta, const float* grad, float* m, float* v, int count, float lr, int step,
                           const unsigned char* frozen, hipStream_t st);
hipError_t launch_mlp_cast(const float* W, int K, int N, bf16_t* Wb, bf16_t* Wt, hipStream_t st);
// input features of the learned polynomial (nb x 16 bf16, from D only) and
// the node mean of the network's 3 outputs (deterministic)
hipError_t launch_poly_features(const int* rowptr, const int* colidx, const double* val64, const float* Dfeat,
                                const float* sdiag, const int* act_node, int nb, bf16_t* X, hipStream_t st);
// partials: mean_rows3_parts(nb) x 3 doubles
int mean_rows3_parts(int nb);
hipError_t launch_mean_rows3(const float* O, int nb, double* part, double* out, hipStream_t st);

}  // namespace hsc
