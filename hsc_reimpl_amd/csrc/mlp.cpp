// Host side of the BF16 graph-MLP (learned preconditioner): forward,
// backward (hand-written kernels in learned_kernels.hip), Adam, and the C ABI
// hsc_mlp_* used to train it on the GPU.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/hsc.h"
#include "common.hpp"
#include "mlp.hpp"

namespace hsc {

namespace {
template <class T>
hipError_t dmalloc(T** p, size_t n) {
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1));
  if (e == hipSuccess) note_alloc();
  return e;
}
#define TRY(x)                           \
  do {                                   \
    hipError_t _e = (x);                 \
    if (_e != hipSuccess) return _e;     \
  } while (0)
constexpr int kRowsPerChunk = 256;
}  // namespace

MlpNet::~MlpNet() {
  for (void* p : {(void*)ptr_, (void*)nbr_, (void*)theta_, (void*)grad_, (void*)m_, (void*)v_, (void*)wb_, (void*)wt_,
                  (void*)H0_, (void*)M1_, (void*)H1_, (void*)M2_, (void*)H2_, (void*)dO_, (void*)dA2_, (void*)dA1_,
                  (void*)dA0_, (void*)O_, (void*)G1_, (void*)G2_, (void*)part_, (void*)lossp_})
    if (p) (void)hipFree(p);
}

hipError_t MlpNet::init(int nb, const int* nbr_ptr_dev, const int* nbr_dev, int n_edges, int max_batch,
                        hipStream_t st) {
  using L = MlpLayout;
  nb_ = nb;
  cap_nb_ = nb;
  cap_edges_ = n_edges;
  max_batch_ = max_batch;
  const size_t M = (size_t)nb * max_batch;
  TRY(dmalloc(&ptr_, nb + 1));
  TRY(dmalloc(&nbr_, n_edges));
  TRY(hipMemcpyAsync(ptr_, nbr_ptr_dev, sizeof(int) * (nb + 1), hipMemcpyDefault, st));
  if (n_edges) TRY(hipMemcpyAsync(nbr_, nbr_dev, sizeof(int) * n_edges, hipMemcpyDefault, st));
  TRY(dmalloc(&theta_, L::count));
  TRY(dmalloc(&grad_, L::count));
  TRY(dmalloc(&m_, L::count));
  TRY(dmalloc(&v_, L::count));
  TRY(dmalloc(&wb_, L::count));
  TRY(dmalloc(&wt_, L::count));
  TRY(hipMemsetAsync(theta_, 0, sizeof(float) * L::count, st));
  TRY(hipMemsetAsync(m_, 0, sizeof(float) * L::count, st));
  TRY(hipMemsetAsync(v_, 0, sizeof(float) * L::count, st));
  for (bf16_t** p : {&H0_, &M1_, &H1_, &M2_, &H2_, &dA2_, &dA1_, &dA0_}) TRY(dmalloc(p, M * 64));
  TRY(dmalloc(&dO_, M * 32));
  TRY(dmalloc(&O_, M * 3));
  TRY(dmalloc(&G1_, M * 64));
  TRY(dmalloc(&G2_, M * 64));
  TRY(dmalloc(&part_, ((M + kRowsPerChunk - 1) / kRowsPerChunk) * 64 * 64));
  TRY(dmalloc(&lossp_, (M + 255) / 256));
  step_ = 0;
  return cast_all(st);
}

hipError_t MlpNet::cast_all(hipStream_t st) {
  using L = MlpLayout;
  const int off[6] = {L::oW0, L::oWs1, L::oWn1, L::oWs2, L::oWn2, L::oWo};
  const int K[6] = {L::IN, L::HID, L::HID, L::HID, L::HID, L::HID};
  const int N[6] = {L::HID, L::HID, L::HID, L::HID, L::HID, L::OUTP};
  for (int i = 0; i < 6; ++i) TRY(launch_mlp_cast(theta_ + off[i], K[i], N[i], wb_ + off[i], wt_ + off[i], st));
  return hipSuccess;
}

hipError_t MlpNet::set_params(const float* theta, bool device_ptr, hipStream_t st) {
  TRY(hipMemcpyAsync(theta_, theta, sizeof(float) * MlpLayout::count,
                     device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
  TRY(hipMemsetAsync(m_, 0, sizeof(float) * MlpLayout::count, st));
  TRY(hipMemsetAsync(v_, 0, sizeof(float) * MlpLayout::count, st));
  step_ = 0;
  return cast_all(st);
}

hipError_t MlpNet::get_params(float* theta_host, hipStream_t st) {
  TRY(hipMemcpyAsync(theta_host, theta_, sizeof(float) * MlpLayout::count, hipMemcpyDeviceToHost, st));
  return hipStreamSynchronize(st);
}

hipError_t MlpNet::set_graph(int nb, const int* nbr_ptr_dev, const int* nbr_dev, int n_edges, hipStream_t st) {
  if (!fits(nb, n_edges)) return hipErrorInvalidValue;
  nb_ = nb;
  gptr_ = gnbr_ = nullptr;
  TRY(hipMemcpyAsync(ptr_, nbr_ptr_dev, sizeof(int) * (nb + 1), hipMemcpyDefault, st));
  if (n_edges) TRY(hipMemcpyAsync(nbr_, nbr_dev, sizeof(int) * n_edges, hipMemcpyDefault, st));
  return hipSuccess;
}

hipError_t MlpNet::forward(const bf16_t* X, int B, float* O, hipStream_t st) {
  using L = MlpLayout;
  if (B > max_batch_ || B <= 0) return hipErrorInvalidValue;
  const int M = nb_ * B;
  MlpGemm g{};
  g.A[0] = X;
  g.K[0] = L::IN;
  g.Bt[0] = wt_ + L::oW0;
  g.nsrc = 1;
  g.bias = theta_ + L::ob0;
  g.M = M;
  g.N = 64;
  g.ldy = 64;
  g.relu = 1;
  g.Y = H0_;
  TRY(launch_mlp_gemm(g, st));
  TRY(launch_mlp_aggregate(H0_, gp(), gn(), nb_, M, M1_, st));
  g.A[0] = H0_;
  g.A[1] = M1_;
  g.K[0] = g.K[1] = L::HID;
  g.Bt[0] = wt_ + L::oWs1;
  g.Bt[1] = wt_ + L::oWn1;
  g.nsrc = 2;
  g.bias = theta_ + L::ob1;
  g.Y = H1_;
  TRY(launch_mlp_gemm(g, st));
  TRY(launch_mlp_aggregate(H1_, gp(), gn(), nb_, M, M2_, st));
  g.A[0] = H1_;
  g.A[1] = M2_;
  g.Bt[0] = wt_ + L::oWs2;
  g.Bt[1] = wt_ + L::oWn2;
  g.bias = theta_ + L::ob2;
  g.Y = H2_;
  TRY(launch_mlp_gemm(g, st));
  MlpGemm o{};
  o.A[0] = H2_;
  o.K[0] = L::HID;
  o.Bt[0] = wt_ + L::oWo;
  o.nsrc = 1;
  o.bias = theta_ + L::obo;
  o.M = M;
  o.N = 32;
  o.ldy = 3;
  o.relu = 0;
  o.Yf = O ? O : O_;
  return launch_mlp_gemm(o, st);
}

hipError_t MlpNet::backward(const bf16_t* X, const float* T, const float* den, int B, float* loss_host,
                            hipStream_t st) {
  using L = MlpLayout;
  TRY(forward(X, B, O_, st));
  const int M = nb_ * B;
  TRY(hipMemsetAsync(grad_, 0, sizeof(float) * L::count, st));
  TRY(launch_mlp_loss(O_, T, den, nb_, M, B, wb_ + L::oWo, H2_, dO_, dA2_, lossp_, st));
  auto wgrad = [&](const bf16_t* A, int lda, int K, const bf16_t* dY, int ldd, int N, int off) {
    MlpWgrad w{A, dY, lda, ldd, K, N, M, kRowsPerChunk, part_};
    return launch_mlp_wgrad(w, grad_ + off, st);
  };
  TRY(wgrad(H2_, 64, 64, dO_, 32, 32, L::oWo));
  TRY(launch_mlp_colsum(dO_, M, 32, 32, kRowsPerChunk, part_, grad_ + L::obo, st));
  // layer 2
  TRY(wgrad(H1_, 64, 64, dA2_, 64, 64, L::oWs2));
  TRY(wgrad(M2_, 64, 64, dA2_, 64, 64, L::oWn2));
  TRY(launch_mlp_colsum(dA2_, M, 64, 64, kRowsPerChunk, part_, grad_ + L::ob2, st));
  MlpGemm g{};
  g.A[0] = dA2_;
  g.K[0] = 64;
  g.nsrc = 1;
  g.M = M;
  g.N = 64;
  g.ldy = 64;
  g.Bt[0] = wb_ + L::oWs2;
  g.Yf = G1_;
  TRY(launch_mlp_gemm(g, st));
  g.Bt[0] = wb_ + L::oWn2;
  g.Yf = G2_;
  TRY(launch_mlp_gemm(g, st));
  TRY(launch_mlp_aggT_mask(G1_, G2_, H1_, gp(), gn(), nb_, M, dA1_, st));
  // layer 1
  TRY(wgrad(H0_, 64, 64, dA1_, 64, 64, L::oWs1));
  TRY(wgrad(M1_, 64, 64, dA1_, 64, 64, L::oWn1));
  TRY(launch_mlp_colsum(dA1_, M, 64, 64, kRowsPerChunk, part_, grad_ + L::ob1, st));
  g.A[0] = dA1_;
  g.Bt[0] = wb_ + L::oWs1;
  g.Yf = G1_;
  TRY(launch_mlp_gemm(g, st));
  g.Bt[0] = wb_ + L::oWn1;
  g.Yf = G2_;
  TRY(launch_mlp_gemm(g, st));
  TRY(launch_mlp_aggT_mask(G1_, G2_, H0_, gp(), gn(), nb_, M, dA0_, st));
  // layer 0
  TRY(wgrad(X, 16, 16, dA0_, 64, 64, L::oW0));
  TRY(launch_mlp_colsum(dA0_, M, 64, 64, kRowsPerChunk, part_, grad_ + L::ob0, st));
  if (loss_host) {
    const int nparts = (M + 255) / 256;
    std::vector<float> lp(nparts);
    TRY(hipMemcpyAsync(lp.data(), lossp_, sizeof(float) * nparts, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    double s = 0.0;
    for (float v : lp) s += v;
    *loss_host = (float)s;
  }
  return hipSuccess;
}

hipError_t MlpNet::adam(float lr, hipStream_t st) {
  ++step_;
  TRY(launch_mlp_adam(theta_, grad_, m_, v_, MlpLayout::count, lr, step_, nullptr, st));
  return cast_all(st);
}

}  // namespace hsc

struct hsc_mlp {
  hsc::MlpNet net;
  int device = 0;
};

namespace {
struct DevGuard {
  int prev = -1;
  explicit DevGuard(int d) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != d) (void)hipSetDevice(d);
  }
  ~DevGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};
int code(hipError_t e) { return e == hipSuccess ? HSC_OK : (e == hipErrorInvalidValue ? HSC_ERR_INVALID_ARGUMENT : HSC_ERR_HIP); }
}  // namespace

extern "C" {

int32_t hsc_mlp_param_count(void) { return hsc::MlpLayout::count; }

int32_t hsc_mlp_create(int32_t device, int32_t nb, const int32_t* nbr_ptr, const int32_t* nbr, int64_t n_edges,
                       int32_t max_batch, hsc_mlp** out) {
  if (!out || nb <= 0 || max_batch <= 0 || !nbr_ptr || (n_edges > 0 && !nbr)) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(device);
  hsc_mlp* m = new hsc_mlp();
  m->device = device;
  hipError_t e = m->net.init(nb, nbr_ptr, nbr, (int)n_edges, max_batch, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    delete m;
    return code(e);
  }
  *out = m;
  return HSC_OK;
}

void hsc_mlp_destroy(hsc_mlp* m) {
  if (!m) return;
  DevGuard g(m->device);
  delete m;
}

int32_t hsc_mlp_set_params(hsc_mlp* m, const float* theta, int32_t device_ptr) {
  if (!m || !theta) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(m->device);
  hipError_t e = m->net.set_params(theta, device_ptr != 0, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  return code(e);
}

int32_t hsc_mlp_get_params(hsc_mlp* m, float* theta_host) {
  if (!m || !theta_host) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(m->device);
  return code(m->net.get_params(theta_host, nullptr));
}

int32_t hsc_mlp_forward(hsc_mlp* m, const uint16_t* X, int32_t B, float* O, hsc_stream stream) {
  if (!m || !X || !O) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(m->device);
  hipError_t e = m->net.forward(X, B, O, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  return code(e);
}

int32_t hsc_mlp_gradients(hsc_mlp* m, const uint16_t* X, const float* T, const float* den, int32_t B,
                          float* grad_host, float* loss, hsc_stream stream) {
  if (!m || !X || !T || !den || B <= 0 || B > m->net.max_batch()) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(m->device);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = m->net.backward(X, T, den, B, loss, st);
  if (e == hipSuccess && grad_host)
    e = hipMemcpyAsync(grad_host, m->net.grad(), sizeof(float) * hsc::MlpLayout::count, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return code(e);
}

int32_t hsc_mlp_train_step(hsc_mlp* m, const uint16_t* X, const float* T, const float* den, int32_t B, float lr,
                           float* loss, hsc_stream stream) {
  if (!m || !X || !T || !den || B <= 0 || B > m->net.max_batch()) return HSC_ERR_INVALID_ARGUMENT;
  DevGuard g(m->device);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = m->net.backward(X, T, den, B, loss, st);
  if (e == hipSuccess) e = m->net.adam(lr, st);
  if (e == hipSuccess && !loss) e = hipStreamSynchronize(st);
  return code(e);
}

}  // extern "C"
