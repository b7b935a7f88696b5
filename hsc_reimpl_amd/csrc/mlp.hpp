// Graph-MLP network object (host side) of the learned preconditioner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "learned.hpp"

namespace hsc {

// flat FP32 parameter layout (Wo/bo padded from 3 to 32 output columns)
struct MlpLayout {
  static constexpr int IN = 16, HID = 64, OUTP = 32;
  static constexpr int oW0 = 0;
  static constexpr int ob0 = oW0 + IN * HID;
  static constexpr int oWs1 = ob0 + HID;
  static constexpr int oWn1 = oWs1 + HID * HID;
  static constexpr int ob1 = oWn1 + HID * HID;
  static constexpr int oWs2 = ob1 + HID;
  static constexpr int oWn2 = oWs2 + HID * HID;
  static constexpr int ob2 = oWn2 + HID * HID;
  static constexpr int oWo = ob2 + HID;
  static constexpr int obo = oWo + HID * OUTP;
  static constexpr int count = obo + OUTP;
};

class MlpNet {
 public:
  MlpNet() = default;
  ~MlpNet();
  MlpNet(const MlpNet&) = delete;
  MlpNet& operator=(const MlpNet&) = delete;

  // graph: neighbour CSR without self loops (device pointers are copied)
  hipError_t init(int nb, const int* nbr_ptr_dev, const int* nbr_dev, int n_edges, int max_batch, hipStream_t st);
  // another graph of at most the nodes and edges init was given (copied; no allocation)
  hipError_t set_graph(int nb, const int* nbr_ptr_dev, const int* nbr_dev, int n_edges, hipStream_t st);
  bool fits(int nb, int n_edges) const { return nb <= cap_nb_ && n_edges <= cap_edges_; }
  // use a graph the caller keeps alive and unchanged while the net uses it
  // (not copied; the forward and backward passes read it in place)
  void borrow_graph(int nb, const int* nbr_ptr_dev, const int* nbr_dev) {
    nb_ = nb;
    gptr_ = nbr_ptr_dev;
    gnbr_ = nbr_dev;
  }
  hipError_t set_params(const float* theta, bool device_ptr, hipStream_t st);
  hipError_t get_params(float* theta_host, hipStream_t st);
  hipError_t forward(const bf16_t* X, int B, float* O, hipStream_t st);
  hipError_t backward(const bf16_t* X, const float* T, const float* den, int B, float* loss_host, hipStream_t st);
  hipError_t adam(float lr, hipStream_t st);
  const float* grad() const { return grad_; }
  int nb() const { return nb_; }
  int max_batch() const { return max_batch_; }

 private:
  hipError_t cast_all(hipStream_t st);
  int nb_ = 0, max_batch_ = 0, step_ = 0, cap_nb_ = 0, cap_edges_ = 0;
  int* ptr_ = nullptr;
  int* nbr_ = nullptr;
  const int* gptr_ = nullptr;  // the graph in use: borrowed, else ptr_ / nbr_
  const int* gnbr_ = nullptr;
  const int* gp() const { return gptr_ ? gptr_ : ptr_; }
  const int* gn() const { return gptr_ ? gnbr_ : nbr_; }
  float* theta_ = nullptr;
  float* grad_ = nullptr;
  float* m_ = nullptr;
  float* v_ = nullptr;
  bf16_t* wb_ = nullptr;  // row-major copies, same offsets as theta (matrices only)
  bf16_t* wt_ = nullptr;  // transposed copies
  bf16_t *H0_ = nullptr, *M1_ = nullptr, *H1_ = nullptr, *M2_ = nullptr, *H2_ = nullptr;
  bf16_t *dO_ = nullptr, *dA2_ = nullptr, *dA1_ = nullptr, *dA0_ = nullptr;
  float *O_ = nullptr, *G1_ = nullptr, *G2_ = nullptr, *part_ = nullptr, *lossp_ = nullptr;
};

}  // namespace hsc
