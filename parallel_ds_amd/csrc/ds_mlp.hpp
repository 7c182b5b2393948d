// Warm-start network: kernel launchers (ds_mlp.hip) and the host-side model.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <vector>

#include "ds_solver.hpp"

namespace pds {

typedef uint16_t bf16_t;  // BF16 bit pattern

constexpr int MLP_WIDTH = 256;   // hidden width
constexpr int MLP_KPAD = 32;     // input features padded to one MFMA K-step
constexpr int MLP_EPI_BIAS_RELU = 0, MLP_EPI_MASK = 1, MLP_EPI_F32 = 2;

constexpr int MLP_SPLITK_WGS = 256;            // target grid of a split-K weight-gradient GEMM
constexpr size_t MLP_SPLITK_FLOATS = 1u << 21;  // its partial-product workspace (8 MiB)
// C = A B^T.  With the FP32 epilogue and a workspace of ws_floats floats, K is
// split over the grid's z dimension and the partials summed in a second launch.
void mlp_gemm_nt(hipStream_t st, int epi, const bf16_t* A, const bf16_t* B, int M, int N, int K, const float* bias,
                 const bf16_t* mask, bf16_t* C16, bf16_t* CT16, float* C32, hipEvent_t e0 = nullptr,
                 hipEvent_t e1 = nullptr, float* ws = nullptr, size_t ws_floats = 0);
void mlp_head(hipStream_t st, const bf16_t* H3, const float* w4, const float* b4, int S, int Sv, const int* rows,
              const float* y, const float* wt, float inv_s, float* z, float* dz, float* loss_part, bf16_t* G3,
              bf16_t* G3T);
void mlp_rowsum(hipStream_t st, const bf16_t* X, int R, int C, const float* scale, float* out);
void mlp_sum(hipStream_t st, const float* v, int n, float* out);
void mlp_adam(hipStream_t st, float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2,
              float eps, float c1, float c2);
void mlp_refresh(hipStream_t st, const float* W, int O, int I, bf16_t* Wb, bf16_t* WTb);
void mlp_pack(hipStream_t st, const float* F, const int* rows, int Sv, int S, int nf, int Kp, const float* mu,
              const float* sd, bf16_t* X, bf16_t* XT);
void mlp_infer(hipStream_t st, const bf16_t* X, int S, int Kp, const bf16_t* W1, const float* b1, const bf16_t* W2,
               const float* b2, const bf16_t* W3, const float* b3, const float* w4, const float* b4, float* z,
               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// Parameter vector layout (FP32): W1 [256 x 32], b1 [256], W2 [256 x 256],
// b2, W3 [256 x 256], b3, w4 [256], b4 [1].  Inputs beyond n_features are
// zero, so the matching columns of W1 get no gradient.
struct MlpLayout {
  static constexpr int W = MLP_WIDTH, K = MLP_KPAD;
  static constexpr size_t oW1 = 0, ob1 = oW1 + (size_t)W * K, oW2 = ob1 + W, ob2 = oW2 + (size_t)W * W,
                          oW3 = ob2 + W, ob3 = oW3 + (size_t)W * W, ow4 = ob3 + W, ob4 = ow4 + W, total = ob4 + 1;
};

class Mlp {
 public:
  Mlp(int device, int n_features, unsigned long long seed);
  ~Mlp();
  int n_features() const { return nf_; }
  void get_params(float* out);
  void set_params(const float* in);
  void set_norm(const float* mu, const float* sd);
  void get_norm(float* mu, float* sd);
  // Logits of S rows of FP32 features (device pointers; rows may index F).
  void infer(const float* F, const int* rows, int S, float* z_dev);
  // Gradient of the mean weighted BCE over S rows (device pointers) into
  // the FP32 gradient buffer; returns the loss.
  double grad(const float* F, const int* rows, int S, const float* y, const float* wt);
  const float* grad_dev() const { return g_.p; }
  void adam_step(float lr);
  // One pass over the rows in `order` (device, length S_total), minibatches of
  // `batch` rows; returns the mean loss.
  double train_epoch(const float* F, const int* order, int S_total, const float* y, int batch, float lr);
  hipStream_t stream() const { return st_; }
  int device() const { return device_; }
  double last_infer_ms() const { return infer_ms_; }
  double gemm_ms() const { return gemm_ms_; }
  long long gemm_flops() const { return gemm_flops_; }
  // Family prior of the learned warm start (feature 16): per variable index,
  // the fraction of the training bases in which it is basic; empty = none.
  void set_prior(std::vector<float> p) { prior_ = std::move(p); }
  const std::vector<float>& prior() const { return prior_; }
  void reset_counters() {
    gemm_ms_ = 0.0;
    gemm_flops_ = 0;
  }

 private:
  int device_ = 0;
  int nf_ = 0;
  hipStream_t st_ = nullptr;
  long long step_ = 0;
  DBuf<float> p_, g_, m1_, m2_, mu_, sd_, z_, dz_, loss_, scal_;
  DBuf<float> ws_;  // split-K partial products of the weight gradients (MLP_SPLITK_FLOATS)
  DBuf<bf16_t> W1b_, W1Tb_, W2b_, W2Tb_, W3b_, W3Tb_;
  DBuf<bf16_t> X_, XT_, H1_, H1T_, H2_, H2T_, H3_, H3T_, G3_, G3T_, G2_, G2T_, G1_, G1T_;
  int cap_ = 0;
  double infer_ms_ = 0.0, gemm_ms_ = 0.0;
  long long gemm_flops_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<float> prior_;
  void refresh();
  void reserve(int S);
};

}  // namespace pds
