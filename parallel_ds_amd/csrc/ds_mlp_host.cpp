// Host side of the warm-start network: parameters, Adam state, activation
// workspace and the launch sequences of inference, gradient and training.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ds_mlp.hpp"
#include "ds_rng.hpp"

namespace pds {

namespace {
constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1e-8f;
int round_up(int v, int q) { return (v + q - 1) / q * q; }
}  // namespace

Mlp::Mlp(int device, int n_features, unsigned long long seed) : device_(device), nf_(n_features) {
  if (n_features <= 0 || n_features > MLP_KPAD) throw std::invalid_argument("n_features must be in 1..32");
  PDS_HIP(hipSetDevice(device_));
  PDS_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  PDS_HIP(hipEventCreate(&ev0_));
  PDS_HIP(hipEventCreate(&ev1_));
  using L = MlpLayout;
  p_.alloc(L::total);
  g_.alloc(L::total);
  m1_.alloc(L::total);
  m2_.alloc(L::total);
  mu_.alloc(MLP_KPAD);
  sd_.alloc(MLP_KPAD);
  scal_.alloc(4);
  ws_.alloc(MLP_SPLITK_FLOATS);
  const int W = MLP_WIDTH, K = MLP_KPAD;
  W1b_.alloc((size_t)W * K);
  W1Tb_.alloc((size_t)W * K);
  W2b_.alloc((size_t)W * W);
  W2Tb_.alloc((size_t)W * W);
  W3b_.alloc((size_t)W * W);
  W3Tb_.alloc((size_t)W * W);
  // He-uniform weights (fan-in of the real inputs), zero biases.
  std::vector<float> p(L::total, 0.f);
  SplitMix rng(seed);
  auto fill = [&](size_t off, int O, int I, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int o = 0; o < O; ++o)
      for (int i = 0; i < I; ++i)
        p[off + (size_t)o * I + i] = (i < fan_in || I != K) ? (float)((2.0 * rng.uniform() - 1.0) * a) : 0.f;
  };
  fill(L::oW1, W, K, nf_);
  fill(L::oW2, W, W, W);
  fill(L::oW3, W, W, W);
  fill(L::ow4, 1, W, W);
  set_params(p.data());
  std::vector<float> zero(L::total, 0.f);
  PDS_HIP(hipMemcpy(m1_.p, zero.data(), sizeof(float) * L::total, hipMemcpyHostToDevice));
  PDS_HIP(hipMemcpy(m2_.p, zero.data(), sizeof(float) * L::total, hipMemcpyHostToDevice));
  std::vector<float> mu(MLP_KPAD, 0.f), sd(MLP_KPAD, 1.f);
  set_norm(mu.data(), sd.data());
}

Mlp::~Mlp() {
  if (st_) {
    (void)hipStreamSynchronize(st_);
    (void)hipStreamDestroy(st_);
  }
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
}

void Mlp::refresh() {
  using L = MlpLayout;
  mlp_refresh(st_, p_.p + L::oW1, MLP_WIDTH, MLP_KPAD, W1b_.p, W1Tb_.p);
  mlp_refresh(st_, p_.p + L::oW2, MLP_WIDTH, MLP_WIDTH, W2b_.p, W2Tb_.p);
  mlp_refresh(st_, p_.p + L::oW3, MLP_WIDTH, MLP_WIDTH, W3b_.p, W3Tb_.p);
  PDS_HIP(hipGetLastError());
}

void Mlp::get_params(float* out) {
  PDS_HIP(hipMemcpyAsync(out, p_.p, sizeof(float) * MlpLayout::total, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}

void Mlp::set_params(const float* in) {
  PDS_HIP(hipMemcpyAsync(p_.p, in, sizeof(float) * MlpLayout::total, hipMemcpyHostToDevice, st_));
  refresh();
  PDS_HIP(hipStreamSynchronize(st_));
}

void Mlp::set_norm(const float* mu, const float* sd) {
  std::vector<float> a(MLP_KPAD, 0.f), b(MLP_KPAD, 1.f);
  for (int i = 0; i < nf_; ++i) {
    a[i] = mu[i];
    b[i] = sd[i] > 0.f ? sd[i] : 1.f;
  }
  PDS_HIP(hipMemcpyAsync(mu_.p, a.data(), sizeof(float) * MLP_KPAD, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(sd_.p, b.data(), sizeof(float) * MLP_KPAD, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}

void Mlp::get_norm(float* mu, float* sd) {
  std::vector<float> a(MLP_KPAD), b(MLP_KPAD);
  PDS_HIP(hipMemcpyAsync(a.data(), mu_.p, sizeof(float) * MLP_KPAD, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipMemcpyAsync(b.data(), sd_.p, sizeof(float) * MLP_KPAD, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
  for (int i = 0; i < nf_; ++i) {
    mu[i] = a[i];
    sd[i] = b[i];
  }
}

void Mlp::reserve(int S) {
  const int Sp = round_up(std::max(S, 1), 64);
  if (Sp <= cap_) return;
  const size_t W = MLP_WIDTH;
  X_.alloc((size_t)Sp * MLP_KPAD);
  XT_.alloc((size_t)Sp * MLP_KPAD);
  for (DBuf<bf16_t>* b : {&H1_, &H1T_, &H2_, &H2T_, &H3_, &H3T_, &G3_, &G3T_, &G2_, &G2T_, &G1_, &G1T_})
    b->alloc((size_t)Sp * W);
  z_.alloc(Sp);
  dz_.alloc(Sp);
  loss_.alloc(Sp / 4 + 1);
  cap_ = Sp;
}

void Mlp::infer(const float* F, const int* rows, int S, float* z_dev) {
  if (S <= 0) return;
  reserve(S);
  mlp_pack(st_, F, rows, S, S, nf_, MLP_KPAD, mu_.p, sd_.p, X_.p, nullptr);
  using L = MlpLayout;
  mlp_infer(st_, X_.p, S, MLP_KPAD, W1b_.p, p_.p + L::ob1, W2b_.p, p_.p + L::ob2, W3b_.p, p_.p + L::ob3,
            p_.p + L::ow4, p_.p + L::ob4, z_dev, ev0_, ev1_);
  PDS_HIP(hipGetLastError());
  PDS_HIP(hipEventSynchronize(ev1_));
  float ms = 0.f;
  PDS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  infer_ms_ = ms;
}

// Forward with stored activations, loss, and the full backward pass into g_.
// The loss of this call is left in scal_[0] (device).
static void enqueue_grad(hipStream_t st, int nf, const float* F, const int* rows, int S, const float* y,
                         const float* wt, const float* p, float* g, const float* mu, const float* sd, bf16_t* X,
                         bf16_t* XT, const bf16_t* W2T, const bf16_t* W3T, const bf16_t* W1, const bf16_t* W2,
                         const bf16_t* W3, bf16_t* H1, bf16_t* H1T, bf16_t* H2, bf16_t* H2T, bf16_t* H3, bf16_t* H3T,
                         bf16_t* G3, bf16_t* G3T, bf16_t* G2, bf16_t* G2T, bf16_t* G1, bf16_t* G1T, float* z,
                         float* dz, float* loss_part, float* loss_out, float* ws) {
  using L = MlpLayout;
  const int W = MLP_WIDTH, K = MLP_KPAD;
  const int Sp = round_up(S, 64);
  mlp_pack(st, F, rows, S, Sp, nf, K, mu, sd, X, XT);
  mlp_gemm_nt(st, MLP_EPI_BIAS_RELU, X, W1, Sp, W, K, p + L::ob1, nullptr, H1, H1T, nullptr);
  mlp_gemm_nt(st, MLP_EPI_BIAS_RELU, H1, W2, Sp, W, W, p + L::ob2, nullptr, H2, H2T, nullptr);
  mlp_gemm_nt(st, MLP_EPI_BIAS_RELU, H2, W3, Sp, W, W, p + L::ob3, nullptr, H3, H3T, nullptr);
  mlp_head(st, H3, p + L::ow4, p + L::ob4, Sp, S, rows, y, wt, 1.f / (float)S, z, dz, loss_part, G3, G3T);
  mlp_rowsum(st, H3T, W, Sp, dz, g + L::ow4);
  mlp_sum(st, dz, Sp, g + L::ob4);
  mlp_gemm_nt(st, MLP_EPI_F32, G3T, H2T, W, W, Sp, nullptr, nullptr, nullptr, nullptr, g + L::oW3, nullptr, nullptr, ws,
              MLP_SPLITK_FLOATS);
  mlp_rowsum(st, G3T, W, Sp, nullptr, g + L::ob3);
  mlp_gemm_nt(st, MLP_EPI_MASK, G3, W3T, Sp, W, W, nullptr, H2, G2, G2T, nullptr);
  mlp_gemm_nt(st, MLP_EPI_F32, G2T, H1T, W, W, Sp, nullptr, nullptr, nullptr, nullptr, g + L::oW2, nullptr, nullptr, ws,
              MLP_SPLITK_FLOATS);
  mlp_rowsum(st, G2T, W, Sp, nullptr, g + L::ob2);
  mlp_gemm_nt(st, MLP_EPI_MASK, G2, W2T, Sp, W, W, nullptr, H1, G1, G1T, nullptr);
  mlp_gemm_nt(st, MLP_EPI_F32, G1T, XT, W, K, Sp, nullptr, nullptr, nullptr, nullptr, g + L::oW1, nullptr, nullptr, ws,
              MLP_SPLITK_FLOATS);
  mlp_rowsum(st, G1T, W, Sp, nullptr, g + L::ob1);
  mlp_sum(st, loss_part, Sp / 4, loss_out);
}

double Mlp::grad(const float* F, const int* rows, int S, const float* y, const float* wt) {
  if (S <= 0) throw std::invalid_argument("empty sample");
  reserve(S);
  enqueue_grad(st_, nf_, F, rows, S, y, wt, p_.p, g_.p, mu_.p, sd_.p, X_.p, XT_.p, W2Tb_.p, W3Tb_.p, W1b_.p, W2b_.p,
               W3b_.p, H1_.p, H1T_.p, H2_.p, H2T_.p, H3_.p, H3T_.p, G3_.p, G3T_.p, G2_.p, G2T_.p, G1_.p, G1T_.p, z_.p,
               dz_.p, loss_.p, scal_.p, ws_.p);
  PDS_HIP(hipGetLastError());
  float loss = 0.f;
  PDS_HIP(hipMemcpyAsync(&loss, scal_.p, sizeof(float), hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
  return loss / (double)S;
}

void Mlp::adam_step(float lr) {
  ++step_;
  const float c1 = 1.f - std::pow(ADAM_B1, (float)step_), c2 = 1.f - std::pow(ADAM_B2, (float)step_);
  mlp_adam(st_, p_.p, g_.p, m1_.p, m2_.p, (int)MlpLayout::total, lr, ADAM_B1, ADAM_B2, ADAM_EPS, c1, c2);
  refresh();
}

double Mlp::train_epoch(const float* F, const int* order, int S_total, const float* y, int batch, float lr) {
  if (S_total <= 0 || batch <= 0) throw std::invalid_argument("empty training set");
  reserve(std::min(batch, S_total));
  const int steps = (S_total + batch - 1) / batch;
  DBuf<float> losses;
  losses.alloc(steps);
  PDS_HIP(hipEventRecord(ev0_, st_));
  for (int s = 0; s < steps; ++s) {
    const int off = s * batch, S = std::min(batch, S_total - off);
    enqueue_grad(st_, nf_, F, order + off, S, y, nullptr, p_.p, g_.p, mu_.p, sd_.p, X_.p, XT_.p, W2Tb_.p, W3Tb_.p,
                 W1b_.p, W2b_.p, W3b_.p, H1_.p, H1T_.p, H2_.p, H2T_.p, H3_.p, H3T_.p, G3_.p, G3T_.p, G2_.p, G2T_.p,
                 G1_.p, G1T_.p, z_.p, dz_.p, loss_.p, losses.p + s, ws_.p);
    adam_step(lr);
    gemm_flops_ += 2LL * round_up(S, 64) * MLP_WIDTH * (2LL * MLP_KPAD + 6LL * MLP_WIDTH);
  }
  PDS_HIP(hipEventRecord(ev1_, st_));
  PDS_HIP(hipGetLastError());
  std::vector<float> h(steps);
  PDS_HIP(hipMemcpyAsync(h.data(), losses.p, sizeof(float) * steps, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
  float ms = 0.f;
  PDS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  gemm_ms_ += ms;
  double tot = 0.0;
  for (float v : h) tot += v;
  return tot / (double)S_total;
}

}  // namespace pds
