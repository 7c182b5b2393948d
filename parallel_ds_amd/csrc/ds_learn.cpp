// Learned warm start: training the network on bases the batched GPU solver
// produced, and starting a batch from the network's predictions.
#include "ds_learn.hpp"

#include <chrono>
#include <vector>

#include "ds_rng.hpp"
#include "ds_warm.hpp"

namespace pds {

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// Family prior (feature 16).  When every LP of the batch has the same m and n
// (instances of one family, sharing A), feature 16 of variable j is the
// fraction of the training bases in which j is basic.  A training row sees the
// fraction over the other training instances (leave one out), so the network
// learns how far to trust it; a prediction sees the fraction over all of them.
static bool family_shape(const BatchSolver& B, int& N) {
  const std::vector<int>& m = B.host_m();
  const std::vector<int>& n = B.host_n();
  if (m.empty()) return false;
  for (size_t k = 1; k < m.size(); ++k)
    if (m[k] != m[0] || n[k] != n[0]) return false;
  N = m[0] + n[0];
  return true;
}

double fit_from_batch(Mlp& net, BatchSolver& B, WarmBuffers& wb, int epochs, int minibatch, float lr,
                      unsigned long long seed, double* epoch_loss) {
  if (net.n_features() != WS_NF) throw std::invalid_argument("network must take the batch features (17)");
  if (epochs <= 0 || minibatch <= 0) throw std::invalid_argument("epochs and minibatch must be positive");
  const long long R = B.vars();
  if (R <= 0 || R > 0x7fffffffLL) throw std::invalid_argument("training set size out of range");
  wb.F.alloc((size_t)R * WS_NF);
  wb.y.alloc(R);
  wb.mu.alloc(WS_NF);
  wb.sd.alloc(WS_NF);
  B.features(wb.F.p);
  B.labels(wb.y.p);
  int N = 0;
  std::vector<float> prior;
  if (B.count() > 1 && family_shape(B, N)) {
    // Label counts are integers below 2^24: exact in FP32.
    const int T = B.count();
    wb.prior.alloc(N);
    launch_prior_count(B.stream(), wb.y.p, T, N, wb.prior.p);
    launch_put_prior(B.stream(), wb.F.p, wb.prior.p, wb.y.p, R, N, (float)(T - 1));
    PDS_HIP(hipGetLastError());
    std::vector<float> cnt(N);
    PDS_HIP(hipMemcpyAsync(cnt.data(), wb.prior.p, sizeof(float) * N, hipMemcpyDeviceToHost, B.stream()));
    PDS_HIP(hipStreamSynchronize(B.stream()));
    prior.resize(N);
    for (int j = 0; j < N; ++j) prior[j] = cnt[j] / (float)T;
  }
  net.set_prior(prior);
  launch_colstats(B.stream(), wb.F.p, R, wb.mu.p, wb.sd.p);
  PDS_HIP(hipGetLastError());
  std::vector<float> mu(WS_NF), sd(WS_NF);
  PDS_HIP(hipMemcpyAsync(mu.data(), wb.mu.p, sizeof(float) * WS_NF, hipMemcpyDeviceToHost, B.stream()));
  PDS_HIP(hipMemcpyAsync(sd.data(), wb.sd.p, sizeof(float) * WS_NF, hipMemcpyDeviceToHost, B.stream()));
  PDS_HIP(hipStreamSynchronize(B.stream()));
  net.set_norm(mu.data(), sd.data());
  // Shuffled row order per epoch (Fisher-Yates on SplitMix64).
  std::vector<int> order(R);
  wb.order.alloc(R);
  SplitMix rng(seed);
  const double t0 = net.gemm_ms();
  for (int e = 0; e < epochs; ++e) {
    for (long long i = 0; i < R; ++i) order[i] = (int)i;
    for (long long i = R - 1; i > 0; --i) {
      const long long j = (long long)(rng.uniform() * (double)(i + 1));
      std::swap(order[i], order[j < i ? j : i]);
    }
    PDS_HIP(hipMemcpyAsync(wb.order.p, order.data(), sizeof(int) * R, hipMemcpyHostToDevice, net.stream()));
    const double loss = net.train_epoch(wb.F.p, wb.order.p, (int)R, wb.y.p, minibatch, lr);
    if (epoch_loss) epoch_loss[e] = loss;
  }
  return net.gemm_ms() - t0;
}

double warm_start_batch(Mlp& net, BatchSolver& B, WarmBuffers& wb, double* infer_ms) {
  if (net.n_features() != WS_NF) throw std::invalid_argument("network must take the batch features (17)");
  const long long R = B.vars();
  if (R <= 0 || R > 0x7fffffffLL) throw std::invalid_argument("batch size out of range");
  wb.F.alloc((size_t)R * WS_NF);
  wb.z.alloc(R);
  const double t0 = now_ms();
  B.features(wb.F.p);
  int N = 0;
  const std::vector<float>& pr = net.prior();
  if (!pr.empty() && family_shape(B, N) && N == (int)pr.size()) {
    wb.prior.alloc(N);
    PDS_HIP(hipMemcpyAsync(wb.prior.p, pr.data(), sizeof(float) * N, hipMemcpyHostToDevice, B.stream()));
    launch_put_prior(B.stream(), wb.F.p, wb.prior.p, nullptr, R, N, 1.f);
    PDS_HIP(hipGetLastError());
  }
  PDS_HIP(hipStreamSynchronize(B.stream()));
  net.infer(wb.F.p, nullptr, (int)R, wb.z.p);  // synchronises its stream
  B.select_warm(wb.z.p);
  PDS_HIP(hipStreamSynchronize(B.stream()));
  if (infer_ms) *infer_ms = net.last_infer_ms();
  return now_ms() - t0;
}

}  // namespace pds
