// extern "C" boundary of libpds.so: argument checks, exception-to-error-code
// translation, and the host/device copies of the hook-level entries.
#include <algorithm>
#include <cstring>
#include <string>

#include "ds_abi.h"
#include "ds_batch_host.hpp"
#include "ds_kernels.hpp"
#include "ds_learn.hpp"
#include "ds_solver.hpp"
#include "ds_warm.hpp"

#include "ds_ctx.hpp"
#include "ds_mps.hpp"

namespace pds {
int chuzr_noalloc_setup(int device);  // ds_chuzr.hip
}

struct ds_mlp {
  pds::Mlp* net = nullptr;
  std::string err;
};

namespace pds {
void launch_gemm_tile_check(hipStream_t s, double* C, const double* P, const double* B, int nb, int use_mfma);
}

namespace {

template <typename F>
int guard(ds_ctx* ctx, F&& f) {
  if (!ctx || !ctx->solver) return DS_EINVAL;
  try {
    f(*ctx->solver);
    return DS_OK;
  } catch (const pds::HipError& e) {
    ctx->err = e.what();
    return DS_EHIP;
  } catch (const pds::SingularError& e) {
    ctx->err = e.what();
    return DS_ESINGULAR;
  } catch (const pds::Unsupported& e) {
    ctx->err = e.what();
    return DS_EUNSUPPORTED;
  } catch (const std::invalid_argument& e) {
    ctx->err = e.what();
    return DS_EINVAL;
  } catch (const std::logic_error& e) {
    ctx->err = e.what();
    return DS_ESTATE;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return DS_EINTERNAL;
  }
}

template <typename F>
int mguard(ds_mlp* m, F&& f) {
  if (!m || !m->net) return DS_EINVAL;
  try {
    f(*m->net);
    return DS_OK;
  } catch (const pds::HipError& e) {
    m->err = e.what();
    return DS_EHIP;
  } catch (const std::invalid_argument& e) {
    m->err = e.what();
    return DS_EINVAL;
  } catch (const std::logic_error& e) {
    m->err = e.what();
    return DS_ESTATE;
  } catch (const std::exception& e) {
    m->err = e.what();
    return DS_EINTERNAL;
  }
}

template <typename F>
int bguard(ds_ctx* ctx, F&& f) {
  if (!ctx) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver&) {
    if (!ctx->batch) ctx->batch = new pds::BatchSolver(ctx->device);
    f(*ctx->batch);
  });
}
}  // namespace

extern "C" {

void ds_default_params(ds_params* p) {
  if (!p) return;
  p->refactor_freq = 150;
  p->max_iterations = 0;
  p->ratio_test = 0;
  p->perturb = 1;
  p->perturb_seed = 0x5eedULL;
  p->chunk = 32;
  p->verbose = 0;
  p->mark_iteration = 0;
  p->time_kernels = 0;
  p->cleanup = 1;
  p->profile = 0;
  p->force_event = DS_FORCE_NONE;
  p->force_at = 0;
  p->time_limit_ms = 0.0;
  p->candidates = 1;
  p->candidate_psi = 0.9;
  p->warm_primal = 0;
  p->warm_shift = 1e-3;
  p->warm_primal_cap = 0;
  p->split_launches = 0;
}

int ds_create(int device, ds_ctx** out) {
  if (!out) return DS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DS_EHIP;
  if (device < 0 || device >= ndev) return DS_EINVAL;
  ds_ctx* c = new ds_ctx();
  try {
    c->solver = new pds::Solver(device);  // makes `device` current
    c->device = device;
  } catch (const std::exception& e) {
    delete c;
    return DS_EHIP;
  }
  if (pds::chuzr_noalloc_setup(device) != DS_OK) {
    delete c->solver;
    delete c;
    return DS_EHIP;
  }
  *out = c;
  return DS_OK;
}

int ds_destroy(ds_ctx* ctx) {
  if (!ctx) return DS_EINVAL;
  if (ctx->chuzr_part_w) (void)hipFree(ctx->chuzr_part_w);
  if (ctx->chuzr_part_i) (void)hipFree(ctx->chuzr_part_i);
  delete ctx->wb;
  delete ctx->batch;
  delete ctx->solver;
  delete ctx;
  return DS_OK;
}

const char* ds_last_error(ds_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ds_load_problem(ds_ctx* ctx, int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                    const double* costs, const double* lower_bound, const double* upper_bound,
                    const int* bound_type, const double* lower_range, const double* upper_range,
                    const int* range_type) {
  if (m <= 0 || n <= 0 || !col_ptr || !costs || !lower_bound || !upper_bound || !bound_type || !lower_range ||
      !upper_range || !range_type)
    return DS_EINVAL;
  if (col_ptr[n] > 0 && (!row_idx || !val)) return DS_EINVAL;
  // Problem::checkConstraints (problem.tpp:230-239): more rows than columns is rejected.
  if (m > n + m) return DS_EINVAL;
  for (int j = 0; j < n; ++j)
    if (col_ptr[j + 1] < col_ptr[j]) return DS_EINVAL;
  for (int e = 0; e < col_ptr[n]; ++e)
    if (row_idx[e] < 0 || row_idx[e] >= m) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) {
    pds::HostModel hm;
    hm.build(m, n, col_ptr, row_idx, val, costs, lower_bound, upper_bound, bound_type, lower_range, upper_range,
             range_type);
    s.load(hm);
  });
}

int ds_set_params(ds_ctx* ctx, const ds_params* p) {
  if (!p) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.set_params(*p); });
}

int ds_init(ds_ctx* ctx) {
  return guard(ctx, [&](pds::Solver& s) { s.initDualSimplex(); });
}

int ds_time_refactor(ds_ctx* ctx, int reps, double* refactor_ms, double* gemm_ms, double* gemm_flops, int* k) {
  if (reps <= 0 || !refactor_ms || !gemm_ms || !gemm_flops || !k) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.time_refactor(reps, *refactor_ms, *gemm_ms, *gemm_flops, *k); });
}

int ds_load_mps(ds_ctx* ctx, const char* path, int* m, int* n) {
  if (!path) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) {
    pds::MpsModel M;
    try {
      pds::read_mps(path, M);
    } catch (const std::runtime_error& e) {
      throw std::invalid_argument(e.what());
    }
    pds::HostModel hm;
    hm.build(M.m, M.n, M.col_ptr.data(), M.row_idx.data(), M.val.data(), M.costs.data(), M.lower_bound.data(),
             M.upper_bound.data(), M.bound_type.data(), M.lower_range.data(), M.upper_range.data(),
             M.range_type.data());
    hm.offset = M.offset;
    s.load(hm);
    if (m) *m = M.m;
    if (n) *n = M.n;
  });
}

int ds_set_offset(ds_ctx* ctx, double offset) {
  return guard(ctx, [&](pds::Solver& s) { s.set_offset(offset); });
}

struct ds_mps {
  pds::MpsModel M;
};

int ds_mps_read(const char* path, ds_mps** out, char* err, int err_len) {
  if (!path || !out) return DS_EINVAL;
  *out = nullptr;
  ds_mps* h = new ds_mps();
  try {
    pds::read_mps(path, h->M);
  } catch (const std::exception& e) {
    if (err && err_len > 0) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = '\0';
    }
    delete h;
    return DS_EINVAL;
  }
  *out = h;
  return DS_OK;
}

int ds_mps_dims(const ds_mps* mps, int* m, int* n, int* nnz) {
  if (!mps) return DS_EINVAL;
  if (m) *m = mps->M.m;
  if (n) *n = mps->M.n;
  if (nnz) *nnz = (int)mps->M.val.size();
  return DS_OK;
}

int ds_mps_get(const ds_mps* mps, int* col_ptr, int* row_idx, double* val, double* costs, double* lower_bound,
               double* upper_bound, int* bound_type, double* lower_range, double* upper_range, int* range_type,
               double* offset) {
  if (!mps) return DS_EINVAL;
  const pds::MpsModel& M = mps->M;
  auto cp = [](auto* dst, const auto& v) {
    if (dst) std::copy(v.begin(), v.end(), dst);
  };
  cp(col_ptr, M.col_ptr);
  cp(row_idx, M.row_idx);
  cp(val, M.val);
  cp(costs, M.costs);
  cp(lower_bound, M.lower_bound);
  cp(upper_bound, M.upper_bound);
  cp(bound_type, M.bound_type);
  cp(lower_range, M.lower_range);
  cp(upper_range, M.upper_range);
  cp(range_type, M.range_type);
  if (offset) *offset = M.offset;
  return DS_OK;
}

int ds_mps_free(ds_mps* mps) {
  delete mps;
  return DS_OK;
}

int ds_init_basis(ds_ctx* ctx, const int* basis) {
  if (!basis) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.init_basis(basis); });
}

int ds_presolve(ds_ctx* ctx, const char* method, int* phase1_status) {
  if (!method || std::strcmp(method, "minInfeas") != 0) return DS_EINVAL;  // stringToPreSolverMethod
  return guard(ctx, [&](pds::Solver& s) {
    int st = s.presolve();
    if (phase1_status) *phase1_status = st;
  });
}

int ds_solve(ds_ctx* ctx) {
  return guard(ctx, [&](pds::Solver& s) { s.solve(); });
}

int ds_get_result(ds_ctx* ctx, ds_result* r) {
  if (!r) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.result(*r); });
}

int ds_get_solution(ds_ctx* ctx, double* Z, int* solved, char* message, int message_len, double* x) {
  return guard(ctx, [&](pds::Solver& s) {
    double z;
    int sv;
    std::string msg;
    std::vector<double> xv;
    s.solution(z, sv, msg, xv);
    if (Z) *Z = z;
    if (solved) *solved = sv;
    if (message && message_len > 0) {
      std::strncpy(message, msg.c_str(), message_len - 1);
      message[message_len - 1] = 0;
    }
    if (x) std::memcpy(x, xv.data(), sizeof(double) * xv.size());
  });
}

int ds_get_basis(ds_ctx* ctx, int* basis) {
  if (!basis) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) {
    auto b = s.basis_host();
    std::memcpy(basis, b.data(), sizeof(int) * b.size());
  });
}

int ds_get_nonbasis(ds_ctx* ctx, int* nonbasis) {
  if (!nonbasis) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) {
    auto b = s.nonbasis_host();
    std::memcpy(nonbasis, b.data(), sizeof(int) * b.size());
  });
}

int ds_solve_lp(ds_ctx* ctx, int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                const double* costs, const double* lower_bound, const double* upper_bound, const int* bound_type,
                const double* lower_range, const double* upper_range, const int* range_type, const ds_params* p,
                ds_result* r, double* x, int* basis) {
  int rc = ds_load_problem(ctx, m, n, col_ptr, row_idx, val, costs, lower_bound, upper_bound, bound_type,
                           lower_range, upper_range, range_type);
  if (rc) return rc;
  if (p && (rc = ds_set_params(ctx, p))) return rc;
  if ((rc = ds_init(ctx))) return rc;
  int st = 0;
  if ((rc = ds_presolve(ctx, "minInfeas", &st))) return rc;
  if (st == 0 && (rc = ds_solve(ctx))) return rc;
  if (r && (rc = ds_get_result(ctx, r))) return rc;
  if (x && (rc = ds_get_solution(ctx, nullptr, nullptr, nullptr, 0, x))) return rc;
  if (basis && (rc = ds_get_basis(ctx, basis))) return rc;
  return DS_OK;
}

int ds_set_basis(ds_ctx* ctx, const int* basis) {
  if (!basis) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.set_basis(basis); });
}

int ds_refactor(ds_ctx* ctx) {
  return guard(ctx, [&](pds::Solver& s) { s.refactor(true); });
}

int ds_btran(ds_ctx* ctx, int p_idx, double* rho) {
  if (!rho) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.btran(p_idx, rho); });
}

int ds_ftran(ds_ctx* ctx, const double* b, double* x) {
  if (!b || !x) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.ftran(b, x); });
}

int ds_ftran_column(ds_ctx* ctx, int j, double* x) {
  if (!x) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.ftran_column(j, x); });
}

int ds_pivot_row(ds_ctx* ctx, const double* rho, double* alpha) {
  if (!rho || !alpha) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.pivot_row(rho, alpha); });
}

int ds_time_pivot_row(ds_ctx* ctx, const double* rho, int reps, double* ms, double* bytes) {
  if (!rho || !ms || !bytes || reps <= 0) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.time_pivot_row(rho, reps, *ms, *bytes); });
}

int ds_time_pivot_row_ext(ds_ctx* ctx, const double* rho, int reps, long long flush_bytes, double* warm_ms,
                          double* cold_ms, double* bytes) {
  if (!rho || !warm_ms || !cold_ms || !bytes || reps <= 0 || flush_bytes < 0) return DS_EINVAL;
  return guard(ctx, [&](pds::Solver& s) { s.time_pivot_row_ext(rho, reps, flush_bytes, *warm_ms, *cold_ms, *bytes); });
}

void ds_batch_default_params(ds_batch_params* p) {
  if (!p) return;
  p->refactor_freq = 150;
  p->max_iterations = 200000;
  p->perturb = 1;
  p->perturb_seed = 0x5eedULL;
  p->block = 256;
  p->lds_inv_max_m = 64;
  p->wgs_per_cu = 2;
  p->use_graph = 1;
  p->cleanup = 1;
}

int ds_batch_load(ds_ctx* ctx, int count, const int* m, const int* n, const long long* nnz_off, const int* col_ptr,
                  const int* row_idx, const double* val, const double* b, const double* c) {
  if (count <= 0 || !m || !n || !nnz_off || !col_ptr || !row_idx || !val || !b || !c) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) { B.load(count, m, n, nnz_off, col_ptr, row_idx, val, b, c); });
}

int ds_batch_set_warm(ds_ctx* ctx, const int* basis) {
  return bguard(ctx, [&](pds::BatchSolver& B) { B.set_warm(basis); });
}

int ds_batch_set_bases(ds_ctx* ctx, const int* basis) {
  if (!basis) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) { B.set_bases(basis); });
}

int ds_batch_solve(ds_ctx* ctx, const ds_batch_params* p, double* ms) {
  if (!p) return DS_EINVAL;
  if (p->block != 128 && p->block != 256 && p->block != 512) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    if (B.count() == 0) throw std::logic_error("no batch loaded");
    double t = B.solve(*p);
    if (ms) *ms = t;
  });
}

int ds_batch_replay(ds_ctx* ctx, double* ms) {
  return bguard(ctx, [&](pds::BatchSolver& B) {
    double t = B.replay();
    if (ms) *ms = t;
  });
}

int ds_batch_results(ds_ctx* ctx, int* status, double* obj, int* iters, int* basis, double* x) {
  return bguard(ctx, [&](pds::BatchSolver& B) { B.results(status, obj, iters, basis, x); });
}

int ds_batch_tune(ds_ctx* ctx, int sample, ds_batch_params* best, double* best_ms, int* tried) {
  if (!best || sample <= 0) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    double ms = 0;
    int t = 0;
    B.tune(sample, *best, ms, t);
    if (best_ms) *best_ms = ms;
    if (tried) *tried = t;
  });
}

int ds_mlp_create(int device, int n_features, uint64_t seed, ds_mlp** out) {
  if (!out) return DS_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DS_EHIP;
  if (device < 0 || device >= ndev || n_features <= 0 || n_features > pds::MLP_KPAD) return DS_EINVAL;
  ds_mlp* m = new ds_mlp;
  try {
    m->net = new pds::Mlp(device, n_features, seed);
  } catch (const std::exception&) {
    delete m;
    return DS_EHIP;
  }
  *out = m;
  return DS_OK;
}

int ds_mlp_destroy(ds_mlp* net) {
  if (!net) return DS_EINVAL;
  delete net->net;
  delete net;
  return DS_OK;
}

const char* ds_mlp_last_error(ds_mlp* net) { return net ? net->err.c_str() : "null network"; }

long long ds_mlp_param_count(void) { return (long long)pds::MlpLayout::total; }

int ds_mlp_get_params(ds_mlp* net, float* params) {
  if (!params) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) { M.get_params(params); });
}

int ds_mlp_set_params(ds_mlp* net, const float* params) {
  if (!params) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) { M.set_params(params); });
}

int ds_mlp_set_norm(ds_mlp* net, const float* mu, const float* sd) {
  if (!mu || !sd) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) { M.set_norm(mu, sd); });
}

int ds_mlp_get_norm(ds_mlp* net, float* mu, float* sd) {
  if (!mu || !sd) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) { M.get_norm(mu, sd); });
}

int ds_mlp_get_prior(ds_mlp* net, float* out, int cap, int* n) {
  if (!n || cap < 0 || (cap > 0 && !out)) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) {
    const std::vector<float>& p = M.prior();
    *n = (int)p.size();
    for (int j = 0; j < cap && j < (int)p.size(); ++j) out[j] = p[j];
  });
}

int ds_mlp_infer(ds_mlp* net, int S, const float* features, float* logits) {
  if (S <= 0 || !features || !logits) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) {
    pds::DBuf<float> F, z;
    F.alloc((size_t)S * M.n_features());
    z.alloc(S);
    PDS_HIP(hipMemcpyAsync(F.p, features, sizeof(float) * S * M.n_features(), hipMemcpyHostToDevice, M.stream()));
    M.infer(F.p, nullptr, S, z.p);
    PDS_HIP(hipMemcpyAsync(logits, z.p, sizeof(float) * S, hipMemcpyDeviceToHost, M.stream()));
    PDS_HIP(hipStreamSynchronize(M.stream()));
  });
}

int ds_mlp_grad(ds_mlp* net, int S, const float* features, const float* labels, const float* weights, double* loss,
                float* grad) {
  if (S <= 0 || !features || !labels || !grad) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) {
    pds::DBuf<float> F, y, w;
    F.alloc((size_t)S * M.n_features());
    y.alloc(S);
    PDS_HIP(hipMemcpyAsync(F.p, features, sizeof(float) * S * M.n_features(), hipMemcpyHostToDevice, M.stream()));
    PDS_HIP(hipMemcpyAsync(y.p, labels, sizeof(float) * S, hipMemcpyHostToDevice, M.stream()));
    if (weights) {
      w.alloc(S);
      PDS_HIP(hipMemcpyAsync(w.p, weights, sizeof(float) * S, hipMemcpyHostToDevice, M.stream()));
    }
    const double l = M.grad(F.p, nullptr, S, y.p, weights ? w.p : nullptr);
    if (loss) *loss = l;
    PDS_HIP(hipMemcpyAsync(grad, M.grad_dev(), sizeof(float) * pds::MlpLayout::total, hipMemcpyDeviceToHost,
                           M.stream()));
    PDS_HIP(hipStreamSynchronize(M.stream()));
  });
}

int ds_mlp_adam_step(ds_mlp* net, float lr) {
  return mguard(net, [&](pds::Mlp& M) {
    M.adam_step(lr);
    PDS_HIP(hipStreamSynchronize(M.stream()));
  });
}

int ds_mlp_fit(ds_mlp* net, ds_ctx* batch, int epochs, int minibatch, float lr, uint64_t seed, double* epoch_loss,
               double* ms) {
  if (!net || !net->net || !batch || !batch->batch) return DS_EINVAL;
  if (epochs <= 0 || minibatch <= 0 || !(lr > 0.f)) return DS_EINVAL;
  return mguard(net, [&](pds::Mlp& M) {
    if (!batch->wb) batch->wb = new pds::WarmBuffers;
    const double t = pds::fit_from_batch(M, *batch->batch, *batch->wb, epochs, minibatch, lr, seed, epoch_loss);
    if (ms) *ms = t;
  });
}

int ds_mlp_train_stats(ds_mlp* net, double* ms, long long* flops) {
  return mguard(net, [&](pds::Mlp& M) {
    if (ms) *ms = M.gemm_ms();
    if (flops) *flops = M.gemm_flops();
  });
}

int ds_batch_features(ds_ctx* ctx, float* features) {
  if (!features) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    if (B.count() == 0) throw std::logic_error("no batch loaded");
    pds::DBuf<float> F;
    F.alloc((size_t)B.vars() * pds::WS_NF);
    B.features(F.p);
    PDS_HIP(hipMemcpyAsync(features, F.p, sizeof(float) * B.vars() * pds::WS_NF, hipMemcpyDeviceToHost,
                           B.stream()));
    PDS_HIP(hipStreamSynchronize(B.stream()));
  });
}

int ds_batch_labels(ds_ctx* ctx, float* labels) {
  if (!labels) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    if (B.count() == 0) throw std::logic_error("no batch loaded");
    pds::DBuf<float> y;
    y.alloc(B.vars());
    B.labels(y.p);
    PDS_HIP(hipMemcpyAsync(labels, y.p, sizeof(float) * B.vars(), hipMemcpyDeviceToHost, B.stream()));
    PDS_HIP(hipStreamSynchronize(B.stream()));
  });
}

int ds_batch_warm_from_scores(ds_ctx* ctx, const float* scores) {
  if (!scores) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    if (B.count() == 0) throw std::logic_error("no batch loaded");
    if (!ctx->wb) ctx->wb = new pds::WarmBuffers;
    ctx->wb->z.alloc(B.vars());
    PDS_HIP(hipMemcpyAsync(ctx->wb->z.p, scores, sizeof(float) * B.vars(), hipMemcpyHostToDevice, B.stream()));
    B.select_warm(ctx->wb->z.p);
    PDS_HIP(hipStreamSynchronize(B.stream()));
  });
}

int ds_batch_warm_start(ds_ctx* ctx, ds_mlp* net, double* ms, double* infer_ms) {
  if (!net || !net->net) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) {
    if (B.count() == 0) throw std::logic_error("no batch loaded");
    if (!ctx->wb) ctx->wb = new pds::WarmBuffers;
    const double t = pds::warm_start_batch(*net->net, B, *ctx->wb, infer_ms);
    if (ms) *ms = t;
  });
}

int ds_batch_get_warm(ds_ctx* ctx, int* basis) {
  if (!basis) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) { B.get_warm(basis); });
}

int ds_batch_set_repair_tol(ds_ctx* ctx, double tol) {
  if (!(tol >= 0.0)) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) { B.set_repair_tol(tol); });
}

int ds_batch_repairs(ds_ctx* ctx, int* count) {
  if (!count) return DS_EINVAL;
  return bguard(ctx, [&](pds::BatchSolver& B) { B.repairs(count); });
}

int ds_debug_copy(long long bytes, double* ms) {
  if (bytes <= 0 || bytes % 16 != 0) return DS_EINVAL;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return DS_EHIP;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return DS_EHIP;
  }
  int rc = DS_OK;
  hipEvent_t e0, e1;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = DS_EHIP;
  } else {
    (void)hipEventRecord(e0, nullptr);
    pds::launch_copy(nullptr, (const double2*)a, (double2*)b, bytes / 16);
    (void)hipEventRecord(e1, nullptr);
    float t = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = DS_EHIP;
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

int ds_debug_gemm_tile(double* C, const double* P, const double* B, int nb, int use_mfma) {
  if (!C || !P || !B || nb <= 0 || nb > pds::GJ_NB) return DS_EINVAL;
  double *dC = nullptr, *dP = nullptr, *dB = nullptr;
  const size_t sc = 64 * 64 * sizeof(double), sp = 64 * pds::GJ_NB * sizeof(double);
  if (hipMalloc((void**)&dC, sc) != hipSuccess) return DS_EHIP;
  if (hipMalloc((void**)&dP, sp) != hipSuccess) {
    (void)hipFree(dC);
    return DS_EHIP;
  }
  if (hipMalloc((void**)&dB, sp) != hipSuccess) {
    (void)hipFree(dC);
    (void)hipFree(dP);
    return DS_EHIP;
  }
  int rc = DS_OK;
  if (hipMemcpy(dC, C, sc, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dP, P, sp, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dB, B, nb * 64 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    rc = DS_EHIP;
  } else {
    pds::launch_gemm_tile_check(nullptr, dC, dP, dB, nb, use_mfma);
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(C, dC, sc, hipMemcpyDeviceToHost) != hipSuccess)
      rc = DS_EHIP;
  }
  (void)hipFree(dC);
  (void)hipFree(dP);
  (void)hipFree(dB);
  return rc;
}

}  // extern "C"
