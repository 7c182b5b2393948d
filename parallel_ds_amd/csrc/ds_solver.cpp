// Host driver of the MI355X dual simplex.
//
// The control flow is that of the reference's sequential solver
// (cpu/sequentialDualSimplex.cpp:253-515 for phase 2,
// common/baseDualSimplex.tpp:799-934 for phase 1); the arithmetic of every
// iteration runs in gfx950 kernels on device-resident state.  The host queues
// up to `chunk` iterations at a time without synchronising; a kernel that
// meets an event (optimality, restart, refactor request, ...) raises the
// device status word, the remaining queued kernels return at entry, and the
// host handles the event exactly where the reference would.
#include "ds_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ds_rng.hpp"

namespace pds {

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void HostModel::build(int m_, int n_, const int* col_ptr, const int* row_idx, const double* val,
                      const double* costs, const double* lb, const double* ub, const int* bound_type,
                      const double* lr, const double* ur, const int* range_type) {
  m = m_;
  n = n_;
  N = n + m;
  const int nnzA = col_ptr[n];
  cptr.assign(N + 1, 0);
  ridx.resize(nnzA + m);
  cval.resize(nnzA + m);
  for (int j = 0; j <= n; ++j) cptr[j] = col_ptr[j];
  for (int e = 0; e < nnzA; ++e) {
    ridx[e] = row_idx[e];
    cval[e] = val[e];
  }
  for (int i = 0; i < m; ++i) {
    ridx[nnzA + i] = i;
    cval[nnzA + i] = 1.0;
    cptr[n + i + 1] = nnzA + i + 1;
  }
  rptr.assign(m + 1, 0);
  for (int e = 0; e < nnzA + m; ++e) rptr[ridx[e] + 1]++;
  for (int i = 0; i < m; ++i) rptr[i + 1] += rptr[i];
  cidx.resize(nnzA + m);
  rval.resize(nnzA + m);
  std::vector<int> fill(rptr.begin(), rptr.end() - 1);
  for (int j = 0; j < N; ++j)
    for (int e = cptr[j]; e < cptr[j + 1]; ++e) {
      int r = ridx[e];
      cidx[fill[r]] = j;
      rval[fill[r]] = cval[e];
      fill[r]++;
    }
  c.resize(N);
  l.resize(N);
  u.resize(N);
  bt.resize(N);
  for (int j = 0; j < n; ++j) {
    c[j] = costs[j];
    l[j] = lb[j];
    u[j] = ub[j];
    bt[j] = (int8_t)bound_type[j];
  }
  // Slack of row i: bounds [-upper_range, -lower_range]; a <= row (range
  // type Upper) gives a Lower slack, >= gives Upper, = Fixed, range Boxed.
  for (int i = 0; i < m; ++i) {
    c[n + i] = 0.0;
    u[n + i] = -lr[i];
    l[n + i] = -ur[i];
    int8_t t;
    switch (range_type[i]) {
      case DS_FIXED: t = BT_FIXED; break;
      case DS_BOXED: t = BT_BOXED; break;
      case DS_LOWER: t = BT_UPPER; break;
      case DS_UPPER: t = BT_LOWER; break;
      default: t = BT_FREE; break;
    }
    bt[n + i] = t;
  }
}

Solver::Solver(int device) : device_(device) {
  PDS_HIP(hipSetDevice(device_));
  PDS_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  PDS_HIP(hipHostMalloc((void**)&hS_, sizeof(DevScalars), hipHostMallocDefault));
  std::memset(hS_, 0, sizeof(DevScalars));
  ds_default_params(&P_);
}

Solver::~Solver() {
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  if (st_) {
    (void)hipStreamSynchronize(st_);
    (void)hipStreamDestroy(st_);
  }
  if (hS_) (void)hipHostFree(hS_);
}

void Solver::set_params(const ds_params& p) {
  P_ = p;
  if (P_.refactor_freq <= 0) P_.refactor_freq = 150;
  if (P_.refactor_freq > ETA_MAX - 8) throw std::invalid_argument("refactor_freq exceeds the eta file capacity");
  if (P_.chunk <= 0) P_.chunk = 1;
  if (P_.candidates < 1 || P_.candidates > CAND_MAX) throw std::invalid_argument("candidates must be 1..8");
  if (P_.candidate_psi < 0.0 || P_.candidate_psi > 1.0) throw std::invalid_argument("candidate_psi must be in [0, 1]");
  if (P_.candidates > 1 && loaded_ && rho_c_.n < (size_t)CAND_MAX * H_.m) rho_c_.alloc((size_t)CAND_MAX * H_.m);
  if (P_.time_kernels && (int)ev_.size() < 2 * P_.chunk) {
    for (int i = (int)ev_.size(); i < 2 * P_.chunk; ++i) {
      hipEvent_t e;
      PDS_HIP(hipEventCreate(&e));
      ev_.push_back(e);
    }
    ev_iter_.assign(P_.chunk, 0);
  }
  if (P_.profile && pev_.empty()) {
    for (int i = 0; i < 16; ++i) {
      hipEvent_t e;
      PDS_HIP(hipEventCreate(&e));
      pev_.push_back(e);
    }
  }
}

// Per-step profile: events recorded between the steps of one queued iteration;
// after the host synchronises, the interval up to each event goes to its step.
void Solver::prof_event(int part) {
  if (!P_.profile) return;
  const int k = (int)pev_part_.size();
  if (k >= (int)pev_.size()) return;
  PDS_HIP(hipEventRecord(pev_[k], st_));
  pev_part_.push_back(part);
}
void Solver::prof_harvest() {
  if (!P_.profile || pev_part_.empty()) return;
  PDS_HIP(hipEventSynchronize(pev_[pev_part_.size() - 1]));
  for (size_t k = 1; k < pev_part_.size(); ++k) {
    float ms = 0.f;
    PDS_HIP(hipEventElapsedTime(&ms, pev_[k - 1], pev_[k]));
    step_ms_[pev_part_[k]] += ms;
  }
  pev_part_.clear();
}
void Solver::prof_host_begin() {
  if (!P_.profile) return;
  sync();
  prof_t0_ = now_ms();
}
void Solver::prof_host_end(int part) {
  if (!P_.profile) return;
  sync();
  step_ms_[part] += now_ms() - prof_t0_;
}

// Sum the SpMV event times of the queued iterations that ran (<= last_done).
void Solver::harvest_events(int last_done) {
  if (!P_.time_kernels) return;
  for (int i = 0; i < ev_used_; ++i) {
    if (ev_iter_[i] > last_done) continue;
    float ms = 0.f;
    PDS_HIP(hipEventElapsedTime(&ms, ev_[2 * i], ev_[2 * i + 1]));
    spmv_ms_ += ms;
    spmv_launches_++;
  }
  ev_used_ = 0;
}

double Solver::nnz_nonbasic() {
  std::vector<int> nb = nonbasis_host();
  double s = 0.0;
  for (int j : nb) s += H_.cptr[j + 1] - H_.cptr[j];
  return s;
}

void Solver::maybe_mark(int done) {
  if (P_.mark_iteration > 0 && !marked_ && done >= P_.mark_iteration) {
    marked_ = true;
    mark_ms_ = now_ms() - t_start_;
    nnzN_mark_ = nnz_nonbasic();
  }
}

DevModel Solver::M() const {
  DevModel m;
  m.m = H_.m;
  m.n = H_.n;
  m.N = H_.N;
  m.nnz = (int)H_.ridx.size();
  m.cptr = cptr_.p;
  m.ridx = ridx_.p;
  m.cval = cval_.p;
  m.rptr = rptr_.p;
  m.cidx = cidx_.p;
  m.rval = rval_.p;
  m.c = c_.p;
  m.l = l_.p;
  m.u = u_.p;
  m.bt = bt_.p;
  return m;
}

DevIter Solver::I() const {
  DevIter it;
  it.x = x_.p;
  it.d = d_.p;
  it.beta = beta_.p;
  it.basis = basis_.p;
  it.nonbasis = nonbasis_.p;
  it.posN = posN_.p;
  it.f = f_.p;
  it.rho = rho_.p;
  it.alpha_r = alpha_r_.p;
  it.alpha_q = alpha_q_.p;
  it.tau = tau_.p;
  it.devex = devex_.p;
  it.w1 = w1_.p;
  it.w2 = w2_.p;
  it.wN = wN_.p;
  it.etas = etas_.p;
  it.eta_pos = eta_pos_.p;
  it.E = E_.p;
  it.MF = MF_.p;
  it.eta_last = eta_last_.p;
  it.x0 = x0_.p;
  it.cand_i = cand_i_.p;
  it.cand_th = cand_th_.p;
  it.cand_a = cand_a_.p;
  it.bmark = bmark_.p;
  it.blist = blist_.p;
  it.flipc = flipc_.p;
  it.flist = flist_.p;
  it.bcc = bcc_.p;
  it.dxb = dxb_.p;
  it.cand_pos = cand_pos_.p;
  it.cand_w = cand_w_.p;
  it.cand_act = cand_act_.p;
  it.rho_c = rho_c_.p;
  it.eta_cap = eta_cap_;
  it.blocked = blocked_.p;
  it.blocked_list = blocked_list_.p;
  it.list_idx = list_idx_.p;
  it.list_val = list_val_.p;
  it.ylist_row = ylist_row_.p;
  it.ylist_val = ylist_val_.p;
  it.list_n = list_n_.p;
  it.list_cap = (int)list_idx_.n;
  it.S = S_.p;
  it.part_d = part_d_.p;
  it.part_i = part_i_.p;
  it.part_cap = 512;  // workgroups of a reduction launch: each one pays an L2 writeback at its release
  return it;
}

DevFactor Solver::F() const {
  DevFactor f;
  f.k = k_;
  f.Kinv = Kinv_.p;
  f.KinvT = KinvT_.p;
  f.pos_P = pos_P_.p;
  f.row_R = row_R_.p;
  f.Ridx_of_row = Ridx_.p;
  f.Pidx_of_pos = Pidx_.p;
  f.slackpos_of_row = slackpos_.p;
  f.slackrow_of_pos = slackrow_.p;
  f.crp = crp_.p;
  f.cci = cci_.p;
  f.ccv = ccv_.p;
  f.ccp = ccp_.p;
  f.ccr = ccr_.p;
  f.ccv2 = ccv2_.p;
  return f;
}

void Solver::pull() {
  PDS_HIP(hipGetLastError());
  PDS_HIP(hipMemcpyAsync(hS_, S_.p, sizeof(DevScalars), hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}
void Solver::push() {
  PDS_HIP(hipMemcpyAsync(S_.p, hS_, sizeof(DevScalars), hipMemcpyHostToDevice, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}
void Solver::sync() { PDS_HIP(hipStreamSynchronize(st_)); }

void Solver::tick_limit(int it) {
  if (P_.max_iterations > 0 && it > P_.max_iterations) {
    iteration_ = P_.max_iterations;
    throw IterLimit{};
  }
  if (P_.time_limit_ms > 0.0 && now_ms() - t_start_ > P_.time_limit_ms) {
    time_hit_ = true;
    throw IterLimit{};
  }
}

template <typename T>
static void upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(v.size());
  if (!v.empty()) PDS_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

void Solver::load(const HostModel& hm) {
  H_ = hm;
  const int m = H_.m, N = H_.N, n = H_.n;
  const size_t nnz = H_.ridx.size();
  upload(cptr_, H_.cptr, st_);
  upload(ridx_, H_.ridx, st_);
  upload(cval_, H_.cval, st_);
  upload(rptr_, H_.rptr, st_);
  upload(cidx_, H_.cidx, st_);
  upload(rval_, H_.rval, st_);
  upload(l_, H_.l, st_);
  upload(u_, H_.u, st_);
  upload(bt_, H_.bt, st_);
  c_host_ = H_.c;
  c_orig_ = H_.c;
  upload(c_, c_host_, st_);
  eta_cap_ = ETA_MAX;
  x_.alloc(N);
  d_.alloc(N);
  beta_.alloc(m);
  f_.alloc(m);
  rho_.alloc(m);
  alpha_r_.alloc(n);
  alpha_q_.alloc(m);
  tau_.alloc(m);
  w1_.alloc(m);
  w2_.alloc(m);
  wN_.alloc(N);
  etas_.alloc((size_t)eta_cap_ * m);
  E_.alloc((size_t)eta_cap_ * eta_cap_);
  MF_.alloc((size_t)eta_cap_ * eta_cap_);
  eta_last_.alloc(eta_cap_);
  x0_.alloc(m);
  cand_i_.alloc(2048 * 8);
  cand_th_.alloc(2048 * 8);
  cand_a_.alloc(2048 * 8);
  eta_pos_.alloc(eta_cap_);
  part_d_.alloc(4096);
  part_i_.alloc(8192);
  basis_.alloc(m);
  nonbasis_.alloc(n);
  posN_.alloc(N);
  blocked_.alloc(N);
  blocked_list_.alloc(std::max(m, 1));
  list_idx_.alloc(1 << 17);
  list_val_.alloc(1 << 17);
  ylist_row_.alloc(ETA_MAX + 1);
  ylist_val_.alloc(ETA_MAX + 1);
  list_n_.alloc(2);
  S_.alloc(1);
  pos_P_.alloc(m);
  row_R_.alloc(m);
  Ridx_.alloc(m);
  Pidx_.alloc(m);
  slackpos_.alloc(m);
  slackrow_.alloc(m);
  crp_.alloc(m + 1);
  cci_.alloc(nnz);
  ccv_.alloc(nnz);
  ccp_.alloc(m + 1);
  ccr_.alloc(nnz);
  ccv2_.alloc(nnz);
  piv_.alloc(m);
  rowscale_.alloc(m);
  perm_.alloc(m);
  flag_.alloc(2 + 2 * GJ_DEP_CAP);
  Pb_.alloc((size_t)m * GJ_NB);
  Pinv_.alloc((size_t)GJ_NB * GJ_NB);
  rowsA_.alloc((size_t)(m / 256 + 2) * GJ_NB);
  rowsB_.alloc((size_t)(m / 256 + 2) * GJ_NB);
  nA_.alloc(m / 256 + 2);
  nB_.alloc(m / 256 + 2);
  has_boxed_ = false;
  primal_ok_ = true;
  for (int j = 0; j < N; ++j) {
    has_boxed_ = has_boxed_ || H_.bt[j] == BT_BOXED;
    primal_ok_ = primal_ok_ && H_.bt[j] != BT_BOXED && H_.bt[j] != BT_FREE;
  }
  devex_.alloc(N);
  const size_t nb = has_boxed_ ? (size_t)n : 1, mb = has_boxed_ ? (size_t)m : 1;
  bmark_.alloc(nb);
  blist_.alloc(nb);
  flipc_.alloc(nb);
  flist_.alloc(nb);
  bcc_.alloc(mb);
  dxb_.alloc(mb);
  cand_pos_.alloc(CAND_MAX);
  cand_act_.alloc(CAND_MAX);
  cand_w_.alloc(CAND_MAX);
  rho_c_.alloc(P_.candidates > 1 ? (size_t)CAND_MAX * m : 1);
  PDS_HIP(hipMemsetAsync(bmark_.p, 0, nb, st_));
  PDS_HIP(hipMemsetAsync(blocked_.p, 0, N, st_));
  PDS_HIP(hipMemsetAsync(E_.p, 0, sizeof(double) * eta_cap_ * eta_cap_, st_));
  sync();
  loaded_ = true;
  inited_ = false;
}

// Kernel buffers grow with the largest kernel seen (k <= m).
static void ensure_kernel_capacity(DBuf<double>& b, int k) {
  size_t need = (size_t)k * (size_t)k;
  if (need > b.n) {
    size_t cap = std::max(need, (size_t)(b.n * 3 / 2));
    b.alloc(cap);
  }
}

void Solver::refactor(bool force) {
  // No basis change since the last factorization (no eta): a new factor of the
  // same basis would be bit-for-bit the one already held (the kernels are
  // deterministic), so the restart paths that refactor again skip the work.
  if (!force && factor_valid_ && t_ == 0 && !(P_.force_event == DS_FORCE_SINGULAR && !force_fired_)) {
    refactors_skipped_++;
    return;
  }
  factor_valid_ = false;
  const double t0 = now_ms();
  // A numerically singular kernel either ends the call (repair off) or has the
  // slacks of the rows chosen for its dependent columns substituted for those
  // columns, after which the basis is factored again.
  for (int round = 0;; ++round) {
    std::vector<std::pair<int, int>> dep;
    factor_once(dep);
    if (dep.empty()) break;
    if (!repair_ || round >= 8) throw SingularError("singular basis kernel in refactorization");
    substitute_slacks(dep);
  }
  refactors_++;
  refactor_ms_ += now_ms() - t0;
  factor_valid_ = true;
}

// Basis position pos_P[g] of each dependent kernel column g takes the slack of
// the original row whose kernel row the tournament chose for that step; the
// leaving structural takes the slack's place in the nonbasic list.
void Solver::substitute_slacks(const std::vector<std::pair<int, int>>& dep) {
  const int m = H_.m, n = H_.n, N = H_.N;
  std::vector<int> basis = basis_host(), nonbasis = nonbasis_host(), posN(N);
  PDS_HIP(hipMemcpyAsync(posN.data(), posN_.p, sizeof(int) * N, hipMemcpyDeviceToHost, st_));
  sync();
  for (const auto& d : dep) {
    const int pos = fac_pos_P_[d.first];
    const int in = n + fac_row_R_[d.second];
    const int out = basis[pos];
    const int q = posN[in];
    if (q < 0 || out >= n) throw SingularError("slack substitution met an inconsistent basis");
    basis[pos] = in;
    nonbasis[q] = out;
    posN[out] = q;
    posN[in] = -1;
  }
  (void)m;
  upload(basis_, basis, st_);
  upload(nonbasis_, nonbasis, st_);
  upload(posN_, posN, st_);
  sync();
  repairs_ += (int)dep.size();
  repaired_ = true;
}

void Solver::factor_once(std::vector<std::pair<int, int>>& dep) {
  const int m = H_.m, n = H_.n;
  std::vector<int> basis(m);
  PDS_HIP(hipMemcpyAsync(basis.data(), basis_.p, sizeof(int) * m, hipMemcpyDeviceToHost, st_));
  sync();
  std::vector<int> pos_P, row_R, Ridx(m, -1), Pidx(m, -1), slackpos(m, -1), slackrow(m, -1);
  for (int i = 0; i < m; ++i) {
    int j = basis[i];
    if (j >= n) {
      slackpos[j - n] = i;
      slackrow[i] = j - n;
    } else {
      Pidx[i] = (int)pos_P.size();
      pos_P.push_back(i);
    }
  }
  for (int r = 0; r < m; ++r)
    if (slackpos[r] < 0) {
      Ridx[r] = (int)row_R.size();
      row_R.push_back(r);
    }
  const int k = (int)pos_P.size();
  if ((int)row_R.size() != k) throw SingularError("basis is structurally singular");
  // Coupling block C = A[covered rows, basic structurals], by kernel columns (CSC) and by rows
  // (CSR). The CSR is a counting sort of the CSC entries by row, with the columns visited in
  // increasing variable order: each row lists its entries as a scan of A's row would, at
  // O(nnz of the basic structurals) instead of O(nnz of the covered rows).
  std::vector<int> crp(m + 1, 0), ccp(k + 1, 0), ccr;
  std::vector<double> ccv2;
  for (int c = 0; c < k; ++c) {
    int j = basis[pos_P[c]];
    for (int e = H_.cptr[j]; e < H_.cptr[j + 1]; ++e) {
      int r = H_.ridx[e];
      if (slackpos[r] >= 0) {
        ccr.push_back(r);
        ccv2.push_back(H_.cval[e]);
        crp[r + 1]++;
      }
    }
    ccp[c + 1] = (int)ccr.size();
  }
  for (int r = 0; r < m; ++r) crp[r + 1] += crp[r];
  std::vector<int> cci(ccr.size()), fill(crp.begin(), crp.end() - 1), by_var(k);
  std::vector<double> ccv(ccr.size());
  for (int c = 0; c < k; ++c) by_var[c] = c;
  std::sort(by_var.begin(), by_var.end(), [&](int a, int b) { return basis[pos_P[a]] < basis[pos_P[b]]; });
  for (int c : by_var)
    for (int q = ccp[c]; q < ccp[c + 1]; ++q) {
      const int slot = fill[ccr[q]]++;
      cci[slot] = c;
      ccv[slot] = ccv2[q];
    }
  pos_P.resize(std::max(k, 1));
  row_R.resize(std::max(k, 1));
  PDS_HIP(hipMemcpyAsync(pos_P_.p, pos_P.data(), sizeof(int) * pos_P.size(), hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(row_R_.p, row_R.data(), sizeof(int) * row_R.size(), hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(Ridx_.p, Ridx.data(), sizeof(int) * m, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(Pidx_.p, Pidx.data(), sizeof(int) * m, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(slackpos_.p, slackpos.data(), sizeof(int) * m, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(slackrow_.p, slackrow.data(), sizeof(int) * m, hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(crp_.p, crp.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice, st_));
  PDS_HIP(hipMemcpyAsync(ccp_.p, ccp.data(), sizeof(int) * (k + 1), hipMemcpyHostToDevice, st_));
  if (!cci.empty()) {
    PDS_HIP(hipMemcpyAsync(cci_.p, cci.data(), sizeof(int) * cci.size(), hipMemcpyHostToDevice, st_));
    PDS_HIP(hipMemcpyAsync(ccv_.p, ccv.data(), sizeof(double) * ccv.size(), hipMemcpyHostToDevice, st_));
    PDS_HIP(hipMemcpyAsync(ccr_.p, ccr.data(), sizeof(int) * ccr.size(), hipMemcpyHostToDevice, st_));
    PDS_HIP(hipMemcpyAsync(ccv2_.p, ccv2.data(), sizeof(double) * ccv2.size(), hipMemcpyHostToDevice, st_));
  }
  k_ = k;
  if (k > 0) {
    ensure_kernel_capacity(GW_, k);
    ensure_kernel_capacity(Kinv_, k);
    ensure_kernel_capacity(KinvT_, k);
    const int zero2[2] = {0, 0};
    PDS_HIP(hipMemcpyAsync(flag_.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, st_));
    launch_build_kernel(st_, M(), basis_.p, F(), GW_.p);
    if (P_.force_event == DS_FORCE_SINGULAR && !force_fired_ && repair_ && iteration_ >= P_.force_at) {
      launch_force_zero_column(st_, GW_.p, k, k / 2);
      force_fired_ = true;
    }
    GJWork G{GW_.p, Pb_.p, piv_.p, perm_.p, flag_.p, Pinv_.p, rowsA_.p, rowsB_.p, nA_.p, nB_.p, gj_mfma_, rowscale_.p};
    if ((int)gemm_ev_.size() >= 2 * ((k + GJ_NB - 1) / GJ_NB)) G.gemm_ev = gemm_ev_.data();
    launch_gj_panels(st_, k, G);
    int flag[2] = {0, 0};
    std::vector<int> piv(k), perm(k);
    PDS_HIP(hipMemcpyAsync(flag, flag_.p, sizeof(flag), hipMemcpyDeviceToHost, st_));
    PDS_HIP(hipMemcpyAsync(piv.data(), piv_.p, sizeof(int) * k, hipMemcpyDeviceToHost, st_));
    sync();
    if (flag[0]) {
      // Dependent steps (step, W row before that panel's exchanges) -> (kernel column,
      // original kernel row), by replaying the row exchanges of the earlier panels.
      if (flag[1] <= 0 || flag[1] > GJ_DEP_CAP) throw SingularError("singular basis kernel in refactorization");
      std::vector<int> ent(2 * flag[1]);
      PDS_HIP(hipMemcpyAsync(ent.data(), flag_.p + 2, sizeof(int) * ent.size(), hipMemcpyDeviceToHost, st_));
      sync();
      std::vector<int> pos2orig(k);
      for (int r = 0; r < k; ++r) pos2orig[r] = r;
      std::vector<std::pair<int, int>> out;
      for (int j0 = 0; j0 < k; j0 += GJ_NB) {
        const int j1 = std::min(k, j0 + GJ_NB);
        for (int d = 0; d < flag[1]; ++d)
          if (ent[2 * d] >= j0 && ent[2 * d] < j1) out.push_back({ent[2 * d], pos2orig[ent[2 * d + 1]]});
        for (int g = j0; g < j1; ++g) std::swap(pos2orig[g], pos2orig[piv[g]]);
      }
      fac_pos_P_ = pos_P;
      fac_row_R_ = row_R;
      dep = out;
      return;
    }
    // Undo the row exchanges on the columns: composed swaps in reverse order.
    for (int c = 0; c < k; ++c) perm[c] = c;
    for (int s2 = k - 1; s2 >= 0; --s2) {
      int r = piv[s2];
      if (r < 0 || r >= k) throw std::runtime_error("pivot index out of range");
      std::swap(perm[s2], perm[r]);
    }
    PDS_HIP(hipMemcpyAsync(perm_.p, perm.data(), sizeof(int) * k, hipMemcpyHostToDevice, st_));
    launch_gj_unscramble(st_, k, G, Kinv_.p, KinvT_.p);
  }
  pull();
  hS_->t = 0;
  hS_->k = k;
  push();
  t_ = 0;
}

void Solver::initBetaWeights(bool identity) {
  if (identity) launch_fill(st_, beta_.p, 1.0, H_.m);
  else launch_dse_exact(st_, M(), I(), F());
}

// y = B^{-T} c_B ; d_N = c_N - A_N^T y  (sequentialDualSimplex.cpp:119-132)
void Solver::initReducedCosts() {
  launch_gather_cb(st_, M(), I(), w1_.p);
  launch_btran_dense(st_, M(), I(), F(), w1_.p, rho_.p, t_);
  launch_spmv(st_, M(), posN_.p, rho_.p, -1.0, 1.0, c_.p, d_.p, 1, S_.p, 0);
}

// sequentialDualSimplex.cpp:230-250
void Solver::dualSimplexInit() {
  launch_set_primal(st_, M(), I());
  launch_mask_nonbasic(st_, M(), I(), x_.p, wN_.p);
  launch_rowspmv(st_, M(), wN_.p, nullptr, -1.0, 1.0, w1_.p);
  launch_ftran_dense(st_, M(), I(), F(), w1_.p, x_.p, t_, 1, 0);
  initReducedCosts();
  launch_dot_obj(st_, M(), I());
}

// minimizeDualInfeasibilityInit (tpp:776-795) + initPhase1PricingVector
void Solver::phase1Init() {
  initReducedCosts();
  launch_phase1_inf(st_, M(), I());
  launch_rowspmv(st_, M(), wN_.p, nullptr, 1.0, 0.0, w1_.p);
  launch_ftran_dense(st_, M(), I(), F(), w1_.p, f_.p, t_, 0, 0);
}

bool Solver::checkPerturbNeed() const {
  std::vector<double> cs(c_host_);
  std::sort(cs.begin(), cs.end());
  int uniq = 0;
  double last = 0.0;
  for (size_t i = 0; i < cs.size(); ++i)
    if (i == 0 || !(std::fabs(last - cs[i]) < EPS_COSTS)) {
      last = cs[i];
      ++uniq;
    }
  return uniq < PERTURB_RATIO * (double)cs.size();
}

// perturbCosts (tpp:238-275) with a seeded stream in place of random_device.
void Solver::perturbCosts() {
  static const double wts[10] = {0.01, 0.1, 1, 2, 5, 10, 20, 30, 40, 100};
  SplitMix rng(P_.perturb_seed);
  rng.s = rng_state_ ? rng_state_ : P_.perturb_seed;
  c_orig_ = c_host_;
  double mean = 0.0;
  for (double v : c_host_) mean = mean + v;
  mean = mean / (double)c_host_.size();
  const double min_p = std::min(1e-2 * EPS_D, PSI);
  const double max_p = std::max(1e+3 * EPS_D, PSI * 10 * mean);
  for (int i = 0; i < H_.N; ++i) {
    double magnitude = 100 * EPS_D + PSI * c_host_[i];
    double pert;
    if (H_.bt[i] != BT_LOWER && H_.bt[i] != BT_FREE) pert = -0.5 * magnitude * (1 + rng.uniform());
    pert = 0.5 * magnitude * (1 + rng.uniform());
    int cnt = H_.cptr[i + 1] - H_.cptr[i];
    int kk = cnt > 10 ? 9 : (cnt == 0 ? 0 : cnt - 1);
    pert = wts[kk] * pert;
    if (pert != 0.0)
      while (std::fabs(pert) > std::fabs(max_p) || std::fabs(pert) < std::fabs(min_p)) {
        if (std::fabs(pert) > std::fabs(max_p)) pert = 0.1 * pert;
        else if (std::fabs(pert) < std::fabs(min_p)) pert = 10 * pert;
      }
    c_host_[i] += pert;
  }
  rng_state_ = rng.s;
  perturbed_ = true;
  upload_costs();
  pull();
  hS_->perturbed = 1;
  push();
}

void Solver::upload_costs() {
  PDS_HIP(hipMemcpyAsync(c_.p, c_host_.data(), sizeof(double) * H_.N, hipMemcpyHostToDevice, st_));
  sync();
}

void Solver::initDualSimplex() {
  if (!loaded_) throw std::logic_error("no problem loaded");
  const int m = H_.m, n = H_.n, N = H_.N;
  std::vector<int> basis(m), nonbasis(n), posN(N, -1);
  for (int i = 0; i < n; ++i) {
    nonbasis[i] = i;
    posN[i] = i;
  }
  for (int i = 0; i < m; ++i) basis[i] = n + i;
  upload(basis_, basis, st_);
  upload(nonbasis_, nonbasis, st_);
  upload(posN_, posN, st_);
  factor_valid_ = false;
  launch_fill(st_, x_.p, 0.0, N);
  launch_fill(st_, d_.p, 0.0, N);
  launch_fill(st_, beta_.p, 1.0, m);
  std::memset(hS_, 0, sizeof(DevScalars));
  push();
  c_host_ = H_.c;
  c_orig_ = H_.c;
  upload_costs();
  perturbed_ = false;
  rng_state_ = 0;
  iteration_ = 0;
  t_ = 0;
  refactors_ = restarts_ = phase1_iters_ = cleanups_ = repairs_ = 0;
  primal_iters_ = shifted_ = 0;
  warm_ = false;
  repaired_ = false;
  repair_ = false;
  force_fired_ = false;
  force_enq_it_ = -1;
  time_hit_ = false;
  last_dualcheck_ = 0;
  for (double& v : step_ms_) v = 0.0;
  pev_part_.clear();
  solve_ms_ = refactor_ms_ = 0.0;
  mark_ms_ = spmv_ms_ = 0.0;
  spmv_launches_ = 0;
  marked_ = false;
  ev_used_ = 0;
  nnzN_mark_ = nnzN_end_ = 0.0;
  status_ = DS_NOT_SOLVED;
  solved_ = false;
  message_.clear();
  refactor();
  inited_ = true;
}

void Solver::enqueue_body(int phase, int it, int skip_top, bool dualcheck, int t) {
  const DevModel m = M();
  const DevIter i = I();
  const DevFactor f = F();
  prof_event(DS_PART_PRICING);
  if (phase == 2 && dualcheck) {
    if (force_due(DS_FORCE_DUAL_RESTART, it)) launch_force(st_, m, i, DS_FORCE_DUAL_RESTART);
    launch_dualcheck(st_, m, i, it);
  }
  // Timed launches carry a start/stop event pair that records the kernel's own
  // execution interval (hipExtLaunchKernelGGL).
  // Every time_kernels-th iteration: a per-launch event pair costs the loop a few percent.
  const bool timed = P_.time_kernels > 0 && marked_ && it % P_.time_kernels == 0 && ev_used_ < (int)ev_iter_.size();
  if (!timed && !P_.profile && (P_.ratio_test == 0 || P_.ratio_test == 1) && !has_boxed_) {
    // CHUZR goes on with BTRAN's set-up, and the DSE FTRAN's stages share the
    // launches of the SpMV, CHUZC and the entering column's FTRAN (same
    // kernels' code, same results), and up to 16,384 update rows the entering
    // column's FTRAN and both sets of etas run in the update's launch: five
    // launches per iteration (six with the Harris ratio test's second pass).
    launch_price_btran(st_, m, i, f, phase, it, skip_top, P_.perturb, t, rho_.p);
    double* tau0 = t > 0 ? w2_.p : tau_.p;
    launch_stage_spmv_gemv(st_, m, i, f, posN_.p, rho_.p, c_.p, alpha_r_.p, tau0);
    launch_stage_chuzc_slack(st_, m, i, f, phase, it, rho_.p, tau0, P_.ratio_test);
    if (phase == 2 && force_due(DS_FORCE_STABILITY, it)) {
      launch_force(st_, m, i, DS_FORCE_STABILITY);
    }
    launch_col_update(st_, m, i, f, phase, it, t, tau0, P_.split_launches);
    return;
  }
  launch_price(st_, m, i, phase, it, skip_top, P_.perturb);
  prof_event(DS_PART_PRICING);
  launch_btran_unit(st_, m, i, f, t, rho_.p, -1);
  prof_event(DS_PART_BTRAN);
  launch_spmv(st_, m, posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 1, timed ? ev_[2 * ev_used_] : nullptr,
              timed ? ev_[2 * ev_used_ + 1] : nullptr);
  if (timed) ev_iter_[ev_used_++] = it;
  prof_event(DS_PART_PIVOT_ROW);
  launch_chuzc(st_, m, i, phase, P_.ratio_test, it);
  // Bound flips (sequentialDualSimplex.cpp:398-422) exist only with boxed variables.
  const bool flips = phase == 2 && has_boxed_;
  if (flips) launch_bfrt(st_, m, i, f, P_.ratio_test, t);
  if (phase == 2 && force_due(DS_FORCE_STABILITY, it)) launch_force(st_, m, i, DS_FORCE_STABILITY);
  prof_event(DS_PART_RATIO_TEST2);
  launch_ftran_col(st_, m, i, f, -1, t, alpha_q_.p, 1);
  prof_event(DS_PART_FTRAN);
  // tau = B^-1 rho first: the dual update and the basis change then go in one launch.
  launch_ftran_dense(st_, m, i, f, rho_.p, tau_.p, t, 0, 1);
  prof_event(DS_PART_FTRAN_BETA);
  if (flips) {
    launch_flips(st_, m, i, f, t);
    prof_event(DS_PART_FTRAN_BFRT);
  }
  launch_dupdate_update(st_, m, i, phase, it);
  prof_event(DS_PART_UPDATE_RED_COSTS);
}

void Solver::enqueue_stageB(int phase, int it, int t) {
  prof_event(DS_PART_BASIS_UPDATE);
  launch_ftran_dense(st_, M(), I(), F(), rho_.p, tau_.p, t, 0, 1);
  launch_update(st_, M(), I(), phase, it);
  if (phase == 2 && P_.candidates > 1) launch_rho_update(st_, M(), I(), P_.candidates);
  prof_event(DS_PART_BASIS_UPDATE);
  prof_harvest();
}

// Queue iterations next, next+1, ... up to the chunk size, stopping before
// the next scheduled refactor and at the iteration cap.
// One minor iteration of a major iteration with several candidate rows: the
// best eligible candidate's BTRAN row replaces the BTRAN of the iteration;
// after the basis change the other candidates' rows are updated.
void Solver::enqueue_minor(int it, int t) {
  const DevModel m = M();
  const DevIter i = I();
  const DevFactor f = F();
  prof_event(DS_PART_PRICING);
  launch_pick(st_, m, i, it, P_.candidate_psi);
  prof_event(DS_PART_PRICING);
  launch_spmv(st_, m, posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 1);
  prof_event(DS_PART_PIVOT_ROW);
  launch_chuzc(st_, m, i, 2, P_.ratio_test, it);
  if (has_boxed_) launch_bfrt(st_, m, i, f, P_.ratio_test, t);
  prof_event(DS_PART_RATIO_TEST2);
  launch_ftran_col(st_, m, i, f, -1, t, alpha_q_.p, 1);
  prof_event(DS_PART_FTRAN);
  launch_ftran_dense(st_, m, i, f, rho_.p, tau_.p, t, 0, 1);
  prof_event(DS_PART_FTRAN_BETA);
  if (has_boxed_) {
    launch_flips(st_, m, i, f, t);
    prof_event(DS_PART_FTRAN_BFRT);
  }
  launch_dupdate_update(st_, m, i, 2, it);
  launch_rho_update(st_, m, i, P_.candidates);
  prof_event(DS_PART_UPDATE_RED_COSTS);
}

// A major iteration (phase 2, several candidates): the dual-feasibility check
// when a multiple of NEED_RESTART iterations has passed since the last one,
// the major pricing, the candidates' BTRANs back to back, then up to
// `candidates` minor iterations, stopping before the next scheduled refactor.
int Solver::plan_major(int next, int t, int skip_first) {
  const int K = P_.refactor_freq;
  int L = P_.candidates;
  L = std::min(L, K - (next % K));
  if (P_.max_iterations > 0) L = std::min(L, P_.max_iterations - next + 1);
  L = std::min(L, eta_cap_ - 1 - t);
  if (!marked_ && P_.mark_iteration > 0 && next <= P_.mark_iteration) L = std::min(L, P_.mark_iteration - next + 1);
  if (L < 1) L = 1;
  const DevModel m = M();
  const DevIter i = I();
  const DevFactor f = F();
  prof_event(DS_PART_PRICING);
  if (!skip_first && next / NEED_RESTART > last_dualcheck_ / NEED_RESTART) {
    launch_dualcheck(st_, m, i, next);
    last_dualcheck_ = next;
  }
  launch_price_multi(st_, m, i, next, skip_first, P_.perturb, P_.candidates);
  prof_event(DS_PART_PRICING);
  for (int c = 0; c < P_.candidates; ++c) launch_btran_cand(st_, m, i, f, t, c);
  prof_event(DS_PART_BTRAN);
  for (int k = 0; k < L; ++k) enqueue_minor(next + k, t + k);
  return L;
}

// Test hook: the forced event is due at iteration `it` (queued once; queued
// again if the chunk holding it stopped before reaching it).
bool Solver::force_due(int event, int it) {
  if (P_.force_event != event || force_fired_ || force_enq_it_ >= 0 || it < P_.force_at) return false;
  force_enq_it_ = it;
  return true;
}

int Solver::plan_chunk(int phase, int next, int t, int skip_first) {
  const int K = P_.refactor_freq;
  if (force_enq_it_ >= 0) {
    if (iteration_ >= force_enq_it_) force_fired_ = true;
    force_enq_it_ = -1;
  }
  if (phase == 2 && P_.candidates > 1) return plan_major(next, t, skip_first);
  int L = 0;
  for (int i = next;; ++i) {
    if (L > 0) {
      if (i % K == 0 || L >= P_.chunk) break;
      if (P_.max_iterations > 0 && i > P_.max_iterations) break;
      if (!marked_ && P_.mark_iteration > 0 && i == P_.mark_iteration + 1) break;
    }
    if (t + L + 1 >= eta_cap_) break;
    bool dc = (phase == 2) && (i % NEED_RESTART == 0) && (L > 0 || skip_first == 0);
    enqueue_body(phase, i, L == 0 ? skip_first : 0, dc, t + L);
    ++L;
    if (P_.profile) break;  // one iteration per synchronisation: the step events are read back each time
  }
  return L;
}

// After a refactor that substituted slacks for dependent columns: the iterate
// belongs to the old basis, so recompute it (exact DSE weights, then x and d,
// or the phase-1 pricing vector).  Returns whether it did.
bool Solver::after_repair(int phase) {
  if (!repaired_) return false;
  repaired_ = false;
  restarts_++;
  initBetaWeights(false);
  if (phase == 2) dualSimplexInit();
  else phase1Init();
  return true;
}

// minimizeDualInfeasibility (baseDualSimplex.tpp:799-934).
int Solver::phase1() {
  const int K = P_.refactor_freq;
  phase1Init();
  pull();
  hS_->cycle = 0;
  hS_->status = ST_RUNNING;
  push();
  int resume = -1, resume_skip = 0;
  while (true) {
    int next, skip_first = 0;
    if (resume >= 0) {
      next = resume;
      skip_first = resume_skip;
      resume = -1;
    } else {
      next = iteration_ + 1;
      tick_limit(next);
      if (next % K == 0) {
        // Top-of-iteration checks in the reference's order, then the refactor.
        pull();
        if (P_.perturb && !hS_->perturbed && hS_->cycle > MAX_CYCLE) {
          perturbCosts();
          phase1Init();
          pull();
        }
        if (hS_->cycle > NEED_RESTART || hS_->obj > BIG || std::isnan(hS_->obj)) {
          phase1Init();
        }
        refactor();
        after_repair(1);
        skip_first = 1;
      }
    }
    int L = plan_chunk(1, next, t_, skip_first);
    pull();
    prof_harvest();
    if (hS_->status == ST_RUNNING && !hS_->refac_mid) {
      iteration_ = next + L - 1;
      t_ = hS_->t;
      harvest_events(iteration_);
      maybe_mark(iteration_);
      continue;
    }
    const int stop = hS_->stop_iter;
    harvest_events(stop - 1);
    iteration_ = stop;
    t_ = hS_->t;
    if (hS_->refac_mid && hS_->status == ST_RUNNING) {
      hS_->refac_mid = 0;
      push();
      refactor();
      // A repaired basis voids the pending basis change: start over from the new basis.
      if (!after_repair(1)) enqueue_stageB(1, stop, 0);
      pull();
      t_ = hS_->t;
      continue;
    }
    const int st = hS_->status;
    hS_->status = ST_RUNNING;
    hS_->refac_mid = 0;
    push();
    if (P_.verbose) std::fprintf(stderr, "[ds] phase 1: event %d at iteration %d\n", st, stop);
    prof_host_begin();
    switch (st) {
      case ST_PH1_SOLVED:
        return 0;
      case ST_PH1_DUALINF:
        solved_ = false;
        message_ = "dual infeasible";
        return 1;
      case ST_NEED_PERTURB:
        perturbCosts();
        phase1Init();
        pull();
        if (hS_->cycle > NEED_RESTART || hS_->obj > BIG || std::isnan(hS_->obj)) phase1Init();
        resume = stop;
        resume_skip = 1;
        break;
      case ST_PH1_REINIT:
        phase1Init();
        resume = stop;
        resume_skip = 1;
        break;
      case ST_PH1_RESTART: {
        restarts_++;
        refactor();
        after_repair(1);
        phase1Init();
        initBetaWeights(false);
        pull();
        if (hS_->nblocked > 0) PDS_HIP(hipMemsetAsync(blocked_.p, 0, H_.N, st_));
        hS_->nblocked = 0;
        hS_->any_blocked = 0;
        push();
        break;
      }
      case ST_NO_RATIO: {
        restarts_++;
        pull();
        int p = hS_->p;
        unsigned char one = 1;
        PDS_HIP(hipMemcpy(blocked_.p + p, &one, 1, hipMemcpyHostToDevice));
        if (hS_->nblocked < H_.m) {  // a blocked variable is not priced again until the set clears
          PDS_HIP(hipMemcpy(blocked_list_.p + hS_->nblocked, &p, sizeof(int), hipMemcpyHostToDevice));
          hS_->nblocked++;
        }
        hS_->any_blocked = 1;
        push();
        refactor();
        after_repair(1);
        phase1Init();
        break;
      }
      case ST_BTRAN_CAP:
        throw std::runtime_error("sparse BTRAN list capacity exceeded");
      default:
        throw std::runtime_error("unexpected device status " + std::to_string(st) + " in phase 1");
    }
    prof_host_end(DS_PART_RESTORE);
  }
}

// callDualSolver (sequentialDualSimplex.cpp:253-515).
bool Solver::phase2() {
  const int K = P_.refactor_freq;
  prof_host_begin();
  dualSimplexInit();
  prof_host_end(DS_PART_INIT);
  pull();
  hS_->cycle = 0;
  hS_->unbound_checked = 0;
  hS_->status = ST_RUNNING;
  push();
  int resume = -1, resume_skip = 0;
  while (true) {
    int next, skip_first = 0;
    if (resume >= 0) {
      next = resume;
      skip_first = resume_skip;
      resume = -1;
    } else {
      next = iteration_ + 1;
      tick_limit(next);
      if (next % K == 0) {
        refactor();
        after_repair(2);
      }
    }
    int L = plan_chunk(2, next, t_, skip_first);
    pull();
    prof_harvest();
    if (hS_->status == ST_RUNNING && !hS_->refac_mid) {
      iteration_ = next + L - 1;
      t_ = hS_->t;
      harvest_events(iteration_);
      maybe_mark(iteration_);
      continue;
    }
    const int stop = hS_->stop_iter;
    harvest_events(stop - 1);
    iteration_ = stop;
    t_ = hS_->t;
    if (hS_->refac_mid && hS_->status == ST_RUNNING) {
      hS_->refac_mid = 0;
      push();
      refactor();
      // A repaired basis voids the pending basis change: start over from the new basis.
      if (!after_repair(2)) enqueue_stageB(2, stop, 0);
      pull();
      t_ = hS_->t;
      continue;
    }
    const int st = hS_->status;
    hS_->status = ST_RUNNING;
    hS_->refac_mid = 0;
    push();
    if (P_.verbose) std::fprintf(stderr, "[ds] phase 2: event %d at iteration %d\n", st, stop);
    switch (st) {
      case ST_OPTIMAL:
        solved_ = true;
        status_ = DS_OPTIMAL;
        message_ = "optimal solution";
        return true;
      case ST_MAJOR_END:
        break;  // no eligible candidate left: the next major iteration prices again
      case ST_NO_PIVOT:
        restarts_++;
        prof_host_begin();
        refactor();
        after_repair(2);
        initBetaWeights(false);
        dualSimplexInit();
        prof_host_end(DS_PART_RESTORE);
        break;
      case ST_DUAL_RESTART: {
        restarts_++;
        prof_host_begin();
        refactor();
        after_repair(2);
        prof_host_end(DS_PART_RESTORE);
        pull();
        int cyc = hS_->cycle;
        phase1();
        prof_host_begin();
        dualSimplexInit();
        prof_host_end(DS_PART_RESTORE);
        pull();
        hS_->cycle = cyc;
        push();
        resume = iteration_;
        resume_skip = 0;
        break;
      }
      case ST_NEED_PERTURB:
        prof_host_begin();
        perturbCosts();
        refactor();
        after_repair(2);
        dualSimplexInit();
        prof_host_end(DS_PART_RESTORE);
        resume = stop;
        resume_skip = 1;
        break;
      case ST_NO_RATIO: {
        pull();
        if (hS_->unbound_checked) {
          solved_ = false;
          status_ = DS_PRIMAL_INFEASIBLE;
          message_ = "dual unbounded";
          return false;
        }
        restarts_++;
        hS_->unbound_checked = 1;
        push();
        prof_host_begin();
        refactor();
        after_repair(2);
        initBetaWeights(false);
        prof_host_end(DS_PART_RESTORE);
        pull();
        int cyc = hS_->cycle;
        phase1();
        prof_host_begin();
        dualSimplexInit();
        prof_host_end(DS_PART_RESTORE);
        pull();
        hS_->cycle = cyc;
        hS_->unbound_checked = 1;
        push();
        break;
      }
      case ST_BTRAN_CAP:
        throw std::runtime_error("sparse BTRAN list capacity exceeded");
      default:
        throw std::runtime_error("unexpected device status " + std::to_string(st) + " in phase 2");
    }
  }
}

// ---------------------------------------------------------------------------
// Primal simplex iterations of a warm start (params.warm_primal; the reference
// always starts from the slack basis).  Bounds of basic variables that the
// starting basis violates are moved warm_shift * (1 + |x_j|) past x_j, which
// makes the basis primal feasible; primal iterations (Devex pricing, two-pass
// ratio test) then run until no reduced cost is dual infeasible.  The bounds
// are restored and the basis refactored on the way out: the dual simplex
// continues from a basis that is dual feasible for the true costs, and its
// phase 2 removes the primal infeasibilities the restored bounds leave.
// ---------------------------------------------------------------------------
// Moves the violated bounds of the current basic variables (on the host copies
// lw_/uw_, then uploaded); returns how many moved.
int Solver::shift_bounds() {
  const int m = H_.m, N = H_.N;
  std::vector<double> x(N);
  PDS_HIP(hipMemcpyAsync(x.data(), x_.p, sizeof(double) * N, hipMemcpyDeviceToHost, st_));
  std::vector<int> b = basis_host();
  int moved = 0;
  for (int i = 0; i < m; ++i) {
    const int j = b[i];
    const double xj = x[j], gap = P_.warm_shift * (1.0 + std::fabs(xj));
    const int8_t t = H_.bt[j];
    if ((t == BT_LOWER || t == BT_FIXED) && xj < lw_[j]) {
      lw_[j] = xj - gap;
      ++moved;
    }
    if ((t == BT_UPPER || t == BT_FIXED) && xj > uw_[j]) {
      uw_[j] = xj + gap;
      ++moved;
    }
  }
  if (moved) {
    PDS_HIP(hipMemcpyAsync(l_.p, lw_.data(), sizeof(double) * N, hipMemcpyHostToDevice, st_));
    PDS_HIP(hipMemcpyAsync(u_.p, uw_.data(), sizeof(double) * N, hipMemcpyHostToDevice, st_));
    sync();
  }
  return moved;
}

void Solver::enqueue_primal(int it, int t) {
  const DevModel m = M();
  const DevIter i = I();
  const DevFactor f = F();
  launch_pprice(st_, m, i, it);
  launch_ftran_col(st_, m, i, f, -1, t, alpha_q_.p, 1);
  launch_pratio_btran(st_, m, i, f, it, t, rho_.p, P_.split_launches);
  launch_spmv(st_, m, posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 1);
  launch_pupdate(st_, m, i, it);
}

void Solver::primal_phase() {
  const int K = P_.refactor_freq;
  const int cap = P_.warm_primal_cap > 0 ? P_.warm_primal_cap : 4 * H_.m;
  lw_ = H_.l;
  uw_ = H_.u;
  const int it0 = iteration_;
  // The bounds stay moved only inside this function.
  struct Restore {
    Solver* s;
    ~Restore() {
      (void)hipMemcpyAsync(s->l_.p, s->H_.l.data(), sizeof(double) * s->H_.N, hipMemcpyHostToDevice, s->st_);
      (void)hipMemcpyAsync(s->u_.p, s->H_.u.data(), sizeof(double) * s->H_.N, hipMemcpyHostToDevice, s->st_);
      (void)hipStreamSynchronize(s->st_);
    }
  } restore{this};
  dualSimplexInit();
  shifted_ += shift_bounds();
  launch_fill(st_, devex_.p, 1.0, H_.N);
  pull();
  hS_->status = ST_RUNNING;
  hS_->refac_mid = 0;
  hS_->cycle = 0;
  push();
  while (iteration_ - it0 < cap) {
    const int next = iteration_ + 1;
    tick_limit(next);
    if (next % K == 0) {
      refactor();
      if (repaired_) {
        repaired_ = false;
        restarts_++;
        dualSimplexInit();
        shifted_ += shift_bounds();
      }
      // A new Devex reference framework (all weights 1) at every refactor.
      launch_fill(st_, devex_.p, 1.0, H_.N);
    }
    int L = 0;
    for (int i = next;; ++i) {
      if (L > 0 && (i % K == 0 || L >= P_.chunk || i - it0 > cap)) break;
      if (P_.max_iterations > 0 && i > P_.max_iterations && L > 0) break;
      if (t_ + L + 1 >= eta_cap_) break;
      enqueue_primal(i, t_ + L);
      ++L;
    }
    if (L == 0) {  // eta file full
      refactor();
      t_ = 0;
      continue;
    }
    pull();
    if (hS_->status == ST_RUNNING && !hS_->refac_mid) {
      iteration_ = next + L - 1;
      t_ = hS_->t;
      continue;
    }
    const int stop = hS_->stop_iter;
    t_ = hS_->t;
    if (hS_->refac_mid && hS_->status == ST_RUNNING) {
      iteration_ = stop;
      hS_->refac_mid = 0;
      push();
      refactor();
      if (repaired_) {
        repaired_ = false;
        restarts_++;
        dualSimplexInit();
        shifted_ += shift_bounds();
      } else {
        enqueue_stageB(2, stop, 0);
      }
      pull();
      t_ = hS_->t;
      continue;
    }
    // ST_OPTIMAL: dual feasible; ST_NO_PIVOT: no blocking row, the dual
    // simplex takes over from this basis either way.
    iteration_ = stop - 1;
    hS_->status = ST_RUNNING;
    hS_->refac_mid = 0;
    push();
    break;
  }
  primal_iters_ += iteration_ - it0;
}

// checkDualFeasible (tpp:362-385) on the current reduced costs.
bool Solver::dual_feasible() {
  pull();
  hS_->status = ST_RUNNING;
  hS_->refac_mid = 0;
  push();
  launch_dualcheck(st_, M(), I(), iteration_);
  pull();
  const bool ok = hS_->dinf == 0;
  hS_->status = ST_RUNNING;
  push();
  return ok;
}

// checkPrimalFeasible (tpp:296-324) on the current x_B: the phase-2 pricing
// kernel raises ST_OPTIMAL exactly when no basic variable is out of bounds.
bool Solver::primal_feasible() {
  pull();
  hS_->status = ST_RUNNING;
  hS_->refac_mid = 0;
  push();
  launch_price(st_, M(), I(), 2, iteration_, 1, 0);
  pull();
  const bool ok = hS_->status == ST_OPTIMAL;
  hS_->status = ST_RUNNING;
  push();
  return ok;
}

int Solver::presolve() {
  if (!inited_) throw std::logic_error("initDualSimplex not called");
  const double t0 = now_ms();
  t_start_ = t0;
  int st = 0;
  try {
    repair_ = true;
    if (P_.perturb && checkPerturbNeed()) perturbCosts();
    if (warm_ && P_.warm_primal && primal_ok_) {
      primal_phase();
      // A fresh factor of the basis the primal iterations reached; the dual
      // simplex's iterate is computed from it.
      refactor();
      repaired_ = false;
    }
    initBetaWeights(!warm_);
    st = phase1();
    phase1_iters_ = iteration_;
    if (st == 1) {
      status_ = DS_DUAL_INFEASIBLE;
      solved_ = false;
      message_ = "dual infeasible";
    } else {
      launch_set_primal(st_, M(), I());
    }
    sync();
  } catch (const IterLimit&) {
    status_ = DS_ITERATION_LIMIT;
    message_ = time_hit_ ? "time limit" : "iteration limit";
    st = 2;
    solve_ms_ += now_ms() - t0;
    if (P_.time_kernels) nnzN_end_ = nnz_nonbasic();
    return st;
  }
  solve_ms_ += now_ms() - t0;
  return st;
}

void Solver::solve() {
  const double t0 = now_ms();
  if (status_ == DS_ITERATION_LIMIT) return;
  try {
    bool ok = phase2();
    // postProcess (tpp:123-159): restore the original costs and recompute.
    if (perturbed_) {
      c_host_ = c_orig_;
      upload_costs();
      dualSimplexInit();
      // dualSimplexInit places boxed nonbasic variables by the sign of the reduced costs it
      // finds on entry (setPrimalVars before initReducedCosts, sequentialDualSimplex.cpp:230-250),
      // here those of the perturbed costs; the cleanup places them again by the restored ones.
      if (P_.cleanup && has_boxed_) dualSimplexInit();
      // The reference's primal correction is a stub (sequentialDualSimplex.cpp:113-116): the
      // final basis may be dual infeasible for the original costs.  With `cleanup` the dual
      // simplex resumes from it (phase 1, then phase 2), as the batched solver does.
      if (P_.cleanup && ok && !dual_feasible()) {
        cleanups_++;
        const int st1 = phase1();
        if (st1 == 1) {
          status_ = DS_DUAL_INFEASIBLE;
          solved_ = false;
          message_ = "dual infeasible";
        } else {
          launch_set_primal(st_, M(), I());
          phase2();
        }
      } else if (P_.cleanup && ok && !primal_feasible()) {
        // Dual feasible for the original costs, but dualSimplexInit moved boxed nonbasic
        // variables whose reduced cost changed sign to their other bound: continue phase 2.
        cleanups_++;
        phase2();
      }
    }
    pull();
    Z_ = hS_->obj;
  } catch (const IterLimit&) {
    status_ = DS_ITERATION_LIMIT;
    message_ = time_hit_ ? "time limit" : "iteration limit";
    pull();
    Z_ = hS_->obj;
  }
  solve_ms_ += now_ms() - t0;
  if (P_.time_kernels) nnzN_end_ = nnz_nonbasic();
}

void Solver::result(ds_result& r) const {
  r.status = status_;
  r.solved = solved_ ? 1 : 0;
  r.iterations = iteration_;
  r.phase1_iterations = phase1_iters_;
  r.perturbed = perturbed_ ? 1 : 0;
  r.refactors = refactors_;
  r.restarts = restarts_;
  r.Z = Z_ - H_.offset;
  r.solve_ms = solve_ms_;
  r.refactor_ms = refactor_ms_;
  r.mark_ms = mark_ms_;
  r.spmv_ms = spmv_ms_;
  r.spmv_launches = spmv_launches_;
  r.kernel_k = k_;
  r.nnz_nonbasic_mark = nnzN_mark_;
  r.nnz_nonbasic_end = nnzN_end_;
  r.cleanups = cleanups_;
  r.repairs = repairs_;
  for (int k = 0; k < DS_NPARTS; ++k) r.step_ms[k] = step_ms_[k];
  r.primal_iterations = primal_iters_;
  r.shifted_bounds = shifted_;
  if (P_.profile) r.step_ms[DS_PART_FACTOR] = refactor_ms_;
}

void Solver::solution(double& Z, int& solved, std::string& message, std::vector<double>& x) {
  x.resize(H_.N);
  PDS_HIP(hipMemcpyAsync(x.data(), x_.p, sizeof(double) * H_.N, hipMemcpyDeviceToHost, st_));
  sync();
  Z = Z_ - H_.offset;
  solved = solved_ ? 1 : 0;
  message = message_;
}

std::vector<int> Solver::basis_host() {
  std::vector<int> b(H_.m);
  PDS_HIP(hipMemcpyAsync(b.data(), basis_.p, sizeof(int) * H_.m, hipMemcpyDeviceToHost, st_));
  sync();
  return b;
}

std::vector<int> Solver::nonbasis_host() {
  std::vector<int> b(H_.n);
  PDS_HIP(hipMemcpyAsync(b.data(), nonbasis_.p, sizeof(int) * H_.n, hipMemcpyDeviceToHost, st_));
  sync();
  return b;
}

// initDualSimplex from a given starting basis: dependent columns give way to
// slacks (the iterate is set up from the repaired basis), and presolve starts
// from the exact DSE weights of that basis.
void Solver::init_basis(const int* basis) {
  initDualSimplex();
  repairs_ = 0;
  install_basis(basis);
  repair_ = true;
  refactor();
  repaired_ = false;
  warm_ = true;
}

void Solver::set_basis(const int* basis) {
  if (!inited_) initDualSimplex();
  install_basis(basis);
  repair_ = false;
  refactor();
}

void Solver::install_basis(const int* basis) {
  const int m = H_.m, n = H_.n, N = H_.N;
  std::vector<int> b(basis, basis + m), nb, posN(N, -1);
  std::vector<char> isb(N, 0);
  for (int i = 0; i < m; ++i) {
    if (b[i] < 0 || b[i] >= N || isb[b[i]]) throw std::invalid_argument("invalid basis");
    isb[b[i]] = 1;
  }
  for (int j = 0; j < N; ++j)
    if (!isb[j]) {
      posN[j] = (int)nb.size();
      nb.push_back(j);
    }
  (void)n;
  upload(basis_, b, st_);
  upload(nonbasis_, nb, st_);
  upload(posN_, posN, st_);
  sync();
  factor_valid_ = false;
}

void Solver::btran(int p_idx, double* rho) {
  if (p_idx < 0 || p_idx >= H_.m) throw std::invalid_argument("p_idx out of range");
  launch_btran_unit(st_, M(), I(), F(), t_, rho_.p, p_idx);
  PDS_HIP(hipMemcpyAsync(rho, rho_.p, sizeof(double) * H_.m, hipMemcpyDeviceToHost, st_));
  sync();
}

void Solver::ftran(const double* b, double* x) {
  PDS_HIP(hipMemcpyAsync(w1_.p, b, sizeof(double) * H_.m, hipMemcpyHostToDevice, st_));
  launch_ftran_dense(st_, M(), I(), F(), w1_.p, tau_.p, t_, 0, 0);
  PDS_HIP(hipMemcpyAsync(x, tau_.p, sizeof(double) * H_.m, hipMemcpyDeviceToHost, st_));
  sync();
}

void Solver::ftran_column(int j, double* x) {
  if (j < 0 || j >= H_.N) throw std::invalid_argument("column out of range");
  launch_ftran_col(st_, M(), I(), F(), j, t_, alpha_q_.p, 0);
  PDS_HIP(hipMemcpyAsync(x, alpha_q_.p, sizeof(double) * H_.m, hipMemcpyDeviceToHost, st_));
  sync();
}

void Solver::pivot_row(const double* rho, double* alpha) {
  PDS_HIP(hipMemcpyAsync(rho_.p, rho, sizeof(double) * H_.m, hipMemcpyHostToDevice, st_));
  launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0);
  PDS_HIP(hipMemcpyAsync(alpha, alpha_r_.p, sizeof(double) * H_.n, hipMemcpyDeviceToHost, st_));
  sync();
}

void Solver::time_pivot_row_ext(const double* rho, int reps, long long flush_bytes, double& warm_ms, double& cold_ms,
                                double& bytes) {
  // Per-launch kernel time from start/stop events around each SpMV launch:
  // warm = back to back (the matrix stays in L2/MALL, as inside the solver);
  // cold = after streaming flush_bytes through a copy kernel before each one.
  PDS_HIP(hipMemcpyAsync(rho_.p, rho, sizeof(double) * H_.m, hipMemcpyHostToDevice, st_));
  launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0);
  hipEvent_t e0, e1;
  PDS_HIP(hipEventCreate(&e0));
  PDS_HIP(hipEventCreate(&e1));
  double tw = 0.0, tc = 0.0;
  float t = 0.f;
  for (int r = 0; r < reps; ++r) {
    launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0, e0, e1);
    PDS_HIP(hipEventSynchronize(e1));
    PDS_HIP(hipEventElapsedTime(&t, e0, e1));
    tw += t;
  }
  if (flush_bytes >= 16) {
    const long long n16 = flush_bytes / 16;
    DBuf<double> fa, fb;
    fa.alloc((size_t)n16 * 2);
    fb.alloc((size_t)n16 * 2);
    PDS_HIP(hipMemsetAsync(fa.p, 0, (size_t)n16 * 16, st_));
    for (int r = 0; r < reps; ++r) {
      launch_copy(st_, (const double2*)fa.p, (double2*)fb.p, n16);
      launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0, e0, e1);
      PDS_HIP(hipEventSynchronize(e1));
      PDS_HIP(hipEventElapsedTime(&t, e0, e1));
      tc += t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  warm_ms = reps > 0 ? tw / reps : 0.0;
  cold_ms = (reps > 0 && flush_bytes >= 16) ? tc / reps : 0.0;
  std::vector<int> nb = nonbasis_host();
  double nnzN = 0;
  for (int j : nb) nnzN += H_.cptr[j + 1] - H_.cptr[j];
  bytes = 12.0 * nnzN + 4.0 * (nb.size() + 1) + 8.0 * nb.size() + 8.0 * H_.m;
}

// Factorizations of the current basis, `reps` times: mean wall time of a
// refactorization, and the device time and executed flops (2 (k - nb) k nb per
// panel of nb columns) of its FP64 MFMA trailing updates from event pairs.
void Solver::time_refactor(int reps, double& refactor_ms, double& gemm_ms, double& gemm_flops, int& k) {
  refactor(true);
  k = k_;
  const int panels = (k_ + GJ_NB - 1) / GJ_NB;
  for (int i = (int)gemm_ev_.size(); i < 2 * panels; ++i) {
    hipEvent_t e;
    PDS_HIP(hipEventCreate(&e));
    gemm_ev_.push_back(e);
  }
  double tr = 0.0, tg = 0.0;
  for (int r = 0; r < reps; ++r) {
    sync();
    const double t0 = now_ms();
    refactor(true);
    sync();
    tr += now_ms() - t0;
    for (int p = 0; p < panels; ++p) {
      float ms = 0.f;
      PDS_HIP(hipEventElapsedTime(&ms, gemm_ev_[2 * p], gemm_ev_[2 * p + 1]));
      tg += ms;
    }
  }
  for (hipEvent_t e : gemm_ev_) (void)hipEventDestroy(e);
  gemm_ev_.clear();
  double fl = 0.0;
  for (int j = 0; j < k_; j += GJ_NB) {
    const double nb = std::min(GJ_NB, k_ - j);
    fl += 2.0 * (k_ - nb) * (double)k_ * nb;
  }
  refactor_ms = reps > 0 ? tr / reps : 0.0;
  gemm_ms = reps > 0 ? tg / reps : 0.0;
  gemm_flops = fl;
}

void Solver::time_pivot_row(const double* rho, int reps, double& ms, double& bytes) {
  PDS_HIP(hipMemcpyAsync(rho_.p, rho, sizeof(double) * H_.m, hipMemcpyHostToDevice, st_));
  launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0);
  hipEvent_t e0, e1;
  PDS_HIP(hipEventCreate(&e0));
  PDS_HIP(hipEventCreate(&e1));
  PDS_HIP(hipEventRecord(e0, st_));
  for (int r = 0; r < reps; ++r) launch_spmv(st_, M(), posN_.p, rho_.p, 1.0, 0.0, c_.p, alpha_r_.p, 0, S_.p, 0);
  PDS_HIP(hipEventRecord(e1, st_));
  PDS_HIP(hipEventSynchronize(e1));
  float total = 0.f;
  PDS_HIP(hipEventElapsedTime(&total, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ms = reps > 0 ? (double)total / reps : 0.0;
  // Algorithmic bytes: 12 per nonzero of the nonbasic columns, their extents,
  // the nonbasic index list, rho once and alpha written once.
  std::vector<int> nb = nonbasis_host();
  double nnzN = 0;
  for (int j : nb) nnzN += H_.cptr[j + 1] - H_.cptr[j];
  bytes = 12.0 * nnzN + 4.0 * (nb.size() + 1) + 8.0 * nb.size() + 8.0 * H_.m;
}

}  // namespace pds
