// Host side of the MI355X dual simplex: the LP model in computational form
// (the reference's Problem, common/problem.{hpp,tpp}) and the solver driving
// the gfx950 kernels (the reference's SequentialDualSimplex /
// BaseDualSimplex control flow, cpu/sequentialDualSimplex.cpp and
// common/baseDualSimplex.tpp).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "ds_abi.h"
#include "ds_kernels.hpp"

namespace pds {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct SingularError : std::runtime_error {
  explicit SingularError(const std::string& s) : std::runtime_error(s) {}
};
struct Unsupported : std::runtime_error {
  explicit Unsupported(const std::string& s) : std::runtime_error(s) {}
};
struct IterLimit {};

#define PDS_HIP(x)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess)                                                                               \
      throw ::pds::HipError(std::string(#x) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                            std::to_string(__LINE__));                                                  \
  } while (0)

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    PDS_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

// LP in computational form: A <- [A | I], slack bounds from row ranges,
// RHS = 0 (problem.tpp:176-226).
struct HostModel {
  int m = 0, n = 0, N = 0;
  std::vector<int> cptr, ridx, rptr, cidx;
  std::vector<double> cval, rval, c, l, u;
  std::vector<int8_t> bt;
  double offset = 0.0;
  void build(int m, int n, const int* col_ptr, const int* row_idx, const double* val, const double* costs,
             const double* lb, const double* ub, const int* bound_type, const double* lr, const double* ur,
             const int* range_type);
};

class Solver {
 public:
  explicit Solver(int device);
  ~Solver();

  void load(const HostModel& hm);
  void set_params(const ds_params& p);
  void initDualSimplex();
  int presolve();      // 0 = Solved, 1 = DualInfeas (Phase1OutStatus)
  void solve();        // phase 2 + postProcess
  void result(ds_result& r) const;
  void solution(double& Z, int& solved, std::string& message, std::vector<double>& x);
  std::vector<int> basis_host();
  std::vector<int> nonbasis_host();

  // hook-level entries
  void set_basis(const int* basis);
  void init_basis(const int* basis);
  void install_basis(const int* basis);
  void refactor(bool force = false);  // force: factor again even when the basis is unchanged
  void btran(int p_idx, double* rho);
  void ftran(const double* b, double* x);
  void ftran_column(int j, double* x);
  void pivot_row(const double* rho, double* alpha);
  void time_pivot_row(const double* rho, int reps, double& ms, double& bytes);
  void time_refactor(int reps, double& refactor_ms, double& gemm_ms, double& gemm_flops, int& k);
  void time_pivot_row_ext(const double* rho, int reps, long long flush_bytes, double& warm_ms, double& cold_ms,
                          double& bytes);

  std::string last_error;
  hipStream_t stream() const { return st_; }
  void set_offset(double offset) { H_.offset = offset; }

 private:
  int device_ = 0;
  hipStream_t st_ = nullptr;
  ds_params P_{};
  HostModel H_;
  bool loaded_ = false, inited_ = false;

  // model
  DBuf<int> cptr_, ridx_, rptr_, cidx_;
  DBuf<double> cval_, rval_, c_, l_, u_;
  DBuf<int8_t> bt_;
  // iterate
  DBuf<double> x_, d_, beta_, f_, rho_, alpha_r_, alpha_q_, tau_, w1_, w2_, wN_, etas_, E_, MF_, x0_, part_d_,
      list_val_, ylist_val_, cand_th_, cand_a_;
  DBuf<int> basis_, nonbasis_, posN_, eta_pos_, eta_last_, blocked_list_, list_idx_, ylist_row_, list_n_, part_i_,
      cand_i_;
  DBuf<unsigned char> blocked_;
  // bound-flipping ratio test (allocated for models with boxed variables)
  DBuf<unsigned char> bmark_;
  DBuf<signed char> flipc_;
  DBuf<int> blist_, flist_;
  DBuf<double> bcc_, dxb_;
  bool has_boxed_ = false;
  bool factor_valid_ = false;  // the factor is that of the current basis (no eta since)
  int refactors_skipped_ = 0;
  // Primal simplex iterations of a warm start (params.warm_primal).
  bool primal_ok_ = true;        // no Free or Boxed variables
  DBuf<double> devex_;
  std::vector<double> lw_, uw_;  // working bounds with the moved ones
  int primal_iters_ = 0, shifted_ = 0;
  int shift_bounds();
  void enqueue_primal(int it, int t);
  void primal_phase();
  // multiple candidate rows (params.candidates > 1)
  DBuf<int> cand_pos_, cand_act_;
  DBuf<double> cand_w_, rho_c_;
  int last_dualcheck_ = 0;
  void enqueue_minor(int it, int t);
  int plan_major(int next, int t, int skip_first);
  DBuf<DevScalars> S_;
  DevScalars* hS_ = nullptr;  // pinned mirror
  // factor
  DBuf<double> rowscale_;
  DBuf<double> Kinv_, KinvT_, GW_, Pb_, Pinv_, ccv_, ccv2_;
  DBuf<int> pos_P_, row_R_, Ridx_, Pidx_, slackpos_, slackrow_, crp_, cci_, ccp_, ccr_, piv_, perm_, flag_, rowsA_,
      rowsB_, nA_, nB_;
  int k_ = 0;
  int eta_cap_ = 0;
  // slack substitution for dependent kernel columns
  bool repair_ = false;    // substitute (warm starts, refactors inside a solve) or fail
  bool repaired_ = false;  // the last refactor changed the basis
  bool warm_ = false;      // the solve starts from a given basis (exact DSE weights)
  // test hook (params.force_event)
  bool force_fired_ = false;
  int force_enq_it_ = -1;
  bool force_due(int event, int it);
  bool time_hit_ = false;  // the wall-clock limit ended the solve
  std::vector<hipEvent_t> gemm_ev_;  // time_refactor: event pairs around the trailing updates
  int repairs_ = 0;
  std::vector<int> fac_pos_P_, fac_row_R_;
  void factor_once(std::vector<std::pair<int, int>>& dep);
  void substitute_slacks(const std::vector<std::pair<int, int>>& dep);
  bool after_repair(int phase);
  int gj_mfma_ = 1;  // trailing update of the inversion on FP64 MFMA

  // host mirrors of the control state
  std::vector<double> c_host_, c_orig_;
  bool perturbed_ = false;
  int iteration_ = 0;
  int t_ = 0;  // etas since the last refactor
  int status_ = DS_NOT_SOLVED;
  bool solved_ = false;
  std::string message_;
  double Z_ = 0.0;
  int refactors_ = 0, restarts_ = 0, phase1_iters_ = 0, cleanups_ = 0;
  // per-step profile (params.profile): one event before the iteration and one after each step
  double step_ms_[DS_NPARTS] = {};
  std::vector<hipEvent_t> pev_;
  std::vector<int> pev_part_;
  void prof_event(int part);
  void prof_harvest();
  double prof_t0_ = 0.0;
  void prof_host_begin();
  void prof_host_end(int part);
  double solve_ms_ = 0.0, refactor_ms_ = 0.0;
  uint64_t rng_state_ = 0;
  // timing
  double t_start_ = 0.0, mark_ms_ = 0.0, spmv_ms_ = 0.0;
  int spmv_launches_ = 0;
  bool marked_ = false;
  double nnzN_mark_ = 0.0, nnzN_end_ = 0.0;
  std::vector<hipEvent_t> ev_;
  std::vector<int> ev_iter_;
  int ev_used_ = 0;
  void harvest_events(int last_done);
  void maybe_mark(int done);
  double nnz_nonbasic();

  DevModel M() const;
  DevIter I() const;
  DevFactor F() const;

  void pull();
  void push();
  void sync();
  void tick_limit(int it);

  void initBetaWeights(bool identity);
  void initReducedCosts();
  void dualSimplexInit();
  void phase1Init();
  bool checkPerturbNeed() const;
  void perturbCosts();
  void upload_costs();
  int phase1();
  bool phase2();
  bool dual_feasible();
  bool primal_feasible();
  void enqueue_body(int phase, int it, int skip_top, bool dualcheck, int t);
  void enqueue_stageB(int phase, int it, int t);
  int plan_chunk(int phase, int next, int t, int skip_first);
};

}  // namespace pds
