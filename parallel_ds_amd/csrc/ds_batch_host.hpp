// Host side of the batched small-LP solver.
#pragma once
#include <vector>

#include "ds_abi.h"
#include "ds_batch.hpp"
#include "ds_solver.hpp"

namespace pds {

class BatchSolver {
 public:
  explicit BatchSolver(int device);
  ~BatchSolver();
  void load(int count, const int* m, const int* n, const long long* nnz_off, const int* col_ptr, const int* row_idx,
            const double* val, const double* b, const double* c);
  void set_warm(const int* basis);
  void set_bases(const int* basis);
  double solve(const ds_batch_params& p);
  double replay();
  void results(int* status, double* obj, int* iters, int* basis, double* x);
  void tune(int sample, ds_batch_params& best, double& best_ms, int& tried);
  int count() const { return count_; }
  long long vars() const { return sum_v_; }
  long long rows() const { return sum_m_; }
  int mmax() const { return mmax_all_; }
  int nmax() const { return nmax_all_; }
  hipStream_t stream() const { return st_; }
  const int* dev_m() const { return m_.p; }
  const int* dev_n() const { return n_.p; }
  const long long* dev_vo() const { return vo_.p; }
  const std::vector<int>& host_m() const { return hm_; }
  const std::vector<int>& host_n() const { return hn_; }
  // Learned warm start: per-variable features (vars() x WS_NF, FP32), 0/1
  // labels of the last solve's bases, starting bases from per-variable scores
  // (top m per LP), and the pivot tolerance below which their first
  // factorization substitutes slacks.  Device pointers, stream-ordered.
  void features(float* F_dev);
  void labels(float* y_dev);
  void select_warm(const float* z_dev);
  void set_repair_tol(double t) { repair_tol_ = t; drop_graph(); }
  void repairs(int* out);
  void get_warm(int* out);

 private:
  struct Class {
    int first, count, mmax, nmax, grid, counter, gj_nb = 1;
    bool inv_in_lds;
  };
  int device_ = 0;
  int cus_ = 1;
  hipStream_t st_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  int count_ = 0;
  long long sum_m_ = 0, sum_v_ = 0;
  int mmax_all_ = 0, nmax_all_ = 0;
  std::vector<int> hm_, hn_;
  bool warm_on_ = false;
  double repair_tol_ = 1e-7;
  bool solved_ = false;  // every LP of the batch has been through a full solve
  std::vector<Class> classes_;
  DBuf<int> m_, n_, colptr_, rowidx_, rowptr_, colidx_, order_, warm_, status_, iters_, basis_, next_, repairs_;
  DBuf<long long> cp_off_, rp_off_, nz_off_, mo_, vo_;
  DBuf<double> cval_, rval_, c_, l_, u_, obj_, x_, gscratch_;
  DBuf<int8_t> bt_;

  void drop_graph();
  BatchDev make_dev() const;
  void plan(const std::vector<int>& items, const ds_batch_params& p);
  void enqueue(const ds_batch_params& p);
  double run(const std::vector<int>& items, const ds_batch_params& p, bool use_graph);
};

}  // namespace pds
