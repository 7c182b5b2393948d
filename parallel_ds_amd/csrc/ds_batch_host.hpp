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
  double solve(const ds_batch_params& p);
  double replay();
  void results(int* status, double* obj, int* iters, int* basis, double* x);
  void tune(int sample, ds_batch_params& best, double& best_ms, int& tried);
  int count() const { return count_; }

 private:
  struct Class {
    int first, count, mmax, nmax, grid, counter;
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
  std::vector<Class> classes_;
  DBuf<int> m_, n_, colptr_, rowidx_, rowptr_, colidx_, order_, warm_, status_, iters_, basis_, next_;
  DBuf<long long> cp_off_, rp_off_, nz_off_, mo_, vo_;
  DBuf<double> cval_, rval_, c_, l_, u_, obj_, x_, gscratch_;
  DBuf<int8_t> bt_;

  void drop_graph();
  void plan(const std::vector<int>& items, const ds_batch_params& p);
  void enqueue(const ds_batch_params& p);
  double run(const std::vector<int>& items, const ds_batch_params& p, bool use_graph);
};

}  // namespace pds
