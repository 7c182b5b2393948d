// Batched small-LP solver: device view of a batch and host launchers.
#pragma once
#include "ds_common.hpp"

namespace pds {

constexpr int GJ_NB_MAX = 8;  // widest panel of the batched factorization

// A batch of LPs  min c^T x, A x <= b, x >= 0  (packing form), each held in
// CSC and CSR with per-LP offsets into concatenated arrays.  Variables of LP k
// are its n_k structurals then its m_k slacks (computational form).
struct BatchDev {
  int count;
  const int* m;
  const int* n;
  const long long* cp_off;   // start of LP k's CSC col_ptr (n+1 entries)
  const long long* rp_off;   // start of LP k's CSR row_ptr (m+1 entries)
  const long long* nz_off;   // start of LP k's nonzeros
  const long long* mo;       // start of LP k's m-sized slices
  const long long* vo;       // start of LP k's (n+m)-sized slices
  const int* colptr;
  const int* rowidx;
  const double* cval;
  const int* rowptr;
  const int* colidx;
  const double* rval;
  const double* c;           // n+m per LP (slack costs 0)
  const double* l;
  const double* u;
  const int8_t* bt;
  const int* order;          // LP indices of this launch's class, in processing order
  int order_count;
  const int* warm;           // optional starting basis (m per LP), nullptr = slack basis
  double repair_tol;         // > 0: slack substitution for pivots below it in the first factorization
  int* repairs;              // per LP: number of substitutions (may be null)
  int* status;
  double* obj;
  int* iters;
  int* basis_out;            // m per LP
  double* x_out;             // n+m per LP
  int* next;                 // work-queue counter of this launch
  double* gscratch;          // per-workgroup global inverse (mmax^2) when not in LDS
  int mmax, nmax;
  int inv_in_lds;
  int gj_nb;                 // factorization panel width (1..GJ_NB_MAX), sized to the LDS left
  int refactor_freq;
  int max_iter;
  int perturb;
  int cleanup;               // resume the dual simplex if the restored costs leave dual infeasibilities
  unsigned long long seed;
};

size_t batch_lds_bytes(int mmax, int nmax, int inv_in_lds, int gj_nb);
// Widest panel (GJ_NB_MAX, 4, 2, 1) whose LDS footprint fits `limit` bytes (0 if none).
int batch_panel_width(int mmax, int nmax, int inv_in_lds, size_t limit);
void launch_batch(hipStream_t s, const BatchDev& D, int grid, int block);
void launch_batch_reset(hipStream_t s, int* next);

}  // namespace pds
