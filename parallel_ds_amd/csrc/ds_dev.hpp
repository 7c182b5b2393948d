// Device-side views of the solver state, passed by value to every kernel.
#pragma once
#include "ds_common.hpp"

namespace pds {

// LP in computational form: A <- [A | I] (CSC and CSR), costs, bounds, types.
struct DevModel {
  int m, n, N, nnz;
  const int* cptr;
  const int* ridx;
  const double* cval;
  const int* rptr;
  const int* cidx;
  const double* rval;
  const double* c;
  const double* l;
  const double* u;
  const int8_t* bt;
};

// Iterate: primal/dual values, basis, work vectors, eta file, scalars.
struct DevIter {
  double* x;         // N, by variable
  double* d;         // N, by variable
  double* beta;      // m, DSE weights by basis position
  int* basis;        // m
  int* nonbasis;     // n
  int* posN;         // N: position in `nonbasis`, -1 when basic
  double* f;         // m, phase-1 pricing vector
  double* rho;       // m, by row
  double* alpha_r;   // n, pivot row by nonbasic position
  double* alpha_q;   // m, entering column by basis position
  double* tau;       // m
  double* w1;        // m work
  double* w2;        // m work
  double* wN;        // N work
  double* etas;      // eta_cap x m, dense eta columns
  int* eta_pos;      // eta_cap pivot positions
  double* E;         // eta_cap x eta_cap, E[k][j] = eta_k[eta_pos[j]]
  double* MF;        // eta_cap x eta_cap, inverse of the unit triangular eta-pivot matrix
  int* eta_last;     // eta_cap, previous eta index on the same pivot position (-1 if none)
  double* x0;        // m, FTRAN result before the etas
  int* cand_i;       // textbook CHUZC candidate lists, TB_LIST per workgroup
  double* cand_th;
  double* cand_a;
  // Bound-flipping ratio test (models with boxed variables only):
  unsigned char* bmark;    // n: positions dropped by the flip loop (cleared after the loop)
  int* blist;              // n: those positions, in drop order
  signed char* flipc;      // n: +1 flips lower -> upper, -1 upper -> lower, 0 none
  int* flist;              // n: flipping positions in ascending order
  double* bcc;             // m: column change sum_j (+-)(u_j - l_j) a_j of the flips
  double* dxb;             // m: B^{-1} bcc
  // Multiple candidate rows per major iteration (params.candidates > 1):
  int* cand_pos;           // CAND_MAX basis positions
  double* cand_w;          // their last weights delta^2 / beta
  int* cand_act;           // still eligible
  double* rho_c;           // CAND_MAX x m: their BTRAN rows, kept current through the minor iterations
  // Primal simplex iterations of a warm start (params.warm_primal):
  double* devex;           // N: Devex reference weights by variable
  int eta_cap;
  unsigned char* blocked;  // N
  int* blocked_list;       // m: rows blocked by phase 1 since the last non-empty ratio test
  int* list_idx;           // BTRAN right-hand-side list (kernel index)
  double* list_val;
  int* ylist_row;          // BTRAN slack-row list
  double* ylist_val;
  int* list_n;             // [0] = rhs list length, [1] = slack list length
  int list_cap;
  DevScalars* S;
  double* part_d;          // per-workgroup partials
  int* part_i;
  int part_cap;
};

// Factor of the basis B0 at the last refactor: basic slack columns are
// eliminated symbolically; the k x k kernel K = A[R, basic structurals] is
// held as an explicit inverse (row-major) and its transpose.
struct DevFactor {
  int k;
  const double* Kinv;
  const double* KinvT;
  const int* pos_P;            // k: basis position of kernel column c
  const int* row_R;            // k: row of kernel row i
  const int* Ridx_of_row;      // m: kernel row index, -1 for rows covered by a basic slack
  const int* Pidx_of_pos;      // m: kernel column index, -1 for slack positions
  const int* slackpos_of_row;  // m: position of the basic slack of a covered row, else -1
  const int* slackrow_of_pos;  // m: row of the slack at a slack position, else -1
  const int* crp;              // m+1: coupling block C = A[covered rows, basic structurals] by rows
  const int* cci;              //   kernel column index
  const double* ccv;
  const int* ccp;              // k+1: C by kernel columns
  const int* ccr;              //   row
  const double* ccv2;
};

}  // namespace pds
