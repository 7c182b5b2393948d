/*
 * Seeded synthetic LP generators (host-only part of libpds.so).
 *
 * The reference reads Kennington MPS files through COIN-OR Clp
 * (src/common/problem.tpp:331-415); the workloads measured here are seeded
 * synthetic LPs instead, built by these functions.  All LPs are
 *   min c^T x  s.t.  rows,  x >= 0
 * with A returned in CSC (int32 indices, FP64 values).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A(i,j) present with probability `density`, value U(0.1,1) (a column left
 * empty gets one entry at a uniform random row, so the LP is bounded);
 * x0 ~ U(0,1)^n; b = A x0 + U(0,1)^m; c = -U(0,1)^n.  Rows are A x <= b.
 * Returns nnz, -2 if `cap` is too small, -1 on bad arguments. */
int ds_gen_bernoulli_lp(int m, int n, double density, uint64_t seed, int* col_ptr, int* row_idx,
                        double* val, int cap, double* b, double* c);

/* Same value recipe with exactly k distinct uniformly random rows per column
 * (row_idx needs n*k entries).  Returns nnz = n*k. */
int ds_gen_kcol_lp(int m, int n, int k, uint64_t seed, int* col_ptr, int* row_idx, double* val,
                   double* b, double* c);

/* Balanced transportation LP: S sources x D sinks, variable x_ij is column
 * i*D + j with a 1 in supply row i and in demand row S+j; supplies U{1..100},
 * demands U{1..100} rescaled to the total supply (largest remainder); costs
 * U{1..1000}.  All S+D rows are equalities with right-hand side `rhs`.
 * Returns nnz = 2*S*D. */
int ds_gen_transport(int S, int D, uint64_t seed, int* col_ptr, int* row_idx, double* val,
                     double* rhs, double* c);

/* out[i] = base[i] * exp(sigma * N(0,1)), normals by Box-Muller. */
int ds_gen_lognormal(int len, double sigma, uint64_t seed, const double* base, double* out);

#ifdef __cplusplus
}
#endif
