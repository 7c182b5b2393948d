/*
 * CPU oracle for the dual simplex iteration — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's sequential solver
 * (src/cpu/sequentialDualSimplex.cpp, src/common/baseDualSimplex.{hpp,tpp},
 * src/cpu/linalg.cpp, src/cpu/matrix.cpp, src/common/problem.tpp).  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links it.
 *
 * Deviations from the reference, all deliberate and flagged by parameters:
 *   - the LU is project-written (slack elimination + dense partial-pivot LU of
 *     the structural kernel) in place of MKL PARDISO;
 *   - cost perturbation draws from a seeded SplitMix64 stream, not
 *     std::random_device (baseDualSimplex.tpp:244-245);
 *   - an optional iteration cap turns the reference's unbounded restart loops
 *     into an ITERATION_LIMIT status (MAX_ITER is defined but unused in the
 *     reference, baseDualSimplex.hpp:45);
 *   - Free rows get a Free slack (the reference pushes no type,
 *     problem.tpp:201-218);
 *   - a perturbation of exactly 0 (magnitude 100*EPS_D + PSI*c_j = 0, i.e.
 *     c_j = -1) is left at 0: the reference's clamping loop
 *     (baseDualSimplex.tpp:262-266) never ends on it;
 *   - a ratio-test candidate set whose ratios are all infinite (a Free
 *     candidate with alpha = 0, sequentialDualSimplex.cpp:361-365) is treated
 *     as empty; the reference indexes with the -1 that simpleRatioTest returns.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bound types, same order as the reference's BoundaryType (types.hpp:48-56). */
enum { ORA_FIXED = 0, ORA_FREE = 1, ORA_BOXED = 2, ORA_UPPER = 3, ORA_LOWER = 4 };

/* Termination status. */
enum { ORA_OPTIMAL = 0, ORA_PRIMAL_INFEASIBLE = 1, ORA_DUAL_INFEASIBLE = 2, ORA_ITERATION_LIMIT = 3 };

typedef struct {
  int refactor_freq;       /* REFACT_FREQ, 150 in the reference */
  int max_iterations;      /* 0 = no cap (reference behaviour) */
  int ratio_test;          /* 0 = textbook min-ratio (reference), 1 = Harris two-pass */
  int tie_rule;            /* 0 = sequential scan (reference), 1 = order-free (min, then max |alpha|, then lowest index) */
  int perturb;             /* 1 = reference rules (presolve check + cycling), 0 = never perturb */
  uint64_t perturb_seed;
  int verbose;
  int cleanup;             /* 0 = reference (stub primal correction after perturbation);
                              1 = resume the dual simplex when the original costs leave dual
                              infeasibilities */
} ora_params;

typedef struct {
  int status;
  int solved;
  int iterations;          /* shared phase-1 + phase-2 counter */
  int perturbed;
  int refactors;
  int phase1_restarts;
  double Z;                /* objective with the offset removed */
  int phase1_iterations;   /* iterations of the initial phase 1 */
  int bfrt_exhausted;      /* bound-flipping loops that flipped every candidate (the last one
                              then enters unflipped, sequentialDualSimplex.cpp:398-422) */
} ora_result;

/* Whole solve.  A is m x n (structurals only) in CSC.  Row types/ranges follow
 * Problem's 8-array constructor (problem.tpp:143-173); infinite bounds are +-inf
 * or |v| > 1e100.  x_out has n+m entries (structurals then slacks), basis_out m. */
int ora_solve(int m, int n, const int* col_ptr, const int* row_idx, const double* val,
              const double* costs, const double* lower_bound, const double* upper_bound,
              const int* bound_type, const double* lower_range, const double* upper_range,
              const int* range_type, const ora_params* params,
              ora_result* result, double* x_out, int* basis_out);
/* ora_solve from the starting basis basis_in (m variable ids, nonbasic set in
 * ascending order, exact DSE weights); NULL = the slack basis.  The reference
 * has no such entry: it replays the same algorithm from another basis, as the
 * batched solver's warm starts do. */
int ora_solve_from(int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                   const double* costs, const double* lower_bound, const double* upper_bound,
                   const int* bound_type, const double* lower_range, const double* upper_range,
                   const int* range_type, const int* basis_in, const ora_params* params, ora_result* result,
                   double* x_out, int* basis_out);

/* Matrix::dotUpdate (matrix.cpp:261-346).  A is m x ncols in CSC+CSR form built
 * from the given CSR.  method 0 = UPDATE (row-wise), 1 = UPDATE_T (column-wise). */
int ora_dot_update(int m, int ncols, const int* row_ptr, const int* col_idx, const double* vals,
                   const double* vec1, const double* vec2, double* sol, double alpha, double beta,
                   const int* cols, int ncols_sel, int method, int set);

/* Solve with the project LU of a dense square matrix (row-major):
 * transpose = 0 solves A x = b, 1 solves A^T x = b (Matrix::solve contract,
 * matrix.cpp:567-582). */
int ora_dense_solve(int k, const double* a, const double* b, double* x, int transpose);

/* linalg::PFIsolve (linalg.cpp:12-46): etas is t x m row-major, piv t. */
int ora_pfi_solve(int m, int t, const double* etas, const int* piv, const double* b, double* sol,
                  int transpose);

/* Textbook ratio test restated (baseDualSimplex.tpp:546-572) over candidate
 * positions F; returns q_idx or -1. */
int ora_simple_ratio_test(int nF, const int* F, const double* d_of_pos, const double* alpha,
                          int tie_rule);

/* Harris two-pass ratio test (ratio_test = 1) over candidate positions F:
 * pass 1 bounds the step by the EPS_D-relaxed ratios, pass 2 takes the largest
 * |alpha| among the candidates whose plain ratio is within that bound (lowest
 * position on ties); returns q_idx or -1. */
int ora_harris_ratio_test(int nF, const int* F, const double* d_of_pos, const double* alpha);

/* DSEPricing (baseDualSimplex.tpp:470-501) over arrays by basis position:
 * *p_idx = argmax delta^2/beta (strict, ascending positions) among the
 * variables checkVarIsInfeas flags with |delta| > 1e-8, or -1. */
int ora_dse_pricing(int m, const signed char* bt, const double* x, const double* l, const double* u,
                    const double* beta, int* p_idx, double* delta);

#ifdef __cplusplus
}
#endif
