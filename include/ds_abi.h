/*
 * C-ABI of the MI355X dual simplex (libpds.so).
 *
 * Plain pointers and sizes only; every function returns 0 (DS_OK) or a
 * negative ds_err and never throws.  One host thread drives one context; a
 * context owns one HIP stream and all device workspaces.  Host-pointer
 * arguments are copied before the call returns.
 *
 * What each entry replaces in the reference (file:line under its src/):
 *   ds_load_problem   Problem(bound_type, range_type, costs, lower_range,
 *                     upper_range, lower_bound, upper_bound, A) +
 *                     transformToComputeForm()       common/problem.tpp:143-226
 *   ds_init           SequentialDualSimplex::initDualSimplex
 *                                                   cpu/sequentialDualSimplex.cpp:4-25
 *                     (GPU analogue: CudaDualSimplex::initDualSimplex, gpu/cudaDualSimplex.cpp:20-81)
 *   ds_presolve       BaseDualSimplex::presolve("minInfeas") common/baseDualSimplex.tpp:5-30
 *   ds_solve          BaseDualSimplex::solve + postProcess  common/baseDualSimplex.tpp:123-180
 *   ds_get_solution   Problem::getSolution / LPsolution    common/problem.tpp:243-259,
 *                                                          common/LPsolution.hpp:22-44
 *   ds_get_basis      basis_indexes (protected in the reference, baseDualSimplex.hpp:80)
 *   ds_load_mps       Problem::readMps (via Clp) + transformToComputeForm
 *                                                   common/problem.tpp:331-415, 176-226
 *   ds_refactor       reFactorize                       cpu/sequentialDualSimplex.cpp:155-160
 *   ds_btran          BTran -> solveLinSys(e_p, rho, true)  cpu/sequentialDualSimplex.cpp:89-92
 *   ds_ftran          solveLinSys(b, x, false)          cpu/sequentialDualSimplex.cpp:28-48
 *   ds_ftran_column   FTran(q, alpha_q)                 cpu/sequentialDualSimplex.cpp:107-110
 *   ds_chuzr          DSEPricing + checkVarIsInfeas     common/baseDualSimplex.tpp:443-501
 *   ds_pivot_row      pivotRow -> Matrix::dotUpdate(UPDATE_T)
 *                                                   cpu/sequentialDualSimplex.cpp:95-104,
 *                                                   cpu/matrix.cpp:273-287; launcher spmvUpdateInc,
 *                                                   gpu/utillities.hpp:130-148
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ds_ctx ds_ctx;

/* Bound / row types, order of the reference's BoundaryType (common/types.hpp:48-56). */
enum { DS_FIXED = 0, DS_FREE = 1, DS_BOXED = 2, DS_UPPER = 3, DS_LOWER = 4 };

/* Termination status; the reference's messages are kept by ds_get_solution. */
enum {
  DS_OPTIMAL = 0,            /* "optimal solution" */
  DS_PRIMAL_INFEASIBLE = 1,  /* "dual unbounded" */
  DS_DUAL_INFEASIBLE = 2,    /* "dual infeasible" */
  DS_ITERATION_LIMIT = 3,    /* max_iterations or time_limit_ms reached */
  DS_NOT_SOLVED = 4
};

/* Error codes. */
enum {
  DS_OK = 0,
  DS_EINVAL = -1,
  DS_EHIP = -2,
  DS_ESINGULAR = -3,
  DS_EUNSUPPORTED = -4,
  DS_ESTATE = -5,
  DS_EINTERNAL = -6
};

typedef struct {
  int refactor_freq;      /* REFACT_FREQ (reference default 150) */
  int max_iterations;     /* 0 = no cap (reference behaviour) */
  int ratio_test;         /* 0 = textbook min-ratio (reference), 1 = Harris two-pass */
  int perturb;            /* 1 = reference perturbation rules, 0 = never perturb */
  uint64_t perturb_seed;  /* seed of the cost perturbation stream */
  int chunk;              /* iterations queued per host synchronisation (>= 1) */
  int verbose;
  int mark_iteration;     /* > 0: synchronise after this iteration and record the time (ds_result.mark_ms) */
  int time_kernels;       /* k > 0: time the pivot-row SpMV of every k-th iteration after the mark with HIP events */
  int cleanup;            /* 1 (default): after a cost perturbation, if the restored costs leave dual
                             infeasibilities, resume the dual simplex from the final basis (phase 1, then
                             phase 2).  0: the reference's behaviour (its primal correction is a stub,
                             sequentialDualSimplex.cpp:113-116, so that basis is returned) */
  int profile;            /* 1: per-step device times in ds_result.step_ms (HIP events between the steps
                             of every iteration, one iteration per host synchronisation) */
  int force_event;        /* test hook, 0 = off: force one recovery path at iteration force_at:
                             DS_FORCE_STABILITY, DS_FORCE_DUAL_RESTART or DS_FORCE_SINGULAR */
  int force_at;
  double time_limit_ms;   /* > 0: end a solve that has run this long (phase 1 + phase 2 wall time) with
                             DS_ITERATION_LIMIT and the message "time limit"; 0 = none (reference) */
  int candidates;         /* phase 2 leaving-row candidates per major iteration, 1..8; 1 (default) = the
                             sequential solver.  > 1: the minor iterations of the reference's parallel
                             dual simplex (parallelDualSimplex.cpp:431-485, 747-1124) on one GPU: the
                             major pricing keeps the rows of largest delta^2/beta, their BTRANs run back to
                             back, each minor iteration takes the best still-eligible row with no BTRAN of
                             its own, and the others' rows follow each basis change */
  double candidate_psi;   /* a candidate drops out when its weight falls below psi times its previous
                             one (the reference's psi, main.cpp:46-53); default 0.9 */
  int warm_primal;        /* starts from ds_init_basis only, no counterpart in the reference: 1 = before
                             the dual simplex, make the starting basis primal feasible by moving each
                             violated bound of a basic variable warm_shift * (1 + |x_j|) past its value,
                             run primal simplex iterations (Devex pricing, two-pass ratio test) until no
                             reduced cost is dual infeasible, restore the bounds, and let the dual
                             simplex (phase 1 if needed, then phase 2) finish from that dual feasible
                             basis.  Models with Free or Boxed variables skip it.  0 (default): the dual
                             simplex starts from the given basis directly */
  double warm_shift;      /* relative bound shift of warm_primal; default 1e-3 */
  int warm_primal_cap;    /* most primal iterations before the dual simplex takes over; 0 = 4 m */
  int split_launches;     /* test hook, 0 = off: 1 runs the iteration stages that share one launch below
                             a size limit (the ratio test's two passes, m <= 8,192; the entering
                             column's FTRAN with the dual update, n <= 16,384) in the launches larger
                             LPs take.  The results are the same bit for bit */
} ds_params;

/* Recovery paths a test can force from the host (sequentialDualSimplex.cpp):
 *  STABILITY     the stability check fails (|alpha_q[p] - alpha_r[q]| test, :428-433): the
 *                recorded alpha_r[q] is moved by 1e-6 relative after CHUZC, so the basis is
 *                refactored before the update, which then uses the exact alpha_q;
 *  DUAL_RESTART  at the next iteration that is a multiple of 40 from force_at on, one nonbasic
 *                Lower variable's reduced cost is made negative before the periodic dual check
 *                (:285-293), which then refactors, reruns phase 1 and re-initialises;
 *  SINGULAR      the first refactorization at or after force_at sees one basic structural
 *                column as zero: the column is replaced by a slack and the iterate recomputed. */
enum { DS_FORCE_NONE = 0, DS_FORCE_STABILITY = 1, DS_FORCE_DUAL_RESTART = 2, DS_FORCE_SINGULAR = 3 };

/* Steps of ds_result.step_ms, the parts of the reference's TimeProfiler
 * (common/types.hpp:12-27, common/timeProfiler.cpp:22-148).  The GPU fuses
 * the candidate set into the choice (both under RATIO_TEST2) and the
 * reduced-cost update with the basis change (under UPDATE_RED_COSTS; a basis
 * change finished after a stability refactor is under BASIS_UPDATE). */
enum {
  DS_PART_INIT = 0, DS_PART_FACTOR, DS_PART_RESTORE, DS_PART_BTRAN, DS_PART_PIVOT_ROW, DS_PART_RATIO_TEST1,
  DS_PART_RATIO_TEST2, DS_PART_FTRAN, DS_PART_FTRAN_BFRT, DS_PART_FTRAN_BETA, DS_PART_PRICING,
  DS_PART_UPDATE_RED_COSTS, DS_PART_BASIS_UPDATE, DS_NPARTS
};

typedef struct {
  int status;
  int solved;
  int iterations;         /* phase 1 + phase 2, the reference's shared counter */
  int phase1_iterations;  /* iterations spent in the initial phase 1 */
  int perturbed;
  int refactors;
  int restarts;
  double Z;               /* objective (offset removed) */
  double solve_ms;        /* device-resident solve wall time (phase 1 + phase 2) */
  double refactor_ms;     /* part of solve_ms spent in refactorization */
  double mark_ms;         /* solve time up to the end of iteration mark_iteration */
  double spmv_ms;         /* summed event time of the timed pivot-row SpMV launches */
  int spmv_launches;
  int kernel_k;           /* basis kernel size at the last refactor */
  double nnz_nonbasic_mark; /* nnz of the nonbasic columns at the mark and at the end */
  double nnz_nonbasic_end;
  int cleanups;           /* dual simplex resumptions after restoring perturbed costs */
  int repairs;            /* structural columns replaced by slacks in singular factorizations */
  double step_ms[DS_NPARTS]; /* with params.profile: time per step (DS_PART_*) */
  int primal_iterations;  /* params.warm_primal: primal simplex iterations (part of `iterations`) */
  int shifted_bounds;     /* params.warm_primal: bounds of basic variables moved for primal feasibility */
} ds_result;

void ds_default_params(ds_params* p);

int ds_create(int device, ds_ctx** out);
int ds_destroy(ds_ctx* ctx);
const char* ds_last_error(ds_ctx* ctx);

/* A is m x n (structurals only) in CSC; infinite bounds as +-inf or |v|>1e100. */
int ds_load_problem(ds_ctx* ctx, int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                    const double* costs, const double* lower_bound, const double* upper_bound,
                    const int* bound_type, const double* lower_range, const double* upper_range,
                    const int* range_type);
int ds_set_params(ds_ctx* ctx, const ds_params* p);

/* Problem::readMps + transformToComputeForm (common/problem.tpp:331-415,
 * 176-226): reads a free- or fixed-format MPS file with this library's own
 * parser (the reference reads it through COIN-OR Clp) and loads the model as
 * ds_load_problem does; the objective constant (minus the RHS entry of the
 * objective row) is removed from Z as getSolution does (problem.tpp:256).
 * *m and *n (may be NULL) receive the model's size. */
int ds_load_mps(ds_ctx* ctx, const char* path, int* m, int* n);
/* Objective constant for a model loaded with ds_load_problem (Problem::_offset). */
int ds_set_offset(ds_ctx* ctx, double offset);

/* The parsed model itself, for callers that keep their own copy (host only, no
 * GPU needed): ds_mps_read parses (on failure err receives the message),
 * ds_mps_dims gives m, n and nnz, ds_mps_get copies the 8 arrays of the
 * reference's Problem constructor (A in CSC) and the objective constant. */
typedef struct ds_mps ds_mps;
int ds_mps_read(const char* path, ds_mps** out, char* err, int err_len);
int ds_mps_dims(const ds_mps* mps, int* m, int* n, int* nnz);
int ds_mps_get(const ds_mps* mps, int* col_ptr, int* row_idx, double* val, double* costs, double* lower_bound,
               double* upper_bound, int* bound_type, double* lower_range, double* upper_range, int* range_type,
               double* offset);
int ds_mps_free(ds_mps* mps);
int ds_init(ds_ctx* ctx);
/* ds_init from a given starting basis (m variable ids; the nonbasic set is the
 * ascending complement) instead of the slack basis.  No counterpart in the
 * reference (its solver always starts from the slack basis,
 * sequentialDualSimplex.cpp:230-250); the algorithm that follows is the same,
 * with exact DSE weights of the starting basis (initBetaWeights(false),
 * baseDualSimplex.tpp:41-68).  A column of the basis that the factorization
 * finds dependent (pivot at or below 1e-13 of its row's scale) is replaced by
 * the slack of the row chosen for it; ds_result.repairs counts them.  The same
 * substitution keeps a solve going when a refactorization inside it meets a
 * numerically singular basis; the iterate is then recomputed from the new
 * basis.  ds_set_basis (hook level) still fails with DS_ESINGULAR. */
int ds_init_basis(ds_ctx* ctx, const int* basis);
int ds_presolve(ds_ctx* ctx, const char* method, int* phase1_status); /* 0 Solved, 1 DualInfeas */
int ds_solve(ds_ctx* ctx);
int ds_get_result(ds_ctx* ctx, ds_result* r);
int ds_get_solution(ds_ctx* ctx, double* Z, int* solved, char* message, int message_len, double* x /* n+m */);
int ds_get_basis(ds_ctx* ctx, int* basis /* m */);

/* Whole solve in one call: load + init + presolve("minInfeas") + solve. */
int ds_solve_lp(ds_ctx* ctx, int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                const double* costs, const double* lower_bound, const double* upper_bound,
                const int* bound_type, const double* lower_range, const double* upper_range,
                const int* range_type, const ds_params* p, ds_result* r, double* x, int* basis);

/* Hook-level entries on the context's current basis (host buffers, synchronous). */
int ds_set_basis(ds_ctx* ctx, const int* basis /* m variable ids; nonbasic set = ascending complement */);
int ds_refactor(ds_ctx* ctx);
int ds_btran(ds_ctx* ctx, int p_idx, double* rho /* m, by row */);
int ds_ftran(ds_ctx* ctx, const double* b /* m, by row */, double* x /* m, by position */);
int ds_ftran_column(ds_ctx* ctx, int j, double* x /* m */);
int ds_pivot_row(ds_ctx* ctx, const double* rho /* m */, double* alpha /* n, by nonbasic position */);
int ds_get_nonbasis(ds_ctx* ctx, int* nonbasis /* n */);

/* CHUZR on caller-owned DEVICE arrays of the m basic variables, by basis
 * position: bound type (DS_FIXED..DS_LOWER), value, bounds and DSE weight.
 * Replaces BaseDualSimplex::DSEPricing + checkVarIsInfeas
 * (common/baseDualSimplex.tpp:443-501): the position maximising delta^2/beta
 * among variables out of bounds (fallthrough rules of checkVarIsInfeas, |delta|
 * above 1e-8), ties to the lowest position.  Writes the position (-1 when no
 * variable is out of bounds) to the device int *p_idx and the signed
 * infeasibility delta (x - bound; delta < 0 is the reference's is_lower) to the
 * device double *delta.  Runs on the context's stream and returns when done;
 * its scratch (per-workgroup partials) is allocated on the first call. */
int ds_chuzr(ds_ctx* ctx, int m, const signed char* bt_B, const double* x_B, const double* l_B, const double* u_B,
             const double* beta, int* p_idx, double* delta);

/* ds_chuzr without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as ds_chuzr; its scratch (per-workgroup partials and an arrival
 * counter, 8 KiB) is a static device array that each GPU holds until the
 * process ends, set up by ds_create.  m above DS_CHUZR_NOALLOC_MAX_M is
 * rejected with DS_EINVAL.  Calls that overlap on one GPU (from several
 * contexts or threads) share that scratch and race; the counter is cleared at
 * the start of every call. */
#define DS_CHUZR_NOALLOC_MAX_M (1 << 28)
int ds_chuzr_noalloc(ds_ctx* ctx, int m, const signed char* bt_B, const double* x_B, const double* l_B,
                     const double* u_B, const double* beta, int* p_idx, double* delta);

/* Timing of the pivot-row SpMV on the context's current basis and a given rho:
 * `reps` launches, HIP events on the context stream; returns the mean ms per
 * launch in *ms and the algorithmic bytes of one launch in *bytes. */
int ds_time_pivot_row(ds_ctx* ctx, const double* rho, int reps, double* ms, double* bytes);

/* Per-launch timing of the pivot-row SpMV (start/stop events around each of
 * `reps` launches): *warm_ms back to back, as inside the solver; *cold_ms with
 * `flush_bytes` streamed through a copy kernel before each launch (0 skips the
 * cold pass and returns 0).  *bytes = algorithmic bytes of one launch.
 * Measurement entry; no counterpart in the reference. */
int ds_time_pivot_row_ext(ds_ctx* ctx, const double* rho, int reps, long long flush_bytes, double* warm_ms,
                          double* cold_ms, double* bytes);

/* ---- Batched small LPs: one persistent launch, one workgroup per LP ----
 * No counterpart in the reference (it solves one LP per process,
 * src/main.cpp:55-92); each LP follows the same algorithm and rules as
 * ds_solve.  LPs are packing LPs  min c^T x, A x <= b, x >= 0.  Arrays are
 * concatenated over the batch: col_ptr holds n_k+1 entries per LP (each
 * starting at 0), row_idx/val hold nnz_off[k]..nnz_off[k+1]-1, b holds m_k
 * entries and c n_k entries per LP. */
typedef struct {
  int refactor_freq;      /* reinversion interval (default 150) */
  int max_iterations;     /* per LP, 0 = no cap */
  int perturb;
  uint64_t perturb_seed;
  int block;              /* workgroup size: 128, 256 or 512 */
  int lds_inv_max_m;      /* LPs with m <= this keep B^{-1} in LDS */
  int wgs_per_cu;         /* resident workgroups per CU of the persistent grid */
  int use_graph;          /* capture the launch in a hipGraph */
  int cleanup;            /* 1 (default): after a cost perturbation, if the original costs leave
                             dual infeasibilities, resume the dual simplex from the final basis.
                             0: the reference's behaviour (its primal correction is a stub,
                             sequentialDualSimplex.cpp:113-116, so that basis is returned) */
} ds_batch_params;

void ds_batch_default_params(ds_batch_params* p);
int ds_batch_load(ds_ctx* ctx, int count, const int* m, const int* n, const long long* nnz_off, const int* col_ptr,
                  const int* row_idx, const double* val, const double* b, const double* c);
/* Starting bases (m_k variable ids per LP, concatenated); NULL restores the slack basis. */
int ds_batch_set_warm(ds_ctx* ctx, const int* basis);
/* Bases of the batch's LPs (m_k variable ids per LP, concatenated) found by
 * another solver, e.g. ds_solve for LPs beyond the batched kernel's LDS limit:
 * they take the place of a batched solve's bases as the labels that
 * ds_batch_labels and ds_mlp_fit read. */
int ds_batch_set_bases(ds_ctx* ctx, const int* basis);
int ds_batch_solve(ds_ctx* ctx, const ds_batch_params* p, double* ms);
int ds_batch_replay(ds_ctx* ctx, double* ms);  /* relaunch the captured graph */
int ds_batch_results(ds_ctx* ctx, int* status, double* obj, int* iters, int* basis /* sum m */,
                     double* x /* sum (n+m) */);
/* Autotuner: times workgroup size x LDS threshold x occupancy on a sample of
 * `sample` LPs, then the refactor interval (150, 300, 600) at the best shape;
 * every candidate must reproduce the statuses and objectives (1e-9) of the
 * first.  Writes the fastest into *best (its other fields are kept). */
int ds_batch_tune(ds_ctx* ctx, int sample, ds_batch_params* best, double* best_ms, int* tried);

/* ---- Learned warm start ----
 * No counterpart in the reference (its solver always starts from the slack
 * basis, sequentialDualSimplex.cpp:230-250).  A per-variable MLP (16 features
 * -> 3 hidden layers of width 256 -> logit of "basic at the optimum"; BF16 on
 * the matrix cores, FP32 accumulation and master weights) scores every
 * variable of every LP of a batch; each LP starts from its m best-scored
 * variables, repaired by slack substitution in its first factorization;
 * the dual simplex then certifies optimality as from any basis. */
typedef struct ds_mlp ds_mlp;

int ds_mlp_create(int device, int n_features /* 1..32 */, uint64_t seed, ds_mlp** out);
int ds_mlp_destroy(ds_mlp* net);
const char* ds_mlp_last_error(ds_mlp* net);
long long ds_mlp_param_count(void);
/* FP32 parameter vector: W1[256x32] b1[256] W2[256x256] b2 W3[256x256] b3 w4[256] b4. */
int ds_mlp_get_params(ds_mlp* net, float* params);
int ds_mlp_set_params(ds_mlp* net, const float* params);
/* Feature standardisation (x - mu) / sd applied before the first layer. */
int ds_mlp_set_norm(ds_mlp* net, const float* mu, const float* sd);
int ds_mlp_get_norm(ds_mlp* net, float* mu, float* sd);
/* The family prior the last ds_mlp_fit found (feature 16: per variable index,
 * the fraction of the training bases in which it is basic; set only when every
 * training LP has the same m and n).  *n receives its length (0 = none); the
 * first min(cap, *n) values go to out. */
int ds_mlp_get_prior(ds_mlp* net, float* out, int cap, int* n);
/* Logits of S feature rows (S x n_features, host buffers). */
int ds_mlp_infer(ds_mlp* net, int S, const float* features, float* logits);
/* Mean weighted BCE-with-logits loss of S rows and its gradient w.r.t. the
 * parameter vector (host buffers; weights may be NULL = all 1). */
int ds_mlp_grad(ds_mlp* net, int S, const float* features, const float* labels, const float* weights, double* loss,
                float* grad);
int ds_mlp_adam_step(ds_mlp* net, float lr);  /* one Adam step on the last gradient */
/* Trains on the batch's last solved bases: features and labels of every
 * variable of every LP, standardisation fitted to them, `epochs` shuffled
 * passes of Adam with minibatches of `minibatch` rows.  epoch_loss (may be
 * NULL) receives the mean loss of each epoch, *ms the device time. */
int ds_mlp_fit(ds_mlp* net, ds_ctx* batch, int epochs, int minibatch, float lr, uint64_t seed, double* epoch_loss,
               double* ms);
/* Device time (ms) and MFMA flops of the training GEMMs so far. */
int ds_mlp_train_stats(ds_mlp* net, double* ms, long long* flops);

/* Per-variable features (sum(n_k+m_k) x 17, FP32; column 16, the family prior,
 * is 0 here and filled only when a trained network predicts) and 0/1 labels of the last
 * solved bases (sum(n_k+m_k)), host copies. */
int ds_batch_features(ds_ctx* ctx, float* features);
int ds_batch_labels(ds_ctx* ctx, float* labels);
/* Starting bases from per-variable scores (sum(n_k+m_k), host): the m_k best
 * of each LP.  ds_batch_warm_start does the same on the device from the
 * network's scores; *ms = features + inference + selection wall time,
 * *infer_ms = the inference kernel alone. */
int ds_batch_warm_from_scores(ds_ctx* ctx, const float* scores);
int ds_batch_warm_start(ds_ctx* ctx, ds_mlp* net, double* ms, double* infer_ms);
int ds_batch_get_warm(ds_ctx* ctx, int* basis /* sum m */);
/* Pivot magnitude below which the first factorization of a starting basis
 * substitutes a slack (default 1e-7; 0 disables: singular bases fail). */
int ds_batch_set_repair_tol(ds_ctx* ctx, double tol);
int ds_batch_repairs(ds_ctx* ctx, int* count /* per LP */);

/* Factorizations of the context's current basis, `reps` times (measurement
 * entry, no counterpart in the reference): *refactor_ms = mean wall time of one
 * refactorization; *gemm_ms = device time of its FP64 MFMA trailing updates
 * (event pairs per panel); *gemm_flops = their executed flops, 2 (k - nb) k nb
 * per panel of nb columns; *k = kernel size. */
int ds_time_refactor(ds_ctx* ctx, int reps, double* refactor_ms, double* gemm_ms, double* gemm_flops, int* k);

/* Counter calibration: one plain copy kernel over `bytes` (a multiple of 16)
 * between two fresh device buffers; *ms = its duration. */
int ds_debug_copy(long long bytes, double* ms);

/* Trailing-update tile check: C(64x64, row-major) -= P(:, :nb) * B(nb x 64), where
 * P is 64x64 row-major (columns >= nb ignored) and B is nb x 64 row-major; the
 * FP64 MFMA kernel when use_mfma, else the plain FMA kernel. */
int ds_debug_gemm_tile(double* C, const double* P, const double* B, int nb, int use_mfma);

#ifdef __cplusplus
}
#endif
