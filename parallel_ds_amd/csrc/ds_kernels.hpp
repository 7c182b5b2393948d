// Host-callable launchers of the gfx950 kernels (asynchronous on `s`).
#pragma once
#include "ds_abi.h"
#include "ds_dev.hpp"

namespace pds {

constexpr int ETA_MAX = 256;  // eta file capacity (refactor interval + margin)

void launch_price(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it, int skip_top, int perturb_on);
// CHUZR whose last workgroup goes on with BTRAN's set-up, then BTRAN's output
// (rho): launch_price + launch_btran_unit in two launches instead of three.
void launch_price_btran(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase, int it,
                        int skip_top, int perturb_on, int t, double* rho);
void launch_dualcheck(hipStream_t s, const DevModel& M, const DevIter& I, int it);
void launch_btran_unit(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t, double* out,
                       int p_fixed);
void launch_spmv(hipStream_t s, const DevModel& M, const int* posN, const double* v, double alpha, double beta,
                 const double* cvec, double* out, int by_col, const DevScalars* S, int check,
                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_chuzc(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int harris, int it);
// Independent stages of one iteration in shared launches (no bound flips):
// pivot-row SpMV + tau's kernel product into tau0; CHUZC + tau's slack rows;
// then launch_col_update (the entering column's FTRAN with its etas, tau's
// etas, the updates).  tau0 is tau itself when t == 0.
void launch_stage_spmv_gemv(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, const int* posN,
                            const double* rho, const double* cvec, double* alpha_r, double* tau0);
// harris = 1: pass 1 beside the slack rows, then pass 2 in a launch of its own.
void launch_stage_chuzc_slack(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase,
                              int it, const double* rho, double* tau0, int harris = 0);
void launch_ftran_col(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int col, int t,
                      double* out, int check);
void launch_ftran_dense(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, const double* b,
                        double* out, int t, int mode, int check);
void launch_btran_dense(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, double* v,
                        double* y, int t);
void launch_update(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it);
void launch_dupdate_update(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it);
// The rest of a fused dual iteration after CHUZC: the entering column's FTRAN
// with its etas, tau's etas (tau0 -> tau), the dual update and the basis
// change; one launch up to 16,384 update rows, else two (the FTRANs, then
// launch_dupdate_update).
// split = 1: the two-launch form at any size (test hook).
void launch_col_update(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase, int it,
                       int t, const double* tau0, int split = 0);
// Primal simplex iterations of a warm start: entering column (Devex), the
// two-pass ratio test over the basic rows, and the update (after FTRAN, BTRAN
// and the pivot-row SpMV).
void launch_pprice(hipStream_t s, const DevModel& M, const DevIter& I, int it);
// The primal ratio test's two passes, the second one going on with BTRAN's
// set-up for the leaving row, then BTRAN's output (rho).
// split = 1: the two passes in launches of their own at any size (test hook).
void launch_pratio_btran(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int it, int t,
                         double* rho, int split = 0);
void launch_pupdate(hipStream_t s, const DevModel& M, const DevIter& I, int it);
// Bound-flipping ratio test: the flip loop after CHUZC, then the flip set, the
// column change, its FTRAN and the x_B / objective correction.
void launch_bfrt(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int harris, int t);
void launch_flips(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t);
// Test hooks (ds_params.force_event): DS_FORCE_STABILITY / DS_FORCE_DUAL_RESTART in the
// iteration chain, and a zeroed kernel column before a factorization.
void launch_force(hipStream_t s, const DevModel& M, const DevIter& I, int event);
void launch_force_zero_column(hipStream_t s, double* W, int k, int c);
// Multiple candidate rows per major iteration (params.candidates > 1).
void launch_price_multi(hipStream_t s, const DevModel& M, const DevIter& I, int it, int skip_top, int perturb_on,
                        int nc);
void launch_btran_cand(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t, int cand);
void launch_pick(hipStream_t s, const DevModel& M, const DevIter& I, int it, double psi);
void launch_rho_update(hipStream_t s, const DevModel& M, const DevIter& I, int nc);
void launch_set_primal(hipStream_t s, const DevModel& M, const DevIter& I);
void launch_mask_nonbasic(hipStream_t s, const DevModel& M, const DevIter& I, const double* src, double* wN);
void launch_rowspmv(hipStream_t s, const DevModel& M, const double* w, const double* b, double alpha, double beta,
                    double* out);
void launch_dot_obj(hipStream_t s, const DevModel& M, const DevIter& I);
void launch_phase1_inf(hipStream_t s, const DevModel& M, const DevIter& I);
void launch_gather_cb(hipStream_t s, const DevModel& M, const DevIter& I, double* out);
void launch_dse_exact(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F);
void launch_fill(hipStream_t s, double* p, double v, int n);

// Kernel inverse by blocked Gauss-Jordan (ds_refactor.hip).
struct GJWork {
  double* W;      // k x k work matrix (row-major, ld = k)
  double* Pb;     // k x NB saved panel
  int* piv;       // k
  int* perm;      // k
  int* flag;      // [0] singular flag, [1] count of dependent steps, [2..] those steps (GJ_DEP_CAP)
  double* Pinv;   // NB x NB inverse of the pivot block
  int* rowsA;     // tournament candidate lists (ceil(k/256)+1) x NB, two buffers
  int* rowsB;
  int* nA;
  int* nB;
  int use_mfma;   // trailing update on FP64 MFMA (1) or plain FMA (0)
  double* rowscale;  // k: largest |entry| of each row of K, kept in step with the row swaps
  hipEvent_t* gemm_ev = nullptr;  // optional: start/stop event pair around each panel's trailing update
};
constexpr int GJ_NB = 64;
constexpr int GJ_DEP_CAP = 4096;  // dependent kernel columns one factorization can report
void launch_build_kernel(hipStream_t s, const DevModel& M, const int* basis, const DevFactor& F, double* W);
void launch_gj_panels(hipStream_t s, int k, const GJWork& G);
void launch_gj_unscramble(hipStream_t s, int k, const GJWork& G, double* Kinv, double* KinvT);

// Plain copy of n 16-byte elements (counter calibration).
void launch_copy(hipStream_t s, const double2* a, double2* b, long long n);

}  // namespace pds
