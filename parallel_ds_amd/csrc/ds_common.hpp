// Definitions shared by the HIP kernels and the host solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pds {

// Solver constants, values of the reference's baseDualSimplex.hpp:23-46.
constexpr double EPS_BOUND = 1e-10;
constexpr double EPS_D = 1e-7;
constexpr double EPS_COSTS = 1e-8;
constexpr double EPS_A = 1e-5;
constexpr double EPS_Z = 1e-10;
constexpr double EPS_R = 1e-9;
constexpr double EPS_F = 1e-10;
constexpr double PERTURB_RATIO = 0.25;
constexpr double PSI = 1e-5;
constexpr int MAX_CYCLE = 5;
constexpr int NEED_RESTART = 40;
constexpr double BIG = 1e+25;        // ratio-test start value and phase-1 blow-up bound
constexpr double DSE_FLOOR = 1e-8;   // |delta| floor of the pricing candidates

// Bound types in the order of the reference's BoundaryType (types.hpp:48-56).
enum : int8_t { BT_FIXED = 0, BT_FREE = 1, BT_BOXED = 2, BT_UPPER = 3, BT_LOWER = 4 };

// Device status word.  Only a kernel's finalising workgroup (or the host)
// changes it, so every workgroup of a launch sees the same value.
enum : int {
  ST_RUNNING = 0,
  ST_OPTIMAL = 1,          // phase 2: primal feasible
  ST_NO_PIVOT = 2,         // phase 2: no leaving row
  ST_NO_RATIO = 3,         // empty ratio-test candidate set
  ST_NEED_PERTURB = 4,     // cycling detected, costs not yet perturbed
  ST_DUAL_RESTART = 5,     // phase 2: periodic dual-feasibility check failed
  ST_PH1_SOLVED = 6,
  ST_PH1_DUALINF = 7,
  ST_PH1_RESTART = 8,      // phase 1: no candidate row or Z > 0
  ST_PH1_REINIT = 9,       // phase 1: cycle > NEED_RESTART or Z blew up
  ST_BFRT = 10,            // (unused: bound flips run on the device)
  ST_BTRAN_CAP = 11,       // sparse BTRAN list overflowed its LDS capacity
  ST_ERROR = 12,
  ST_MAJOR_END = 13        // multiple candidates: no candidate row left in this major iteration
};

// Scalars of one iteration, resident in device memory.
struct DevScalars {
  int status;
  int stop_iter;     // iteration number at which `status` was raised
  int refac_mid;     // stability check asked for a refactor before the basis update
  int p_idx, p, q_idx, q;
  int is_lower;
  int neg;           // pivot row was negated for the ratio test
  int cycle;
  int perturbed;
  int t;             // number of etas since the last refactor
  int dinf;          // dual infeasibility count over the nonbasic set
  int nF;
  int unbound_checked;
  int any_blocked;
  int nblocked;
  int k;             // kernel size of the current factor
  int bfrt;          // the entering candidate is boxed and flips: the bound-flipping loop runs
  int nflip;         // boxed nonbasic variables that flip in this iteration
  int ncand;         // multiple candidates: rows chosen by the last major pricing
  int cand_best;     // the candidate the current minor iteration takes
  int pad0;
  double delta, theta, obj, f_p, corr, alpha_p, alpha_rq, theta_P;
  double tmp0, tmp1;
  double dq, wq;     // primal iterations: the entering column's reduced cost and Devex weight
  double dz;         // objective change of the flips' costs (first part of delta_z)
  int ctr[16];       // last-workgroup counters, reset by the finaliser
};

// Counter slots in DevScalars::ctr.
enum { CTR_PRICE = 0, CTR_CHUZC1, CTR_CHUZC2, CTR_DUPD, CTR_UPD, CTR_DCHK, CTR_DOT, CTR_INF, CTR_SPMV, CTR_FLIP, CTR_BXB };

constexpr int CAND_MAX = 8;  // candidate rows per major iteration

}  // namespace pds
