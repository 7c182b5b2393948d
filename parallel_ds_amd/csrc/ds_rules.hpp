// Per-variable feasibility rules of the reference solver, shared by the
// single-LP kernels and the batched kernel.
#pragma once
#include "ds_common.hpp"

namespace pds {

// checkVarIsInfeas with its case fallthrough (baseDualSimplex.tpp:443-466).
__device__ inline bool var_infeas(int8_t t, double x, double l, double u, double& delta) {
  switch (t) {
    case BT_LOWER:
      if (x < l - EPS_BOUND) { delta = x - l; return true; }
      [[fallthrough]];
    case BT_UPPER:
      if (x > u + EPS_BOUND) { delta = x - u; return true; }
      [[fallthrough]];
    case BT_FIXED:
      if (fabs(x - l) < EPS_BOUND) { delta = x - l; return true; }
      [[fallthrough]];
    case BT_BOXED:
      if (x > u + EPS_BOUND) { delta = x - u; return true; }
      else if (x < l - EPS_BOUND) { delta = x - l; return true; }
      break;
    default:
      break;
  }
  delta = 0.0;
  return false;
}

// Negation of checkPrimalFeasible for one basic variable (tpp:296-324).
__device__ inline bool primal_infeas(int8_t t, double x, double l, double u) {
  switch (t) {
    case BT_FIXED:
      if (fabs(x - l) > EPS_BOUND) return true;
      break;
    case BT_FREE:
      break;
    case BT_BOXED:
      if (x > u + EPS_R * fabs(u) + EPS_BOUND || x < l - EPS_R * fabs(l) - EPS_BOUND) return true;
      [[fallthrough]];
    case BT_UPPER:
      if (x > u + EPS_R * fabs(u) + EPS_BOUND) return true;
      [[fallthrough]];
    case BT_LOWER:
      if (x < l - EPS_R * fabs(l) - EPS_BOUND) return true;
      break;
    default:
      break;
  }
  return false;
}

// counterDualInfeasible's per-variable test (tpp:415-440).
__device__ inline int dual_infeas(int8_t t, double d) {
  switch (t) {
    case BT_FREE: return (fabs(d) < EPS_D || isnan(d)) ? 1 : 0;
    case BT_UPPER: return (d > EPS_D || isnan(d)) ? 1 : 0;
    case BT_LOWER: return (d < -EPS_D || isnan(d)) ? 1 : 0;
    default: return 0;
  }
}

}  // namespace pds
