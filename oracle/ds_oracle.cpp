/*
 * CPU oracle: restatement of the reference's sequential dual simplex.
 * TEST INFRASTRUCTURE ONLY — loaded by tests/, the smoke check and the
 * benchmark's CPU-baseline leg, never by the product library.
 *
 * Built with -ffp-contract=off so that every a*b+c is two roundings, as in the
 * reference's scalar loops (matrix.cpp:276-279, linalg.cpp:30).
 */
#include "ds_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

namespace {

/* Constants of baseDualSimplex.hpp:23-46. */
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
constexpr double BIG = 1e+25;          /* INF of baseDualSimplex.hpp:46 */

constexpr double DSE_FLOOR = 1e-8;     /* baseDualSimplex.tpp:484 */

struct Abort {};

/* SplitMix64 -> U[0,1) with 53 mantissa bits. */
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next_u64() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

/* Problem in computational form: A <- [A | I], slack bounds from the row
 * ranges, RHS = 0 (problem.tpp:176-226). */
struct Model {
  int m = 0, n = 0, N = 0;
  std::vector<int> cptr, ridx;
  std::vector<double> cval;
  std::vector<int> rptr, cidx;
  std::vector<double> rval;
  std::vector<double> c, l, u;
  std::vector<int> bt;
  std::vector<double> rhs;
};


void build_model(Model& M, int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                 const double* costs, const double* lb, const double* ub, const int* bt,
                 const double* lr, const double* ur, const int* rt) {
  M.m = m;
  M.n = n;
  M.N = n + m;
  int nnz = col_ptr[n];
  M.cptr.assign(M.N + 1, 0);
  M.ridx.resize(nnz + m);
  M.cval.resize(nnz + m);
  for (int j = 0; j <= n; ++j) M.cptr[j] = col_ptr[j];
  for (int k = 0; k < nnz; ++k) {
    M.ridx[k] = row_idx[k];
    M.cval[k] = val[k];
  }
  for (int i = 0; i < m; ++i) {
    M.ridx[nnz + i] = i;
    M.cval[nnz + i] = 1.0;
    M.cptr[n + i + 1] = nnz + i + 1;
  }
  /* CSR of the full matrix, entries of each row in ascending column order
   * (matrix.cpp:349-372 appends the slack after the row's structurals). */
  M.rptr.assign(m + 1, 0);
  for (int k = 0; k < nnz + m; ++k) M.rptr[M.ridx[k] + 1]++;
  for (int i = 0; i < m; ++i) M.rptr[i + 1] += M.rptr[i];
  M.cidx.resize(nnz + m);
  M.rval.resize(nnz + m);
  std::vector<int> fill(M.rptr.begin(), M.rptr.end() - 1);
  for (int j = 0; j < M.N; ++j)
    for (int k = M.cptr[j]; k < M.cptr[j + 1]; ++k) {
      int r = M.ridx[k];
      M.cidx[fill[r]] = j;
      M.rval[fill[r]] = M.cval[k];
      fill[r]++;
    }
  M.c.resize(M.N);
  M.l.resize(M.N);
  M.u.resize(M.N);
  M.bt.resize(M.N);
  for (int j = 0; j < n; ++j) {
    M.c[j] = costs[j];
    M.l[j] = lb[j];
    M.u[j] = ub[j];
    M.bt[j] = bt[j];
  }
  for (int i = 0; i < m; ++i) {
    M.c[n + i] = 0.0;
    M.u[n + i] = -lr[i];
    M.l[n + i] = -ur[i];
    switch (rt[i]) {
      case ORA_FIXED: M.bt[n + i] = ORA_FIXED; break;
      case ORA_BOXED: M.bt[n + i] = ORA_BOXED; break;
      case ORA_LOWER: M.bt[n + i] = ORA_UPPER; break;
      case ORA_UPPER: M.bt[n + i] = ORA_LOWER; break;
      default: M.bt[n + i] = ORA_FREE; break;
    }
  }
  M.rhs.assign(m, 0.0);
}

/* Project LU replacing PARDISO.  Basic slack columns are unit columns: they
 * are eliminated symbolically, leaving the k x k kernel K = A[R, basic
 * structurals] (R = rows whose slack is nonbasic), which gets a dense
 * partial-pivot LU.  Contract of Matrix::solve (matrix.cpp:567-582):
 * solve(b) gives B x = b, solveT(b) gives B^T y = b. */
struct Factor {
  int m = 0, k = 0;
  std::vector<int> basis0, pos_P, row_R, Ridx_of_row, slackpos_of_row;
  std::vector<double> LU;
  std::vector<int> piv;
  const Model* M = nullptr;

  bool build(const Model& Mod, const std::vector<int>& basis) {
    M = &Mod;
    m = Mod.m;
    basis0 = basis;
    pos_P.clear();
    row_R.clear();
    Ridx_of_row.assign(m, -1);
    slackpos_of_row.assign(m, -1);
    for (int i = 0; i < m; ++i) {
      int j = basis[i];
      if (j >= Mod.n) slackpos_of_row[j - Mod.n] = i;
      else pos_P.push_back(i);
    }
    for (int r = 0; r < m; ++r)
      if (slackpos_of_row[r] < 0) {
        Ridx_of_row[r] = (int)row_R.size();
        row_R.push_back(r);
      }
    k = (int)pos_P.size();
    if ((int)row_R.size() != k) return false;
    LU.assign((size_t)k * k, 0.0);
    for (int c = 0; c < k; ++c) {
      int j = basis[pos_P[c]];
      for (int e = Mod.cptr[j]; e < Mod.cptr[j + 1]; ++e) {
        int ri = Ridx_of_row[Mod.ridx[e]];
        if (ri >= 0) LU[(size_t)ri * k + c] = Mod.cval[e];
      }
    }
    piv.resize(k);
    for (int s = 0; s < k; ++s) {
      int pr = s;
      double best = std::fabs(LU[(size_t)s * k + s]);
      for (int r = s + 1; r < k; ++r) {
        double v = std::fabs(LU[(size_t)r * k + s]);
        if (v > best) { best = v; pr = r; }
      }
      piv[s] = pr;
      if (best == 0.0) return false;
      if (pr != s)
        for (int c = 0; c < k; ++c) std::swap(LU[(size_t)s * k + c], LU[(size_t)pr * k + c]);
      double inv = LU[(size_t)s * k + s];
      for (int r = s + 1; r < k; ++r) {
        double f = LU[(size_t)r * k + s] / inv;
        LU[(size_t)r * k + s] = f;
        if (f != 0.0) {
          double* rr = &LU[(size_t)r * k];
          const double* sr = &LU[(size_t)s * k];
          for (int c = s + 1; c < k; ++c) rr[c] = rr[c] - f * sr[c];
        }
      }
    }
    return true;
  }

  /* B x = b; x indexed by basis position. */
  void solve(const std::vector<double>& b, std::vector<double>& x) const {
    std::vector<double> z(k);
    for (int i = 0; i < k; ++i) z[i] = b[row_R[i]];
    for (int s = 0; s < k; ++s)
      if (piv[s] != s) std::swap(z[s], z[piv[s]]);
    for (int r = 0; r < k; ++r) {
      double acc = z[r];
      const double* lr = &LU[(size_t)r * k];
      for (int c = 0; c < r; ++c) acc = acc - lr[c] * z[c];
      z[r] = acc;
    }
    for (int r = k - 1; r >= 0; --r) {
      double acc = z[r];
      const double* ur = &LU[(size_t)r * k];
      for (int c = r + 1; c < k; ++c) acc = acc - ur[c] * z[c];
      z[r] = acc / ur[r];
    }
    x.assign(m, 0.0);
    std::vector<double> acc(m, 0.0);
    for (int c = 0; c < k; ++c) {
      x[pos_P[c]] = z[c];
      int j = basis0[pos_P[c]];
      for (int e = M->cptr[j]; e < M->cptr[j + 1]; ++e) {
        int r = M->ridx[e];
        if (slackpos_of_row[r] >= 0) acc[r] = acc[r] + M->cval[e] * z[c];
      }
    }
    for (int r = 0; r < m; ++r)
      if (slackpos_of_row[r] >= 0) x[slackpos_of_row[r]] = b[r] - acc[r];
  }

  /* B^T y = e; e indexed by basis position, y by row. */
  void solveT(const std::vector<double>& e, std::vector<double>& y) const {
    y.assign(m, 0.0);
    for (int r = 0; r < m; ++r)
      if (slackpos_of_row[r] >= 0) y[r] = e[slackpos_of_row[r]];
    std::vector<double> w(k);
    for (int c = 0; c < k; ++c) {
      double acc = e[pos_P[c]];
      int j = basis0[pos_P[c]];
      for (int q = M->cptr[j]; q < M->cptr[j + 1]; ++q) {
        int r = M->ridx[q];
        if (slackpos_of_row[r] >= 0) acc = acc - M->cval[q] * y[r];
      }
      w[c] = acc;
    }
    /* K^T z = w with P K = L U: U^T v = w, L^T t = v, z = P^T t. */
    for (int c = 0; c < k; ++c) {
      double acc = w[c];
      for (int r = 0; r < c; ++r) acc = acc - LU[(size_t)r * k + c] * w[r];
      w[c] = acc / LU[(size_t)c * k + c];
    }
    for (int c = k - 1; c >= 0; --c) {
      double acc = w[c];
      for (int r = c + 1; r < k; ++r) acc = acc - LU[(size_t)r * k + c] * w[r];
      w[c] = acc;
    }
    for (int s = k - 1; s >= 0; --s)
      if (piv[s] != s) std::swap(w[s], w[piv[s]]);
    for (int i = 0; i < k; ++i) y[row_R[i]] = w[i];
  }
};

struct Eta {
  std::vector<double> v;
  int p;
};

/* checkVarIsInfeas with its fallthrough (tpp:443-466). */
bool var_infeas_rule(int t, double xv, double lo, double up, double& delta) {
  switch (t) {
    case ORA_LOWER:
      if (xv < lo - EPS_BOUND) { delta = xv - lo; return true; }
      /* fallthrough */
    case ORA_UPPER:
      if (xv > up + EPS_BOUND) { delta = xv - up; return true; }
      /* fallthrough */
    case ORA_FIXED:
      if (std::fabs(xv - lo) < EPS_BOUND) { delta = xv - lo; return true; }
      /* fallthrough */
    case ORA_BOXED:
      if (xv > up + EPS_BOUND) { delta = xv - up; return true; }
      else if (xv < lo - EPS_BOUND) { delta = xv - lo; return true; }
  }
  delta = 0.0;
  return false;
}


/* Harris two-pass selection over candidate positions F (ascending): pass 1
 * bounds the step by the EPS_D-relaxed ratios, pass 2 takes the largest
 * |alpha| among the candidates whose plain ratio is within the bound; the
 * strict comparison keeps the lowest position on ties. */
template <typename DOf>
int harris_pick(int nF, const int* F, DOf dof, const double* alpha) {
  double tmax = BIG;
  for (int s = 0; s < nF; ++s) {
    const int i = F[s];
    const double dj = dof(i);
    const double t = alpha[i] > 0 ? (dj + EPS_D) / alpha[i] : (dj - EPS_D) / alpha[i];
    tmax = std::min(tmax, t);
  }
  int q_idx = -1;
  double amax = -1.0;
  for (int s = 0; s < nF; ++s) {
    const int i = F[s];
    const double t = dof(i) / alpha[i];
    if (t <= tmax && std::fabs(alpha[i]) > amax) {
      amax = std::fabs(alpha[i]);
      q_idx = i;
    }
  }
  return q_idx;
}

/* linalg::PFIsolve (linalg.cpp:12-46). */
void pfi_apply(const std::vector<Eta>& etas, std::vector<double>& sol, bool transpose) {
  int m = (int)sol.size();
  if (!transpose) {
    for (const Eta& E : etas) {
      double s = sol[E.p];
      for (int i = 0; i < m; ++i) sol[i] = sol[i] + s * E.v[i];
      sol[E.p] = s * E.v[E.p];
    }
  } else {
    for (auto it = etas.rbegin(); it != etas.rend(); ++it) {
      double acc = 0.0;
      for (int i = 0; i < m; ++i) acc = acc + it->v[i] * sol[i];
      sol[it->p] = acc;
    }
  }
}

struct Solver {
  const Model& M;
  ora_params P;
  int m, n, N;
  std::vector<double> c, original_costs, x, d, beta;
  std::vector<int> basis, nonbasis;
  Factor F;
  std::vector<Eta> etas;
  bool perturbed = false;
  bool warm = false;
  int phase1_iterations = 0;
  int iteration = 0;
  double obj = 0.0;
  int status = ORA_OPTIMAL;
  bool solved = false;
  int refactors = 0, phase1_restarts = 0;
  int bfrt_exhausted = 0; /* BFRT loops that flipped every candidate */
  SplitMix rng;

  Solver(const Model& mod, const ora_params& prm)
      : M(mod), P(prm), m(mod.m), n(mod.n), N(mod.N), rng(prm.perturb_seed) {
    c = M.c;
    x.assign(N, 0.0);
    d.assign(N, 0.0);
    beta.assign(m, 0.0);
    nonbasis.resize(N - m);
    basis.resize(m);
    /* Slack basis (baseDualSimplex.tpp:184-204). */
    for (int i = 0; i < N - m; ++i) nonbasis[i] = i;
    for (int i = 0; i < m; ++i) basis[i] = i + (N - m);
    if (P.refactor_freq <= 0) P.refactor_freq = 150;
  }

  /* Starting basis other than the slack basis: the m given variables, the
   * nonbasic set in ascending order (no counterpart in the reference, whose
   * solver always starts from the slack basis). */
  void set_start(const int* b) {
    std::vector<char> in(N, 0);
    for (int i = 0; i < m; ++i) {
      if (b[i] < 0 || b[i] >= N || in[b[i]]) throw std::runtime_error("invalid starting basis");
      in[b[i]] = 1;
      basis[i] = b[i];
    }
    int k = 0;
    for (int j = 0; j < N; ++j)
      if (!in[j]) nonbasis[k++] = j;
    warm = true;
  }

  void tick() {
    iteration += 1;
    if (P.max_iterations > 0 && iteration > P.max_iterations) {
      status = ORA_ITERATION_LIMIT;
      throw Abort{};
    }
  }

  void reFactorize() {
    etas.clear();
    refactors++;
    if (!F.build(M, basis)) throw std::runtime_error("singular basis");
  }

  void solveLinSys(const std::vector<double>& rhs, std::vector<double>& sol, bool transpose) {
    if (transpose) {
      std::vector<double> t = rhs;
      pfi_apply(etas, t, true);
      F.solveT(t, sol);
    } else {
      F.solve(rhs, sol);
      pfi_apply(etas, sol, false);
    }
  }

  /* Matrix::dotUpdate UPDATE_T over the nonbasic columns (matrix.cpp:273-287). */
  void dot_update_T(const std::vector<double>& v1, const std::vector<double>& v2,
                    std::vector<double>& sol, double a, double b, bool set) {
    for (size_t i = 0; i < nonbasis.size(); ++i) {
      int j = nonbasis[i];
      double acc = 0.0;
      for (int e = M.cptr[j]; e < M.cptr[j + 1]; ++e) acc = acc + M.cval[e] * v1[M.ridx[e]] * a;
      sol[set ? j : (int)i] = acc + v2[j] * b;
    }
  }

  void initBetaWeights(bool identity) {
    if (identity) {
      for (int i = 0; i < m; ++i) beta[i] = 1.0;
      return;
    }
    std::vector<double> e(m), y(m);
    for (int i = 0; i < m; ++i) {
      std::fill(e.begin(), e.end(), 0.0);
      e[i] = 1.0;
      solveLinSys(e, y, true);
      double s = 0.0;
      for (double v : y) s = s + v * v;   /* ValuesVector::norm is the squared norm */
      beta[i] = s;
    }
  }

  void setPrimalVars() {
    for (int j : nonbasis) {
      switch (M.bt[j]) {
        case ORA_FIXED: x[j] = M.l[j]; break;
        case ORA_FREE: x[j] = 0.0; break;
        case ORA_BOXED: x[j] = (d[j] > 0) ? M.l[j] : M.u[j]; break;
        case ORA_UPPER: x[j] = M.u[j]; break;
        case ORA_LOWER: x[j] = M.l[j]; break;
      }
    }
  }

  bool checkPerturbNeed() const {
    std::vector<double> cs(c);
    std::sort(cs.begin(), cs.end());
    int uniq = 0;
    double last = 0.0;
    for (size_t i = 0; i < cs.size(); ++i) {
      if (i == 0 || !(std::fabs(last - cs[i]) < EPS_COSTS)) {
        last = cs[i];
        uniq++;
      }
    }
    return uniq < PERTURB_RATIO * (double)cs.size();
  }

  static double getWeight(int cnt) {
    static const double w[10] = {0.01, 0.1, 1, 2, 5, 10, 20, 30, 40, 100};
    int k = cnt > 10 ? 9 : (cnt == 0 ? 0 : cnt - 1);
    return w[k];
  }

  /* baseDualSimplex.tpp:238-275 with a seeded stream. */
  void perturbCosts() {
    original_costs = c;
    double mean = 0.0;
    for (double v : c) mean = mean + v;
    mean = mean / (double)c.size();
    double min_p = std::min(1e-2 * EPS_D, PSI);
    double max_p = std::max(1e+3 * EPS_D, PSI * 10 * mean);
    for (int i = 0; i < N; ++i) {
      double magnitude = 100 * EPS_D + PSI * c[i];
      double pert;
      if (M.bt[i] != ORA_LOWER && M.bt[i] != ORA_FREE) pert = -0.5 * magnitude * (1 + rng.uniform());
      pert = 0.5 * magnitude * (1 + rng.uniform());
      pert = getWeight(M.cptr[i + 1] - M.cptr[i]) * pert;
      if (pert != 0.0)
        while (std::fabs(pert) > std::fabs(max_p) || std::fabs(pert) < std::fabs(min_p)) {
          if (std::fabs(pert) > std::fabs(max_p)) pert = 0.1 * pert;
          else if (std::fabs(pert) < std::fabs(min_p)) pert = 10 * pert;
        }
      c[i] += pert;
    }
    perturbed = true;
  }

  /* checkPrimalFeasible with its switch fallthrough (tpp:296-324). */
  bool checkPrimalFeasible() const {
    for (int j : basis) {
      double xv = x[j], lo = M.l[j], up = M.u[j];
      switch (M.bt[j]) {
        case ORA_FIXED:
          if (std::fabs(xv - lo) > EPS_BOUND) return false;
          /* fallthrough */
        case ORA_FREE:
          break;
        case ORA_BOXED:
          if (xv > up + EPS_R * std::fabs(up) + EPS_BOUND || xv < lo - EPS_R * std::fabs(lo) - EPS_BOUND)
            return false;
          /* fallthrough */
        case ORA_UPPER:
          if (xv > up + EPS_R * std::fabs(up) + EPS_BOUND) return false;
          /* fallthrough */
        case ORA_LOWER:
          if (xv < lo - EPS_R * std::fabs(lo) - EPS_BOUND) return false;
      }
    }
    return true;
  }

  bool dualInfeasAt(int j) const {
    switch (M.bt[j]) {
      case ORA_FREE: return std::fabs(d[j]) < EPS_D || std::isnan(d[j]);
      case ORA_UPPER: return d[j] > EPS_D || std::isnan(d[j]);
      case ORA_LOWER: return d[j] < -EPS_D || std::isnan(d[j]);
      default: return false;
    }
  }
  bool checkDualFeasible() const {
    for (int j : nonbasis)
      if (dualInfeasAt(j)) return false;
    return true;
  }
  int counterDualInfeasible() const {
    int k = 0;
    for (int j : nonbasis) k += dualInfeasAt(j) ? 1 : 0;
    return k;
  }

  /* checkVarIsInfeas with its fallthrough (tpp:443-466). */
  bool varInfeas(int j, double& delta) const { return var_infeas_rule(M.bt[j], x[j], M.l[j], M.u[j], delta); }

  /* DSEPricing (tpp:470-501). */
  void dsePricing(const std::vector<char>& blocked, int& p, int& p_idx, double& delta, bool& is_lower) {
    p = -1;
    p_idx = -1;
    delta = 0.0;
    is_lower = false;
    double best = 0.0;
    for (int i = 0; i < m; ++i) {
      int j = basis[i];
      double dt;
      if (varInfeas(j, dt) && std::fabs(dt) > DSE_FLOOR) {
        double w = (dt * dt) / beta[i];
        if (w > best && !blocked[j]) {
          p = j;
          p_idx = i;
          delta = dt;
          best = w;
          is_lower = dt < 0;
        }
      }
    }
  }

  /* phase1Pricing (tpp:746-772). */
  bool phase1Pricing(const std::vector<double>& f, const std::vector<char>& blocked, int& p, int& p_idx) {
    double best = 0.0;
    bool found = false;
    p = -1;
    p_idx = -1;
    for (int i = 0; i < m; ++i) {
      int j = basis[i];
      int t = M.bt[j];
      if ((t == ORA_LOWER && f[i] > EPS_F) || (t == ORA_UPPER && f[i] < -EPS_F) || t == ORA_BOXED ||
          t == ORA_FIXED) {
        double w = (f[i] * f[i]) / beta[i];
        if (w > best && !blocked[j]) {
          p = j;
          p_idx = i;
          best = w;
          found = true;
        }
      }
    }
    return found;
  }

  /* simpleRatioTest (tpp:546-572) and its order-free variant. */
  int ratioTest(const std::vector<int>& Fs, const std::vector<double>& alpha) const {
    int q_idx = -1;
    double theta = BIG;
    if (P.tie_rule == 0) {
      for (int i : Fs) {
        double t = d[nonbasis[i]] / alpha[i];
        if (t < theta) {
          theta = t;
          q_idx = i;
        } else if (q_idx >= 0 && std::fabs(t - theta) < EPS_Z) {
          if (std::fabs(alpha[i]) > std::fabs(alpha[q_idx])) {
            theta = t;
            q_idx = i;
          }
        }
      }
      return q_idx;
    }
    for (int i : Fs) theta = std::min(theta, d[nonbasis[i]] / alpha[i]);
    double best = -1.0;
    for (int i : Fs) {
      double t = d[nonbasis[i]] / alpha[i];
      if (std::fabs(t - theta) < EPS_Z && std::fabs(alpha[i]) > best) {
        best = std::fabs(alpha[i]);
        q_idx = i;
      }
    }
    return q_idx;
  }

  /* Harris two-pass (the method behind tpp:576-611): pass 1 bounds the step
   * with the EPS_D-relaxed ratios, pass 2 takes the largest |alpha| under it. */
  int harrisTest(const std::vector<int>& Fs, const std::vector<double>& alpha) const {
    return harris_pick((int)Fs.size(), Fs.data(), [&](int i) { return d[nonbasis[i]]; }, alpha.data());
  }

  int chooseColumn(const std::vector<int>& Fs, const std::vector<double>& alpha) const {
    return P.ratio_test == 1 ? harrisTest(Fs, alpha) : ratioTest(Fs, alpha);
  }

  void initReducedCosts() {
    std::vector<double> cb(m), y(m);
    for (int i = 0; i < m; ++i) cb[i] = c[basis[i]];
    solveLinSys(cb, y, true);
    dot_update_T(y, c, d, -1.0, 1.0, true);
  }

  void dualSimplexInit() {
    setPrimalVars();
    /* rhs = RHS - A_N x_N by Matrix::dotUpdate UPDATE, set = true (matrix.cpp:330-341). */
    std::vector<double> buff(N, 0.0), rhs(m), xb(m);
    for (int j : nonbasis) buff[j] = x[j];
    for (int i = 0; i < m; ++i) {
      double s = M.rhs[i] * 1.0;
      for (int e = M.rptr[i]; e < M.rptr[i + 1]; ++e) s = s + M.rval[e] * buff[M.cidx[e]] * -1.0;
      rhs[i] = s;
    }
    solveLinSys(rhs, xb, false);
    for (int i = 0; i < m; ++i) x[basis[i]] = xb[i];
    initReducedCosts();
    double s = 0.0;
    for (int j = 0; j < N; ++j) s = s + c[j] * x[j];
    obj = s;
  }

  /* minimizeDualInfeasibilityInit + initPhase1PricingVector
   * (tpp:776-795, sequentialDualSimplex.cpp:135-152). */
  void phase1Init(std::vector<double>& f) {
    initReducedCosts();
    std::vector<int> inf_u, inf_l;
    for (int j : nonbasis) {
      int t = M.bt[j];
      if ((t == ORA_UPPER || t == ORA_FREE) && d[j] > EPS_D) inf_u.push_back(j);
      else if ((t == ORA_LOWER || t == ORA_FREE) && d[j] < -EPS_D) inf_l.push_back(j);
    }
    std::vector<double> cc(m, 0.0);
    double Z = 0.0;
    for (int j : inf_l) {
      Z += d[j];
      for (int e = M.cptr[j]; e < M.cptr[j + 1]; ++e) cc[M.ridx[e]] += M.cval[e];
    }
    for (int j : inf_u) {
      Z -= d[j];
      for (int e = M.cptr[j]; e < M.cptr[j + 1]; ++e) cc[M.ridx[e]] -= M.cval[e];
    }
    obj = Z;
    solveLinSys(cc, f, false);
  }

  void column(int q, std::vector<double>& a) const {
    a.assign(m, 0.0);
    for (int e = M.cptr[q]; e < M.cptr[q + 1]; ++e) a[M.ridx[e]] = M.cval[e];
  }

  /* Basis change shared by both phases (sequentialDualSimplex.cpp:174-227). */
  void changeBasis(std::vector<double>* f, const std::vector<double>& rho,
                   const std::vector<double>& aq, int p_idx, int p, int q_idx, int q, double theta_P) {
    std::vector<double> tau(m);
    solveLinSys(rho, tau, false);
    Eta E;
    E.v.resize(m);
    E.p = p_idx;
    double ap = aq[p_idx];
    double bp = beta[p_idx];
    for (int i = 0; i < m; ++i) {
      if (f) {
        if (i != p_idx) (*f)[i] = (*f)[i] - aq[i] * theta_P;
      } else {
        int j = basis[i];
        x[j] = x[j] - theta_P * aq[i];
      }
      E.v[i] = (i != p_idx) ? -aq[i] / ap : 1 / ap;
      if (i != p_idx) {
        double r = aq[i] / ap;
        beta[i] = beta[i] - 2 * aq[i] / ap * tau[i] + (r * r) * bp;
      }
    }
    beta[p_idx] = beta[p_idx] / (ap * ap);
    if (!f) x[q] = x[q] + theta_P;
    basis[p_idx] = q;
    nonbasis[q_idx] = p;
    etas.push_back(std::move(E));
  }

  /* minimizeDualInfeasibility (tpp:799-934). */
  int phase1() {
    std::vector<double> rho(m), f(m), alpha(N - m), aq(m);
    std::vector<char> blocked(N, 0);
    bool any_blocked = false;
    phase1Init(f);
    int cycle = 0;
    int st;
    while (true) {
      tick();
      if (!perturbed && P.perturb && cycle > MAX_CYCLE) {
        perturbCosts();
        phase1Init(f);
      }
      if (cycle > NEED_RESTART || obj > BIG || std::isnan(obj)) phase1Init(f);
      if (iteration % P.refactor_freq == 0) reFactorize();

      int p, p_idx;
      bool found = phase1Pricing(f, blocked, p, p_idx);
      if (!found || !counterDualInfeasible()) {
        if (std::fabs(obj) < EPS_A || !counterDualInfeasible()) { st = 0; break; }
        else if (obj > 0) { st = 1; break; }
      }
      if (p < 0 || p_idx < 0 || obj > EPS_BOUND) {
        phase1_restarts++;
        reFactorize();
        phase1Init(f);
        initBetaWeights(false);
        if (any_blocked) { std::fill(blocked.begin(), blocked.end(), 0); any_blocked = false; }
        continue;
      }
      std::vector<double> e(m, 0.0);
      e[p_idx] = 1.0;
      solveLinSys(e, rho, true);
      dot_update_T(rho, c, alpha, 1.0, 0.0, false);
      bool neg = f[p_idx] > 0;
      if (neg) for (double& a : alpha) a = -a;
      std::vector<int> Fs;
      for (int i = 0; i < N - m; ++i) {
        int j = nonbasis[i];
        if ((d[j] >= 0 && alpha[i] > EPS_A) || (d[j] <= 0 && alpha[i] < -EPS_A)) Fs.push_back(i);
      }
      if (Fs.empty()) {
        blocked[p] = 1;
        any_blocked = true;
        phase1_restarts++;
        reFactorize();
        phase1Init(f);
        continue;
      }
      if (any_blocked) { std::fill(blocked.begin(), blocked.end(), 0); any_blocked = false; }
      int q_idx = ratioTest(Fs, alpha);
      int q = nonbasis[q_idx];
      if (neg) for (double& a : alpha) a = -a;
      double theta = d[q] / alpha[q_idx];
      std::vector<double> a;
      column(q, a);
      solveLinSys(a, aq, false);
      if (std::fabs(aq[p_idx] - alpha[q_idx]) > EPS_R * (1 + std::fabs(aq[p_idx]))) reFactorize();
      double corr = 0.0;
      int tq = M.bt[q];
      if ((tq == ORA_UPPER || tq == ORA_FREE) && d[q] > EPS_D) corr = 1;
      else if ((tq == ORA_LOWER || tq == ORA_FREE) && d[q] < -EPS_D) corr = -1;
      obj = obj - theta * f[p_idx];
      for (int i = 0; i < N - m; ++i) {
        int j = nonbasis[i];
        d[j] -= theta * alpha[i];
      }
      d[p] = -theta;
      d[q] = 0;
      double theta_P = f[p_idx] / aq[p_idx];
      changeBasis(&f, rho, aq, p_idx, p, q_idx, q, theta_P);
      f[p_idx] = f[p_idx] / aq[p_idx] + corr;
      cycle = (std::fabs(theta) < EPS_A) ? cycle + 1 : 0;
    }
    return st;
  }

  /* callDualSolver (sequentialDualSimplex.cpp:253-515). */
  bool phase2() {
    std::vector<double> rho(m), alpha(N - m), aq(m);
    std::vector<char> blocked(N, 0);
    dualSimplexInit();
    int cycle = 0;
    bool unbound_checked = false;
    while (true) {
      tick();
      if (iteration % P.refactor_freq == 0) reFactorize();
      if (iteration % NEED_RESTART == 0 && !checkDualFeasible()) {
        reFactorize();
        phase1();
        dualSimplexInit();
      }
      if (!perturbed && P.perturb && cycle > MAX_CYCLE) {
        perturbCosts();
        reFactorize();
        dualSimplexInit();
      }
      if (checkPrimalFeasible()) {
        solved = true;
        status = ORA_OPTIMAL;
        break;
      }
      int p, p_idx;
      double delta;
      bool is_lower;
      dsePricing(blocked, p, p_idx, delta, is_lower);
      if (p < 0 || p_idx < 0) {
        reFactorize();
        initBetaWeights(false);
        dualSimplexInit();
        continue;
      }
      std::vector<double> e(m, 0.0);
      e[p_idx] = 1.0;
      solveLinSys(e, rho, true);
      dot_update_T(rho, c, alpha, 1.0, 0.0, false);
      delta = std::fabs(delta);
      int sgn = is_lower ? -1 : 1;
      if (is_lower) for (double& a : alpha) a = -a;
      std::vector<int> Fs;
      for (int i = 0; i < N - m; ++i) {
        double a = alpha[i];
        int j = nonbasis[i];
        bool cond = false;
        if (a > EPS_A) {
          cond = std::fabs(x[j] - M.l[j]) < EPS_Z && (M.bt[j] == ORA_LOWER || M.bt[j] == ORA_BOXED);
        } else if (a < -EPS_A) {
          cond = std::fabs(x[j] - M.u[j]) < EPS_Z && (M.bt[j] == ORA_UPPER || M.bt[j] == ORA_BOXED);
        } else {
          cond = M.bt[j] == ORA_FREE;
        }
        if (cond) Fs.push_back(i);
      }
      /* A Free candidate enters F with |alpha| <= EPS_A (seq.cpp:361-365); when alpha is 0 its
       * ratio is infinite and simpleRatioTest returns -1, which the reference then uses as an
       * index (undefined behaviour).  Here a candidate set without a finite ratio is treated
       * as an empty one, as the GPU ratio test does. */
      if (!Fs.empty() && chooseColumn(Fs, alpha) < 0) Fs.clear();
      if (Fs.empty()) {
        if (unbound_checked) {
          solved = false;
          status = ORA_PRIMAL_INFEASIBLE;
          break;
        }
        unbound_checked = true;
        reFactorize();
        initBetaWeights(false);
        phase1();
        dualSimplexInit();
        continue;
      }
      unbound_checked = false;
      int q_idx = -1, q = -1;
      /* When every candidate flips, the loop ends with F empty and q still the last flipped
       * candidate, which then enters the basis without its flip: the row's remaining
       * infeasibility is left on the leaving variable (sequentialDualSimplex.cpp:398-422).
       * Restated as is; counted in bfrt_exhausted. */
      while (!Fs.empty()) {
        const int qi = chooseColumn(Fs, alpha);
        if (qi < 0) break; /* only infinite ratios left: keep the last flip candidate */
        q_idx = qi;
        q = nonbasis[q_idx];
        if (M.bt[q] == ORA_BOXED) {
          double span = M.u[q] - M.l[q];
          if (delta - span * std::fabs(alpha[q_idx]) <= 0) break;
          delta -= span * std::fabs(alpha[q_idx]);
          Fs.erase(std::remove(Fs.begin(), Fs.end(), q_idx), Fs.end());
          if (Fs.empty()) bfrt_exhausted++;
        } else {
          break;
        }
      }
      if (is_lower) for (double& a : alpha) a = -a;
      double theta = d[q] / alpha[q_idx];
      delta = sgn * delta;
      std::vector<double> a;
      column(q, a);
      solveLinSys(a, aq, false);
      if (std::fabs(aq[p_idx] - alpha[q_idx]) > EPS_R * (1 + std::fabs(aq[p_idx]))) reFactorize();

      std::vector<int> low_inf, up_inf;
      std::vector<double> cc(m, 0.0);
      double delta_z = 0.0;
      for (int i = 0; i < N - m; ++i) {
        int j = nonbasis[i];
        d[j] = d[j] - theta * alpha[i];
        if (M.bt[j] == ORA_BOXED) {
          if (std::fabs(x[j] - M.l[j]) < EPS_Z && d[j] < -EPS_D) {
            low_inf.push_back(j);
            delta_z += c[j] * (M.u[j] - M.l[j]);
          } else if (std::fabs(x[j] - M.u[j]) < EPS_Z && d[j] > EPS_D) {
            up_inf.push_back(j);
            delta_z -= c[j] * (M.u[j] - M.l[j]);
          }
        }
      }
      for (int j : low_inf)
        for (int e2 = M.cptr[j]; e2 < M.cptr[j + 1]; ++e2) cc[M.ridx[e2]] += M.cval[e2] * (M.u[j] - M.l[j]);
      for (int j : up_inf)
        for (int e2 = M.cptr[j]; e2 < M.cptr[j + 1]; ++e2) cc[M.ridx[e2]] -= M.cval[e2] * (M.u[j] - M.l[j]);
      d[p] = -theta;
      d[q] = 0;
      if (!low_inf.empty() || !up_inf.empty()) {
        std::vector<double> dxb(m);
        solveLinSys(cc, dxb, false);
        for (int i = 0; i < m; ++i) x[basis[i]] = x[basis[i]] - dxb[i];
        for (int i = 0; i < m; ++i) delta_z -= c[basis[i]] * dxb[i];
      }
      obj += delta_z;
      double theta_P = delta / aq[p_idx];
      changeBasis(nullptr, rho, aq, p_idx, p, q_idx, q, theta_P);
      for (int j : low_inf) x[j] = M.u[j];
      for (int j : up_inf) x[j] = M.l[j];
      obj += theta * delta;
      if (std::fabs(theta) < EPS_A) cycle += 1;
      if (P.verbose && iteration % 500 == 0) std::fprintf(stderr, "%d : Z = %.12g\n", iteration, obj);
    }
    return solved;
  }

  /* initDualSimplex + presolve("minInfeas") + solve + postProcess. */
  void run(double& Z) {
    reFactorize();
    if (P.perturb && checkPerturbNeed()) perturbCosts();
    initBetaWeights(!warm);  /* exact weights for a given starting basis */
    int st = phase1();
    phase1_iterations = iteration;
    if (st == 1) {
      status = ORA_DUAL_INFEASIBLE;
      solved = false;
      Z = 0.0;
      return;
    }
    setPrimalVars();
    bool ok = phase2();
    if (perturbed) {
      c = original_costs;
      dualSimplexInit();
      /* dualSimplexInit places boxed nonbasic variables by the sign of the reduced costs it
       * finds on entry (setPrimalVars runs before initReducedCosts, seq.cpp:230-250), here
       * those of the perturbed costs; the cleanup places them again by the restored ones. */
      if (P.cleanup) dualSimplexInit();
      /* postProcess (tpp:123-159) calls callPrimalSolver when the restored
       * costs leave dual infeasibilities; that correction is a stub in the
       * reference (sequentialDualSimplex.cpp:113-116).  cleanup = 1 resumes
       * the dual simplex from the final basis instead: phase 1 + phase 2 when the
       * original costs leave dual infeasibilities, phase 2 when only x_B is out of bounds. */
      if (P.cleanup && ok && !checkDualFeasible()) {
        int st2 = phase1();
        if (st2 == 1) {
          status = ORA_DUAL_INFEASIBLE;
          solved = false;
          Z = 0.0;
          return;
        }
        setPrimalVars();
        ok = phase2();
      } else if (P.cleanup && ok && !checkPrimalFeasible()) {
        /* Dual feasible for the original costs, but dualSimplexInit moved boxed nonbasic
         * variables whose reduced cost changed sign to their other bound: continue phase 2. */
        ok = phase2();
      }
    }
    (void)ok;
    Z = obj;
  }
};

}  // namespace

extern "C" int ora_solve(int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                         const double* costs, const double* lower_bound, const double* upper_bound,
                         const int* bound_type, const double* lower_range, const double* upper_range,
                         const int* range_type, const ora_params* params, ora_result* result,
                         double* x_out, int* basis_out) {
  return ora_solve_from(m, n, col_ptr, row_idx, val, costs, lower_bound, upper_bound, bound_type, lower_range,
                        upper_range, range_type, nullptr, params, result, x_out, basis_out);
}

extern "C" int ora_solve_from(int m, int n, const int* col_ptr, const int* row_idx, const double* val,
                              const double* costs, const double* lower_bound, const double* upper_bound,
                              const int* bound_type, const double* lower_range, const double* upper_range,
                              const int* range_type, const int* basis_in, const ora_params* params,
                              ora_result* result, double* x_out, int* basis_out) {
  try {
    Model M;
    build_model(M, m, n, col_ptr, row_idx, val, costs, lower_bound, upper_bound, bound_type,
                lower_range, upper_range, range_type);
    ora_params P = *params;
    Solver S(M, P);
    if (basis_in) S.set_start(basis_in);
    double Z = 0.0;
    try {
      S.run(Z);
    } catch (const Abort&) {
      Z = S.obj;
    }
    result->status = S.status;
    result->solved = S.solved ? 1 : 0;
    result->iterations = S.iteration;
    result->perturbed = S.perturbed ? 1 : 0;
    result->refactors = S.refactors;
    result->phase1_restarts = S.phase1_restarts;
    result->Z = Z;
    result->phase1_iterations = S.phase1_iterations;
    result->bfrt_exhausted = S.bfrt_exhausted;
    if (x_out)
      for (int j = 0; j < M.N; ++j) x_out[j] = S.x[j];
    if (basis_out)
      for (int i = 0; i < m; ++i) basis_out[i] = S.basis[i];
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ora_solve: %s\n", e.what());
    return -1;
  }
}

extern "C" int ora_dot_update(int m, int ncols, const int* row_ptr, const int* col_idx, const double* vals,
                              const double* vec1, const double* vec2, double* sol, double alpha,
                              double beta, const int* cols, int ncols_sel, int method, int set) {
  /* Matrix::dotUpdate (matrix.cpp:261-346) on an m x ncols CSR matrix. */
  std::vector<int> cptr(ncols + 1, 0);
  int nnz = row_ptr[m];
  for (int k = 0; k < nnz; ++k) cptr[col_idx[k] + 1]++;
  for (int j = 0; j < ncols; ++j) cptr[j + 1] += cptr[j];
  std::vector<int> crow(nnz);
  std::vector<double> cval(nnz);
  std::vector<int> fill(cptr.begin(), cptr.end() - 1);
  for (int i = 0; i < m; ++i)
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      int j = col_idx[k];
      crow[fill[j]] = i;
      cval[fill[j]] = vals[k];
      fill[j]++;
    }
  std::vector<double> buff(ncols, 0.0);
  if (method == 1) {
    for (int s = 0; s < ncols_sel; ++s) {
      int j = cols[s];
      for (int k = cptr[j]; k < cptr[j + 1]; ++k) buff[j] = buff[j] + cval[k] * vec1[crow[k]] * alpha;
    }
    for (int s = 0; s < ncols_sel; ++s) {
      int j = cols[s];
      sol[set ? j : s] = buff[j] + vec2[j] * beta;
    }
    return 0;
  }
  if (method == 0) {
    for (int s = 0; s < ncols_sel; ++s) buff[cols[s]] = vec1[set ? cols[s] : s];
    for (int i = 0; i < m; ++i) {
      sol[i] = vec2[i] * beta;
      for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sol[i] = sol[i] + vals[k] * buff[col_idx[k]] * alpha;
    }
    return 0;
  }
  return -1;
}

extern "C" int ora_dense_solve(int k, const double* a, const double* b, double* x, int transpose) {
  /* Wrap the dense matrix as the kernel of a basis with no slack columns. */
  Model M;
  M.m = k;
  M.n = k;
  M.N = 2 * k;
  M.cptr.assign(2 * k + 1, 0);
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i < k; ++i)
      if (a[(size_t)i * k + j] != 0.0) {
        M.ridx.push_back(i);
        M.cval.push_back(a[(size_t)i * k + j]);
      }
    M.cptr[j + 1] = (int)M.ridx.size();
  }
  for (int j = k; j < 2 * k; ++j) M.cptr[j + 1] = M.cptr[j];
  std::vector<int> basis(k);
  for (int i = 0; i < k; ++i) basis[i] = i;
  Factor F;
  if (!F.build(M, basis)) return -1;
  std::vector<double> rhs(b, b + k), out;
  if (transpose) F.solveT(rhs, out);
  else F.solve(rhs, out);
  for (int i = 0; i < k; ++i) x[i] = out[i];
  return 0;
}

extern "C" int ora_pfi_solve(int m, int t, const double* etas, const int* piv, const double* b,
                             double* sol, int transpose) {
  std::vector<Eta> E(t);
  for (int s = 0; s < t; ++s) {
    E[s].v.assign(etas + (size_t)s * m, etas + (size_t)(s + 1) * m);
    E[s].p = piv[s];
  }
  std::vector<double> v(b, b + m);
  pfi_apply(E, v, transpose != 0);
  for (int i = 0; i < m; ++i) sol[i] = v[i];
  return 0;
}

extern "C" int ora_simple_ratio_test(int nF, const int* F, const double* d_of_pos, const double* alpha,
                                     int tie_rule) {
  int q_idx = -1;
  double theta = BIG;
  if (tie_rule == 0) {
    for (int s = 0; s < nF; ++s) {
      int i = F[s];
      double t = d_of_pos[i] / alpha[i];
      if (t < theta) {
        theta = t;
        q_idx = i;
      } else if (q_idx >= 0 && std::fabs(t - theta) < EPS_Z && std::fabs(alpha[i]) > std::fabs(alpha[q_idx])) {
        theta = t;
        q_idx = i;
      }
    }
    return q_idx;
  }
  for (int s = 0; s < nF; ++s) theta = std::min(theta, d_of_pos[F[s]] / alpha[F[s]]);
  double best = -1.0;
  for (int s = 0; s < nF; ++s) {
    int i = F[s];
    double t = d_of_pos[i] / alpha[i];
    if (std::fabs(t - theta) < EPS_Z && std::fabs(alpha[i]) > best) {
      best = std::fabs(alpha[i]);
      q_idx = i;
    }
  }
  return q_idx;
}

extern "C" int ora_harris_ratio_test(int nF, const int* F, const double* d_of_pos, const double* alpha) {
  return harris_pick(nF, F, [&](int i) { return d_of_pos[i]; }, alpha);
}

extern "C" int ora_dse_pricing(int m, const signed char* bt, const double* x, const double* l, const double* u,
                               const double* beta, int* p_idx, double* delta) {
  /* DSEPricing (tpp:470-501) over arrays by basis position. */
  double best = 0.0;
  *p_idx = -1;
  *delta = 0.0;
  for (int i = 0; i < m; ++i) {
    double dt;
    if (var_infeas_rule(bt[i], x[i], l[i], u[i], dt) && std::fabs(dt) > DSE_FLOOR) {
      double w = (dt * dt) / beta[i];
      if (w > best) {
        *p_idx = i;
        *delta = dt;
        best = w;
      }
    }
  }
  return 0;
}
