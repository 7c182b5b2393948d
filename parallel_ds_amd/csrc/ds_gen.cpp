// Seeded synthetic LP generators (host code).  They produce the workloads the
// benchmark and the tests run: random sparse packing LPs, a balanced
// transportation LP and lognormal perturbations of a base instance.
//
// Random stream: SplitMix64, doubles in [0,1) from the top 53 bits, so any
// other implementation of the same recipe reproduces the instances bit for bit.
#include "ds_gen.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "ds_rng.hpp"

using pds::SplitMix;

extern "C" int ds_gen_bernoulli_lp(int m, int n, double density, uint64_t seed, int* col_ptr,
                                   int* row_idx, double* val, int cap, double* b, double* c) {
  if (m <= 0 || n <= 0 || !col_ptr || !row_idx || !val || !b || !c) return -1;
  SplitMix rng(seed);
  int nnz = 0;
  col_ptr[0] = 0;
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      if (rng.uniform() < density) {
        if (nnz >= cap) return -2;
        row_idx[nnz] = i;
        val[nnz] = 0.1 + 0.9 * rng.uniform();
        ++nnz;
      }
    }
    if (nnz == col_ptr[j]) {
      // An empty column with a negative cost would make the LP unbounded;
      // give it one entry at a uniform random row so every instance is bounded.
      if (nnz >= cap) return -2;
      int r = (int)(rng.uniform() * m);
      row_idx[nnz] = r < m ? r : m - 1;
      val[nnz] = 0.1 + 0.9 * rng.uniform();
      ++nnz;
    }
    col_ptr[j + 1] = nnz;
  }
  std::vector<double> x0(n);
  for (int j = 0; j < n; ++j) x0[j] = rng.uniform();
  std::vector<double> ax(m, 0.0);
  for (int j = 0; j < n; ++j)
    for (int e = col_ptr[j]; e < col_ptr[j + 1]; ++e) ax[row_idx[e]] += val[e] * x0[j];
  for (int i = 0; i < m; ++i) b[i] = ax[i] + rng.uniform();
  for (int j = 0; j < n; ++j) c[j] = -rng.uniform();
  return nnz;
}

extern "C" int ds_gen_kcol_lp(int m, int n, int k, uint64_t seed, int* col_ptr, int* row_idx, double* val,
                              double* b, double* c) {
  if (m <= 0 || n <= 0 || k <= 0 || k > m || !col_ptr || !row_idx || !val || !b || !c) return -1;
  SplitMix rng(seed);
  std::vector<int> rows(k);
  col_ptr[0] = 0;
  for (int j = 0; j < n; ++j) {
    // k distinct rows by rejection, then ascending order.
    for (int t = 0; t < k; ++t) {
      while (true) {
        int r = (int)(rng.uniform() * m);
        if (r >= m) r = m - 1;
        bool dup = false;
        for (int s = 0; s < t; ++s) dup |= (rows[s] == r);
        if (!dup) { rows[t] = r; break; }
      }
    }
    std::sort(rows.begin(), rows.end());
    for (int t = 0; t < k; ++t) {
      row_idx[(size_t)j * k + t] = rows[t];
      val[(size_t)j * k + t] = 0.1 + 0.9 * rng.uniform();
    }
    col_ptr[j + 1] = (j + 1) * k;
  }
  std::vector<double> x0(n);
  for (int j = 0; j < n; ++j) x0[j] = rng.uniform();
  std::vector<double> ax(m, 0.0);
  for (int j = 0; j < n; ++j)
    for (int e = col_ptr[j]; e < col_ptr[j + 1]; ++e) ax[row_idx[e]] += val[e] * x0[j];
  for (int i = 0; i < m; ++i) b[i] = ax[i] + rng.uniform();
  for (int j = 0; j < n; ++j) c[j] = -rng.uniform();
  return n * k;
}

extern "C" int ds_gen_transport(int S, int D, uint64_t seed, int* col_ptr, int* row_idx, double* val,
                                double* rhs, double* c) {
  if (S <= 0 || D <= 0 || !col_ptr || !row_idx || !val || !rhs || !c) return -1;
  SplitMix rng(seed);
  std::vector<long long> sup(S), dem(D);
  long long total = 0;
  for (int i = 0; i < S; ++i) {
    sup[i] = 1 + (long long)(rng.uniform() * 100.0);
    if (sup[i] > 100) sup[i] = 100;
    total += sup[i];
  }
  std::vector<double> raw(D);
  double rsum = 0.0;
  for (int j = 0; j < D; ++j) {
    long long v = 1 + (long long)(rng.uniform() * 100.0);
    if (v > 100) v = 100;
    raw[j] = (double)v;
    rsum += raw[j];
  }
  // Rescale demands to the total supply; largest-remainder rounding keeps
  // them integral (ties broken by lower sink index).
  std::vector<double> frac(D);
  long long assigned = 0;
  for (int j = 0; j < D; ++j) {
    double exact = raw[j] * (double)total / rsum;
    dem[j] = (long long)std::floor(exact);
    frac[j] = exact - (double)dem[j];
    assigned += dem[j];
  }
  std::vector<int> order(D);
  for (int j = 0; j < D; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return frac[a] > frac[b2]; });
  for (long long r = 0; r < total - assigned; ++r) dem[order[(size_t)r % D]] += 1;
  // Variable x_ij = column i*D + j: one entry in supply row i, one in demand row S + j.
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < D; ++j) {
      int col = i * D + j;
      col_ptr[col] = 2 * col;
      row_idx[2 * col] = i;
      row_idx[2 * col + 1] = S + j;
      val[2 * col] = 1.0;
      val[2 * col + 1] = 1.0;
      long long cost = 1 + (long long)(rng.uniform() * 1000.0);
      if (cost > 1000) cost = 1000;
      c[col] = (double)cost;
    }
  col_ptr[S * D] = 2 * S * D;
  for (int i = 0; i < S; ++i) rhs[i] = (double)sup[i];
  for (int j = 0; j < D; ++j) rhs[S + j] = (double)dem[j];
  return 2 * S * D;
}

extern "C" int ds_gen_lognormal(int len, double sigma, uint64_t seed, const double* base, double* out) {
  if (len < 0 || !base || !out) return -1;
  SplitMix rng(seed);
  for (int i = 0; i < len; ++i) {
    // Box-Muller; u1 in (0,1].
    double u1 = 1.0 - rng.uniform();
    double u2 = rng.uniform();
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    out[i] = base[i] * std::exp(sigma * z);
  }
  return 0;
}
