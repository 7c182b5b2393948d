// Learned warm start, LP side: per-variable features of every LP of a batch,
// basis labels from solved bases, and the top-m selection of a starting
// basis from the network's scores.  One workgroup per LP.
#include "ds_reduce.hpp"
#include "ds_warm.hpp"

namespace pds {

namespace {

constexpr int WBS = 256;

// Features of LP `lp` (packing form  min c^T x, A x <= b, x >= 0; variables
// are the n structurals then the m slacks).  All in FP32.
//
// Structural j (column a_j, cost c_j):
//   0 slack flag (0)      1 c_j / C            2 sum a_j       3 |a_j|_2
//   4 nnz(a_j)            5 max a_j            6 |c_j| / sum a_j / C
//   7 u_j / B  with u_j = min_i b_i / a_ij  (largest x_j on its own)
//   8 |c_j| u_j / (C B)   9 f_j = sum_i a_ij / b_i  (budget share per unit)
//  10 |c_j| / f_j / C    11 d_j / C  with d_j = c_j + sum_i a_ij y_i
//  12 rank of feature 10 among the structurals / n
//  13 rank of d_j among the structurals / n
//  14 max_i t_i over the rows of a_j     15 max_i y_i / C over those rows
//  16 family prior (0 here; see ds_learn.cpp)
// Slack of row i:
//   0 slack flag (1)      1 0                  2 sum a_i.      3 |a_i.|_2
//   4 nnz(a_i.)           5 max a_i.           6 b_i / B
//   7 t_i = sum a_i. / b_i (row load at x = 1)   8 y_i / C
//   9 sum_j a_ij |c_j| / b_i / C   10 rank of t_i / m   11 rank of y_i / m
//  12-15 zero                            16 family prior (0 here)
// with C = mean |c_j|, B = mean b_i and the row price estimate
// y_i = mean over the row's columns of |c_j| / sum a_j.
__global__ __launch_bounds__(WBS) void k_features(BatchDev D, float* __restrict__ F) {
  extern __shared__ float lds_f[];
  const int lp = blockIdx.x;
  const int m = D.m[lp], n = D.n[lp];
  const int* cp = D.colptr + D.cp_off[lp];
  const int* ri = D.rowidx + D.nz_off[lp];
  const double* cv = D.cval + D.nz_off[lp];
  const int* rp = D.rowptr + D.rp_off[lp];
  const int* ci = D.colidx + D.nz_off[lp];
  const double* rv = D.rval + D.nz_off[lp];
  const double* c = D.c + D.vo[lp];
  const double* l = D.l + D.vo[lp];
  float* out = F + (size_t)D.vo[lp] * WS_NF;
  float* colsum = lds_f;        // n
  float* key1 = colsum + n;     // n: bang for buck
  float* key2 = key1 + n;       // n: reduced-cost estimate
  float* bvec = key2 + n;       // m
  float* yv = bvec + m;         // m
  float* tv = yv + m;           // m
  const int tid = threadIdx.x;
  double sc = 0.0, sb = 0.0;
  for (int j = tid; j < n; j += WBS) sc += fabs(c[j]);
  for (int i = tid; i < m; i += WBS) sb += -l[n + i];
  sc = block_sum<WBS>(sc);
  sb = block_sum<WBS>(sb);
  const float Cm = (float)(sc / n) > 0.f ? (float)(sc / n) : 1.f;
  const float Bm = (float)(sb / m) > 0.f ? (float)(sb / m) : 1.f;
  for (int i = tid; i < m; i += WBS) bvec[i] = (float)(-l[n + i]);
  for (int j = tid; j < n; j += WBS) {
    float s = 0.f;
    for (int e = cp[j]; e < cp[j + 1]; ++e) s += (float)cv[e];
    colsum[j] = s;
  }
  __syncthreads();
  // Rows.
  for (int i = tid; i < m; i += WBS) {
    float s = 0.f, s2 = 0.f, mx = 0.f, yy = 0.f, vp = 0.f;
    const int nz = rp[i + 1] - rp[i];
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const float a = (float)rv[e];
      const int j = ci[e];
      const float cj = fabsf((float)c[j]);
      s += a;
      s2 += a * a;
      mx = fmaxf(mx, a);
      yy += colsum[j] > 0.f ? cj / colsum[j] : 0.f;
      vp += a * cj;
    }
    const float b = bvec[i] > 0.f ? bvec[i] : 1e-30f;
    yy = nz > 0 ? yy / nz : 0.f;
    yv[i] = yy;
    tv[i] = s / b;
    float* f = out + (size_t)(n + i) * WS_NF;
    f[0] = 1.f;
    f[1] = 0.f;
    f[2] = s;
    f[3] = sqrtf(s2);
    f[4] = (float)nz;
    f[5] = mx;
    f[6] = bvec[i] / Bm;
    f[7] = s / b;
    f[8] = yy / Cm;
    f[9] = vp / b / Cm;
  }
  __syncthreads();
  // Columns.
  for (int j = tid; j < n; j += WBS) {
    float s2 = 0.f, mx = 0.f, u = 3.0e38f, fj = 0.f, dj = (float)c[j], tmax = 0.f, ymax = 0.f;
    for (int e = cp[j]; e < cp[j + 1]; ++e) {
      const float a = (float)cv[e];
      const int i = ri[e];
      const float b = bvec[i] > 0.f ? bvec[i] : 1e-30f;
      s2 += a * a;
      mx = fmaxf(mx, a);
      if (a > 0.f) u = fminf(u, b / a);
      fj += a / b;
      dj += a * yv[i];
      tmax = fmaxf(tmax, tv[i]);
      ymax = fmaxf(ymax, yv[i]);
    }
    const float cj = fabsf((float)c[j]);
    if (u > 1.0e38f) u = 0.f;
    const float bb = fj > 0.f ? cj / fj : 0.f;
    key1[j] = bb;
    key2[j] = dj;
    float* f = out + (size_t)j * WS_NF;
    f[0] = 0.f;
    f[1] = (float)c[j] / Cm;
    f[2] = colsum[j];
    f[3] = sqrtf(s2);
    f[4] = (float)(cp[j + 1] - cp[j]);
    f[5] = mx;
    f[6] = colsum[j] > 0.f ? cj / colsum[j] / Cm : 0.f;
    f[7] = u / Bm;
    f[8] = cj * u / (Cm * Bm);
    f[9] = fj;
    f[10] = bb / Cm;
    f[11] = dj / Cm;
    f[14] = tmax;
    f[15] = ymax / Cm;
    f[16] = 0.f;
  }
  __syncthreads();
  // Ranks (ties broken by index), O(n^2) comparisons over the workgroup.
  for (int j = tid; j < n; j += WBS) {
    const float a = key1[j], b = key2[j];
    int r1 = 0, r2 = 0;
    for (int k = 0; k < n; ++k) {
      r1 += (key1[k] < a || (key1[k] == a && k < j)) ? 1 : 0;
      r2 += (key2[k] < b || (key2[k] == b && k < j)) ? 1 : 0;
    }
    float* f = out + (size_t)j * WS_NF;
    f[12] = (float)r1 / n;
    f[13] = (float)r2 / n;
  }
  for (int i = tid; i < m; i += WBS) {
    const float a = tv[i], b = yv[i];
    int r1 = 0, r2 = 0;
    for (int k = 0; k < m; ++k) {
      r1 += (tv[k] < a || (tv[k] == a && k < i)) ? 1 : 0;
      r2 += (yv[k] < b || (yv[k] == b && k < i)) ? 1 : 0;
    }
    float* f = out + (size_t)(n + i) * WS_NF;
    f[10] = (float)r1 / m;
    f[11] = (float)r2 / m;
    f[12] = f[13] = f[14] = f[15] = f[16] = 0.f;
  }
}

// y[v] = 1 for the variables of LP lp's basis, 0 otherwise.
__global__ __launch_bounds__(WBS) void k_labels(BatchDev D, const int* __restrict__ basis, float* __restrict__ y) {
  const int lp = blockIdx.x;
  const int m = D.m[lp], N = m + D.n[lp];
  float* yl = y + D.vo[lp];
  for (int j = threadIdx.x; j < N; j += WBS) yl[j] = 0.f;
  __syncthreads();
  const int* b = basis + D.mo[lp];
  for (int i = threadIdx.x; i < m; i += WBS)
    if (b[i] >= 0 && b[i] < N) yl[b[i]] = 1.f;
}

// Starting basis of LP lp: its m variables of highest score (ties to the
// lower index), by a bitonic sort of (score, index) in LDS.
__global__ __launch_bounds__(WBS) void k_select(BatchDev D, const float* __restrict__ z, int* __restrict__ warm) {
  extern __shared__ float lds_s[];
  const int lp = blockIdx.x;
  const int m = D.m[lp], N = m + D.n[lp];
  int P2 = 1;
  while (P2 < N) P2 <<= 1;
  float* key = lds_s;
  int* idx = (int*)(key + P2);
  const float* zl = z + D.vo[lp];
  for (int j = threadIdx.x; j < P2; j += WBS) {
    key[j] = j < N ? zl[j] : -3.0e38f;
    idx[j] = j;
  }
  __syncthreads();
  // Descending by key, ascending by index among equal keys.
  for (int k = 2; k <= P2; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = threadIdx.x; i < P2; i += WBS) {
        const int o = i ^ jj;
        if (o > i) {
          const bool first_before = (key[i] > key[o]) || (key[i] == key[o] && idx[i] < idx[o]);
          const bool up = (i & k) == 0;
          if (first_before != up) {
            const float tk = key[i];
            key[i] = key[o];
            key[o] = tk;
            const int ti = idx[i];
            idx[i] = idx[o];
            idx[o] = ti;
          }
        }
      }
      __syncthreads();
    }
  int* w = warm + D.mo[lp];
  for (int i = threadIdx.x; i < m; i += WBS) w[i] = idx[i];
}

// Column mean and standard deviation of an R x WS_NF FP32 matrix: one
// workgroup per column, FP64 accumulation.
__global__ __launch_bounds__(WBS) void k_colstats(const float* __restrict__ F, long long R, float* __restrict__ mu,
                                                  float* __restrict__ sd) {
  const int col = blockIdx.x;
  double s = 0.0, s2 = 0.0;
  for (long long r = threadIdx.x; r < R; r += WBS) {
    const double v = F[r * WS_NF + col];
    s += v;
    s2 += v * v;
  }
  s = block_sum<WBS>(s);
  s2 = block_sum<WBS>(s2);
  if (threadIdx.x == 0) {
    const double mean = s / (double)R;
    const double var = s2 / (double)R - mean * mean;
    mu[col] = (float)mean;
    sd[col] = var > 1e-12 ? (float)sqrt(var) : 1.f;
  }
}

// Family prior (last feature) of a batch of T LPs with N variables each:
// cnt[j] = sum over the LPs of y[k N + j].
__global__ void k_prior_count(const float* __restrict__ y, int T, int N, float* __restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int k = 0; k < T; ++k) s += y[(size_t)k * N + j];
  cnt[j] = s;
}

// F[r][WS_NF - 1] = (base[r mod N] - y[r]) / denom (y may be null: 0).
__global__ void k_put_prior(float* __restrict__ F, const float* __restrict__ base, const float* __restrict__ y,
                            long long R, int N, float denom) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const float v = base[r % N] - (y ? y[r] : 0.f);
  F[r * WS_NF + (WS_NF - 1)] = v / denom;
}

}  // namespace

size_t features_lds_bytes(int mmax, int nmax) { return sizeof(float) * (3 * (size_t)nmax + 3 * (size_t)mmax); }

size_t select_lds_bytes(int Nmax) {
  int P2 = 1;
  while (P2 < Nmax) P2 <<= 1;
  return (sizeof(float) + sizeof(int)) * (size_t)P2;
}

void launch_features(hipStream_t s, const BatchDev& D, int count, int mmax, int nmax, float* F) {
  k_features<<<count, WBS, features_lds_bytes(mmax, nmax), s>>>(D, F);
}
void launch_labels(hipStream_t s, const BatchDev& D, int count, const int* basis, float* y) {
  k_labels<<<count, WBS, 0, s>>>(D, basis, y);
}
void launch_select(hipStream_t s, const BatchDev& D, int count, int Nmax, const float* z, int* warm) {
  k_select<<<count, WBS, select_lds_bytes(Nmax), s>>>(D, z, warm);
}
void launch_prior_count(hipStream_t s, const float* y, int T, int N, float* cnt) {
  k_prior_count<<<(N + WBS - 1) / WBS, WBS, 0, s>>>(y, T, N, cnt);
}
void launch_put_prior(hipStream_t s, float* F, const float* base, const float* y, long long R, int N, float denom) {
  k_put_prior<<<(unsigned)((R + WBS - 1) / WBS), WBS, 0, s>>>(F, base, y, R, N, denom);
}
void launch_colstats(hipStream_t s, const float* F, long long R, float* mu, float* sd) {
  k_colstats<<<WS_NF, WBS, 0, s>>>(F, R, mu, sd);
}

}  // namespace pds
