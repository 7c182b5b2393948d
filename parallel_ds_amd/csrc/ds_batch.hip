// Batched dual simplex on gfx950: thousands of small LPs in one persistent
// launch, one workgroup per LP at a time, instances taken from an atomic
// work queue until it is exhausted.
//
// Each workgroup runs the reference's algorithm (phase 1 = minimize the sum
// of dual infeasibilities, baseDualSimplex.tpp:799-934; phase 2 with dual
// steepest-edge pricing and the textbook ratio test,
// sequentialDualSimplex.cpp:253-515; the same restart, perturbation and
// refactor rules) on an explicit dense basis inverse: rows of B^{-1} give
// BTRAN, columns give FTRAN, and each basis change is the product-form row
// update B^{-1} <- E B^{-1}.  "Refactor" recomputes B^{-1} by Gauss-Jordan
// with partial pivoting.  All vectors of the LP live in LDS; the inverse
// lives in LDS for the small size class and in the workgroup's slice of a
// global scratch buffer (L2/MALL-resident) otherwise.
#include "ds_batch.hpp"
#include "ds_reduce.hpp"
#include "ds_rules.hpp"

namespace pds {

namespace {

__device__ inline double dev_uniform(unsigned long long& s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}


// Dense kernels on the explicit inverse (row-major, m x m).  They take
// __restrict__ pointers and keep several loads in flight per thread: with
// plain pointers the compiler must assume a store to B may change the LDS
// vectors and serialises every element behind the previous store.

// Orders the LDS (and global) accesses of one wavefront: every lane's
// earlier stores are visible to every lane's later loads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row of flat index e < m*m (exact for m <= 4096: e/m is at least 0.5/m away
// from the next integer, far above the float rounding error).
__device__ __forceinline__ int flat_row(int e, float rinv, float half) {
  return (int)((float)e * rinv + half);
}

// One Gauss-Jordan / product-form step on B:
//   row p    <- r
//   row i!=p <- B[i] - a[i] r, with column zc read as zero (zc < 0: none).
template <int BS>
__device__ __forceinline__ void inv_eliminate(double* __restrict__ B, const double* __restrict__ a,
                                              const double* __restrict__ r, int m, int p, int zc, int tid) {
  const int mm = m * m;
  const float rinv = 1.0f / (float)m, half = 0.5f * rinv;
  for (int e0 = tid; e0 < mm; e0 += 4 * BS) {
    double v[4], ai[4], rc[4];
    bool piv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k * BS;
      const bool in = e < mm;
      const int i = flat_row(e, rinv, half), cc = e - i * m;
      piv[k] = i == p;
      v[k] = (in && cc != zc) ? B[e] : 0.0;
      ai[k] = in ? a[i] : 0.0;
      rc[k] = in ? r[cc] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k * BS;
      if (e < mm) B[e] = piv[k] ? rc[k] : v[k] - ai[k] * rc[k];
    }
  }
}

// out[i] = sum_r B[i][r] v[r]: one wavefront per row, two rows at a time.
template <int BS>
__device__ __forceinline__ void inv_matvec(const double* __restrict__ B, const double* __restrict__ v,
                                           double* __restrict__ out, int m, int tid) {
  constexpr int NW = BS / WAVE;
  const int lane = tid % WAVE, wv = tid / WAVE;
  for (int i = wv; i < m; i += 2 * NW) {
    const int i2 = i + NW;
    const bool two = i2 < m;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 4
    for (int r = lane; r < m; r += WAVE) {
      const double vr = v[r];
      s0 += B[i * m + r] * vr;
      if (two) s1 += B[i2 * m + r] * vr;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane == 0) {
      out[i] = s0;
      if (two) out[i2] = s1;
    }
  }
}

// out[i] = sum_k B[i][cmap[idx[k]]] val[k] over the nz entries of one
// column: one wavefront per row, four rows in flight, the column's (mapped)
// entries held in registers across rows.
template <int BS>
__device__ __forceinline__ void inv_spcol(const double* __restrict__ B, const int* __restrict__ cmap,
                                          const int* __restrict__ idx, const double* __restrict__ val, int nz,
                                          double* __restrict__ out, int m, int tid) {
  constexpr int NW = BS / WAVE;
  const int lane = tid % WAVE, wv = tid / WAVE;
  if (nz <= 2 * WAVE) {
    int ix[2];
    double vx[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = lane + k * WAVE;
      ix[k] = e < nz ? cmap[idx[e]] : 0;
      vx[k] = e < nz ? val[e] : 0.0;
    }
    for (int i0 = wv; i0 < m; i0 += 4 * NW) {
      double s[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r * NW;
        s[r] = 0.0;
        if (i < m) {
          const double* Bi = B + (size_t)i * m;
#pragma unroll
          for (int k = 0; k < 2; ++k)
            if (lane + k * WAVE < nz) s[r] += Bi[ix[k]] * vx[k];
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r * NW;
        const double t = wave_sum(s[r]);
        if (lane == 0 && i < m) out[i] = t;
      }
    }
  } else {
    for (int i = wv; i < m; i += NW) {
      const double* Bi = B + (size_t)i * m;
      double s = 0.0;
      for (int e = lane; e < nz; e += WAVE) s += Bi[cmap[idx[e]]] * val[e];
      s = wave_sum(s);
      if (lane == 0) out[i] = s;
    }
  }
}

// The basis change on B in one pass.  For every row i, from the old row:
//   FT:  a[i] = B[i] . ad   (FTRAN of the entering column, ad dense in storage
//        column order; a[p] is set to ap, computed the same way beforehand)
//   tau[i] = B[i] . rs      (rs = old row p: the DSE vector B^{-1} rho)
// then row p <- pr and row i != p <- B[i] - a[i] pr.  One wavefront per row,
// four rows in flight; rows of m <= 256 are held in registers.
template <int BS, bool FT>
__device__ __forceinline__ void inv_update_fused(double* __restrict__ B, const double* __restrict__ rs,
                                                 const double* __restrict__ pr, const double* __restrict__ ad,
                                                 double ap, double* __restrict__ a, double* __restrict__ tau,
                                                 int m, int p, int tid) {
  constexpr int NW = BS / WAVE;
  const int lane = tid % WAVE, wv = tid / WAVE;
  if (m <= 4 * WAVE) {
    double rr[4], pp[4], dd[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = lane + k * WAVE;
      rr[k] = c < m ? rs[c] : 0.0;
      pp[k] = c < m ? pr[c] : 0.0;
      dd[k] = (FT && c < m) ? ad[c] : 0.0;
    }
    for (int i0 = wv; i0 < m; i0 += 4 * NW) {
      double v[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r * NW;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int c = lane + k * WAVE;
          v[r][k] = (i < m && c < m) ? B[(size_t)i * m + c] : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r * NW;
        if (i >= m) continue;
        double ai;
        if (FT) {
          double sa = 0.0;
#pragma unroll
          for (int k = 0; k < 4; ++k) sa += v[r][k] * dd[k];
          sa = wave_sum(sa);
          ai = (i == p) ? ap : sa;
        } else {
          ai = a[i];
        }
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int c = lane + k * WAVE;
          s += v[r][k] * rr[k];
          if (c < m) B[(size_t)i * m + c] = (i == p) ? pp[k] : v[r][k] - ai * pp[k];
        }
        s = wave_sum(s);
        if (lane == 0) {
          tau[i] = s;
          if (FT) a[i] = ai;
        }
      }
    }
  } else {
    for (int i = wv; i < m; i += NW) {
      double ai = 0.0;
      if (FT) {
        double sa = 0.0;
        for (int c = lane; c < m; c += WAVE) sa += B[(size_t)i * m + c] * ad[c];
        sa = wave_sum(sa);
        ai = (i == p) ? ap : sa;
      } else {
        ai = a[i];
      }
      double s = 0.0;
      for (int c = lane; c < m; c += WAVE) {
        const double v = B[(size_t)i * m + c];
        s += v * rs[c];
        B[(size_t)i * m + c] = (i == p) ? pr[c] : v - ai * pr[c];
      }
      s = wave_sum(s);
      if (lane == 0) {
        tau[i] = s;
        if (FT) a[i] = ai;
      }
    }
  }
}

// v . d over m entries in the lane order the fused pass uses for one row.
__device__ __forceinline__ double wave_rowdot(const double* __restrict__ v, const double* __restrict__ d, int m,
                                              int lane) {
  double s = 0.0;
  for (int c = lane; c < m; c += WAVE) s += v[c] * d[c];
  return wave_sum(s);
}

// out[r] = sum_i v[i] B[i][r]: one column per thread, coalesced across threads.
template <int BS>
__device__ __forceinline__ void inv_tmatvec(const double* __restrict__ B, const double* __restrict__ v,
                                            double* __restrict__ out, int m, int tid) {
  for (int r = tid; r < m; r += BS) {
    double y = 0.0;
#pragma unroll 8
    for (int i = 0; i < m; ++i) y += v[i] * B[i * m + r];
    out[r] = y;
  }
}

// acc_j = alpha * sum_e val[e] v[idx[e]] over column j of a CSC matrix, in
// CSC order, one column per thread.
__device__ __forceinline__ double csc_dot(const int* __restrict__ cp, const int* __restrict__ ri,
                                          const double* __restrict__ cv, const double* __restrict__ v, int j,
                                          double alpha) {
  double acc = 0.0;
  const int e1 = cp[j + 1];
#pragma unroll 4
  for (int e = cp[j]; e < e1; ++e) acc = acc + cv[e] * v[ri[e]] * alpha;
  return acc;
}

struct BShared {
  int status;
  int iteration;
  int perturbed;
  int cycle;
  int p, p_idx, q, q_idx;
  int neg;
  int nF;
  int dinf;
  int sing;
  int nblocked;
  int nrep;
  int pn, pnext, nperm;
  int pl[GJ_NB_MAX];
  int prow_of[2 * GJ_NB_MAX], psrc[2 * GJ_NB_MAX];
  double prp[GJ_NB_MAX];
  int blocked[64];
  double delta, theta, obj, f_p, corr, tmp;
  unsigned long long rng;
};

constexpr int ST_OK = -1;  // internal: keep iterating

// Every member is force-inlined into k_batch: as an out-of-line call the
// object would live in scratch memory and each field would be reloaded after
// every store through a generic pointer.
template <int BS>
struct LpSolve {
  int m, n, N;
  const int* cp;
  const int* ri;
  const double* cv;
  const int* rp;
  const int* ci;
  const double* rv;
  const double* corig;
  const double* l;
  const double* u;
  const int8_t* bt;
  double* c;
  double* x;
  double* d;
  double* wN;
  double* beta;
  double* f;
  double* rho;
  double* aq;
  double* tau;
  double* w;
  double* prow;
  double* ar;
  double* sortbuf;
  int* basis;
  int* nonbasis;
  int* posN;
  int* piv;
  int* cmap;   // B^{-1}[i][c] = Binv[i * m + cmap[c]]
  int* rpos;
  int* rflag;
  int* rat;
  double* Pn;          // panel columns (m x gj_nb), LDS
  double* Vb;          // row-operation vectors of the panel (m x gj_nb)
  double* Rb;          // deferred multipliers (gj_nb x m)
  int gj_nb;           // panel width of the factorization
  int repair_on;       // slack substitution for (near-)singular columns
  double repair_tol;
  double sing_tol;     // a pivot at or below it is singular
  double* Binv;
  BShared* sh;
  int K, max_iter, perturb_on;
  unsigned long long seed;
  int tid;

  __device__ __forceinline__ void sync() { __syncthreads(); }

  // out = B^{-1} v for v indexed by constraint row (prow is the staging
  // buffer in storage column order).
  __device__ __forceinline__ void rowdot(const double* v, double* out) {
    for (int c = tid; c < m; c += BS) prow[cmap[c]] = v[c];
    sync();
    inv_matvec<BS>(Binv, prow, out, m, tid);
    sync();
  }

  // Column j of [A | I] dotted with v, in CSC order (matrix.cpp:273-287 with alpha).
  __device__ __forceinline__ double col_dot(int j, const double* v, double alpha) const {
    double acc = 0.0;
    if (j < n) {
      acc = csc_dot(cp, ri, cv, v, j, alpha);
    } else {
      acc = acc + 1.0 * v[j - n] * alpha;
    }
    return acc;
  }

  // B^{-1} of the current basis by in-place Gauss-Jordan with partial
  // pivoting, columns taken slacks first.  A slack column is a unit vector,
  // so pivoting on it only exchanges two rows: thread 0 composes those
  // exchanges into the row placement rpos[] used to build B, and only the
  // structural columns cost an elimination pass.  The column exchanges an
  // in-place Gauss-Jordan inverse ends with are not carried out; they are
  // folded into cmap[] (B^{-1}[i][c] = Binv[i][cmap[c]]).
  __device__ __forceinline__ bool reinvert() {
    const int mm = m * m;
    if (tid == 0) {
      int* at = rat;  // original row held at each position
      for (int r = 0; r < m; ++r) {
        rpos[r] = r;
        at[r] = r;
      }
      for (int s = 0; s < m; ++s) {
        const int j = basis[s];
        if (j < n) continue;
        const int r = j - n, cur = rpos[r];
        piv[s] = cur;
        if (cur != s) {
          const int r2 = at[s];
          at[s] = r;
          at[cur] = r2;
          rpos[r] = s;
          rpos[r2] = cur;
        }
      }
      for (int s = 0; s < m; ++s) rflag[s] = basis[s] >= n ? 1 : 0;
      sh->sing = 0;
    }
    for (int e = tid; e < mm; e += BS) Binv[e] = 0.0;
    sync();
    for (int i = tid; i < m; i += BS) {
      const int j = basis[i];
      if (j < n) {
        for (int e = cp[j]; e < cp[j + 1]; ++e) Binv[rpos[ri[e]] * m + i] = cv[e];
      } else {
        Binv[rpos[j - n] * m + i] = 1.0;
      }
    }
    sync();
    // Structural columns in panels of up to gj_nb positions.  Inside a panel
    // the steps run on an LDS copy of the panel's columns (Pn, m x nb); row
    // exchanges are applied at once to the stored matrix; the elimination of
    // every other column is deferred: step t is the row operation
    // E_t = I + v_t e_{s_t}^T (v_t[i] = -w_t[i]/pv_t, v_t[s_t] = 1/pv_t - 1),
    // and a later exchange of rows a, b turns it into I + (P v_t) e_{s_t}^T
    // (s_t is used, a and b are not), so the stored v's are exchanged too.  At
    // the end of the panel each other column g becomes g + V r with
    // r_t = g[s_t] + sum_{t'<t} v_t'[s_t] r_t': one pass over the matrix per
    // panel instead of one per step.  cmap serves as the panel-index map of a
    // position while the factorization runs (it is rebuilt at the end).
    const int NB = gj_nb;
    for (int c = tid; c < m; c += BS) cmap[c] = -1;
    if (tid == 0) sh->pnext = 0;
    sync();
    while (true) {
      if (tid == 0) {
        int cnt = 0, pos = sh->pnext;
        for (; pos < m && cnt < NB; ++pos)
          if (basis[pos] < n) {
            sh->pl[cnt] = pos;
            cmap[pos] = cnt;
            ++cnt;
          }
        sh->pn = cnt;
        sh->pnext = pos;
        sh->nperm = 0;
      }
      sync();
      const int nb = sh->pn;
      if (nb == 0) break;
      for (int e = tid; e < m * nb; e += BS) {
        const int i = e / nb, j = e - i * nb;
        Pn[e] = Binv[i * m + sh->pl[j]];
      }
      sync();
      // The nb steps of the panel run in wavefront 0 alone (the panel is
      // m x nb values); the other wavefronts wait at the barrier after it.
      if (tid < WAVE) {
        const int lane = tid;
        for (int j = 0; j < nb; ++j) {
          const int sj = sh->pl[j];
          double bv = -1.0;
          int bi = INT_MAX;
          for (int r = lane; r < m; r += WAVE)
            if (!rflag[r]) {
              const double a = fabs(Pn[r * nb + j]);
              if (a >= 0.0) argmax_merge(bv, bi, a, r);
            }
          wave_argmax(bv, bi);
          if (repair_on && !(bi != INT_MAX && bv > repair_tol)) {
            // Slack substitution: column sj depends (numerically) on the
            // columns already pivoted.  Every position below sj is used (slack
            // positions by the slack pass, structural ones by their steps), so
            // position sj is the lowest free one; the slack of the original row
            // held there enters.  Transformed by the steps so far, that slack's
            // column is the unit vector at sj: row operations only add
            // multiples of pivot rows, and the row at sj has not been one.
            if (lane == 0) {
              const int jold = basis[sj], jnew = n + rat[sj];
              const int k = posN[jnew];
              nonbasis[k] = jold;
              posN[jold] = k;
              posN[jnew] = -1;
              basis[sj] = jnew;
              sh->nrep += 1;
            }
            for (int i = lane; i < m; i += WAVE) Pn[i * nb + j] = (i == sj) ? 1.0 : 0.0;
            bi = sj;
            bv = 1.0;
            wave_lds_sync();
          }
          if (bi == INT_MAX || !(bv > sing_tol)) {
            if (lane == 0) sh->sing = 1;
            break;
          }
          if (lane == 0) {
            piv[sj] = bi;
            rflag[sj] = 1;
            const int t = rat[sj];
            rat[sj] = rat[bi];
            rat[bi] = t;
          }
          if (bi != sj) {
            // The exchange of rows sj and bi in the columns outside the panel
            // is recorded (row -> source row) and applied once per panel.
            if (lane == 0) {
              int ka = -1, kb = -1;
              for (int k = 0; k < sh->nperm; ++k) {
                if (sh->prow_of[k] == sj) ka = k;
                if (sh->prow_of[k] == bi) kb = k;
              }
              if (ka < 0) {
                ka = sh->nperm++;
                sh->prow_of[ka] = sj;
                sh->psrc[ka] = sj;
              }
              if (kb < 0) {
                kb = sh->nperm++;
                sh->prow_of[kb] = bi;
                sh->psrc[kb] = bi;
              }
              const int t = sh->psrc[ka];
              sh->psrc[ka] = sh->psrc[kb];
              sh->psrc[kb] = t;
            }
            for (int j2 = lane; j2 < nb; j2 += WAVE) {
              const double a = Pn[sj * nb + j2];
              Pn[sj * nb + j2] = Pn[bi * nb + j2];
              Pn[bi * nb + j2] = a;
            }
            for (int t = lane; t < j; t += WAVE) {
              const double a = Vb[sj * nb + t];
              Vb[sj * nb + t] = Vb[bi * nb + t];
              Vb[bi * nb + t] = a;
            }
          }
          wave_lds_sync();
          const double pv = Pn[sj * nb + j];
          const double ipv = 1.0 / pv;
          for (int i = lane; i < m; i += WAVE) Vb[i * nb + j] = (i == sj) ? ipv - 1.0 : -Pn[i * nb + j] / pv;
          for (int j2 = lane; j2 < nb; j2 += WAVE) sh->prp[j2] = Pn[sj * nb + j2];
          wave_lds_sync();
          for (int i = lane; i < m; i += WAVE) {
            const double vi = Vb[i * nb + j];
            for (int j2 = 0; j2 < nb; ++j2) {
              const int e = i * nb + j2;
              if (j2 == j) Pn[e] = (i == sj) ? ipv : vi;
              else Pn[e] = Pn[e] + vi * sh->prp[j2];
            }
          }
          wave_lds_sync();
        }
      }
      sync();
      if (sh->sing) return false;
      // Row exchanges of the panel on the columns outside it: each thread
      // owns columns, reads every source row first, then writes.
      {
        const int np = sh->nperm;
        if (np > 0) {
          for (int c = tid; c < m; c += BS) {
            if (cmap[c] >= 0) continue;
            double v[2 * GJ_NB_MAX];
#pragma unroll
            for (int k = 0; k < 2 * GJ_NB_MAX; ++k)
              if (k < np) v[k] = Binv[sh->psrc[k] * m + c];
#pragma unroll
            for (int k = 0; k < 2 * GJ_NB_MAX; ++k)
              if (k < np) Binv[sh->prow_of[k] * m + c] = v[k];
          }
          sync();
        }
      }
      // Deferred elimination of the columns outside the panel.
      for (int c = tid; c < m; c += BS) {
        if (cmap[c] >= 0) continue;
        double r[GJ_NB_MAX];
#pragma unroll
        for (int t = 0; t < GJ_NB_MAX; ++t) {
          if (t >= nb) break;
          double v = Binv[sh->pl[t] * m + c];
          for (int t2 = 0; t2 < t; ++t2) v += Vb[sh->pl[t] * nb + t2] * r[t2];
          r[t] = v;
          Rb[t * m + c] = v;
        }
      }
      sync();
      {
        // G += V R over the columns outside the panel, the panel columns
        // copied back; four elements in flight per thread.
        const float rinv = 1.0f / (float)m, half = 0.5f * rinv;
        for (int e0 = tid; e0 < mm; e0 += 4 * BS) {
          double g[4];
          int ii[4], cc[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int e = e0 + k * BS;
            const int i = flat_row(e, rinv, half);
            ii[k] = i;
            cc[k] = e - i * m;
            g[k] = (e < mm && cmap[cc[k]] < 0) ? Binv[e] : 0.0;
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int e = e0 + k * BS;
            if (e >= mm) continue;
            const int pj = cmap[cc[k]];
            double v;
            if (pj >= 0) {
              v = Pn[ii[k] * nb + pj];
            } else {
              v = g[k];
              for (int t = 0; t < nb; ++t) v += Vb[ii[k] * nb + t] * Rb[t * m + cc[k]];
            }
            Binv[e] = v;
          }
        }
      }
      sync();
      for (int j = tid; j < nb; j += BS) cmap[sh->pl[j]] = -1;
      sync();
    }
    if (tid == 0) {
      for (int c = 0; c < m; ++c) cmap[c] = c;
      for (int pass = 0; pass < 2; ++pass)
        for (int s = m - 1; s >= 0; --s) {
          const bool slack = basis[s] >= n;
          if (slack != (pass == 1) || piv[s] == s) continue;
          const int t = cmap[piv[s]];
          cmap[piv[s]] = cmap[s];
          cmap[s] = t;
        }
    }
    sync();
    return true;
  }

  // initBetaWeights (tpp:34-70): identity or exact row norms of B^{-1}.
  __device__ __forceinline__ void init_beta(bool identity) {
    if (identity) {
      for (int i = tid; i < m; i += BS) beta[i] = 1.0;
    } else {
      const int lane = tid % WAVE, wv = tid / WAVE;
      for (int i = wv; i < m; i += BS / WAVE) {
        double s2 = 0.0;
        for (int r = lane; r < m; r += WAVE) s2 += Binv[i * m + r] * Binv[i * m + r];
        s2 = wave_sum(s2);
        if (lane == 0) beta[i] = s2;
      }
    }
    sync();
  }

  // y = B^{-T} c_B ; d_N = c_N - A_N^T y  (sequentialDualSimplex.cpp:119-132)
  __device__ __forceinline__ void init_reduced_costs() {
    for (int i = tid; i < m; i += BS) tau[i] = c[basis[i]];
    sync();
    inv_tmatvec<BS>(Binv, tau, prow, m, tid);
    sync();
    for (int r = tid; r < m; r += BS) w[r] = prow[cmap[r]];
    sync();
    for (int i = tid; i < n; i += BS) {
      int j = nonbasis[i];
      d[j] = col_dot(j, w, -1.0) + c[j] * 1.0;
    }
    sync();
  }

  // setPrimalVars (tpp:74-101)
  __device__ __forceinline__ void set_primal() {
    for (int i = tid; i < n; i += BS) {
      int j = nonbasis[i];
      double v = 0.0;
      switch (bt[j]) {
        case BT_FIXED: v = l[j]; break;
        case BT_FREE: v = 0.0; break;
        case BT_BOXED: v = (d[j] > 0) ? l[j] : u[j]; break;
        case BT_UPPER: v = u[j]; break;
        case BT_LOWER: v = l[j]; break;
      }
      x[j] = v;
    }
    sync();
  }

  // dualSimplexInit (sequentialDualSimplex.cpp:230-250)
  __device__ __forceinline__ void dual_init() {
    set_primal();
    for (int r = tid; r < m; r += BS) {
      double s = 0.0 * 1.0;
      for (int e = rp[r]; e < rp[r + 1]; ++e) {
        int col = ci[e];
        double bf = (posN[col] >= 0) ? x[col] : 0.0;
        s = s + rv[e] * bf * -1.0;
      }
      double bf = (posN[n + r] >= 0) ? x[n + r] : 0.0;
      s = s + 1.0 * bf * -1.0;
      w[r] = s;
    }
    sync();
    rowdot(w, tau);
    for (int i = tid; i < m; i += BS) x[basis[i]] = tau[i];
    sync();
    init_reduced_costs();
    double s = 0.0;
    for (int j = tid; j < N; j += BS) s += c[j] * x[j];
    s = block_sum<BS>(s);
    if (tid == 0) sh->obj = s;
    sync();
  }

  // minimizeDualInfeasibilityInit + initPhase1PricingVector
  __device__ __forceinline__ void phase1_init() {
    init_reduced_costs();
    double z = 0.0;
    int cnt = 0;
    for (int j = tid; j < N; j += BS) {
      double wt = 0.0;
      if (posN[j] >= 0) {
        int8_t t = bt[j];
        double dj = d[j];
        if ((t == BT_UPPER || t == BT_FREE) && dj > EPS_D) wt = -1.0;
        else if ((t == BT_LOWER || t == BT_FREE) && dj < -EPS_D) wt = 1.0;
        z += wt * dj;
        cnt += dual_infeas(t, dj);
      }
      wN[j] = wt;
    }
    z = block_sum<BS>(z);
    cnt = block_sumi<BS>(cnt);
    if (tid == 0) {
      sh->obj = z;
      sh->dinf = cnt;
    }
    sync();
    for (int r = tid; r < m; r += BS) {
      double s = 0.0;
      for (int e = rp[r]; e < rp[r + 1]; ++e) s += rv[e] * wN[ci[e]];
      s += wN[n + r];
      w[r] = s;
    }
    sync();
    rowdot(w, f);
  }

  // checkPerturbNeed (tpp:223-234): bitonic sort of the costs, then the
  // tolerance-unique count of the sorted sequence.
  __device__ __forceinline__ bool perturb_need() {
    int P2 = 1;
    while (P2 < N) P2 <<= 1;
    for (int j = tid; j < P2; j += BS) sortbuf[j] = (j < N) ? c[j] : 1e300;
    sync();
    for (int k = 2; k <= P2; k <<= 1)
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int i = tid; i < P2; i += BS) {
          int ixj = i ^ jj;
          if (ixj > i) {
            bool up = (i & k) == 0;
            double a = sortbuf[i], b = sortbuf[ixj];
            if ((a > b) == up) {
              sortbuf[i] = b;
              sortbuf[ixj] = a;
            }
          }
        }
        sync();
      }
    if (tid == 0) {
      int uniq = 0;
      double last = 0.0;
      for (int i = 0; i < N; ++i)
        if (i == 0 || !(fabs(last - sortbuf[i]) < EPS_COSTS)) {
          last = sortbuf[i];
          ++uniq;
        }
      sh->tmp = (uniq < PERTURB_RATIO * (double)N) ? 1.0 : 0.0;
    }
    sync();
    return sh->tmp != 0.0;
  }

  // perturbCosts (tpp:238-275) with the seeded stream.
  __device__ __forceinline__ void perturb_costs() {
    double s = 0.0;
    for (int j = tid; j < N; j += BS) s += c[j];
    s = block_sum<BS>(s);
    if (tid == 0) {
      const double wts[10] = {0.01, 0.1, 1, 2, 5, 10, 20, 30, 40, 100};
      const double mean = s / (double)N;
      const double min_p = fmin(1e-2 * EPS_D, PSI);
      const double max_p = fmax(1e+3 * EPS_D, PSI * 10 * mean);
      unsigned long long st = sh->rng;
      for (int i = 0; i < N; ++i) {
        double magnitude = 100 * EPS_D + PSI * c[i];
        double pert;
        if (bt[i] != BT_LOWER && bt[i] != BT_FREE) pert = -0.5 * magnitude * (1 + dev_uniform(st));
        pert = 0.5 * magnitude * (1 + dev_uniform(st));
        int cnt = (i < n) ? (cp[i + 1] - cp[i]) : 1;
        int kk = cnt > 10 ? 9 : (cnt == 0 ? 0 : cnt - 1);
        pert = wts[kk] * pert;
        if (pert != 0.0)
          for (int guard = 0; guard < 4000 && (fabs(pert) > fabs(max_p) || fabs(pert) < fabs(min_p)); ++guard) {
            if (fabs(pert) > fabs(max_p)) pert = 0.1 * pert;
            else if (fabs(pert) < fabs(min_p)) pert = 10 * pert;
          }
        c[i] += pert;
      }
      sh->rng = st;
      sh->perturbed = 1;
    }
    sync();
  }

  // One basis change shared by both phases (sequentialDualSimplex.cpp:174-227).
  // alpha_p = B^{-1}[p] . a_q with the entering column scattered densely in
  // storage order into rho (free once the pivot row is formed) and row p of
  // B^{-1} staged in w; the result goes to sh->tmp.
  __device__ __forceinline__ double pivot_alpha() {
    const int q = sh->q, p_idx = sh->p_idx;
    for (int c = tid; c < m; c += BS) {
      rho[c] = 0.0;
      w[c] = Binv[p_idx * m + c];
    }
    sync();
    if (q < n) {
      for (int e = cp[q] + tid; e < cp[q + 1]; e += BS) rho[cmap[ri[e]]] = cv[e];
    } else if (tid == 0) {
      rho[cmap[q - n]] = 1.0;
    }
    sync();
    if (tid < WAVE) {
      const double ap = wave_rowdot(w, rho, m, tid);
      if (tid == 0) sh->tmp = ap;
    }
    sync();
    return sh->tmp;
  }

  // Basis change: FTRAN (when FT; otherwise aq holds alpha_q), DSE vector,
  // product-form update of B^{-1} in one pass, then the primal / phase-1 and
  // weight updates (sequentialDualSimplex.cpp:174-227).
  template <bool FT>
  __device__ __forceinline__ void change_basis(int phase, double ap_in) {
    const int p_idx = sh->p_idx, q = sh->q, q_idx = sh->q_idx, p = sh->p;
    if (!FT) {
      for (int cc = tid; cc < m; cc += BS) w[cc] = Binv[p_idx * m + cc];
      sync();
    }
    const double ap = FT ? ap_in : aq[p_idx];
    const double bp = beta[p_idx];
    for (int cc = tid; cc < m; cc += BS) prow[cc] = w[cc] / ap;
    sync();
    inv_update_fused<BS, FT>(Binv, w, prow, rho, ap, aq, tau, m, p_idx, tid);
    sync();
    const double thP = (phase == 2) ? sh->delta / ap : sh->f_p / ap;
    for (int i = tid; i < m; i += BS) {
      double ai = aq[i];
      if (phase == 2) {
        int j = basis[i];
        x[j] = x[j] - thP * ai;
      } else if (i != p_idx) {
        f[i] = f[i] - ai * thP;
      }
      if (i != p_idx) {
        double r = ai / ap;
        beta[i] = beta[i] - 2 * ai / ap * tau[i] + (r * r) * bp;
      }
    }
    sync();
    if (tid == 0) {
      beta[p_idx] = bp / (ap * ap);
      if (phase == 2) x[q] = x[q] + thP;
      else f[p_idx] = f[p_idx] / ap + sh->corr;
      basis[p_idx] = q;
      nonbasis[q_idx] = p;
      posN[q] = -1;
      posN[p] = q_idx;
    }
    sync();
  }

  // BTRAN (row of B^{-1}) and the pivot row over the nonbasic columns.
  __device__ __forceinline__ void pivot_row() {
    const int p_idx = sh->p_idx;
    for (int r = tid; r < m; r += BS) rho[r] = Binv[p_idx * m + cmap[r]];
    sync();
    for (int i = tid; i < n; i += BS) ar[i] = col_dot(nonbasis[i], rho, 1.0);
    sync();
  }

  // Entering column by the order-free textbook rule: minimum ratio, then the
  // largest |alpha| within EPS_Z of it, then the lowest position.
  __device__ __forceinline__ int chuzc(int phase) {
    const int neg = sh->neg;
    double mn = BIG;
    int cnt = 0;
    for (int i = tid; i < n; i += BS) {
      double a = neg ? -ar[i] : ar[i];
      int j = nonbasis[i];
      if (!candidate(phase, j, a)) continue;
      ++cnt;
      mn = fmin(mn, d[j] / a);
    }
    mn = block_min<BS>(mn);
    cnt = block_sumi<BS>(cnt);
    if (cnt == 0) return -1;
    double best = -1.0;
    int bi = INT_MAX;
    for (int i = tid; i < n; i += BS) {
      double a = neg ? -ar[i] : ar[i];
      int j = nonbasis[i];
      if (!candidate(phase, j, a)) continue;
      if (fabs(d[j] / a - mn) < EPS_Z) argmax_merge(best, bi, fabs(a), i);
    }
    block_argmax<BS>(best, bi);
    return bi == INT_MAX ? -1 : bi;
  }

  __device__ __forceinline__ bool candidate(int phase, int j, double a) const {
    if (phase == 2) {
      int8_t t = bt[j];
      if (a > EPS_A) return fabs(x[j] - l[j]) < EPS_Z && (t == BT_LOWER || t == BT_BOXED);
      if (a < -EPS_A) return fabs(x[j] - u[j]) < EPS_Z && (t == BT_UPPER || t == BT_BOXED);
      return t == BT_FREE;
    }
    double dj = d[j];
    return (dj >= 0 && a > EPS_A) || (dj <= 0 && a < -EPS_A);
  }

  // FTRAN of column q: alpha_q = B^{-1} a_q.
  __device__ __forceinline__ void ftran_q() {
    const int q = sh->q;
    if (q < n) {
      inv_spcol<BS>(Binv, cmap, ri + cp[q], cv + cp[q], cp[q + 1] - cp[q], aq, m, tid);
    } else {
      const int cq = cmap[q - n];
      for (int i = tid; i < m; i += BS) aq[i] = Binv[i * m + cq];
    }
    sync();
  }

  __device__ __forceinline__ bool tick() {
    if (tid == 0) sh->iteration += 1;
    sync();
    return !(max_iter > 0 && sh->iteration > max_iter);
  }

  __device__ __forceinline__ void clear_blocked() {
    if (tid == 0) sh->nblocked = 0;
    sync();
  }
  __device__ __forceinline__ bool is_blocked(int j) const {
    for (int b = 0; b < sh->nblocked; ++b)
      if (sh->blocked[b] == j) return true;
    return false;
  }

  // minimizeDualInfeasibility (tpp:799-934).  Returns 0 Solved, 1 DualInfeas,
  // 2 iteration limit, 3 singular basis.
  __device__ __forceinline__ int phase1() {
    phase1_init();
    if (tid == 0) sh->cycle = 0;
    clear_blocked();
    while (true) {
      if (!tick()) return 2;
      if (perturb_on && !sh->perturbed && sh->cycle > MAX_CYCLE) {
        perturb_costs();
        phase1_init();
      }
      if (sh->cycle > NEED_RESTART || sh->obj > BIG || isnan(sh->obj)) phase1_init();
      if (sh->iteration % K == 0 && !reinvert()) return 3;
      double bw = -1.0;
      int bi = INT_MAX;
      for (int i = tid; i < m; i += BS) {
        int j = basis[i];
        int8_t t = bt[j];
        double fi = f[i];
        if ((t == BT_LOWER && fi > EPS_F) || (t == BT_UPPER && fi < -EPS_F) || t == BT_BOXED || t == BT_FIXED) {
          double wgt = (fi * fi) / beta[i];
          if (wgt > 0.0 && !is_blocked(j)) argmax_merge(bw, bi, wgt, i);
        }
      }
      block_argmax<BS>(bw, bi);
      const bool found = bi != INT_MAX;
      const int dinf = sh->dinf;
      const double obj = sh->obj;
      if (!found || dinf == 0) {
        if (fabs(obj) < EPS_A || dinf == 0) return 0;
        if (obj > 0) return 1;
      }
      if (!found || obj > EPS_BOUND) {
        if (!reinvert()) return 3;
        phase1_init();
        init_beta(false);
        clear_blocked();
        continue;
      }
      if (tid == 0) {
        sh->p_idx = bi;
        sh->p = basis[bi];
        sh->f_p = f[bi];
        sh->neg = f[bi] > 0 ? 1 : 0;
      }
      sync();
      pivot_row();
      const int qi = chuzc(1);
      if (qi < 0) {
        if (tid == 0 && sh->nblocked < 64) sh->blocked[sh->nblocked++] = sh->p;
        sync();
        if (!reinvert()) return 3;
        phase1_init();
        continue;
      }
      clear_blocked();
      if (tid == 0) {
        int q = nonbasis[qi];
        sh->q_idx = qi;
        sh->q = q;
        sh->theta = d[q] / ar[qi];
        int8_t tq = bt[q];
        double dq = d[q];
        double corr = 0.0;
        if ((tq == BT_UPPER || tq == BT_FREE) && dq > EPS_D) corr = 1.0;
        else if ((tq == BT_LOWER || tq == BT_FREE) && dq < -EPS_D) corr = -1.0;
        sh->corr = corr;
      }
      sync();
      const double ap = pivot_alpha();
      if (fabs(ap - ar[sh->q_idx]) > EPS_R * (1 + fabs(ap))) {
        // Unstable pivot: alpha_q from the old inverse is kept across the refactor.
        ftran_q();
        if (!reinvert()) return 3;
        update_duals(1);
        change_basis<false>(1, 0.0);
      } else {
        update_duals(1);
        change_basis<true>(1, ap);
      }
      if (tid == 0) sh->cycle = (fabs(sh->theta) < EPS_A) ? sh->cycle + 1 : 0;
      sync();
    }
  }

  // d_N -= theta alpha_r; d_p = -theta; d_q = 0; objective update; recount of
  // dual infeasibilities over the next nonbasic set.
  __device__ __forceinline__ void update_duals(int phase) {
    const double theta = sh->theta;
    const int q_idx = sh->q_idx;
    int cnt = 0;
    for (int i = tid; i < n; i += BS) {
      int j = nonbasis[i];
      double dj = d[j] - theta * ar[i];
      d[j] = dj;
      if (i != q_idx) cnt += dual_infeas(bt[j], dj);
    }
    cnt = block_sumi<BS>(cnt);
    if (tid == 0) {
      d[sh->p] = -theta;
      d[sh->q] = 0.0;
      sh->dinf = cnt + dual_infeas(bt[sh->p], -theta);
      if (phase == 2) sh->obj = sh->obj + theta * sh->delta;
      else sh->obj = sh->obj - theta * sh->f_p;
    }
    sync();
  }

  __device__ __forceinline__ bool dual_feasible() {
    int cnt = 0;
    for (int i = tid; i < n; i += BS) {
      int j = nonbasis[i];
      cnt += dual_infeas(bt[j], d[j]);
    }
    cnt = block_sumi<BS>(cnt);
    return cnt == 0;
  }

  // callDualSolver (sequentialDualSimplex.cpp:253-515).  Returns the status.
  __device__ __forceinline__ int phase2() {
    dual_init();
    if (tid == 0) sh->cycle = 0;
    sync();
    int unbound_checked = 0;
    while (true) {
      if (!tick()) return DS_ITER;
      if (sh->iteration % K == 0 && !reinvert()) return DS_SING;
      if (sh->iteration % NEED_RESTART == 0 && !dual_feasible()) {
        if (!reinvert()) return DS_SING;
        int cyc = sh->cycle;
        int st = phase1();
        if (st >= 2) return st == 2 ? DS_ITER : DS_SING;
        dual_init();
        if (tid == 0) sh->cycle = cyc;
        sync();
      }
      if (perturb_on && !sh->perturbed && sh->cycle > MAX_CYCLE) {
        perturb_costs();
        if (!reinvert()) return DS_SING;
        dual_init();
      }
      double bw = -1.0;
      int bi = INT_MAX, pinf = 0;
      for (int i = tid; i < m; i += BS) {
        int j = basis[i];
        int8_t t = bt[j];
        double xv = x[j];
        pinf |= primal_infeas(t, xv, l[j], u[j]) ? 1 : 0;
        double dl;
        if (var_infeas(t, xv, l[j], u[j], dl) && fabs(dl) > DSE_FLOOR) {
          double wgt = (dl * dl) / beta[i];
          if (wgt > 0.0) argmax_merge(bw, bi, wgt, i);
        }
      }
      block_argmax<BS>(bw, bi);
      pinf = block_sumi<BS>(pinf);
      if (!pinf) return DS_OPT;
      if (bi == INT_MAX) {
        if (!reinvert()) return DS_SING;
        init_beta(false);
        dual_init();
        continue;
      }
      if (tid == 0) {
        int j = basis[bi];
        double dl;
        var_infeas(bt[j], x[j], l[j], u[j], dl);
        sh->p_idx = bi;
        sh->p = j;
        sh->delta = dl;
        sh->neg = dl < 0 ? 1 : 0;
      }
      sync();
      pivot_row();
      const int qi = chuzc(2);
      if (qi < 0) {
        if (unbound_checked) return DS_UNB;
        unbound_checked = 1;
        if (!reinvert()) return DS_SING;
        init_beta(false);
        int st = phase1();
        if (st >= 2) return st == 2 ? DS_ITER : DS_SING;
        dual_init();
        continue;
      }
      unbound_checked = 0;
      if (tid == 0) {
        int q = nonbasis[qi];
        sh->q_idx = qi;
        sh->q = q;
        sh->theta = d[q] / ar[qi];
        sh->tmp = (bt[q] == BT_BOXED && fabs(sh->delta) - (u[q] - l[q]) * fabs(ar[qi]) > 0) ? 1.0 : 0.0;
      }
      sync();
      if (sh->tmp != 0.0) return DS_BFRT;
      const double ap = pivot_alpha();
      if (fabs(ap - ar[sh->q_idx]) > EPS_R * (1 + fabs(ap))) {
        // Unstable pivot: alpha_q from the old inverse is kept across the refactor.
        ftran_q();
        if (!reinvert()) return DS_SING;
        update_duals(2);
        change_basis<false>(2, 0.0);
      } else {
        update_duals(2);
        change_basis<true>(2, ap);
      }
      if (tid == 0 && fabs(sh->theta) < EPS_A) sh->cycle = sh->cycle + 1;
      sync();
    }
  }

  static constexpr int DS_OPT = 0, DS_UNB = 1, DS_DINF = 2, DS_ITER = 3, DS_SING = 5, DS_BFRT = 6;
};

}  // namespace

size_t batch_lds_bytes(int mmax, int nmax, int inv_in_lds, int gj_nb) {
  const int N = mmax + nmax;
  // c, x, d, wN (N each); beta, f, rho, aq, tau, w, prow (m each); ar (n).
  // The sort buffer of checkPerturbNeed aliases wN..ar.
  size_t dbl = (size_t)4 * N + (size_t)7 * mmax + nmax;
  // factorization panel: Pn, Vb (m x nb each), Rb (nb x m)
  dbl += (size_t)3 * gj_nb * mmax;
  // basis (m), nonbasis (n), posN (N), piv, cmap, rpos, rflag, rat (m each).
  size_t ints = (size_t)mmax + nmax + N + 5 * (size_t)mmax;
  size_t bytes = dbl * 8 + ints * 4 + sizeof(BShared) + 64;
  if (inv_in_lds) bytes += (size_t)mmax * mmax * 8;
  return bytes;
}

int batch_panel_width(int mmax, int nmax, int inv_in_lds, size_t limit) {
  for (int nb = GJ_NB_MAX; nb >= 1; nb /= 2)
    if (batch_lds_bytes(mmax, nmax, inv_in_lds, nb) <= limit) return nb;
  return 0;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch(BatchDev D) {
  extern __shared__ double lds_d[];
  const int mmax = D.mmax, nmax = D.nmax, Nmax = mmax + nmax;
  double* base = lds_d;
  LpSolve<BS> S = {};
  S.tid = threadIdx.x;
  S.c = base;
  S.x = S.c + Nmax;
  S.d = S.x + Nmax;
  S.wN = S.d + Nmax;
  S.beta = S.wN + Nmax;
  S.f = S.beta + mmax;
  S.rho = S.f + mmax;
  S.aq = S.rho + mmax;
  S.tau = S.aq + mmax;
  S.w = S.tau + mmax;
  S.prow = S.w + mmax;
  S.ar = S.prow + mmax;
  // checkPerturbNeed runs once per LP before wN, beta, f, rho, aq, tau, w,
  // prow and ar hold anything; its sort buffer (a power of two below
  // 2(m+n) <= m+n + 7m + n doubles) reuses that region.
  S.sortbuf = S.wN;
  double* after = S.ar + nmax;
  if (D.inv_in_lds) {
    S.Binv = after;
    after += (size_t)mmax * mmax;
  } else {
    S.Binv = D.gscratch + (size_t)blockIdx.x * mmax * mmax;
  }
  S.gj_nb = D.gj_nb;
  S.Pn = after;
  S.Vb = S.Pn + (size_t)D.gj_nb * mmax;
  S.Rb = S.Vb + (size_t)D.gj_nb * mmax;
  after = S.Rb + (size_t)D.gj_nb * mmax;
  int* ip = (int*)after;
  S.basis = ip;
  S.nonbasis = S.basis + mmax;
  S.posN = S.nonbasis + nmax;
  S.piv = S.posN + Nmax;
  S.cmap = S.piv + mmax;
  S.rpos = S.cmap + mmax;
  S.rflag = S.rpos + mmax;
  S.rat = S.rflag + mmax;
  S.sh = (BShared*)(((uintptr_t)(S.rat + mmax) + 15) & ~(uintptr_t)15);
  S.K = D.refactor_freq;
  S.repair_tol = D.repair_tol;
  S.repair_on = 0;
  S.max_iter = D.max_iter;
  S.perturb_on = D.perturb;
  S.seed = D.seed;
  __shared__ int s_item;
  while (true) {
    if (threadIdx.x == 0) s_item = atomicAdd(D.next, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= D.order_count) break;
    const int lp = D.order[item];
    const int m = D.m[lp], n = D.n[lp], N = m + n;
    S.m = m;
    S.n = n;
    S.N = N;
    S.cp = D.colptr + D.cp_off[lp];
    S.ri = D.rowidx + D.nz_off[lp];
    S.cv = D.cval + D.nz_off[lp];
    S.rp = D.rowptr + D.rp_off[lp];
    S.ci = D.colidx + D.nz_off[lp];
    S.rv = D.rval + D.nz_off[lp];
    S.corig = D.c + D.vo[lp];
    S.l = D.l + D.vo[lp];
    S.u = D.u + D.vo[lp];
    S.bt = D.bt + D.vo[lp];
    // Slack basis, or the given starting basis.
    for (int j = threadIdx.x; j < N; j += BS) {
      S.c[j] = S.corig[j];
      S.x[j] = 0.0;
      S.d[j] = 0.0;
      S.posN[j] = -1;
    }
    if (threadIdx.x == 0) {
      BShared* sh = S.sh;
      sh->status = ST_OK;
      sh->iteration = 0;
      sh->perturbed = 0;
      sh->cycle = 0;
      sh->nblocked = 0;
      sh->nrep = 0;
      sh->sing = 0;
      sh->obj = 0.0;
      sh->rng = D.seed;
      if (D.warm) {
        const int* wb = D.warm + D.mo[lp];
        for (int i = 0; i < m; ++i) S.basis[i] = wb[i];
      } else {
        for (int i = 0; i < m; ++i) S.basis[i] = n + i;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += BS) S.posN[S.basis[i]] = -2;
    __syncthreads();
    if (threadIdx.x == 0) {
      int k = 0;
      for (int j = 0; j < N; ++j)
        if (S.posN[j] == -1) {
          S.nonbasis[k] = j;
          S.posN[j] = k++;
        }
      for (int j = 0; j < N; ++j)
        if (S.posN[j] == -2) S.posN[j] = -1;
    }
    __syncthreads();
    int status;
    // A given starting basis may be repaired by slack substitution in its
    // first factorization; its DSE weights are the exact row norms.
    // Without repair a given start is singular once a pivot is numerically
    // zero (the blocked elimination does not cancel exactly).
    S.repair_on = (D.warm && D.repair_tol > 0.0) ? 1 : 0;
    S.sing_tol = (D.warm && !S.repair_on) ? 1e-12 : 0.0;
    bool ok = S.reinvert();
    S.repair_on = 0;
    S.sing_tol = 0.0;
    if (!ok) {
      status = 5;
    } else {
      if (D.perturb && S.perturb_need()) S.perturb_costs();
      S.init_beta(D.warm == nullptr);
      int st = S.phase1();
      if (st == 1) status = 2;
      else if (st == 2) status = 3;
      else if (st == 3) status = 5;
      else {
        S.set_primal();
        status = S.phase2();
        if (S.sh->perturbed && (status == 0 || status == 1)) {
          // postProcess (tpp:123-159): original costs, primal values and duals
          // recomputed.  The reference then calls a primal correction that is
          // a stub (sequentialDualSimplex.cpp:113-116), leaving a basis that
          // may be dual infeasible for the original costs.  With `cleanup` the
          // dual simplex resumes from that basis instead (the perturbation
          // flag stays set, so no second perturbation).
          for (int j = threadIdx.x; j < N; j += BS) S.c[j] = S.corig[j];
          __syncthreads();
          S.dual_init();
          if (D.cleanup && status == 0 && !S.dual_feasible()) {
            const int st2 = S.phase1();
            if (st2 == 1) status = 2;
            else if (st2 == 2) status = 3;
            else if (st2 == 3) status = 5;
            else {
              S.set_primal();
              status = S.phase2();
            }
          }
        }
      }
    }
    // Results.
    if (threadIdx.x == 0) {
      D.status[lp] = status;
      D.obj[lp] = S.sh->obj;
      D.iters[lp] = S.sh->iteration;
      if (D.repairs) D.repairs[lp] = S.sh->nrep;
    }
    for (int i = threadIdx.x; i < m; i += BS) D.basis_out[D.mo[lp] + i] = S.basis[i];
    for (int j = threadIdx.x; j < N; j += BS) D.x_out[D.vo[lp] + j] = S.x[j];
    __syncthreads();
  }
}

__global__ void k_batch_reset(int* next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *next = 0;
}

void launch_batch(hipStream_t s, const BatchDev& D, int grid, int block) {
  size_t lds = batch_lds_bytes(D.mmax, D.nmax, D.inv_in_lds, D.gj_nb);
  switch (block) {
    case 128: k_batch<128><<<grid, 128, lds, s>>>(D); break;
    case 512: k_batch<512><<<grid, 512, lds, s>>>(D); break;
    default: k_batch<256><<<grid, 256, lds, s>>>(D); break;
  }
}

void launch_batch_reset(hipStream_t s, int* next) { k_batch_reset<<<1, 64, 0, s>>>(next); }

}  // namespace pds
