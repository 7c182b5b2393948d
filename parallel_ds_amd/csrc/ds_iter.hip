// gfx950 kernels of one dual simplex iteration and of its (re)initialisation.
//
// Every kernel reads the device status word at entry and returns when an
// earlier kernel of the queued chunk raised an event; the host then handles
// the event and resumes.  Kernels that reduce across workgroups finish in the
// last-arriving workgroup (ds_reduce.hpp), which alone may change the status.
#include "ds_dev.hpp"
#include "ds_kernels.hpp"
#include "ds_reduce.hpp"
#include "ds_rules.hpp"

#include <hip/hip_ext.h>

namespace pds {

constexpr int BS = 256;

__device__ inline bool stopped(const DevScalars* S) { return S->status != ST_RUNNING || S->refac_mid != 0; }

// Launch guard of the queued kernels: check 0 = always run, 1 = not after an
// event, 2 = also not when the iteration has no bound flips (the BFRT FTRAN).
__device__ inline bool skip_launch(const DevScalars* S, int check) {
  return check && (stopped(S) || (check == 2 && S->nflip == 0));
}

__device__ inline void raise_status(DevScalars* S, int st, int it) {
  S->status = st;
  S->stop_iter = it;
}

// ---------------------------------------------------------------------------
// CHUZR: DSE pricing (tpp:470-501) in phase 2, phase1Pricing (tpp:746-772) in
// phase 1, fused with the primal feasibility test of callDualSolver
// (sequentialDualSimplex.cpp:305-310) and the phase-1 termination tests
// (tpp:841-875).  Argmax of delta^2/beta, ties to the lowest position.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void btran_prep_body(const DevModel& M, const DevIter& I, const DevFactor& F, int t,
                                                int p_fixed, int cand);

// BTRAN = true: the last workgroup goes on with BTRAN's set-up for the chosen
// row (k_btran_prep's work, one workgroup), so the iteration needs one launch
// fewer; the launch then carries BTRAN's dynamic LDS.
template <bool BTRAN>
__global__ __launch_bounds__(BS) void k_price(DevModel M, DevIter I, int phase, int it, int skip_top,
                                              int perturb_on, DevFactor F, int t) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  if (!skip_top) {
    int st = ST_RUNNING;
    if (perturb_on && !S->perturbed && S->cycle > MAX_CYCLE) st = ST_NEED_PERTURB;
    else if (phase == 1 && (S->cycle > NEED_RESTART || S->obj > BIG || isnan(S->obj))) st = ST_PH1_REINIT;
    if (st != ST_RUNNING) {
      if (blockIdx.x == 0 && threadIdx.x == 0) raise_status(S, st, it);
      return;
    }
  }
  double bw = -1.0;
  int bi = INT_MAX;
  int pinf = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    int j = I.basis[i];
    int8_t t = M.bt[j];
    if (phase == 2) {
      double x = I.x[j], l = M.l[j], u = M.u[j];
      pinf |= primal_infeas(t, x, l, u) ? 1 : 0;
      double dl;
      if (var_infeas(t, x, l, u, dl) && fabs(dl) > DSE_FLOOR && !I.blocked[j]) {
        double w = (dl * dl) / I.beta[i];
        if (w > 0.0) argmax_merge(bw, bi, w, i);
      }
    } else {
      double fi = I.f[i];
      if ((t == BT_LOWER && fi > EPS_F) || (t == BT_UPPER && fi < -EPS_F) || t == BT_BOXED || t == BT_FIXED) {
        double w = (fi * fi) / I.beta[i];
        if (w > 0.0 && !I.blocked[j]) argmax_merge(bw, bi, w, i);
      }
    }
  }
  block_argmax<BS>(bw, bi);
  pinf = block_sumi<BS>(pinf);
  if (threadIdx.x == 0) {
    I.part_d[blockIdx.x] = bw;
    I.part_i[2 * blockIdx.x] = bi;
    I.part_i[2 * blockIdx.x + 1] = pinf;
  }
  if (!last_block(&S->ctr[CTR_PRICE])) return;
  double fw = -1.0;
  int fi = INT_MAX, fp = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) {
    argmax_merge(fw, fi, I.part_d[b], I.part_i[2 * b]);
    fp += I.part_i[2 * b + 1];
  }
  block_argmax<BS>(fw, fi);
  fp = block_sumi<BS>(fp);
  __shared__ int s_row;  // the chosen row, or -1 when an event was raised
  if (threadIdx.x == 0) {
    s_row = -1;
    S->ctr[CTR_PRICE] = 0;
    bool found = fi != INT_MAX;
    if (phase == 2) {
      if (!fp) {
        raise_status(S, ST_OPTIMAL, it);
      } else if (!found) {
        raise_status(S, ST_NO_PIVOT, it);
      } else {
        int j = I.basis[fi];
        double dl;
        var_infeas(M.bt[j], I.x[j], M.l[j], M.u[j], dl);
        S->p_idx = fi;
        S->p = j;
        S->delta = dl;
        S->is_lower = dl < 0 ? 1 : 0;
        S->neg = S->is_lower;
        s_row = fi;
      }
    } else {
      int dinf = S->dinf;
      double obj = S->obj;
      int st = ST_RUNNING;
      if (!found || dinf == 0) {
        if (fabs(obj) < EPS_A || dinf == 0) st = ST_PH1_SOLVED;
        else if (obj > 0) st = ST_PH1_DUALINF;
      }
      if (st == ST_RUNNING && (!found || obj > EPS_BOUND)) st = ST_PH1_RESTART;
      if (st != ST_RUNNING) {
        raise_status(S, st, it);
      } else {
        S->p_idx = fi;
        S->p = I.basis[fi];
        S->f_p = I.f[fi];
        S->neg = I.f[fi] > 0 ? 1 : 0;
        s_row = fi;
      }
    }
  }
  if constexpr (BTRAN) {
    __syncthreads();
    if (s_row >= 0) btran_prep_body(M, I, F, t, s_row, -1);
  }
}

// Periodic dual feasibility check of phase 2 (sequentialDualSimplex.cpp:285-293).
__global__ __launch_bounds__(BS) void k_dualcheck(DevModel M, DevIter I, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  int cnt = 0;
  int nN = M.N - M.m;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    int j = I.nonbasis[i];
    cnt += dual_infeas(M.bt[j], I.d[j]);
  }
  cnt = block_sumi<BS>(cnt);
  if (threadIdx.x == 0) I.part_i[blockIdx.x] = cnt;
  if (!last_block(&S->ctr[CTR_DCHK])) return;
  int tot = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) tot += I.part_i[b];
  tot = block_sumi<BS>(tot);
  if (threadIdx.x == 0) {
    S->ctr[CTR_DCHK] = 0;
    S->dinf = tot;
    if (tot > 0) raise_status(S, ST_DUAL_RESTART, it);
  }
}

// ---------------------------------------------------------------------------
// BTRAN for a unit vector: rho = B0^{-T} E_1^T ... E_t^T e_p
// (sequentialDualSimplex.cpp:89-92, linalg.cpp:34-44).  The transposed etas
// only write the pivot positions; their values u solve a unit triangular
// system whose inverse MF is maintained by the basis update, so u = MF^T b is
// one parallel matvec (b_k = eta_k[p] while position p still holds its unit
// entry).  The kernel-inverse gather then runs on all rows in k_btran_out.
// ---------------------------------------------------------------------------
// MF's lower triangle, packed by rows (row r from r(r+1)/2 on), goes to
// dynamic LDS when the eta count is at most MF_TRI_MAX (125 KB at 176).
constexpr int MF_TRI_MAX = 176;
constexpr int MF_SEQ_MAX = 64;  // up to this eta count BTRAN sums MF's columns in one chain, beyond in four
__host__ __device__ constexpr int tri(int r) { return r * (r + 1) / 2; }
static size_t mf_lds_bytes(int t) { return t > 0 && t <= MF_TRI_MAX ? sizeof(double) * tri(t) : 0; }

// Every thread of the workgroup: rows 0..t-1 of MF's lower triangle into mft,
// eight independent loads in flight per thread.
__device__ __forceinline__ void load_mf_tri(const DevIter& I, int t, double* mft) {
  constexpr int U = 8;
  const int n = tri(t), cap = I.eta_cap;
  for (int e0 = threadIdx.x; e0 < n; e0 += BS * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * BS;
      v[u] = 0.0;
      if (e < n) {
        int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
        while (tri(r + 1) <= e) ++r;
        while (tri(r) > e) --r;
        v[u] = I.MF[(size_t)r * cap + (e - tri(r))];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (e0 + u * BS < n) mft[e0 + u * BS] = v[u];
  }
}

__device__ __forceinline__ void btran_prep_body(const DevModel& M, const DevIter& I, const DevFactor& F, int t,
                                                int p_fixed, int cand) {
  DevScalars* S = I.S;
  if (p_fixed < 0 && stopped(S)) return;
  if (cand >= 0 && cand >= S->ncand) return;
  const int p = cand >= 0 ? I.cand_pos[cand] : (p_fixed >= 0 ? p_fixed : S->p_idx);
  const int m = M.m;
  const int cap = I.eta_cap;
  __shared__ int pos[ETA_MAX + 1];
  __shared__ int slot[ETA_MAX + 1];
  __shared__ double v[ETA_MAX + 1];
  __shared__ double bv[ETA_MAX];
  __shared__ double u[ETA_MAX];
  __shared__ int cnt[ETA_MAX + 1];
  __shared__ int off[ETA_MAX + 2];
  __shared__ int pidx[ETA_MAX + 1];  // kernel column of the position, -1 at a slack position
  __shared__ int srow[ETA_MAX + 1];  // slack row of a slack position
  __shared__ int ccb[ETA_MAX + 1], cce[ETA_MAX + 1];  // its coupling-row extent
  extern __shared__ double mft[];  // MF's lower triangle (t <= MF_TRI_MAX)
  __shared__ int kmax_p, kmin_p;
  const int ns = t + 1;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  const bool tile = t <= MF_TRI_MAX;
  if (tid == 0) {
    kmax_p = -1;
    kmin_p = INT_MAX;
  }
  for (int s = tid; s < ns; s += BS) pos[s] = (s == 0) ? p : I.eta_pos[s - 1];
  // MF's lower triangle (the part the sweep reads) into LDS when it fits.
  if (tile) load_mf_tri(I, t, mft);
  __syncthreads();
  for (int k = tid; k < t; k += BS)
    if (pos[1 + k] == p) {
      atomicMax(&kmax_p, k);
      atomicMin(&kmin_p, k);
    }
  for (int s = tid; s < ns; s += BS) {
    const int ps = pos[s];
    const int c = F.Pidx_of_pos[ps];
    pidx[s] = c;
    if (c < 0) {
      const int r = F.slackrow_of_pos[ps];
      srow[s] = r;
      ccb[s] = F.crp[r];
      cce[s] = F.crp[r + 1];
    }
  }
  // slot[s] = the first s2 <= s at the same position (one s per wave at a time).
  for (int s = wv; s < ns; s += BS / WAVE) {
    int f = s;
    for (int b0 = 0; b0 < s; b0 += WAVE) {
      const int s2 = b0 + lane;
      const unsigned long long bal = __ballot(s2 < s && pos[s2] == pos[s]);
      if (bal) {
        f = b0 + __ffsll((long long)bal) - 1;
        break;
      }
    }
    if (lane == 0) slot[s] = f;
  }
  __syncthreads();
  for (int k = tid; k < t; k += BS) bv[k] = (k >= kmax_p) ? I.etas[(size_t)k * m + p] : 0.0;
  __syncthreads();
  for (int k = tid; k < t; k += BS) {
    double acc = 0.0;
    if (t <= MF_SEQ_MAX) {
      for (int k2 = k; k2 < t; ++k2) acc += mft[tri(k2) + k] * bv[k2];
    } else {
      // four independent partial sums (from LDS or, beyond MF_TRI_MAX, with
      // four MF loads in flight)
      double a4[4] = {0.0, 0.0, 0.0, 0.0};
      int k2 = k;
      if (tile) {
        for (; k2 + 3 < t; k2 += 4) {
#pragma unroll
          for (int q = 0; q < 4; ++q) a4[q] += mft[tri(k2 + q) + k] * bv[k2 + q];
        }
        for (; k2 < t; ++k2) a4[0] += mft[tri(k2) + k] * bv[k2];
      } else {
        for (; k2 + 3 < t; k2 += 4) {
#pragma unroll
          for (int q = 0; q < 4; ++q) a4[q] += I.MF[(size_t)(k2 + q) * cap + k] * bv[k2 + q];
        }
        for (; k2 < t; ++k2) a4[0] += I.MF[(size_t)k2 * cap + k] * bv[k2];
      }
      acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
    }
    u[k] = acc;
  }
  __syncthreads();
  // Final values at the support: each position keeps the last assignment of
  // the reverse sweep, i.e. that of its smallest eta index.
  for (int s = tid; s < ns; s += BS) {
    double val = 0.0;
    if (slot[s] == s) {
      if (s == 0) val = (kmax_p < 0) ? 1.0 : u[kmin_p];
      else val = u[s - 1];
    }
    v[s] = val;
  }
  __syncthreads();
  // Right-hand side of the kernel solve: rhs_P = v_P - C^T y_S, with
  // y_S = v at slack positions (ordered by slot, prefix-summed offsets).
  for (int s = tid; s < ns; s += BS) {
    int c = 0;
    if (slot[s] == s && v[s] != 0.0) c = pidx[s] >= 0 ? 1 : 1 + (cce[s] - ccb[s]);
    cnt[s] = c;
  }
  __syncthreads();
  if (wv == 0) {  // exclusive prefix sum of cnt
    int carry = 0;
    for (int b0 = 0; b0 < ns; b0 += WAVE) {
      const int s = b0 + lane;
      const int c = s < ns ? cnt[s] : 0;
      int x = c;
      for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
      }
      if (s < ns) off[s] = carry + x - c;
      carry += __shfl(x, WAVE - 1, WAVE);
    }
    if (lane == 0) off[ns] = carry;
  }
  __syncthreads();
  const int total = off[ns];
  if (total > I.list_cap) {
    if (tid == 0) raise_status(S, ST_BTRAN_CAP, -1);
    return;
  }
  for (int s = tid; s < ns; s += BS) {
    if (cnt[s] == 0) continue;
    const double vs = v[s];
    int o = off[s];
    const int c = pidx[s];
    if (c >= 0) {
      I.list_idx[o] = c;
      I.list_val[o] = vs;
    } else {
      const int r = srow[s];
      I.list_idx[o] = -1 - r;   // marks a slack row entry (no kernel contribution)
      I.list_val[o] = vs;
      for (int e = ccb[s]; e < cce[s]; ++e) {
        ++o;
        I.list_idx[o] = F.cci[e];
        I.list_val[o] = -F.ccv[e] * vs;
      }
    }
  }
  if (wv == 0) {  // slack positions in order, compacted by ballot
    int a = 0;
    for (int b0 = 0; b0 < ns; b0 += WAVE) {
      const int s = b0 + lane;
      const bool take = s < ns && cnt[s] != 0 && pidx[s] < 0;
      const unsigned long long bal = __ballot(take);
      if (take) {
        const int idx = a + __popcll(bal & ((1ull << lane) - 1ull));
        I.ylist_row[idx] = srow[s];
        I.ylist_val[idx] = v[s];
      }
      a += __popcll(bal);
    }
    if (lane == 0) {
      I.list_n[0] = total;
      I.list_n[1] = a;
    }
  }
}

__global__ __launch_bounds__(BS) void k_btran_prep(DevModel M, DevIter I, DevFactor F, int t, int p_fixed,
                                                   int cand) {
  btran_prep_body(M, I, F, t, p_fixed, cand);
}

// A workgroup takes BO_ROWS consecutive rows; its BS / BO_ROWS row groups
// split the right-hand-side list between them (each a contiguous quarter), so
// the serial walk over the list per thread is a quarter as long, and the
// partial sums meet in LDS.
constexpr int BO_ROWS = 64;
__global__ __launch_bounds__(BS) void k_btran_out(DevModel M, DevIter I, DevFactor F, double* out, int check,
                                                  int cand) {
  DevScalars* S = I.S;
  if (skip_launch(S, check)) return;
  if (cand >= 0 && cand >= S->ncand) return;
  constexpr int NG = BS / BO_ROWS;
  constexpr int CH = 1024;
  __shared__ int sidx[CH];
  __shared__ double sval[CH];
  __shared__ int yrow[ETA_MAX + 1];
  __shared__ double yval[ETA_MAX + 1];
  __shared__ double part[NG][BO_ROWS];
  const int total = I.list_n[0];
  const int ny = I.list_n[1];
  for (int s = threadIdx.x; s < ny; s += BS) {
    yrow[s] = I.ylist_row[s];
    yval[s] = I.ylist_val[s];
  }
  const int g = threadIdx.x / BO_ROWS, lr = threadIdx.x % BO_ROWS;
  const int r = blockIdx.x * BO_ROWS + lr;
  const bool live = r < M.m;
  const int ri = live ? F.Ridx_of_row[r] : -1;
  double acc = 0.0;
  for (int base = 0; base < total; base += CH) {
    const int nchunk = min(CH, total - base);
    __syncthreads();
    for (int e = threadIdx.x; e < nchunk; e += BS) {
      sidx[e] = I.list_idx[base + e];
      sval[e] = I.list_val[base + e];
    }
    __syncthreads();
    const int per = (nchunk + NG - 1) / NG;
    const int e0 = g * per, e1 = min(nchunk, e0 + per);
    if (live && ri >= 0) {
      const double* col = F.Kinv + ri;
      for (int e = e0; e < e1; ++e) {
        const int c = sidx[e];
        if (c >= 0) acc += sval[e] * col[(size_t)c * F.k];
      }
    }
  }
  part[g][lr] = acc;
  __syncthreads();
  if (g != 0 || !live) return;
  double sum = part[0][lr];
#pragma unroll
  for (int q = 1; q < NG; ++q) sum += part[q][lr];
  if (ri < 0) {
    sum = 0.0;
    for (int s2 = 0; s2 < ny; ++s2)
      if (yrow[s2] == r) sum = yval[s2];
  }
  out[r] = sum;
}

// ---------------------------------------------------------------------------
// Pivot-row SpMV: Matrix::dotUpdate UPDATE_T (matrix.cpp:273-287) over the
// CSC of [A | I] (= CSR of A^T).  One 256-thread workgroup per SPMV_COLS
// consecutive columns: the tile's column pointers and nonbasic positions go
// to LDS, the workgroup streams the tile's contiguous (row, value) segment
// with SPMV_ILP independent loads in flight per thread and writes the
// products to LDS, and one lane per column then sums its column in CSC order,
// so the result is bitwise that of the sequential loop.  Basic columns are
// skipped at the store; so is everything when the solver has stopped (the
// check sits after the loads, off their dependency chain).  out index =
// nonbasic position, or the column id when by_col (the `set` flag).
// ---------------------------------------------------------------------------
constexpr int SPMV_THREADS = 256;
constexpr int SPMV_COLS_MAX = 512;
constexpr int SPMV_SEG = 2048;
constexpr int SPMV_ILP = 4;
constexpr int SPMV_GRID = 1024;  // tile width: the fewest columns that keep the grid within this
// bid: the tile (workgroup) index within the SpMV's share of the launch.
__device__ __forceinline__ void spmv_body(const DevModel& M, const int* __restrict__ posN, const double* __restrict__ v,
                                          double alpha, double beta, const double* __restrict__ cvec,
                                          double* __restrict__ out, int by_col, const DevScalars* S, int check,
                                          int cols, int bid) {
  __shared__ int cp[SPMV_COLS_MAX + 1];
  __shared__ int pn[SPMV_COLS_MAX];
  __shared__ double cb[SPMV_COLS_MAX];
  __shared__ double prod[SPMV_SEG];
  __shared__ int halt;
  const int tid = threadIdx.x;
  const int j0 = bid * cols;
  const int ncol = min(cols, M.N - j0);
  // Everything that does not depend on the entries is fetched in one round.
  for (int c = tid; c <= ncol; c += SPMV_THREADS) cp[c] = M.cptr[j0 + c];
  for (int c = tid; c < ncol; c += SPMV_THREADS) {
    pn[c] = posN[j0 + c];
    cb[c] = cvec[j0 + c];
  }
  if (tid == 0) halt = check && stopped(S);
  __syncthreads();
  const int e0 = cp[0], e1 = cp[ncol];
  const bool staged = e1 - e0 <= SPMV_SEG;
  if (staged) {
    for (int base = e0; base < e1; base += SPMV_ILP * SPMV_THREADS) {
      int ri[SPMV_ILP];
      double cv[SPMV_ILP];
#pragma unroll
      for (int u = 0; u < SPMV_ILP; ++u) {
        const int e = base + u * SPMV_THREADS + tid;
        ri[u] = e < e1 ? M.ridx[e] : 0;
        cv[u] = e < e1 ? M.cval[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < SPMV_ILP; ++u) {
        const int e = base + u * SPMV_THREADS + tid;
        if (e < e1) prod[e - e0] = cv[u] * v[ri[u]] * alpha;
      }
    }
  }
  __syncthreads();
  if (halt) return;
  for (int c = tid; c < ncol; c += SPMV_THREADS) {
    if (pn[c] < 0) continue;
    const int eb = cp[c], ee = cp[c + 1];
    double acc = 0.0;
    if (staged) {
      for (int e = eb; e < ee; ++e) acc = acc + prod[e - e0];
    } else {
      for (int e = eb; e < ee; ++e) acc = acc + M.cval[e] * v[M.ridx[e]] * alpha;
    }
    out[by_col ? j0 + c : pn[c]] = acc + cb[c] * beta;
  }
}

__global__ __launch_bounds__(SPMV_THREADS) void k_spmv(DevModel M, const int* __restrict__ posN,
                                                       const double* __restrict__ v, double alpha, double beta,
                                                       const double* __restrict__ cvec, double* __restrict__ out,
                                                       int by_col, const DevScalars* S, int check, int cols) {
  spmv_body(M, posN, v, alpha, beta, cvec, out, by_col, S, check, cols, blockIdx.x);
}

// Tile width: the smallest power of two in [64, SPMV_COLS_MAX] whose grid has
// at most SPMV_GRID workgroups (about four per CU).
static int spmv_cols(const DevModel& M) {
  int c = 64;
  while (c < SPMV_COLS_MAX && (M.N + c - 1) / c > SPMV_GRID) c *= 2;
  return c;
}

// ---------------------------------------------------------------------------
// CHUZC.  Pass 1 bounds the step: textbook = min d_j/alpha_j over the
// candidate set (sequentialDualSimplex.cpp:342-367, tpp:531-542, tpp:546-572);
// Harris = min of the EPS_D-relaxed ratios.  Pass 2 picks, among candidates
// within the bound, the largest |alpha_j| (lowest position on ties).
// ---------------------------------------------------------------------------
__device__ inline bool ratio_candidate(const DevModel& M, const DevIter& I, int phase, int j, double a) {
  if (phase == 2) {
    int8_t t = M.bt[j];
    if (a > EPS_A) return fabs(I.x[j] - M.l[j]) < EPS_Z && (t == BT_LOWER || t == BT_BOXED);
    if (a < -EPS_A) return fabs(I.x[j] - M.u[j]) < EPS_Z && (t == BT_UPPER || t == BT_BOXED);
    return t == BT_FREE;
  }
  double dj = I.d[j];
  return (dj >= 0 && a > EPS_A) || (dj <= 0 && a < -EPS_A);
}

// Shared tail of CHUZC once the entering position fi is known.
__device__ inline void chuzc_finish(const DevModel& M, const DevIter& I, int phase, int it, int fi) {
  DevScalars* S = I.S;
  if (fi == INT_MAX) {
    raise_status(S, ST_NO_RATIO, it);
    return;
  }
  const int q = I.nonbasis[fi];
  const double arq = I.alpha_r[fi];
  S->q_idx = fi;
  S->q = q;
  S->alpha_rq = arq;
  S->theta = I.d[q] / arq;
  if (phase == 2) {
    S->unbound_checked = 0;
    S->bfrt = 0;
    if (M.bt[q] == BT_BOXED) {
      // The bound-flipping loop (k_bfrt) continues from this candidate.
      double span = M.u[q] - M.l[q];
      if (fabs(S->delta) - span * fabs(arq) > 0) S->bfrt = 1;
    }
  } else {
    int8_t tq = M.bt[q];
    double dq = I.d[q];
    double corr = 0.0;
    if ((tq == BT_UPPER || tq == BT_FREE) && dq > EPS_D) corr = 1.0;
    else if ((tq == BT_LOWER || tq == BT_FREE) && dq < -EPS_D) corr = -1.0;
    S->corr = corr;
    if (S->any_blocked) {
      for (int b = 0; b < S->nblocked; ++b) I.blocked[I.blocked_list[b]] = 0;
      S->nblocked = 0;
      S->any_blocked = 0;
    }
  }
}

// Textbook ratio test in one launch.  Each workgroup keeps the candidates
// within EPS_Z of its own minimum ratio (every candidate within EPS_Z of the
// global minimum is among them); the last workgroup merges the lists.  A
// workgroup whose list overflowed and may hold global candidates makes the
// last workgroup rescan all positions.
constexpr int TB_LIST = 8;
// bid, nb: this workgroup and the workgroup count of the CHUZC share of the launch.
__device__ __forceinline__ void chuzc_tb_body(const DevModel& M, const DevIter& I, int phase, int it, int bid, int nb) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int neg = S->neg;
  const int nN = M.N - M.m;
  double mn = BIG;
  int nF = 0;
  for (int i = bid * BS + threadIdx.x; i < nN; i += nb * BS) {
    double a = I.alpha_r[i];
    if (neg) a = -a;
    int j = I.nonbasis[i];
    if (!ratio_candidate(M, I, phase, j, a)) continue;
    ++nF;
    mn = fmin(mn, I.d[j] / a);
  }
  mn = block_min<BS>(mn);
  nF = block_sumi<BS>(nF);
  __shared__ int lc;
  __shared__ int li[TB_LIST];
  __shared__ double lt[TB_LIST], la[TB_LIST];
  if (threadIdx.x == 0) lc = 0;
  __syncthreads();
  for (int i = bid * BS + threadIdx.x; i < nN; i += nb * BS) {
    double a = I.alpha_r[i];
    if (neg) a = -a;
    int j = I.nonbasis[i];
    if (!ratio_candidate(M, I, phase, j, a)) continue;
    double th = I.d[j] / a;
    if (th < mn + EPS_Z) {
      int slot = atomicAdd(&lc, 1);
      if (slot < TB_LIST) {
        li[slot] = i;
        lt[slot] = th;
        la[slot] = fabs(a);
      }
    }
  }
  __syncthreads();
  const int b = bid;
  if (threadIdx.x == 0) {
    I.part_d[b] = mn;
    I.part_i[2 * b] = nF;
    I.part_i[2 * b + 1] = lc;
  }
  for (int e = threadIdx.x; e < TB_LIST; e += BS)
    if (e < lc) {
      I.cand_i[b * TB_LIST + e] = li[e];
      I.cand_th[b * TB_LIST + e] = lt[e];
      I.cand_a[b * TB_LIST + e] = la[e];
    }
  if (!last_block(&S->ctr[CTR_CHUZC1], nb)) return;
  double g = BIG;
  int tot = 0;
  for (int w = threadIdx.x; w < nb; w += BS) {
    g = fmin(g, I.part_d[w]);
    tot += I.part_i[2 * w];
  }
  g = block_min<BS>(g);
  tot = block_sumi<BS>(tot);
  int rescan = 0;
  double best = -1.0;
  int bi = INT_MAX;
  for (int w = threadIdx.x; w < nb; w += BS) {
    int n = I.part_i[2 * w + 1];
    if (n > TB_LIST && I.part_d[w] < g + EPS_Z) rescan = 1;
    n = n < TB_LIST ? n : TB_LIST;
    for (int e = 0; e < n; ++e)
      if (fabs(I.cand_th[w * TB_LIST + e] - g) < EPS_Z) argmax_merge(best, bi, I.cand_a[w * TB_LIST + e], I.cand_i[w * TB_LIST + e]);
  }
  rescan = block_sumi<BS>(rescan);
  if (rescan) {
    best = -1.0;
    bi = INT_MAX;
    for (int i = threadIdx.x; i < nN; i += BS) {
      double a = I.alpha_r[i];
      if (neg) a = -a;
      int j = I.nonbasis[i];
      if (!ratio_candidate(M, I, phase, j, a)) continue;
      if (fabs(I.d[j] / a - g) < EPS_Z) argmax_merge(best, bi, fabs(a), i);
    }
  }
  block_argmax<BS>(best, bi);
  if (threadIdx.x != 0) return;
  S->ctr[CTR_CHUZC1] = 0;
  S->nF = tot;
  S->tmp0 = g;
  if (tot == 0) {
    raise_status(S, ST_NO_RATIO, it);
    return;
  }
  chuzc_finish(M, I, phase, it, bi);
}

__global__ __launch_bounds__(BS) void k_chuzc_tb(DevModel M, DevIter I, int phase, int it) {
  chuzc_tb_body(M, I, phase, it, blockIdx.x, gridDim.x);
}

// bid, nb: this workgroup and the number taking part (all of k_chuzc1's grid,
// or the CHUZC share of a fused stage launch).
__device__ __forceinline__ void chuzc1_body(const DevModel& M, const DevIter& I, int phase, int harris, int it,
                                            int bid, int nb) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int neg = S->neg;
  const int nN = M.N - M.m;
  double mn = BIG;
  int nF = 0;
  for (int i = bid * BS + threadIdx.x; i < nN; i += nb * BS) {
    double a = I.alpha_r[i];
    if (neg) a = -a;
    int j = I.nonbasis[i];
    if (!ratio_candidate(M, I, phase, j, a)) continue;
    ++nF;
    double dj = I.d[j];
    double th = harris ? (a > 0 ? (dj + EPS_D) / a : (dj - EPS_D) / a) : dj / a;
    mn = fmin(mn, th);
  }
  mn = block_min<BS>(mn);
  nF = block_sumi<BS>(nF);
  if (threadIdx.x == 0) {
    I.part_d[bid] = mn;
    I.part_i[bid] = nF;
  }
  if (!last_block(&S->ctr[CTR_CHUZC1], nb)) return;
  double g = BIG;
  int tot = 0;
  for (int b = threadIdx.x; b < nb; b += BS) {
    g = fmin(g, I.part_d[b]);
    tot += I.part_i[b];
  }
  g = block_min<BS>(g);
  tot = block_sumi<BS>(tot);
  if (threadIdx.x == 0) {
    S->ctr[CTR_CHUZC1] = 0;
    S->nF = tot;
    S->tmp0 = g;
    if (tot == 0) raise_status(S, ST_NO_RATIO, it);
  }
}
__global__ __launch_bounds__(BS) void k_chuzc1(DevModel M, DevIter I, int phase, int harris, int it) {
  chuzc1_body(M, I, phase, harris, it, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BS) void k_chuzc2(DevModel M, DevIter I, int phase, int harris, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int neg = S->neg;
  const int nN = M.N - M.m;
  const double bound = S->tmp0;
  double best = -1.0;
  int bi = INT_MAX;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    double a = I.alpha_r[i];
    if (neg) a = -a;
    int j = I.nonbasis[i];
    if (!ratio_candidate(M, I, phase, j, a)) continue;
    double th = I.d[j] / a;
    bool ok = harris ? (th <= bound) : (fabs(th - bound) < EPS_Z);
    if (ok) argmax_merge(best, bi, fabs(a), i);
  }
  block_argmax<BS>(best, bi);
  if (threadIdx.x == 0) {
    I.part_d[blockIdx.x] = best;
    I.part_i[blockIdx.x] = bi;
  }
  if (!last_block(&S->ctr[CTR_CHUZC2])) return;
  double fb = -1.0;
  int fi = INT_MAX;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) argmax_merge(fb, fi, I.part_d[b], I.part_i[b]);
  block_argmax<BS>(fb, fi);
  if (threadIdx.x != 0) return;
  S->ctr[CTR_CHUZC2] = 0;
  chuzc_finish(M, I, phase, it, fi);
}

// ---------------------------------------------------------------------------
// FTRAN of a sparse column: x0 = B0^{-1} a_j.  Kernel rows of the column
// select rows of K^{-T}; slack positions subtract the coupling block.
// ---------------------------------------------------------------------------
constexpr int FT_MAXCOL = 512;
// Row i's factor data, which does not depend on the column: its kernel column
// c (or -1 for a slack-covered row), that row r and its coupling range.
struct ColRow {
  int c, r, qb, qe;
};
__device__ __forceinline__ ColRow col_row(const DevFactor& F, int i) {
  ColRow w;
  w.c = F.Pidx_of_pos[i];
  w.r = w.c < 0 ? F.slackrow_of_pos[i] : 0;
  w.qb = w.c < 0 ? F.crp[w.r] : 0;
  w.qe = w.c < 0 ? F.crp[w.r + 1] : 0;
  return w;
}
// Position i of x0 = B0^{-1} a for the column held in er/eri/ev (ne entries).
__device__ __forceinline__ double ftran_col_value(const DevFactor& F, const ColRow& w, int ne, const int* er,
                                                  const int* eri, const double* ev) {
  const int k = F.k;
  double x;
  if (w.c >= 0) {
    x = 0.0;
    for (int e = 0; e < ne; ++e)
      if (eri[e] >= 0) x += ev[e] * F.KinvT[(size_t)eri[e] * k + w.c];
  } else {
    double ar = 0.0;
    for (int e = 0; e < ne; ++e)
      if (er[e] == w.r) ar = ev[e];
    // Coupling entries four at a time, so their loads overlap; each x_P value
    // and the sum over q keep the order of a one-at-a-time loop.
    constexpr int U = 4;
    double s = 0.0;
    for (int q0 = w.qb; q0 < w.qe; q0 += U) {
      int c2[U];
      double cv[U], xp[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = q0 + u < w.qe;
        c2[u] = ok ? F.cci[q0 + u] : 0;
        cv[u] = ok ? F.ccv[q0 + u] : 0.0;
        xp[u] = 0.0;
      }
      for (int e = 0; e < ne; ++e)
        if (eri[e] >= 0) {
          const double* row = F.KinvT + (size_t)eri[e] * k;
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (q0 + u < w.qe) xp[u] += ev[e] * row[c2[u]];
        }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (q0 + u < w.qe) s += cv[u] * xp[u];
    }
    x = ar - s;
  }
  return x;
}
// The column's entries into LDS (row, its kernel row or -1, value); returns
// their count.  No barrier.
__device__ __forceinline__ int load_col(const DevModel& M, const DevFactor& F, int j, int* er, int* eri,
                                        double* ev) {
  const int cb = M.cptr[j];
  const int ne = min(M.cptr[j + 1] - cb, FT_MAXCOL);
  for (int e = threadIdx.x; e < ne; e += BS) {
    int rr = M.ridx[cb + e];
    er[e] = rr;
    eri[e] = F.Ridx_of_row[rr];
    ev[e] = M.cval[cb + e];
  }
  return ne;
}

__device__ __forceinline__ void ftran_col_body(const DevModel& M, const DevIter& I, const DevFactor& F, int col,
                                               double* out, int check, int bid) {
  DevScalars* S = I.S;
  if (skip_launch(S, check)) return;
  const int j = col >= 0 ? col : S->q;
  __shared__ int er[FT_MAXCOL];
  __shared__ int eri[FT_MAXCOL];
  __shared__ double ev[FT_MAXCOL];
  // The row's own factor data does not depend on the column: fetch it first.
  const int i = bid * BS + threadIdx.x;
  const bool live = i < M.m;
  const ColRow w = live ? col_row(F, i) : ColRow{0, 0, 0, 0};
  const int ne = load_col(M, F, j, er, eri, ev);
  __syncthreads();
  if (!live) return;
  out[i] = ftran_col_value(F, w, ne, er, eri, ev);
}

__global__ __launch_bounds__(BS) void k_ftran_col(DevModel M, DevIter I, DevFactor F, int col, double* out,
                                                  int check) {
  ftran_col_body(M, I, F, col, out, check, blockIdx.x);
}

// The eta file's coefficients for an FTRAN whose eta-free part is x0:
// xm[k] = x0 at eta k's pivot when no earlier eta has that pivot (else 0),
// sk = MF xm, pk[k] = eta k's pivot.  Every thread of the workgroup takes part;
// ends with a barrier.
__device__ __forceinline__ void eta_solve(const DevIter& I, int t, const double* xm, double* sk, const double* mft);
__device__ __forceinline__ void eta_coeffs(const DevIter& I, const double* __restrict__ x0, int t, double* xm,
                                           double* sk, int* pk, double* mft) {
  for (int k = threadIdx.x; k < t; k += BS) {
    int q = I.eta_pos[k];
    pk[k] = q;
    xm[k] = (I.eta_last[k] < 0) ? x0[q] : 0.0;
  }
  if (t <= MF_TRI_MAX) load_mf_tri(I, t, mft);
  __syncthreads();
  eta_solve(I, t, xm, sk, mft);
}
// sk = MF xm (MF's triangle in LDS when t <= MF_TRI_MAX); ends with a barrier.
__device__ __forceinline__ void eta_solve(const DevIter& I, int t, const double* xm, double* sk, const double* mft) {
  const int cap = I.eta_cap;
  const bool tile = t <= MF_TRI_MAX;
  // s = MF xm: one thread per row of MF with four independent partial sums;
  // the row comes from LDS, or from global memory in 16-byte pairs.
  for (int k = threadIdx.x; k < t; k += BS) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int j = 0;
    if (tile) {
      const double* row = mft + tri(k);
      for (; j + 3 <= k; j += 4) {
        a0 += row[j] * xm[j];
        a1 += row[j + 1] * xm[j + 1];
        a2 += row[j + 2] * xm[j + 2];
        a3 += row[j + 3] * xm[j + 3];
      }
      for (; j <= k; ++j) a0 += row[j] * xm[j];
    } else {
      const double* row = I.MF + (size_t)k * cap;  // 16-byte aligned: cap is even
      for (; j + 3 <= k; j += 4) {
        const double2 r0 = *reinterpret_cast<const double2*>(row + j);
        const double2 r1 = *reinterpret_cast<const double2*>(row + j + 2);
        a0 += r0.x * xm[j];
        a1 += r0.y * xm[j + 1];
        a2 += r1.x * xm[j + 2];
        a3 += r1.y * xm[j + 3];
      }
      for (; j <= k; ++j) a0 += row[j] * xm[j];
    }
    sk[k] = (a0 + a1) + (a2 + a3);
  }
  __syncthreads();
}

// Row i of the FTRAN result: x0[i] with the etas applied in order (an eta on
// row i replaces the value, the others add to it).  The entries of eta k at
// row i are read from stride-m storage (`e`, eta k at e[k * ld]), EA of them
// loaded ahead of their (order-dependent) application.
template <int EA = 32>
__device__ __forceinline__ double eta_apply(double x, const double* __restrict__ e, size_t ld, const double* sk,
                                            const int* pk, int i, int t) {
  int k = 0;
  for (; k + EA <= t; k += EA) {
    double v[EA];
#pragma unroll
    for (int q = 0; q < EA; ++q) v[q] = e[(size_t)(k + q) * ld];
#pragma unroll
    for (int q = 0; q < EA; ++q) {
      const double s = sk[k + q];
      if (pk[k + q] == i) x = s * v[q];
      else if (s != 0.0) x = x + s * v[q];
    }
  }
  for (; k < t; ++k) {
    const double s = sk[k];
    if (pk[k] == i) x = s * e[(size_t)k * ld];
    else if (s != 0.0) x = x + s * e[(size_t)k * ld];
  }
  return x;
}

// Forward etas (linalg.cpp:21-32) on x0 = B0^{-1} a.  The multiplier of eta
// k is the value at its pivot just before it is applied, s = MF (mask .* x0
// at the pivots); every workgroup forms s from MF, then updates its slice:
// x = x0 + sum_k s_k eta_k, with a pivot position overwritten by s_k eta_k[p_k].
// mode 0: out[i] = x; mode 1: x_var[basis[i]] = x (dualSimplexInit's scatter).
__device__ __forceinline__ void ftran_out_body(const DevModel& M, const DevIter& I, const double* __restrict__ x0,
                                               double* out, int t, int mode, int check, int bid) {
  DevScalars* S = I.S;
  if (skip_launch(S, check)) return;
  __shared__ double xm[ETA_MAX];
  __shared__ double sk[ETA_MAX];
  __shared__ int pk[ETA_MAX];
  extern __shared__ double mft[];  // MF's lower triangle (t <= MF_TRI_MAX)
  eta_coeffs(I, x0, t, xm, sk, pk, mft);
  const int i = bid * BS + threadIdx.x;
  if (i >= M.m) return;
  const double x = eta_apply(x0[i], I.etas + i, M.m, sk, pk, i, t);
  if (mode == 0) out[i] = x;
  else out[I.basis[i]] = x;
}

__global__ __launch_bounds__(BS) void k_ftran_out(DevModel M, DevIter I, const double* __restrict__ x0,
                                                  double* out, int t, int mode, int check) {
  ftran_out_body(M, I, x0, out, t, mode, check, blockIdx.x);
}

// FTRAN of a column with the etas in one launch (t > 0): every workgroup
// computes x0 = B0^{-1} a at the etas' pivots itself (the same arithmetic as
// k_ftran_col), forms the eta coefficients, then writes x0 and the result for
// its own rows.  Results are bit-for-bit those of k_ftran_col + k_ftran_out.
__device__ __forceinline__ void ftran_col_eta_body(const DevModel& M, const DevIter& I, const DevFactor& F, int col,
                                                   double* out, int t, int check, int bid) {
  DevScalars* S = I.S;
  if (skip_launch(S, check)) return;
  const int j = col >= 0 ? col : S->q;
  __shared__ int er[FT_MAXCOL];
  __shared__ int eri[FT_MAXCOL];
  __shared__ double ev[FT_MAXCOL];
  __shared__ double xm[ETA_MAX];
  __shared__ double sk[ETA_MAX];
  __shared__ int pk[ETA_MAX];
  extern __shared__ double mft[];
  const int i = bid * BS + threadIdx.x;
  const bool live = i < M.m;
  const ColRow w = live ? col_row(F, i) : ColRow{0, 0, 0, 0};
  const int ne = load_col(M, F, j, er, eri, ev);
  if (t <= MF_TRI_MAX) load_mf_tri(I, t, mft);
  __syncthreads();
  for (int k = threadIdx.x; k < t; k += BS) {
    const int q = I.eta_pos[k];
    pk[k] = q;
    xm[k] = (I.eta_last[k] < 0) ? ftran_col_value(F, col_row(F, q), ne, er, eri, ev) : 0.0;
  }
  __syncthreads();
  eta_solve(I, t, xm, sk, mft);
  if (!live) return;
  const double x0 = ftran_col_value(F, w, ne, er, eri, ev);
  I.x0[i] = x0;
  out[i] = eta_apply<64>(x0, I.etas + i, M.m, sk, pk, i, t);
}
__global__ __launch_bounds__(BS) void k_ftran_col_eta(DevModel M, DevIter I, DevFactor F, int col, double* out,
                                                      int t, int check) {
  ftran_col_eta_body(M, I, F, col, out, t, check, blockIdx.x);
}
// The entering column's FTRAN with its etas (x0 and alpha_q), beside tau's
// etas (tau0 -> tau) in workgroups [nb1, ...).
__global__ __launch_bounds__(BS) void k_stage_col_eta_out(DevModel M, DevIter I, DevFactor F, int nb1,
                                                          const double* tau0, double* tau, int t) {
  if ((int)blockIdx.x < nb1) ftran_col_eta_body(M, I, F, -1, I.alpha_q, t, 1, blockIdx.x);
  else ftran_out_body(M, I, tau0, tau, t, 0, 1, blockIdx.x - nb1);
}

// Dense FTRAN pieces: gather b_R, y = K^{-1} b_R (one wave per row), slack rows.
__global__ __launch_bounds__(BS) void k_gather(const double* src, const int* map, double* dst, int len,
                                               const DevScalars* S, int check) {
  if (skip_launch(S, check)) return;
  int i = blockIdx.x * BS + threadIdx.x;
  if (i < len) dst[i] = src[map[i]];
}

__device__ __forceinline__ void gemv_body(const double* __restrict__ A, int rows, int cols,
                                          const double* __restrict__ xin, const int* imap, double* __restrict__ dst,
                                          const int* omap, const DevScalars* S, int check, int bid) {
  if (skip_launch(S, check)) return;
  const int w = (bid * BS + threadIdx.x) / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (w >= rows) return;
  const double* row = A + (size_t)w * cols;
  double acc = 0.0;
  if (imap)
    for (int jj = lane; jj < cols; jj += WAVE) acc += row[jj] * xin[imap[jj]];
  else
    for (int jj = lane; jj < cols; jj += WAVE) acc += row[jj] * xin[jj];
  acc = wave_sum(acc);
  if (lane == 0) dst[omap ? omap[w] : w] = acc;
}

__global__ __launch_bounds__(BS) void k_gemv(const double* __restrict__ A, int rows, int cols,
                                             const double* __restrict__ xin, const int* imap,
                                             double* __restrict__ dst, const int* omap, const DevScalars* S,
                                             int check) {
  gemv_body(A, rows, cols, xin, imap, dst, omap, S, check, blockIdx.x);
}

__device__ __forceinline__ void ftran_slack_body(const DevModel& M, const DevFactor& F, const double* b, double* dst,
                                                 const DevScalars* S, int check, int bid) {
  if (skip_launch(S, check)) return;
  int r = bid * BS + threadIdx.x;
  if (r >= M.m) return;
  int sp = F.slackpos_of_row[r];
  if (sp < 0) return;
  double s = 0.0;
  for (int q = F.crp[r]; q < F.crp[r + 1]; ++q) s += F.ccv[q] * dst[F.pos_P[F.cci[q]]];
  dst[sp] = b[r] - s;
}

__global__ __launch_bounds__(BS) void k_ftran_slack(DevModel M, DevFactor F, const double* b, double* dst,
                                                    const DevScalars* S, int check) {
  ftran_slack_body(M, F, b, dst, S, check, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Stages of one iteration that do not depend on each other, in one launch:
// workgroups [0, nb1) run the first stage and the rest the second.  The DSE
// vector's FTRAN (tau = B^{-1} rho) needs only rho, so its three steps ride
// along with the pivot-row SpMV, CHUZC (textbook) and the entering column's
// FTRAN: the chain of a phase-1 or phase-2 iteration without bound flips has
// eight launches instead of eleven.  Each stage runs its kernel's code
// unchanged (same summation orders); tau's pre-eta values go to w2, so the
// entering column keeps x0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_stage_spmv_gemv(DevModel M, const int* __restrict__ posN,
                                                        const double* __restrict__ rho,
                                                        const double* __restrict__ cvec, double* __restrict__ alpha_r,
                                                        const DevScalars* S, int cols, int nb1, DevFactor F,
                                                        double* __restrict__ tau0) {
  if ((int)blockIdx.x < nb1) spmv_body(M, posN, rho, 1.0, 0.0, cvec, alpha_r, 0, S, 1, cols, blockIdx.x);
  else gemv_body(F.Kinv, F.k, F.k, rho, F.row_R, tau0, F.pos_P, S, 1, blockIdx.x - nb1);
}

__global__ __launch_bounds__(BS) void k_stage_chuzc_slack(DevModel M, DevIter I, int phase, int it, int nb1,
                                                          DevFactor F, const double* rho, double* tau0) {
  if ((int)blockIdx.x < nb1) chuzc_tb_body(M, I, phase, it, blockIdx.x, nb1);
  else ftran_slack_body(M, F, rho, tau0, I.S, 1, blockIdx.x - nb1);
}

// Harris: pass 1 of CHUZC beside tau's slack rows (pass 2 follows in k_chuzc2).
__global__ __launch_bounds__(BS) void k_stage_chuzc1_slack(DevModel M, DevIter I, int phase, int it, int nb1,
                                                           DevFactor F, const double* rho, double* tau0) {
  if ((int)blockIdx.x < nb1) chuzc1_body(M, I, phase, 1, it, blockIdx.x, nb1);
  else ftran_slack_body(M, F, rho, tau0, I.S, 1, blockIdx.x - nb1);
}

// ---------------------------------------------------------------------------
// Reduced-cost update (sequentialDualSimplex.cpp:428-467, and
// simpleReducedCostsUpate :163-172 in phase 1) with the stability check.
// ---------------------------------------------------------------------------
// Grid-stride part of the dual update: d_j -= theta * alpha_r over the
// nonbasic set, with the dual-infeasibility and bound-flip counts.
__device__ __forceinline__ void dupdate_body(const DevModel& M, const DevIter& I, int phase, double theta, int q_idx,
                                             int& dinf) {
  const int nN = M.N - M.m;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    int j = I.nonbasis[i];
    double dj = I.d[j] - theta * I.alpha_r[i];
    I.d[j] = dj;
    if (i != q_idx) dinf += dual_infeas(M.bt[j], dj);
  }
}

// Thread 0 of the last workgroup: entering/leaving duals, the stability check
// of alpha_q[p] against alpha_r[q], bound flips, objective.
// aqp: alpha_q at the leaving row.
__device__ __forceinline__ void dupdate_final(const DevModel& M, const DevIter& I, int phase, int it, double theta,
                                              int td, double aqp) {
  DevScalars* S = I.S;
  const int p = S->p, q = S->q;
  I.d[p] = -theta;
  I.d[q] = 0.0;
  S->dinf = td + dual_infeas(M.bt[p], -theta);
  if (fabs(aqp - S->alpha_rq) > EPS_R * (1 + fabs(aqp))) {
    S->refac_mid = 1;
    S->stop_iter = it;
  }
  if (phase == 2) {
    S->obj = S->obj + theta * S->delta;
  } else {
    S->obj = S->obj - theta * S->f_p;
  }
  S->alpha_p = aqp;
}

// ---------------------------------------------------------------------------
// Basis change: x_B (or f) update, new eta, DSE weights, swap
// (sequentialDualSimplex.cpp:174-227).  The finaliser appends the eta's
// pivot and its row/column of the E table.
// ---------------------------------------------------------------------------
struct UpdateArgs {
  int p_idx, t;
  double ap, bp, thP;
  double* eta;
};
__device__ __forceinline__ UpdateArgs update_args(const DevModel& M, const DevIter& I, int phase) {
  const DevScalars* S = I.S;
  UpdateArgs u;
  u.p_idx = S->p_idx;
  u.ap = I.alpha_q[u.p_idx];
  u.bp = I.beta[u.p_idx];
  u.thP = (phase == 2 ? S->delta : S->f_p) / u.ap;
  u.t = S->t;
  u.eta = I.etas + (size_t)u.t * M.m;
  return u;
}

// Grid-stride part of the basis change: x_B (or f), the new eta, DSE weights.
__device__ __forceinline__ void update_body(const DevModel& M, const DevIter& I, int phase, const UpdateArgs& u,
                                            bool dse = true) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    double ai = I.alpha_q[i];
    if (phase == 2) {
      int j = I.basis[i];
      I.x[j] = I.x[j] - u.thP * ai;
    } else if (i != u.p_idx) {
      I.f[i] = I.f[i] - ai * u.thP;
    }
    if (i != u.p_idx) {
      u.eta[i] = -ai / u.ap;
      double r = ai / u.ap;
      if (dse) I.beta[i] = I.beta[i] - 2 * ai / u.ap * I.tau[i] + (r * r) * u.bp;
    } else {
      u.eta[i] = 1 / u.ap;
    }
  }
}

// Entry i of the new eta, -alpha_q[i]/alpha_p off the pivot and 1/alpha_p at
// it: the value update_body writes, recomputed from alpha_q (written by an
// earlier launch) so that the last workgroup never reads eta entries that
// other workgroups of the same launch wrote.
__device__ __forceinline__ double eta_value(const UpdateArgs& u, int i, double ai) {
  return (i != u.p_idx) ? -ai / u.ap : 1 / u.ap;
}
__device__ __forceinline__ double eta_entry(const DevIter& I, const UpdateArgs& u, int i) {
  return eta_value(u, i, i != u.p_idx ? I.alpha_q[i] : 0.0);
}

// Last workgroup (every thread): new E entries, row t (eta_t at the older
// pivots) and column t (older etas at the new pivot); then row t of
// MF = (I - L)^{-1}, where L[t][i] = E[i][t] for i at or after the previous
// eta on the same position; then (thread 0) the eta bookkeeping and the swap.
// aq_pos (optional): alpha_q at the older etas' pivots, aq_pos[j] at eta j's;
// mft (optional): MF's lower triangle in LDS, rows packed (row i at tri(i)).
__device__ __forceinline__ void update_final(const DevModel& M, const DevIter& I, int phase, const UpdateArgs& u,
                                             const double* aq_pos = nullptr, const double* mft = nullptr) {
  DevScalars* S = I.S;
  const int t = u.t, p_idx = u.p_idx;
  const int cap = I.eta_cap;
  __shared__ double ecol[ETA_MAX];
  __shared__ int s_last;
  if (threadIdx.x == 0) s_last = -1;
  __syncthreads();
  for (int jj = threadIdx.x; jj < t; jj += BS) {
    I.E[(size_t)t * cap + jj] =
        aq_pos ? eta_value(u, I.eta_pos[jj], aq_pos[jj]) : eta_entry(I, u, I.eta_pos[jj]);
    double e = I.etas[(size_t)jj * M.m + p_idx];
    I.E[(size_t)jj * cap + t] = e;
    ecol[jj] = e;
    if (I.eta_pos[jj] == p_idx) atomicMax(&s_last, jj);
  }
  __syncthreads();
  const int lastt = s_last;
  for (int j = threadIdx.x; j < t; j += BS) {
    double acc = 0.0;
    if (mft)
      for (int i = (j > lastt ? j : lastt); i < t; ++i) acc += ecol[i] * mft[tri(i) + j];
    else
      for (int i = (j > lastt ? j : lastt); i < t; ++i) acc += ecol[i] * I.MF[(size_t)i * cap + j];
    I.MF[(size_t)t * cap + j] = acc;
  }
  if (threadIdx.x != 0) return;
  I.MF[(size_t)t * cap + t] = 1.0;
  I.eta_last[t] = lastt;
  I.E[(size_t)t * cap + t] = eta_value(u, p_idx, 0.0);
  I.eta_pos[t] = p_idx;
  S->t = t + 1;
  I.beta[p_idx] = u.bp / (u.ap * u.ap);
  const int p = S->p, q = S->q, q_idx = S->q_idx;
  if (phase == 2) {
    I.x[q] = I.x[q] + u.thP;
  } else {
    I.f[p_idx] = I.f[p_idx] / u.ap + S->corr;
  }
  I.basis[p_idx] = q;
  I.nonbasis[q_idx] = p;
  I.posN[q] = -1;
  I.posN[p] = q_idx;
  S->theta_P = u.thP;
  const double theta = S->theta;
  if (phase == 2) {
    if (fabs(theta) < EPS_A) S->cycle = S->cycle + 1;
  } else {
    S->cycle = (fabs(theta) < EPS_A) ? S->cycle + 1 : 0;
  }
}

__global__ __launch_bounds__(BS) void k_update(DevModel M, DevIter I, int phase, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const UpdateArgs u = update_args(M, I, phase);
  update_body(M, I, phase, u);
  if (!last_block(&S->ctr[CTR_UPD])) return;
  if (threadIdx.x == 0) S->ctr[CTR_UPD] = 0;
  update_final(M, I, phase, u);
}

// Dual update and basis change in one launch. The basis change only waits for
// the dual update's stability check (alpha_q[p] against alpha_r[q]), which
// every workgroup evaluates at entry from the same two values; when it fails
// the basis change is left out and the refactor request is raised, as the
// separate kernels do. The DSE vector tau must be ready before the launch.
__global__ __launch_bounds__(BS) void k_dupdate_update(DevModel M, DevIter I, int phase, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const double theta = S->theta;
  const UpdateArgs u = update_args(M, I, phase);
  const bool stable = !(fabs(u.ap - S->alpha_rq) > EPS_R * (1 + fabs(u.ap)));
  int dinf = 0;
  dupdate_body(M, I, phase, theta, S->q_idx, dinf);
  if (stable) update_body(M, I, phase, u);
  dinf = block_sumi<BS>(dinf);
  if (threadIdx.x == 0) I.part_i[blockIdx.x] = dinf;
  if (!last_block(&S->ctr[CTR_DUPD])) return;
  int td = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) td += I.part_i[b];
  td = block_sumi<BS>(td);
  if (threadIdx.x == 0) {
    S->ctr[CTR_DUPD] = 0;
    dupdate_final(M, I, phase, it, theta, td, I.alpha_q[S->p_idx]);
  }
  if (stable) update_final(M, I, phase, u);
}

// The second half of a dual iteration in one launch: the entering column's
// FTRAN, its etas (alpha_q) and the DSE vector's etas (tau0 -> tau), then the
// dual update and the basis change of k_dupdate_update.  Every workgroup
// computes x0 = B0^{-1} a_q at the eta pivots (the arithmetic of k_ftran_col)
// and reads tau0 there, forms both sets of eta coefficients from MF's
// triangle (in LDS), and writes x0, alpha_q and tau for its rows as it
// updates them.  The last workgroup takes alpha_q at the older etas' pivots
// from the E table (E[k][j] holds eta k at eta j's pivot) in the same order
// of operations, so that it reads no entry another workgroup of this launch
// wrote.  Results are bit-for-bit those of the separate launches
// (k_stage_col_eta_out, then k_dupdate_update).  tau0 is tau itself when
// t == 0.
__global__ __launch_bounds__(BS) void k_dupdate_update_col(DevModel M, DevIter I, DevFactor F, int phase, int it,
                                                           const double* tau0, int t) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  __shared__ int er[FT_MAXCOL];
  __shared__ int eri[FT_MAXCOL];
  __shared__ double ev[FT_MAXCOL];
  __shared__ double xm[ETA_MAX];  // x0 at the first etas' pivots, then (last workgroup) alpha_q there
  __shared__ double sk[ETA_MAX];
  __shared__ double xt[ETA_MAX];  // tau0 at the pivots
  __shared__ double st[ETA_MAX];
  __shared__ int pk[ETA_MAX];
  __shared__ double x0p[ETA_MAX + 1];  // x0 at every eta's pivot, [t] at the leaving row
  __shared__ double ep[ETA_MAX];       // the etas' entries at the leaving row
  extern __shared__ double mft[];
  const int p_idx = S->p_idx;
  // The first row's factor data does not depend on the column: fetch it first.
  const int i0 = blockIdx.x * BS + threadIdx.x;
  const ColRow w0 = i0 < M.m ? col_row(F, i0) : ColRow{0, 0, 0, 0};
  const int ne = load_col(M, F, S->q, er, eri, ev);
  if (t <= MF_TRI_MAX) load_mf_tri(I, t, mft);
  __syncthreads();
  for (int k = threadIdx.x; k <= t; k += BS) {
    const int q = k < t ? I.eta_pos[k] : p_idx;
    const double x = ftran_col_value(F, col_row(F, q), ne, er, eri, ev);
    x0p[k] = x;
    if (k < t) {
      const bool first = I.eta_last[k] < 0;
      pk[k] = q;
      xm[k] = first ? x : 0.0;
      xt[k] = first ? tau0[q] : 0.0;
      ep[k] = I.etas[(size_t)k * M.m + p_idx];
    }
  }
  __syncthreads();
  eta_solve(I, t, xm, sk, mft);
  eta_solve(I, t, xt, st, mft);
  const double theta = S->theta;
  UpdateArgs u;
  u.p_idx = p_idx;
  {
    // alpha_q[p]: the etas applied in order, from the entries staged in LDS.
    double x = x0p[t];
    for (int k = 0; k < t; ++k) {
      const double sv = sk[k];
      if (pk[k] == p_idx) x = sv * ep[k];
      else if (sv != 0.0) x = x + sv * ep[k];
    }
    u.ap = x;
  }
  u.bp = I.beta[u.p_idx];
  u.thP = (phase == 2 ? S->delta : S->f_p) / u.ap;
  u.t = S->t;
  u.eta = I.etas + (size_t)u.t * M.m;
  const bool stable = !(fabs(u.ap - S->alpha_rq) > EPS_R * (1 + fabs(u.ap)));
  int dinf = 0;
  dupdate_body(M, I, phase, theta, S->q_idx, dinf);
  for (int i = i0; i < M.m; i += gridDim.x * BS) {
    const double x0 = ftran_col_value(F, i == i0 ? w0 : col_row(F, i), ne, er, eri, ev);
    const double ai = eta_apply<64>(x0, I.etas + i, M.m, sk, pk, i, t);
    const double ti = eta_apply<64>(tau0[i], I.etas + i, M.m, st, pk, i, t);
    I.x0[i] = x0;
    I.alpha_q[i] = ai;
    I.tau[i] = ti;
    if (!stable) continue;
    if (phase == 2) {
      int j = I.basis[i];
      I.x[j] = I.x[j] - u.thP * ai;
    } else if (i != u.p_idx) {
      I.f[i] = I.f[i] - ai * u.thP;
    }
    if (i != u.p_idx) {
      u.eta[i] = -ai / u.ap;
      double r = ai / u.ap;
      I.beta[i] = I.beta[i] - 2 * ai / u.ap * ti + (r * r) * u.bp;
    } else {
      u.eta[i] = 1 / u.ap;
    }
  }
  dinf = block_sumi<BS>(dinf);
  if (threadIdx.x == 0) I.part_i[blockIdx.x] = dinf;
  if (!last_block(&S->ctr[CTR_DUPD])) return;
  int td = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) td += I.part_i[b];
  td = block_sumi<BS>(td);
  if (threadIdx.x == 0) {
    S->ctr[CTR_DUPD] = 0;
    dupdate_final(M, I, phase, it, theta, td, u.ap);
  }
  if (!stable) return;
  // alpha_q at eta j's pivot, into xm (free after the coefficients).
  const int cap = I.eta_cap;
  for (int j = threadIdx.x; j < t; j += BS) xm[j] = eta_apply<64>(x0p[j], I.E + j, cap, sk, pk, pk[j], t);
  __syncthreads();
  update_final(M, I, phase, u, xm, t <= MF_TRI_MAX ? mft : nullptr);
}

// ---------------------------------------------------------------------------
// Primal simplex iterations of a warm start (params.warm_primal; no
// counterpart in the reference).  The host makes the starting basis primal
// feasible by moving each violated bound of a basic variable just past its
// value; these iterations then remove the dual infeasibilities: Devex pricing
// over the nonbasic set, FTRAN of the entering column, a two-pass (Harris)
// ratio test over the basic rows, BTRAN of the leaving row, the pivot-row SpMV
// and the update of d, x_B and the eta file.  With d feasible the host
// restores the bounds and the dual simplex finishes from that basis.  One
// iteration: k_pprice -> FTRAN -> k_pratio1/2 -> BTRAN -> SpMV -> k_pupdate.
// Models with Free or Boxed variables do not take this path.
// ---------------------------------------------------------------------------
constexpr double PRIMAL_PIV = 1e-7;  // smallest |alpha_q[i]| of a leaving row
constexpr double PRIMAL_TOL = 1e-9;  // feasibility slack of the first ratio pass

// Entering column: the largest d_j^2 / w_j over nonbasic Lower variables with
// d_j < -EPS_D and Upper ones with d_j > EPS_D (ties to the lowest position).
// None left: ST_OPTIMAL (the basis is dual feasible for the true costs).
// S->neg records the direction: 0 = x_q increases, 1 = it decreases.
__global__ __launch_bounds__(BS) void k_pprice(DevModel M, DevIter I, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int nN = M.N - M.m;
  double bw = -1.0;
  int bi = INT_MAX;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    const int j = I.nonbasis[i];
    const int8_t t = M.bt[j];
    const double dj = I.d[j];
    if ((t == BT_LOWER && dj < -EPS_D) || (t == BT_UPPER && dj > EPS_D))
      argmax_merge(bw, bi, dj * dj / I.devex[j], i);
  }
  block_argmax<BS>(bw, bi);
  if (threadIdx.x == 0) {
    I.part_d[blockIdx.x] = bw;
    I.part_i[blockIdx.x] = bi;
  }
  if (!last_block(&S->ctr[CTR_PRICE])) return;
  double fw = -1.0;
  int fi = INT_MAX;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) argmax_merge(fw, fi, I.part_d[b], I.part_i[b]);
  block_argmax<BS>(fw, fi);
  if (threadIdx.x != 0) return;
  S->ctr[CTR_PRICE] = 0;
  if (fi == INT_MAX) {
    raise_status(S, ST_OPTIMAL, it);
    return;
  }
  const int q = I.nonbasis[fi];
  S->q_idx = fi;
  S->q = q;
  S->neg = I.d[q] < 0 ? 0 : 1;
  S->dq = I.d[q];
  S->wq = I.devex[q];
}

// Ratio of basic row i for the step of x_q in direction dir: the row blocks
// at its lower bound when a = dir * alpha_q[i] > PRIMAL_PIV, at its upper
// bound when a < -PRIMAL_PIV (x_B moves by -step * alpha_q).  Returns false
// when the row does not block; bound receives the blocking bound.
__device__ __forceinline__ bool primal_row(const DevModel& M, const DevIter& I, double dir, int i, double& a,
                                           double& x, double& bound) {
  a = dir * I.alpha_q[i];
  const int j = I.basis[i];
  const int8_t t = M.bt[j];
  x = I.x[j];
  if (a > PRIMAL_PIV && t != BT_UPPER && t != BT_FREE) {
    bound = M.l[j];
    return true;
  }
  if (a < -PRIMAL_PIV && (t == BT_UPPER || t == BT_FIXED || t == BT_BOXED)) {
    bound = M.u[j];
    return true;
  }
  return false;
}

// Pass 1: theta_max = min over blocking rows of the ratio with the bound
// relaxed by PRIMAL_TOL (at least 0).  No blocking row: ST_NO_PIVOT (the
// direction is unbounded for the shifted bounds; the host gives up the
// primal iterations).
__global__ __launch_bounds__(BS) void k_pratio1(DevModel M, DevIter I, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const double dir = S->neg ? -1.0 : 1.0;
  double mn = BIG;
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    double a, x, bd;
    if (!primal_row(M, I, dir, i, a, x, bd)) continue;
    const double r = a > 0 ? (x - bd + PRIMAL_TOL) / a : (x - bd - PRIMAL_TOL) / a;
    mn = fmin(mn, fmax(r, 0.0));
  }
  mn = block_min<BS>(mn);
  if (threadIdx.x == 0) I.part_d[blockIdx.x] = mn;
  if (!last_block(&S->ctr[CTR_CHUZC1])) return;
  double g = BIG;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) g = fmin(g, I.part_d[b]);
  g = block_min<BS>(g);
  if (threadIdx.x != 0) return;
  S->ctr[CTR_CHUZC1] = 0;
  S->tmp0 = g;
  if (!(g < BIG)) raise_status(S, ST_NO_PIVOT, it);
}

// Pass 2: among blocking rows whose exact ratio is within theta_max, the
// largest |alpha_q[i]| (ties to the lowest position) leaves; delta = x_p -
// its bound, so the basis change moves x_q by delta / alpha_q[p].
__device__ __forceinline__ void pratio2_scan(const DevModel& M, const DevIter& I, double dir, double g, int i0,
                                             int stride, double& best, int& bi) {
  for (int i = i0; i < M.m; i += stride) {
    double a, x, bd;
    if (!primal_row(M, I, dir, i, a, x, bd)) continue;
    if ((x - bd) / a <= g) argmax_merge(best, bi, fabs(a), i);
  }
}
// The same scan by one workgroup over every row, U rows per thread at a time
// with their loads issued together (the rows are independent; the arg-max does
// not depend on the order).
__device__ __forceinline__ void pratio2_scan_wg(const DevModel& M, const DevIter& I, double dir, double g,
                                                double& best, int& bi) {
  constexpr int U = 8;
  for (int i0 = threadIdx.x; i0 < M.m; i0 += BS * U) {
    double a[U], x[U], lb[U], ub[U];
    int j[U];
    int8_t bt[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * BS;
      a[u] = i < M.m ? dir * I.alpha_q[i] : 0.0;
      j[u] = i < M.m ? I.basis[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bt[u] = M.bt[j[u]];
      x[u] = I.x[j[u]];
      lb[u] = M.l[j[u]];
      ub[u] = M.u[j[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * BS;
      if (i >= M.m) continue;
      const int8_t t = bt[u];
      double bd;
      if (a[u] > PRIMAL_PIV && t != BT_UPPER && t != BT_FREE) bd = lb[u];
      else if (a[u] < -PRIMAL_PIV && (t == BT_UPPER || t == BT_FIXED || t == BT_BOXED)) bd = ub[u];
      else continue;
      if ((x[u] - bd) / a[u] <= g) argmax_merge(best, bi, fabs(a[u]), i);
    }
  }
}
// The leaving row fi (INT_MAX: none, ST_NO_PIVOT) is recorded; with BTRAN the
// workgroup goes on with BTRAN's set-up for it, as k_price<true> does.  Every
// thread of the workgroup calls it.
template <bool BTRAN>
__device__ __forceinline__ void pratio_leave(const DevModel& M, const DevIter& I, int it, const DevFactor& F, int t,
                                             double dir, int fi) {
  DevScalars* S = I.S;
  __shared__ int s_row;  // the leaving row, or -1 when an event was raised
  if (threadIdx.x == 0) {
    s_row = -1;
    if (fi == INT_MAX) {
      raise_status(S, ST_NO_PIVOT, it);
    } else {
      double a, x, bd;
      primal_row(M, I, dir, fi, a, x, bd);
      S->p_idx = fi;
      S->p = I.basis[fi];
      S->delta = x - bd;
      S->is_lower = a > 0 ? 1 : 0;
      s_row = fi;
    }
  }
  if constexpr (BTRAN) {
    __syncthreads();
    if (s_row >= 0) btran_prep_body(M, I, F, t, s_row, -1);
  }
}
template <bool BTRAN>
__global__ __launch_bounds__(BS) void k_pratio2(DevModel M, DevIter I, int it, DevFactor F, int t) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const double dir = S->neg ? -1.0 : 1.0;
  double best = -1.0;
  int bi = INT_MAX;
  pratio2_scan(M, I, dir, S->tmp0, blockIdx.x * BS + threadIdx.x, gridDim.x * BS, best, bi);
  block_argmax<BS>(best, bi);
  if (threadIdx.x == 0) {
    I.part_d[blockIdx.x] = best;
    I.part_i[blockIdx.x] = bi;
  }
  if (!last_block(&S->ctr[CTR_CHUZC2])) return;
  double fb = -1.0;
  int fi = INT_MAX;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) argmax_merge(fb, fi, I.part_d[b], I.part_i[b]);
  block_argmax<BS>(fb, fi);
  if (threadIdx.x == 0) S->ctr[CTR_CHUZC2] = 0;
  pratio_leave<BTRAN>(M, I, it, F, t, dir, fi);
}

// Both passes and BTRAN's set-up in one launch, for m up to PRATIO_ONE_WG:
// the last workgroup of pass 1 scans every row for pass 2 itself (the same
// minimum and the same arg-max, whatever the split of the rows).
constexpr int PRATIO_ONE_WG = 8192;
__global__ __launch_bounds__(BS) void k_pratio_btran(DevModel M, DevIter I, int it, DevFactor F, int t) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const double dir = S->neg ? -1.0 : 1.0;
  double mn = BIG;
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    double a, x, bd;
    if (!primal_row(M, I, dir, i, a, x, bd)) continue;
    const double r = a > 0 ? (x - bd + PRIMAL_TOL) / a : (x - bd - PRIMAL_TOL) / a;
    mn = fmin(mn, fmax(r, 0.0));
  }
  mn = block_min<BS>(mn);
  if (threadIdx.x == 0) I.part_d[blockIdx.x] = mn;
  if (!last_block(&S->ctr[CTR_CHUZC1])) return;
  double g = BIG;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) g = fmin(g, I.part_d[b]);
  g = block_min<BS>(g);
  if (threadIdx.x == 0) {
    S->ctr[CTR_CHUZC1] = 0;
    S->tmp0 = g;
    if (!(g < BIG)) raise_status(S, ST_NO_PIVOT, it);
  }
  if (!(g < BIG)) return;
  double best = -1.0;
  int bi = INT_MAX;
  pratio2_scan_wg(M, I, dir, g, best, bi);
  block_argmax<BS>(best, bi);
  pratio_leave<true>(M, I, it, F, t, dir, bi);
}

// d_N -= theta alpha_r with the Devex weights w_j = max(w_j, (alpha_j /
// alpha_q)^2 w_q), then the basis change of k_dupdate_update without the DSE
// weights (they are recomputed exactly when the dual simplex takes over).
// Every workgroup forms, from values earlier launches wrote, the dual step
// theta = d_q / alpha_q[p] (alpha_q[p] passed the pivot tolerance) and
// alpha_r[q] for the stability check.
__global__ __launch_bounds__(BS) void k_pupdate(DevModel M, DevIter I, int it) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int q_idx = S->q_idx;
  const double arq = I.alpha_r[q_idx], wq = S->wq;
  const double theta = S->dq / I.alpha_q[S->p_idx];
  const UpdateArgs u = update_args(M, I, 2);
  const bool stable = !(fabs(u.ap - arq) > EPS_R * (1 + fabs(u.ap)));
  const int nN = M.N - M.m;
  int dinf = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    const int j = I.nonbasis[i];
    const double a = I.alpha_r[i];
    const double dj = I.d[j] - theta * a;
    I.d[j] = dj;
    if (i != q_idx) {
      dinf += dual_infeas(M.bt[j], dj);
      const double r = a / arq;
      I.devex[j] = fmax(I.devex[j], r * r * wq);
    }
  }
  if (stable) update_body(M, I, 2, u, false);
  dinf = block_sumi<BS>(dinf);
  if (threadIdx.x == 0) I.part_i[blockIdx.x] = dinf;
  if (!last_block(&S->ctr[CTR_DUPD])) return;
  int td = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) td += I.part_i[b];
  td = block_sumi<BS>(td);
  if (threadIdx.x == 0) {
    S->ctr[CTR_DUPD] = 0;
    S->alpha_rq = arq;
    S->theta = theta;
    dupdate_final(M, I, 2, it, theta, td, I.alpha_q[S->p_idx]);
    I.devex[S->p] = fmax(wq / (arq * arq), 1.0);
  }
  if (stable) update_final(M, I, 2, u);
}

// ---------------------------------------------------------------------------
// Multiple candidate rows per major iteration (the minor iterations of the
// reference's parallel dual simplex, parallelDualSimplex.cpp:431-485,
// :747-1124, on one GPU): the major pricing keeps the `nc` rows of largest
// delta^2/beta, their BTRAN rows are computed back to back, and each minor
// iteration takes the best still-eligible candidate without a BTRAN of its
// own; after each basis change the other candidates' rows are brought up to
// date with the product-form row update.
// ---------------------------------------------------------------------------
// Major pricing: the top-of-iteration checks and the optimality test of
// k_price (phase 2), then the nc best rows (weight, then lowest position).
__global__ __launch_bounds__(BS) void k_price_multi(DevModel M, DevIter I, int it, int skip_top, int perturb_on,
                                                    int nc) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  if (!skip_top && perturb_on && !S->perturbed && S->cycle > MAX_CYCLE) {
    if (blockIdx.x == 0 && threadIdx.x == 0) raise_status(S, ST_NEED_PERTURB, it);
    return;
  }
  __shared__ int taken[CAND_MAX];
  int pinf = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    const int j = I.basis[i];
    pinf |= primal_infeas(M.bt[j], I.x[j], M.l[j], M.u[j]) ? 1 : 0;
  }
  pinf = block_sumi<BS>(pinf);
  // this workgroup's nc best, one argmax round per candidate
  for (int c = 0; c < nc; ++c) {
    double bw = -1.0;
    int bi = INT_MAX;
    for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
      bool used = false;
      for (int c2 = 0; c2 < c; ++c2) used |= taken[c2] == i;
      if (used) continue;
      const int j = I.basis[i];
      double dl;
      if (var_infeas(M.bt[j], I.x[j], M.l[j], M.u[j], dl) && fabs(dl) > DSE_FLOOR && !I.blocked[j]) {
        const double w = (dl * dl) / I.beta[i];
        if (w > 0.0) argmax_merge(bw, bi, w, i);
      }
    }
    block_argmax<BS>(bw, bi);
    if (threadIdx.x == 0) {
      taken[c] = bi;
      I.part_d[blockIdx.x * CAND_MAX + c] = bw;
      I.part_i[blockIdx.x * CAND_MAX + c] = bi;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) I.part_i[4096 + blockIdx.x] = pinf;
  if (!last_block(&S->ctr[CTR_PRICE])) return;
  int fp = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) fp += I.part_i[4096 + b];
  fp = block_sumi<BS>(fp);
  const int total = gridDim.x * nc;
  int got = 0;
  for (int c = 0; c < nc; ++c) {
    double bw = -1.0;
    int bi = INT_MAX;
    for (int e = threadIdx.x; e < total; e += BS) {
      const int b = e / nc, c2 = e % nc;
      const int i = I.part_i[b * CAND_MAX + c2];
      if (i == INT_MAX) continue;
      bool used = false;
      for (int c3 = 0; c3 < c; ++c3) used |= taken[c3] == i;
      if (!used) argmax_merge(bw, bi, I.part_d[b * CAND_MAX + c2], i);
    }
    block_argmax<BS>(bw, bi);
    if (threadIdx.x == 0) {
      taken[c] = bi;
      if (bi != INT_MAX) {
        const int j = I.basis[bi];
        double dl;
        var_infeas(M.bt[j], I.x[j], M.l[j], M.u[j], dl);
        I.cand_pos[c] = bi;
        I.cand_w[c] = bw;
        I.cand_act[c] = dl < 0 ? 2 : 1;  // eligible, and on which side of its bounds
      }
    }
    __syncthreads();
    if (bi != INT_MAX) got = c + 1;
  }
  if (threadIdx.x != 0) return;
  S->ctr[CTR_PRICE] = 0;
  S->ncand = got;
  if (!fp) raise_status(S, ST_OPTIMAL, it);
  else if (got == 0) raise_status(S, ST_NO_PIVOT, it);
}

// Minor pricing (one workgroup): each eligible candidate's current
// infeasibility; it drops out when its row has become feasible or changed
// side, or when its weight fell below psi times its previous one
// (parallelDualSimplex.cpp:842-874).  The best remaining one is this
// iteration's leaving row, and its BTRAN row becomes rho.
__global__ __launch_bounds__(BS) void k_pick(DevModel M, DevIter I, int it, double psi) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  __shared__ int s_best;
  if (threadIdx.x == 0) {
    double bw = -1.0;
    int best = -1;
    double bdl = 0.0;
    for (int c = 0; c < S->ncand; ++c) {
      if (!I.cand_act[c]) continue;
      const int pos = I.cand_pos[c];
      const int j = I.basis[pos];
      double dl;
      const bool inf = var_infeas(M.bt[j], I.x[j], M.l[j], M.u[j], dl) && fabs(dl) > DSE_FLOOR;
      const double w = inf ? (dl * dl) / I.beta[pos] : 0.0;
      const bool same_side = (I.cand_act[c] == 2) == (dl < 0);  // 2: it was below its lower bound
      if (!inf || !same_side || w < psi * I.cand_w[c]) {
        I.cand_act[c] = 0;
        continue;
      }
      I.cand_w[c] = w;
      if (w > bw) {
        bw = w;
        best = c;
        bdl = dl;
      }
    }
    s_best = best;
    if (best < 0) {
      raise_status(S, ST_MAJOR_END, it);
    } else {
      const int pos = I.cand_pos[best];
      I.cand_act[best] = 0;
      S->cand_best = best;
      S->p_idx = pos;
      S->p = I.basis[pos];
      S->delta = bdl;
      S->is_lower = bdl < 0 ? 1 : 0;
      S->neg = S->is_lower;
    }
  }
  __syncthreads();
  if (s_best < 0) return;
  const double* src = I.rho_c + (size_t)s_best * M.m;
  for (int r = threadIdx.x; r < M.m; r += BS) I.rho[r] = src[r];
}

// After the basis change at row p: row c of the new inverse is
// row c - (alpha_q[c] / alpha_q[p]) row p, for every still-eligible candidate.
__global__ __launch_bounds__(BS) void k_rho_update(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const int c = blockIdx.y;
  if (c >= S->ncand || !I.cand_act[c]) return;
  const int pc = I.cand_pos[c], p = S->p_idx;
  const double f = I.alpha_q[pc] / I.alpha_q[p];
  double* rc = I.rho_c + (size_t)c * M.m;
  for (int r = blockIdx.x * BS + threadIdx.x; r < M.m; r += gridDim.x * BS) rc[r] = rc[r] - f * I.rho[r];
}

// ---------------------------------------------------------------------------
// Bound-flipping ratio test and BFRT FTRAN (sequentialDualSimplex.cpp:398-482),
// queued only for models with boxed variables.
// ---------------------------------------------------------------------------
constexpr int BFRT_BS = 1024;

// Exclusive prefix count of `flag` over the 1024 threads of a workgroup, in
// thread order; *total receives the count of the whole workgroup.
__device__ inline int block_excl_count(int flag, int* total) {
  __shared__ int wsum[BFRT_BS / WAVE];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  const unsigned long long b = __ballot(flag);
  const int pre = __popcll(b & ((1ull << lane) - 1ull));
  __syncthreads();
  if (lane == 0) wsum[w] = __popcll(b);
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < BFRT_BS / WAVE; ++k) {
    off += (k < w) ? wsum[k] : 0;
    tot += wsum[k];
  }
  *total = tot;
  return off + pre;
}

// The flip loop, one workgroup.  CHUZC chose a boxed candidate q whose flip
// alone does not absorb the leaving row's infeasibility delta: drop it
// (delta -= (u_q - l_q)|alpha_q|) and choose again among the remaining
// candidates with the same rule (textbook or Harris), until the choice is not
// boxed, or its flip would absorb what is left, or no candidate remains (then
// the last dropped one enters, as in the reference).  Writes q, theta and the
// signed remaining delta.
__global__ __launch_bounds__(BFRT_BS) void k_bfrt(DevModel M, DevIter I, int harris) {
  DevScalars* S = I.S;
  if (stopped(S) || !S->bfrt) return;
  const int neg = S->neg;
  const int nN = M.N - M.m;
  __shared__ double s_delta;
  __shared__ int s_nrm, s_qi, s_go;
  if (threadIdx.x == 0) {
    const int q0 = S->q_idx, j0 = S->q;
    s_delta = fabs(S->delta) - (M.u[j0] - M.l[j0]) * fabs(I.alpha_r[q0]);
    I.bmark[q0] = 1;
    I.blist[0] = q0;
    s_nrm = 1;
    s_qi = q0;
    s_go = 1;
  }
  __syncthreads();
  while (s_go) {
    // pass 1: bound of the step over the remaining candidates
    double mn = BIG;
    for (int i = threadIdx.x; i < nN; i += BFRT_BS) {
      if (I.bmark[i]) continue;
      double a = I.alpha_r[i];
      if (neg) a = -a;
      const int j = I.nonbasis[i];
      if (!ratio_candidate(M, I, 2, j, a)) continue;
      const double dj = I.d[j];
      mn = fmin(mn, harris ? (a > 0 ? (dj + EPS_D) / a : (dj - EPS_D) / a) : dj / a);
    }
    mn = block_min<BFRT_BS>(mn);
    // pass 2: largest |alpha| within the bound, lowest position on ties
    double best = -1.0;
    int bi = INT_MAX;
    for (int i = threadIdx.x; i < nN; i += BFRT_BS) {
      if (I.bmark[i]) continue;
      double a = I.alpha_r[i];
      if (neg) a = -a;
      const int j = I.nonbasis[i];
      if (!ratio_candidate(M, I, 2, j, a)) continue;
      const double th = I.d[j] / a;
      if (harris ? (th <= mn) : (fabs(th - mn) < EPS_Z)) argmax_merge(best, bi, fabs(a), i);
    }
    block_argmax<BFRT_BS>(best, bi);
    if (threadIdx.x == 0) {
      if (bi == INT_MAX) {
        s_go = 0;  // every candidate dropped: the last one enters
      } else {
        const int j = I.nonbasis[bi];
        s_qi = bi;
        const double span = M.u[j] - M.l[j];
        if (M.bt[j] == BT_BOXED && s_delta - span * fabs(I.alpha_r[bi]) > 0) {
          s_delta -= span * fabs(I.alpha_r[bi]);
          I.bmark[bi] = 1;
          I.blist[s_nrm++] = bi;
        } else {
          s_go = 0;
        }
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < s_nrm; k += BFRT_BS) I.bmark[I.blist[k]] = 0;
  if (threadIdx.x != 0) return;
  const int qi = s_qi, q = I.nonbasis[qi];
  const double arq = I.alpha_r[qi];
  S->q_idx = qi;
  S->q = q;
  S->alpha_rq = arq;
  S->theta = I.d[q] / arq;
  S->delta = (S->delta < 0) ? -s_delta : s_delta;
  S->bfrt = 0;
}

// Boxed nonbasic variables whose reduced cost changes sign in the dual update
// (the same d_j - theta alpha_j the update computes): +1 at the lower bound
// with d_j < -EPS_D (flips to the upper), -1 at the upper with d_j > EPS_D.
__global__ __launch_bounds__(BS) void k_flip_detect(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  const double theta = S->theta;
  const int nN = M.N - M.m;
  int cnt = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < nN; i += gridDim.x * BS) {
    const int j = I.nonbasis[i];
    signed char code = 0;
    if (M.bt[j] == BT_BOXED) {
      const double dj = I.d[j] - theta * I.alpha_r[i];
      const double xj = I.x[j];
      if (fabs(xj - M.l[j]) < EPS_Z && dj < -EPS_D) code = 1;
      else if (fabs(xj - M.u[j]) < EPS_Z && dj > EPS_D) code = -1;
    }
    I.flipc[i] = code;
    cnt += code != 0;
  }
  cnt = block_sumi<BS>(cnt);
  if (threadIdx.x == 0) I.part_i[blockIdx.x] = cnt;
  if (!last_block(&S->ctr[CTR_FLIP])) return;
  int tot = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) tot += I.part_i[b];
  tot = block_sumi<BS>(tot);
  if (threadIdx.x == 0) {
    S->ctr[CTR_FLIP] = 0;
    S->nflip = tot;
  }
}

// One workgroup: the flips in ascending position (the reference's loop
// order), their cost change dz = sum c_j (u_j - l_j) (+ for lower -> upper,
// - for upper -> lower), and the column change
//   bcc = sum_{lower->upper} (u_j - l_j) a_j - sum_{upper->lower} (u_j - l_j) a_j
// accumulated column after column (Matrix::addSparseCol, matrix.cpp:606-616).
__global__ __launch_bounds__(BFRT_BS) void k_flip_cc(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  if (skip_launch(S, 2)) return;
  const int nN = M.N - M.m;
  for (int r = threadIdx.x; r < M.m; r += BFRT_BS) I.bcc[r] = 0.0;
  __shared__ int s_base;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int i0 = 0; i0 < nN; i0 += BFRT_BS) {
    const int i = i0 + threadIdx.x;
    const int f = (i < nN && I.flipc[i] != 0) ? 1 : 0;
    int tot;
    const int off = block_excl_count(f, &tot);
    if (f) I.flist[s_base + off] = i;
    __syncthreads();
    if (threadIdx.x == 0) s_base += tot;
    __syncthreads();
  }
  // The sums run in one thread, in the reference's order (a column's entries
  // may repeat a row); a flip touches a handful of entries.
  if (threadIdx.x != 0) return;
  const int nf = s_base;
  double dz = 0.0;
  for (int k = 0; k < nf; ++k) {
    const int i = I.flist[k], j = I.nonbasis[i];
    const double span = M.u[j] - M.l[j];
    if (I.flipc[i] > 0) dz += M.c[j] * span;
    else dz -= M.c[j] * span;
  }
  S->dz = dz;
  for (int sgn = 1; sgn >= -1; sgn -= 2) {
    const double mult = (double)sgn;
    for (int k = 0; k < nf; ++k) {
      const int i = I.flist[k];
      if (I.flipc[i] != sgn) continue;
      const int j = I.nonbasis[i];
      const double span = M.u[j] - M.l[j];
      for (int e = M.cptr[j]; e < M.cptr[j + 1]; ++e) I.bcc[M.ridx[e]] += M.cval[e] * span * mult;
    }
  }
}

// x_B -= B^{-1} bcc, delta_z = dz - c_B . dx_B added to the objective, and the
// flipped variables moved to their other bound.
__global__ __launch_bounds__(BS) void k_bfrt_xb(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  if (skip_launch(S, 2)) return;
  double s = 0.0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.m; i += gridDim.x * BS) {
    const int j = I.basis[i];
    const double v = I.dxb[i];
    I.x[j] = I.x[j] - v;
    s += M.c[j] * v;
  }
  const int nf = S->nflip;
  for (int k = blockIdx.x * BS + threadIdx.x; k < nf; k += gridDim.x * BS) {
    const int i = I.flist[k], j = I.nonbasis[i];
    I.x[j] = I.flipc[i] > 0 ? M.u[j] : M.l[j];
  }
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) I.part_d[blockIdx.x] = s;
  if (!last_block(&S->ctr[CTR_BXB])) return;
  if (threadIdx.x == 0) {
    double g = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) g += I.part_d[b];
    S->ctr[CTR_BXB] = 0;
    S->obj = S->obj + (S->dz - g);
  }
}

// ---------------------------------------------------------------------------
// Test hooks that force a recovery path from the host; every launch ends
// normally.
// ---------------------------------------------------------------------------
// The recorded alpha_r[q] moves by 1e-6 relative: the stability check fails.
__global__ void k_force_stability(DevIter I) {
  DevScalars* S = I.S;
  if (stopped(S)) return;
  S->alpha_rq = S->alpha_rq + 1e-6 * (1.0 + fabs(S->alpha_rq));
}
// The first nonbasic Lower variable with d_j > EPS_D gets d_j = -1: the basis is
// no longer dual feasible.
__global__ void k_force_dual(DevModel M, DevIter I) {
  if (stopped(I.S)) return;
  for (int i = 0; i < M.N - M.m; ++i) {
    const int j = I.nonbasis[i];
    if (M.bt[j] == BT_LOWER && I.d[j] > EPS_D) {
      I.d[j] = -1.0;
      return;
    }
  }
}
// Kernel column c of the work matrix reads zero: exactly dependent.
__global__ void k_force_zero_column(double* W, int k, int c) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < k) W[(size_t)i * k + c] = 0.0;
}

// ---------------------------------------------------------------------------
// Initialisation helpers.
// ---------------------------------------------------------------------------

// setPrimalVars (tpp:74-101) over the nonbasic set.
__global__ __launch_bounds__(BS) void k_set_primal(DevModel M, DevIter I) {
  int nN = M.N - M.m;
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= nN) return;
  int j = I.nonbasis[i];
  double v = 0.0;
  switch (M.bt[j]) {
    case BT_FIXED: v = M.l[j]; break;
    case BT_FREE: v = 0.0; break;
    case BT_BOXED: v = (I.d[j] > 0) ? M.l[j] : M.u[j]; break;
    case BT_UPPER: v = M.u[j]; break;
    case BT_LOWER: v = M.l[j]; break;
  }
  I.x[j] = v;
}

// wN[j] = x[j] for nonbasic j, else 0 (Matrix::dotUpdate UPDATE's buffer).
__global__ __launch_bounds__(BS) void k_mask_nonbasic(DevModel M, DevIter I, const double* src, double* wN) {
  int j = blockIdx.x * BS + threadIdx.x;
  if (j >= M.N) return;
  wN[j] = (I.posN[j] >= 0) ? src[j] : 0.0;
}

// Row-wise SpMV (Matrix::dotUpdate UPDATE, matrix.cpp:330-341):
// out[r] = b[r]*beta + sum_e val_e * w[col_e] * alpha in CSR order.
__global__ __launch_bounds__(BS) void k_rowspmv(DevModel M, const double* w, const double* b, double alpha,
                                                double beta, double* out) {
  int r = blockIdx.x * BS + threadIdx.x;
  if (r >= M.m) return;
  double s = (b ? b[r] : 0.0) * beta;
  for (int e = M.rptr[r]; e < M.rptr[r + 1]; ++e) s = s + M.rval[e] * w[M.cidx[e]] * alpha;
  out[r] = s;
}

// obj = c . x over all variables.
__global__ __launch_bounds__(BS) void k_dot_obj(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  double s = 0.0;
  for (int j = blockIdx.x * BS + threadIdx.x; j < M.N; j += gridDim.x * BS) s += M.c[j] * I.x[j];
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) I.part_d[blockIdx.x] = s;
  if (!last_block(&S->ctr[CTR_DOT])) return;
  double g = 0.0;
  if (threadIdx.x == 0)
    for (int b = 0; b < (int)gridDim.x; ++b) g += I.part_d[b];
  if (threadIdx.x == 0) {
    S->ctr[CTR_DOT] = 0;
    S->obj = g;
  }
}

// Phase-1 infeasibility sets (tpp:776-795, sequentialDualSimplex.cpp:135-152):
// wN[j] = +1 for inf_l, -1 for inf_u, else 0; Z = sum w_j d_j; dinf count.
__global__ __launch_bounds__(BS) void k_phase1_inf(DevModel M, DevIter I) {
  DevScalars* S = I.S;
  double z = 0.0;
  int cnt = 0;
  for (int j = blockIdx.x * BS + threadIdx.x; j < M.N; j += gridDim.x * BS) {
    double w = 0.0;
    if (I.posN[j] >= 0) {
      int8_t t = M.bt[j];
      double dj = I.d[j];
      if ((t == BT_UPPER || t == BT_FREE) && dj > EPS_D) w = -1.0;
      else if ((t == BT_LOWER || t == BT_FREE) && dj < -EPS_D) w = 1.0;
      z += w * dj;
      cnt += dual_infeas(t, dj);
    }
    I.wN[j] = w;
  }
  z = block_sum<BS>(z);
  cnt = block_sumi<BS>(cnt);
  if (threadIdx.x == 0) {
    I.part_d[blockIdx.x] = z;
    I.part_i[blockIdx.x] = cnt;
  }
  if (!last_block(&S->ctr[CTR_INF])) return;
  if (threadIdx.x == 0) {
    double g = 0.0;
    int c = 0;
    for (int b = 0; b < (int)gridDim.x; ++b) {
      g += I.part_d[b];
      c += I.part_i[b];
    }
    S->ctr[CTR_INF] = 0;
    S->obj = g;
    S->dinf = c;
  }
}

// w1[i] = c[basis[i]]
__global__ __launch_bounds__(BS) void k_gather_cb(DevModel M, DevIter I, double* out) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i < M.m) out[i] = M.c[I.basis[i]];
}

// Transposed etas on a dense vector (linalg.cpp:34-44), one workgroup.  The
// modified pivot entries live in LDS so no thread reads a value another
// thread of the workgroup wrote to global memory.
__global__ __launch_bounds__(1024) void k_btran_dense_eta(DevModel M, DevIter I, double* v, int t) {
  __shared__ int mpos[ETA_MAX];
  __shared__ double mval[ETA_MAX];
  __shared__ double morig[ETA_MAX];
  __shared__ int nmod;
  __shared__ double red[16];
  if (threadIdx.x == 0) nmod = 0;
  __syncthreads();
  for (int k = t - 1; k >= 0; --k) {
    const double* eta = I.etas + (size_t)k * M.m;
    double s = 0.0;
    for (int i = threadIdx.x; i < M.m; i += 1024) s += eta[i] * v[i];
    s = wave_sum(s);
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double dot = 0.0;
      for (int w = 0; w < 16; ++w) dot += red[w];
      for (int q = 0; q < nmod; ++q) dot += eta[mpos[q]] * (mval[q] - morig[q]);
      int pk = I.eta_pos[k];
      int f = -1;
      for (int q = 0; q < nmod; ++q)
        if (mpos[q] == pk) f = q;
      if (f < 0) {
        f = nmod++;
        mpos[f] = pk;
        morig[f] = v[pk];
      }
      mval[f] = dot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < nmod; ++q) v[mpos[q]] = mval[q];
}

// rhs_P[c] = v[pos_P[c]] - sum_{(r,val) in C column c} val * v[slackpos(r)]
__global__ __launch_bounds__(BS) void k_btran_dense_rhs(DevFactor F, const double* v, double* rhs) {
  int c = blockIdx.x * BS + threadIdx.x;
  if (c >= F.k) return;
  double s = v[F.pos_P[c]];
  for (int q = F.ccp[c]; q < F.ccp[c + 1]; ++q) s = s - F.ccv2[q] * v[F.slackpos_of_row[F.ccr[q]]];
  rhs[c] = s;
}

// y[r] = v[slackpos(r)] for covered rows (kernel rows come from the GEMV).
__global__ __launch_bounds__(BS) void k_btran_dense_slack(DevModel M, DevFactor F, const double* v, double* y) {
  int r = blockIdx.x * BS + threadIdx.x;
  if (r >= M.m) return;
  int sp = F.slackpos_of_row[r];
  if (sp >= 0) y[r] = v[sp];
}

// Exact DSE weights right after a refactor (initBetaWeights(false),
// tpp:41-68): beta_i = ||e_i^T B0^{-1}||^2.  One workgroup per position.
__global__ __launch_bounds__(BS) void k_dse_exact(DevModel M, DevIter I, DevFactor F) {
  const int i = blockIdx.x;
  if (i >= M.m) return;
  const int k = F.k;
  int c = F.Pidx_of_pos[i];
  double s = 0.0;
  if (c >= 0) {
    const double* row = F.Kinv + (size_t)c * k;
    for (int jj = threadIdx.x; jj < k; jj += BS) s += row[jj] * row[jj];
  } else {
    int r = F.slackrow_of_pos[i];
    for (int jj = threadIdx.x; jj < k; jj += BS) {
      double g = 0.0;
      for (int q = F.crp[r]; q < F.crp[r + 1]; ++q) g += F.ccv[q] * F.Kinv[(size_t)F.cci[q] * k + jj];
      s += g * g;
    }
  }
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) I.beta[i] = (c >= 0) ? s : 1.0 + s;
}

__global__ __launch_bounds__(BS) void k_fill(double* p, double v, int n) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) p[i] = v;
}

// Plain streaming copy (16 B per lane per step): the reference kernel for
// scaling the FETCH_SIZE / WRITE_SIZE hardware counters.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }
static inline int cap_grid(long n, int bs, int cap) {
  int g = nblk(n, bs);
  if (g < 1) g = 1;
  return g > cap ? cap : g;
}

void launch_price(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it, int skip_top, int perturb_on) {
  k_price<false><<<cap_grid(M.m, BS, I.part_cap), BS, 0, s>>>(M, I, phase, it, skip_top, perturb_on, DevFactor{}, 0);
}
void launch_price_btran(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase, int it,
                        int skip_top, int perturb_on, int t, double* rho) {
  k_price<true><<<cap_grid(M.m, BS, I.part_cap), BS, mf_lds_bytes(t), s>>>(M, I, phase, it, skip_top, perturb_on, F,
                                                                         t);
  k_btran_out<<<nblk(M.m, BO_ROWS), BS, 0, s>>>(M, I, F, rho, 1, -1);
}
void launch_dualcheck(hipStream_t s, const DevModel& M, const DevIter& I, int it) {
  k_dualcheck<<<cap_grid(M.N - M.m, BS, I.part_cap), BS, 0, s>>>(M, I, it);
}
void launch_btran_unit(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t, double* out,
                       int p_fixed) {
  k_btran_prep<<<1, BS, mf_lds_bytes(t), s>>>(M, I, F, t, p_fixed, -1);
  k_btran_out<<<nblk(M.m, BO_ROWS), BS, 0, s>>>(M, I, F, out, p_fixed < 0 ? 1 : 0, -1);
}
void launch_btran_cand(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t, int cand) {
  k_btran_prep<<<1, BS, mf_lds_bytes(t), s>>>(M, I, F, t, -1, cand);
  k_btran_out<<<nblk(M.m, BO_ROWS), BS, 0, s>>>(M, I, F, I.rho_c + (size_t)cand * M.m, 1, cand);
}
void launch_spmv(hipStream_t s, const DevModel& M, const int* posN, const double* v, double alpha, double beta,
                 const double* cvec, double* out, int by_col, const DevScalars* S, int check, hipEvent_t ev0,
                 hipEvent_t ev1) {
  const int cols = spmv_cols(M);
  if (ev0 && ev1)
    hipExtLaunchKernelGGL(k_spmv, dim3(nblk(M.N, cols)), dim3(SPMV_THREADS), 0, s, ev0, ev1, 0, M, posN, v, alpha,
                          beta, cvec, out, by_col, S, check, cols);
  else
    k_spmv<<<nblk(M.N, cols), SPMV_THREADS, 0, s>>>(M, posN, v, alpha, beta, cvec, out, by_col, S, check, cols);
}
void launch_stage_spmv_gemv(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, const int* posN,
                            const double* rho, const double* cvec, double* alpha_r, double* tau0) {
  const int cols = spmv_cols(M);
  const int nb1 = nblk(M.N, cols);
  const int nb2 = F.k > 0 ? nblk((long)F.k * WAVE, BS) : 0;
  k_stage_spmv_gemv<<<nb1 + nb2, BS, 0, s>>>(M, posN, rho, cvec, alpha_r, I.S, cols, nb1, F, tau0);
}
void launch_stage_chuzc_slack(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase,
                              int it, const double* rho, double* tau0, int harris) {
  const int nb1 = cap_grid(M.N - M.m, BS, I.part_cap);
  if (!harris) {
    k_stage_chuzc_slack<<<nb1 + nblk(M.m, BS), BS, 0, s>>>(M, I, phase, it, nb1, F, rho, tau0);
    return;
  }
  k_stage_chuzc1_slack<<<nb1 + nblk(M.m, BS), BS, 0, s>>>(M, I, phase, it, nb1, F, rho, tau0);
  k_chuzc2<<<nb1, BS, 0, s>>>(M, I, phase, 1, it);
}
void launch_chuzc(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int harris, int it) {
  int g = cap_grid(M.N - M.m, BS, I.part_cap);
  if (!harris) {
    k_chuzc_tb<<<g, BS, 0, s>>>(M, I, phase, it);
    return;
  }
  k_chuzc1<<<g, BS, 0, s>>>(M, I, phase, harris, it);
  k_chuzc2<<<g, BS, 0, s>>>(M, I, phase, harris, it);
}
void launch_ftran_col(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int col, int t,
                      double* out, int check) {
  if (t > 0) {
    k_ftran_col_eta<<<nblk(M.m, BS), BS, mf_lds_bytes(t), s>>>(M, I, F, col, out, t, check);
  } else {
    k_ftran_col<<<nblk(M.m, BS), BS, 0, s>>>(M, I, F, col, out, check);
  }
}
void launch_ftran_dense(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, const double* b,
                        double* out, int t, int mode, int check) {
  // out (by position) = B^{-1} b; mode 1 scatters the result into x via basis.
  double* tmp = (mode == 0 && t == 0) ? out : I.x0;
  if (F.k > 0)
    k_gemv<<<nblk((long)F.k * WAVE, BS), BS, 0, s>>>(F.Kinv, F.k, F.k, b, F.row_R, tmp, F.pos_P, I.S, check);
  k_ftran_slack<<<nblk(M.m, BS), BS, 0, s>>>(M, F, b, tmp, I.S, check);
  if (t > 0 || mode == 1)
    k_ftran_out<<<nblk(M.m, BS), BS, mf_lds_bytes(t), s>>>(M, I, tmp, out, t, mode, check);
}
void launch_btran_dense(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, double* v,
                        double* y, int t) {
  // v (by position, overwritten) -> y (by row) = B^{-T} v
  if (t > 0) k_btran_dense_eta<<<1, 1024, 0, s>>>(M, I, v, t);
  if (F.k > 0) {
    k_btran_dense_rhs<<<nblk(F.k, BS), BS, 0, s>>>(F, v, I.wN);
    k_gemv<<<nblk((long)F.k * WAVE, BS), BS, 0, s>>>(F.KinvT, F.k, F.k, I.wN, nullptr, y, F.row_R, I.S, 0);
  }
  k_btran_dense_slack<<<nblk(M.m, BS), BS, 0, s>>>(M, F, v, y);
}
void launch_update(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it) {
  k_update<<<cap_grid(M.m, BS, I.part_cap), BS, 0, s>>>(M, I, phase, it);
}
void launch_dupdate_update(hipStream_t s, const DevModel& M, const DevIter& I, int phase, int it) {
  const int n = (M.N - M.m) > M.m ? (M.N - M.m) : M.m;
  k_dupdate_update<<<cap_grid(n, BS, I.part_cap), BS, 0, s>>>(M, I, phase, it);
}
// Above this many update rows the workgroups' redundant eta coefficients
// (each reads MF's triangle) cost more than the launches they save.
constexpr int UPDATE_ETA_MAX_N = 16384;
void launch_col_update(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int phase, int it,
                       int t, const double* tau0, int split) {
  const int n = (M.N - M.m) > M.m ? (M.N - M.m) : M.m;
  if (split || n > UPDATE_ETA_MAX_N) {
    const int nb1 = nblk(M.m, BS);
    k_stage_col_eta_out<<<nb1 + (t > 0 ? nb1 : 0), BS, mf_lds_bytes(t), s>>>(M, I, F, nb1, tau0, I.tau, t);
    launch_dupdate_update(s, M, I, phase, it);
    return;
  }
  k_dupdate_update_col<<<cap_grid(n, BS, I.part_cap), BS, mf_lds_bytes(t), s>>>(M, I, F, phase, it, tau0, t);
}
void launch_pprice(hipStream_t s, const DevModel& M, const DevIter& I, int it) {
  k_pprice<<<cap_grid(M.N - M.m, BS, I.part_cap), BS, 0, s>>>(M, I, it);
}
void launch_pratio_btran(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int it, int t,
                         double* rho, int split) {
  const int g = cap_grid(M.m, BS, I.part_cap);
  if (!split && M.m <= PRATIO_ONE_WG) {
    k_pratio_btran<<<g, BS, mf_lds_bytes(t), s>>>(M, I, it, F, t);
  } else {
    k_pratio1<<<g, BS, 0, s>>>(M, I, it);
    k_pratio2<true><<<g, BS, mf_lds_bytes(t), s>>>(M, I, it, F, t);
  }
  k_btran_out<<<nblk(M.m, BO_ROWS), BS, 0, s>>>(M, I, F, rho, 1, -1);
}
void launch_pupdate(hipStream_t s, const DevModel& M, const DevIter& I, int it) {
  const int n = (M.N - M.m) > M.m ? (M.N - M.m) : M.m;
  k_pupdate<<<cap_grid(n, BS, I.part_cap), BS, 0, s>>>(M, I, it);
}
void launch_bfrt(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int harris, int t) {
  k_bfrt<<<1, BFRT_BS, 0, s>>>(M, I, harris);
}
void launch_flips(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F, int t) {
  k_flip_detect<<<cap_grid(M.N - M.m, BS, I.part_cap), BS, 0, s>>>(M, I);
  k_flip_cc<<<1, BFRT_BS, 0, s>>>(M, I);
  launch_ftran_dense(s, M, I, F, I.bcc, I.dxb, t, 0, 2);
  const int n = (M.N - M.m) > M.m ? (M.N - M.m) : M.m;
  k_bfrt_xb<<<cap_grid(n, BS, I.part_cap), BS, 0, s>>>(M, I);
}
void launch_force(hipStream_t s, const DevModel& M, const DevIter& I, int event) {
  if (event == DS_FORCE_STABILITY) k_force_stability<<<1, 1, 0, s>>>(I);
  else if (event == DS_FORCE_DUAL_RESTART) k_force_dual<<<1, 1, 0, s>>>(M, I);
}
void launch_force_zero_column(hipStream_t s, double* W, int k, int c) {
  if (k > 0 && c >= 0 && c < k) k_force_zero_column<<<(k + 255) / 256, 256, 0, s>>>(W, k, c);
}
void launch_price_multi(hipStream_t s, const DevModel& M, const DevIter& I, int it, int skip_top, int perturb_on,
                        int nc) {
  k_price_multi<<<cap_grid(M.m, BS, I.part_cap), BS, 0, s>>>(M, I, it, skip_top, perturb_on, nc);
}
void launch_pick(hipStream_t s, const DevModel& M, const DevIter& I, int it, double psi) {
  k_pick<<<1, BS, 0, s>>>(M, I, it, psi);
}
void launch_rho_update(hipStream_t s, const DevModel& M, const DevIter& I, int nc) {
  k_rho_update<<<dim3(cap_grid(M.m, BS, 64), nc), BS, 0, s>>>(M, I);
}
void launch_set_primal(hipStream_t s, const DevModel& M, const DevIter& I) {
  k_set_primal<<<nblk(M.N - M.m, BS), BS, 0, s>>>(M, I);
}
void launch_mask_nonbasic(hipStream_t s, const DevModel& M, const DevIter& I, const double* src, double* wN) {
  k_mask_nonbasic<<<nblk(M.N, BS), BS, 0, s>>>(M, I, src, wN);
}
void launch_rowspmv(hipStream_t s, const DevModel& M, const double* w, const double* b, double alpha, double beta,
                    double* out) {
  k_rowspmv<<<nblk(M.m, BS), BS, 0, s>>>(M, w, b, alpha, beta, out);
}
void launch_dot_obj(hipStream_t s, const DevModel& M, const DevIter& I) {
  k_dot_obj<<<cap_grid(M.N, BS, I.part_cap), BS, 0, s>>>(M, I);
}
void launch_phase1_inf(hipStream_t s, const DevModel& M, const DevIter& I) {
  k_phase1_inf<<<cap_grid(M.N, BS, I.part_cap), BS, 0, s>>>(M, I);
}
void launch_gather_cb(hipStream_t s, const DevModel& M, const DevIter& I, double* out) {
  k_gather_cb<<<nblk(M.m, BS), BS, 0, s>>>(M, I, out);
}
void launch_dse_exact(hipStream_t s, const DevModel& M, const DevIter& I, const DevFactor& F) {
  k_dse_exact<<<M.m, BS, 0, s>>>(M, I, F);
}
void launch_fill(hipStream_t s, double* p, double v, int n) {
  if (n > 0) k_fill<<<nblk(n, BS), BS, 0, s>>>(p, v, n);
}

void launch_copy(hipStream_t s, const double2* a, double2* b, long long n) {
  k_copy<<<2048, 256, 0, s>>>(a, b, n);
}

}  // namespace pds
