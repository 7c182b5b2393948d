// Refactorization on gfx950: explicit inverse of the basis kernel K by
// blocked Gauss-Jordan elimination with partial pivoting.
//
// Per panel of NB columns: pivot rows chosen by tournament pivoting (LDS LU
// of row chunks, merged in rounds) -> row swaps across the matrix -> the
// panel is saved and its columns reset to identity columns -> W = P_mid^{-1}
// * (mid rows) with the explicit inverse of the pivot block -> every other
// row gets the rank-NB update
// rows -= P_saved * W, an FP64 MFMA GEMM (v_mfma_f64_16x16x4f64).  After the
// last panel the column permutation is undone while Kinv and its transpose
// are written.  Total work 2k^3 flops, all but O(k^2 NB) in the MFMA update.
#include "ds_dev.hpp"
#include "ds_kernels.hpp"

#include <hip/hip_ext.h>
#include "ds_reduce.hpp"

namespace pds {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NB = GJ_NB;

// Dense kernel K[i][c] = A[row_R[i], basis[pos_P[c]]] into zeroed W.
__global__ __launch_bounds__(256) void k_build_kernel(DevModel M, const int* basis, DevFactor F, double* W) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= F.k) return;
  int j = basis[F.pos_P[c]];
  for (int e = M.cptr[j]; e < M.cptr[j + 1]; ++e) {
    int ri = F.Ridx_of_row[M.ridx[e]];
    if (ri >= 0) W[(size_t)ri * F.k + c] = M.cval[e];
  }
}

// Pivot rows of a panel by tournament: each workgroup eliminates up to
// TS_ROWS rows of the panel W[., j:j+nb] (their current values) with partial
// pivoting and keeps its nb pivot rows in pivot order; rounds of merges (up to
// four lists per workgroup) reduce the lists to one.  Level 0 reads rows
// j..k-1 directly (in_rows == nullptr).  The final workgroup also writes the
// inverse of the winners' block (Pinv) and the row exchanges piv[j..j+nb)
// that bring the winners to rows j..j+nb-1 in pivot order.
//
// Layout: lane = row, a[x] = the row's panel column x (registers).  Steps run
// in blocks of TS_BLK fully unrolled, and the register frame shifts by TS_BLK
// columns after each block, so every register index is a constant.  A step:
// wave max of |a[s]| over live rows, ballot for the lowest such row, the
// wave's candidate lane writes its row to LDS (double-buffered by step
// parity), one barrier, and every lane subtracts its multiple of the winner's
// row.  Only the winners are used (the panel itself is eliminated by
// k_mid_gemm / k_gemm_sub with Pinv), so the multipliers use the pivot's
// reciprocal and pivoted rows are not kept.
constexpr int TS_ROWS = 256;  // rows per group of the multi-round tournament (4 waves)
constexpr int TS_BLK = 16;
constexpr double SING_REL = 1e-13;  // relative pivot below which the kernel counts as singular

// A final-round pivot step without a usable pivot: the singular flag, and the
// step (kernel column) with the row the tournament chose for it (its position
// in W before this panel's exchanges), which the host uses to substitute that
// row's slack for the column.
__device__ inline void mark_dependent(int* flag, int step, int row) {
  flag[0] = 1;
  const int c = flag[1];
  if (c < GJ_DEP_CAP) {
    flag[2 + 2 * c] = step;
    flag[3 + 2 * c] = row;
  }
  flag[1] = c + 1;
}
static_assert(NB == WAVE, "one pivot-block row per lane");
static_assert(NB % TS_BLK == 0, "unrolled blocks tile the panel");

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform_d(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
// DPP lane permutation of a double (both halves, every lane has a source).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// Wave maximum: quad_perm [1,0,3,2] and [2,3,0,1], row_ror 4 and 8 give each
// row of 16 lanes its maximum; the four rows are combined from lanes 0/16/32/48.
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_d<0xB1>(v));
  v = fmax(v, dpp_d<0x4E>(v));
  v = fmax(v, dpp_d<0x124>(v));
  v = fmax(v, dpp_d<0x128>(v));
  return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)), fmax(readlane_d(v, 32), readlane_d(v, 48)));
}
// 1/p from the hardware reciprocal and one Newton step (pivot choice only).
__device__ __forceinline__ double recip_nr(double p) {
  double r = __builtin_amdgcn_rcp(p);
  return fma(r, fma(-p, r, 1.0), r);
}

// WV waves, 64 WV rows per group. WV = 4 stages the rows in an LDS tile. WV = 8
// (one group of up to 512 rows, a single round) loads them directly and
// reads the winners' block for the Gauss-Jordan from W.
template <int WV>
__global__ __launch_bounds__(WV * WAVE) void k_tslu(const double* W, int k, int j, int nb, const int* in_rows,
                                                     const int* in_n, int in_groups, int* out_rows, int* out_n,
                                                     int final_, double* Pinv, int* piv, int* flag,
                                                     const double* rowscale) {
  constexpr int TS_WAVES = WV;
  constexpr int TS_THREADS = WV * WAVE;
  constexpr int ROWS = WV * WAVE;
  constexpr int GJ_COLS = 2 * NB / TS_WAVES;  // [B | I] columns per wave in the final Gauss-Jordan
  constexpr bool TILE = WV == 4;
  static_assert(NB % GJ_COLS == 0, "unrolled blocks tile the panel");
  __shared__ int rid[ROWS];
  __shared__ int goff[5];
  __shared__ double cand[2][TS_WAVES][NB];
  __shared__ double cv[2][TS_WAVES];
  __shared__ int ci[2][TS_WAVES];
  __shared__ double cs[2][TS_WAVES];  // the candidate row's scale (largest |entry| of its panel row, as loaded)
  __shared__ int win[NB];
  __shared__ int wrow[NB];
  __shared__ unsigned char depf[NB];  // final round: the step had no usable pivot
  __shared__ double tile[TILE ? ROWS : 1][NB + 1];  // the group's rows of the panel, as loaded
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  const int wv = uniform_i(tid / WAVE), lane = tid % WAVE;
  if (tid == 0) {
    int n = 0;
    goff[0] = 0;
    for (int g = 0; g < 4; ++g) {
      int gi = 4 * w + g;
      n += (in_rows && gi < in_groups) ? in_n[gi] : 0;
      goff[g + 1] = n;
    }
  }
  __syncthreads();
  int R;
  if (!in_rows) {
    const int r0 = j + w * ROWS;
    R = min(k, r0 + ROWS) - r0;
    for (int t = tid; t < R; t += TS_THREADS) rid[t] = r0 + t;
  } else {
    R = goff[4];
    for (int t = tid; t < R; t += TS_THREADS) {
      int g = 0;
      while (t >= goff[g + 1]) ++g;
      rid[t] = in_rows[(4 * w + g) * NB + (t - goff[g])];
    }
  }
  R = uniform_i(R);
  __syncthreads();
  bool live = tid < R;
  double a[NB];
  if constexpr (TILE) {
    // Coalesced load (lane = column) into LDS, then lane = row into registers.
#pragma unroll 8
    for (int r = wv; r < R; r += TS_WAVES) tile[r][lane] = lane < nb ? W[(size_t)rid[r] * k + j + lane] : 0.0;
    __syncthreads();
#pragma unroll
    for (int x = 0; x < NB; ++x) a[x] = live ? tile[tid][x] : 0.0;
  } else {
    const double* src = W + (size_t)(live ? rid[tid] : 0) * k + j;
#pragma unroll
    for (int x = 0; x < NB; ++x) a[x] = (live && x < nb) ? src[x] : 0.0;
  }
  // Row scale for the final round's singularity test: the row's largest
  // |entry| in K before any elimination (a dependent row shrinks to rounding
  // level in every later panel).
  const double rmax = live ? rowscale[rid[tid]] : 0.0;
  const int steps = min(nb, R);
  for (int cb = 0; cb * TS_BLK < steps; ++cb) {
#pragma unroll
    for (int s = 0; s < TS_BLK; ++s) {
      const int st = cb * TS_BLK + s;
      if (st < steps) {
        const int b = st & 1;
        // This wave's candidate: the live row with the largest |a[s]|, lowest
        // row on ties (a NaN counts as 0, so a live row always qualifies).
        double v = fabs(a[s]);
        v = live ? (v == v ? v : 0.0) : -1.0;
        const double m = uniform_d(wave_max(v));
        const unsigned long long bal = __ballot(live && v == m);
        const int cl = bal ? __ffsll((long long)bal) - 1 : -1;
        if (lane == cl) {
#pragma unroll
          for (int x = s; x < NB; ++x) cand[b][wv][x] = a[x];
        }
        const double crow = cl >= 0 ? readlane_d(rmax, cl) : 0.0;
        if (lane == 0) {
          cv[b][wv] = m;
          ci[b][wv] = cl >= 0 ? wv * WAVE + cl : INT_MAX;
          cs[b][wv] = crow;
        }
        __syncthreads();
        // The waves' candidates: every LDS load first, then the comparisons.
        double cvq[TS_WAVES];
        int ciq[TS_WAVES];
#pragma unroll
        for (int q = 0; q < TS_WAVES; ++q) {
          cvq[q] = cv[b][q];
          ciq[q] = ci[b][q];
        }
        double gv = -2.0;
        int gr = INT_MAX;
#pragma unroll
        for (int q = 0; q < TS_WAVES; ++q) {
          const bool better = cvq[q] > gv || (cvq[q] == gv && ciq[q] < gr);
          gv = better ? cvq[q] : gv;
          gr = better ? ciq[q] : gr;
        }
        gv = uniform_d(gv);
        gr = uniform_i(gr);
        // steps <= R, so some row is live and gr is a row of this group.
        if (tid == 0) win[st] = gr;
        if (tid == gr) live = false;
        // Final round: a pivot at or below SING_REL of its row's scale in K (or no
        // nonzero pivot) marks the column dependent: rounding leaves dependent rows
        // near, not at, zero.  The step then proceeds as if the column were the
        // unit vector of the chosen row (multipliers 0), so that later columns are
        // judged on the remaining rows and not on rounding noise.
        const bool dep = final_ && (gv <= 0.0 || gv <= SING_REL * cs[b][gr / WAVE]);
        if (tid == 0) {
          depf[st] = dep ? 1 : 0;
          if (dep) mark_dependent(flag, j + st, rid[gr]);
        }
        if (gv > 0.0 && !dep) {
          const double* pr = cand[b][gr / WAVE];
          const double l = a[s] * recip_nr(pr[s]);
          // The winner's row in register batches of NB / 4: each batch's LDS
          // loads issue back to back before its FMAs, instead of one or two
          // ahead of them.  Columns at or before the pivot column in a batch
          // are updated too: they were eliminated earlier in this block, are
          // never read again, and their LDS entries hold finite rows of
          // earlier steps.
          constexpr int H = NB / 4;
#pragma unroll
          for (int h = 0; h < NB / H; ++h) {
            if ((h + 1) * H <= s + 1) continue;
            double pv[H];
#pragma unroll
            for (int y = 0; y < H; ++y) pv[y] = pr[h * H + y];
#pragma unroll
            for (int y = 0; y < H; ++y) a[h * H + y] = a[h * H + y] - l * pv[y];
          }
        }
      }
    }
#pragma unroll
    for (int x = 0; x < NB - TS_BLK; ++x) a[x] = a[x + TS_BLK];
#pragma unroll
    for (int x = NB - TS_BLK; x < NB; ++x) a[x] = 0.0;
  }
  __syncthreads();
  for (int t = tid; t < steps; t += TS_THREADS) {
    wrow[t] = rid[win[t]];
    out_rows[w * NB + t] = rid[win[t]];
  }
  if (tid == 0) out_n[w] = steps;
  if (!final_) return;
  if (steps < nb && tid == 0) *flag = 1;
  __syncthreads();
  // Gauss-Jordan on [B | I], B = the winners' original block in pivot order
  // (the tournament's pivots, so no further exchanges).  Lane = row s; wave
  // wv holds columns GJ_COLS*wv + x of [B | I].  Column t is in wave
  // t / GJ_COLS, register t % GJ_COLS, a constant inside the unrolled block.
  __shared__ double colt[2][NB];
  __shared__ double rowt[2][2 * NB];
  double g[GJ_COLS];
  {
    const int sl = lane < steps ? win[lane] : 0;
#pragma unroll
    for (int x = 0; x < GJ_COLS; ++x) {
      const int c = GJ_COLS * wv + x;
      double v = 0.0;
      if (lane < steps) {
        if (c < NB && c < steps && depf[c]) v = (c == lane) ? 1.0 : 0.0;  // dependent: unit column
        else if (c < NB) {
          if constexpr (TILE) v = tile[sl][c];
          else v = c < nb ? W[(size_t)rid[sl] * k + j + c] : 0.0;
        }
        else v = (c - NB == lane) ? 1.0 : 0.0;
      }
      g[x] = v;
    }
  }
  for (int tb = 0; tb * GJ_COLS < nb; ++tb) {
#pragma unroll
    for (int x = 0; x < GJ_COLS; ++x) {
      const int t = tb * GJ_COLS + x;
      if (t < nb) {
        const int b = t & 1;
        if (wv == tb) colt[b][lane] = g[x];
        if (lane == t) {
#pragma unroll
          for (int y = 0; y < GJ_COLS; ++y) rowt[b][GJ_COLS * wv + y] = g[y];
        }
        __syncthreads();
        // Row t: f = 1 and pc == g[y], so the first FMA gives exactly 0 and
        // the second pc / pv (as pc * rp); other rows add 0.
        const double pv = rowt[b][t];
        const double rp = 1.0 / pv;
        const bool own = lane == t;
        const double f = own ? 1.0 : colt[b][lane] * rp;
        const double wt = own ? rp : 0.0;
#pragma unroll
        for (int y = 0; y < GJ_COLS; ++y) {
          const double pc = rowt[b][GJ_COLS * wv + y];
          g[y] = fma(wt, pc, fma(-f, pc, g[y]));
        }
      }
    }
  }
#pragma unroll
  for (int x = 0; x < GJ_COLS; ++x) {
    const int c = GJ_COLS * wv + x - NB;
    if (c >= 0 && c < nb && lane < nb) Pinv[(size_t)lane * NB + c] = g[x];
  }
  // Sequential exchanges: position j+s2 receives winner s2.  Lane i tracks
  // winner i's current position (wh) and the row now at position j+i (oc).
  if (wv == 0) {
    const int myrow = lane < nb ? wrow[lane] : -1;
    int wh = myrow;
    int oc = j + lane;
    for (int s2 = 0; s2 < nb; ++s2) {
      const int loc = __builtin_amdgcn_readlane(wh, s2);
      const int occ = __builtin_amdgcn_readlane(oc, s2);
      if (lane == 0) piv[j + s2] = loc;
      // the row at j+s2 moves to loc; if it is a later winner, so does its position
      if (lane > s2 && lane < nb && myrow == occ) wh = loc;
      if (lane == loc - j) oc = occ;
      if (lane == s2) oc = myrow;
    }
  }
}

// Dependent columns of panel j (final round): column step := unit vector of its
// chosen row in W (before the panel's exchanges), matching the tournament's
// treatment, so the trailing update leaves the other columns as they would be
// for a basis with that row's slack in place of the column.
// Thread per column c of W: a column that the tournament found dependent in
// this panel becomes the unit column of the row it was given; then the
// panel's row exchanges, in order (rowscale follows them in thread 0).
__global__ __launch_bounds__(256) void k_swap_rows(double* W, int k, int j, int nb, const int* piv,
                                                   double* rowscale, const int* flag) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= k) return;
  const int n = min(flag[1], GJ_DEP_CAP);
  for (int d = 0; d < n; ++d) {
    if (flag[2 + 2 * d] != c || c < j || c >= j + nb) continue;
    const int row = flag[3 + 2 * d];
    for (int i = 0; i < k; ++i) W[(size_t)i * k + c] = (i == row) ? 1.0 : 0.0;
  }
  for (int s = 0; s < nb; ++s) {
    int r = piv[j + s];
    if (r != j + s && r > j + s && r < k) {
      double a = W[(size_t)(j + s) * k + c];
      W[(size_t)(j + s) * k + c] = W[(size_t)r * k + c];
      W[(size_t)r * k + c] = a;
      if (c == 0) {
        double g = rowscale[j + s];
        rowscale[j + s] = rowscale[r];
        rowscale[r] = g;
      }
    }
  }
}
// rowscale[r] = max_c |W[r][c]| (one wave per row).
__global__ __launch_bounds__(256) void k_row_scale(const double* W, int k, double* rowscale) {
  const int r = (blockIdx.x * 256 + threadIdx.x) / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (r >= k) return;
  double v = 0.0;
  for (int c = lane; c < k; c += WAVE) v = fmax(v, fabs(W[(size_t)r * k + c]));
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  if (lane == 0) rowscale[r] = v;
}

// Mid rows: the pivot rows j..j+nb of W become Pinv times themselves, their panel
// columns taken as the identity; and the panel (columns j..j+nb) of the other
// rows moves to Pb, leaving zeros (the pivot rows of Pb are zero: the
// trailing update skips them).  The two parts touch disjoint entries of W.
__global__ __launch_bounds__(256) void k_mid_gemm(double* W, int k, int j, int nb, const double* Pinv, double* Pb) {
  __shared__ double Pi[NB][NB + 1];
  __shared__ double X[NB][64 + 1];
  const int c0 = blockIdx.x * 64;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)k * nb; e += (long)gridDim.x * 256) {
    const int r = (int)(e / nb), cc = (int)(e % nb);
    if (r >= j && r < j + nb) {
      Pb[(size_t)r * NB + cc] = 0.0;
    } else {
      double* w = &W[(size_t)r * k + j + cc];
      Pb[(size_t)r * NB + cc] = *w;
      *w = 0.0;
    }
  }
  for (int e = threadIdx.x; e < NB * NB; e += 256) {
    int r = e / NB, t = e % NB;
    Pi[r][t] = (r < nb && t < nb) ? Pinv[(size_t)r * NB + t] : 0.0;
  }
  for (int e = threadIdx.x; e < NB * 64; e += 256) {
    int r = e / 64, cc = e % 64;
    const int c = c0 + cc;
    X[r][cc] = (r < nb && c < k) ? ((c >= j && c < j + nb) ? (c - j == r ? 1.0 : 0.0) : W[(size_t)(j + r) * k + c])
                                 : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  double acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int t = 0; t < nb; ++t) {
    double pv[4], xv[4];
    for (int a = 0; a < 4; ++a) pv[a] = Pi[tr + 16 * a][t];
    for (int b = 0; b < 4; ++b) xv[b] = X[t][tc + 16 * b];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(pv[a], xv[b], acc[a][b]);
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      int r = tr + 16 * a, c = c0 + tc + 16 * b;
      if (r < nb && c < k) W[(size_t)(j + r) * k + c] = acc[a][b];
    }
}

// C[r][c] -= sum_t P[r][t] * B[t][c] for rows outside [skip0, skip1).
// 64x64 output tile per workgroup; each wave owns a 32x32 quadrant as 2x2
// MFMA 16x16 blocks, accumulating over NB in steps of 4.
// Lane layout of v_mfma_f64_16x16x4f64 (measured on gfx950): A[i][kk] in
// lane i + 16*kk, B[kk][j] in lane j + 16*kk, D[i][j] in lane j + 16*(i%4),
// element i/4.
__global__ __launch_bounds__(256) void k_gemm_sub(double* C, int ldc, int rows, int cols, const double* P,
                                                  const double* B, int ldb, int nb, int skip0, int skip1) {
  __shared__ double As[64][NB + 1];
  __shared__ double Bs[NB][64 + 1];
  const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * NB; e += 256) {
    int r = e / NB, t = e % NB;
    int gr = row0 + r;
    As[r][t] = (gr < rows && t < nb) ? P[(size_t)gr * NB + t] : 0.0;
  }
  for (int e = threadIdx.x; e < NB * 64; e += 256) {
    int t = e / 64, cc = e % 64;
    int gc = col0 + cc;
    Bs[t][cc] = (t < nb && gc < cols) ? B[(size_t)t * ldb + gc] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int li = lane % 16, lk = lane / 16;
  for (int kk = 0; kk < NB; kk += 4) {
    double av[2], bv[2];
    for (int a = 0; a < 2; ++a) av[a] = As[wr + a * 16 + li][kk + lk];
    for (int b = 0; b < 2; ++b) bv[b] = Bs[kk + lk][wc + b * 16 + li];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int e = 0; e < 4; ++e) {
        int gr = row0 + wr + a * 16 + lk + 4 * e;
        int gc = col0 + wc + b * 16 + li;
        if (gr < rows && gc < cols && (gr < skip0 || gr >= skip1)) C[(size_t)gr * ldc + gc] -= acc[a][b][e];
      }
}

// Same update with plain FP64 FMAs (16x16 threads, 4x4 outputs each).
__global__ __launch_bounds__(256) void k_gemm_sub_fma(double* C, int ldc, int rows, int cols, const double* P,
                                                      const double* B, int ldb, int nb, int skip0, int skip1) {
  __shared__ double As[64][NB + 1];
  __shared__ double Bs[NB][64 + 1];
  const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * NB; e += 256) {
    int r = e / NB, t = e % NB;
    int gr = row0 + r;
    As[r][t] = (gr < rows && t < nb) ? P[(size_t)gr * NB + t] : 0.0;
  }
  for (int e = threadIdx.x; e < NB * 64; e += 256) {
    int t = e / 64, cc = e % 64;
    int gc = col0 + cc;
    Bs[t][cc] = (t < nb && gc < cols) ? B[(size_t)t * ldb + gc] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  double acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int t = 0; t < nb; ++t) {
    double av[4], bv[4];
    for (int a = 0; a < 4; ++a) av[a] = As[tr + 16 * a][t];
    for (int b = 0; b < 4; ++b) bv[b] = Bs[t][tc + 16 * b];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      int gr = row0 + tr + 16 * a, gc = col0 + tc + 16 * b;
      if (gr < rows && gc < cols && (gr < skip0 || gr >= skip1)) C[(size_t)gr * ldc + gc] -= acc[a][b];
    }
}

// Kinv[r][c] = W[r][perm[c]] and KinvT[c][r] = the same value (tiled transpose).
__global__ __launch_bounds__(256) void k_unscramble(const double* W, int k, const int* perm, double* Kinv,
                                                    double* KinvT) {
  __shared__ double tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int rr = ty; rr < 32; rr += 8) {
    int r = r0 + rr, c = c0 + tx;
    double v = 0.0;
    if (r < k && c < k) {
      v = W[(size_t)r * k + perm[c]];
      Kinv[(size_t)r * k + c] = v;
    }
    tile[rr][tx] = v;
  }
  __syncthreads();
  for (int cc = ty; cc < 32; cc += 8) {
    int c = c0 + cc, r = r0 + tx;
    if (r < k && c < k) KinvT[(size_t)c * k + r] = tile[tx][cc];
  }
}

void launch_build_kernel(hipStream_t s, const DevModel& M, const int* basis, const DevFactor& F, double* W) {
  if (F.k <= 0) return;
  launch_fill(s, W, 0.0, F.k * F.k);
  k_build_kernel<<<(F.k + 255) / 256, 256, 0, s>>>(M, basis, F, W);
}

void launch_gj_panels(hipStream_t s, int k, const GJWork& G) {
  if (k <= 0) return;
  k_row_scale<<<(int)(((long)k * WAVE + 255) / 256), 256, 0, s>>>(G.W, k, G.rowscale);
  for (int j = 0; j < k; j += NB) {
    int nb = (k - j < NB) ? (k - j) : NB;
    int groups = (k - j + TS_ROWS - 1) / TS_ROWS;
    int* in_rows = G.rowsA;
    int* in_n = G.nA;
    if (groups == 2) {
      // 257..512 rows: one 8-wave group in a single round instead of two groups and a merge.
      k_tslu<8><<<1, 8 * WAVE, 0, s>>>(G.W, k, j, nb, nullptr, nullptr, 0, in_rows, in_n, 1, G.Pinv, G.piv, G.flag,
                                        G.rowscale);
      groups = 1;
    } else {
      k_tslu<4><<<groups, 4 * WAVE, 0, s>>>(G.W, k, j, nb, nullptr, nullptr, 0, in_rows, in_n, groups == 1 ? 1 : 0,
                                            G.Pinv, G.piv, G.flag, G.rowscale);
    }
    int* out_rows = G.rowsB;
    int* out_n = G.nB;
    while (groups > 1) {
      int g2 = (groups + 3) / 4;
      k_tslu<4><<<g2, 4 * WAVE, 0, s>>>(G.W, k, j, nb, in_rows, in_n, groups, out_rows, out_n, g2 == 1 ? 1 : 0,
                                        G.Pinv, G.piv, G.flag, G.rowscale);
      int* tr = in_rows;
      in_rows = out_rows;
      out_rows = tr;
      int* tn = in_n;
      in_n = out_n;
      out_n = tn;
      groups = g2;
    }
    k_swap_rows<<<(k + 255) / 256, 256, 0, s>>>(G.W, k, j, nb, G.piv, G.rowscale, G.flag);
    k_mid_gemm<<<(k + 63) / 64, 256, 0, s>>>(G.W, k, j, nb, G.Pinv, G.Pb);
    dim3 grid((k + 63) / 64, (k + 63) / 64);
    if (G.use_mfma && G.gemm_ev)
      hipExtLaunchKernelGGL(k_gemm_sub, grid, dim3(256), 0, s, G.gemm_ev[2 * (j / NB)], G.gemm_ev[2 * (j / NB) + 1], 0,
                            G.W, k, k, k, (const double*)G.Pb, (const double*)(G.W + (size_t)j * k), k, nb, j, j + nb);
    else if (G.use_mfma)
      k_gemm_sub<<<grid, 256, 0, s>>>(G.W, k, k, k, G.Pb, G.W + (size_t)j * k, k, nb, j, j + nb);
    else
      k_gemm_sub_fma<<<grid, 256, 0, s>>>(G.W, k, k, k, G.Pb, G.W + (size_t)j * k, k, nb, j, j + nb);
  }
}

void launch_gj_unscramble(hipStream_t s, int k, const GJWork& G, double* Kinv, double* KinvT) {
  if (k <= 0) return;
  dim3 g2((k + 31) / 32, (k + 31) / 32);
  k_unscramble<<<g2, 256, 0, s>>>(G.W, k, G.perm, Kinv, KinvT);
}

// Lane-layout check of the FP64 MFMA tile: C (64x64) -= P (64xNB) * B (NBx64).
void launch_gemm_tile_check(hipStream_t s, double* C, const double* P, const double* B, int nb, int use_mfma) {
  if (use_mfma) k_gemm_sub<<<dim3(1, 1), 256, 0, s>>>(C, 64, 64, 64, P, B, 64, nb, 0, 0);
  else k_gemm_sub_fma<<<dim3(1, 1), 256, 0, s>>>(C, 64, 64, 64, P, B, 64, nb, 0, 0);
}

}  // namespace pds
