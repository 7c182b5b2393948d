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

// Pivot rows of a panel by tournament: each workgroup factors up to TS_ROWS
// rows of the panel W[., j:j+nb] (their current values) in LDS with partial
// pivoting and keeps its nb pivot rows; rounds of merges (up to four lists
// per workgroup) reduce the lists to one.  Level 0 reads rows j..k-1
// directly (in_rows == nullptr).  The final workgroup also writes the LU of
// the winners' block (Pf), its inverse (Pinv) and the row exchanges
// piv[j..j+nb) that bring the winners to rows j..j+nb-1 in pivot order.
constexpr int TS_ROWS = 256;
constexpr int TS_THREADS = 1024;

// Pivot rows of a panel by tournament: each workgroup factors up to TS_ROWS
// rows of the panel W[., j:j+nb] (their current values) in LDS with partial
// pivoting and keeps its nb pivot rows; rounds of merges (up to four lists
// per workgroup) reduce the lists to one.  Level 0 reads rows j..k-1
// directly (in_rows == nullptr).  The final workgroup also writes the LU of
// the winners' block (Pf), its inverse (Pinv) and the row exchanges
// piv[j..j+nb) that bring the winners to rows j..j+nb-1 in pivot order.
__global__ __launch_bounds__(TS_THREADS) void k_tslu(const double* W, int k, int j, int nb, const int* in_rows,
                                                     const int* in_n, int in_groups, int* out_rows, int* out_n,
                                                     int final_, double* Pf, double* Pinv, int* piv, int* flag) {
  __shared__ double A[TS_ROWS][NB + 1];
  __shared__ int rid[TS_ROWS];
  __shared__ int goff[5];
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
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
    const int r0 = j + w * TS_ROWS;
    R = min(k, r0 + TS_ROWS) - r0;
    for (int t = tid; t < R; t += TS_THREADS) rid[t] = r0 + t;
  } else {
    R = goff[4];
    for (int t = tid; t < R; t += TS_THREADS) {
      int g = 0;
      while (t >= goff[g + 1]) ++g;
      rid[t] = in_rows[(4 * w + g) * NB + (t - goff[g])];
    }
  }
  __syncthreads();
  for (int e = tid; e < R * nb; e += TS_THREADS) {
    int r = e / nb, c = e % nb;
    A[r][c] = W[(size_t)rid[r] * k + j + c];
  }
  __syncthreads();
  const int steps = min(nb, R);
  for (int st = 0; st < steps; ++st) {
    double bv = -1.0;
    int bi = INT_MAX;
    for (int r = st + tid; r < R; r += TS_THREADS) {
      double a = fabs(A[r][st]);
      if (a >= 0.0) argmax_merge(bv, bi, a, r);
    }
    block_argmax<TS_THREADS>(bv, bi);
    if (bi == INT_MAX || !(bv > 0.0)) {
      // No nonzero pivot in this column among these rows (common for a
      // sparse kernel): leave the column, keep row st in place.  Only the
      // final round makes this an error.
      if (final_ && tid == 0) *flag = 1;
      __syncthreads();
      continue;
    }
    if (bi != st) {
      for (int c = tid; c < nb; c += TS_THREADS) {
        double x = A[st][c];
        A[st][c] = A[bi][c];
        A[bi][c] = x;
      }
      if (tid == 0) {
        int x = rid[st];
        rid[st] = rid[bi];
        rid[bi] = x;
      }
    }
    __syncthreads();
    const double dv = A[st][st];
    for (int r = st + 1 + tid; r < R; r += TS_THREADS) A[r][st] = A[r][st] / dv;
    __syncthreads();
    const int wdt = nb - st - 1;
    if (wdt > 0)
      for (int e = tid; e < (R - st - 1) * wdt; e += TS_THREADS) {
        int r = st + 1 + e / wdt, c = st + 1 + e % wdt;
        A[r][c] = A[r][c] - A[r][st] * A[st][c];
      }
    __syncthreads();
  }
  for (int t = tid; t < steps; t += TS_THREADS) out_rows[w * NB + t] = rid[t];
  if (tid == 0) out_n[w] = steps;
  if (!final_) return;
  if (steps < nb && tid == 0) *flag = 1;
  // Final round: LU factors, inverse of the winners' block, row exchanges.
  for (int e = tid; e < nb * nb; e += TS_THREADS) Pf[(size_t)(e / nb) * NB + (e % nb)] = A[e / nb][e % nb];
  __syncthreads();
  // Gauss-Jordan on [B | I], B = the winners' original block in pivot order
  // (same pivots as the LU above, so no further exchanges).
  double(*G)[NB + 1] = &A[NB];       // rows NB..2NB-1: B
  double(*H)[NB + 1] = &A[2 * NB];   // rows 2NB..3NB-1: I -> B^{-1}
  __shared__ double prow[2 * NB];
  __shared__ double fcol[NB];
  for (int e = tid; e < nb * nb; e += TS_THREADS) {
    int r = e / nb, c = e % nb;
    G[r][c] = W[(size_t)rid[r] * k + j + c];
    H[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int t = 0; t < nb; ++t) {
    const double pv = G[t][t];
    for (int c = tid; c < 2 * nb; c += TS_THREADS) prow[c] = (c < nb) ? G[t][c] : H[t][c - nb];
    for (int r = tid; r < nb; r += TS_THREADS) fcol[r] = G[r][t] / pv;
    __syncthreads();
    for (int e = tid; e < nb * 2 * nb; e += TS_THREADS) {
      int r = e / (2 * nb), c = e % (2 * nb);
      double x = (c < nb) ? G[r][c] : H[r][c - nb];
      x = (r == t) ? x / pv : x - fcol[r] * prow[c];
      if (c < nb) G[r][c] = x;
      else H[r][c - nb] = x;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += TS_THREADS) Pinv[(size_t)(e / nb) * NB + (e % nb)] = H[e / nb][e % nb];
  // Sequential exchanges: position j+s receives winner rid[s].  Touched
  // positions live in a small LDS map (position -> row now there), searched
  // by wave 0.
  __shared__ int key[2 * NB], val[2 * NB];
  __shared__ int nk;
  if (tid == 0) nk = 0;
  __syncthreads();
  if (tid < WAVE) {
    const int lane = tid;
    for (int s2 = 0; s2 < nb; ++s2) {
      const int target = rid[s2];
      const int pj = j + s2;
      const int n = nk;
      int fv = -1, fj = -1, fl = -1;
      for (int q = lane; q < n; q += WAVE) {
        if (val[q] == target) fv = q;
        if (key[q] == pj) fj = q;
      }
      for (int o = 32; o > 0; o >>= 1) {
        fv = max(fv, __shfl_xor(fv, o, WAVE));
        fj = max(fj, __shfl_xor(fj, o, WAVE));
      }
      const int loc = (fv >= 0) ? key[fv] : target;
      for (int q = lane; q < n; q += WAVE)
        if (key[q] == loc) fl = q;
      for (int o = 32; o > 0; o >>= 1) fl = max(fl, __shfl_xor(fl, o, WAVE));
      if (lane == 0) {
        const int cj = (fj >= 0) ? val[fj] : pj;
        piv[pj] = loc;
        int kj = fj, kl = fl, cnt = n;
        if (kj < 0) { key[cnt] = pj; kj = cnt++; }
        val[kj] = target;
        if (loc != pj) {
          if (kl < 0) { key[cnt] = loc; kl = cnt++; }
          val[kl] = cj;
        }
        nk = cnt;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
}

// Apply the panel's row swaps to every column.
__global__ __launch_bounds__(256) void k_swap_rows(double* W, int k, int j, int nb, const int* piv) {
  int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= k) return;
  for (int s = 0; s < nb; ++s) {
    int r = piv[j + s];
    if (r != j + s && r > j + s && r < k) {
      double a = W[(size_t)(j + s) * k + c];
      W[(size_t)(j + s) * k + c] = W[(size_t)r * k + c];
      W[(size_t)r * k + c] = a;
    }
  }
}

// Save the (swapped) panel columns and reset them to identity columns.
__global__ __launch_bounds__(256) void k_panel_save(double* W, int k, int j, int nb, double* Pb) {
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)k * nb) return;
  int r = (int)(e / nb), cc = (int)(e % nb);
  double* w = &W[(size_t)r * k + j + cc];
  Pb[(size_t)r * NB + cc] = *w;
  *w = (r == j + cc) ? 1.0 : 0.0;
}

// Mid rows: X <- Pinv X for every column, Pinv = inverse of the pivot block.
__global__ __launch_bounds__(256) void k_mid_gemm(double* W, int k, int j, int nb, const double* Pinv) {
  __shared__ double Pi[NB][NB + 1];
  __shared__ double X[NB][64 + 1];
  const int c0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < NB * NB; e += 256) {
    int r = e / NB, t = e % NB;
    Pi[r][t] = (r < nb && t < nb) ? Pinv[(size_t)r * NB + t] : 0.0;
  }
  for (int e = threadIdx.x; e < NB * 64; e += 256) {
    int r = e / 64, cc = e % 64;
    X[r][cc] = (r < nb && c0 + cc < k) ? W[(size_t)(j + r) * k + c0 + cc] : 0.0;
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
  for (int j = 0; j < k; j += NB) {
    int nb = (k - j < NB) ? (k - j) : NB;
    int groups = (k - j + TS_ROWS - 1) / TS_ROWS;
    int* in_rows = G.rowsA;
    int* in_n = G.nA;
    k_tslu<<<groups, TS_THREADS, 0, s>>>(G.W, k, j, nb, nullptr, nullptr, 0, in_rows, in_n, groups == 1 ? 1 : 0, G.Pf,
                                   G.Pinv, G.piv, G.flag);
    int* out_rows = G.rowsB;
    int* out_n = G.nB;
    while (groups > 1) {
      int g2 = (groups + 3) / 4;
      k_tslu<<<g2, TS_THREADS, 0, s>>>(G.W, k, j, nb, in_rows, in_n, groups, out_rows, out_n, g2 == 1 ? 1 : 0, G.Pf,
                                 G.Pinv, G.piv, G.flag);
      int* tr = in_rows;
      in_rows = out_rows;
      out_rows = tr;
      int* tn = in_n;
      in_n = out_n;
      out_n = tn;
      groups = g2;
    }
    k_swap_rows<<<(k + 255) / 256, 256, 0, s>>>(G.W, k, j, nb, G.piv);
    k_panel_save<<<(int)(((long)k * nb + 255) / 256), 256, 0, s>>>(G.W, k, j, nb, G.Pb);
    k_mid_gemm<<<(k + 63) / 64, 256, 0, s>>>(G.W, k, j, nb, G.Pinv);
    dim3 grid((k + 63) / 64, (k + 63) / 64);
    if (G.use_mfma)
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
