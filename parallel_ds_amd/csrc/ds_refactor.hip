// Refactorization on gfx950: explicit inverse of the basis kernel K by
// blocked Gauss-Jordan elimination with partial pivoting.
//
// Per panel of NB columns: LU of the panel (one workgroup, partial
// pivoting) -> row swaps across the matrix -> the panel is saved and its
// columns reset to identity columns -> W = P_mid^{-1} * (mid rows) by two
// small triangular solves -> every other row gets the rank-NB update
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

// LU with partial pivoting of the panel W[j:k, j:j+nb] into Pf (ld NB).
__global__ __launch_bounds__(1024) void k_panel_lu(const double* W, int k, int j, int nb, double* Pf, int* piv,
                                                   int* flag) {
  const int R = k - j;
  for (int e = threadIdx.x; e < R * nb; e += 1024) {
    int r = e / nb, cc = e % nb;
    Pf[(size_t)r * NB + cc] = W[(size_t)(j + r) * k + j + cc];
  }
  __syncthreads();
  __shared__ int s_piv;
  for (int s = 0; s < nb; ++s) {
    double bv = -1.0;
    int bi = INT_MAX;
    for (int r = s + threadIdx.x; r < R; r += 1024) argmax_merge(bv, bi, fabs(Pf[(size_t)r * NB + s]), r);
    block_argmax<1024>(bv, bi);
    if (threadIdx.x == 0) {
      // No finite nonzero pivot (zero or NaN column): keep the row, flag the
      // factor as singular; every index stays inside the panel.
      if (bi == INT_MAX || !(bv > 0.0)) {
        if (bi == INT_MAX || bi < s || bi >= R) bi = s;
        *flag = 1;
      }
      s_piv = bi;
      piv[j + s] = j + bi;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != s)
      for (int cc = threadIdx.x; cc < nb; cc += 1024) {
        double a = Pf[(size_t)s * NB + cc];
        Pf[(size_t)s * NB + cc] = Pf[(size_t)pr * NB + cc];
        Pf[(size_t)pr * NB + cc] = a;
      }
    __syncthreads();
    const double dv = Pf[(size_t)s * NB + s];
    for (int r = s + 1 + threadIdx.x; r < R; r += 1024) Pf[(size_t)r * NB + s] = Pf[(size_t)r * NB + s] / dv;
    __syncthreads();
    const int w = nb - s - 1;
    if (w > 0)
      for (int e = threadIdx.x; e < (R - s - 1) * w; e += 1024) {
        int r = s + 1 + e / w, cc = s + 1 + e % w;
        Pf[(size_t)r * NB + cc] = Pf[(size_t)r * NB + cc] - Pf[(size_t)r * NB + s] * Pf[(size_t)s * NB + cc];
      }
    __syncthreads();
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

// Mid rows: X <- U^{-1} L^{-1} X for every column (one lane per column).
__global__ __launch_bounds__(64) void k_mid_solve(double* W, int k, int j, int nb, const double* Pf) {
  __shared__ double LU[NB][NB + 1];
  __shared__ double X[NB][64 + 1];
  const int lane = threadIdx.x;
  for (int e = lane; e < nb * nb; e += 64) LU[e / nb][e % nb] = Pf[(size_t)(e / nb) * NB + (e % nb)];
  const int c = blockIdx.x * 64 + lane;
  const bool live = c < k;
  for (int s = 0; s < nb; ++s) X[s][lane] = live ? W[(size_t)(j + s) * k + c] : 0.0;
  __syncthreads();
  for (int s = 1; s < nb; ++s) {
    double a = X[s][lane];
    for (int t = 0; t < s; ++t) a = a - LU[s][t] * X[t][lane];
    X[s][lane] = a;
  }
  for (int s = nb - 1; s >= 0; --s) {
    double a = X[s][lane];
    for (int t = s + 1; t < nb; ++t) a = a - LU[s][t] * X[t][lane];
    X[s][lane] = a / LU[s][s];
  }
  if (live)
    for (int s = 0; s < nb; ++s) W[(size_t)(j + s) * k + c] = X[s][lane];
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

// perm = composition of the column swaps applied in reverse order.
__global__ void k_perm(const int* piv, int* perm, int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int c = 0; c < k; ++c) perm[c] = c;
  for (int s = k - 1; s >= 0; --s) {
    int r = piv[s];
    int a = perm[s];
    perm[s] = perm[r];
    perm[r] = a;
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

void launch_gj_invert(hipStream_t s, int k, const GJWork& G, double* Kinv, double* KinvT) {
  if (k <= 0) return;
  for (int j = 0; j < k; j += NB) {
    int nb = (k - j < NB) ? (k - j) : NB;
    k_panel_lu<<<1, 1024, 0, s>>>(G.W, k, j, nb, G.Pf, G.piv, G.flag);
    k_swap_rows<<<(k + 255) / 256, 256, 0, s>>>(G.W, k, j, nb, G.piv);
    k_panel_save<<<(int)(((long)k * nb + 255) / 256), 256, 0, s>>>(G.W, k, j, nb, G.Pb);
    k_mid_solve<<<(k + 63) / 64, 64, 0, s>>>(G.W, k, j, nb, G.Pf);
    dim3 grid((k + 63) / 64, (k + 63) / 64);
    if (G.use_mfma)
      k_gemm_sub<<<grid, 256, 0, s>>>(G.W, k, k, k, G.Pb, G.W + (size_t)j * k, k, nb, j, j + nb);
    else
      k_gemm_sub_fma<<<grid, 256, 0, s>>>(G.W, k, k, k, G.Pb, G.W + (size_t)j * k, k, nb, j, j + nb);
  }
  k_perm<<<1, 1, 0, s>>>(G.piv, G.perm, k);
  dim3 g2((k + 31) / 32, (k + 31) / 32);
  k_unscramble<<<g2, 256, 0, s>>>(G.W, k, G.perm, Kinv, KinvT);
}

// Lane-layout check of the FP64 MFMA tile: C (64x64) -= P (64xNB) * B (NBx64).
void launch_gemm_tile_check(hipStream_t s, double* C, const double* P, const double* B, int nb, int use_mfma) {
  if (use_mfma) k_gemm_sub<<<dim3(1, 1), 256, 0, s>>>(C, 64, 64, 64, P, B, 64, nb, 0, 0);
  else k_gemm_sub_fma<<<dim3(1, 1), 256, 0, s>>>(C, 64, 64, 64, P, B, 64, nb, 0, 0);
}

}  // namespace pds
