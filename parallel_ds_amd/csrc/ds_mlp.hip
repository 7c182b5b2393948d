// The warm-start network on gfx950: a per-variable MLP, BF16 operands on the
// matrix cores with FP32 accumulation, FP32 master weights.
//
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), h3 = relu(W3 h2 + b3),
//   z  = w4 . h3 + b4          (logit of "variable is basic at the optimum")
//
// Every product is an "NT" GEMM  C[M x N] = A[M x K] . B[N x K]^T  with both
// operands contiguous along K, so each MFMA fragment is one 16-byte load:
//   forward   H = X W^T        A = X  (samples x in),  B = W   (out x in)
//   input grad  G W            A = G  (samples x out), B = W^T (in x out)
//   weight grad G^T X          A = G^T (out x samples), B = X^T (in x samples)
// The forward pass stores activations in both orientations for that reason.
//
// v_mfma_f32_16x16x32_bf16 lane layout (measured on the MI355X): lane l holds
// A[l%16][8(l/16)+e] and B[8(l/16)+e][l%16] for e < 8, and D[4(l/16)+e][l%16]
// for e < 4.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ds_mlp.hpp"
#include "ds_reduce.hpp"

namespace pds {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf8 load8(const bf16_t* p) {
  uint4 u = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf8, u);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
  __bf16 b = (__bf16)f;  // round to nearest even
  return __builtin_bit_cast(bf16_t, b);
}

// 256 threads, a 64 x 64 tile of C per workgroup, 32 x 32 per wavefront
// (2 x 2 MFMA tiles).  M, N multiples of 32 (16-row tiles beyond M or N are
// skipped per wavefront), K a multiple of 32.  Workgroup z takes K columns
// [z kc, (z + 1) kc): with the FP32 epilogue and gridDim.z > 1 it writes the
// partial product of its slice to C32 + z M N (split K, summed by k_splitk_sum).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(const bf16_t* __restrict__ A, const bf16_t* __restrict__ B, int M,
                                                 int N, int K, int kc, const float* __restrict__ bias,
                                                 const bf16_t* __restrict__ mask, bf16_t* __restrict__ C16,
                                                 bf16_t* __restrict__ CT16, float* __restrict__ C32) {
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const int m0 = blockIdx.y * 64 + (wv / 2) * 32, n0 = blockIdx.x * 64 + (wv % 2) * 32;
  const int lr = lane % 16, lk = 8 * (lane / 16);
  bool mt[2], nt[2];
  for (int t = 0; t < 2; ++t) {
    mt[t] = m0 + 16 * t < M;
    nt[t] = n0 + 16 * t < N;
  }
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
  if (EPI == MLP_EPI_F32) C32 += (size_t)blockIdx.z * M * N;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf8 fa[2], fb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (mt[t]) fa[t] = load8(A + (size_t)(m0 + 16 * t + lr) * K + k0 + lk);
      if (nt[t]) fb[t] = load8(B + (size_t)(n0 + 16 * t + lr) * K + k0 + lk);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        if (mt[a] && nt[b]) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (!(mt[a] && nt[b])) continue;
      const int col = n0 + 16 * b + lr;
      const int row0 = m0 + 16 * a + 4 * (lane / 16);
      bf16_t t4[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = row0 + e;
        float v = acc[a][b][e];
        if (EPI == MLP_EPI_F32) {
          C32[(size_t)row * N + col] = v;
          continue;
        }
        if (EPI == MLP_EPI_BIAS_RELU) v = fmaxf(v + bias[col], 0.f);
        if (EPI == MLP_EPI_MASK) v = (bf2f(mask[(size_t)row * N + col]) > 0.f) ? v : 0.f;
        const bf16_t h = f2bf(v);
        C16[(size_t)row * N + col] = h;
        t4[e] = h;
      }
      if (EPI != MLP_EPI_F32 && CT16) {
        // four consecutive rows of one column: 8 contiguous bytes of C^T
        uint2 pk;
        pk.x = (unsigned)t4[0] | ((unsigned)t4[1] << 16);
        pk.y = (unsigned)t4[2] | ((unsigned)t4[3] << 16);
        *reinterpret_cast<uint2*>(CT16 + (size_t)col * M + row0) = pk;
      }
    }
}

// C[i] = sum over z < splits of P[z MN + i], in ascending z.
__global__ void k_splitk_sum(const float* __restrict__ P, int splits, int MN, float* __restrict__ C) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += P[(size_t)z * MN + i];
  C[i] = s;
}

// Output layer and loss.  One wavefront per sample: z = w4 . h3 + b4; with
// labels, dz = wt (sigmoid(z) - y) / S, the BCE-with-logits loss term, and
// the gradient into h3's pre-activation G3 = dz w4 (h3 > 0), in both
// orientations.
// Rows s >= Sv are padding: no loss, zero gradient.  Labels and weights of
// row s are y[rows[s]], wt[rows[s]] (or y[s], wt[s] without rows).
__global__ __launch_bounds__(256) void k_head(const bf16_t* __restrict__ H3, const float* __restrict__ w4,
                                              const float* __restrict__ b4, int S, int Sv, int W,
                                              const int* __restrict__ rows, const float* __restrict__ y,
                                              const float* __restrict__ wt, float inv_s, float* __restrict__ z,
                                              float* __restrict__ dz, float* __restrict__ loss_part,
                                              bf16_t* __restrict__ G3, bf16_t* __restrict__ G3T) {
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const int s = blockIdx.x * 4 + wv;
  __shared__ float sl[4];
  float lterm = 0.f;
  if (s < S) {
    float acc = 0.f;
    for (int o = lane; o < W; o += 64) acc += bf2f(H3[(size_t)s * W + o]) * w4[o];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    const float zs = acc + b4[0];
    if (lane == 0) z[s] = zs;
    if (y) {
      const bool live = s < Sv;
      const int r = live ? (rows ? rows[s] : s) : 0;
      const float ys = live ? y[r] : 0.f, ws = live ? (wt ? wt[r] : 1.f) : 0.f;
      const float sg = 1.f / (1.f + __expf(-zs));
      const float d = ws * (sg - ys) * inv_s;
      if (lane == 0) {
        dz[s] = d;
        lterm = ws * (fmaxf(zs, 0.f) - zs * ys + log1pf(__expf(-fabsf(zs))));
      }
      for (int o = lane; o < W; o += 64) {
        const float h = bf2f(H3[(size_t)s * W + o]);
        const bf16_t g = f2bf(h > 0.f ? d * w4[o] : 0.f);
        G3[(size_t)s * W + o] = g;
        G3T[(size_t)o * S + s] = g;
      }
    }
  }
  if (lane == 0) sl[wv] = lterm;
  __syncthreads();
  if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = sl[0] + sl[1] + sl[2] + sl[3];
}

// out[r] = sum_c scale(c) X[r][c] over rows of a K-contiguous matrix: one
// wavefront per row (bias gradients from G^T; the w4 gradient from H3^T with
// scale = dz).
__global__ __launch_bounds__(256) void k_rowsum(const bf16_t* __restrict__ X, int R, int Ccols,
                                                const float* __restrict__ scale, float* __restrict__ out) {
  const int lane = threadIdx.x % 64, r = blockIdx.x * 4 + threadIdx.x / 64;
  if (r >= R) return;
  float acc = 0.f;
  for (int c = lane; c < Ccols; c += 64) acc += bf2f(X[(size_t)r * Ccols + c]) * (scale ? scale[c] : 1.f);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) out[r] = acc;
}

__global__ void k_sum(const float* __restrict__ v, int n, float* __restrict__ out) {
  float acc = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += v[i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ float s[16];
  if (threadIdx.x % 64 == 0) s[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int k = 0; k < (int)blockDim.x / 64; ++k) t += s[k];
    *out = t;
  }
}

// Adam (Kingma & Ba, 2015) on the flat FP32 parameter vector.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies of a row-major FP32 weight matrix (out x in) and its transpose.
__global__ void k_refresh(const float* __restrict__ W, int O, int I, bf16_t* __restrict__ Wb,
                          bf16_t* __restrict__ WTb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= O * I) return;
  const int o = i / I, c = i - o * I;
  const bf16_t h = f2bf(W[i]);
  Wb[i] = h;
  WTb[(size_t)c * O + o] = h;
}

// FP32 rows (features) -> BF16 rows padded to Kp columns, standardised by
// (x - mu) / sd, plus the transpose (Kp x S) when XT is given.  Rows Sv..S-1
// are zero padding.
__global__ void k_pack(const float* __restrict__ F, const int* __restrict__ rows, int Sv, int S, int nf, int Kp,
                       const float* __restrict__ mu, const float* __restrict__ sd, bf16_t* __restrict__ X,
                       bf16_t* __restrict__ XT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S * Kp) return;
  const int s = i / Kp, c = i - s * Kp;
  float v = 0.f;
  if (c < nf && s < Sv) {
    const int r = rows ? rows[s] : s;
    if (r >= 0) v = (F[(size_t)r * nf + c] - mu[c]) / sd[c];
  }
  const bf16_t h = f2bf(v);
  X[i] = h;
  if (XT) XT[(size_t)c * S + s] = h;
}

// Fused inference: 64 samples per workgroup, hidden layers kept in LDS as
// BF16, weights read from L2.  Writes the logit of every sample.
__global__ __launch_bounds__(256) void k_infer(const bf16_t* __restrict__ X, int S, int Kp,
                                               const bf16_t* __restrict__ W1, const float* __restrict__ b1,
                                               const bf16_t* __restrict__ W2, const float* __restrict__ b2,
                                               const bf16_t* __restrict__ W3, const float* __restrict__ b3,
                                               const float* __restrict__ w4, const float* __restrict__ b4,
                                               float* __restrict__ z) {
  constexpr int W = MLP_WIDTH, T = 64, LDH = MLP_WIDTH + 8;  // padded rows against bank conflicts
  __shared__ bf16_t h[2][T * LDH];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const int s0 = blockIdx.x * T;
  const int lr = lane % 16, lk = 8 * (lane / 16);
  // One layer: out[T x W] = relu(in . Wl^T + bl); wavefront wv owns output
  // columns [64 wv, 64 wv + 64) for all 64 samples: 4 x 4 MFMA tiles.
  auto layer = [&](auto load_a, int K, const bf16_t* Wl, const float* bl, bf16_t* out) {
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
      bf8 fa[4], fb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        fa[t] = load_a(16 * t + lr, k0 + lk);
        fb[t] = load8(Wl + (size_t)(64 * wv + 16 * t + lr) * K + k0 + lk);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();  // everyone is done reading the buffer `out` may alias
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int col = 64 * wv + 16 * b + lr;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * a + 4 * (lane / 16) + e;
          out[row * LDH + col] = f2bf(fmaxf(acc[a][b][e] + bl[col], 0.f));
        }
      }
    __syncthreads();
  };
  auto from_x = [&](int r, int k) -> bf8 {
    const int s = s0 + r;
    if (s < S) return load8(X + (size_t)s * Kp + k);
    bf8 zero;
    for (int e = 0; e < 8; ++e) zero[e] = (__bf16)0.f;
    return zero;
  };
  auto from_h0 = [&](int r, int k) -> bf8 { return load8(&h[0][r * LDH + k]); };
  auto from_h1 = [&](int r, int k) -> bf8 { return load8(&h[1][r * LDH + k]); };
  layer(from_x, Kp, W1, b1, h[0]);
  layer(from_h0, W, W2, b2, h[1]);
  layer(from_h1, W, W3, b3, h[0]);
  // Output layer: 16 samples per wavefront.
  for (int r = wv * 16; r < wv * 16 + 16; ++r) {
    float acc = 0.f;
    for (int o = lane; o < W; o += 64) acc += bf2f(h[0][r * LDH + o]) * w4[o];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0 && s0 + r < S) z[s0 + r] = acc + b4[0];
  }
}

}  // namespace

void mlp_gemm_nt(hipStream_t st, int epi, const bf16_t* A, const bf16_t* B, int M, int N, int K, const float* bias,
                 const bf16_t* mask, bf16_t* C16, bf16_t* CT16, float* C32, hipEvent_t e0, hipEvent_t e1, float* ws,
                 size_t ws_floats) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  switch (epi) {
    case MLP_EPI_BIAS_RELU:
      hipExtLaunchKernelGGL(k_gemm_nt<MLP_EPI_BIAS_RELU>, grid, dim3(256), 0, st, e0, e1, 0, A, B, M, N, K, K, bias,
                            mask, C16, CT16, C32);
      break;
    case MLP_EPI_MASK:
      hipExtLaunchKernelGGL(k_gemm_nt<MLP_EPI_MASK>, grid, dim3(256), 0, st, e0, e1, 0, A, B, M, N, K, K, bias, mask,
                            C16, CT16, C32);
      break;
    default: {
      // Weight gradients (long K = samples, few output tiles): split K so that
      // the grid has about MLP_SPLITK_WGS workgroups, each slice a multiple of 32.
      const int tiles = (int)(grid.x * grid.y);
      int splits = (MLP_SPLITK_WGS + tiles - 1) / tiles;
      splits = std::min(splits, K / 32);
      while (splits > 1 && (size_t)splits * M * N > ws_floats) --splits;
      if (!ws || splits <= 1) {
        hipExtLaunchKernelGGL(k_gemm_nt<MLP_EPI_F32>, grid, dim3(256), 0, st, e0, e1, 0, A, B, M, N, K, K, bias, mask,
                              C16, CT16, C32);
        break;
      }
      const int kc = ((K + splits - 1) / splits + 31) / 32 * 32;
      splits = (K + kc - 1) / kc;
      grid.z = splits;
      hipExtLaunchKernelGGL(k_gemm_nt<MLP_EPI_F32>, grid, dim3(256), 0, st, e0, nullptr, 0, A, B, M, N, K, kc, bias,
                            mask, C16, CT16, ws);
      hipExtLaunchKernelGGL(k_splitk_sum, dim3((M * N + 255) / 256), dim3(256), 0, st, nullptr, e1, 0, ws, splits,
                            M * N, C32);
      break;
    }
  }
}

void mlp_head(hipStream_t st, const bf16_t* H3, const float* w4, const float* b4, int S, int Sv, const int* rows,
              const float* y, const float* wt, float inv_s, float* z, float* dz, float* loss_part, bf16_t* G3,
              bf16_t* G3T) {
  k_head<<<(S + 3) / 4, 256, 0, st>>>(H3, w4, b4, S, Sv, MLP_WIDTH, rows, y, wt, inv_s, z, dz, loss_part, G3, G3T);
}

void mlp_rowsum(hipStream_t st, const bf16_t* X, int R, int C, const float* scale, float* out) {
  k_rowsum<<<(R + 3) / 4, 256, 0, st>>>(X, R, C, scale, out);
}

void mlp_sum(hipStream_t st, const float* v, int n, float* out) { k_sum<<<1, 1024, 0, st>>>(v, n, out); }

void mlp_adam(hipStream_t st, float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2,
              float eps, float c1, float c2) {
  k_adam<<<(n + 255) / 256, 256, 0, st>>>(p, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
}

void mlp_refresh(hipStream_t st, const float* W, int O, int I, bf16_t* Wb, bf16_t* WTb) {
  k_refresh<<<(O * I + 255) / 256, 256, 0, st>>>(W, O, I, Wb, WTb);
}

void mlp_pack(hipStream_t st, const float* F, const int* rows, int Sv, int S, int nf, int Kp, const float* mu,
              const float* sd, bf16_t* X, bf16_t* XT) {
  k_pack<<<(S * Kp + 255) / 256, 256, 0, st>>>(F, rows, Sv, S, nf, Kp, mu, sd, X, XT);
}

void mlp_infer(hipStream_t st, const bf16_t* X, int S, int Kp, const bf16_t* W1, const float* b1, const bf16_t* W2,
               const float* b2, const bf16_t* W3, const float* b3, const float* w4, const float* b4, float* z,
               hipEvent_t e0, hipEvent_t e1) {
  hipExtLaunchKernelGGL(k_infer, dim3((S + 63) / 64), dim3(256), 0, st, e0, e1, 0, X, S, Kp, W1, b1, W2, b2, W3, b3,
                        w4, b4, z);
}

}  // namespace pds
