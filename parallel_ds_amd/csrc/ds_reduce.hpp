// Wavefront (64-lane) and workgroup reductions used by the iteration kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>

namespace pds {

constexpr int WAVE = 64;

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ inline int wave_sumi(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
  return v;
}

// (value, index) pairs: larger value wins, equal values go to the lower index.
__device__ inline void argmax_merge(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}
__device__ inline void wave_argmax(double& v, int& i) {
  for (int o = 32; o > 0; o >>= 1) {
    double v2 = __shfl_xor(v, o, WAVE);
    int i2 = __shfl_xor(i, o, WAVE);
    argmax_merge(v, i, v2, i2);
  }
}

// Workgroup reductions; BS threads, result valid in every thread.
template <int BS>
__device__ inline void block_argmax(double& v, int& i) {
  __shared__ double sv[BS / WAVE];
  __shared__ int si[BS / WAVE];
  wave_argmax(v, i);
  int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) { sv[w] = v; si[w] = i; }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int k = 1; k < BS / WAVE; ++k) argmax_merge(v, i, sv[k], si[k]);
}
template <int BS>
__device__ inline double block_sum(double v) {
  __shared__ double sv[BS / WAVE];
  v = wave_sum(v);
  int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) sv[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < BS / WAVE; ++k) r += sv[k];
  return r;
}
template <int BS>
__device__ inline int block_sumi(int v) {
  __shared__ int sv[BS / WAVE];
  v = wave_sumi(v);
  int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) sv[w] = v;
  __syncthreads();
  int r = 0;
  for (int k = 0; k < BS / WAVE; ++k) r += sv[k];
  return r;
}
template <int BS>
__device__ inline double block_min(double v) {
  __shared__ double sv[BS / WAVE];
  v = wave_min(v);
  int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  __syncthreads();
  if (l == 0) sv[w] = v;
  __syncthreads();
  double r = sv[0];
  for (int k = 1; k < BS / WAVE; ++k) r = fmin(r, sv[k]);
  return r;
}

// Last-workgroup detection: every workgroup of the launch must call it once.
// Returns true in the workgroup that arrives last; that workgroup then sees
// all partials the others wrote before arriving.  The partials must be
// written by wave 0 (thread 0's wave), which issues the release: an
// agent-scope release fence (L2 writeback) before a relaxed count, and the
// last workgroup's agent-scope acquire fence after it.  The producers do not
// invalidate their L2 (a sequentially consistent fence would, in every
// workgroup, and the following loads of the launch and of the next kernels
// would miss).
// nb: the number of workgroups that take part (all of the launch, or the
// share of one stage in a launch that runs several stages side by side).
__device__ inline bool last_block(int* ctr, int nb) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int prev = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == nb - 1);
  }
  __syncthreads();
  if (s_last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return s_last != 0;
}
__device__ inline bool last_block(int* ctr) { return last_block(ctr, (int)gridDim.x); }

}  // namespace pds
