// Learned warm start, LP side: feature, label and selection kernels over a
// batch (one workgroup per LP).
#pragma once
#include "ds_batch.hpp"

namespace pds {

constexpr int WS_NF = 17;  // features per variable (the last one is the family prior, filled by the host)

size_t features_lds_bytes(int mmax, int nmax);
size_t select_lds_bytes(int Nmax);
void launch_features(hipStream_t s, const BatchDev& D, int count, int mmax, int nmax, float* F);
void launch_labels(hipStream_t s, const BatchDev& D, int count, const int* basis, float* y);
void launch_select(hipStream_t s, const BatchDev& D, int count, int Nmax, const float* z, int* warm);
// Family prior (feature WS_NF - 1) of T LPs of N variables each: counts of
// basic labels per variable index, and F[r][WS_NF - 1] = (base[r mod N] -
// y[r]) / denom (y null: 0).
void launch_prior_count(hipStream_t s, const float* y, int T, int N, float* cnt);
void launch_put_prior(hipStream_t s, float* F, const float* base, const float* y, long long R, int N, float denom);
void launch_colstats(hipStream_t s, const float* F, long long R, float* mu, float* sd);

}  // namespace pds
