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
void launch_colstats(hipStream_t s, const float* F, long long R, float* mu, float* sd);

}  // namespace pds
