// CHUZR as a standalone operation on caller-owned device arrays: dual
// steepest-edge pricing over the basic variables (BaseDualSimplex::DSEPricing,
// common/baseDualSimplex.tpp:470-501, with checkVarIsInfeas :443-466).  The
// whole-solve path fuses the same rule into its pricing kernel; this entry
// serves a caller that keeps its own basis data in HBM.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ds_abi.h"
#include "ds_ctx.hpp"
#include "ds_reduce.hpp"
#include "ds_rules.hpp"

namespace pds {
namespace {

constexpr int CZ_BS = 256;
constexpr int CZ_GRID_MAX = 512;  // workgroups of one launch = partials in the scratch

// Argmax over i of delta_i^2 / beta_i among the basic variables that are out of
// bounds by more than the DSE floor (ties to the lowest position).  Each
// workgroup leaves its winner in part_w/part_i; the last one to arrive merges
// them, writes (p_idx, delta) or (-1, 0), and resets the arrival counter.
__global__ __launch_bounds__(CZ_BS) void k_chuzr(int m, const signed char* __restrict__ bt,
                                                 const double* __restrict__ x, const double* __restrict__ l,
                                                 const double* __restrict__ u, const double* __restrict__ beta,
                                                 double* part_w, int* part_i, int* ctr, int* out_idx,
                                                 double* out_delta) {
  double bw = -1.0;
  int bi = INT_MAX;
  for (int i = blockIdx.x * CZ_BS + threadIdx.x; i < m; i += gridDim.x * CZ_BS) {
    double dl;
    if (var_infeas((int8_t)bt[i], x[i], l[i], u[i], dl) && fabs(dl) > DSE_FLOOR) {
      const double w = (dl * dl) / beta[i];
      if (w > 0.0) argmax_merge(bw, bi, w, i);
    }
  }
  block_argmax<CZ_BS>(bw, bi);
  if (threadIdx.x == 0) {
    part_w[blockIdx.x] = bw;
    part_i[blockIdx.x] = bi;
  }
  if (!last_block(ctr)) return;
  double fw = -1.0;
  int fi = INT_MAX;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += CZ_BS) argmax_merge(fw, fi, part_w[b], part_i[b]);
  block_argmax<CZ_BS>(fw, fi);
  if (threadIdx.x != 0) return;
  *ctr = 0;
  if (fi == INT_MAX) {
    *out_idx = -1;
    *out_delta = 0.0;
    return;
  }
  double dl;
  var_infeas((int8_t)bt[fi], x[fi], l[fi], u[fi], dl);
  *out_idx = fi;
  *out_delta = dl;
}

// Zeroes the arrival counter before a launch (a kernel, so that the call can be
// captured in a graph).
__global__ void k_chuzr_reset(int* ctr) { *ctr = 0; }

int chuzr_grid(int m) {
  int g = (m + CZ_BS - 1) / CZ_BS;
  return g < 1 ? 1 : (g > CZ_GRID_MAX ? CZ_GRID_MAX : g);
}

bool chuzr_args_ok(int m, const signed char* bt, const double* x, const double* l, const double* u,
                   const double* beta, const int* p_idx, const double* delta) {
  return m > 0 && bt && x && l && u && beta && p_idx && delta;
}

}  // namespace
}  // namespace pds

using namespace pds;

extern "C" int ds_chuzr(ds_ctx* ctx, int m, const signed char* bt_B, const double* x_B, const double* l_B,
                        const double* u_B, const double* beta, int* p_idx, double* delta) {
  if (!ctx) return DS_EINVAL;
  if (!chuzr_args_ok(m, bt_B, x_B, l_B, u_B, beta, p_idx, delta)) {
    ctx->err = "ds_chuzr: m must be positive and every pointer set";
    return DS_EINVAL;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return DS_EHIP;
  if (!ctx->chuzr_part_w) {
    // partials of CZ_GRID_MAX workgroups; the last int of part_i is the arrival counter
    if (hipMalloc((void**)&ctx->chuzr_part_w, sizeof(double) * CZ_GRID_MAX) != hipSuccess ||
        hipMalloc((void**)&ctx->chuzr_part_i, sizeof(int) * (CZ_GRID_MAX + 1)) != hipSuccess) {
      ctx->err = "ds_chuzr: device allocation failed";
      return DS_EHIP;
    }
  }
  hipStream_t s = ctx->solver->stream();
  int* ctr = ctx->chuzr_part_i + CZ_GRID_MAX;
  k_chuzr_reset<<<1, 1, 0, s>>>(ctr);
  k_chuzr<<<chuzr_grid(m), CZ_BS, 0, s>>>(m, bt_B, x_B, l_B, u_B, beta, ctx->chuzr_part_w, ctx->chuzr_part_i, ctr,
                                          p_idx, delta);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    ctx->err = std::string("ds_chuzr: ") + hipGetErrorString(e);
    return DS_EHIP;
  }
  return DS_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant.  Its scratch (the partials of CZ_GRID_MAX
// workgroups and the arrival counter) is a file-scope device array; every GPU
// the library runs on holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol): chuzr_noalloc_setup() looks up the device address once per
// device, with that device current, and the entry passes that address.
// ---------------------------------------------------------------------------
namespace pds {
namespace {

constexpr size_t CZ_SCRATCH_BYTES = 8192;  // 512 doubles + 513 ints, rounded up
static_assert(CZ_GRID_MAX * (sizeof(double) + sizeof(int)) + sizeof(int) <= CZ_SCRATCH_BYTES, "scratch too small");

static __device__ __align__(256) unsigned char scratch[CZ_SCRATCH_BYTES];

constexpr int CZ_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[CZ_MAX_DEVICES] = {};

}  // namespace

// Called by ds_create with `device` current.
int chuzr_noalloc_setup(int device) {
  if (device < 0 || device >= CZ_MAX_DEVICES) return DS_EINVAL;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return DS_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DS_EHIP;
  g_scratch_ptr[device] = p;
  return DS_OK;
}

}  // namespace pds

extern "C" int ds_chuzr_noalloc(ds_ctx* ctx, int m, const signed char* bt_B, const double* x_B, const double* l_B,
                                const double* u_B, const double* beta, int* p_idx, double* delta) {
  if (!ctx) return DS_EINVAL;
  if (!chuzr_args_ok(m, bt_B, x_B, l_B, u_B, beta, p_idx, delta)) {
    ctx->err = "ds_chuzr_noalloc: m must be positive and every pointer set";
    return DS_EINVAL;
  }
  if (m > DS_CHUZR_NOALLOC_MAX_M) {
    ctx->err = "ds_chuzr_noalloc: m = " + std::to_string(m) + " exceeds DS_CHUZR_NOALLOC_MAX_M = " +
               std::to_string(DS_CHUZR_NOALLOC_MAX_M);
    return DS_EINVAL;
  }
  void* base = nullptr;
  if (ctx->device >= 0 && ctx->device < CZ_MAX_DEVICES) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  if (!base) {
    ctx->err = "ds_chuzr_noalloc: scratch address of this device was not set up by ds_create";
    return DS_ESTATE;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return DS_EHIP;
  double* part_w = (double*)base;
  int* part_i = (int*)((unsigned char*)base + sizeof(double) * CZ_GRID_MAX);
  int* ctr = part_i + CZ_GRID_MAX;
  hipStream_t s = ctx->solver->stream();
  k_chuzr_reset<<<1, 1, 0, s>>>(ctr);
  k_chuzr<<<chuzr_grid(m), CZ_BS, 0, s>>>(m, bt_B, x_B, l_B, u_B, beta, part_w, part_i, ctr, p_idx, delta);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    ctx->err = std::string("ds_chuzr_noalloc: ") + hipGetErrorString(e);
    return DS_EHIP;
  }
  return DS_OK;
}
