// The opaque ds_ctx of the C-ABI: one device, one single-LP solver (which owns
// the context's HIP stream), an optional batch and its warm-start buffers.
#pragma once
#include <string>

#include "ds_batch_host.hpp"
#include "ds_learn.hpp"
#include "ds_solver.hpp"
#include "ds_warm.hpp"

struct ds_ctx {
  pds::Solver* solver = nullptr;
  pds::BatchSolver* batch = nullptr;
  pds::WarmBuffers* wb = nullptr;
  int device = 0;
  std::string err;
  // CHUZR entry (ds_chuzr): per-workgroup partials and the arrival counter,
  // allocated on first use.
  double* chuzr_part_w = nullptr;
  int* chuzr_part_i = nullptr;
};
