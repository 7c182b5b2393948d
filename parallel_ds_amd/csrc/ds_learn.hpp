// Learned warm start: training the network on a solved batch and starting a
// batch from its predictions.
#pragma once
#include "ds_batch_host.hpp"
#include "ds_mlp.hpp"

namespace pds {

struct WarmBuffers {
  DBuf<float> F, y, z, mu, sd, prior;
  DBuf<int> order;
};

// Features and labels of every variable of the batch's last solve,
// standardisation fitted to the features, then `epochs` shuffled Adam passes.
// Returns the device time of the training steps (ms).
double fit_from_batch(Mlp& net, BatchSolver& B, WarmBuffers& wb, int epochs, int minibatch, float lr,
                      unsigned long long seed, double* epoch_loss);

// Features -> logits -> top-m starting bases, all on the device.  Returns the
// wall time of the three steps (ms); *infer_ms gets the inference kernel time.
double warm_start_batch(Mlp& net, BatchSolver& B, WarmBuffers& wb, double* infer_ms);

}  // namespace pds
