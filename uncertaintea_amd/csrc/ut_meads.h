// ut_meads.h -- MEADS kernels (internal to libuncertaintea_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ut_lockstep.h"

namespace ut {

// Device arrays of the MEADS ensemble; chains are split into K contiguous folds.
struct MeadsState {
  float* pos = nullptr;     // C x P (shared with the workspace chain state)
  float* grad = nullptr;
  float* lj = nullptr;      // C
  float* mom = nullptr;     // C x P persistent whitened momentum
  double* slice = nullptr;  // C persistent slice variable in [-1, 1)
  float* wq = nullptr;      // C x P trajectory position
  float* wp = nullptr;      // C x P momentum (flipped after the step)
  float* wg = nullptr;      // C x P gradient at wq
  float* wlj = nullptr;     // C
  float* refreshed = nullptr;  // C x P refreshed whitened momentum
  double* h0 = nullptr;     // C
  uint8_t* valid = nullptr; // C
  // per fold
  double* sigma = nullptr;  // K x P neighbour-fold standard deviations
  float* sigma_f = nullptr; // K x P
  double* white = nullptr;  // 2 x K x Nmax x P whitened positions / gradients (member-major)
  double* gram = nullptr;      // 2 x K x Nmax x Nmax Gram matrices of the whitened ensembles
  double* gram_row = nullptr;  // 2 x K x Nmax row partials of the off-diagonal squared Gram sum
  double* fold = nullptr;   // K x 4: step, alpha, delta, (unused)
  int* fold_start = nullptr;   // K + 1 chain offsets
};

struct MeadsRecord {
  double* accept_prob = nullptr;
  uint8_t* flags = nullptr;
  double* energy_error = nullptr;
};

// Fold parameters from the neighbour folds' current states (iteration t >= 1):
// sigma, whitened ensembles, Gram trace ratios, step / damping / drift.
hipError_t launch_meads_fold_params(const MeadsState& m, int P, int K, int nmax, int t, hipStream_t stream);
// Momentum refresh + half kick + drift (one leapfrog step starts)
hipError_t launch_meads_begin(const MeadsState& m, const int* fold_of, int P, int C, uint32_t itc,
                              hipStream_t stream);
// Final half kick, momentum flip, persistent-slice accept; skip_fold keeps its state
hipError_t launch_meads_end(const MeadsState& m, const MeadsRecord& r, const int* fold_of, int P, int C,
                            int skip_fold, double div_thr, hipStream_t stream);
// Record iteration `it` (1-based) of W warmup + S samples into the sampler outputs
hipError_t launch_meads_record(const MeadsState& m, const MeadsRecord& r, const int* fold_of,
                               const SamplerParams& sp, int it, int skip_fold, hipStream_t stream);
// Initial persistent momentum (unit normals) and slice (uniform in [-1, 1))
hipError_t launch_meads_init(const MeadsState& m, int P, int C, uint32_t itc, hipStream_t stream);

}  // namespace ut
