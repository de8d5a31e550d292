// ut_lockstep.h -- lock-step HMC trajectory kernels and ChEES reductions
// (internal to libuncertaintea_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ut_kernels.h"

namespace ut {

// Device arrays of a lock-step trajectory; chain-major, reference parameter order.
struct LockstepState {
  float* pos = nullptr;   // C x P current state
  float* grad = nullptr;
  float* lj = nullptr;    // C
  float* wq = nullptr;    // C x P trajectory position / proposal
  float* wp = nullptr;    // C x P momentum (flipped end momentum after k_ls_end)
  float* wg = nullptr;    // C x P gradient at wq
  float* wlj = nullptr;   // C log density at wq (final step)
  double* h0 = nullptr;   // C initial Hamiltonian
  uint8_t* valid = nullptr;
  float* initials = nullptr;  // C x P positions at the start of the iteration
  const float* im_f = nullptr;
};

struct LockstepRecord {
  double* accept_prob = nullptr;  // C
  uint8_t* flags = nullptr;       // C, FLAG_DIVERGENT | FLAG_ACCEPTED
  uint8_t* pfinite = nullptr;     // proposal position and momentum finite
  uint8_t* qfinite = nullptr;     // proposal position finite
  uint8_t* ifinite = nullptr;     // initial position finite
  double* energy = nullptr;
  double* error = nullptr;
  double* imean = nullptr;        // P
  double* pmean = nullptr;        // P
  double* inv_sqrt = nullptr;     // P
  double* term = nullptr;         // C
  uint8_t* include = nullptr;     // C
};

hipError_t launch_ls_begin(const LockstepState& s, int P, int C, uint32_t itc, uint32_t mom_stream, float eps,
                           int save_initials, hipStream_t stream);
hipError_t launch_ls_step(const LockstepState& s, int P, int C, float eps, hipStream_t stream);
hipError_t launch_ls_end(const LockstepState& s, const LockstepRecord& r, int P, int C, uint32_t itc, float eps,
                         double div_thr, int mode, hipStream_t stream);
hipError_t launch_ls_record(const LockstepState& s, const LockstepRecord& r, const SamplerParams& sp, int it, int L,
                            hipStream_t stream);
// stats[0..3]: ChEES numerator, denominator, harmonic-mean acceptance, mean acceptance;
// vsign multiplies the stored (flipped) end momentum in the numerator
hipError_t launch_chees_stats(const LockstepState& s, const LockstepRecord& r, const double* im_d, int P, int C,
                              bool trajectory_terms, double vsign, double* stats, hipStream_t stream);

}  // namespace ut
