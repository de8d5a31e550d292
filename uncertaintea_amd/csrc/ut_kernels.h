// ut_kernels.h -- host-side interface between the C ABI layer (ut_api.hip) and the
// device kernels (ut_kernels.hip).  Internal to libuncertaintea_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ut_plan.h"

namespace ut {

// MODEL_GLM_MFMA has no wavefront-per-chain evaluator: its density and gradient run
// for all chains at once on the matrix cores (ut_glm.h).
// MODEL_GP (latent GP Poisson, ut_gp.h) likewise evaluates all chains at once.
// MODEL_PLAN is the general static execution plan, run by the plan interpreter.
enum ModelKind { MODEL_LOGISTIC = 1, MODEL_HIER = 2, MODEL_FUNNEL = 3, MODEL_GLM_MFMA = 4, MODEL_GP = 5, MODEL_PLAN = 6 };

// Philox stream ids.  NUTS uses 0..3 exactly as the reference's persistent kernel
// (src/device/persistent_nuts.jl:70-73); the remaining purposes get their own ids.
enum Stream : uint32_t {
  STREAM_MOMENTUM = 0,
  STREAM_DIRECTION = 1,
  STREAM_LEAF = 2,
  STREAM_MERGE = 3,
  STREAM_STEP_SEARCH = 4,
  STREAM_INIT = 5,
  STREAM_HMC_ACCEPT = 6,
  STREAM_MEADS_REFRESH = 7,
  STREAM_MEADS_INIT = 8
};

// Per-iteration flag bits stored next to warmup history and draws.
enum : uint8_t { FLAG_DIVERGENT = 1, FLAG_ACCEPTED = 2 };

// Device description of a model: pointers into plan-owned HBM plus scalars.  Each
// kernel launch builds the matching device model struct from it.
struct DevModelDesc {
  int kind = 0;
  int P = 0;
  int K = 0;
  // hierarchical
  const float* y_f = nullptr;
  const double* y_d = nullptr;
  const int* nobs = nullptr;
  const float* lane_f = nullptr;
  const double* lane_d = nullptr;
  int J = 0;
  int max_n = 0;
  double mu_loc = 0, mu_scale = 1, tau_scale = 1;
  double inv_mu_scale = 1, inv_tau_scale = 1;
  // logistic
  const float* xT_f = nullptr;
  const double* xT_d = nullptr;
  const float* yv_f = nullptr;
  const double* yv_d = nullptr;
  long long n = 0;
  int D = 0;
  double loc = 0, scale = 1;
  // wavefronts per chain in the persistent samplers and the density kernel (1, or
  // 2, 4, 8 for the logistic model: one chain's rows split over its workgroup)
  int waves = 1;
  // funnel
  int dim = 0;
  double v_scale = 1;
  double total_const = 0;
  // general plan (device arrays owned by the plan; the host description for the
  // samplers' prior draws)
  PlanView plan;
  const ut_plan_desc* host_plan = nullptr;
};

// Feature rows of the logistic plan's staged X (feature-major, zero rows past D): the
// register bound of the model instance that runs it (32 up to 32 covariates, else 64 per
// slot), so the row loop needs no per-feature masks.
inline int logistic_feature_rows(int D, int K) { return D <= 32 ? 32 : 64 * K; }

// LDS of one wavefront's NUTS checkpoint stack (positions and momenta of the
// dyadic checkpoints, max_depth + 1 of each).
inline size_t nuts_checkpoint_bytes(int K, int max_depth) {
  return (size_t)(max_depth + 1) * 2 * K * WAVE * sizeof(float);
}
// LDS of the persistent NUTS kernel for a desc's model and chain layout (the
// checkpoint stacks of every wavefront plus a workgroup-per-chain exchange area).
inline size_t nuts_lds_bytes(const DevModelDesc& d, int max_depth) {
  const size_t w = d.waves > 1 ? (size_t)d.waves : 1;
  size_t b = w * nuts_checkpoint_bytes(d.K, max_depth);
  if (d.waves > 1) b = (b + 15) / 16 * 16 + w * ((size_t)d.K * WAVE + 1) * sizeof(float);
  return b;
}

// Device pointers of the chain state (api order, P x C column-major).
struct ChainState {
  float* pos = nullptr;
  float* grad = nullptr;
  float* lj = nullptr;
  double* step = nullptr;   // C
  double* dual = nullptr;   // C x 5: mu, log_step, log_step_avg, hbar, iteration
  float* im_f = nullptr;    // P
};

struct SamplerParams {
  uint32_t run_seed = 0;
  int P = 0, C = 0;
  int max_depth = 10;
  int num_leapfrog = 10;
  double div_thr = 1000.0;
  double target_accept = 0.8;
  int adapt_step = 1;
  int num_warmup = 0;
  int num_samples = 0;
  float* hist_pos = nullptr;      // num_warmup x C x P
  double* hist_acc = nullptr;     // num_warmup x C
  uint8_t* hist_flags = nullptr;  // num_warmup x C
  float* draws = nullptr;         // C x S x P
  double* o_lj = nullptr;         // C x S
  double* o_acc = nullptr;
  double* o_energy = nullptr;
  double* o_err = nullptr;
  uint8_t* o_accepted = nullptr;
  uint8_t* o_divergent = nullptr;
  int* o_depth = nullptr;
  int* o_steps = nullptr;
  unsigned long long* eval_counter = nullptr;  // leapfrog steps while sampling
};

// Launchers (all asynchronous on `stream`).  They return hipSuccess or the launch
// error.
hipError_t launch_density_f32(const DevModelDesc& m, const float* params, float* values,
                              float* grads, int C, hipStream_t stream);
hipError_t launch_density_f64(const DevModelDesc& m, const double* params, double* values,
                              double* grads, int C, hipStream_t stream);
hipError_t launch_nuts(const DevModelDesc& m, const ChainState& st, const SamplerParams& sp,
                       int it_begin, int it_end, hipStream_t stream);
hipError_t launch_hmc(const DevModelDesc& m, const ChainState& st, const SamplerParams& sp,
                      int it_begin, int it_end, hipStream_t stream);
// Per-chain reasonable-step-size search (do_search) and dual-averaging reset.
hipError_t launch_step_search(const DevModelDesc& m, const ChainState& st, int C,
                              uint32_t iteration_coord, int do_search, double target_accept,
                              hipStream_t stream);
// Replays the pooled running-variance updates of warmup iterations [it_lo, it_hi]
// (1-based) and writes the window's inverse mass.  summary[0..4] = count, weight
// sum, weight-square sum, effective count, updated flag.
hipError_t launch_mass_replay(const float* hist_pos, const double* hist_acc,
                              const uint8_t* hist_flags, int C, int P, int it_lo, int it_hi,
                              int window_length, int use_accepted, double regularization,
                              int min_samples, double* im_d, float* im_f, double* summary,
                              hipStream_t stream);
// ---- masked NUTS (tree_strategy = masked): every chain advances one leaf of its
// own tree per step and the gradient of all chains comes from one batched call.
struct MnChain {
  double H0, cont_w, accept_sum, prop_energy, prop_error, sub_w, sub_asum, sub_energy, sub_error, step;
  float ljL, ljR, ljP, ljS, cur_lj, lj0, lj;
  int accept_count, depth, steps, sub_int, sub_acount, leaf, forward, divergent, turning, sub_div, sub_turn;
  int phase, it;
  uint32_t leaf_counter;
  // per-chain step search
  double search_s, search_next, search_dir, search_H0;
  int search_t, search_valid;
};
struct MnState {
  float* vec = nullptr;   // C x MN_NV x K x 64 (tree vectors)
  float* ck = nullptr;    // C x (max_depth + 1) x 2 x K x 64 (dyadic checkpoints)
  MnChain* ch = nullptr;  // C
  int* active = nullptr;  // chains not finished after a step
  float* wq = nullptr;    // C x P: rows whose gradient the next batched call computes
  const float* wg = nullptr;
  const float* wlj = nullptr;
};
// Number of tree vectors per chain and the layout used for plans without a
// wavefront-per-chain model (parameter a in lane a % 64, slot a / 64).
constexpr int MN_NV = 15;
int masked_slots(const DevModelDesc& d);
// Range [it_begin, it_end] of NUTS iterations: mark every chain to start at it_begin.
hipError_t launch_mn_reset(const MnState& ms, int C, int it_begin, hipStream_t stream);
// One step: consume the gradient of the last rows written (when pending), advance
// each chain's tree by one leaf (finishing and starting iterations as they end),
// write the next rows to ms.wq and count the chains still running.
hipError_t launch_mn_step(const DevModelDesc& d, const ChainState& st, const SamplerParams& sp, const MnState& ms,
                          int it_end, hipStream_t stream);
// Per-chain step search in lock step: start (do_search 0 only resets the dual
// averaging) and one trial step per call; done when no chain is searching.
hipError_t launch_mn_search_begin(const DevModelDesc& d, const ChainState& st, const MnState& ms, int C, int P,
                                  uint32_t itc, int do_search, hipStream_t stream);
hipError_t launch_mn_search_step(const DevModelDesc& d, const ChainState& st, const MnState& ms, int C, int P,
                                 uint32_t itc, hipStream_t stream);

hipError_t launch_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream_id,
                              uint32_t first, int64_t n, uint32_t* blocks, float* uf,
                              float* nf, double* ud, hipStream_t stream);

}  // namespace ut
