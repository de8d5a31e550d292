// ut_flow.h -- affine-coupling flow preconditioner (internal to libuncertaintea_hip).
//
// q = f(z): x0 = mu + exp(ls) * z, then for layers l = 0 .. L-1 the passive
// coordinates (parity of index + l even, the reference's alternating masks,
// src/inference/vi/guides.jl:282-300) stay and the active ones become
// x[a] = x[a] * exp(s) + t, where [s_raw, t] = W2 tanh(W1 x[p] + b1) + b2 and
// s = c tanh(s_raw / c).  log |det df/dz| = sum(ls) + sum over layers of sum(s).
// Weights are fp32 masters; the two GEMMs of every conditioner run on the matrix
// cores with bf16 operands (inputs, hidden activations and bf16 weight copies) and
// fp32 accumulation.  Everything else is fp32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace ut {

constexpr int FLOW_PAD = 128;  // passive inputs and conditioner outputs are padded to 128
constexpr int FLOW_WSPLITS = 32;  // split-K of the weight-gradient GEMMs (K = rows)
constexpr int FACTOR_SPLITS = 64;  // row splits of the full-rank factor gradient

struct FlowLayer {
  int np = 0, na = 0;      // passive / active counts (2 na <= 128, np <= 128)
  int* pidx = nullptr;     // device, np passive indices
  int* aidx = nullptr;     // device, na active indices
  // offsets (floats) into the flat parameter vector
  long long oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0;
  uint16_t* W1b = nullptr;   // H x 128 bf16
  uint16_t* W1Tb = nullptr;  // 128 x H
  uint16_t* W2b = nullptr;   // 128 x H
  uint16_t* W2Tb = nullptr;  // H x 128
  // the same four matrices in the all-layer kernels' fragment order (the B-operand
  // pieces of one wavefront contiguous), filled when the layer's widths are the fused ones
  uint16_t* W1f = nullptr;
  uint16_t* W2f = nullptr;
  uint16_t* W2Tf = nullptr;
  uint16_t* W1Tf = nullptr;
};

// Device descriptors of the layers served by the LDS-staged kernels (rebuilt whenever
// the scratch buffers move): weights, biases and the activations each direction uses.
struct FusedFwdLayer {
  const uint16_t* W1b;
  const float* b1;
  const uint16_t* W2b;
  const float* b2;
  float* h;   // the layer's tanh activations, in the fused kernels' fragment order
  float* s;   // B x na
  float* xs;  // B x na
  int np, na, par;
  const uint16_t* W1f;  // W1b / W2b in fragment order (the all-layer kernels)
  const uint16_t* W2f;
};
struct FusedBwdLayer {
  const uint16_t* W2Tb;
  const uint16_t* W1Tb;
  const float* h;
  const float* s;
  const float* xs;
  int np, na, par;
  const uint16_t* W2Tf;  // W2Tb / W1Tb in fragment order
  const uint16_t* W1Tf;
};

// Layer descriptor of the 16x16x32-MFMA all-layer kernels (ut_flow16.hip): fragment-order
// weight copies (that kernel's layout), biases, and the per-batch values the forward keeps for
// the backward (s and x_a e^s, B x na fp32; the bf16 passive inputs in fragment order, B x 64).
struct Flow16Layer {
  const uint16_t* W1f;
  const uint16_t* W2f;
  const uint16_t* W2Tf;
  const uint16_t* W1Tf;
  const float* b1;
  const float* b2;
  float* s;
  float* xs;
  uint16_t* xpg;
  int np, na, par;
};

// Per-batch activations kept for the backward pass.
struct FlowScratch {
  int B = 0;
  std::vector<uint16_t*> xp;  // per layer B x 128 bf16 passive inputs
  std::vector<float*> h;      // per layer B x H tanh activations
  std::vector<uint16_t*> hb;  // per layer B x H bf16
  std::vector<float*> s;      // per layer B x na
  std::vector<float*> xa;     // per layer B x na: x_a e^s (forward) / the inverted x_a (inverse)
  float* O = nullptr;         // B x 128
  uint16_t* gO = nullptr;     // B x 128
  float* gOf = nullptr;       // B x 128
  uint16_t* gU = nullptr;     // B x H
  float* gUf = nullptr;       // B x H
  uint16_t* tA = nullptr;     // transposes for the weight gradients (128 or H) x B
  uint16_t* tB = nullptr;
  float* x = nullptr;         // B x P work positions
  float* g = nullptr;         // B x P work gradient
  float* aux = nullptr;       // B x P per-element terms of the base-parameter gradients
  float* logdet = nullptr;    // B
  float* colsum = nullptr;    // column-sum partials (64 chunks x widest row)
  float* wpart = nullptr;     // split-K partials of the weight-gradient GEMMs (FLOW_WSPLITS x H x widest)
  // per layer: the last forward ran the LDS-staged kernel (its h is in that kernel's
  // private fragment order, readable only by the matching backward kernel)
  std::vector<char> fwd_fused;
  // the last forward ran the 16x16x32-MFMA all-layer kernel (its saved values are readable
  // only by the matching backward kernel)
  bool fwd16 = false;
};

struct Flow {
  int P = 0, L = 0, H = 0;
  float clamp = 5.0f;
  long long nparams = 0;
  float* params = nullptr;  // device flat: mu[P], ls[P], per layer W1[H x 128], b1[H], W2[128 x H], b2[128]
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long long adam_t = 0;
  std::vector<FlowLayer> layers;
  FlowScratch sc;           // sized for sc.B rows
  std::vector<void*> owned;
  hipStream_t stream = nullptr;
  unsigned gen = 0;         // bumped whenever the scratch buffers move (captured graphs go stale)
  FusedFwdLayer* dfwd = nullptr;  // device, L descriptors
  FusedBwdLayer* dbwd = nullptr;
  Flow16Layer* d16 = nullptr;     // device, L descriptors of the 16x16x32-MFMA kernels
  bool m16 = false;               // the sampling path runs those kernels (flow16_eligible)
  float* ldbase = nullptr;        // device: sum of the base log-scales, refreshed with the weights
  bool fused = true;        // LDS-staged layer kernels on the sampling path where the widths allow
  // full-rank Gaussian guide (no coupling layers): q = mu + e^ls z + F z with F the
  // strictly lower P x P factor stored row-major at params + ofac
  bool factor = false;
  long long ofac = 0;
  double* fpart = nullptr;  // fsplits x P x P split partials of F's gradient
  int fsplits = 0;          // FACTOR_SPLITS, fewer for large P (the partials stay <= 128 MiB)
};

// The sampling path's 16x16x32-MFMA all-layer kernels (ut_flow16.hip): eligible when every
// coupling layer has 1..64 passive and <= 64 active coordinates and H = 256.
bool flow16_eligible(const Flow& f);
// fragment-order weight copies of those kernels and the base log-determinant
hipError_t flow16_sync_weights(Flow& f, hipStream_t st);
// q = f(z) of B rows (B a multiple of 16) into x (and q when not null), log-determinants into
// logdet (and logdet_out when not null); keeps the backward's values in the descriptors' buffers
hipError_t flow16_forward(Flow& f, const float* z, float* x, float* q, float* logdet, float* logdet_out, int B,
                          hipStream_t st);
// g_z = (dq/dz)^T g_q + d logdet / dz for the rows of the last flow16_forward
hipError_t flow16_backward(Flow& f, const float* g_q, float* g_z, int B, hipStream_t st);
// A model whose log density and gradient one row's coordinates give by themselves (Neal's
// funnel), evaluated in the flow forward's epilogue on the rows it produces, with the device
// model's own arithmetic (FunnelModel<K>, one wavefront per row, as in the density kernel).
struct FlowRowModel {
  int K = 0;  // register slots per lane of the density kernel's instance (1 or 2)
  int dim = 0;
  double v_scale = 1.0, total_const = 0.0;
};
// flow16_forward followed by the row model's gradient (g, P x B chain-major as the density
// kernel writes it) and log density (lj, without the flow's log-determinant)
hipError_t flow16_forward_model(Flow& f, const float* z, float* x, float* logdet, const FlowRowModel& m, float* g,
                                float* lj, int B, hipStream_t st);
// The sampling-path forward with the row model fused; hipErrorNotSupported unless the
// 16-row kernels run this flow.
hipError_t flow_forward_model(Flow& f, const float* z, const FlowRowModel& m, float* g, float* lj, int B,
                              hipStream_t st);

// A lock-step leapfrog's full kick and drift applied to the rows of a backward pass (the
// lock-step samplers' k_ls_step, same arithmetic): chain-major positions and momenta,
// the shared inverse mass and the per-chain validity flags.
struct Flow16Step {
  float* wq = nullptr;
  float* wp = nullptr;
  const float* im_f = nullptr;
  uint8_t* valid = nullptr;
  float eps = 0.0f;
};
// flow16_backward with the leapfrog step in its epilogue: for every row of a still-valid chain
// either the step is taken with the gradient flow16_backward would write (all of the row's
// entries finite) or the chain is marked invalid; g_z is not written (nothing reads a
// mid-trajectory gradient after its kick).
hipError_t flow16_backward_step(Flow& f, const float* g_q, float* g_z, int B, const Flow16Step& stp,
                                hipStream_t st);
// The whole flow-wrapped gradient of a row model in one kernel per 16-row tile (forward with its
// saved values in LDS, the row model, backward): with a step, followed by the leapfrog step
// (the arithmetic of flow16_forward_model + flow16_backward_step; lj receives the model's log
// density, g_z is not written), without one (stp null) lj receives the model's log density
// plus the flow's log-determinant and g_z the gradient rows (flow_forward_model +
// flow16_backward + flow_add_logdet).
// hipErrorNotSupported when the layers' saved values do not fit in LDS.
hipError_t flow16_grad_step(Flow& f, const float* z, const FlowRowModel& m, float* lj, float* g_z, int B,
                           const Flow16Step* stp, hipStream_t st);

// factor = true: the full-rank guide (L must be 0), mu, ls and the strictly lower
// factor as its weights
hipError_t flow_alloc(Flow& f, int P, int L, int H, bool factor = false);
void flow_free(Flow& f);
hipError_t flow_scratch(Flow& f, int B);
// refresh the bf16 weight copies after the fp32 masters changed
hipError_t flow_sync_weights(Flow& f, hipStream_t st);
// q = f(z) for B rows (B a multiple of 128, B <= scratch rows); keeps activations
hipError_t flow_forward(Flow& f, const float* z, float* q, float* logdet, int B, hipStream_t st);
// g_z = (dq/dz)^T g_q + d logdet / dz, from the activations of the last forward
hipError_t flow_backward(Flow& f, const float* g_q, float* g_z, int B, hipStream_t st);
// z = f^-1(q), logdet_inv = log |det df^-1/dq| (keeps activations for flow_loss_backward)
hipError_t flow_inverse(Flow& f, const float* q, float* z, float* logdet_inv, int B, hipStream_t st);
// gradient of the mean negative log-likelihood  0.5|z|^2 - logdet_inv  (+ const) of
// the last inverse over the parameters, into f.grads
hipError_t flow_loss_backward(Flow& f, const float* z, int B, hipStream_t st);
// Adam step on the mean over B rows of the last flow_loss_backward gradients
hipError_t flow_adam(Flow& f, int B, double lr, double b1, double b2, double eps, hipStream_t st);
// lj += logdet (z-space log density) / lj -= logdet (back to the model's log density)
hipError_t flow_add_logdet(float* lj, const float* logdet, int B, hipStream_t st);
hipError_t flow_sub_logdet(double* lj, const float* logdet, int B, hipStream_t st);
// rows of src picked by Philox draws (key = (seed, step)) -> dst (B x P)
hipError_t flow_gather_rows(const float* src, long long n, int P, float* dst, int B, uint32_t seed, uint32_t step,
                            hipStream_t st);
// per-row loss 0.5 |z|^2 - logdet_inv + P/2 log(2 pi), summed into *out (device double)
hipError_t flow_loss_value(const float* z, const float* logdet_inv, int B, int P, double* out, hipStream_t st);

// ---- reparameterised ELBO training (the reference's flow ADVI, src/inference/vi.jl:750-840)
// z ~ N(0, I): B x P Philox normals, key = (seed, step)
hipError_t flow_normal_fill(float* z, int B, int P, uint32_t seed, uint32_t step, hipStream_t st);
// forward pass that keeps every activation the weight gradient needs (never the
// LDS-staged sampling kernels, which keep only what the input gradient needs)
hipError_t flow_forward_train(Flow& f, const float* z, int B, hipStream_t st);
// Per row of the last flow_forward_train with the model's logjoint lj and gradient
// in sc.g: valid = lj, logdet and the gradient finite; an invalid row's gradient is
// zeroed and rw[b] = 0 (else 1); logw[b] = lj + 0.5|z|^2 + P/2 log 2 pi + logdet.
hipError_t flow_elbo_rows(Flow& f, const float* z, const float* lj, int B, float* rw, double* logw,
                          hipStream_t st);
// Sum over rows of the weight gradient of rw[b] (g_q . f(z) + logdet) with g_q = sc.g
// held fixed (invalid rows contribute nothing), into f.grads
hipError_t flow_elbo_backward(Flow& f, const float* z, const float* rw, int B, hipStream_t st);
// stats (device, 5 doubles) = {divisor, objective, gradient norm of the mean, update
// scale = min(1, clip / norm) / divisor, standard ELBO}: divisor = valid rows and
// objective = the standard ELBO (mean logw over them), or with group statistics
// (gstat, ngroups) divisor = groups used and objective = the importance-weighted
// bound; part: >= 256 doubles
hipError_t flow_elbo_stats(Flow& f, const float* rw, const double* logw, int B, const double* gstat, int ngroups,
                           double clip, double* part, double* stats, hipStream_t st);
// importance weights of groups of K particles (K divides B): wts[b] within-group
// self-normalised weights (0 when invalid), gstat[2g] the group's log-mean-exp bound,
// gstat[2g + 1] whether it had a valid particle; scales sc.g's rows by wts
hipError_t flow_iwae_weights(Flow& f, const float* rw, const double* logw, int B, int K, float* wts, double* gstat,
                             hipStream_t st);
// Adam ascent on f.grads scaled by stats[3]; no update when stats[0] == 0 or, with
// track, once track->failed is set
struct ElboTrack {
  double best;    // best objective so far (-inf at the start)
  int failed;     // set at the first iteration without a valid particle (weight)
  int fail_iter;  // that iteration
  int take;       // this iteration's objective is a new best
  int pad;
};
hipError_t flow_adam_ascent(Flow& f, const double* stats, double lr, double b1, double b2, double eps,
                            hipStream_t st, const ElboTrack* track = nullptr);
// per-iteration bookkeeping of an ELBO fit, on the device: hist[3 it + {0, 1, 2}] =
// objective, standard ELBO, gradient norm; the best objective and the failure flag in
// *track; then best_w = params when the objective is a new best (params as evaluated)
hipError_t flow_elbo_track(Flow& f, const double* stats, int it, double* hist, ElboTrack* track, float* best_w,
                           hipStream_t st);

}  // namespace ut
