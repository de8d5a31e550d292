// ut_gp.h -- latent Gaussian-process Poisson model (internal to libuncertaintea_hip).
//
//   log l_d ~ normal(ls_loc, ls_scale)  (l_d ~ lognormal, log transform), d = 1..D
//   log a  : a ~ halfnormal(amp_scale)  (log transform)
//   z ~ normal(0, 1)^N;  f = chol(K) z;  y_i ~ poisson(exp(f_i))
//   K_ij = a^2 exp(-0.5 sum_d (x_di - x_dj)^2 / l_d^2) + (noise^2 + 1e-8) delta_ij
// Parameter order: (log l_1 .. log l_D, log a, z_1 .. z_N).  The gradient with
// respect to the D + 1 kernel hyperparameters propagates their tangents through the
// Cholesky factorisation (forward mode, three tangents for D = 2); everything
// about K and L is binary64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut {

constexpr int GP_TILE = 32;
constexpr int GP_MAX_T = 3;  // hyperparameter tangents carried through the factorisation

struct GpPlan {
  const double* X = nullptr;  // N x D inputs
  const double* y = nullptr;  // N counts
  int N = 0, D = 0;
  double noise2 = 0.0;        // noise^2 + 1e-8
  double ls_loc = 0.0, ls_scale = 0.5, amp_scale = 1.0;
  double total_const = 0.0;   // every constant of the log density
};

struct GpScratch {
  double* A = nullptr;        // C x (1 + T) x N x N dual matrices (row-major, lower triangle used)
  double* logdet = nullptr;   // C
};

// Builds K and its tangents and factorises them for all chains; log-determinants
// 2 sum log L_ii into s.logdet.  hyper(c, k) = log-hyperparameter k of chain c at
// hyper[c * stride + k].
hipError_t launch_gp_factor(const GpPlan& p, const float* q, int stride, const double* hyper_d, int C,
                            GpScratch& s, hipStream_t stream);
// Logjoint and gradient of all chains at positions q (C x P, fp32)
hipError_t launch_gp_grad(const GpPlan& p, const float* q, float* grad, float* lj, int C, GpScratch& s,
                          hipStream_t stream);

}  // namespace ut
