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
#include <stddef.h>
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

// Factor storage is tile-major: per chain and component (value, then one tangent per
// log hyperparameter) the nb (nb + 1) / 2 lower tiles (i >= j, index i (i + 1) / 2 + j)
// of 32 x 32 doubles, each tile column-major (element (row, col) at col * 32 + row).
// The upper triangle of the diagonal tiles is stored as zero.
struct GpScratch {
  double* A = nullptr;        // C x (T + 1) x tiles x 1024
  double* pf = nullptr;       // C x nb x (T + 1) x N partial products (L z, dL z) per tile column
  double* ldp = nullptr;      // C x nb partial sums of log L_ii
  double* r = nullptr;        // C x N residuals y - exp(f)
  double* logdet = nullptr;   // C log det K
};

inline size_t gp_tiles(int N) {
  const size_t nb = (size_t)(N / GP_TILE);
  return nb * (nb + 1) / 2;
}
// offset (in doubles) of tile (i, j), i >= j, of component comp of chain c
inline size_t gp_tile_offset(int N, int T, int c, int comp, int i, int j) {
  return (((size_t)c * (T + 1) + comp) * gp_tiles(N) + (size_t)i * (i + 1) / 2 + j) * (GP_TILE * GP_TILE);
}

hipError_t gp_scratch_alloc(GpScratch& s, const GpPlan& p, int C);
void gp_scratch_free(GpScratch& s);

// Builds K and its tangents tile by tile and factorises them for all chains.
// hyper(c, k) = log-hyperparameter k of chain c at hyper_d[c * stride + k] when
// hyper_d is given, else q[c * stride + k]; with q given the partial products of
// L and dL with z = q[c * stride + D + 1 ..] are formed as well.
hipError_t launch_gp_factor(const GpPlan& p, const float* q, int stride, const double* hyper_d, int C,
                            GpScratch& s, hipStream_t stream);
// Logjoint and gradient of all chains at positions q (C x P, fp32)
hipError_t launch_gp_grad(const GpPlan& p, const float* q, float* grad, float* lj, int C, GpScratch& s,
                          hipStream_t stream);

}  // namespace ut
