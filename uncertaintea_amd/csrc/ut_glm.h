// ut_glm.h -- fused matrix-core logistic GLM (internal to libuncertaintea_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut {

// Data of a logistic GLM staged for the matrix cores: X (N x D, bf16, row-major),
// outcomes y (fp32), normal(loc, scale) prior.
struct GlmPlan {
  const uint16_t* X = nullptr;
  const float* y = nullptr;
  int N = 0;
  int D = 0;
  double loc = 0.0, scale = 1.0, total_const = 0.0;
};

// Per-workspace scratch: bf16 hi/lo positions (C x D), per-split gradient
// partials (S x C x D, fp32) and log-likelihood partials (S x C, binary64).
struct GlmScratch {
  int splits = 8;
  uint16_t* bhi = nullptr;
  uint16_t* blo = nullptr;
  float* gpart = nullptr;
  double* llpart = nullptr;
};

// Row splits used for N rows and C chains (fills the chip, multiple of 8 when possible).
int glm_splits(int N, int C);

// grad (C x D) and, when need_lj, lj (C) at positions q (C x D, fp32).
// Requires D == 256, N % 32 == 0, C % 128 == 0.
hipError_t launch_glm_grad(const GlmPlan& p, const float* q, float* grad, float* lj, int C, bool need_lj,
                           GlmScratch& s, hipStream_t stream);

}  // namespace ut
