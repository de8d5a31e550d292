// ut_chees.h -- ChEES trajectory-length state machine (host scalars), shared by
// the library's sampler driver and its CPU restatement so both take the same
// decisions.  Restates src/inference/api_batched_chees.jl:25-116,237-275
// (Halton jitter, jittered step count, Adam ascent on log T, clipped update,
// moving average, clamp to [eps, max_steps * eps]).
#pragma once
#include <math.h>

#include "ut_math.h"

namespace ut {

constexpr double CHEES_LOG_UPDATE_CLIP = 0.35;
constexpr double CHEES_EPS = 1e-20;

// base-2 van der Corput radical inverse of index >= 1
inline double halton_base2(long long index) {
  double result = 0.0, fraction = 0.5;
  long long remaining = index;
  while (remaining > 0) {
    result += fraction * (double)(remaining & 1);
    remaining >>= 1;
    fraction *= 0.5;
  }
  return result;
}

inline double chees_jitter(double amount, long long iteration) {
  return halton_base2(iteration) * amount + (1.0 - amount);
}

// max(1, ceil(jitter * T / eps)) with a snap to the nearest integer within 1e-9
inline int chees_steps(double jitter, double T, double eps) {
  const double raw = jitter * T / eps;
  const double nearest = nearbyint(raw);
  const double tol = 1e-9 * (fabs(raw) > 1.0 ? fabs(raw) : 1.0);
  double steps = fabs(raw - nearest) <= tol ? nearest : ceil(raw);
  if (!(steps >= 1.0)) steps = 1.0;
  if (steps > 1e9) steps = 1e9;
  return (int)steps;
}

struct CheesAdam {
  double beta1, beta2, epsilon, lr, m = 0.0, v = 0.0;
  long long t = 0;
  double ascend(double g) {
    t += 1;
    m = beta1 * m + (1.0 - beta1) * g;
    v = beta2 * v + (1.0 - beta2) * g * g;
    double b1t = 1.0, b2t = 1.0;
    for (long long i = 0; i < t; ++i) { b1t *= beta1; b2t *= beta2; }
    const double mhat = m / (1.0 - b1t);
    const double vhat = v / (1.0 - b2t);
    return lr * mhat / (sqrt_d(vhat) + epsilon);
  }
};

struct CheesTrajectory {
  double T, log_T_ma;
  CheesAdam adam;
  long long step = 1;
  double decay;
  int max_steps;
  // one cross-chain update from the reduced numerator / denominator
  double apply(double numerator, double denominator, double jitter, double eps_clip) {
    const double grad = denominator > 0.0 ? jitter * T * numerator / denominator : 0.0;
    const double logT = log_d(T);
    const double sm = adam.m, sv = adam.v;
    const long long st = adam.t;
    double upd = adam.ascend(grad);
    upd = upd < -CHEES_LOG_UPDATE_CLIP ? -CHEES_LOG_UPDATE_CLIP : (upd > CHEES_LOG_UPDATE_CLIP ? CHEES_LOG_UPDATE_CLIP : upd);
    double logT_new = logT + upd;
    if (!isfinite_d(logT_new)) {
      logT_new = logT;
      adam.m = sm;
      adam.v = sv;
      adam.t = st;
    }
    const double w = pow_d((double)step, -decay);
    log_T_ma = (1.0 - w) * log_T_ma + w * logT_new;
    double nt = exp_d(log_T_ma);
    const double hi = (double)max_steps * eps_clip;
    nt = nt < eps_clip ? eps_clip : (nt > hi ? hi : nt);
    T = nt;
    step += 1;
    return nt;
  }
};

}  // namespace ut
