// ut_models.h -- device models for the wavefront-per-chain kernels.
//
// Layout: one 64-lane wavefront evaluates one chain.  Each model fixes an internal
// parameter order; internal index i lives in lane i % 64, register slot i / 64, and
// api(lane, slot) maps it back to the reference's parameter order (the row of the
// P x C parameter matrix), or -1 for a padding slot (value 0, inverse mass 0).
//
// Every reduction over lanes is wave_sum(): four DPP butterfly steps inside each
// 16-lane row (lane l adds lane l^1, l^2, then the half-row mirror l^7 and the row
// mirror l^15), after which every lane of row r holds the row total R_r; the
// wavefront total is (R_0 + R_1) + (R_2 + R_3), read with v_readlane.  Every lane
// ends with the same bit pattern.  DPP moves and readlane stay in the VALU
// pipeline (no LDS round trip per step).  Per-lane partial sums run over slots in
// increasing order starting from zero.  A CPU restatement that reproduces this
// order reproduces the kernels bit for bit.
#pragma once
#include "ut_math.h"

namespace ut {

// DPP controls (GCN/CDNA encoding): quad_perm [1,0,3,2] and [2,3,0,1],
// row_half_mirror, row_mirror.
constexpr int DPP_XOR1 = 0xB1;
constexpr int DPP_XOR2 = 0x4E;
constexpr int DPP_HALF_MIRROR = 0x141;
constexpr int DPP_MIRROR = 0x140;

template <int CTRL>
__device__ inline int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ inline float dpp(float v) {
  return __builtin_bit_cast(float, dpp_i<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL>
__device__ inline double dpp(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)dpp_i<CTRL>((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)dpp_i<CTRL>((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline float readlane(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ inline double readlane(double v, int lane) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <class T>
__device__ inline T wave_sum(T v) {
  v = v + dpp<DPP_XOR1>(v);
  v = v + dpp<DPP_XOR2>(v);
  v = v + dpp<DPP_HALF_MIRROR>(v);
  v = v + dpp<DPP_MIRROR>(v);
  const T r0 = readlane(v, 0), r1 = readlane(v, 16), r2 = readlane(v, 32), r3 = readlane(v, 48);
  return (r0 + r1) + (r2 + r3);
}
template <class T>
__device__ inline T wave_bcast(T v, int src) { return readlane(v, src); }
__device__ inline bool wave_all(bool b) { return __all(b ? 1 : 0) != 0; }
__device__ inline bool wave_any(bool b) { return __any(b ? 1 : 0) != 0; }

template <class T> __device__ inline T t_exp(T x);
template <> __device__ inline float t_exp<float>(float x) { return exp_f(x); }
template <> __device__ inline double t_exp<double>(double x) { return exp_d(x); }
template <class T> __device__ inline T t_log1p(T x);
template <> __device__ inline float t_log1p<float>(float x) { return log1p_f(x); }
template <> __device__ inline double t_log1p<double>(double x) { return log1p_d(x); }
template <class T> __device__ inline T t_log1p_exp(T x);
template <> __device__ inline float t_log1p_exp<float>(float x) { return log1p_exp_f(x); }
template <> __device__ inline double t_log1p_exp<double>(double x) { return log1p_exp_d(x); }
template <class T> __device__ inline T t_logistic(T x);
template <> __device__ inline float t_logistic<float>(float x) { return logistic_f(x); }
template <> __device__ inline double t_logistic<double>(double x) { return logistic_d(x); }
template <class T> __device__ inline T t_fma(T a, T b, T c);
template <> __device__ inline float t_fma<float>(float a, float b, float c) { return fma_f(a, b, c); }
template <> __device__ inline double t_fma<double>(double a, double b, double c) { return fma_d(a, b, c); }
template <class T> __device__ inline bool t_isfinite(T x);
template <> __device__ inline bool t_isfinite<float>(float x) { return isfinite_f(x); }
template <> __device__ inline bool t_isfinite<double>(double x) { return isfinite_d(x); }

// Model constants are staged once per plan in both precisions; a kernel of working
// precision T reads the matching copy.
template <class T> struct Pick;
template <> struct Pick<float> {
  __device__ static float get(const float* f, const double*, int i) { return f[i]; }
};
template <> struct Pick<double> {
  __device__ static double get(const float*, const double* d, int i) { return d[i]; }
};

// ----------------------------------------------------------------------------------
// Non-centred hierarchical normal.  Internal order: slot 0 of lane j is theta_raw_j
// (api 2 + j) for j < J; slot 1 of lane 0 is mu (api 0), of lane 1 is log tau (api 1).
// Lane j scores group j's observations:
//   sum_i log N(y_ij; theta_j, sigma_j) = c_j - hinv_j * sum_i (y_ij - theta_j)^2
//   with c_j = -n_j log sigma_j - n_j log(2 pi)/2 and hinv_j = 1 / (2 sigma_j^2),
//   d/dtheta_j = inv_j * sum_i (y_ij - theta_j), inv_j = 1 / sigma_j^2.
// Observations are stored y[i * 64 + j] (zero padded to max_n rows) so that the lanes
// of one row read consecutive words.  With NMAX > 0 every lane keeps its group's
// observations in registers for the whole launch (prepare()); NMAX == 0 streams
// them from memory on every evaluation.  DENSE instances serve plans whose groups all
// hold exactly NMAX observations and carry no per-observation masks.  The arithmetic
// order is the same in every instance.
template <int NMAX, bool DENSE = false>
struct HierModel {
  static constexpr int K = 2;
  const float* y_f;
  const double* y_d;
  const int* nobs;         // [64], 0 for lanes >= J
  const float* lane_f;     // [3][64]: c_j, hinv_j, inv_j
  const double* lane_d;
  int J;
  int max_n;
  double mu_loc, mu_scale, tau_scale, total_const;
  double inv_mu_scale, inv_tau_scale;

  template <class T>
  struct Lane {
    T y[NMAX > 0 ? NMAX : 1];
    T cj, hinv, inv;
    int n;
  };

  __device__ int api(int lane, int slot) const {
    if (slot == 0) return lane < J ? 2 + lane : -1;
    return lane == 0 ? 0 : (lane == 1 ? 1 : -1);
  }

  template <class T>
  __device__ void prepare(Lane<T>& ld) const {
    const int lane = threadIdx.x & (WAVE - 1);
    ld.n = nobs[lane];
    if (NMAX > 0) {
#pragma unroll
      for (int i = 0; i < (NMAX > 0 ? NMAX : 1); ++i)
        ld.y[i] = i < ld.n ? Pick<T>::get(y_f, y_d, i * WAVE + lane) : T(0);
    }
    ld.cj = Pick<T>::get(lane_f, lane_d, lane);
    ld.hinv = Pick<T>::get(lane_f, lane_d, WAVE + lane);
    ld.inv = Pick<T>::get(lane_f, lane_d, 2 * WAVE + lane);
  }

  template <class T>
  __device__ T eval(const Lane<T>& ld, const T (&q)[K], T (&g)[K], bool& ok) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const T mu = wave_bcast(q[1], 0);
    const T ltau = wave_bcast(q[1], 1);
    const T tau = t_exp<T>(ltau);
    const T qr = q[0];
    const T theta = mu + tau * qr;
    const int n = ld.n;
    T sd = T(0), ss = T(0);
    if constexpr (NMAX > 0 && DENSE) {
      // every group has exactly NMAX observations: no per-observation mask; lanes
      // past the last group are zeroed after the loop
#pragma unroll
      for (int i = 0; i < (NMAX > 0 ? NMAX : 1); ++i) {
        const T d = ld.y[i] - theta;
        sd = sd + d;
        ss = t_fma<T>(d, d, ss);
      }
    } else if (NMAX > 0) {
#pragma unroll
      for (int i = 0; i < (NMAX > 0 ? NMAX : 1); ++i) {
        if (i < n) {
          const T d = ld.y[i] - theta;
          sd = sd + d;
          ss = t_fma<T>(d, d, ss);
        }
      }
    } else {
      for (int i = 0; i < max_n; ++i) {
        if (i < n) {
          const T d = Pick<T>::get(y_f, y_d, i * WAVE + lane) - theta;
          sd = sd + d;
          ss = t_fma<T>(d, d, ss);
        }
      }
    }
    T lp = ld.cj - ss * ld.hinv;
    T gth = sd * ld.inv;
    if (lane >= J) {
      lp = T(0);
      gth = T(0);
    }
    lp = lp - T(0.5) * qr * qr;
    g[0] = tau * gth - qr;
    const T s1 = wave_sum(gth);
    const T s2 = wave_sum(gth * qr);
    const T sl = wave_sum(lp);
    // mu ~ normal(mu_loc, mu_scale)
    const T zmu = (mu - T(mu_loc)) * T(inv_mu_scale);
    const T lp_mu = -(zmu * zmu) * T(0.5);
    const T g_mu = s1 - zmu * T(inv_mu_scale);
    // tau ~ halfcauchy(tau_scale), tau = exp(log tau), log-abs-det = log tau
    const T zt = tau * T(inv_tau_scale);
    const T z2 = zt * zt;
    const T lp_tau = ltau - t_log1p<T>(z2);
    const T g_ltau = tau * s2 - (T(2) * z2) / (T(1) + z2) + T(1);
    const T total = ((sl + lp_mu) + lp_tau) + T(total_const);
    g[1] = lane == 0 ? g_mu : (lane == 1 ? g_ltau : T(0));
    ok = wave_all(t_isfinite(g[0]) && t_isfinite(g[1]) && t_isfinite(total));
    return total;
  }
};

// ----------------------------------------------------------------------------------
// Logistic regression, beta_d ~ normal(loc, scale), y_i ~ bernoullilogit(x_i . beta).
// Internal order = api order (lane d % 64, slot d / 64).  Lane l scores rows
// i = l, l + 64, ...; X is stored feature-major xT[d * N + i] so a row step of the
// 64 lanes reads 64 consecutive words of each feature.  The D gradient partials are
// folded with one butterfly each and lane d keeps feature d's total.
template <int KK>
struct LogisticModel {
  static constexpr int K = KK;
  static constexpr int DMAX = KK * WAVE;
  const float* xT_f;
  const double* xT_d;
  const float* y_f;
  const double* y_d;
  long long n;
  int D;
  double loc, scale, total_const;

  template <class T>
  struct Lane {};
  template <class T>
  __device__ void prepare(Lane<T>&) const {}

  __device__ int api(int lane, int slot) const {
    int i = slot * WAVE + lane;
    return i < D ? i : -1;
  }

  template <class T>
  __device__ T eval(const Lane<T>&, const T (&q)[K], T (&g)[K], bool& ok) const {
    const int lane = threadIdx.x & (WAVE - 1);
    T beta[DMAX];
    T gp[DMAX];
#pragma unroll
    for (int d = 0; d < DMAX; ++d) {
      beta[d] = wave_bcast(q[d / WAVE], d % WAVE);
      gp[d] = T(0);
    }
    T lp = T(0);
    for (long long i = lane; i < n; i += WAVE) {
      T eta = T(0);
#pragma unroll
      for (int d = 0; d < DMAX; ++d)
        if (d < D) eta = eta + Pick<T>::get(xT_f, xT_d, (int)(d * n + i)) * beta[d];
      const T yi = Pick<T>::get(y_f, y_d, (int)i);
      lp = lp + (yi * eta - t_log1p_exp<T>(eta));
      const T r = yi - t_logistic<T>(eta);
#pragma unroll
      for (int d = 0; d < DMAX; ++d)
        if (d < D) gp[d] = gp[d] + Pick<T>::get(xT_f, xT_d, (int)(d * n + i)) * r;
    }
    // prior terms for this lane's slots
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (api(lane, k) >= 0) {
        const T z = (q[k] - T(loc)) / T(scale);
        lp = lp - (z * z) / T(2);
      }
    }
    T gl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) gl[k] = T(0);
#pragma unroll
    for (int d = 0; d < DMAX; ++d) {
      if (d < D) {
        const T tot = wave_sum(gp[d]);
        if (lane == d % WAVE) gl[d / WAVE] = tot;
      }
    }
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (api(lane, k) >= 0) {
        const T z = (q[k] - T(loc)) / T(scale);
        g[k] = gl[k] - z / T(scale);
      } else {
        g[k] = T(0);
      }
      fin = fin && t_isfinite(g[k]);
    }
    const T total = wave_sum(lp) + T(total_const);
    ok = wave_all(fin && t_isfinite(total));
    return total;
  }
};

// ----------------------------------------------------------------------------------
// Neal's funnel, v ~ normal(0, v_scale), x_i ~ normal(0, exp(v/2)).
// Internal order: index i < dim is x_{i+1} (api 1 + i), index dim is v (api 0).
// With s_i = x_i^2 e^{-v}:
//   logjoint = -S/2 - dim (v/2 + log(2 pi)/2) + log N(v; 0, v_scale),  S = sum s_i,
//   d/dx_i = -x_i e^{-v},  d/dv = S/2 - dim/2 - v / v_scale^2.
template <int KK>
struct FunnelModel {
  static constexpr int K = KK;
  int dim;
  double v_scale, total_const;

  template <class T>
  struct Lane {};
  template <class T>
  __device__ void prepare(Lane<T>&) const {}

  __device__ int api(int lane, int slot) const {
    int i = slot * WAVE + lane;
    if (i < dim) return 1 + i;
    return i == dim ? 0 : -1;
  }

  template <class T>
  __device__ T eval(const Lane<T>&, const T (&q)[K], T (&g)[K], bool& ok) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const T v = wave_bcast(q[dim / WAVE], dim % WAVE);
    const T ev = t_exp<T>(-v);
    T sp = T(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = k * WAVE + lane;
      if (i < dim) {
        const T s = (q[k] * q[k]) * ev;
        sp = sp + s;
        g[k] = -(q[k] * ev);
      }
    }
    const T S = wave_sum(sp);
    const T zv = v / T(v_scale);
    const T gv = (S / T(2) - T(dim) / T(2)) - zv / T(v_scale);
    const T total = ((-(S / T(2)) - T(dim) * (v / T(2))) - (zv * zv) / T(2)) + T(total_const);
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = k * WAVE + lane;
      if (i == dim) g[k] = gv;
      else if (i > dim) g[k] = T(0);
      fin = fin && t_isfinite(g[k]);
    }
    ok = wave_all(fin && t_isfinite(total));
    return total;
  }
};

}  // namespace ut
