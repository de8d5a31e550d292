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
#include "ut_plan.h"

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
  __device__ static float get(const float* f, const double*, long long i) { return f[i]; }
};
template <> struct Pick<double> {
  __device__ static double get(const float*, const double* d, long long i) { return d[i]; }
};

// Models with LDS of their own (the plan interpreter) take a byte offset into the
// kernel's dynamic LDS; the fused models use none.
struct NoModelLds {
  template <class T>
  size_t lds_bytes() const { return 0; }
  void set_lds_offset(size_t) {}
};

// LDS written by one lane and read by another lane of the same wavefront: order
// the accesses (a wavefront's LDS operations complete in program order).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------------------------
// Non-centred hierarchical normal.  Internal order: slot 0 of lane j is theta_raw_j
// (api 2 + j) for j < J; slot 1 of lane 0 is mu (api 0), of lane 1 is log tau (api 1).
// Lane j scores group j's n_j observations through the group's sufficient statistics
// (mean ybar_j and centred sum of squares S_j, formed in binary64 when the plan is
// staged):
//   sum_i (y_ij - theta_j)^2 = S_j + n_j (ybar_j - theta_j)^2,
//   sum_i log N(y_ij; theta_j, sigma_j) = c_j - hinv_j [S_j + n_j (ybar_j - theta_j)^2]
//   with c_j = -n_j log sigma_j - n_j log(2 pi)/2 and hinv_j = 1 / (2 sigma_j^2),
//   d/dtheta_j = inv_j n_j (ybar_j - theta_j), inv_j = 1 / sigma_j^2,
// an exact rewriting of the per-observation sums (and better conditioned than them:
// no cancellation between observations), so one evaluation costs O(1) per lane
// whatever the group sizes.  Empty groups and lanes >= J contribute zero.
struct HierModel : NoModelLds {
  static constexpr int K = 2;
  const float* lane_f;     // [6][64]: c_j, hinv_j, inv_j, ybar_j, S_j, n_j
  const double* lane_d;
  int J;
  double mu_loc, mu_scale, tau_scale, total_const;
  double inv_mu_scale, inv_tau_scale;

  template <class T>
  struct Lane {
    T cj, hinv, inv, ybar, ss0, n;
  };

  __device__ int api(int lane, int slot) const {
    if (slot == 0) return lane < J ? 2 + lane : -1;
    return lane == 0 ? 0 : (lane == 1 ? 1 : -1);
  }

  template <class T>
  __device__ void prepare(Lane<T>& ld) const {
    const int lane = threadIdx.x & (WAVE - 1);
    ld.cj = Pick<T>::get(lane_f, lane_d, lane);
    ld.hinv = Pick<T>::get(lane_f, lane_d, WAVE + lane);
    ld.inv = Pick<T>::get(lane_f, lane_d, 2 * WAVE + lane);
    ld.ybar = Pick<T>::get(lane_f, lane_d, 3 * WAVE + lane);
    ld.ss0 = Pick<T>::get(lane_f, lane_d, 4 * WAVE + lane);
    ld.n = Pick<T>::get(lane_f, lane_d, 5 * WAVE + lane);
  }

  template <class T>
  __device__ T eval(const Lane<T>& ld, const T (&q)[K], T (&g)[K], bool& ok) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const T mu = wave_bcast(q[1], 0);
    const T ltau = wave_bcast(q[1], 1);
    const T tau = t_exp<T>(ltau);
    const T qr = q[0];
    const T theta = mu + tau * qr;
    const T dlt = ld.ybar - theta;
    const T ss = ld.ss0 + ld.n * (dlt * dlt);
    T lp = ld.cj - ss * ld.hinv;
    T gth = (ld.n * dlt) * ld.inv;
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
//
// One lane's rows (logistic_rows): per row eta = fma(x_{D-1}, b_{D-1}, ... fma(x_0, b_0, 0)),
// lp += y eta - log1p_exp(eta), r = y - logistic(eta), gp_d = fma(x_d, r, gp_d).  DB (>= D) is
// the compile-time bound of the register arrays; the plan stages X with DB feature rows,
// rows D.. zero, and the padded coefficients are zero, so the loop runs over all DB
// features without masks (fma(0, 0, eta) leaves eta's value unchanged, up to the sign of
// a zero eta, which no later step distinguishes).  The covariate loads of a row (or of a
// 32-feature chunk) are issued together so their latencies overlap; with DB <= 32 the
// next row's covariates are loaded while the current row is scored (the row index
// clamped to n - 1 on the last step).
template <int DB, class T>
__device__ inline void logistic_rows(const float* xT_f, const double* xT_d, const float* y_f,
                                     const double* y_d, long long n, int D, long long first,
                                     long long step, const T (&beta)[DB], T (&gp)[DB], T& lp) {
  if constexpr (DB <= 32) {
    T xv[DB];
    T yv = T(0);
    auto load = [&](long long r) {
#pragma unroll
      for (int d = 0; d < DB; ++d) xv[d] = Pick<T>::get(xT_f, xT_d, (long long)d * n + r);
      yv = Pick<T>::get(y_f, y_d, r);
    };
    if (first < n) load(first);
    for (long long i = first; i < n; i += step) {
      T xc[DB];
#pragma unroll
      for (int d = 0; d < DB; ++d) xc[d] = xv[d];
      const T yi = yv;
      load(i + step < n ? i + step : n - 1);
      T eta = T(0);
#pragma unroll
      for (int d = 0; d < DB; ++d) eta = t_fma<T>(xc[d], beta[d], eta);
      lp = lp + (yi * eta - t_log1p_exp<T>(eta));
      const T r = yi - t_logistic<T>(eta);
#pragma unroll
      for (int d = 0; d < DB; ++d) gp[d] = t_fma<T>(xc[d], r, gp[d]);
    }
  } else {
    constexpr int CH = 32;
    for (long long i = first; i < n; i += step) {
      const T yi = Pick<T>::get(y_f, y_d, i);
      T eta = T(0);
#pragma unroll
      for (int c = 0; c < DB; c += CH) {
        if (c < D) {
          T xv[CH];
#pragma unroll
          for (int j = 0; j < CH; ++j) xv[j] = Pick<T>::get(xT_f, xT_d, (long long)(c + j) * n + i);
#pragma unroll
          for (int j = 0; j < CH; ++j) eta = t_fma<T>(xv[j], beta[c + j], eta);
        }
      }
      lp = lp + (yi * eta - t_log1p_exp<T>(eta));
      const T r = yi - t_logistic<T>(eta);
#pragma unroll
      for (int c = 0; c < DB; c += CH) {
        if (c < D) {
          T xv[CH];
#pragma unroll
          for (int j = 0; j < CH; ++j) xv[j] = Pick<T>::get(xT_f, xT_d, (long long)(c + j) * n + i);
#pragma unroll
          for (int j = 0; j < CH; ++j) gp[c + j] = t_fma<T>(xv[j], r, gp[c + j]);
        }
      }
    }
  }
}

template <int KK, int DBB = KK * WAVE>
struct LogisticModel : NoModelLds {
  static constexpr int K = KK;
  static constexpr int DB = DBB;  // D <= DB <= 64 K
  static_assert(DB <= KK * WAVE, "feature bound beyond the slots");
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
    T beta[DB];
    T gp[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      beta[d] = wave_bcast(q[d / WAVE], d % WAVE);
      gp[d] = T(0);
    }
    T lp = T(0);
    logistic_rows<DB, T>(xT_f, xT_d, y_f, y_d, n, D, lane, WAVE, beta, gp, lp);
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
    for (int d = 0; d < DB; ++d) {
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
// Logistic regression with one chain's observations split over the W wavefronts of
// its workgroup (W = 2, 4, 8), for few chains over many rows.  Wavefront w, lane l
// scores rows i = 64 (w + W k) + l, k = 0, 1, ...; wavefront 0's lanes add the prior
// terms of their slots to the partial log density after the rows, as LogisticModel
// does.  Each wavefront folds its partials with wave_sum into one total per
// feature (and one log-density total), writes them to LDS, and after a workgroup
// barrier every wavefront adds the W totals in wavefront order
// ((T_0 + T_1) + T_2) + ... and holds the same values; a second barrier frees the
// exchange area for the next evaluation.  With W = 1 this is LogisticModel's
// arithmetic.  The sampler kernels run the same transition in every wavefront on
// the same values, so all of them reach eval() equally often; wavefront 0 writes
// the results.
template <class M, class = void>
struct WavesOf {
  static constexpr int value = 1;
};
template <class M>
struct WavesOf<M, decltype((void)M::WAVES)> {
  static constexpr int value = M::WAVES;
};

template <int KK, int DBB, int W>
struct TiledLogisticModel : LogisticModel<KK, DBB> {
  static constexpr int K = KK;
  static constexpr int DB = DBB;
  static constexpr int WAVES = W;
  static constexpr int DMAX = KK * WAVE;
  static constexpr int XS = DMAX + 1;  // per-wavefront exchange row: slot totals, log density
  size_t lds_offset = 0;

  template <class T>
  size_t lds_bytes() const { return (size_t)W * XS * sizeof(T); }
  void set_lds_offset(size_t off) { lds_offset = (off + 15) / 16 * 16; }

  template <class T>
  __device__ T eval(const typename LogisticModel<KK, DBB>::template Lane<T>&, const T (&q)[K], T (&g)[K],
                    bool& ok) const {
    extern __shared__ __attribute__((aligned(16))) unsigned char ut_model_lds[];
    T* xs = (T*)(ut_model_lds + lds_offset);
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int D = this->D;
    T beta[DB];
    T gp[DB];
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      beta[d] = wave_bcast(q[d / WAVE], d % WAVE);
      gp[d] = T(0);
    }
    T lp = T(0);
    logistic_rows<DB, T>(this->xT_f, this->xT_d, this->y_f, this->y_d, this->n, D, (long long)wave * WAVE + lane,
                         (long long)W * WAVE, beta, gp, lp);
    if (wave == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (this->api(lane, k) >= 0) {
          const T z = (q[k] - T(this->loc)) / T(this->scale);
          lp = lp - (z * z) / T(2);
        }
      }
    }
    T gl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) gl[k] = T(0);
#pragma unroll
    for (int d = 0; d < DB; ++d) {
      if (d < D) {
        const T tot = wave_sum(gp[d]);
        if (lane == d % WAVE) gl[d / WAVE] = tot;
      }
    }
    const T tl = wave_sum(lp);
    T* mine = xs + wave * XS;
#pragma unroll
    for (int k = 0; k < K; ++k) mine[k * WAVE + lane] = gl[k];
    if (lane == 0) mine[DMAX] = tl;
    __syncthreads();
    T G[K];
#pragma unroll
    for (int k = 0; k < K; ++k) G[k] = xs[k * WAVE + lane];
    T total = xs[DMAX];
#pragma unroll
    for (int w = 1; w < W; ++w) {
#pragma unroll
      for (int k = 0; k < K; ++k) G[k] = G[k] + xs[w * XS + k * WAVE + lane];
      total = total + xs[w * XS + DMAX];
    }
    __syncthreads();
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (this->api(lane, k) >= 0) {
        const T z = (q[k] - T(this->loc)) / T(this->scale);
        g[k] = G[k] - z / T(this->scale);
      } else {
        g[k] = T(0);
      }
      fin = fin && t_isfinite(g[k]);
    }
    total = total + T(this->total_const);
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
struct FunnelModel : NoModelLds {
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

// ----------------------------------------------------------------------------------
// The general static execution plan (ut_plan_desc), interpreted by one wavefront per
// chain.  Parameter row i lives in lane i % 64, register slot i / 64.
//
// eval() runs the plan forward step by step and differentiates it in reverse:
//  1. Top-level steps run uniformly (every lane computes the same values; uniform
//     slot values and adjoints sit in LDS).  A LOOP runs its iterations 64 at a
//     time, iteration t on lane t % 64 in increasing t: each iteration runs the
//     body forward (per-lane slot values), then backward, accumulating per-lane
//     adjoints of the slots the body reads from outside and per-lane partials of
//     the parameters it reads (PARAM nodes, GLM coefficients).  The loop's log
//     density is the wave_sum of the per-lane partial sums, added to the running
//     total at the loop's place; the outer-slot adjoints are wave_summed into the
//     uniform adjoints.
//  2. The parameter partials of the loop bodies are wave_summed into the gradient
//     (rows in increasing order).
//  3. The top-level steps other than loops are revisited in reverse order: each
//     recomputes its arguments, adds its own partials to the gradient and pushes
//     the adjoints of its arguments back through their expressions.
// A top-level DIAG_NORMAL scores its rows on the lanes that hold them and
// wave_sums the partial sums.  The twin of this order is in the oracle.
template <int KK>
struct GenericModel {
  static constexpr int K = KK;
  PlanView pv;
  size_t lds_offset = 0;
  double total_const = 0.0;

  template <class T>
  size_t lds_bytes() const { return plan_lds_bytes<T>(pv); }
  void set_lds_offset(size_t off) { lds_offset = (off + 15) / 16 * 16; }

  template <class T>
  struct Lane {
    T* su;  // [S] uniform slot values
    T* au;  // [S] uniform slot adjoints
    T* sl;  // [S][64] per-lane slot values (loop bodies)
    T* al;  // [S][64] per-lane slot adjoints
    T* nv;  // [N][64] expression node values
    T* na;  // [N][64] expression node adjoints
    T* ap;  // [PL][64] per-lane parameter partials of loop bodies
    T* qu;  // [P] parameters (uniform copy)
  };

  template <class T>
  __device__ void prepare(Lane<T>& ld) const {
    extern __shared__ __attribute__((aligned(16))) unsigned char ut_model_lds[];
    T* p = (T*)(ut_model_lds + lds_offset);
    const int S = pv.num_slots, N = pv.max_nodes, PL = pv.loop_params;
    ld.su = p; p += S;
    ld.au = p; p += S;
    ld.sl = p; p += S * WAVE;
    ld.al = p; p += S * WAVE;
    ld.nv = p; p += N * WAVE;
    ld.na = p; p += N * WAVE;
    ld.ap = p; p += PL * WAVE;
    ld.qu = p;
  }

  __device__ int api(int lane, int slot) const {
    const int i = slot * WAVE + lane;
    return i < pv.num_params ? i : -1;
  }

  // ---- register access by a uniform row
  template <class T>
  __device__ static void add_row(T (&g)[K], int row, T v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (row == k * WAVE + lane) g[k] = g[k] + v;
  }

  // ---- expressions.  `lp` selects the loop-body context (per-lane slots, OBS
  // rows at `cur`, parameter partials to ap); otherwise the top level.
  template <class T>
  __device__ T slot_read(const Lane<T>& ld, int s, bool lp) const {
    const int lane = threadIdx.x & (WAVE - 1);
    return (lp && pv.slot_local[s]) ? ld.sl[s * WAVE + lane] : ld.su[s];
  }

  template <class T>
  __device__ T expr_forward(const Lane<T>& ld, int root, bool lp, long long cur) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const int first = pv.first_node[root];
    for (int n = first; n <= root; ++n) {
      const ut_plan_expr e = pv.exprs[n];
      T v;
      if (e.op == UT_OP_LITERAL) {
        v = (T)e.value;
      } else if (e.op == UT_OP_SLOT) {
        v = slot_read(ld, e.index - 1, lp);
      } else if (e.op == UT_OP_OBS) {
        v = PObs<T>::get(pv, cur + e.index);
      } else if (e.op == UT_OP_PARAM) {
        const int off = e.a >= 0 ? (int)(ld.nv[(e.a - first) * WAVE + lane] + T(0.5)) : 0;
        v = ld.qu[e.index - 1 + off];
      } else {
        const T va = ld.nv[(e.a - first) * WAVE + lane];
        const T vb = plan_op_binary(e.op) ? ld.nv[(e.b - first) * WAVE + lane] : T(0);
        v = plan_node_forward<T>(e.op, va, vb);
      }
      ld.nv[(n - first) * WAVE + lane] = v;
    }
    return ld.nv[(root - first) * WAVE + lane];
  }

  template <class T>
  __device__ void expr_backward(const Lane<T>& ld, int root, T seed, bool lp, long long cur, T (&g)[K]) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const int first = pv.first_node[root];
    expr_forward(ld, root, lp, cur);
    for (int n = first; n <= root; ++n) ld.na[(n - first) * WAVE + lane] = T(0);
    ld.na[(root - first) * WAVE + lane] = seed;
    for (int n = root; n >= first; --n) {
      const ut_plan_expr e = pv.exprs[n];
      const T adj = ld.na[(n - first) * WAVE + lane];
      if (e.op == UT_OP_LITERAL || e.op == UT_OP_OBS) continue;
      if (e.op == UT_OP_SLOT) {
        const int sidx = e.index - 1;
        if (lp) ld.al[sidx * WAVE + lane] = ld.al[sidx * WAVE + lane] + adj;
        else ld.au[sidx] = ld.au[sidx] + adj;
        continue;
      }
      if (e.op == UT_OP_PARAM) {
        const int off = e.a >= 0 ? (int)(ld.nv[(e.a - first) * WAVE + lane] + T(0.5)) : 0;
        const int row = e.index - 1 + off;
        if (lp) ld.ap[row * WAVE + lane] = ld.ap[row * WAVE + lane] + adj;
        else add_row(g, row, adj);
        continue;  // the offset is data: no adjoint
      }
      const T va = ld.nv[(e.a - first) * WAVE + lane];
      const bool bin = plan_op_binary(e.op);
      const T vb = bin ? ld.nv[(e.b - first) * WAVE + lane] : T(0);
      const T v = ld.nv[(n - first) * WAVE + lane];
      T da, db;
      plan_node_partials<T>(e.op, va, vb, v, da, db);
      ld.na[(e.a - first) * WAVE + lane] = ld.na[(e.a - first) * WAVE + lane] + adj * da;
      if (bin) ld.na[(e.b - first) * WAVE + lane] = ld.na[(e.b - first) * WAVE + lane] + adj * db;
    }
  }

  template <class T>
  __device__ void args_forward(const Lane<T>& ld, const ut_plan_step& st, bool lp, long long cur, T (&a)[3]) const {
    const int na = plan_family_arity(st.family);
#pragma unroll
    for (int j = 0; j < 3; ++j) a[j] = j < na ? expr_forward(ld, st.args[j], lp, cur) : T(0);
  }

  // ---- one loop: forward and backward of every iteration
  template <class T>
  __device__ T run_loop(const Lane<T>& ld, int si, T (&g)[K]) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const ut_plan_step lst = pv.steps[si];
    const int count = pv.trip[lst.loop_id - 1], start = pv.start[lst.loop_id - 1];
    const long long base_row = pv.cursor[si];
    const int* info = pv.loop_info + 4 * si;
    const int body0 = si + 1, body1 = si + lst.body_length;
    for (int i = 0; i < info[1]; ++i) ld.al[pv.lists[info[0] + i] * WAVE + lane] = T(0);
    T lpl = T(0);
    for (int t0 = 0; t0 < count; t0 += WAVE) {
      const int t = t0 + lane;
      if (t < count) {
        const long long cur = base_row + (long long)t * lst.stride;
        if (lst.iterator_slot > 0) ld.sl[(lst.iterator_slot - 1) * WAVE + lane] = (T)(start + t);
        // forward
        for (int bi = body0; bi <= body1; ++bi) {
          const ut_plan_step st = pv.steps[bi];
          if (st.kind == UT_STEP_DETERMINISTIC) {
            ld.sl[(st.binding_slot - 1) * WAVE + lane] = expr_forward(ld, st.args[0], true, cur);
            continue;
          }
          if (st.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
            T eta = expr_forward(ld, st.args[0], true, cur);
            for (int d = 0; d < st.dimension; ++d)
              eta = eta + ld.qu[st.coef_source - 1 + d] * PObs<T>::get(pv, cur + st.value_offset + d);
            const T y = PObs<T>::get(pv, cur + st.value_offset + st.dimension);
            const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_BERNOULLI_LOGIT, y, &eta);
            lpl = lpl + fo.lp;
            continue;
          }
          T a[3];
          args_forward(ld, st, true, cur, a);
          const T x = PObs<T>::get(pv, cur + st.value_offset);
          const FamilyOut<T> fo = plan_family<T>(st.family, x, a);
          lpl = lpl + fo.lp;
          if (st.binding_slot > 0) ld.sl[(st.binding_slot - 1) * WAVE + lane] = x;
        }
        // backward: the body's own slots start at zero adjoint every iteration
        for (int i = 0; i < info[3]; ++i) ld.al[pv.lists[info[2] + i] * WAVE + lane] = T(0);
        for (int bi = body1; bi >= body0; --bi) {
          const ut_plan_step st = pv.steps[bi];
          if (st.kind == UT_STEP_DETERMINISTIC) {
            const T adj = ld.al[(st.binding_slot - 1) * WAVE + lane];
            expr_backward(ld, st.args[0], adj, true, cur, g);
            continue;
          }
          if (st.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
            T eta = expr_forward(ld, st.args[0], true, cur);
            for (int d = 0; d < st.dimension; ++d)
              eta = eta + ld.qu[st.coef_source - 1 + d] * PObs<T>::get(pv, cur + st.value_offset + d);
            const T y = PObs<T>::get(pv, cur + st.value_offset + st.dimension);
            const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_BERNOULLI_LOGIT, y, &eta);
            const T r = fo.da[0];
            for (int d = 0; d < st.dimension; ++d) {
              const int row = st.coef_source - 1 + d;
              ld.ap[row * WAVE + lane] = ld.ap[row * WAVE + lane] + r * PObs<T>::get(pv, cur + st.value_offset + d);
            }
            expr_backward(ld, st.args[0], r, true, cur, g);
            continue;
          }
          T a[3];
          args_forward(ld, st, true, cur, a);
          const T x = PObs<T>::get(pv, cur + st.value_offset);
          const FamilyOut<T> fo = plan_family<T>(st.family, x, a);
          const int na = plan_family_arity(st.family);
          for (int j = 0; j < na; ++j) expr_backward(ld, st.args[j], fo.da[j], true, cur, g);
        }
      }
    }
    for (int i = 0; i < info[1]; ++i) {
      const int sidx = pv.lists[info[0] + i];
      const T tot = wave_sum(ld.al[sidx * WAVE + lane]);
      ld.au[sidx] = ld.au[sidx] + tot;
    }
    return wave_sum(lpl);
  }

  // ---- a top-level DIAG_NORMAL: lanes score the rows they hold
  template <class T>
  __device__ T diag_forward(const Lane<T>& ld, const ut_plan_step& st, int si, const T (&q)[K], T mu, T sig) const {
    const int lane = threadIdx.x & (WAVE - 1);
    T part = T(0);
    const T a[3] = {mu, sig, T(0)};
    if (st.value_source > 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r = k * WAVE + lane;
        if (r >= st.value_source - 1 && r < st.value_source - 1 + st.dimension)
          part = part + plan_family<T>(UT_FAMILY_NORMAL, q[k], a).lp;
      }
    } else {
      for (int d = lane; d < st.dimension; d += WAVE)
        part = part + plan_family<T>(UT_FAMILY_NORMAL, PObs<T>::get(pv, pv.cursor[si] + d), a).lp;
    }
    return wave_sum(part);
  }

  template <class T>
  __device__ T eval(const Lane<T>& ld, const T (&q)[K], T (&g)[K], bool& ok) const {
    const int lane = threadIdx.x & (WAVE - 1);
    const int S = pv.num_slots, P = pv.num_params, PL = pv.loop_params;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      g[k] = T(0);
      const int r = k * WAVE + lane;
      if (r < P) ld.qu[r] = q[k];
    }
    for (int sidx = lane; sidx < S; sidx += WAVE) ld.au[sidx] = T(0);
    for (int r = 0; r < PL; ++r) ld.ap[r * WAVE + lane] = T(0);
    wave_lds_sync();
    T total = T(0);
    // 1. forward (loops run forward and backward)
    for (int si = 0; si < pv.num_steps; ++si) {
      const ut_plan_step st = pv.steps[si];
      if (st.kind == UT_STEP_LOOP) {
        total = total + run_loop(ld, si, g);
        si += st.body_length;
        continue;
      }
      if (st.kind == UT_STEP_DETERMINISTIC) {
        const T v = expr_forward(ld, st.args[0], false, 0);
        ld.su[st.binding_slot - 1] = v;
        continue;
      }
      T a[3];
      args_forward(ld, st, false, 0, a);
      if (st.family == UT_FAMILY_DIAG_NORMAL) {
        total = total + diag_forward(ld, st, si, q, a[0], a[1]);
        continue;
      }
      T x, lad = T(0);
      FamilyOut<T> fo;
      if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
        const T z = ld.qu[st.value_source - 1];
        const T std_args[3] = {T(0), T(1), T(0)};
        fo = plan_family<T>(UT_FAMILY_NORMAL, z, std_args);
        x = a[0] + a[1] * z;
      } else {
        if (st.value_source > 0) {
          const TransformOut<T> tr = plan_transform<T>(st.transform, ld.qu[st.value_source - 1]);
          x = tr.c;
          lad = tr.lad;
        } else {
          x = PObs<T>::get(pv, pv.cursor[si]);
        }
        fo = plan_family<T>(st.family, x, a);
      }
      total = total + (fo.lp + lad);
      if (st.binding_slot > 0) ld.su[st.binding_slot - 1] = x;
    }
    // 2. parameter partials of the loop bodies
    for (int r = 0; r < PL; ++r) {
      const T tot = wave_sum(ld.ap[r * WAVE + lane]);
      add_row(g, r, tot);
    }
    // 3. top-level steps in reverse
    for (int si = pv.num_steps - 1; si >= 0; --si) {
      const ut_plan_step st = pv.steps[si];
      if (st.kind == UT_STEP_LOOP) continue;
      // a body step (inside a loop) is skipped: its loop ran it
      if (pv.cursor[si] == -2) continue;
      if (st.kind == UT_STEP_DETERMINISTIC) {
        const T adj = ld.au[st.binding_slot - 1];
        expr_backward(ld, st.args[0], adj, false, 0, g);
        continue;
      }
      T a[3];
      args_forward(ld, st, false, 0, a);
      const int na = plan_family_arity(st.family);
      T seed[3] = {T(0), T(0), T(0)};
      if (st.family == UT_FAMILY_DIAG_NORMAL) {
        T pmu = T(0), psg = T(0);
        const T fa[3] = {a[0], a[1], T(0)};
        if (st.value_source > 0) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const int r = k * WAVE + lane;
            if (r >= st.value_source - 1 && r < st.value_source - 1 + st.dimension) {
              const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_NORMAL, q[k], fa);
              g[k] = g[k] + fo.dx;
              pmu = pmu + fo.da[0];
              psg = psg + fo.da[1];
            }
          }
        } else {
          for (int d = lane; d < st.dimension; d += WAVE) {
            const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_NORMAL, PObs<T>::get(pv, pv.cursor[si] + d), fa);
            pmu = pmu + fo.da[0];
            psg = psg + fo.da[1];
          }
        }
        seed[0] = wave_sum(pmu);
        seed[1] = wave_sum(psg);
      } else if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
        const T z = ld.qu[st.value_source - 1];
        const T std_args[3] = {T(0), T(1), T(0)};
        const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_NORMAL, z, std_args);
        const T adj = st.binding_slot > 0 ? ld.au[st.binding_slot - 1] : T(0);
        add_row(g, st.value_source - 1, fo.dx + adj * a[1]);
        seed[0] = adj;
        seed[1] = adj * z;
      } else {
        if (st.value_source > 0) {
          const TransformOut<T> tr = plan_transform<T>(st.transform, ld.qu[st.value_source - 1]);
          const FamilyOut<T> fo = plan_family<T>(st.family, tr.c, a);
          const T adj_x = st.binding_slot > 0 ? fo.dx + ld.au[st.binding_slot - 1] : fo.dx;
          add_row(g, st.value_source - 1, adj_x * tr.dc + tr.dlad);
          for (int j = 0; j < 3; ++j) seed[j] = fo.da[j];
        } else {
          const FamilyOut<T> fo = plan_family<T>(st.family, PObs<T>::get(pv, pv.cursor[si]), a);
          for (int j = 0; j < 3; ++j) seed[j] = fo.da[j];
        }
      }
      for (int j = 0; j < na; ++j) expr_backward(ld, st.args[j], seed[j], false, 0, g);
    }
    bool fin = t_isfinite(total);
#pragma unroll
    for (int k = 0; k < K; ++k) fin = fin && t_isfinite(g[k]);
    ok = wave_all(fin);
    return total;
  }
};

}  // namespace ut
