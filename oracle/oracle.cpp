// oracle.cpp -- CPU restatement of the reference's batched density and batched MCMC
// algorithms.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
// benchmark's CPU-baseline leg load this library as the checker; the product
// (libuncertaintea_hip) never calls it.
//
// Two evaluation policies share one sampler restatement:
//   * RefF64  -- binary64 throughout, models written with the reference's own
//                per-observation formulas (src/device/math.jl:88-130,202-224) and
//                plain sequential sums.  This is the accuracy oracle and the
//                multi-threaded "reference CPU path" timed by the benchmark.
//   * TwinF32 -- binary32 positions/momenta/gradients with binary64 energies,
//                evaluated in exactly the arithmetic order of the gfx950 kernels
//                (per-lane partial sums over register slots, then the xor
//                butterfly over 64 lanes).  Driven by the same Philox coordinates,
//                it must reproduce the kernels bit for bit, including every NUTS
//                tree-depth and divergence decision.
//
// Reference algorithms restated (paths relative to the reference repository):
//   NUTS transition       src/device/persistent_nuts.jl:226-492
//                          (= src/inference/nuts/tree_dynamics.jl:1-175,
//                             tree_expand.jl:64-133, state_init.jl:284-330)
//   HMC transition        src/device/hmc_kernels.jl:218-280 (leapfrog),
//                          :517-577 (Hamiltonian, accept, divergence)
//   step-size search      src/inference/nuts/adaptation.jl:437-520
//   dual averaging        src/inference/nuts/adaptation.jl:1-28
//   warmup schedule       src/inference/nuts/adaptation.jl:103-138
//   pooled mass           src/inference/warmup.jl:262-472,
//                          src/inference/nuts/adaptation.jl:30-41,140-260,
//                          src/inference/data_types.jl:189-193
//   sampler drivers       src/inference/api_batched_nuts.jl:761-980,
//                          src/device/hmc_kernels.jl:306-723
//   full-rank guide       src/inference/vi/guides.jl:96-117,164-199,459-489
//                          (orc_fullrank_*: pinned by the Gaussian-target optimum,
//                          finite differences and the matrix forms of L, L^-1, L^T)
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/uncertaintea_hip.h"
#include "../uncertaintea_amd/csrc/ut_math.h"
#include "oracle_plan.h"

using namespace ut;

namespace {

constexpr double HALF_LOG_2PI = 0.9189385332046728;
constexpr double LOG_PI = 1.1447298858494002;
constexpr double LOG_2 = 0.6931471805599453;
enum { KIND_LOGISTIC = 1, KIND_HIER = 2, KIND_FUNNEL = 3, KIND_GP = 5, KIND_PLAN = 6 };
enum { S_MOM = 0, S_DIR = 1, S_LEAF = 2, S_MERGE = 3, S_SEARCH = 4, S_INIT = 5, S_ACCEPT = 6 };

// ------------------------------------------------------------------ model data
struct Model {
  int kind = 0;
  int P = 0;
  // logistic
  std::vector<float> x;  // N x D row-major
  std::vector<float> y;
  long long n = 0;
  int D = 0;
  double loc = 0, scale = 1;
  int tiles = 1;  // wavefronts per chain (the workgroup-per-chain summation order)
  // hierarchical
  std::vector<float> hy;
  std::vector<int> offsets;
  std::vector<double> sigma;
  int J = 0;
  double mu_loc = 0, mu_scale = 1, tau_scale = 1;
  // funnel
  int dim = 0;
  double v_scale = 1;
  // latent GP Poisson: inputs (N x D), counts, kernel nugget, priors
  std::vector<double> gx, gy;
  int gn = 0, gd = 0;
  double noise2 = 0, ls_loc = 0, ls_scale = 0.5, amp_scale = 1;
  // general plan (interpreter restatements in oracle_plan.h)
  std::shared_ptr<plan_oracle::PlanData> plan;
  // twin layout
  int K = 1;
  double total_const = 0;
  std::vector<int> tnobs;
  std::vector<double> tlane;    // [6][64], binary64 (binary32 kernels round each)
  int max_n = 0;

  int api(int lane, int slot) const {
    const int i = slot * WAVE + lane;
    switch (kind) {
      case KIND_LOGISTIC:
        return i < D ? i : -1;
      case KIND_HIER:
        if (slot == 0) return lane < J ? 2 + lane : -1;
        if (slot == 1) return lane == 0 ? 0 : (lane == 1 ? 1 : -1);
        return -1;
      case KIND_FUNNEL:
        if (i < dim) return 1 + i;
        return i == dim ? 0 : -1;
      case KIND_PLAN:
        return i < P ? i : -1;
    }
    return -1;
  }
};

int slots_for(int P) {
  int k = (P + WAVE - 1) / WAVE;
  if (k <= 1) return 1;
  if (k <= 2) return 2;
  return 4;
}

// ------------------------------------------------------------------ RefF64 models
// log N(x; mu, sigma) as the reference writes it (math.jl:88-91)
inline double normal_lpdf(double mu, double sigma, double x) {
  const double z = (x - mu) / sigma;
  return -std::log(sigma) - HALF_LOG_2PI - z * z / 2.0;
}
inline double log1p_exp_ref(double eta) {
  return eta > 0.0 ? eta + std::log1p(std::exp(-eta)) : std::log1p(std::exp(eta));
}

// logjoint + gradient in api order (binary64, reference formulas)
// latent GP Poisson in binary64: K and its tangents along each log hyperparameter,
// an unblocked Cholesky carrying the tangents (dL_jj = dA_jj / (2 L_jj),
// dL_ij = (dA_ij - L_ij dL_jj) / L_jj, d(A - L L^T) = dA - dL L^T - L dL^T), then
// f = L z, the Poisson terms and the priors (reference gp_cholesky,
// distributions/gaussian_process.jl:180-188,283-294; poisson math.jl:230-235)
double gp_eval(const Model& m, const double* q, double* g, double* logdet_out = nullptr) {
  const int N = m.gn, D = m.gd, T = D + 1;
  std::vector<double> A((size_t)(T + 1) * N * N, 0.0);
  auto at = [&](int comp, int i, int j) -> double& { return A[((size_t)comp * N + i) * N + j]; };
  for (int i = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = 0.0, terms[2] = {0.0, 0.0};
      for (int d = 0; d < D; ++d) {
        const double diff = (m.gx[(size_t)i * D + d] - m.gx[(size_t)j * D + d]) * std::exp(-q[d]);
        terms[d] = diff * diff;
        s += terms[d];
      }
      const double k = std::exp(2.0 * q[D]) * std::exp(-0.5 * s);
      at(0, i, j) = k + (i == j ? m.noise2 : 0.0);
      for (int d = 0; d < D; ++d) at(1 + d, i, j) = k * terms[d];
      at(T, i, j) = 2.0 * k;
    }
  for (int j = 0; j < N; ++j) {
    const double l = std::sqrt(at(0, j, j));
    at(0, j, j) = l;
    for (int k = 1; k <= T; ++k) at(k, j, j) /= 2.0 * l;
    for (int i = j + 1; i < N; ++i) {
      const double v = at(0, i, j) / l;
      for (int k = 1; k <= T; ++k) at(k, i, j) = (at(k, i, j) - v * at(k, j, j)) / l;
      at(0, i, j) = v;
    }
    for (int i = j + 1; i < N; ++i)
      for (int kk = j + 1; kk <= i; ++kk) {
        const double lij = at(0, i, j), lkj = at(0, kk, j);
        at(0, i, kk) -= lij * lkj;
        for (int k = 1; k <= T; ++k) at(k, i, kk) -= at(k, i, j) * lkj + lij * at(k, kk, j);
      }
  }
  const double* z = q + D + 1;
  std::vector<double> r(N);
  double ll = 0.0, zz = 0.0, ld = 0.0, gh[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < N; ++i) {
    double f = 0.0, df[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j <= i; ++j) {
      f += at(0, i, j) * z[j];
      for (int k = 0; k < T; ++k) df[k] += at(k + 1, i, j) * z[j];
    }
    const double ef = std::exp(f);
    r[i] = m.gy[i] - ef;
    ll += m.gy[i] * f - ef - std::lgamma(m.gy[i] + 1.0);
    for (int k = 0; k < T; ++k) gh[k] += r[i] * df[k];
    zz += z[i] * z[i];
    ld += std::log(at(0, i, i));
  }
  for (int j = 0; j < N; ++j) {
    double s = 0.0;
    for (int i = j; i < N; ++i) s += at(0, i, j) * r[i];
    g[D + 1 + j] = s - z[j];
  }
  double lp = 0.0;
  for (int d = 0; d < D; ++d) {
    const double w = (q[d] - m.ls_loc) / m.ls_scale;
    lp += -0.5 * w * w - std::log(m.ls_scale) - HALF_LOG_2PI;
    g[d] = gh[d] - w / m.ls_scale;
  }
  const double a = std::exp(q[D]), wa = a / m.amp_scale;
  lp += LOG_2 - HALF_LOG_2PI - std::log(m.amp_scale) - 0.5 * wa * wa + q[D];
  g[D] = gh[D] - wa * wa + 1.0;
  if (logdet_out) *logdet_out = 2.0 * ld;
  return ll + lp - 0.5 * zz - N * HALF_LOG_2PI;
}

double ref_eval(const Model& m, const double* q, double* g, bool& ok) {
  double lp = 0.0;
  const int P = m.P;
  for (int a = 0; a < P; ++a) g[a] = 0.0;
  if (m.kind == KIND_LOGISTIC) {
    for (int d = 0; d < m.D; ++d) {
      lp += normal_lpdf(m.loc, m.scale, q[d]);
      g[d] += -(q[d] - m.loc) / (m.scale * m.scale);
    }
    for (long long i = 0; i < m.n; ++i) {
      double eta = 0.0;
      for (int d = 0; d < m.D; ++d) eta += (double)m.x[i * m.D + d] * q[d];
      const double yi = m.y[i];
      lp += yi * eta - log1p_exp_ref(eta);
      const double r = yi - 1.0 / (1.0 + std::exp(-eta));
      for (int d = 0; d < m.D; ++d) g[d] += (double)m.x[i * m.D + d] * r;
    }
  } else if (m.kind == KIND_HIER) {
    const double mu = q[0], ltau = q[1], tau = std::exp(ltau);
    lp += normal_lpdf(m.mu_loc, m.mu_scale, mu);
    g[0] += -(mu - m.mu_loc) / (m.mu_scale * m.mu_scale);
    const double zt = tau / m.tau_scale;
    lp += LOG_2 - LOG_PI - std::log(m.tau_scale) - std::log1p(zt * zt) + ltau;
    g[1] += -2.0 * zt * zt / (1.0 + zt * zt) + 1.0;
    for (int j = 0; j < m.J; ++j) {
      const double r = q[2 + j];
      lp += normal_lpdf(0.0, 1.0, r);
      g[2 + j] += -r;
      const double theta = mu + tau * r;
      double gth = 0.0;
      for (int i = m.offsets[j]; i < m.offsets[j + 1]; ++i) {
        lp += normal_lpdf(theta, m.sigma[j], m.hy[i]);
        gth += (m.hy[i] - theta) / (m.sigma[j] * m.sigma[j]);
      }
      g[0] += gth;
      g[1] += gth * tau * r;
      g[2 + j] += gth * tau;
    }
  } else if (m.kind == KIND_GP) {
    lp = gp_eval(m, q, g);
  } else if (m.kind == KIND_PLAN) {
    lp = plan_oracle::plan_ref_eval(*m.plan, q, g);
  } else if (m.kind == KIND_FUNNEL) {
    const double v = q[0];
    lp += normal_lpdf(0.0, m.v_scale, v);
    g[0] += -v / (m.v_scale * m.v_scale);
    const double s = std::exp(v / 2.0);
    for (int i = 0; i < m.dim; ++i) {
      const double x = q[1 + i];
      lp += normal_lpdf(0.0, s, x);
      g[1 + i] += -x / (s * s);
      g[0] += 0.5 * x * x / (s * s) - 0.5;
    }
  }
  ok = std::isfinite(lp);
  for (int a = 0; a < P && ok; ++a) ok = std::isfinite(g[a]);
  return lp;
}

// ------------------------------------------------------------------ TwinF32 models
// Vectors are K x 64 (slot-major).  bfly() is the kernels' wavefront reduction:
// lane l adds lane l^1, l^2, l^7 and l^15 (DPP quad permutes, half-row mirror and
// row mirror inside each 16-lane row), then the four row totals combine as
// (R0 + R1) + (R2 + R3).
template <class T>
T bfly(const T* in) {
  T v[WAVE], w[WAVE];
  for (int l = 0; l < WAVE; ++l) v[l] = in[l];
  const int partner_mask[4] = {1, 2, 7, 15};
  for (int s = 0; s < 4; ++s) {
    for (int l = 0; l < WAVE; ++l) w[l] = v[l] + v[l ^ partner_mask[s]];
    for (int l = 0; l < WAVE; ++l) v[l] = w[l];
  }
  return (v[0] + v[16]) + (v[32] + v[48]);
}

template <class T> T texp(T x);
template <> float texp<float>(float x) { return exp_f(x); }
template <> double texp<double>(double x) { return exp_d(x); }
template <class T> T tlog1p(T x);
template <> float tlog1p<float>(float x) { return log1p_f(x); }
template <> double tlog1p<double>(double x) { return log1p_d(x); }
template <class T> T tl1pe(T x);
template <> float tl1pe<float>(float x) { return log1p_exp_f(x); }
template <> double tl1pe<double>(double x) { return log1p_exp_d(x); }
template <class T> T tfma(T a, T b, T c);
template <> float tfma<float>(float a, float b, float c) { return fma_f(a, b, c); }
template <> double tfma<double>(double a, double b, double c) { return fma_d(a, b, c); }
template <class T> T tlogistic(T x);
template <> float tlogistic<float>(float x) { return logistic_f(x); }
template <> double tlogistic<double>(double x) { return logistic_d(x); }
template <class T> bool tfin(T x) { return std::isfinite(x); }

template <class T>
T twin_eval(const Model& m, const T* q, T* g, bool& ok) {
  const int K = m.K;
  bool fin = true;
  if (m.kind == KIND_PLAN) return plan_oracle::plan_twin_eval<T>(*m.plan, K, q, g, ok);

  if (m.kind == KIND_HIER) {
    const T mu = q[1 * WAVE + 0], ltau = q[1 * WAVE + 1];
    const T tau = texp<T>(ltau);
    T gth[WAVE], gq[WAVE], lpl[WAVE];
    for (int l = 0; l < WAVE; ++l) {
      // the group's observations through their sufficient statistics (ut_models.h HierModel)
      const T qr = q[l];
      const T theta = mu + tau * qr;
      const T cj = T(m.tlane[l]), hinv = T(m.tlane[WAVE + l]), inv = T(m.tlane[2 * WAVE + l]);
      const T ybar = T(m.tlane[3 * WAVE + l]), ss0 = T(m.tlane[4 * WAVE + l]), n = T(m.tlane[5 * WAVE + l]);
      const T dlt = ybar - theta;
      const T ss = ss0 + n * (dlt * dlt);
      T lp = cj - ss * hinv;
      gth[l] = (n * dlt) * inv;
      if (l >= m.J) {
        lp = T(0);
        gth[l] = T(0);
      }
      lp = lp - T(0.5) * qr * qr;
      g[l] = tau * gth[l] - qr;
      gq[l] = gth[l] * qr;
      lpl[l] = lp;
    }
    const T s1 = bfly(gth), s2 = bfly(gq), sl = bfly(lpl);
    const T zmu = (mu - T(m.mu_loc)) * T(1.0 / m.mu_scale);
    const T lp_mu = -(zmu * zmu) * T(0.5);
    const T g_mu = s1 - zmu * T(1.0 / m.mu_scale);
    const T zt = tau * T(1.0 / m.tau_scale);
    const T z2 = zt * zt;
    const T lp_tau = ltau - tlog1p<T>(z2);
    const T g_ltau = tau * s2 - (T(2) * z2) / (T(1) + z2) + T(1);
    const T total = ((sl + lp_mu) + lp_tau) + T(m.total_const);
    for (int l = 0; l < WAVE; ++l) g[WAVE + l] = l == 0 ? g_mu : (l == 1 ? g_ltau : T(0));
    for (int i = 0; i < 2 * WAVE; ++i) fin = fin && tfin(g[i]);
    ok = fin && tfin(total);
    return total;
  }
  if (m.kind == KIND_LOGISTIC) {
    // W = m.tiles wavefronts per chain (ut_models.h TiledLogisticModel; W = 1 is
    // LogisticModel): wavefront w, lane l takes rows 64 (w + W k) + l; wavefront 0's
    // lanes add the prior terms; per-wavefront butterfly totals are added in
    // wavefront order.
    const int DMAX = K * WAVE;
    const int NW = m.tiles;
    std::vector<T> beta(DMAX), gp((size_t)DMAX * WAVE), lpl(WAVE);
    for (int d = 0; d < DMAX; ++d) beta[d] = q[(d / WAVE) * WAVE + d % WAVE];
    std::vector<T> gl((size_t)K * WAVE, T(0));
    T lp_total = T(0);
    for (int w = 0; w < NW; ++w) {
      std::fill(gp.begin(), gp.end(), T(0));
      for (int l = 0; l < WAVE; ++l) {
        T lp = T(0);
        T* gpl = &gp[(size_t)l * DMAX];
        for (long long i = (long long)w * WAVE + l; i < m.n; i += (long long)NW * WAVE) {
          T eta = T(0);
          for (int d = 0; d < m.D; ++d) eta = tfma<T>(T(m.x[i * m.D + d]), beta[d], eta);
          const T yi = T(m.y[i]);
          lp = lp + (yi * eta - tl1pe<T>(eta));
          const T r = yi - tlogistic<T>(eta);
          for (int d = 0; d < m.D; ++d) gpl[d] = tfma<T>(T(m.x[i * m.D + d]), r, gpl[d]);
        }
        if (w == 0) {
          for (int k = 0; k < K; ++k) {
            if (m.api(l, k) >= 0) {
              const T z = (q[k * WAVE + l] - T(m.loc)) / T(m.scale);
              lp = lp - (z * z) / T(2);
            }
          }
        }
        lpl[l] = lp;
      }
      T col[WAVE];
      for (int d = 0; d < m.D; ++d) {
        for (int l = 0; l < WAVE; ++l) col[l] = gp[(size_t)l * DMAX + d];
        const T tot = bfly(col);
        T& dst = gl[(d / WAVE) * WAVE + d % WAVE];
        dst = w == 0 ? tot : dst + tot;
      }
      const T tl = bfly(lpl.data());
      lp_total = w == 0 ? tl : lp_total + tl;
    }
    for (int k = 0; k < K; ++k)
      for (int l = 0; l < WAVE; ++l) {
        const int i = k * WAVE + l;
        if (m.api(l, k) >= 0) {
          const T z = (q[i] - T(m.loc)) / T(m.scale);
          g[i] = gl[i] - z / T(m.scale);
        } else {
          g[i] = T(0);
        }
        fin = fin && tfin(g[i]);
      }
    const T total = lp_total + T(m.total_const);
    ok = fin && tfin(total);
    return total;
  }
  // funnel
  const T v = q[(m.dim / WAVE) * WAVE + m.dim % WAVE];
  const T ev = texp<T>(-v);
  T spl[WAVE];
  for (int l = 0; l < WAVE; ++l) {
    T sp = T(0);
    for (int k = 0; k < K; ++k) {
      const int i = k * WAVE + l;
      if (i < m.dim) {
        const T s = (q[i] * q[i]) * ev;
        sp = sp + s;
        g[i] = -(q[i] * ev);
      }
    }
    spl[l] = sp;
  }
  const T S = bfly(spl);
  const T zv = v / T(m.v_scale);
  const T gv = (S / T(2) - T(m.dim) / T(2)) - zv / T(m.v_scale);
  const T total = ((-(S / T(2)) - T(m.dim) * (v / T(2))) - (zv * zv) / T(2)) + T(m.total_const);
  for (int k = 0; k < K; ++k)
    for (int l = 0; l < WAVE; ++l) {
      const int i = k * WAVE + l;
      if (i == m.dim) g[i] = gv;
      else if (i > m.dim) g[i] = T(0);
      fin = fin && tfin(g[i]);
    }
  ok = fin && tfin(total);
  return total;
}

// ------------------------------------------------------------------ policies
// A policy owns the vector layout, the reductions and the model evaluation.
struct TwinF32 {
  using T = float;
  const Model* m;
  int n() const { return m->K * WAVE; }
  int api(int i) const { return m->api(i % WAVE, i / WAVE); }
  T eval(const T* q, T* g, bool& ok) const { return twin_eval<float>(*m, q, g, ok); }
  double kinetic(const T* p, const T* im) const {
    double lane[WAVE];
    for (int l = 0; l < WAVE; ++l) {
      double acc = 0.0;
      for (int k = 0; k < m->K; ++k) {
        const int i = k * WAVE + l;
        acc = acc + ((double)p[i] * (double)p[i]) * (double)im[i];
      }
      lane[l] = acc;
    }
    return bfly(lane) * 0.5;
  }
  // sum_i (a_i * im_i) * b_i and (a_i * im_i) * c_i with a = x - y, per-lane then butterfly
  bool turning(const T* x, const T* y, const T* pl, const T* pr, const T* im) const {
    float la[WAVE], ra[WAVE];
    for (int l = 0; l < WAVE; ++l) {
      float a = 0.0f, b = 0.0f;
      for (int k = 0; k < m->K; ++k) {
        const int i = k * WAVE + l;
        const float delta = x[i] - y[i];
        a = a + (delta * im[i]) * pl[i];
        b = b + (delta * im[i]) * pr[i];
      }
      la[l] = a;
      ra[l] = b;
    }
    const float ld = bfly(la), rd = bfly(ra);
    return ld <= 0.0f || rd <= 0.0f;
  }
  // lock-step kernels keep parameter a (reference order) in lane a % 64, slot a / 64
  double ls_kinetic(const T* p, const T* im) const {
    double lane[WAVE];
    for (int l = 0; l < WAVE; ++l) {
      double acc = 0.0;
      for (int a = l; a < m->P; a += WAVE) acc = acc + ((double)p[a] * (double)p[a]) * (double)im[a];
      lane[l] = acc;
    }
    return bfly(lane) * 0.5;
  }
  // the density kernel on reference-order vectors (P entries)
  T ls_eval(const T* q_ref, T* g_ref) const {
    if (m->kind == KIND_GP) {
      // no kernel-order twin for the GP (blocked device Cholesky): binary64, rounded
      std::vector<double> qd(m->P), gd(m->P);
      for (int a = 0; a < m->P; ++a) qd[a] = (double)q_ref[a];
      const double v = gp_eval(*m, qd.data(), gd.data());
      for (int a = 0; a < m->P; ++a) g_ref[a] = (T)gd[a];
      return (T)v;
    }
    std::vector<T> q(n(), T(0)), g(n(), T(0));
    for (int i = 0; i < n(); ++i) {
      const int a = api(i);
      if (a >= 0) q[i] = q_ref[a];
    }
    bool ok;
    const T v = eval(q.data(), g.data(), ok);
    for (int i = 0; i < n(); ++i) {
      const int a = api(i);
      if (a >= 0) g_ref[a] = g[i];
    }
    return v;
  }
};

struct RefF64 {
  using T = double;
  const Model* m;
  int n() const { return m->P; }
  int api(int i) const { return i; }
  T eval(const T* q, T* g, bool& ok) const { return ref_eval(*m, q, g, ok); }
  double kinetic(const T* p, const T* im) const {
    double acc = 0.0;
    for (int i = 0; i < m->P; ++i) acc += p[i] * p[i] * im[i];
    return acc * 0.5;
  }
  bool turning(const T* x, const T* y, const T* pl, const T* pr, const T* im) const {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < m->P; ++i) {
      const double delta = x[i] - y[i];
      a += delta * im[i] * pl[i];
      b += delta * im[i] * pr[i];
    }
    return a <= 0.0 || b <= 0.0;
  }
  double ls_kinetic(const T* p, const T* im) const { return kinetic(p, im); }
  T ls_eval(const T* q, T* g) const {
    bool ok;
    return eval(q, g, ok);
  }
};

template <class T> T sqrt_t(T x);
template <> float sqrt_t<float>(float x) { return sqrt_f(x); }
template <> double sqrt_t<double>(double x) { return sqrt_d(x); }
template <class T> T draw_normal(uint32_t c, uint32_t itc, uint32_t s, uint32_t a) {
  return (T)rand_normal_f(c, itc, s, a);
}

struct TransOut {
  double accept_prob = 0, energy = 0, error = 0;
  int depth = 0, steps = 0;
  bool divergent = false, moved = false;
};

// ------------------------------------------------------------------ NUTS
// One chain's multinomial NUTS transition; vectors have pol.n() entries.
// One NUTS subtree of nleaves = 2^d leapfrog leaves grown from (cq, p, wg) in one direction, with
// the U-turn test of every dyadic sub-subtree evaluated from checkpoints instead of recursion:
// leaf li (0-based) that is even stores its (q, p) in slot popcount(li); an odd leaf closes the
// 2^k-leaf subtrees for k = 1..ctz(~li), whose leftmost leaf li - 2^k + 1 sits in slot
// popcount(li - 2^k + 1). This visits the same subtrees, in the same order, as the recursive
// build_tree that test/uncertaintea/core/nuts_uturn_turning.jl:25-55 restates; the reference's own
// _build_nuts_subtree (src/inference/nuts/tree_dynamics.jl:1-93) keeps the same checkpoints. leaf(lj) runs the
// per-leaf energy/multinomial work and returns false for a divergent leaf.
template <class Pol, class Leaf>
void dyadic_subtree(const Pol& pol, const typename Pol::T* im, typename Pol::T ss, bool fwd, int nleaves,
                    typename Pol::T* cq, typename Pol::T* p, typename Pol::T* wg, typename Pol::T& cur_lj,
                    typename Pol::T* ck, int& sub_int, bool& sub_div, bool& sub_turn, Leaf&& leaf) {
  using T = typename Pol::T;
  const int n = pol.n();
  const T half = ss / T(2);
  auto cp = [n](T* d, const T* s) { std::memcpy(d, s, sizeof(T) * n); };
  for (int l = 0; l < nleaves; ++l) {
    for (int i = 0; i < n; ++i) {
      p[i] = p[i] + half * wg[i];
      cq[i] = cq[i] + ss * (im[i] * p[i]);
    }
    bool ok;
    const T ljc = pol.eval(cq, wg, ok);
    if (!ok) { sub_div = true; return; }
    for (int i = 0; i < n; ++i) p[i] = p[i] + half * wg[i];
    cur_lj = ljc;
    sub_int += 1;
    if (!leaf(ljc)) { sub_div = true; return; }
    const int li = sub_int - 1;
    if ((li & 1) == 0) {
      const int slot = __builtin_popcount((unsigned)li);
      cp(&ck[(size_t)(slot * 2 + 0) * n], cq);
      cp(&ck[(size_t)(slot * 2 + 1) * n], p);
    } else {
      const int nk = __builtin_ctz(~(unsigned)li);
      bool turned = false;
      for (int k = 1; k <= nk && !turned; ++k) {
        const int bs = li - (1 << k) + 1;
        const int slot = __builtin_popcount((unsigned)bs);
        const T* cpos = &ck[(size_t)(slot * 2 + 0) * n];
        const T* cmom = &ck[(size_t)(slot * 2 + 1) * n];
        if (fwd) turned = pol.turning(cq, cpos, cmom, p, im);
        else turned = pol.turning(cpos, cq, p, cmom, im);
      }
      if (turned) { sub_turn = true; return; }
    }
  }
}

template <class Pol>
void nuts_transition(const Pol& pol, const typename Pol::T* im, typename Pol::T h, uint32_t chain,
                     uint32_t itc, int max_depth, double div_thr, typename Pol::T* q,
                     typename Pol::T* g, typename Pol::T& lj, TransOut& out) {
  using T = typename Pol::T;
  const int n = pol.n();
  std::vector<T> buf((size_t)n * 16);
  T* p = &buf[0 * n];
  T* q0 = &buf[1 * n];
  T* Lq = &buf[2 * n]; T* Lp = &buf[3 * n]; T* Lg = &buf[4 * n];
  T* Rq = &buf[5 * n]; T* Rp = &buf[6 * n]; T* Rg = &buf[7 * n];
  T* sq = &buf[8 * n]; T* sg = &buf[9 * n];
  T* pq = &buf[10 * n]; T* pg = &buf[11 * n];
  T* wg = &buf[12 * n];
  std::vector<T> ck((size_t)(max_depth + 1) * 2 * n);
  auto cp = [n](T* d, const T* s) { std::memcpy(d, s, sizeof(T) * n); };

  for (int i = 0; i < n; ++i) {
    const int a = pol.api(i);
    p[i] = a >= 0 ? draw_normal<T>(chain, itc, S_MOM, (uint32_t)a) / sqrt_t<T>(im[i]) : T(0);
  }
  const double H0 = pol.kinetic(p, im) - (double)lj;
  cp(q0, q);
  cp(wg, g);
  cp(Lq, q); cp(Lp, p); cp(Lg, g);
  cp(Rq, q); cp(Rp, p); cp(Rg, g);
  cp(pq, q); cp(pg, g);
  cp(sq, q); cp(sg, g);
  T ljL = lj, ljR = lj, ljP = lj, ljS = lj;
  double cont_w = -H0, accept_sum = 0.0;
  int accept_count = 0, depth = 1, steps = 0;
  bool divergent = false, turning = false;
  double prop_energy = H0, prop_error = 0.0;
  uint32_t leaf_counter = 0;
  T* cq = q;  // working position

  {
    const float u = rand_uniform_f(chain, itc, S_DIR, 0u);
    const bool fwd = !(u < 0.5f);
    const T ss = fwd ? h : -h;
    const T half = ss / T(2);
    for (int i = 0; i < n; ++i) {
      p[i] = p[i] + half * wg[i];
      cq[i] = cq[i] + ss * (im[i] * p[i]);
    }
    bool ok;
    const T lj1 = pol.eval(cq, wg, ok);
    if (!ok) {
      divergent = true;
    } else {
      for (int i = 0; i < n; ++i) p[i] = p[i] + half * wg[i];
      if (fwd) { cp(Rq, cq); cp(Rp, p); cp(Rg, wg); ljR = lj1; }
      else { cp(Lq, cq); cp(Lp, p); cp(Lg, wg); ljL = lj1; }
      const double pe = pol.kinetic(p, im) - (double)lj1;
      const double delta = pe - H0;
      steps = 1;
      if (!std::isfinite(delta) || delta > div_thr) {
        divergent = true;
      } else {
        const double e = exp_d(std::fmin(0.0, -delta));
        accept_sum += e < 1.0 ? e : 1.0;
        accept_count += 1;
        const double cand = -pe;
        const double comb = logaddexp_d(cont_w, cand);
        const double ul = rand_uniform_d(chain, itc, S_LEAF, leaf_counter);
        if (!std::isfinite(cont_w) || log_d(ul) < cand - comb) {
          cp(pq, cq); cp(pg, wg); ljP = lj1;
          prop_energy = pe;
          prop_error = delta;
        }
        cont_w = comb;
        turning = pol.turning(Rq, Lq, Lp, Rp, im);
      }
      leaf_counter += 1;
    }
  }

  while (depth < max_depth && !divergent && !turning) {
    const uint32_t round = (uint32_t)depth;
    const float ud = rand_uniform_f(chain, itc, S_DIR, round);
    const bool fwd = !(ud < 0.5f);
    const T ss = fwd ? h : -h;
    if (fwd) { cp(cq, Rq); cp(p, Rp); cp(wg, Rg); }
    else { cp(cq, Lq); cp(p, Lp); cp(wg, Lg); }
    T cur_lj = fwd ? ljR : ljL;
    double sub_w = -inf_d(), sub_asum = 0.0, sub_energy = 0.0, sub_error = 0.0;
    int sub_int = 0, sub_acount = 0;
    bool sub_div = false, sub_turn = false;
    const int nleaves = 1 << depth;
    dyadic_subtree(pol, im, ss, fwd, nleaves, cq, p, wg, cur_lj, ck.data(), sub_int, sub_div, sub_turn,
                   [&](T ljc) {
                     const double pe = pol.kinetic(p, im) - (double)ljc;
                     const double delta = pe - H0;
                     if (!std::isfinite(delta) || delta > div_thr) return false;
                     const double e = exp_d(std::fmin(0.0, -delta));
                     sub_asum += e < 1.0 ? e : 1.0;
                     sub_acount += 1;
                     const double cand = -pe;
                     const double comb = logaddexp_d(sub_w, cand);
                     const double ul = rand_uniform_d(chain, itc, S_LEAF, leaf_counter);
                     if (!std::isfinite(sub_w) || log_d(ul) < cand - comb) {
                       cp(sq, cq); cp(sg, wg); ljS = ljc;
                       sub_energy = pe;
                       sub_error = delta;
                     }
                     sub_w = comb;
                     leaf_counter += 1;
                     return true;
                   });
    depth += 1;
    if (sub_int == 0) { divergent = sub_div; break; }
    if (sub_turn || sub_div) {
      steps += sub_int;
      accept_sum += sub_asum;
      accept_count += sub_acount;
      turning = sub_turn;
      divergent = sub_div;
      break;
    }
    if (fwd) { cp(Rq, cq); cp(Rp, p); cp(Rg, wg); ljR = cur_lj; }
    else { cp(Lq, cq); cp(Lp, p); cp(Lg, wg); ljL = cur_lj; }
    double comb_cw = cont_w;
    if (std::isfinite(sub_w)) {
      comb_cw = logaddexp_d(cont_w, sub_w);
      const double um = rand_uniform_d(chain, itc, S_MERGE, round);
      if (log_d(um) < sub_w - cont_w) {
        cp(pq, sq); cp(pg, sg); ljP = ljS;
        prop_energy = sub_energy;
        prop_error = sub_error;
      }
    }
    steps += sub_int;
    accept_sum += sub_asum;
    accept_count += sub_acount;
    if (std::isfinite(sub_w)) cont_w = comb_cw;
    turning = sub_turn || pol.turning(Rq, Lq, Lp, Rp, im);
  }
  bool moved = false;
  for (int i = 0; i < n; ++i) moved = moved || (pq[i] != q0[i]);
  out.moved = moved;
  out.accept_prob = accept_count == 0 ? 0.0 : accept_sum / (double)accept_count;
  out.energy = prop_energy;
  out.error = prop_error;
  out.divergent = divergent;
  out.depth = depth;
  out.steps = steps;
  if (moved) {
    cp(q, pq); cp(g, pg); lj = ljP;
  } else {
    cp(q, q0);
  }
}

// ------------------------------------------------------------------ HMC
template <class Pol>
void hmc_transition(const Pol& pol, const typename Pol::T* im, typename Pol::T h, uint32_t chain,
                    uint32_t itc, int L, double div_thr, uint32_t mom_stream, uint32_t acc_stream,
                    typename Pol::T* q, typename Pol::T* g, typename Pol::T& lj, TransOut& out,
                    double* lar_out) {
  using T = typename Pol::T;
  const int n = pol.n();
  std::vector<T> p(n), nq(q, q + n), ng(g, g + n);
  for (int i = 0; i < n; ++i) {
    const int a = pol.api(i);
    p[i] = a >= 0 ? draw_normal<T>(chain, itc, mom_stream, (uint32_t)a) / sqrt_t<T>(im[i]) : T(0);
  }
  const double H0 = pol.kinetic(p.data(), im) - (double)lj;
  T nlj = lj;
  bool valid = true;
  for (int i = 0; i < n; ++i) valid = valid && std::isfinite(g[i]);
  const T half = T(0.5) * h;
  if (valid)
    for (int i = 0; i < n; ++i) p[i] = p[i] + half * ng[i];
  for (int s = 1; s <= L; ++s) {
    if (!valid) break;
    for (int i = 0; i < n; ++i) nq[i] = nq[i] + (h * im[i]) * p[i];
    bool ok;
    nlj = pol.eval(nq.data(), ng.data(), ok);
    if (s < L) {
      bool gfin = true;
      for (int i = 0; i < n; ++i) gfin = gfin && std::isfinite(ng[i]);
      valid = valid && gfin;
      if (valid)
        for (int i = 0; i < n; ++i) p[i] = p[i] + (T(1) * h) * ng[i];
    } else {
      valid = valid && ok;
    }
  }
  if (valid)
    for (int i = 0; i < n; ++i) p[i] = -(p[i] + half * ng[i]);
  out.depth = 0;
  out.steps = L;
  out.moved = false;
  if (valid) {
    const double H1 = pol.kinetic(p.data(), im) - (double)nlj;
    const double lar = H0 - H1;
    const double err = H1 - H0;
    out.error = err;
    out.divergent = !std::isfinite(err) || err > div_thr;
    out.accept_prob = !std::isfinite(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
    if (lar_out) *lar_out = lar;
    if (acc_stream != 0xffffffffu) {
      const double u = rand_uniform_d(chain, itc, acc_stream, 0u);
      out.moved = log_d(u) < std::fmin(0.0, lar);
    }
    out.energy = out.moved ? H1 : H0;
  } else {
    out.error = inf_d();
    out.divergent = true;
    out.accept_prob = 0.0;
    out.energy = H0;
    if (lar_out) *lar_out = -inf_d();
  }
  if (out.moved) {
    std::copy(nq.begin(), nq.end(), q);
    std::copy(ng.begin(), ng.end(), g);
    lj = nlj;
  }
}

// ------------------------------------------------------------------ adaptation
struct Dual {
  double mu, log_step, log_step_avg, hbar, it;
};
Dual dual_init(double step) {
  const double ls = log_d(step);
  return Dual{log_d(10.0 * step), ls, ls, 0.0, 0.0};
}
double dual_update(Dual& d, double accept_prob, double target) {
  const double gamma = 0.05, t0 = 10.0, kappa = 0.75;
  d.it = d.it + 1.0;
  const double it = d.it;
  const double eta = 1.0 / (it + t0);
  d.hbar = (1.0 - eta) * d.hbar + eta * (target - accept_prob);
  d.log_step = d.mu - sqrt_d(it) / gamma * d.hbar;
  const double eta_bar = pow_d(it, -kappa);
  d.log_step_avg = eta_bar * d.log_step + (1.0 - eta_bar) * d.log_step_avg;
  return exp_d(d.log_step);
}

struct Schedule {
  int initial_buffer = 0;
  std::vector<int> ends;
  int terminal_buffer = 0;
};
Schedule warmup_schedule(int n) {
  Schedule s;
  if (n <= 0) return s;
  if (n < 20) {
    const int ib = std::min(5, std::max(n - 1, 0));
    s.initial_buffer = ib;
    if (ib == n) return s;
    s.ends.push_back(n);
    return s;
  }
  const int ib = std::min(std::max(5, n / 10), n);
  const int tb = std::min(std::max(5, n / 10), std::max(n - ib, 0));
  const int slow = n - ib - tb;
  if (slow <= 0) { s.initial_buffer = n; return s; }
  s.initial_buffer = ib;
  s.terminal_buffer = tb;
  int next = std::min(25, slow), start = ib + 1, remaining = slow;
  while (remaining > 0) {
    const int size = remaining <= (3 * next) / 2 ? remaining : next;
    const int end = start + size - 1;
    s.ends.push_back(end);
    remaining -= size;
    start = end + 1;
    next *= 2;
  }
  return s;
}

// clipped weighted running variance (reference adaptation.jl:30-41,140-260)
struct RunVar {
  std::vector<double> mean, m2, clipped;
  int window_length = 20;
  long long count = 0;
  double wsum = 0, wsq = 0;
  void reset(int P, int wl) {
    mean.assign(P, 0.0); m2.assign(P, 0.0); clipped.assign(P, 0.0);
    window_length = wl; count = 0; wsum = 0; wsq = 0;
  }
  static double progress(long long count, int wl) {
    if (wl <= 4) return 1.0;
    const double r = (double)std::max(count - 4, 0ll) / (double)(wl - 4);
    return r < 1.0 ? r : 1.0;
  }
  double clip_scale() const {
    if (count <= 4) return 8.0;
    return 8.0 + (5.0 - 8.0) * progress(count, window_length);
  }
  void update(const double* x, double w) {
    if (w == 0.0) return;
    const int P = (int)mean.size();
    if (count < 4) {
      for (int i = 0; i < P; ++i) clipped[i] = x[i];
    } else {
      const double cs = clip_scale();
      const double dc = (double)std::max(count - 1, 1ll);
      for (int i = 0; i < P; ++i) {
        const double var = m2[i] / dc;
        const double bound = cs * sqrt_d(var > 1e-3 ? var : 1e-3);
        double delta = x[i] - mean[i];
        delta = delta > bound ? bound : (delta < -bound ? -bound : delta);
        clipped[i] = mean[i] + delta;
      }
    }
    count += 1;
    const double nw = wsum + w;
    const double frac = w / nw;
    for (int i = 0; i < P; ++i) {
      const double delta = clipped[i] - mean[i];
      mean[i] += frac * delta;
      const double delta2 = clipped[i] - mean[i];
      m2[i] += (w * delta) * delta2;
    }
    wsum = nw;
    wsq += w * w;
  }
  double eff() const { return wsq <= 0.0 ? 0.0 : (wsum * wsum) / wsq; }
  void inverse_mass(double reg, std::vector<double>& im) const {
    const int P = (int)mean.size();
    const double e = eff();
    const double denom = wsum - wsq / wsum;
    const double shrink = e / (e + 5.0);
    for (int i = 0; i < P; ++i) {
      if (e < 2.0 || denom <= 0.0) { im[i] = 1.0; continue; }
      const double v = shrink * (m2[i] / denom) + (1.0 - shrink);
      im[i] = v > reg ? v : reg;
    }
  }
};

double mass_weight(const RunVar& st, bool accepted, double ap, bool divergent) {
  if (divergent) return 0.0;
  if (accepted) return 1.0;
  if (!std::isfinite(ap)) return 0.0;
  const double prog = RunVar::progress(st.count, st.window_length);
  const double rej = ap < 0.0 ? 0.0 : (ap > 1.0 ? 1.0 : ap);
  return 1.0 + (rej - 1.0) * prog;
}

// Reasonable step-size search of one chain from step s: one leapfrog step with the momentum of
// stream S_SEARCH at (chain, itc), then doubling (or halving) until the log accept ratio first
// crosses log 0.5, at most 20 trials, clamped to [1e-8, 1e3] (adaptation.jl:516-571 per chain;
// the batched per-chain form :438-514 and the window-end re-search call it with each chain's
// current step as the anchor).
template <class Pol>
double reasonable_step(const Pol& pol, const typename Pol::T* im, const typename Pol::T* q,
                       const typename Pol::T* g, typename Pol::T lj, uint32_t c, uint32_t itc, double s) {
  using T = typename Pol::T;
  const int n = pol.n();
  const double log_target = -0.6931471805599453;
  auto trial = [&](double st) {
    std::vector<T> tq(q, q + n), tg(g, g + n);
    T tlj = lj;
    TransOut to;
    double lar;
    hmc_transition(pol, im, (T)st, c, itc, 1, 0.0, S_SEARCH, 0xffffffffu, tq.data(), tg.data(), tlj, to, &lar);
    return lar;
  };
  const double lar0 = trial(s);
  const double dir = lar0 > log_target ? 1.0 : -1.0;
  for (int t = 0; t < 20; ++t) {
    const double next = s * (dir > 0.0 ? 2.0 : 0.5);
    if (next < 1e-8 || next > 1e3) break;
    const double lar = trial(next);
    s = next;
    if ((dir > 0.0 && lar <= log_target) || (dir < 0.0 && lar >= log_target)) break;
  }
  return s < 1e-8 ? 1e-8 : (s > 1e3 ? 1e3 : s);
}

// persistent-slice drift ((u + 1 + delta) mod 2) - 1 with a floored mod, in [-1, 1)
// (api_batched_meads.jl:141)
double meads_slice_drift(double u, double delta) {
  const double x = u + 1.0 + delta;
  return (x - 2.0 * std::floor(x / 2.0)) - 1.0;
}

uint32_t run_seed_of(uint64_t seed) {
  return philox4x32((uint32_t)seed, (uint32_t)(seed >> 32), 0x5EEDu, 0u, 0u, 0u).w[0];
}

void draw_initial_column(const Model& m, int strategy, uint32_t chain, uint32_t key, double* col) {
  const int P = m.P;
  if (strategy == UT_INIT_UNIFORM) {
    for (int a = 0; a < P; ++a) col[a] = 4.0 * rand_uniform_d(chain, key, S_INIT, (uint32_t)a) - 2.0;
    return;
  }
  auto z = [&](int idx) { return (double)rand_normal_f(chain, key, S_INIT, (uint32_t)idx); };
  if (m.kind == KIND_LOGISTIC) {
    for (int a = 0; a < P; ++a) col[a] = m.loc + m.scale * z(a);
  } else if (m.kind == KIND_HIER) {
    col[0] = m.mu_loc + m.mu_scale * z(0);
    const double tau = m.tau_scale * std::fabs(z(1) / z(P + 1));
    col[1] = log_d(tau);
    for (int j = 0; j < m.J; ++j) col[2 + j] = z(2 + j);
  } else if (m.kind == KIND_PLAN) {
    plan_prior_draw(m.plan->desc, strategy, chain, key, S_INIT, col);
  } else {
    const double v = m.v_scale * z(0);
    col[0] = v;
    const double s = exp_d(v / 2.0);
    for (int i = 0; i < m.dim; ++i) col[1 + i] = s * z(1 + i);
  }
}

// ------------------------------------------------------------------ sampler driver
// Mirrors the device driver: per-chain step with dual averaging, pooled diagonal
// mass fed per iteration in chain order, window-end re-search, sampling.
template <class Pol>
int run_sampler(const Pol& pol, const ut_sampler_opts* o, int C, bool nuts, int threads,
                void* samples, ut_chain_out* out, const int* stop_after) {
  using T = typename Pol::T;
  const Model& m = *pol.m;
  const int P = m.P, n = pol.n(), S = o->num_samples, W = o->num_warmup;
  const uint32_t rs = run_seed_of(o->seed);
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<double> pos((size_t)P * C);
  if (o->initial_params) std::memcpy(pos.data(), o->initial_params, pos.size() * 8);
  else
    for (int c = 0; c < C; ++c) draw_initial_column(m, o->init_strategy, (uint32_t)c, rs, &pos[(size_t)c * P]);
  std::vector<T> q((size_t)n * C, T(0)), g((size_t)n * C, T(0)), lj(C);
  auto load = [&](int c) {
    for (int i = 0; i < n; ++i) {
      const int a = pol.api(i);
      q[(size_t)c * n + i] = a >= 0 ? (T)(float)pos[(size_t)c * P + a] : T(0);
    }
    if (std::is_same<T, double>::value)
      for (int i = 0; i < n; ++i) {
        const int a = pol.api(i);
        q[(size_t)c * n + i] = a >= 0 ? (T)pos[(size_t)c * P + a] : T(0);
      }
  };
  for (int attempt = 0;; ++attempt) {
    std::vector<int> bad;
    for (int c = 0; c < C; ++c) {
      load(c);
      bool ok;
      lj[c] = pol.eval(&q[(size_t)c * n], &g[(size_t)c * n], ok);
      if (!ok) bad.push_back(c);
    }
    if (bad.empty()) break;
    if (o->initial_params || attempt >= o->init_max_retries) return UT_E_NONFINITE_INIT;
    for (int c : bad)
      draw_initial_column(m, o->init_strategy, (uint32_t)c, rs + (uint32_t)(attempt + 1), &pos[(size_t)c * P]);
  }
  std::vector<T> im(n, T(0));
  std::vector<double> im_d(P, 1.0);
  auto set_im = [&]() {
    for (int i = 0; i < n; ++i) {
      const int a = pol.api(i);
      im[i] = a >= 0 ? (T)(float)im_d[a] : T(0);
      if (std::is_same<T, double>::value && a >= 0) im[i] = (T)im_d[a];
    }
  };
  set_im();
  std::vector<double> step(C, o->step_size);
  std::vector<Dual> dual(C);
  auto search = [&](uint32_t itc, bool do_search) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < C; ++c) {
      double s = step[c];
      if (do_search) s = reasonable_step(pol, im.data(), &q[(size_t)c * n], &g[(size_t)c * n], lj[c], (uint32_t)c, itc, s);
      step[c] = s;
      dual[c] = dual_init(s);
    }
  };
  const bool do_search = o->find_reasonable_step_size || (W > 0 && o->adapt_step_size);
  search(rs, do_search);

  const Schedule sch = warmup_schedule(W);
  RunVar rv;
  size_t win = 0;
  rv.reset(P, sch.ends.empty() ? 20 : sch.ends[0] - sch.initial_buffer);
  std::vector<TransOut> outs(C);
  std::vector<double> colbuf(P);
  const int last = stop_after ? std::min(*stop_after, W + S) : W + S;
  // phase wall times (the reference harness times warmup and sampling apart)
  using clk = std::chrono::steady_clock;
  const clk::time_point t_start = clk::now();
  clk::time_point t_warm = t_start;
  for (int it = 1; it <= last; ++it) {
    if (it == W + 1) t_warm = clk::now();
    const uint32_t itc = rs + (uint32_t)it;
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < C; ++c) {
      const T h = (T)(float)step[c];
      const T hh = std::is_same<T, double>::value ? (T)step[c] : h;
      T* qc = &q[(size_t)c * n];
      T* gc = &g[(size_t)c * n];
      std::vector<T> gk(gc, gc + n);
      const T ljk = lj[c];
      if (nuts) {
        nuts_transition(pol, im.data(), hh, (uint32_t)c, itc, o->max_tree_depth, o->divergence_threshold, qc, gc, lj[c], outs[c]);
        if (!outs[c].moved) { std::copy(gk.begin(), gk.end(), gc); lj[c] = ljk; }
      } else {
        hmc_transition(pol, im.data(), hh, (uint32_t)c, itc, o->num_leapfrog_steps, o->divergence_threshold,
                       S_MOM, S_ACCEPT, qc, gc, lj[c], outs[c], nullptr);
      }
      if (it <= W && o->adapt_step_size) {
        step[c] = dual_update(dual[c], outs[c].divergent ? 0.0 : outs[c].accept_prob, o->target_accept);
        if (it == W) step[c] = exp_d(dual[c].log_step_avg);
      }
    }
    if (it <= W) {
      // pooled mass update, chain order, weights from the state at iteration start
      const bool in_window = o->adapt_mass_matrix && win < sch.ends.size() && it > sch.initial_buffer;
      if (in_window) {
        std::vector<double> w(C);
        for (int c = 0; c < C; ++c) w[c] = mass_weight(rv, !nuts && outs[c].moved, outs[c].accept_prob, outs[c].divergent);
        for (int c = 0; c < C; ++c) {
          for (int i = 0; i < n; ++i) {
            const int a = pol.api(i);
            if (a >= 0) colbuf[a] = (double)q[(size_t)c * n + i];
          }
          rv.update(colbuf.data(), w[c]);
        }
        if (it == sch.ends[win]) {
          if (rv.eff() >= (double)o->mass_matrix_min_samples) {
            rv.inverse_mass(o->mass_matrix_regularization, im_d);
            set_im();
          }
          ++win;
          if (win < sch.ends.size()) rv.reset(P, sch.ends[win] - sch.ends[win - 1]);
          else rv.reset(P, 20);
          if (o->adapt_step_size && it < W) search(rs + (uint32_t)it, true);
        }
      }
    } else if (samples) {
      const int s = it - W - 1;
      for (int c = 0; c < C; ++c) {
        const size_t row = (size_t)c * S + s;
        for (int i = 0; i < n; ++i) {
          const int a = pol.api(i);
          if (a < 0) continue;
          if (std::is_same<T, double>::value) ((double*)samples)[row * P + a] = (double)q[(size_t)c * n + i];
          else ((float*)samples)[row * P + a] = (float)q[(size_t)c * n + i];
        }
        if (out) {
          if (out->logjoint) out->logjoint[row] = (double)lj[c];
          if (out->accept_stat) out->accept_stat[row] = outs[c].accept_prob;
          if (out->energy) out->energy[row] = outs[c].energy;
          if (out->energy_error) out->energy_error[row] = outs[c].error;
          if (out->accepted) out->accepted[row] = outs[c].moved ? 1 : 0;
          if (out->divergent) out->divergent[row] = outs[c].divergent ? 1 : 0;
          if (out->tree_depth) out->tree_depth[row] = outs[c].depth;
          if (out->integration_steps) out->integration_steps[row] = outs[c].steps;
        }
      }
    }
  }
  if (out) {
    if (out->step_size) for (int c = 0; c < C; ++c) out->step_size[c] = step[c];
    if (out->inverse_mass) for (int a = 0; a < P; ++a) out->inverse_mass[a] = im_d[a];
    const clk::time_point t_end = clk::now();
    if (last <= W) t_warm = t_end;
    out->warmup_seconds = std::chrono::duration<double>(t_warm - t_start).count();
    out->sampling_seconds = std::chrono::duration<double>(t_end - t_warm).count();
  }
  return UT_OK;
}

// Sampling-phase NUTS iterations from a given state, without adaptation; returns
// the number of leapfrog gradient evaluations.  Used to time the CPU baseline.
template <class Pol>
int64_t sample_from(const Pol& pol, int C, int threads, double* positions, const double* steps,
                    const double* im_api, uint32_t rs, int first_it, int iters, int max_depth) {
  using T = typename Pol::T;
  const Model& m = *pol.m;
  const int P = m.P, n = pol.n();
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<T> im(n, T(0));
  for (int i = 0; i < n; ++i) {
    const int a = pol.api(i);
    im[i] = a >= 0 ? (T)im_api[a] : T(0);
  }
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (int c = 0; c < C; ++c) {
    std::vector<T> q(n, T(0)), g(n, T(0)), gk(n);
    for (int i = 0; i < n; ++i) {
      const int a = pol.api(i);
      if (a >= 0) q[i] = (T)positions[(size_t)c * P + a];
    }
    bool ok;
    T lj = pol.eval(q.data(), g.data(), ok);
    for (int it = first_it; it < first_it + iters; ++it) {
      TransOut o;
      const T ljk = lj;
      gk = g;
      nuts_transition(pol, im.data(), (T)steps[c], (uint32_t)c, rs + (uint32_t)it, max_depth, 1000.0,
                      q.data(), g.data(), lj, o);
      if (!o.moved) { g = gk; lj = ljk; }
      total += o.steps;
    }
    for (int i = 0; i < n; ++i) {
      const int a = pol.api(i);
      if (a >= 0) positions[(size_t)c * P + a] = (double)q[i];
    }
  }
  return total;
}

// ------------------------------------------------------------------ ChEES
// ChEES-HMC restated from src/inference/api_batched_chees.jl: Halton jitter
// (:29-51), jittered step count with the integer snap (:53-70), scalar Adam ascent
// (:86-110), trajectory update (:157-275), harmonic-mean accept (:277-295), the host
// loop (:435-700) and the shared batched step-size search
// (src/inference/nuts/adaptation.jl:303-375).  Trajectories run in lock step in
// reference parameter order, as the device kernels do.
double chees_halton(long long index) {
  double r = 0.0, f = 0.5;
  for (long long k = index; k > 0; k >>= 1) {
    if (k & 1) r += f;
    f *= 0.5;
  }
  return r;
}

int chees_leapfrog_steps(double jitter, double T, double eps) {
  const double raw = jitter * T / eps;
  const double nearest = std::nearbyint(raw);
  double steps;
  if (std::fabs(raw - nearest) <= 1e-9 * std::max(1.0, std::fabs(raw))) steps = nearest;
  else steps = std::ceil(raw);
  if (steps < 1.0) steps = 1.0;
  return (int)std::min(steps, 1e9);
}

struct CheesState {
  double T, log_T_ma;
  double b1, b2, aeps, lr, m = 0.0, v = 0.0;
  long long t = 0, step = 1;
  double decay;
  int max_steps;
  double adam(double g) {
    t += 1;
    m = b1 * m + (1.0 - b1) * g;
    v = b2 * v + (1.0 - b2) * g * g;
    double p1 = 1.0, p2 = 1.0;
    for (long long i = 0; i < t; ++i) {
      p1 *= b1;
      p2 *= b2;
    }
    return lr * (m / (1.0 - p1)) / (sqrt_d(v / (1.0 - p2)) + aeps);
  }
  void update(double numerator, double denominator, double jitter, double eps) {
    const double grad = denominator > 0.0 ? jitter * T * numerator / denominator : 0.0;
    const double lt = log_d(T);
    const double sm = m, sv = v;
    const long long st = t;
    const double u = std::min(std::max(adam(grad), -0.35), 0.35);
    double ltn = lt + u;
    if (!std::isfinite(ltn)) {
      ltn = lt;
      m = sm;
      v = sv;
      t = st;
    }
    const double w = pow_d((double)step, -decay);
    log_T_ma = (1.0 - w) * log_T_ma + w * ltn;
    T = std::min(std::max(exp_d(log_T_ma), eps), (double)max_steps * eps);
    step += 1;
  }
};

// hmc = true: fixed-length HMC with shared adaptation (reference batched_hmc,
// per_chain_adaptation = false, api_batched_hmc.jl:184-330): L = num_leapfrog_steps,
// dual averaging to the mean acceptance, no trajectory adaptation.
template <class Pol>
int run_chees(const Pol& pol, const ut_chees_opts* co, int C, int threads, void* samples, ut_chain_out* out,
              double* trace, bool hmc) {
  using T = typename Pol::T;
  constexpr bool F32 = std::is_same<T, float>::value;
  const ut_sampler_opts* o = &co->base;
  const Model& m = *pol.m;
  const int P = m.P, S = o->num_samples, W = o->num_warmup;
  const uint32_t rs = run_seed_of(o->seed);
  if (threads > 0) omp_set_num_threads(threads);
  const size_t PC = (size_t)P * C;
  std::vector<double> pos(PC);
  if (o->initial_params) std::memcpy(pos.data(), o->initial_params, PC * 8);
  else
    for (int c = 0; c < C; ++c) draw_initial_column(m, o->init_strategy, (uint32_t)c, rs, &pos[(size_t)c * P]);
  std::vector<T> q(PC), g(PC), lj(C);
  for (int attempt = 0;; ++attempt) {
    std::vector<int> bad;
    for (int c = 0; c < C; ++c) {
      for (int a = 0; a < P; ++a) q[(size_t)c * P + a] = F32 ? (T)(float)pos[(size_t)c * P + a] : (T)pos[(size_t)c * P + a];
      lj[c] = pol.ls_eval(&q[(size_t)c * P], &g[(size_t)c * P]);
      bool ok = std::isfinite(lj[c]);
      for (int a = 0; a < P; ++a) ok = ok && std::isfinite(g[(size_t)c * P + a]);
      if (!ok) bad.push_back(c);
    }
    if (bad.empty()) break;
    if (o->initial_params || attempt >= o->init_max_retries) return UT_E_NONFINITE_INIT;
    for (int c : bad)
      draw_initial_column(m, o->init_strategy, (uint32_t)c, rs + (uint32_t)(attempt + 1), &pos[(size_t)c * P]);
  }
  std::vector<double> im_d(P, 1.0);
  std::vector<T> im(P, T(1));
  // trajectory buffers and per-chain records
  std::vector<T> wq(PC), wp(PC), wg(PC), wlj(C), init(PC);
  std::vector<double> h0(C), ap(C), energy(C), err(C);
  std::vector<char> valid(C), div(C), acc(C), pfin(C);

  auto begin = [&](uint32_t itc, uint32_t stream, T h, bool save) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
      T* qc = &wq[(size_t)c * P];
      T* pc = &wp[(size_t)c * P];
      T* gc = &wg[(size_t)c * P];
      bool fin = true;
      for (int a = 0; a < P; ++a) {
        qc[a] = q[(size_t)c * P + a];
        gc[a] = g[(size_t)c * P + a];
        pc[a] = draw_normal<T>((uint32_t)c, itc, stream, (uint32_t)a) / sqrt_t<T>(im[a]);
        fin = fin && std::isfinite(gc[a]);
        if (save) init[(size_t)c * P + a] = qc[a];
      }
      h0[c] = pol.ls_kinetic(pc, im.data()) - (double)lj[c];
      valid[c] = fin;
      if (fin) {
        const T half = T(0.5) * h;
        for (int a = 0; a < P; ++a) {
          pc[a] = pc[a] + half * gc[a];
          qc[a] = qc[a] + (h * im[a]) * pc[a];
        }
      }
    }
  };
  auto gradient = [&]() {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) wlj[c] = pol.ls_eval(&wq[(size_t)c * P], &wg[(size_t)c * P]);
  };
  auto inner = [&](T h) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
      if (!valid[c]) continue;
      T* qc = &wq[(size_t)c * P];
      T* pc = &wp[(size_t)c * P];
      const T* gc = &wg[(size_t)c * P];
      bool fin = true;
      for (int a = 0; a < P; ++a) fin = fin && std::isfinite(gc[a]);
      if (!fin) {
        valid[c] = 0;
        continue;
      }
      for (int a = 0; a < P; ++a) {
        pc[a] = pc[a] + (T(1) * h) * gc[a];
        qc[a] = qc[a] + (h * im[a]) * pc[a];
      }
    }
  };
  // mode 0: accept / reject; mode 1: acceptance probability only
  auto finish = [&](uint32_t itc, T h, int mode) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
      T* qc = &wq[(size_t)c * P];
      T* pc = &wp[(size_t)c * P];
      const T* gc = &wg[(size_t)c * P];
      bool fin = true;
      for (int a = 0; a < P; ++a) fin = fin && std::isfinite(gc[a]);
      const bool ok = valid[c] && fin && std::isfinite(wlj[c]);
      if (ok) {
        const T half = T(0.5) * h;
        for (int a = 0; a < P; ++a) pc[a] = -(pc[a] + half * gc[a]);
      }
      double a_p = 0.0, e = inf_d(), en = h0[c];
      bool d = true, moved = false;
      if (ok) {
        const double H1 = pol.ls_kinetic(pc, im.data()) - (double)wlj[c];
        const double lar = h0[c] - H1;
        e = H1 - h0[c];
        d = !std::isfinite(e) || e > o->divergence_threshold;
        a_p = !std::isfinite(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
        if (mode == 0) moved = log_d(rand_uniform_d((uint32_t)c, itc, S_ACCEPT, 0u)) < std::fmin(0.0, lar);
        en = moved ? H1 : h0[c];
      }
      bool pf = true;
      for (int a = 0; a < P; ++a) pf = pf && std::isfinite(qc[a]) && std::isfinite(pc[a]);
      ap[c] = a_p;
      err[c] = e;
      energy[c] = en;
      div[c] = d;
      acc[c] = moved;
      pfin[c] = pf;
      if (mode == 0 && moved) {
        std::copy(qc, qc + P, &q[(size_t)c * P]);
        std::copy(gc, gc + P, &g[(size_t)c * P]);
        lj[c] = wlj[c];
      }
    }
  };
  auto mean_accept = [&]() {
    double s = 0.0;
    for (int c = 0; c < C; ++c) s += div[c] ? 0.0 : ap[c];
    return s / (double)C;
  };
  auto harmonic = [&]() {
    double s = 0.0;
    long long n = 0;
    for (int c = 0; c < C; ++c) {
      if (div[c]) continue;
      s += 1.0 / ap[c];
      n += 1;
    }
    if (n == 0) return 0.0;
    const double hm = (double)n / s;
    return std::isfinite(hm) ? hm : 0.0;
  };
  auto cols_finite = [&](const std::vector<T>& v, int c) {
    for (int a = 0; a < P; ++a)
      if (!std::isfinite(v[(size_t)c * P + a])) return false;
    return true;
  };
  // ChEES numerator and denominator, with the iteration-start inverse mass
  const double vsign = co->forward_velocity ? -1.0 : 1.0;
  auto criterion = [&](const std::vector<double>& sigma, double* num, double* den) {
    std::vector<double> imean(P, 0.0), pmean(P, 0.0), isq(P);
    double wsum = 0.0;
    long long icount = 0;
    for (int a = 0; a < P; ++a) {
      double is = 0.0, ps = 0.0;
      long long ic = 0;
      double ws = 0.0;
      for (int c = 0; c < C; ++c) {
        if (cols_finite(init, c)) {
          is += (double)init[(size_t)c * P + a];
          ic += 1;
        }
        if (!div[c] && cols_finite(wq, c)) {
          ws += ap[c];
          ps += ap[c] * (double)wq[(size_t)c * P + a];
        }
      }
      imean[a] = ic > 0 ? is / (double)ic : 0.0;
      pmean[a] = ps * (1.0 / (ws + 1e-20));
      isq[a] = 1.0 / sqrt_d(sigma[a]);
      wsum = ws;
      icount = ic;
    }
    (void)wsum;
    (void)icount;
    double n = 0.0, d = 0.0;
    for (int c = 0; c < C; ++c) {
      if (div[c] || !pfin[c]) continue;
      double se = 0.0, ss = 0.0, ev = 0.0;
      for (int a = 0; a < P; ++a) {
        const double we = ((double)wq[(size_t)c * P + a] - pmean[a]) * isq[a];
        const double ws = ((double)init[(size_t)c * P + a] - imean[a]) * isq[a];
        const double wv = vsign * ((double)wp[(size_t)c * P + a] * sigma[a] * isq[a]);
        se += we * we;
        ss += ws * ws;
        ev += we * wv;
      }
      n += ap[c] * (se - ss) * ev;
      d += ap[c] + 1e-20;
    }
    *num = n;
    *den = d;
  };
  auto search = [&](double step, uint32_t itc) {
    double direction = 0.0;
    for (int trial = 0; trial <= 20; ++trial) {
      const T h = F32 ? (T)(float)step : (T)step;
      begin(itc, S_SEARCH, h, false);
      gradient();
      finish(itc, h, 1);
      const double ma = mean_accept();
      if (trial == 0) direction = ma > 0.5 ? 1.0 : -1.0;
      else if ((direction > 0 && ma <= 0.5) || (direction < 0 && ma >= 0.5)) break;
      const double next = step * (direction > 0 ? 2.0 : 0.5);
      if (next < 1e-8 || next > 1e3) break;
      step = next;
    }
    return std::min(std::max(step, 1e-8), 1e3);
  };
  auto set_im = [&]() {
    for (int a = 0; a < P; ++a) im[a] = F32 ? (T)(float)im_d[a] : (T)im_d[a];
  };

  double eps = o->step_size;
  if (o->find_reasonable_step_size || (W > 0 && o->adapt_step_size)) eps = search(eps, rs);
  Dual dual = dual_init(eps);
  const double T0 = (double)o->num_leapfrog_steps * eps;
  CheesState cs{T0, log_d(T0), co->trajectory_adam_beta1, co->trajectory_adam_beta2, co->trajectory_adam_epsilon,
                co->trajectory_learning_rate};
  cs.decay = co->trajectory_decay_rate;
  cs.max_steps = co->max_leapfrog_steps;
  const bool adapt_traj = !hmc && W > 0 && co->adapt_trajectory_length;
  const Schedule sch = warmup_schedule(W);
  RunVar rv;
  size_t win = 0;
  rv.reset(P, sch.ends.empty() ? 20 : sch.ends[0] - sch.initial_buffer);
  for (int it = 1; it <= W + S; ++it) {
    const double jitter = chees_halton(it) * co->jitter_amount + (1.0 - co->jitter_amount);
    const int L = hmc ? o->num_leapfrog_steps : chees_leapfrog_steps(jitter, cs.T, eps);
    const uint32_t itc = rs + (uint32_t)it;
    const T h = F32 ? (T)(float)eps : (T)eps;
    const bool warm = it <= W;
    begin(itc, S_MOM, h, warm && adapt_traj);
    for (int l = 1; l < L; ++l) {
      gradient();
      inner(h);
    }
    gradient();
    finish(itc, h, 0);
    if (warm) {
      double num = 0.0, den = 0.0;
      if (adapt_traj) criterion(im_d, &num, &den);
      const double stat = hmc ? mean_accept() : harmonic();
      if (o->adapt_step_size) eps = dual_update(dual, stat, o->target_accept);
      if (o->adapt_mass_matrix && win < sch.ends.size() && it > sch.initial_buffer) {
        std::vector<double> w(C), col(P);
        for (int c = 0; c < C; ++c) w[c] = mass_weight(rv, acc[c], ap[c], div[c]);
        for (int c = 0; c < C; ++c) {
          for (int a = 0; a < P; ++a) col[a] = (double)q[(size_t)c * P + a];
          rv.update(col.data(), w[c]);
        }
        if (it == sch.ends[win]) {
          if (rv.eff() >= (double)o->mass_matrix_min_samples) {
            rv.inverse_mass(o->mass_matrix_regularization, im_d);
            set_im();
          }
          ++win;
          if (win < sch.ends.size()) rv.reset(P, sch.ends[win] - sch.ends[win - 1]);
          else rv.reset(P, 20);
          if (o->adapt_step_size && it < W) {
            eps = search(eps, rs + (uint32_t)it);
            dual = dual_init(eps);
          }
        }
      }
      if (adapt_traj) cs.update(num, den, jitter, eps);
      if (trace && !hmc) trace[it - 1] = cs.T;
      if (it == W && o->adapt_step_size) eps = exp_d(dual.log_step_avg);
    } else {
      const int s = it - W - 1;
      for (int c = 0; c < C; ++c) {
        const size_t row = (size_t)c * S + s;
        if (samples)
          for (int a = 0; a < P; ++a) {
            if (F32) ((float*)samples)[row * P + a] = (float)q[(size_t)c * P + a];
            else ((double*)samples)[row * P + a] = (double)q[(size_t)c * P + a];
          }
        if (out) {
          if (out->logjoint) out->logjoint[row] = (double)lj[c];
          if (out->accept_stat) out->accept_stat[row] = ap[c];
          if (out->energy) out->energy[row] = energy[c];
          if (out->energy_error) out->energy_error[row] = err[c];
          if (out->accepted) out->accepted[row] = acc[c] ? 1 : 0;
          if (out->divergent) out->divergent[row] = div[c] ? 1 : 0;
          if (out->tree_depth) out->tree_depth[row] = 0;
          if (out->integration_steps) out->integration_steps[row] = L;
        }
      }
    }
  }
  if (out) {
    if (out->step_size)
      for (int c = 0; c < C; ++c) out->step_size[c] = eps;
    if (out->inverse_mass)
      for (int a = 0; a < P; ++a) out->inverse_mass[a] = im_d[a];
  }
  return UT_OK;
}

// ------------------------------------------------------------------ MEADS
// Restates src/inference/api_batched_meads.jl (constants :32-34, _meads_max_eig
// :48-77, _meads_fold_parameters! :89-136, slice drift :141, iteration :272-470)
// with the device's reduction orders (member norms and off-diagonal squared Gram
// entries per member, then members in order) and Philox draws.
enum { S_MEADS_REFRESH = 7, S_MEADS_INIT = 8 };

double meads_max_eig(const std::vector<double>& x, int n, int P) {
  if (n < 2) return 0.0;
  std::vector<double> norms(n), rows(n);
  for (int a = 0; a < n; ++a) {
    const double* xa = &x[(size_t)a * P];
    double nn = 0.0;
    for (int p = 0; p < P; ++p) nn += xa[p] * xa[p];
    norms[a] = nn;
    double r = 0.0;
    for (int b = 0; b < n; ++b) {
      if (b == a) continue;
      const double* xb = &x[(size_t)b * P];
      double g = 0.0;
      for (int p = 0; p < P; ++p) g += xa[p] * xb[p];
      r += g * g;
    }
    rows[a] = r;
  }
  double trace = 0.0, off = 0.0;
  for (int a = 0; a < n; ++a) trace += norms[a];
  for (int a = 0; a < n; ++a) off += rows[a];
  const double lb = trace / (double)n;
  if (!(lb > 0.0)) return 0.0;
  return (off / ((double)n * (double)(n - 1))) / lb;
}

template <class Pol>
int run_meads(const Pol& pol, const ut_meads_opts* mo, int C, int threads, void* samples, ut_chain_out* out,
              double* trace) {
  using T = typename Pol::T;
  constexpr bool F32 = std::is_same<T, float>::value;
  const Model& m = *pol.m;
  const int P = m.P, S = mo->num_samples, W = mo->num_warmup, K = mo->num_folds;
  const uint32_t rs = run_seed_of(mo->seed);
  if (threads > 0) omp_set_num_threads(threads);
  const size_t PC = (size_t)P * C;
  std::vector<double> pos(PC);
  if (mo->initial_params) std::memcpy(pos.data(), mo->initial_params, PC * 8);
  else
    for (int c = 0; c < C; ++c) draw_initial_column(m, mo->init_strategy, (uint32_t)c, rs, &pos[(size_t)c * P]);
  std::vector<T> q(PC), g(PC), lj(C);
  for (int attempt = 0;; ++attempt) {
    std::vector<int> bad;
    for (int c = 0; c < C; ++c) {
      for (int a = 0; a < P; ++a) q[(size_t)c * P + a] = F32 ? (T)(float)pos[(size_t)c * P + a] : (T)pos[(size_t)c * P + a];
      lj[c] = pol.ls_eval(&q[(size_t)c * P], &g[(size_t)c * P]);
      bool ok = std::isfinite(lj[c]);
      for (int a = 0; a < P; ++a) ok = ok && std::isfinite(g[(size_t)c * P + a]);
      if (!ok) bad.push_back(c);
    }
    if (bad.empty()) break;
    if (mo->initial_params || attempt >= mo->init_max_retries) return UT_E_NONFINITE_INIT;
    for (int c : bad)
      draw_initial_column(m, mo->init_strategy, (uint32_t)c, rs + (uint32_t)(attempt + 1), &pos[(size_t)c * P]);
  }
  std::vector<int> start(K + 1, 0), fold_of(C);
  const int base = C / K, rem = C % K;
  for (int k = 0; k < K; ++k) start[k + 1] = start[k] + base + (k < rem ? 1 : 0);
  for (int k = 0; k < K; ++k)
    for (int c = start[k]; c < start[k + 1]; ++c) fold_of[c] = k;
  std::vector<T> mom(PC), refreshed(PC), wq(PC), wp(PC), wg(PC), wlj(C);
  std::vector<double> slice(C), h0(C), ap(C), err(C);
  std::vector<char> valid(C), div(C), acc(C);
  for (int c = 0; c < C; ++c) {
    for (int a = 0; a < P; ++a) mom[(size_t)c * P + a] = (T)rand_normal_f((uint32_t)c, rs, S_MEADS_INIT, (uint32_t)a);
    slice[c] = 2.0 * rand_uniform_d((uint32_t)c, rs, S_MEADS_INIT, (uint32_t)P) - 1.0;
  }
  std::vector<double> sigma((size_t)K * P), fold((size_t)4 * K);
  std::vector<T> sigma_t((size_t)K * P), im_t((size_t)K * P);
  double final_step = 0.0;
  for (int it = 1; it <= W + S; ++it) {
    const uint32_t itc = rs + (uint32_t)it;
    const int skip = (it - 1) % K;
    for (int k = 0; k < K; ++k) {
      const int nb = (k - 1 + K) % K;
      const int c0 = start[nb], n = start[nb + 1] - c0;
      std::vector<double> wpos((size_t)n * P), wgrd((size_t)n * P);
      double mx = 0.0;
      for (int p = 0; p < P; ++p) {
        double mean = 0.0;
        for (int j = 0; j < n; ++j) mean += (double)q[(size_t)(c0 + j) * P + p];
        mean /= (double)n;
        double var = 0.0;
        for (int j = 0; j < n; ++j) {
          const double d = (double)q[(size_t)(c0 + j) * P + p] - mean;
          var += d * d;
        }
        var /= (double)n;
        double sd = sqrt_d(var);
        sd = sd > 1e-10 ? sd : 1e-10;
        sigma[(size_t)k * P + p] = sd;
        sigma_t[(size_t)k * P + p] = F32 ? (T)(float)sd : (T)sd;
        im_t[(size_t)k * P + p] = F32 ? (T)(float)(sd * sd) : (T)(sd * sd);
        const double inv = 1.0 / sd;
        for (int j = 0; j < n; ++j) {
          wpos[(size_t)j * P + p] = ((double)q[(size_t)(c0 + j) * P + p] - mean) * inv;
          wgrd[(size_t)j * P + p] = (double)g[(size_t)(c0 + j) * P + p] * sd;
        }
        mx = std::fmax(mx, sd);
      }
      const double eg = std::max(meads_max_eig(wgrd, n, P), 1e-12);
      const double ep = std::max(meads_max_eig(wpos, n, P), 1e-12);
      double step = 0.5 / sqrt_d(eg);
      step = step < 1.0 ? step : 1.0;
      const double floor_g = 1.0 / ((double)it * step);
      const double inv_p = 1.0 / sqrt_d(ep);
      const double gamma = floor_g > inv_p ? floor_g : inv_p;
      double alpha = 1.0 - exp_d(-2.0 * step * gamma);
      alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
      fold[4 * k + 0] = step;
      fold[4 * k + 1] = alpha;
      fold[4 * k + 2] = alpha / 2.0;
      fold[4 * k + 3] = mx;
    }
    double ssum = 0.0, asum = 0.0, smax = 0.0;
    for (int k = 0; k < K; ++k) {
      ssum += fold[4 * k];
      asum += fold[4 * k + 1];
      smax = std::max(smax, fold[4 * k + 3]);
    }
    final_step = ssum / K;
    if (trace) {
      trace[3 * (size_t)(it - 1) + 0] = ssum / K;
      trace[3 * (size_t)(it - 1) + 1] = asum / K;
      trace[3 * (size_t)(it - 1) + 2] = smax;
    }
    // refresh + half kick + drift
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
      const int f = fold_of[c];
      const double alpha = fold[4 * f + 1];
      const T keep = F32 ? (T)(float)sqrt_d(1.0 - alpha > 0.0 ? 1.0 - alpha : 0.0) : (T)sqrt_d(std::max(1.0 - alpha, 0.0));
      const T refresh = F32 ? (T)(float)sqrt_d(alpha > 0.0 ? alpha : 0.0) : (T)sqrt_d(std::max(alpha, 0.0));
      const T eps = F32 ? (T)(float)fold[4 * f] : (T)fold[4 * f];
      T* pc = &wp[(size_t)c * P];
      T* qc = &wq[(size_t)c * P];
      const T* im = &im_t[(size_t)f * P];
      bool fin = true;
      for (int a = 0; a < P; ++a) {
        const T xi = draw_normal<T>((uint32_t)c, itc, S_MEADS_REFRESH, (uint32_t)a);
        const T mt = keep * mom[(size_t)c * P + a] + refresh * xi;
        refreshed[(size_t)c * P + a] = mt;
        pc[a] = mt / sigma_t[(size_t)f * P + a];
        qc[a] = q[(size_t)c * P + a];
        fin = fin && std::isfinite(g[(size_t)c * P + a]);
      }
      h0[c] = pol.ls_kinetic(pc, im) - (double)lj[c];
      valid[c] = fin;
      if (fin) {
        const T half = T(0.5) * eps;
        for (int a = 0; a < P; ++a) {
          pc[a] = pc[a] + half * g[(size_t)c * P + a];
          qc[a] = qc[a] + (eps * im[a]) * pc[a];
        }
      }
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) wlj[c] = pol.ls_eval(&wq[(size_t)c * P], &wg[(size_t)c * P]);
    // final half kick, flip, persistent-slice accept
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
      const int f = fold_of[c];
      const T eps = F32 ? (T)(float)fold[4 * f] : (T)fold[4 * f];
      const double delta = fold[4 * f + 2];
      T* pc = &wp[(size_t)c * P];
      const T* gc = &wg[(size_t)c * P];
      const T* im = &im_t[(size_t)f * P];
      bool fin = true;
      for (int a = 0; a < P; ++a) fin = fin && std::isfinite(gc[a]);
      const bool ok = valid[c] && fin && std::isfinite(wlj[c]);
      if (ok) {
        const T half = T(0.5) * eps;
        for (int a = 0; a < P; ++a) pc[a] = -(pc[a] + half * gc[a]);
      }
      double a_p = 0.0, e = inf_d(), lar = -inf_d();
      bool d = true;
      if (ok) {
        const double H1 = pol.ls_kinetic(pc, im) - (double)wlj[c];
        lar = h0[c] - H1;
        e = H1 - h0[c];
        d = !std::isfinite(e) || e > mo->divergence_threshold;
        a_p = !std::isfinite(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
      }
      ap[c] = a_p;
      err[c] = e;
      div[c] = d;
      acc[c] = false;
      if (f == skip) continue;
      const double drifted = meads_slice_drift(slice[c], delta);
      const bool moved = ok && log_d(std::fabs(drifted)) <= lar;
      acc[c] = moved;
      for (int a = 0; a < P; ++a) {
        if (moved) {
          q[(size_t)c * P + a] = wq[(size_t)c * P + a];
          g[(size_t)c * P + a] = gc[a];
          mom[(size_t)c * P + a] = -pc[a] * sigma_t[(size_t)f * P + a];
        } else {
          mom[(size_t)c * P + a] = -refreshed[(size_t)c * P + a];
        }
      }
      if (moved) {
        lj[c] = wlj[c];
        slice[c] = drifted * exp_d(e);
      } else {
        slice[c] = drifted;
      }
    }
    if (it <= W) continue;
    const int s = it - W - 1;
    for (int c = 0; c < C; ++c) {
      const size_t row = (size_t)c * S + s;
      const bool skipped = fold_of[c] == skip;
      if (samples)
        for (int a = 0; a < P; ++a) {
          if (F32) ((float*)samples)[row * P + a] = (float)q[(size_t)c * P + a];
          else ((double*)samples)[row * P + a] = (double)q[(size_t)c * P + a];
        }
      if (!out) continue;
      double kin;
      if (F32) {
        double lane[WAVE];
        for (int l = 0; l < WAVE; ++l) {
          double acc2 = 0.0;
          for (int a = l; a < P; a += WAVE) acc2 = acc2 + (double)mom[(size_t)c * P + a] * (double)mom[(size_t)c * P + a];
          lane[l] = acc2;
        }
        kin = bfly(lane);
      } else {
        kin = 0.0;
        for (int a = 0; a < P; ++a) kin += (double)mom[(size_t)c * P + a] * (double)mom[(size_t)c * P + a];
      }
      if (out->logjoint) out->logjoint[row] = (double)lj[c];
      if (out->energy) out->energy[row] = -(double)lj[c] + 0.5 * kin;
      if (out->accept_stat) out->accept_stat[row] = skipped ? 1.0 : ap[c];
      if (out->accepted) out->accepted[row] = skipped ? 1 : (acc[c] ? 1 : 0);
      if (out->divergent) out->divergent[row] = skipped ? 0 : (div[c] ? 1 : 0);
      if (out->energy_error) out->energy_error[row] = skipped ? 0.0 : err[c];
      if (out->tree_depth) out->tree_depth[row] = 0;
      if (out->integration_steps) out->integration_steps[row] = skipped ? 0 : 1;
    }
  }
  if (out) {
    if (out->step_size)
      for (int c = 0; c < C; ++c) out->step_size[c] = final_step;
    if (out->inverse_mass)
      for (int a = 0; a < P; ++a) out->inverse_mass[a] = sigma[a] * sigma[a];
  }
  return UT_OK;
}

// ------------------------------------------------------------------ flow
// Affine-coupling flow, binary64 restatement of the device flow (structure of the
// reference FlowGuide, src/inference/vi/guides.jl:282-369, with MLP conditioners):
// x0 = mu + e^ls z; layer l: passive i with (i + l) even; [s_raw, t] =
// W2 tanh(W1 x_p + b1) + b2; s = c tanh(s_raw / c); x_a <- x_a e^s + t.  With
// `bf16` the GEMM operands (passive inputs, weights, hidden activations and the
// backward's output and hidden gradients) are rounded to bfloat16 as the device's
// matrix-core operands are; accumulation stays binary64.
struct FlowSpec {
  int P, L, H;
  double c;
  bool bf16;
  const float* w;  // flat weights, device layout
  struct Layer {
    std::vector<int> p, a;
    int kp, op;
    long long oW1, ob1, oW2, ob2;
  };
  std::vector<Layer> layers;
  FlowSpec(int P_, int L_, int H_, double c_, bool bf, const float* w_) : P(P_), L(L_), H(H_), c(c_), bf16(bf), w(w_) {
    long long off = 2LL * P;
    for (int l = 0; l < L; ++l) {
      Layer ly;
      for (int i = 0; i < P; ++i) ((i + l) % 2 == 0 ? ly.p : ly.a).push_back(i);
      ly.kp = ((int)ly.p.size() + 127) / 128 * 128;
      ly.op = (2 * (int)ly.a.size() + 127) / 128 * 128;
      ly.oW1 = off;
      off += (long long)H * ly.kp;
      ly.ob1 = off;
      off += H;
      ly.oW2 = off;
      off += (long long)ly.op * H;
      ly.ob2 = off;
      off += ly.op;
      layers.push_back(ly);
    }
  }
  double r(double v) const {  // operand rounding
    if (!bf16) return v;
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float o;
    std::memcpy(&o, &u, 4);
    return (double)o;
  }
  double W1(const Layer& ly, int j, int k) const { return r(w[ly.oW1 + (long long)j * ly.kp + k]); }
  double W2(const Layer& ly, int o, int j) const { return r(w[ly.oW2 + (long long)o * H + j]); }
  // conditioner: h (H), s, t (na) from the passive coordinates of x
  void cond(const Layer& ly, const double* x, std::vector<double>& h, std::vector<double>& s,
            std::vector<double>& t) const {
    const int np = (int)ly.p.size(), na = (int)ly.a.size();
    h.assign(H, 0.0);
    for (int j = 0; j < H; ++j) {
      double u = 0.0;
      for (int k = 0; k < np; ++k) u += r(x[ly.p[k]]) * W1(ly, j, k);
      h[j] = std::tanh(u + (double)w[ly.ob1 + j]);
    }
    s.assign(na, 0.0);
    t.assign(na, 0.0);
    for (int i = 0; i < na; ++i) {
      double os = 0.0, ot = 0.0;
      for (int j = 0; j < H; ++j) {
        os += r(h[j]) * W2(ly, i, j);
        ot += r(h[j]) * W2(ly, na + i, j);
      }
      s[i] = c * std::tanh((os + (double)w[ly.ob2 + i]) / c);
      t[i] = ot + (double)w[ly.ob2 + na + i];
    }
  }
  // backward through the conditioner: given dS (d/ds) and dT (d/dt), add d/dx_p into g;
  // optionally accumulate weight gradients
  void cond_bwd(const Layer& ly, const double* x, const std::vector<double>& h, const std::vector<double>& s,
                const std::vector<double>& dS, const std::vector<double>& dT, double* g, double* gw) const {
    const int np = (int)ly.p.size(), na = (int)ly.a.size();
    std::vector<double> dO(2 * na), dU(H);
    for (int i = 0; i < na; ++i) {
      const double q = s[i] / c;
      dO[i] = dS[i] * (1.0 - q * q);
      dO[na + i] = dT[i];
    }
    for (int j = 0; j < H; ++j) {
      double gh = 0.0;
      for (int o = 0; o < 2 * na; ++o) gh += r(dO[o]) * W2(ly, o, j);
      dU[j] = gh * (1.0 - h[j] * h[j]);
    }
    for (int k = 0; k < np; ++k) {
      double gx = 0.0;
      for (int j = 0; j < H; ++j) gx += r(dU[j]) * W1(ly, j, k);
      g[ly.p[k]] += gx;
    }
    if (!gw) return;
    for (int o = 0; o < 2 * na; ++o) {
      for (int j = 0; j < H; ++j) gw[ly.oW2 + (long long)o * H + j] += r(dO[o]) * r(h[j]);
      gw[ly.ob2 + o] += dO[o];
    }
    for (int j = 0; j < H; ++j) {
      for (int k = 0; k < np; ++k) gw[ly.oW1 + (long long)j * ly.kp + k] += r(dU[j]) * r(x[ly.p[k]]);
      gw[ly.ob1 + j] += dU[j];
    }
  }
  // q = f(z); returns log|det df/dz|.  With g_q, also overwrites g_q by
  // d/dz [g_q . f(z) + logdet]; with gw as well, accumulates the gradient of that
  // quantity (g_q held fixed) over the weights (the reparameterised ELBO gradient
  // of the reference's flow ADVI, src/inference/vi.jl:807-826).
  double forward(const double* z, double* q, double* g_q, double* gw = nullptr) const {
    std::vector<double> x(P);
    double ld = 0.0;
    for (int i = 0; i < P; ++i) {
      x[i] = (double)w[i] + std::exp((double)w[P + i]) * z[i];
      ld += (double)w[P + i];
    }
    std::vector<std::vector<double>> X(L), Hs(L), S(L), XS(L);
    for (int l = 0; l < L; ++l) {
      const Layer& ly = layers[l];
      X[l] = x;
      std::vector<double> t;
      cond(ly, x.data(), Hs[l], S[l], t);
      XS[l].resize(ly.a.size());
      for (size_t i = 0; i < ly.a.size(); ++i) {
        const double xs = x[ly.a[i]] * std::exp(S[l][i]);
        XS[l][i] = xs;
        x[ly.a[i]] = xs + t[i];
        ld += S[l][i];
      }
    }
    for (int i = 0; i < P; ++i) q[i] = x[i];
    if (!g_q) return ld;
    std::vector<double> g(g_q, g_q + P);
    for (int l = L - 1; l >= 0; --l) {
      const Layer& ly = layers[l];
      const int na = (int)ly.a.size();
      std::vector<double> dS(na), dT(na);
      for (int i = 0; i < na; ++i) {
        const double ga = g[ly.a[i]];
        dS[i] = ga * XS[l][i] + 1.0;
        dT[i] = ga;
        g[ly.a[i]] = ga * std::exp(S[l][i]);
      }
      cond_bwd(ly, X[l].data(), Hs[l], S[l], dS, dT, g.data(), gw);
    }
    for (int i = 0; i < P; ++i) {
      const double e = std::exp((double)w[P + i]);
      if (gw) {
        gw[i] += g[i];
        gw[P + i] += g[i] * e * z[i] + 1.0;
      }
      g_q[i] = g[i] * e;
    }
    return ld;
  }
  // z = f^-1(q); returns log|det df^-1/dq|.  With gw, accumulates the gradient of
  // 0.5|z|^2 - logdet_inv over the weights and returns that loss in *loss.
  double inverse(const double* q, double* z, double* gw, double* loss) const {
    std::vector<double> x(q, q + P);
    double ld = 0.0;
    std::vector<std::vector<double>> X(L), Hs(L), S(L), XO(L);
    for (int l = L - 1; l >= 0; --l) {
      const Layer& ly = layers[l];
      std::vector<double> t;
      X[l] = x;  // conditioner input (passive coordinates are unchanged by the layer)
      cond(ly, x.data(), Hs[l], S[l], t);
      XO[l].resize(ly.a.size());
      for (size_t i = 0; i < ly.a.size(); ++i) {
        const double xo = (x[ly.a[i]] - t[i]) * std::exp(-S[l][i]);
        x[ly.a[i]] = xo;
        XO[l][i] = xo;
        ld -= S[l][i];
      }
    }
    double zz = 0.0;
    for (int i = 0; i < P; ++i) {
      z[i] = (x[i] - (double)w[i]) * std::exp(-(double)w[P + i]);
      ld -= (double)w[P + i];
      zz += z[i] * z[i];
    }
    if (loss) *loss = 0.5 * zz - ld + 0.5 * P * 1.8378770664093453;
    if (!gw) return ld;
    std::vector<double> g(P);
    for (int i = 0; i < P; ++i) {
      const double e = std::exp(-(double)w[P + i]);
      g[i] = z[i] * e;
      gw[i] += -z[i] * e;
      gw[P + i] += 1.0 - z[i] * z[i];
    }
    for (int l = 0; l < L; ++l) {
      const Layer& ly = layers[l];
      const int na = (int)ly.a.size();
      std::vector<double> dS(na), dT(na);
      for (int i = 0; i < na; ++i) {
        const double go = g[ly.a[i]];
        const double e = std::exp(-S[l][i]);
        dT[i] = -go * e;
        dS[i] = -go * XO[l][i] + 1.0;
        g[ly.a[i]] = go * e;
      }
      cond_bwd(ly, X[l].data(), Hs[l], S[l], dS, dT, g.data(), gw);
    }
    return ld;
  }
  long long nparams() const { return layers.empty() ? 2LL * P : layers.back().ob2 + layers.back().op; }
};

void finish_twin_layout(Model& m) {
  m.K = m.kind == KIND_HIER ? 2 : slots_for(m.P);
  if (m.kind == KIND_HIER) {
    m.max_n = 0;
    m.tnobs.assign(WAVE, 0);
    // [6][64]: c_j, hinv_j, inv_j, then the group's sufficient statistics ybar_j,
    // S_j = sum_i (y_ij - ybar_j)^2 (binary64, observation order, two passes), n_j
    m.tlane.assign(6 * WAVE, 0.0);
    for (int j = 0; j < m.J; ++j) m.max_n = std::max(m.max_n, m.offsets[j + 1] - m.offsets[j]);
    for (int j = 0; j < m.J; ++j) {
      const int nj = m.offsets[j + 1] - m.offsets[j];
      m.tnobs[j] = nj;
      double sy = 0.0;
      for (int i = 0; i < nj; ++i) sy = sy + (double)m.hy[m.offsets[j] + i];
      const double ybar = nj > 0 ? sy / (double)nj : 0.0;
      double ss = 0.0;
      for (int i = 0; i < nj; ++i) {
        const double dv = (double)m.hy[m.offsets[j] + i] - ybar;
        ss = ss + dv * dv;
      }
      const double s = m.sigma[j];
      m.tlane[j] = -(double)nj * log_d(s) - (double)nj * HALF_LOG_2PI;
      m.tlane[WAVE + j] = 1.0 / (2.0 * s * s);
      m.tlane[2 * WAVE + j] = 1.0 / (s * s);
      m.tlane[3 * WAVE + j] = ybar;
      m.tlane[4 * WAVE + j] = ss;
      m.tlane[5 * WAVE + j] = (double)nj;
    }
    m.total_const = (double)m.J * (-HALF_LOG_2PI) + (-log_d(m.mu_scale) - HALF_LOG_2PI) +
                    (LOG_2 - LOG_PI - log_d(m.tau_scale));
  } else if (m.kind == KIND_LOGISTIC) {
    m.total_const = (double)m.D * (-log_d(m.scale) - HALF_LOG_2PI);
  } else if (m.kind == KIND_PLAN) {
    m.total_const = 0.0;
  } else {
    m.total_const = -(double)m.dim * HALF_LOG_2PI + (-log_d(m.v_scale) - HALF_LOG_2PI);
  }
}

}  // namespace

// ====================================================================== C API
extern "C" {

void orc_philox4x32(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]) {
  philox_block b = philox4x32(key[0], key[1], ctr[0], ctr[1], ctr[2], ctr[3]);
  for (int i = 0; i < 4; ++i) out[i] = b.w[i];
}

void orc_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t first, int64_t n,
                     uint32_t* blocks, float* uf, float* nf, double* ud) {
  for (int64_t i = 0; i < n; ++i) {
    const philox_block b = draw_block(chain, iteration, stream, first + (uint32_t)i);
    if (blocks) for (int w = 0; w < 4; ++w) blocks[4 * i + w] = b.w[w];
    if (uf) uf[i] = u01_f(b.w[0]);
    if (nf) nf[i] = box_muller_f(b);
    if (ud) ud[i] = u01_d(b.w[0], b.w[1]);
  }
}

// elementwise shared math, for accuracy tests: fn 0 exp 1 log 2 log1p 3 cos 4 log1p_exp 5 logistic
void orc_math_f32(int fn, const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: y[i] = exp_f(x[i]); break;
      case 1: y[i] = log_f(x[i]); break;
      case 2: y[i] = log1p_f(x[i]); break;
      case 3: y[i] = cos_f(x[i]); break;
      case 4: y[i] = log1p_exp_f(x[i]); break;
      case 5: y[i] = logistic_f(x[i]); break;
    }
  }
}
void orc_math_f64(int fn, const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: y[i] = exp_d(x[i]); break;
      case 1: y[i] = log_d(x[i]); break;
      case 2: y[i] = log1p_d(x[i]); break;
      case 4: y[i] = log1p_exp_d(x[i]); break;
      case 5: y[i] = logistic_d(x[i]); break;
      case 6: y[i] = pow_d(x[i], -0.75); break;
      default: y[i] = 0.0;
    }
  }
}

void* orc_model_logistic(const float* x, const float* y, int64_t n, int32_t d, double loc, double scale) {
  auto* m = new Model();
  m->kind = KIND_LOGISTIC;
  m->x.assign(x, x + n * d);
  m->y.assign(y, y + n);
  m->n = n;
  m->D = d;
  m->P = d;
  m->loc = loc;
  m->scale = scale;
  finish_twin_layout(*m);
  return m;
}

// Wavefronts per chain of the logistic model's device layout (1, 2, 4, 8).
int32_t orc_model_set_tiles(void* h, int32_t tiles) {
  auto* m = (Model*)h;
  if (m->kind != KIND_LOGISTIC || (tiles != 1 && tiles != 2 && tiles != 4 && tiles != 8)) return -1;
  m->tiles = tiles;
  return 0;
}

void* orc_model_hier(const float* y, const int32_t* offsets, const double* sigma, int32_t J,
                     double mu_loc, double mu_scale, double tau_scale) {
  auto* m = new Model();
  m->kind = KIND_HIER;
  m->J = J;
  m->offsets.assign(offsets, offsets + J + 1);
  m->hy.assign(y, y + offsets[J]);
  m->sigma.assign(sigma, sigma + J);
  m->mu_loc = mu_loc;
  m->mu_scale = mu_scale;
  m->tau_scale = tau_scale;
  m->P = J + 2;
  finish_twin_layout(*m);
  return m;
}

// a general plan; NULL (with the reason in orc_plan_error) when the audit fails
static std::string g_plan_error;
void* orc_model_plan(const ut_plan_desc* desc) {
  auto data = plan_oracle::copy_plan(*desc);
  const int rc = plan_audit(&data->desc, data->audit);
  if (rc != UT_OK) {
    g_plan_error = data->audit.error;
    return nullptr;
  }
  auto* m = new Model();
  m->kind = KIND_PLAN;
  m->P = desc->num_params;
  m->plan = data;
  finish_twin_layout(*m);
  return m;
}
const char* orc_plan_error(void) { return g_plan_error.c_str(); }
// the prior draw of a plan column (the samplers' UT_INIT_PRIOR start)
void orc_plan_prior(void* h, int32_t strategy, uint32_t chain, uint32_t key, double* col) {
  const Model& m = *(Model*)h;
  plan_prior_draw(m.plan->desc, strategy, chain, key, S_INIT, col);
}

void* orc_model_funnel(int32_t dim, double v_scale) {
  auto* m = new Model();
  m->kind = KIND_FUNNEL;
  m->dim = dim;
  m->v_scale = v_scale;
  m->P = dim + 1;
  finish_twin_layout(*m);
  return m;
}

void* orc_model_gp(const double* inputs, const int32_t* counts, int32_t n, int32_t d, double noise, double ls_loc,
                   double ls_scale, double amp_scale) {
  auto* m = new Model();
  m->kind = KIND_GP;
  m->gn = n;
  m->gd = d;
  m->gx.assign(inputs, inputs + (size_t)n * d);
  m->gy.resize(n);
  for (int i = 0; i < n; ++i) m->gy[i] = (double)counts[i];
  m->noise2 = noise * noise + 1e-8;
  m->ls_loc = ls_loc;
  m->ls_scale = ls_scale;
  m->amp_scale = amp_scale;
  m->P = d + 1 + n;
  m->K = 1;  // the twin layout is the reference order for this model
  return m;
}

// One binary64 subtree of 2^depth leaves from (q, p) with the checkpoint U-turn test of
// nuts_transition. Writes the leaves integrated before the subtree stopped and whether it turned,
// and, independently, the full 2^depth-leaf leapfrog trajectory (traj_q, traj_p: 2^depth x P,
// leaf order) so that a caller can rebuild the recursive tree over the same points.
void orc_nuts_subtree(void* h, const double* q_in, const double* p_in, const double* im, double step,
                      int32_t direction, int32_t depth, int32_t* steps, int32_t* turning, int32_t* divergent,
                      double* traj_q, double* traj_p) {
  const Model& m = *(const Model*)h;
  RefF64 pol{&m};
  const int n = m.P;
  const bool fwd = direction > 0;
  const double ss = fwd ? step : -step;
  std::vector<double> q(q_in, q_in + n), p(p_in, p_in + n), g(n);
  bool ok;
  double lj = pol.eval(q.data(), g.data(), ok);
  std::vector<double> ck((size_t)(depth + 2) * 2 * n);
  int sub_int = 0;
  bool sub_div = false, sub_turn = false;
  std::vector<double> cq = q, cpm = p, cg = g;
  double cur = lj;
  dyadic_subtree(pol, im, ss, fwd, 1 << depth, cq.data(), cpm.data(), cg.data(), cur, ck.data(), sub_int,
                 sub_div, sub_turn, [](double) { return true; });
  *steps = sub_int;
  *turning = sub_turn ? 1 : 0;
  *divergent = sub_div ? 1 : 0;
  cq = q; cpm = p; cg = g;
  const double half = ss / 2.0;
  for (int l = 0; l < (1 << depth); ++l) {
    for (int i = 0; i < n; ++i) {
      cpm[i] = cpm[i] + half * cg[i];
      cq[i] = cq[i] + ss * (im[i] * cpm[i]);
    }
    pol.eval(cq.data(), cg.data(), ok);
    for (int i = 0; i < n; ++i) cpm[i] = cpm[i] + half * cg[i];
    std::memcpy(traj_q + (size_t)l * n, cq.data(), sizeof(double) * n);
    std::memcpy(traj_p + (size_t)l * n, cpm.data(), sizeof(double) * n);
  }
}

// binary64 per-chain reasonable step-size search (the samplers' initial and window-end search):
// params C x P, inverse_mass P, anchors C, Philox iteration coordinate itc; writes C steps.
void orc_step_search(void* h, int32_t C, const double* params, const double* inverse_mass, const double* anchors,
                     uint32_t itc, double* steps) {
  const Model& m = *(const Model*)h;
  RefF64 pol{&m};
  const int P = m.P;
  for (int c = 0; c < C; ++c) {
    std::vector<double> q(params + (size_t)c * P, params + (size_t)(c + 1) * P), g(P);
    bool ok;
    const double lj = pol.eval(q.data(), g.data(), ok);
    steps[c] = reasonable_step(pol, inverse_mass, q.data(), g.data(), lj, (uint32_t)c, itc, anchors[c]);
  }
}

// binary64 GP log density and gradient with the log-determinant of K (C chains)
void orc_gp_eval(void* h, const double* params, int32_t C, double* lj, double* grad, double* logdet) {
  const Model& m = *(Model*)h;
#pragma omp parallel for
  for (int c = 0; c < C; ++c) lj[c] = gp_eval(m, params + (size_t)c * m.P, grad + (size_t)c * m.P, logdet + c);
}

// OpenMP team size of the entry points without a threads argument (n <= 0 keeps it)
void orc_set_threads(int32_t n) {
  if (n > 0) omp_set_num_threads(n);
}

void orc_model_free(void* h) { delete (Model*)h; }
int32_t orc_model_num_params(void* h) { return ((Model*)h)->P; }

// binary64 reference-formula logjoint/gradient, params P x C (api order)
void orc_logjoint_grad_f64(void* h, const double* params, int32_t C, double* lj, double* grad) {
  const Model& m = *(Model*)h;
#pragma omp parallel for
  for (int c = 0; c < C; ++c) {
    bool ok;
    lj[c] = ref_eval(m, params + (size_t)c * m.P, grad + (size_t)c * m.P, ok);
  }
}

// kernel-exact logjoint/gradient in binary32 (twin=1) or binary64 (twin=2, the
// precision=Float64 density kernels), params P x C (api order)
void orc_logjoint_grad_twin(void* h, int32_t precision, const void* params, int32_t C, void* lj, void* grad) {
  const Model& m = *(Model*)h;
  const int n = m.K * WAVE;
  for (int c = 0; c < C; ++c) {
    if (precision == UT_FLOAT32) {
      std::vector<float> q(n, 0.0f), g(n, 0.0f);
      for (int i = 0; i < n; ++i) {
        const int a = m.api(i % WAVE, i / WAVE);
        if (a >= 0) q[i] = ((const float*)params)[(size_t)c * m.P + a];
      }
      bool ok;
      ((float*)lj)[c] = twin_eval<float>(m, q.data(), g.data(), ok);
      for (int i = 0; i < n; ++i) {
        const int a = m.api(i % WAVE, i / WAVE);
        if (a >= 0) ((float*)grad)[(size_t)c * m.P + a] = g[i];
      }
    } else {
      std::vector<double> q(n, 0.0), g(n, 0.0);
      for (int i = 0; i < n; ++i) {
        const int a = m.api(i % WAVE, i / WAVE);
        if (a >= 0) q[i] = ((const double*)params)[(size_t)c * m.P + a];
      }
      bool ok;
      ((double*)lj)[c] = twin_eval<double>(m, q.data(), g.data(), ok);
      for (int i = 0; i < n; ++i) {
        const int a = m.api(i % WAVE, i / WAVE);
        if (a >= 0) ((double*)grad)[(size_t)c * m.P + a] = g[i];
      }
    }
  }
}

// Full sampler run.  mode 0: TwinF32 (kernel-exact, samples are float);
// mode 1: RefF64 (samples are double).  kind 0 NUTS, 1 HMC.  stop_after < 0 runs
// every iteration; otherwise stops after that many iterations (timing samples).
int orc_sampler_run(void* h, int32_t kind, int32_t mode, const ut_sampler_opts* o, int32_t C,
                    int32_t threads, void* samples, ut_chain_out* out, int32_t stop_after) {
  const Model* m = (const Model*)h;
  const int* sa = stop_after >= 0 ? &stop_after : nullptr;
  if (mode == 0) {
    TwinF32 pol{m};
    return run_sampler(pol, o, C, kind == 0, threads, samples, out, sa);
  }
  RefF64 pol{m};
  return run_sampler(pol, o, C, kind == 0, threads, samples, out, sa);
}

// ChEES run.  mode 0: TwinF32 (kernel-exact, samples float); mode 1: RefF64.
int orc_chees_run(void* h, int32_t mode, const ut_chees_opts* co, int32_t C, int32_t threads, void* samples,
                  ut_chain_out* out, double* trace, int32_t hmc) {
  const Model* m = (const Model*)h;
  if (mode == 0) return run_chees(TwinF32{m}, co, C, threads, samples, out, trace, hmc != 0);
  return run_chees(RefF64{m}, co, C, threads, samples, out, trace, hmc != 0);
}

// MEADS run.  mode 0: TwinF32 (kernel-exact); mode 1: RefF64.
int orc_meads_run(void* h, int32_t mode, const ut_meads_opts* mo, int32_t C, int32_t threads, void* samples,
                  ut_chain_out* out, double* trace) {
  const Model* m = (const Model*)h;
  if (mode == 0) return run_meads(TwinF32{m}, mo, C, threads, samples, out, trace);
  return run_meads(RefF64{m}, mo, C, threads, samples, out, trace);
}
// the trace-ratio eigenvalue estimate on an n x P member-major ensemble
double orc_meads_slice_drift(double u, double delta) { return meads_slice_drift(u, delta); }
double orc_meads_max_eig(const double* x, int32_t n, int32_t P) {
  std::vector<double> v(x, x + (size_t)n * P);
  return meads_max_eig(v, n, P);
}

// ChEES scalar pieces for known-answer tests: Halton values, step counts and a
// sequence of trajectory updates from given (numerator, denominator, jitter, eps)
void orc_chees_halton(int32_t n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = chees_halton(i + 1);
}
int32_t orc_chees_steps(double jitter, double T, double eps) { return chees_leapfrog_steps(jitter, T, eps); }
void orc_chees_trajectory(double T0, double lr, double b1, double b2, double aeps, double decay, int32_t max_steps,
                          const double* num, const double* den, const double* jitter, const double* eps, int32_t n,
                          double* T_out) {
  CheesState cs{T0, log_d(T0), b1, b2, aeps, lr};
  cs.decay = decay;
  cs.max_steps = max_steps;
  for (int i = 0; i < n; ++i) {
    cs.update(num[i], den[i], jitter[i], eps[i]);
    T_out[i] = cs.T;
  }
}


// ---- flow (binary64; bf16 = 1 rounds the matrix-core operands as the device does)
void orc_flow_forward(int32_t P, int32_t L, int32_t H, double clamp, int32_t bf16, const float* weights,
                      const double* z, int64_t n, double* q, double* logdet, double* g) {
  FlowSpec fs(P, L, H, clamp, bf16 != 0, weights);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r)
    logdet[r] = fs.forward(z + r * P, q + r * P, g ? g + r * P : nullptr);
}
void orc_flow_inverse(int32_t P, int32_t L, int32_t H, double clamp, int32_t bf16, const float* weights,
                      const double* q, int64_t n, double* z, double* logdet_inv) {
  FlowSpec fs(P, L, H, clamp, bf16 != 0, weights);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) logdet_inv[r] = fs.inverse(q + r * P, z + r * P, nullptr, nullptr);
}
// mean loss and mean weight gradient over n rows
double orc_flow_loss_grad(int32_t P, int32_t L, int32_t H, double clamp, int32_t bf16, const float* weights,
                          const double* q, int64_t n, double* grad) {
  FlowSpec fs(P, L, H, clamp, bf16 != 0, weights);
  const long long np = fs.nparams();
  std::vector<double> total((size_t)np, 0.0);
  double loss_total = 0.0;
#pragma omp parallel
  {
    std::vector<double> gw((size_t)np, 0.0), z(P);
    double lt = 0.0;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
      double loss;
      fs.inverse(q + r * P, z.data(), gw.data(), &loss);
      lt += loss;
    }
#pragma omp critical
    {
      for (long long i = 0; i < np; ++i) total[i] += gw[i];
      loss_total += lt;
    }
  }
  for (long long i = 0; i < np; ++i) grad[i] = total[i] / (double)n;
  return loss_total / (double)n;
}
// Mean over the valid rows (valid[r] != 0, or all rows when valid is NULL) of the
// weight gradient of g_q . f(z) + log|det df/dz| with g_q held fixed; with
// row_weights, instead the sum over rows of row_weights[r] times that gradient (the
// importance-weighted bound's gradient, src/inference/vi/guides.jl:373-381).
void orc_flow_elbo_grad(int32_t P, int32_t L, int32_t H, double clamp, int32_t bf16, const float* weights,
                        const double* z, const double* g_q, const int32_t* valid, const double* row_weights,
                        int64_t n, double* grad) {
  FlowSpec fs(P, L, H, clamp, bf16 != 0, weights);
  const long long np = fs.nparams();
  std::vector<double> total((size_t)np, 0.0);
  int64_t nv = 0;
#pragma omp parallel
  {
    std::vector<double> gw((size_t)np, 0.0), gr((size_t)np), q(P), g(P);
    int64_t cnt = 0;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
      if (valid && !valid[r]) continue;
      std::copy(g_q + r * P, g_q + (r + 1) * P, g.begin());
      if (row_weights) {
        std::fill(gr.begin(), gr.end(), 0.0);
        fs.forward(z + r * P, q.data(), g.data(), gr.data());
        for (long long i = 0; i < np; ++i) gw[i] += row_weights[r] * gr[i];
      } else {
        fs.forward(z + r * P, q.data(), g.data(), gw.data());
      }
      ++cnt;
    }
#pragma omp critical
    {
      for (long long i = 0; i < np; ++i) total[i] += gw[i];
      nv += cnt;
    }
  }
  for (long long i = 0; i < np; ++i) grad[i] = row_weights ? total[i] : (nv ? total[i] / (double)nv : 0.0);
}
// ---- full-rank Gaussian guide (binary64): q = mu + e^ls z + F z with F the strict
// lower triangle of w[2P:] (P x P row-major), src/inference/vi/guides.jl:96-117;
// log|det dq/dz| = sum ls.
void orc_fullrank_forward(int32_t P, const float* w, const double* z, int64_t n, double* q, double* logdet) {
  const float* F = w + 2LL * P;
  double ld = 0.0;
  for (int i = 0; i < P; ++i) ld += (double)w[P + i];
  for (int64_t r = 0; r < n; ++r) {
    const double* zr = z + r * P;
    for (int i = 0; i < P; ++i) {
      double v = (double)w[i] + std::exp((double)w[P + i]) * zr[i];
      for (int j = 0; j < i; ++j) v += (double)F[(long long)i * P + j] * zr[j];
      q[r * P + i] = v;
    }
    logdet[r] = ld;
  }
}
// z = L^-1 (q - mu) by forward substitution; logdet_inv = -sum ls
void orc_fullrank_inverse(int32_t P, const float* w, const double* q, int64_t n, double* z, double* logdet_inv) {
  const float* F = w + 2LL * P;
  double ld = 0.0;
  for (int i = 0; i < P; ++i) ld -= (double)w[P + i];
  for (int64_t r = 0; r < n; ++r) {
    double* zr = z + r * P;
    for (int i = 0; i < P; ++i) {
      double v = q[r * P + i] - (double)w[i];
      for (int j = 0; j < i; ++j) v -= (double)F[(long long)i * P + j] * zr[j];
      zr[i] = v / std::exp((double)w[P + i]);
    }
    logdet_inv[r] = ld;
  }
}
// g_z = L^T g_q (the log-determinant does not depend on z)
void orc_fullrank_vjp(int32_t P, const float* w, const double* g_q, int64_t n, double* g_z) {
  const float* F = w + 2LL * P;
  for (int64_t r = 0; r < n; ++r) {
    const double* g = g_q + r * P;
    for (int j = 0; j < P; ++j) {
      double v = std::exp((double)w[P + j]) * g[j];
      for (int i = j + 1; i < P; ++i) v += (double)F[(long long)i * P + j] * g[i];
      g_z[r * P + j] = v;
    }
  }
}
// The weight gradient of g_q . q(z) + sum ls with g_q held fixed, averaged over the
// valid rows, or summed with row_weights (as orc_flow_elbo_grad):
// d/dmu = g, d/dls_i = g_i e^ls_i z_i + 1, d/dF[i, j] = g_i z_j for j < i, 0 on and
// above the diagonal (src/inference/vi/guides.jl:164-199, 459-489).
void orc_fullrank_elbo_grad(int32_t P, const float* w, const double* z, const double* g_q, const int32_t* valid,
                            const double* row_weights, int64_t n, double* grad) {
  const long long np = 2LL * P + (long long)P * P;
  std::vector<double> total((size_t)np, 0.0);
  double nv = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    if (valid && !valid[r]) continue;
    const double wr = row_weights ? row_weights[r] : 1.0;
    const double* g = g_q + r * P;
    const double* zr = z + r * P;
    for (int i = 0; i < P; ++i) {
      total[i] += wr * g[i];
      total[P + i] += wr * (g[i] * std::exp((double)w[P + i]) * zr[i] + 1.0);
      for (int j = 0; j < i; ++j) total[2LL * P + (long long)i * P + j] += wr * g[i] * zr[j];
    }
    nv += 1.0;
  }
  for (long long i = 0; i < np; ++i) grad[i] = row_weights ? total[i] : (nv > 0 ? total[i] / nv : 0.0);
}

int64_t orc_flow_num_weights(int32_t P, int32_t L, int32_t H) {
  FlowSpec fs(P, L, H, 1.0, false, nullptr);
  return fs.nparams();
}

int64_t orc_nuts_sample_from(void* h, int32_t mode, int32_t C, int32_t threads, double* positions,
                             const double* steps, const double* inverse_mass, int32_t first_iteration,
                             int32_t iters, uint64_t seed, int32_t max_depth) {
  const Model* m = (const Model*)h;
  const uint32_t rs = run_seed_of(seed);
  if (mode == 0) return sample_from(TwinF32{m}, C, threads, positions, steps, inverse_mass, rs, first_iteration, iters, max_depth);
  return sample_from(RefF64{m}, C, threads, positions, steps, inverse_mass, rs, first_iteration, iters, max_depth);
}

// warmup schedule: writes initial buffer, terminal buffer and up to cap window ends
int32_t orc_warmup_schedule(int32_t n, int32_t* initial_buffer, int32_t* terminal_buffer,
                            int32_t* ends, int32_t cap) {
  Schedule s = warmup_schedule(n);
  *initial_buffer = s.initial_buffer;
  *terminal_buffer = s.terminal_buffer;
  for (int i = 0; i < (int)s.ends.size() && i < cap; ++i) ends[i] = s.ends[i];
  return (int32_t)s.ends.size();
}

// dual-averaging sequence for KATs: runs `n` updates from `step` with the given
// accept statistics, writing the step after each update and the final averaged step
void orc_dual_averaging(double step, double target, const double* accepts, int32_t n, double* steps,
                        double* final_step) {
  Dual d = dual_init(step);
  for (int i = 0; i < n; ++i) steps[i] = dual_update(d, accepts[i], target);
  *final_step = exp_d(d.log_step_avg);
}

// pooled running variance over a (iterations x C x P) history with per-entry
// accept probabilities and flags (bit0 divergent, bit1 accepted), as the mass
// adaptation of one window consumes it; writes the inverse mass and returns the
// effective count.
double orc_running_variance(const double* hist, const double* acc, const uint8_t* flags, int32_t iters,
                            int32_t C, int32_t P, int32_t window_length, int32_t use_accepted,
                            double reg, double* inverse_mass) {
  RunVar rv;
  rv.reset(P, window_length);
  for (int it = 0; it < iters; ++it) {
    std::vector<double> w(C);
    for (int c = 0; c < C; ++c) {
      const uint8_t f = flags[(size_t)it * C + c];
      w[c] = mass_weight(rv, use_accepted && (f & 2), acc[(size_t)it * C + c], f & 1);
    }
    for (int c = 0; c < C; ++c) rv.update(&hist[((size_t)it * C + c) * P], w[c]);
  }
  std::vector<double> im(P);
  rv.inverse_mass(reg, im);
  for (int i = 0; i < P; ++i) inverse_mass[i] = im[i];
  return rv.eff();
}

}  // extern "C"
