// ut_chain.h -- the wavefront-per-chain (or workgroup-per-chain) kernels of the
// batched density and the persistent samplers, as templates over the device
// models: model construction from a DevModelDesc and dispatch, the NUTS and HMC
// transitions, k_density, k_nuts, k_hmc and k_step_search.  The launchers that
// instantiate them live in ut_chain.hip (density, HMC, step search) and ut_nuts.hip
// (NUTS), so the instantiations compile in parallel translation units.
//
// One 64-lane wavefront (one workgroup) per chain; the logistic model can also split
// one chain's rows over the W wavefronts of its workgroup, each of which then runs the
// same transition on the same values.  The whole per-chain state of a sampler
// iteration -- position, momentum, gradient, the NUTS frontiers and proposal -- lives
// in registers; the NUTS dyadic checkpoint stack lives in LDS.  Kernels run a whole
// range of sampler iterations per launch, so within a range the host is not involved
// at all.  Working precision is binary32 for positions, momenta, gradients and the
// log density; Hamiltonians, energy errors, acceptance probabilities, multinomial
// log-weights and adaptation state are binary64.
#pragma once
#include <algorithm>
#include <utility>

#include "ut_kernels.h"
#include "ut_models.h"

namespace ut {

// ------------------------------------------------------------------ model dispatch
template <int K, int DB = K * WAVE>
static LogisticModel<K, DB> make_logistic(const DevModelDesc& d) {
  LogisticModel<K, DB> m;
  m.xT_f = d.xT_f; m.xT_d = d.xT_d; m.y_f = d.yv_f; m.y_d = d.yv_d;
  m.n = d.n; m.D = d.D; m.loc = d.loc; m.scale = d.scale; m.total_const = d.total_const;
  return m;
}
template <int K, int DB, int W>
static TiledLogisticModel<K, DB, W> make_logistic_tiled(const DevModelDesc& d) {
  TiledLogisticModel<K, DB, W> m;
  m.xT_f = d.xT_f; m.xT_d = d.xT_d; m.y_f = d.yv_f; m.y_d = d.yv_d;
  m.n = d.n; m.D = d.D; m.loc = d.loc; m.scale = d.scale; m.total_const = d.total_const;
  return m;
}
template <int K>
static FunnelModel<K> make_funnel(const DevModelDesc& d) {
  FunnelModel<K> m;
  m.dim = d.dim; m.v_scale = d.v_scale; m.total_const = d.total_const;
  return m;
}
static HierModel make_hier(const DevModelDesc& d) {
  HierModel m;
  m.lane_f = d.lane_f; m.lane_d = d.lane_d;
  m.J = d.J; m.mu_loc = d.mu_loc; m.mu_scale = d.mu_scale;
  m.tau_scale = d.tau_scale; m.total_const = d.total_const;
  m.inv_mu_scale = d.inv_mu_scale; m.inv_tau_scale = d.inv_tau_scale;
  return m;
}

template <int K>
static GenericModel<K> make_generic(const DevModelDesc& d) {
  GenericModel<K> m;
  m.pv = d.plan;
  m.total_const = 0.0;
  return m;
}

// dynamic LDS of a launch: `base` bytes of the kernel's own, then the model's
// region; kernels whose dynamic LDS exceeds 64 KB are allowed up to 160 KB once
template <class T, class M>
static size_t model_lds(M& m, size_t base, const void* kernel) {
  const size_t mb = m.template lds_bytes<T>();
  if (mb == 0) return base;
  m.set_lds_offset(base);
  const size_t total = (base + 15) / 16 * 16 + mb;
  if (total > 65536) (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  return total;
}

template <class F>
static hipError_t dispatch(const DevModelDesc& d, F&& f) {
  switch (d.kind) {
    case MODEL_HIER:
      return f(make_hier(d));
    case MODEL_LOGISTIC:
      // up to 32 covariates: register arrays of 32 and the next row prefetched
      if (d.K == 1 && d.D <= 32) return f(make_logistic<1, 32>(d));
      if (d.K == 1) return f(make_logistic<1>(d));
      if (d.K == 2) return f(make_logistic<2>(d));
      if (d.K == 4) return f(make_logistic<4>(d));
      return hipErrorInvalidValue;
    case MODEL_FUNNEL:
      if (d.K == 1) return f(make_funnel<1>(d));
      if (d.K == 2) return f(make_funnel<2>(d));
      if (d.K == 4) return f(make_funnel<4>(d));
      return hipErrorInvalidValue;
    case MODEL_PLAN:
      if (d.K == 1) return f(make_generic<1>(d));
      if (d.K == 2) return f(make_generic<2>(d));
      if (d.K == 4) return f(make_generic<4>(d));
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
}

// Workgroup-per-chain instances (d.waves wavefronts per chain): the logistic model.
template <int W, class F>
static hipError_t dispatch_tiled_w(const DevModelDesc& d, F&& f) {
  if (d.K == 1 && d.D <= 32) return f(make_logistic_tiled<1, 32, W>(d));
  if (d.K == 1) return f(make_logistic_tiled<1, 64, W>(d));
  if (d.K == 2) return f(make_logistic_tiled<2, 128, W>(d));
  if (d.K == 4) return f(make_logistic_tiled<4, 256, W>(d));
  return hipErrorInvalidValue;
}
template <class F>
static hipError_t dispatch_tiled(const DevModelDesc& d, F&& f) {
  if (d.kind != MODEL_LOGISTIC) return hipErrorInvalidValue;
  switch (d.waves) {
    case 2: return dispatch_tiled_w<2>(d, f);
    case 4: return dispatch_tiled_w<4>(d, f);
    case 8: return dispatch_tiled_w<8>(d, f);
    default: return hipErrorInvalidValue;
  }
}
// The persistent samplers and the density kernel take either.
template <class F>
static hipError_t dispatch_chain(const DevModelDesc& d, F&& f) {
  return d.waves > 1 ? dispatch_tiled(d, f) : dispatch(d, f);
}

// A model bound to its per-lane data (prepared once per launch).
template <class M, class T>
struct Bound {
  static constexpr int K = M::K;
  const M& m;
  const typename M::template Lane<T>& ld;
  __device__ int api(int lane, int slot) const { return m.api(lane, slot); }
  __device__ T eval(const T (&q)[K], T (&g)[K], bool& ok) const { return m.eval(ld, q, g, ok); }
};

// ------------------------------------------------------------------ wave helpers
template <int K>
__device__ inline float lane_dot(const float (&a)[K], const float (&im)[K], const float (&b)[K]) {
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = acc + (a[k] * im[k]) * b[k];
  return acc;
}

// 0.5 * sum p^2 m^-1 in binary64 (kinetic energy, reference hmc_support.jl:497-511)
template <int K>
__device__ inline double kinetic(const float (&p)[K], const float (&im)[K]) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = acc + ((double)p[k] * (double)p[k]) * (double)im[k];
  return wave_sum(acc) * 0.5;
}

template <int K>
__device__ inline bool any_differs(const float (&a)[K], const float (&b)[K]) {
  bool d = false;
#pragma unroll
  for (int k = 0; k < K; ++k) d = d || (a[k] != b[k]);
  return wave_any(d);
}

// Whole-trajectory U-turn between frontiers (reference persistent_nuts.jl:109-119).
template <int K>
__device__ inline bool is_turning_lr(const float (&Lq)[K], const float (&Lp)[K],
                                     const float (&Rq)[K], const float (&Rp)[K],
                                     const float (&im)[K]) {
  float la = 0.0f, ra = 0.0f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float delta = Rq[k] - Lq[k];
    la = la + (delta * im[k]) * Lp[k];
    ra = ra + (delta * im[k]) * Rp[k];
  }
  const float ld = wave_sum(la);
  const float rd = wave_sum(ra);
  return ld <= 0.0f || rd <= 0.0f;
}

// Checkpoint stack in LDS: ck[((slot * 2 + which) * K + k) * 64 + lane].
template <int K>
__device__ inline bool dyadic_turning(const float* ck, const float (&q)[K], const float (&p)[K],
                                      int leaf_index, bool forward, const float (&im)[K]) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int nk = __builtin_ctz(~(unsigned)leaf_index);  // trailing ones
  for (int k = 1; k <= nk; ++k) {
    const int block_start = leaf_index - (1 << k) + 1;
    const int slot = __builtin_popcount((unsigned)block_start);
    float la = 0.0f, ra = 0.0f;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const float cp = ck[((slot * 2 + 0) * K + s) * WAVE + lane];
      const float cm = ck[((slot * 2 + 1) * K + s) * WAVE + lane];
      if (forward) {
        const float delta = q[s] - cp;
        la = la + (delta * im[s]) * cm;
        ra = ra + (delta * im[s]) * p[s];
      } else {
        const float delta = cp - q[s];
        la = la + (delta * im[s]) * p[s];
        ra = ra + (delta * im[s]) * cm;
      }
    }
    const float ld = wave_sum(la);
    const float rd = wave_sum(ra);
    if (ld <= 0.0f || rd <= 0.0f) return true;
  }
  return false;
}

template <int K>
__device__ inline void copyv(float (&dst)[K], const float (&src)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) dst[k] = src[k];
}

struct TransitionOut {
  double accept_prob;
  double energy;
  double error;
  int depth;
  int steps;
  bool divergent;
  bool moved;
};

// Dual-averaging step-size update (reference nuts/adaptation.jl:16-24).
// dual = {mu, log_step, log_step_avg, hbar, iteration}.
__device__ inline double dual_update(double (&dual)[5], double accept_prob, double target) {
  const double gamma = 0.05, t0 = 10.0, kappa = 0.75;
  dual[4] = dual[4] + 1.0;
  const double it = dual[4];
  const double eta = 1.0 / (it + t0);
  dual[3] = (1.0 - eta) * dual[3] + eta * (target - accept_prob);
  dual[1] = dual[0] - sqrt_d(it) / gamma * dual[3];
  const double eta_bar = pow_d(it, -kappa);
  dual[2] = eta_bar * dual[1] + (1.0 - eta_bar) * dual[2];
  return exp_d(dual[1]);
}

// ------------------------------------------------------------------ NUTS transition
// Canonical multinomial NUTS for one chain, transcribing the reference persistent
// kernel _persist_nuts_chain! (src/device/persistent_nuts.jl:226-492) with
// binary64 energies.  On return (q, g, lj) hold the selected proposal when `moved`.
template <class M>
__device__ void nuts_transition(const M& m, const float (&im)[M::K], float h, uint32_t chain,
                                uint32_t itc, int max_depth, double div_thr, float (&q)[M::K],
                                float (&g)[M::K], float& lj, float* ck, TransitionOut& out) {
  constexpr int K = M::K;
  const int lane = threadIdx.x & (WAVE - 1);
  float p[K], q0[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = m.api(lane, k);
    if (a >= 0) {
      const float z = rand_normal_f(chain, itc, STREAM_MOMENTUM, (uint32_t)a);
      p[k] = z / sqrt_f(im[k]);
    } else {
      p[k] = 0.0f;
    }
  }
  const double H0 = kinetic(p, im) - (double)lj;
  // Lane-parallel draws for this iteration: lane r holds the direction uniform and
  // the log merge uniform of doubling round r; leaf uniforms come in blocks of 64
  // (lane l: leaf 64 b + l).  Each value is the same function of its Philox
  // coordinate as a per-use draw; only the evaluation is batched across lanes.
  const float dir_u = rand_uniform_f(chain, itc, STREAM_DIRECTION, (uint32_t)lane);
  const double merge_lu = log_d(rand_uniform_d(chain, itc, STREAM_MERGE, (uint32_t)lane));
  double leaf_lu = 0.0;
  int leaf_block = -1;
  float Lq[K], Lp[K], Lg[K], Rq[K], Rp[K], Rg[K], sq[K], sg[K], pq[K], pg[K];
  copyv(q0, q);
  copyv(Lq, q); copyv(Lp, p); copyv(Lg, g);
  copyv(Rq, q); copyv(Rp, p); copyv(Rg, g);
  copyv(pq, q); copyv(pg, g);
  copyv(sq, q); copyv(sg, g);
  float ljL = lj, ljR = lj, ljP = lj, ljS = lj;

  double cont_w = -H0;
  double accept_sum = 0.0;
  int accept_count = 0;
  int depth = 1;
  int steps = 0;
  bool divergent = false;
  bool turning = false;
  double prop_energy = H0;
  double prop_error = 0.0;
  uint32_t leaf_counter = 0;

  // first single-leaf trajectory
  {
    const float u = readlane(dir_u, 0);
    const bool forward = !(u < 0.5f);
    const float ss = forward ? h : -h;
    const float half = ss / 2.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      p[k] = p[k] + half * g[k];
      q[k] = q[k] + ss * (im[k] * p[k]);
    }
    bool ok;
    const float lj1 = m.eval(q, g, ok);
    if (!ok) {
      divergent = true;
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) p[k] = p[k] + half * g[k];
      if (forward) { copyv(Rq, q); copyv(Rp, p); copyv(Rg, g); ljR = lj1; }
      else { copyv(Lq, q); copyv(Lp, p); copyv(Lg, g); ljL = lj1; }
      const double pe = kinetic(p, im) - (double)lj1;
      const double delta = pe - H0;
      steps = 1;
      if (!isfinite_d(delta) || delta > div_thr) {
        divergent = true;
      } else {
        const double cand = -pe;
        double hi, arg;
        lae_split(cont_w, cand, hi, arg);
        const double ex = exp_d(lane == 0 ? fmin(0.0, -delta) : arg);
        const double e = readlane(ex, 0);
        accept_sum += e < 1.0 ? e : 1.0;
        accept_count += 1;
        const double comb = lae_join(hi, readlane(ex, 1));
        leaf_lu = log_d(rand_uniform_d(chain, itc, STREAM_LEAF, (uint32_t)lane));
        leaf_block = 0;
        if (!isfinite_d(cont_w) || readlane(leaf_lu, 0) < cand - comb) {
          copyv(pq, q); copyv(pg, g); ljP = lj1;
          prop_energy = pe;
          prop_error = delta;
        }
        cont_w = comb;
        turning = is_turning_lr(Lq, Lp, Rq, Rp, im);
      }
      leaf_counter += 1;
    }
  }

  // doubling loop (reference _continue_nuts_proposal!)
  while (depth < max_depth && !divergent && !turning) {
    const uint32_t round = (uint32_t)depth;
    const float ud = readlane(dir_u, (int)round);
    const bool forward = !(ud < 0.5f);
    const float ss = forward ? h : -h;
    const float half = ss / 2.0f;
    if (forward) { copyv(q, Rq); copyv(p, Rp); copyv(g, Rg); }
    else { copyv(q, Lq); copyv(p, Lp); copyv(g, Lg); }
    float cur_lj = forward ? ljR : ljL;

    double sub_w = -inf_d();
    int sub_int = 0;
    double sub_asum = 0.0;
    int sub_acount = 0;
    bool sub_div = false, sub_turn = false;
    double sub_energy = 0.0, sub_error = 0.0;

    const int nleaves = 1 << depth;
    for (int leaf = 0; leaf < nleaves; ++leaf) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        p[k] = p[k] + half * g[k];
        q[k] = q[k] + ss * (im[k] * p[k]);
      }
      bool ok;
      const float ljc = m.eval(q, g, ok);
      if (!ok) { sub_div = true; break; }
#pragma unroll
      for (int k = 0; k < K; ++k) p[k] = p[k] + half * g[k];
      cur_lj = ljc;
      sub_int += 1;
      const double pe = kinetic(p, im) - (double)ljc;
      const double delta = pe - H0;
      if (!isfinite_d(delta) || delta > div_thr) { sub_div = true; break; }
      const double cand = -pe;
      double hi, arg;
      lae_split(sub_w, cand, hi, arg);
      const double ex = exp_d(lane == 0 ? fmin(0.0, -delta) : arg);
      const double e = readlane(ex, 0);
      sub_asum += e < 1.0 ? e : 1.0;
      sub_acount += 1;
      const double comb = lae_join(hi, readlane(ex, 1));
      const int blk = (int)(leaf_counter >> 6);
      if (blk != leaf_block) {
        leaf_lu = log_d(rand_uniform_d(chain, itc, STREAM_LEAF, ((uint32_t)blk << 6) + (uint32_t)lane));
        leaf_block = blk;
      }
      if (!isfinite_d(sub_w) || readlane(leaf_lu, (int)(leaf_counter & 63u)) < cand - comb) {
        copyv(sq, q); copyv(sg, g); ljS = ljc;
        sub_energy = pe;
        sub_error = delta;
      }
      sub_w = comb;
      leaf_counter += 1;
      const int li = sub_int - 1;
      if ((li & 1) == 0) {
        const int slot = __builtin_popcount((unsigned)li);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          ck[((slot * 2 + 0) * K + k) * WAVE + lane] = q[k];
          ck[((slot * 2 + 1) * K + k) * WAVE + lane] = p[k];
        }
      } else if (dyadic_turning(ck, q, p, li, forward, im)) {
        sub_turn = true;
        break;
      }
    }

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
    if (forward) { copyv(Rq, q); copyv(Rp, p); copyv(Rg, g); ljR = cur_lj; }
    else { copyv(Lq, q); copyv(Lp, p); copyv(Lg, g); ljL = cur_lj; }
    double comb_cw = cont_w;
    if (isfinite_d(sub_w)) {
      comb_cw = logaddexp_d(cont_w, sub_w);
      if (readlane(merge_lu, (int)round) < sub_w - cont_w) {
        copyv(pq, sq); copyv(pg, sg); ljP = ljS;
        prop_energy = sub_energy;
        prop_error = sub_error;
      }
    }
    steps += sub_int;
    accept_sum += sub_asum;
    accept_count += sub_acount;
    if (isfinite_d(sub_w)) cont_w = comb_cw;
    turning = sub_turn || is_turning_lr(Lq, Lp, Rq, Rp, im);
  }

  out.moved = any_differs(pq, q0);
  out.accept_prob = accept_count == 0 ? 0.0 : accept_sum / (double)accept_count;
  out.energy = prop_energy;
  out.error = prop_error;
  out.divergent = divergent;
  out.depth = depth;
  out.steps = steps;
  if (out.moved) {
    copyv(q, pq); copyv(g, pg); lj = ljP;
  } else {
    copyv(q, q0);
    // gradient of the start point is the caller's; restored by the caller
  }
}

// ------------------------------------------------------------------ HMC transition
// Fixed-length leapfrog with per-chain step (reference
// _device_leapfrog_integrate_perchain!, src/device/hmc_kernels.jl:252-280) and the
// host accept rule (hmc_kernels.jl:555-577), both on the device: the momentum and
// the accept uniform are Philox draws.
template <class M>
__device__ void hmc_transition(const M& m, const float (&im)[M::K], float h, uint32_t chain,
                               uint32_t itc, int L, double div_thr, uint32_t momentum_stream,
                               uint32_t accept_stream, float (&q)[M::K], float (&g)[M::K],
                               float& lj, TransitionOut& out, double* lar_out) {
  constexpr int K = M::K;
  const int lane = threadIdx.x & (WAVE - 1);
  float p[K], nq[K], ng[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = m.api(lane, k);
    if (a >= 0) {
      const float z = rand_normal_f(chain, itc, momentum_stream, (uint32_t)a);
      p[k] = z / sqrt_f(im[k]);
    } else {
      p[k] = 0.0f;
    }
  }
  const double H0 = kinetic(p, im) - (double)lj;
  copyv(nq, q);
  copyv(ng, g);
  float nlj = lj;
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) fin = fin && isfinite_f(g[k]);
  bool valid = wave_all(fin);
  const float half = 0.5f * h;
  if (valid) {
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = p[k] + half * ng[k];
  }
  for (int s = 1; s <= L; ++s) {
    if (!valid) break;
#pragma unroll
    for (int k = 0; k < K; ++k) nq[k] = nq[k] + (h * im[k]) * p[k];
    bool ok;
    nlj = m.eval(nq, ng, ok);
    if (s < L) {
      bool gfin = true;
#pragma unroll
      for (int k = 0; k < K; ++k) gfin = gfin && isfinite_f(ng[k]);
      valid = valid && wave_all(gfin);
      if (valid) {
#pragma unroll
        for (int k = 0; k < K; ++k) p[k] = p[k] + (1.0f * h) * ng[k];
      }
    } else {
      valid = valid && ok;
    }
  }
  if (valid) {
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = -(p[k] + half * ng[k]);
  }
  out.depth = 0;
  out.steps = L;
  out.moved = false;
  if (valid) {
    const double H1 = kinetic(p, im) - (double)nlj;
    const double lar = H0 - H1;
    const double err = H1 - H0;
    out.error = err;
    out.divergent = !isfinite_d(err) || err > div_thr;
    out.accept_prob = !isfinite_d(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
    if (lar_out) *lar_out = lar;
    if (accept_stream != 0xffffffffu) {
      const double u = rand_uniform_d(chain, itc, accept_stream, 0u);
      out.moved = log_d(u) < fmin(0.0, lar);
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
    copyv(q, nq); copyv(g, ng); lj = nlj;
  }
}

// ------------------------------------------------------------------ kernels
template <class M, class T>
__global__ __launch_bounds__(64 * WavesOf<M>::value) void k_density(M model, const T* __restrict__ params,
                                                                    T* __restrict__ values,
                                                                    T* __restrict__ grads, int P) {
  constexpr int K = M::K;
  typename M::template Lane<T> ld;
  model.prepare(ld);
  const Bound<M, T> m{model, ld};
  const int lane = threadIdx.x & (WAVE - 1);
  const bool lead = threadIdx.x < WAVE;  // wavefront 0 writes the results
  const long long c = blockIdx.x;
  T q[K], g[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = m.api(lane, k);
    q[k] = a >= 0 ? params[c * P + a] : T(0);
  }
  bool ok;
  const T v = m.eval(q, g, ok);
  if (grads && lead) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      if (a >= 0) grads[c * P + a] = g[k];
    }
  }
  if (threadIdx.x == 0) values[c] = v;
}

template <class M>
__device__ inline void load_chain(const M& m, const ChainState& st, int P, long long c,
                                  float (&q)[M::K], float (&g)[M::K], float (&im)[M::K]) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int k = 0; k < M::K; ++k) {
    const int a = m.api(lane, k);
    q[k] = a >= 0 ? st.pos[c * P + a] : 0.0f;
    g[k] = a >= 0 ? st.grad[c * P + a] : 0.0f;
    im[k] = a >= 0 ? st.im_f[a] : 0.0f;
  }
}

template <class M>
__device__ inline void store_vec(const M& m, float* dst, const float (&v)[M::K]) {
  if (threadIdx.x >= WAVE) return;  // wavefront 0 writes a workgroup-per-chain state
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int k = 0; k < M::K; ++k) {
    const int a = m.api(lane, k);
    if (a >= 0) dst[a] = v[k];
  }
}

// Runs iterations [it_begin, it_end] (1-based, inclusive) of batched NUTS for one
// chain per workgroup: transition, accept, and -- during warmup -- the chain's
// dual-averaging update and its warmup-history record for the pooled mass.  Each
// wavefront of a workgroup-per-chain model keeps its own checkpoint stack.
template <class M>
__global__ __launch_bounds__(64 * WavesOf<M>::value) void k_nuts(M model, ChainState st, SamplerParams sp,
                                                                 int it_begin, int it_end) {
  extern __shared__ float ck_all[];
  float* ck = ck_all + (size_t)(threadIdx.x >> 6) * ((sp.max_depth + 1) * 2 * M::K * WAVE);
  constexpr int K = M::K;
  typename M::template Lane<float> ld;
  model.prepare(ld);
  const Bound<M, float> m{model, ld};
  const int lane = threadIdx.x;
  const long long c = blockIdx.x;
  const int P = sp.P;
  float q[K], g[K], im[K];
  load_chain(m, st, P, c, q, g, im);
  float lj = st.lj[c];
  double step = st.step[c];
  double dual[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) dual[i] = st.dual[c * 5 + i];
  const uint32_t chain = (uint32_t)c;
  unsigned long long sample_steps = 0;
  for (int it = it_begin; it <= it_end; ++it) {
    const uint32_t itc = sp.run_seed + (uint32_t)it;
    const float h = (float)step;
    float g_keep[K];
    copyv(g_keep, g);
    const float lj_keep = lj;
    TransitionOut o;
    nuts_transition(m, im, h, chain, itc, sp.max_depth, sp.div_thr, q, g, lj, ck, o);
    if (!o.moved) { copyv(g, g_keep); lj = lj_keep; }
    const uint8_t flags = (uint8_t)((o.divergent ? FLAG_DIVERGENT : 0) | (o.moved ? FLAG_ACCEPTED : 0));
    if (it <= sp.num_warmup) {
      if (sp.adapt_step) {
        step = dual_update(dual, o.divergent ? 0.0 : o.accept_prob, sp.target_accept);
        if (it == sp.num_warmup) step = exp_d(dual[2]);
      }
      const long long row = (long long)(it - 1) * sp.C + c;
      store_vec(m, sp.hist_pos + row * P, q);
      if (lane == 0) {
        sp.hist_acc[row] = o.accept_prob;
        sp.hist_flags[row] = flags;
      }
    } else {
      const long long s = it - sp.num_warmup - 1;
      const long long row = c * sp.num_samples + s;
      store_vec(m, sp.draws + row * P, q);
      if (lane == 0) {
        sp.o_lj[row] = (double)lj;
        sp.o_acc[row] = o.accept_prob;
        sp.o_energy[row] = o.energy;
        sp.o_err[row] = o.error;
        sp.o_accepted[row] = o.moved ? 1 : 0;
        sp.o_divergent[row] = o.divergent ? 1 : 0;
        sp.o_depth[row] = o.depth;
        sp.o_steps[row] = o.steps;
      }
      sample_steps += o.steps;
    }
  }
  if (lane == 0 && sample_steps > 0) atomicAdd(sp.eval_counter, sample_steps);
  store_vec(m, st.pos + c * P, q);
  store_vec(m, st.grad + c * P, g);
  if (lane == 0) {
    st.lj[c] = lj;
    st.step[c] = step;
#pragma unroll
    for (int i = 0; i < 5; ++i) st.dual[c * 5 + i] = dual[i];
  }
}

template <class M>
__global__ __launch_bounds__(64 * WavesOf<M>::value) void k_hmc(M model, ChainState st, SamplerParams sp,
                                                                int it_begin, int it_end) {
  constexpr int K = M::K;
  typename M::template Lane<float> ld;
  model.prepare(ld);
  const Bound<M, float> m{model, ld};
  const int lane = threadIdx.x;
  const long long c = blockIdx.x;
  const int P = sp.P;
  float q[K], g[K], im[K];
  load_chain(m, st, P, c, q, g, im);
  float lj = st.lj[c];
  double step = st.step[c];
  double dual[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) dual[i] = st.dual[c * 5 + i];
  const uint32_t chain = (uint32_t)c;
  unsigned long long sample_steps = 0;
  for (int it = it_begin; it <= it_end; ++it) {
    const uint32_t itc = sp.run_seed + (uint32_t)it;
    const float h = (float)step;
    TransitionOut o;
    hmc_transition(m, im, h, chain, itc, sp.num_leapfrog, sp.div_thr, STREAM_MOMENTUM,
                   STREAM_HMC_ACCEPT, q, g, lj, o, nullptr);
    const uint8_t flags = (uint8_t)((o.divergent ? FLAG_DIVERGENT : 0) | (o.moved ? FLAG_ACCEPTED : 0));
    if (it <= sp.num_warmup) {
      if (sp.adapt_step) {
        step = dual_update(dual, o.divergent ? 0.0 : o.accept_prob, sp.target_accept);
        if (it == sp.num_warmup) step = exp_d(dual[2]);
      }
      const long long row = (long long)(it - 1) * sp.C + c;
      store_vec(m, sp.hist_pos + row * P, q);
      if (lane == 0) {
        sp.hist_acc[row] = o.accept_prob;
        sp.hist_flags[row] = flags;
      }
    } else {
      const long long s = it - sp.num_warmup - 1;
      const long long row = c * sp.num_samples + s;
      store_vec(m, sp.draws + row * P, q);
      if (lane == 0) {
        sp.o_lj[row] = (double)lj;
        sp.o_acc[row] = o.accept_prob;
        sp.o_energy[row] = o.energy;
        sp.o_err[row] = o.error;
        sp.o_accepted[row] = o.moved ? 1 : 0;
        sp.o_divergent[row] = o.divergent ? 1 : 0;
        sp.o_depth[row] = 0;
        sp.o_steps[row] = sp.num_leapfrog;
      }
      sample_steps += (unsigned long long)sp.num_leapfrog;
    }
  }
  if (lane == 0 && sample_steps > 0) atomicAdd(sp.eval_counter, sample_steps);
  store_vec(m, st.pos + c * P, q);
  store_vec(m, st.grad + c * P, g);
  if (lane == 0) {
    st.lj[c] = lj;
    st.step[c] = step;
#pragma unroll
    for (int i = 0; i < 5; ++i) st.dual[c * 5 + i] = dual[i];
  }
}

// Per-chain reasonable step size (reference
// _find_reasonable_batched_step_size_per_chain, nuts/adaptation.jl:437-520): one
// momentum draw, a one-step leapfrog at the current step, then step doubling or
// halving until the log acceptance ratio crosses log(1/2), at most 20 trials.
// Afterwards the chain's dual-averaging state restarts from its step
// (adaptation.jl:1-13).
template <class M>
__global__ __launch_bounds__(64 * WavesOf<M>::value) void k_step_search(M model, ChainState st, int P,
                                                                        uint32_t itc, int do_search,
                                                                        double target) {
  typename M::template Lane<float> ld;
  model.prepare(ld);
  const Bound<M, float> m{model, ld};
  constexpr int K = M::K;
  const int lane = threadIdx.x;
  const long long c = blockIdx.x;
  double s = st.step[c];
  if (do_search) {
    float q[K], g[K], im[K];
    load_chain(m, st, P, c, q, g, im);
    const float lj = st.lj[c];
    const double log_target = -0.6931471805599453;  // log(0.5)
    const double min_step = 1e-8, max_step = 1e3;
    auto trial_lar = [&](double step) {
      float tq[K], tg[K];
      copyv(tq, q);
      copyv(tg, g);
      float tlj = lj;
      TransitionOut o;
      double lar;
      hmc_transition(m, im, (float)step, (uint32_t)c, itc, 1, 0.0, STREAM_STEP_SEARCH,
                     0xffffffffu, tq, tg, tlj, o, &lar);
      return lar;
    };
    const double lar0 = trial_lar(s);
    const double dir = lar0 > log_target ? 1.0 : -1.0;
    for (int t = 0; t < 20; ++t) {
      const double next = s * (dir > 0.0 ? 2.0 : 0.5);
      if (next < min_step || next > max_step) break;
      const double lar = trial_lar(next);
      s = next;
      if ((dir > 0.0 && lar <= log_target) || (dir < 0.0 && lar >= log_target)) break;
    }
    s = s < min_step ? min_step : (s > max_step ? max_step : s);
  }
  if (lane == 0) {
    st.step[c] = s;
    const double ls = log_d(s);
    st.dual[c * 5 + 0] = log_d(10.0 * s);
    st.dual[c * 5 + 1] = ls;
    st.dual[c * 5 + 2] = ls;
    st.dual[c * 5 + 3] = 0.0;
    st.dual[c * 5 + 4] = 0.0;
  }
  (void)target;
}

}  // namespace ut
