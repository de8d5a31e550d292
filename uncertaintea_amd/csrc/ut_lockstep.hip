// ut_lockstep.hip -- lock-step (all chains, same number of leapfrog steps) HMC
// kernels for ChEES, and ChEES's cross-chain reductions.
//
// The trajectory is split into kernels around the batched gradient, so any
// gradient engine fits between them (the wavefront-per-chain density kernel, or
// the fused matrix-core GLM that evaluates all chains at once):
//   k_ls_begin  momentum draw, H0, half kick, first drift
//   [gradient]  k_ls_step  full kick + drift          (L - 1 times)
//   [gradient]  k_ls_end   final half kick, momentum flip, H1, Metropolis accept
// One 64-lane wavefront per chain; parameters in the reference order, parameter
// a = lane + 64 k in register slot k.  Arithmetic follows the reference's batched
// leapfrog (src/inference/integrator.jl:146-226; device twin
// src/device/hmc_kernels.jl:252-280) with the shared ChEES step size; energies,
// acceptance and all cross-chain reductions are binary64 and sequential in chain
// order, as the reference's host loops (src/inference/api_batched_chees.jl:158-290).
#include "ut_kernels.h"
#include "ut_lockstep.h"
#include "ut_models.h"

namespace ut {

namespace {

template <int K>
__device__ inline double ls_kinetic(const float (&p)[K], const float (&im)[K]) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = acc + ((double)p[k] * (double)p[k]) * (double)im[k];
  return wave_sum(acc) * 0.5;
}

template <int K>
__global__ __launch_bounds__(64) void k_ls_begin(LockstepState s, int P, uint32_t itc, uint32_t mom_stream,
                                                 float eps, int save_initials) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  float q[K], g[K], p[K], im[K];
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    const bool in = a < P;
    q[k] = in ? s.pos[c * P + a] : 0.0f;
    g[k] = in ? s.grad[c * P + a] : 0.0f;
    im[k] = in ? s.im_f[a] : 0.0f;
    p[k] = in ? rand_normal_f((uint32_t)c, itc, mom_stream, (uint32_t)a) / sqrt_f(im[k]) : 0.0f;
    fin = fin && isfinite_f(g[k]);
    if (in && save_initials) s.initials[c * P + a] = q[k];
  }
  const double H0 = ls_kinetic(p, im) - (double)s.lj[c];
  const bool valid = wave_all(fin);
  if (valid) {
    const float half = 0.5f * eps;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      p[k] = p[k] + half * g[k];
      q[k] = q[k] + (eps * im[k]) * p[k];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    if (a < P) {
      s.wq[c * P + a] = q[k];
      s.wp[c * P + a] = p[k];
      s.wg[c * P + a] = g[k];
    }
  }
  if (lane == 0) {
    s.h0[c] = H0;
    s.valid[c] = valid ? 1 : 0;
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_ls_step(LockstepState s, int P, float eps) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  if (!s.valid[c]) return;  // uniform per workgroup
  float q[K], g[K], p[K], im[K];
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    const bool in = a < P;
    q[k] = in ? s.wq[c * P + a] : 0.0f;
    g[k] = in ? s.wg[c * P + a] : 0.0f;
    p[k] = in ? s.wp[c * P + a] : 0.0f;
    im[k] = in ? s.im_f[a] : 0.0f;
    fin = fin && isfinite_f(g[k]);
  }
  const bool valid = wave_all(fin);
  if (!valid) {
    if (lane == 0) s.valid[c] = 0;
    return;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    p[k] = p[k] + (1.0f * eps) * g[k];
    q[k] = q[k] + (eps * im[k]) * p[k];
    const int a = lane + 64 * k;
    if (a < P) {
      s.wq[c * P + a] = q[k];
      s.wp[c * P + a] = p[k];
    }
  }
}

// mode 0: ChEES iteration (accept, record);  mode 1: step-size search trial
// (accept probability only, state untouched).
template <int K>
__global__ __launch_bounds__(64) void k_ls_end(LockstepState s, LockstepRecord r, int P, uint32_t itc,
                                               float eps, double div_thr, int mode) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  float q[K], g[K], p[K], im[K];
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    const bool in = a < P;
    q[k] = in ? s.wq[c * P + a] : 0.0f;
    g[k] = in ? s.wg[c * P + a] : 0.0f;
    p[k] = in ? s.wp[c * P + a] : 0.0f;
    im[k] = in ? s.im_f[a] : 0.0f;
    fin = fin && isfinite_f(g[k]);
  }
  const float nlj = s.wlj[c];
  bool valid = s.valid[c] && wave_all(fin) && isfinite_f(nlj);
  if (valid) {
    const float half = 0.5f * eps;
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = -(p[k] + half * g[k]);
  }
  const double H0 = s.h0[c];
  double ap = 0.0, err = inf_d(), energy = H0, lar = -inf_d();
  bool div = true, acc = false;
  if (valid) {
    const double H1 = ls_kinetic(p, im) - (double)nlj;
    lar = H0 - H1;
    err = H1 - H0;
    div = !isfinite_d(err) || err > div_thr;
    ap = !isfinite_d(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
    if (mode == 0) {
      const double u = rand_uniform_d((uint32_t)c, itc, STREAM_HMC_ACCEPT, 0u);
      acc = log_d(u) < fmin(0.0, lar);
    }
    energy = acc ? H1 : H0;
  }
  bool pfin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    if (a < P) {
      pfin = pfin && isfinite_f(q[k]) && isfinite_f(p[k]);
      s.wp[c * P + a] = p[k];  // the flipped end momentum feeds the ChEES criterion
    }
  }
  pfin = wave_all(pfin);
  if (lane == 0) {
    r.accept_prob[c] = ap;
    r.flags[c] = (uint8_t)((div ? FLAG_DIVERGENT : 0) | (acc ? FLAG_ACCEPTED : 0));
    r.pfinite[c] = pfin ? 1 : 0;
    r.energy[c] = energy;
    r.error[c] = err;
  }
  if (mode == 0 && acc) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = lane + 64 * k;
      if (a < P) {
        s.pos[c * P + a] = q[k];
        s.grad[c * P + a] = g[k];
      }
    }
    if (lane == 0) s.lj[c] = nlj;
  }
}

// copies positions / statistics of iteration `it` into the warmup history or the
// sample buffers after k_ls_end
__global__ void k_ls_record(LockstepState s, LockstepRecord r, SamplerParams sp, int it, int L) {
  const long long c = blockIdx.x;
  const int P = sp.P;
  if (it <= sp.num_warmup) {
    const long long row = (long long)(it - 1) * sp.C + c;
    for (int a = threadIdx.x; a < P; a += blockDim.x) sp.hist_pos[row * P + a] = s.pos[c * P + a];
    if (threadIdx.x == 0) {
      sp.hist_acc[row] = r.accept_prob[c];
      sp.hist_flags[row] = r.flags[c];
    }
  } else {
    const long long srow = c * sp.num_samples + (it - sp.num_warmup - 1);
    for (int a = threadIdx.x; a < P; a += blockDim.x) sp.draws[srow * P + a] = s.pos[c * P + a];
    if (threadIdx.x == 0) {
      sp.o_lj[srow] = (double)s.lj[c];
      sp.o_acc[srow] = r.accept_prob[c];
      sp.o_energy[srow] = r.energy[c];
      sp.o_err[srow] = r.error[c];
      sp.o_accepted[srow] = (r.flags[c] & FLAG_ACCEPTED) ? 1 : 0;
      sp.o_divergent[srow] = (r.flags[c] & FLAG_DIVERGENT) ? 1 : 0;
      sp.o_depth[srow] = 0;
      sp.o_steps[srow] = L;
    }
  }
}

// ---- ChEES reductions (binary64, chain order) ------------------------------------
// phase 1: per parameter, the plain mean of the initial positions and the
// acceptance-weighted, divergence-masked mean of the proposals
// (api_batched_chees.jl:171-203)
__global__ void k_chees_means(LockstepState s, LockstepRecord r, const double* im_d, int P, int C,
                              double* stats) {
  // the sums run in chain order; each group of 8 chains is loaded before it is added
  // so that the loads overlap instead of each waiting for the previous addition
  constexpr int G = 8;
  for (int a = threadIdx.x; a < P; a += blockDim.x) {
    double isum = 0.0, psum = 0.0, wsum = 0.0;
    long long icount = 0;
    for (int c0 = 0; c0 < C; c0 += G) {
      float iv[G], qv[G];
      double wv[G];
      bool iok[G], qok[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int c = c0 + k;
        if (c < C) {
          iok[k] = r.ifinite[c] != 0;
          qok[k] = !(r.flags[c] & FLAG_DIVERGENT) && r.qfinite[c];
          iv[k] = s.initials[(long long)c * P + a];
          qv[k] = s.wq[(long long)c * P + a];
          wv[k] = r.accept_prob[c];
        } else {
          iok[k] = qok[k] = false;
        }
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        if (iok[k]) {
          isum += (double)iv[k];
          icount += 1;
        }
        if (qok[k]) {
          wsum += wv[k];
          psum += wv[k] * (double)qv[k];
        }
      }
    }
    r.imean[a] = icount > 0 ? isum / (double)icount : 0.0;
    r.pmean[a] = psum * (1.0 / (wsum + 1e-20));
    r.inv_sqrt[a] = 1.0 / sqrt_d(im_d[a]);
  }
  (void)stats;
}

// phase 2: per chain, accept * (|dx'|^2 - |dx|^2) * <dx', v'> in the whitened
// metric (api_batched_chees.jl:205-229); v' is the stored end momentum times vsign
__global__ void k_chees_terms(LockstepState s, LockstepRecord r, const double* im_d, int P, int C, double vsign) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  if ((r.flags[c] & FLAG_DIVERGENT) || !r.pfinite[c]) {
    r.term[c] = 0.0;
    r.include[c] = 0;
    return;
  }
  double se = 0.0, ss = 0.0, ev = 0.0;
  for (int a = 0; a < P; ++a) {
    const double sigma = im_d[a];
    const double inv_sqrt = r.inv_sqrt[a];
    const double we = ((double)s.wq[c * P + a] - r.pmean[a]) * inv_sqrt;
    const double ws = ((double)s.initials[c * P + a] - r.imean[a]) * inv_sqrt;
    const double wv = vsign * ((double)s.wp[c * P + a] * sigma * inv_sqrt);
    se += we * we;
    ss += ws * ws;
    ev += we * wv;
  }
  r.term[c] = r.accept_prob[c] * (se - ss) * ev;
  r.include[c] = 1;
}

// phase 3: sequential totals -> stats[0] numerator, [1] denominator,
// [2] harmonic-mean acceptance of non-divergent chains, [3] mean acceptance with
// divergent chains counted as 0
__global__ void k_chees_final(LockstepRecord r, int C, double* stats) {
  // chunks of the per-chain values are staged in LDS by the whole workgroup; one
  // thread then adds them in chain order
  constexpr int CH = 1024;
  __shared__ double s_term[CH], s_acc[CH];
  __shared__ uint8_t s_flag[CH], s_inc[CH];
  double num = 0.0, den = 0.0, hsum = 0.0, msum = 0.0;
  long long hcount = 0;
  for (int c0 = 0; c0 < C; c0 += CH) {
    const int nb = C - c0 < CH ? C - c0 : CH;
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
      const int c = c0 + k;
      s_acc[k] = r.accept_prob[c];
      s_flag[k] = r.flags[c];
      s_inc[k] = (r.include && r.include[c]) ? 1 : 0;
      s_term[k] = s_inc[k] ? r.term[c] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < nb; ++k) {
        const bool div = s_flag[k] & FLAG_DIVERGENT;
        if (s_inc[k]) {
          num += s_term[k];
          den += s_acc[k] + 1e-20;
        }
        if (!div) {
          hsum += 1.0 / s_acc[k];
          hcount += 1;
        }
        msum += div ? 0.0 : s_acc[k];
      }
    }
  }
  if (threadIdx.x != 0) return;
  double hm = 0.0;
  if (hcount > 0) {
    hm = (double)hcount / hsum;
    if (!isfinite_d(hm)) hm = 0.0;
  }
  stats[0] = num;
  stats[1] = den;
  stats[2] = hm;
  stats[3] = msum / (double)C;
}

// initial-position finiteness flags (ChEES initials mean)
__global__ void k_ls_initial_flags(LockstepState s, LockstepRecord r, int P, int C) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  bool f = true, fq = true;
  for (int a = 0; a < P; ++a) {
    f = f && isfinite_f(s.initials[c * P + a]);
    fq = fq && isfinite_f(s.wq[c * P + a]);
  }
  r.ifinite[c] = f ? 1 : 0;
  r.qfinite[c] = fq ? 1 : 0;
}

template <class F>
hipError_t by_k(int P, F&& f) {
  const int k = (P + 63) / 64;
  if (k <= 1) return f(std::integral_constant<int, 1>());
  if (k <= 2) return f(std::integral_constant<int, 2>());
  if (k <= 4) return f(std::integral_constant<int, 4>());
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_ls_begin(const LockstepState& s, int P, int C, uint32_t itc, uint32_t mom_stream, float eps,
                           int save_initials, hipStream_t stream) {
  return by_k(P, [&](auto kk) {
    hipLaunchKernelGGL((k_ls_begin<decltype(kk)::value>), dim3(C), dim3(64), 0, stream, s, P, itc, mom_stream, eps,
                       save_initials);
    return hipGetLastError();
  });
}

hipError_t launch_ls_step(const LockstepState& s, int P, int C, float eps, hipStream_t stream) {
  return by_k(P, [&](auto kk) {
    hipLaunchKernelGGL((k_ls_step<decltype(kk)::value>), dim3(C), dim3(64), 0, stream, s, P, eps);
    return hipGetLastError();
  });
}

hipError_t launch_ls_end(const LockstepState& s, const LockstepRecord& r, int P, int C, uint32_t itc, float eps,
                         double div_thr, int mode, hipStream_t stream) {
  return by_k(P, [&](auto kk) {
    hipLaunchKernelGGL((k_ls_end<decltype(kk)::value>), dim3(C), dim3(64), 0, stream, s, r, P, itc, eps, div_thr,
                       mode);
    return hipGetLastError();
  });
}

hipError_t launch_ls_record(const LockstepState& s, const LockstepRecord& r, const SamplerParams& sp, int it, int L,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_ls_record, dim3(sp.C), dim3(64), 0, stream, s, r, sp, it, L);
  return hipGetLastError();
}

hipError_t launch_chees_stats(const LockstepState& s, const LockstepRecord& r, const double* im_d, int P, int C,
                              bool trajectory_terms, double vsign, double* stats, hipStream_t stream) {
  if (trajectory_terms) {
    hipLaunchKernelGGL(k_ls_initial_flags, dim3((C + 255) / 256), dim3(256), 0, stream, s, r, P, C);
    hipLaunchKernelGGL(k_chees_means, dim3(1), dim3(256), 0, stream, s, r, im_d, P, C, stats);
    hipLaunchKernelGGL(k_chees_terms, dim3((C + 255) / 256), dim3(256), 0, stream, s, r, im_d, P, C, vsign);
  }
  LockstepRecord rr = r;
  if (!trajectory_terms) rr.include = nullptr;
  hipLaunchKernelGGL(k_chees_final, dim3(1), dim3(256), 0, stream, rr, C, stats);
  return hipGetLastError();
}

}  // namespace ut
