// ut_meads.hip -- MEADS (reference batched_meads, src/inference/api_batched_meads.jl:
// constants :32-34, max-eigenvalue estimate :48-77, fold parameters :89-136, slice
// drift :141, iteration :337-470) as lock-step kernels around the batched gradient.
//
// Every iteration: per fold k, the neighbour fold (k - 1) mod K supplies the
// diagonal preconditioner (population std per parameter), the whitened position
// and gradient ensembles, and through their Gram trace ratios the step size,
// damping and slice drift (k_meads_white, k_meads_gram, k_meads_fold_params; binary64,
// fixed reduction orders).  Then every chain refreshes its persistent whitened
// momentum, takes ONE leapfrog step with its fold's step and inverse mass, and the
// persistent-slice Metropolis test decides (k_meads_begin / gradient /
// k_meads_end).  The rotating skipped fold keeps its state.  No host
// synchronisation: fold parameters live on the device and are traced there.
#include "ut_kernels.h"
#include "ut_meads.h"
#include "ut_models.h"

namespace ut {

namespace {

constexpr double MEADS_STEP_MARGIN = 0.5;
constexpr double MEADS_EIG_FLOOR = 1e-12;
constexpr double MEADS_STD_FLOOR = 1e-10;
constexpr int FOLD_THREADS = 256;

// Whitening of the neighbour fold (k - 1) mod K for fold k, one thread per parameter:
// population mean and std over the members in member order, sigma, the whitened
// positions (x - mean) / sd and gradients g * sd (member-major).
__global__ __launch_bounds__(64) void k_meads_white(MeadsState m, int P, int K, int nmax) {
  const int k = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int nb = (k - 1 + K) % K;
  const int c0 = m.fold_start[nb], n = m.fold_start[nb + 1] - c0;
  double* wpos = m.white + ((long long)(0 * K + k) * nmax) * P;
  double* wgrd = m.white + ((long long)(1 * K + k) * nmax) * P;
  double mean = 0.0;
  for (int j = 0; j < n; ++j) mean += (double)m.pos[(long long)(c0 + j) * P + p];
  mean /= (double)n;
  double var = 0.0;
  for (int j = 0; j < n; ++j) {
    const double d = (double)m.pos[(long long)(c0 + j) * P + p] - mean;
    var += d * d;
  }
  var /= (double)n;
  double sd = sqrt_d(var);
  sd = sd > MEADS_STD_FLOOR ? sd : MEADS_STD_FLOOR;
  m.sigma[(long long)k * P + p] = sd;
  m.sigma_f[(long long)k * P + p] = (float)sd;
  const double inv = 1.0 / sd;
  for (int j = 0; j < n; ++j) {
    wpos[(long long)j * P + p] = ((double)m.pos[(long long)(c0 + j) * P + p] - mean) * inv;
    wgrd[(long long)j * P + p] = (double)m.grad[(long long)(c0 + j) * P + p] * sd;
  }
}

// Gram matrices G = W W^T of the whitened ensembles, 64 x 64 entries per workgroup,
// 4 x 4 per thread; every entry is the sum over parameters in ascending order
// (separate multiply and add), the order of the binary64 oracle.
constexpr int GT = 64, GK = 16;
__global__ __launch_bounds__(256) void k_meads_gram(MeadsState m, int P, int K, int nmax) {
  const int z = blockIdx.z, k = z % K;
  const int nb = (k - 1 + K) % K;
  const int n = m.fold_start[nb + 1] - m.fold_start[nb];
  const int a0 = blockIdx.x * GT, b0 = blockIdx.y * GT;
  if (a0 >= n || b0 >= n) return;
  const double* W = m.white + ((long long)z * nmax) * P;  // z = mat * K + k
  double* G = m.gram + (long long)z * nmax * nmax;
  __shared__ double As[GK][GT], Bs[GK][GT];
  const int t = threadIdx.x, ta = t & 15, tb = t >> 4;
  double acc[4][4];
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int p0 = 0; p0 < P; p0 += GK) {
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
      const int e = t + 256 * s, row = e >> 4, pp = e & 15;
      const bool pin = p0 + pp < P;
      As[pp][row] = (a0 + row < n && pin) ? W[(long long)(a0 + row) * P + p0 + pp] : 0.0;
      Bs[pp][row] = (b0 + row < n && pin) ? W[(long long)(b0 + row) * P + p0 + pp] : 0.0;
    }
    __syncthreads();
    for (int pp = 0; pp < GK; ++pp) {
      double a[4], b[4];
      for (int u = 0; u < 4; ++u) a[u] = As[pp][ta + 16 * u];
      for (int v = 0; v < 4; ++v) b[v] = Bs[pp][tb + 16 * v];
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v) acc[u][v] = acc[u][v] + a[u] * b[v];
    }
  }
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v) {
      const int a = a0 + ta + 16 * u, b = b0 + tb + 16 * v;
      if (a < n && b < n) G[(long long)a * nmax + b] = acc[u][v];
    }
}

// Trace ratio lambda2_bar / lambda_bar from the Gram matrix G (n x n, row stride
// nmax), or 0 for a degenerate ensemble.  Fixed orders: squared off-diagonal entries
// over partner members, then norms and rows in member order.
__device__ double max_eig(const double* G, int n, int nmax, double* rows, double* norms) {
  for (int a = threadIdx.x; a < n; a += blockDim.x) {
    const double* ga = G + (long long)a * nmax;
    norms[a] = ga[a];
    double r = 0.0;
    for (int b = 0; b < n; ++b) {
      if (b == a) continue;
      r += ga[b] * ga[b];
    }
    rows[a] = r;
  }
  __syncthreads();
  __shared__ double result;
  if (threadIdx.x == 0) {
    double trace = 0.0, off = 0.0;
    for (int a = 0; a < n; ++a) trace += norms[a];
    for (int a = 0; a < n; ++a) off += rows[a];
    double e = 0.0;
    if (n >= 2) {
      const double lb = trace / (double)n;
      if (lb > 0.0) e = (off / ((double)n * (double)(n - 1))) / lb;
    }
    result = e;
  }
  __syncthreads();
  return result;
}

__global__ __launch_bounds__(FOLD_THREADS) void k_meads_fold_params(MeadsState m, int P, int K, int nmax, int t) {
  const int k = blockIdx.x;                 // fold receiving the parameters
  const int nb = (k - 1 + K) % K;           // its neighbour
  const int n = m.fold_start[nb + 1] - m.fold_start[nb];
  double* rows = m.gram_row + (long long)(0 * K + k) * nmax;
  double* norms = m.gram_row + (long long)(1 * K + k) * nmax;
  const double eig_g = max_eig(m.gram + (long long)(1 * K + k) * nmax * nmax, n, nmax, rows, norms);
  const double eig_p = max_eig(m.gram + (long long)(0 * K + k) * nmax * nmax, n, nmax, rows, norms);
  if (threadIdx.x == 0) {
    double mx = 0.0;
    for (int p = 0; p < P; ++p) mx = fmax(mx, m.sigma[(long long)k * P + p]);
    const double eg = eig_g > MEADS_EIG_FLOOR ? eig_g : MEADS_EIG_FLOOR;
    const double ep = eig_p > MEADS_EIG_FLOOR ? eig_p : MEADS_EIG_FLOOR;
    double step = MEADS_STEP_MARGIN / sqrt_d(eg);
    step = step < 1.0 ? step : 1.0;
    const double floor_g = 1.0 / ((double)t * step);
    const double inv_p = 1.0 / sqrt_d(ep);
    const double gamma = floor_g > inv_p ? floor_g : inv_p;
    double alpha = 1.0 - exp_d(-2.0 * step * gamma);
    alpha = alpha < 0.0 ? 0.0 : (alpha > 1.0 ? 1.0 : alpha);
    m.fold[k * 4 + 0] = step;
    m.fold[k * 4 + 1] = alpha;
    m.fold[k * 4 + 2] = alpha / 2.0;
    m.fold[k * 4 + 3] = mx;
  }
}

template <int K>
__device__ inline double meads_kinetic(const float (&p)[K], const float (&im)[K]) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = acc + ((double)p[k] * (double)p[k]) * (double)im[k];
  return wave_sum(acc) * 0.5;
}

template <int K>
__global__ __launch_bounds__(64) void k_meads_begin(MeadsState m, const int* __restrict__ fold_of, int P,
                                                    uint32_t itc) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  const int f = fold_of[c];
  const double alpha = m.fold[f * 4 + 1];
  const float keep = (float)sqrt_d(1.0 - alpha > 0.0 ? 1.0 - alpha : 0.0);
  const float refresh = (float)sqrt_d(alpha > 0.0 ? alpha : 0.0);
  const float eps = (float)m.fold[f * 4 + 0];
  float q[K], g[K], p[K], im[K];
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    const bool in = a < P;
    q[k] = in ? m.pos[c * P + a] : 0.0f;
    g[k] = in ? m.grad[c * P + a] : 0.0f;
    const double sd = in ? m.sigma[(long long)f * P + a] : 1.0;
    im[k] = in ? (float)(sd * sd) : 0.0f;
    p[k] = 0.0f;
    if (in) {
      const float xi = rand_normal_f((uint32_t)c, itc, STREAM_MEADS_REFRESH, (uint32_t)a);
      const float mt = keep * m.mom[c * P + a] + refresh * xi;
      m.refreshed[c * P + a] = mt;
      p[k] = mt / m.sigma_f[(long long)f * P + a];
    }
    fin = fin && isfinite_f(g[k]);
  }
  const double H0 = meads_kinetic(p, im) - (double)m.lj[c];
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
      m.wq[c * P + a] = q[k];
      m.wp[c * P + a] = p[k];
    }
  }
  if (lane == 0) {
    m.h0[c] = H0;
    m.valid[c] = valid ? 1 : 0;
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_meads_end(MeadsState m, MeadsRecord r, const int* __restrict__ fold_of,
                                                  int P, int skip_fold, double div_thr) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  const int f = fold_of[c];
  const float eps = (float)m.fold[f * 4 + 0];
  const double delta = m.fold[f * 4 + 2];
  float g[K], p[K], im[K];
  bool fin = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    const bool in = a < P;
    g[k] = in ? m.wg[c * P + a] : 0.0f;
    p[k] = in ? m.wp[c * P + a] : 0.0f;
    const double sd = in ? m.sigma[(long long)f * P + a] : 1.0;
    im[k] = in ? (float)(sd * sd) : 0.0f;
    fin = fin && isfinite_f(g[k]);
  }
  const float nlj = m.wlj[c];
  const bool valid = m.valid[c] && wave_all(fin) && isfinite_f(nlj);
  if (valid) {
    const float half = 0.5f * eps;
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = -(p[k] + half * g[k]);
  }
  const double H0 = m.h0[c];
  double ap = 0.0, err = inf_d(), lar = -inf_d();
  bool div = true;
  if (valid) {
    const double H1 = meads_kinetic(p, im) - (double)nlj;
    lar = H0 - H1;
    err = H1 - H0;
    div = !isfinite_d(err) || err > div_thr;
    ap = !isfinite_d(lar) ? 0.0 : (lar >= 0.0 ? 1.0 : exp_d(lar));
  }
  if (f == skip_fold) {
    if (lane == 0) {
      r.accept_prob[c] = ap;
      r.flags[c] = (uint8_t)(div ? FLAG_DIVERGENT : 0);
      r.energy_error[c] = err;
    }
    return;
  }
  // persistent slice: u~ = ((u + 1 + delta) mod 2) - 1, accept iff log|u~| <= log ratio
  const double x = m.slice[c] + 1.0 + delta;
  const double drifted = (x - 2.0 * floor(x / 2.0)) - 1.0;
  const bool acc = valid && log_d(fabs(drifted)) <= lar;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    if (a < P) {
      if (acc) {
        m.pos[c * P + a] = m.wq[c * P + a];
        m.grad[c * P + a] = g[k];
        m.mom[c * P + a] = -p[k] * m.sigma_f[(long long)f * P + a];
      } else {
        m.mom[c * P + a] = -m.refreshed[c * P + a];
      }
    }
  }
  if (lane == 0) {
    if (acc) {
      m.lj[c] = nlj;
      m.slice[c] = drifted * exp_d(err);
    } else {
      m.slice[c] = drifted;
    }
    r.accept_prob[c] = ap;
    r.flags[c] = (uint8_t)((div ? FLAG_DIVERGENT : 0) | (acc ? FLAG_ACCEPTED : 0));
    r.energy_error[c] = err;
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_meads_record(MeadsState m, MeadsRecord r, const int* __restrict__ fold_of,
                                                     SamplerParams sp, int it, int skip_fold) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  const int P = sp.P;
  if (it <= sp.num_warmup) return;
  const long long srow = c * sp.num_samples + (it - sp.num_warmup - 1);
  float u[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = lane + 64 * k;
    u[k] = a < P ? m.mom[c * P + a] : 0.0f;
    if (a < P) sp.draws[srow * P + a] = m.pos[c * P + a];
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = acc + (double)u[k] * (double)u[k];
  const double kin = wave_sum(acc);
  if (lane == 0) {
    const bool skipped = fold_of[c] == skip_fold;
    sp.o_lj[srow] = (double)m.lj[c];
    sp.o_energy[srow] = -(double)m.lj[c] + 0.5 * kin;
    sp.o_acc[srow] = skipped ? 1.0 : r.accept_prob[c];
    sp.o_accepted[srow] = skipped ? 1 : ((r.flags[c] & FLAG_ACCEPTED) ? 1 : 0);
    sp.o_divergent[srow] = skipped ? 0 : ((r.flags[c] & FLAG_DIVERGENT) ? 1 : 0);
    sp.o_err[srow] = skipped ? 0.0 : r.energy_error[c];
    sp.o_depth[srow] = 0;
    sp.o_steps[srow] = skipped ? 0 : 1;
  }
}

__global__ void k_meads_init(MeadsState m, int P, int C, uint32_t itc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)C * P) return;
  const int c = (int)(i / P), a = (int)(i % P);
  m.mom[i] = rand_normal_f((uint32_t)c, itc, STREAM_MEADS_INIT, (uint32_t)a);
  if (a == 0) m.slice[c] = 2.0 * rand_uniform_d((uint32_t)c, itc, STREAM_MEADS_INIT, (uint32_t)P) - 1.0;
}

template <class F>
hipError_t by_slots(int P, F&& f) {
  const int k = (P + 63) / 64;
  if (k <= 1) return f(std::integral_constant<int, 1>());
  if (k <= 2) return f(std::integral_constant<int, 2>());
  if (k <= 4) return f(std::integral_constant<int, 4>());
  if (k <= 8) return f(std::integral_constant<int, 8>());
  if (k <= 16) return f(std::integral_constant<int, 16>());
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_meads_fold_params(const MeadsState& m, int P, int K, int nmax, int t, hipStream_t stream) {
  hipLaunchKernelGGL(k_meads_white, dim3((P + 63) / 64, K), dim3(64), 0, stream, m, P, K, nmax);
  const int nt = (nmax + GT - 1) / GT;
  hipLaunchKernelGGL(k_meads_gram, dim3(nt, nt, 2 * K), dim3(256), 0, stream, m, P, K, nmax);
  hipLaunchKernelGGL(k_meads_fold_params, dim3(K), dim3(FOLD_THREADS), 0, stream, m, P, K, nmax, t);
  return hipGetLastError();
}

hipError_t launch_meads_begin(const MeadsState& m, const int* fold_of, int P, int C, uint32_t itc,
                              hipStream_t stream) {
  return by_slots(P, [&](auto kk) {
    hipLaunchKernelGGL((k_meads_begin<decltype(kk)::value>), dim3(C), dim3(64), 0, stream, m, fold_of, P, itc);
    return hipGetLastError();
  });
}

hipError_t launch_meads_end(const MeadsState& m, const MeadsRecord& r, const int* fold_of, int P, int C,
                            int skip_fold, double div_thr, hipStream_t stream) {
  return by_slots(P, [&](auto kk) {
    hipLaunchKernelGGL((k_meads_end<decltype(kk)::value>), dim3(C), dim3(64), 0, stream, m, r, fold_of, P,
                       skip_fold, div_thr);
    return hipGetLastError();
  });
}

hipError_t launch_meads_record(const MeadsState& m, const MeadsRecord& r, const int* fold_of,
                               const SamplerParams& sp, int it, int skip_fold, hipStream_t stream) {
  return by_slots(sp.P, [&](auto kk) {
    hipLaunchKernelGGL((k_meads_record<decltype(kk)::value>), dim3(sp.C), dim3(64), 0, stream, m, r, fold_of, sp,
                       it, skip_fold);
    return hipGetLastError();
  });
}

hipError_t launch_meads_init(const MeadsState& m, int P, int C, uint32_t itc, hipStream_t stream) {
  const long long n = (long long)C * P;
  hipLaunchKernelGGL(k_meads_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, m, P, C, itc);
  return hipGetLastError();
}

}  // namespace ut
