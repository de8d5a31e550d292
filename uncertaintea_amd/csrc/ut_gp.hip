// ut_gp.hip -- latent-GP Poisson log density and gradient for all chains, with a
// batched binary64 Cholesky factorisation that carries the tangents of the kernel
// hyperparameters (reference model pieces: gp_cholesky and _gp_kernel_covariance,
// src/distributions/gaussian_process.jl:180-188,283-294; poisson
// src/device/math.jl:230-235; lognormal :93-97; halfnormal :120-124; log transform
// :1144-1159).
//
// Per chain the dual matrix (K, dK/dlog l_1, .., dK/dlog a) is factorised
// left-looking, one 32-wide tile column k at a time, two launches per column:
//   k_gp_column  A_ik = K_ik - sum_{j<k} L_ij L_kj^T for every tile i >= k, with K
//                built on the fly from the inputs; the diagonal tile's workgroup then
//                factors it in LDS
//   k_gp_panel   L_ik = A_ik L_kk^-T for the tiles below (one row per 8 lanes, the
//                substitution runs in registers)
// with the tangent rules  dL_jj = dA_jj / (2 L_jj),  dL_ij = (dA_ij - L_ij dL_jj) / L_jj,
// and d(A - L M^T) = dA - dL M^T - L dM^T.  Each finished tile also leaves its
// share of f = L z and df = dL z.  k_gp_resid sums those shares in tile order and forms
// the Poisson terms, grad_hyper = r . df plus the priors and the log-determinant;
// k_gp_ltr forms grad_z = L^T r - z.  Every sum has a fixed order.
#include <math.h>

#include "ut_gp.h"
#include "ut_math.h"

namespace ut {

namespace {

constexpr int TS = GP_TILE;
constexpr int TT = TS * TS;
constexpr int TSP = TS + 1;  // LDS column stride (doubles)
constexpr int MC = 16;       // inner columns per staged chunk

__device__ inline size_t tri(int i) { return (size_t)i * (i + 1) / 2; }

// linear block index -> work item, so that consecutive items (the tiles of one chain)
// run on one XCD and share its L2; G is a multiple of 8
__device__ inline int xcd_item(int G) {
  const int b = blockIdx.x;
  return (b & 7) * (G >> 3) + (b >> 3);
}

inline int round8(long long n) { return (int)((n + 7) / 8 * 8); }

template <int T>
__device__ inline double* tile_ptr(double* A, int ntiles, int c, int comp, int i, int j) {
  return A + (((size_t)c * (T + 1) + comp) * ntiles + tri(i) + j) * TT;
}

#define DS(comp, col, row) Ds[((comp) * TS + (col)) * TSP + (row)]

template <int T>
__global__ __launch_bounds__(256) void k_gp_column(GpPlan p, const float* __restrict__ q, int stride,
                                                   const double* __restrict__ hyper_d, double* __restrict__ A,
                                                   double* __restrict__ pf, double* __restrict__ ldp, int k, int C,
                                                   int G) {
  constexpr int D = T - 1;
  constexpr int SH = (T + 1) * TS * TSP > 2 * (T + 1) * MC * TS ? (T + 1) * TS * TSP : 2 * (T + 1) * MC * TS;
  __shared__ double sh[SH];
  const int N = p.N, nb = N / TS, nt = nb - k;
  const int w = xcd_item(G);
  if (w >= C * nt) return;
  const int c = w / nt, i = k + w % nt;
  const int ntiles = nb * (nb + 1) / 2;
  const int t = threadIdx.x;
  const int tr = t & 15, tc = t >> 4;  // rows 2tr, 2tr+1 and columns 2tc, 2tc+1 of the tile
  double h[T];
  for (int kk = 0; kk < T; ++kk)
    h[kk] = hyper_d ? hyper_d[(size_t)c * stride + kk] : (double)q[(size_t)c * stride + kk];
  double inv_l[D];
  for (int d = 0; d < D; ++d) inv_l[d] = exp(-h[d]);
  const double a2 = exp(2.0 * h[D]);
  double acc[2][2][T + 1];
  for (int u = 0; u < 2; ++u)
    for (int v = 0; v < 2; ++v) {
      const int gi = i * TS + 2 * tr + u, gj = k * TS + 2 * tc + v;
      double s = 0.0, terms[D];
      for (int d = 0; d < D; ++d) {
        const double diff = (p.X[(size_t)gi * D + d] - p.X[(size_t)gj * D + d]) * inv_l[d];
        terms[d] = diff * diff;
        s += terms[d];
      }
      const double kv = a2 * exp(-0.5 * s);
      acc[u][v][0] = kv + (gi == gj ? p.noise2 : 0.0);
      for (int d = 0; d < D; ++d) acc[u][v][1 + d] = kv * terms[d];
      acc[u][v][T] = 2.0 * kv;
    }
  double* As = sh;                      // [T+1][MC][TS]
  double* Bs = sh + (T + 1) * MC * TS;  // [T+1][MC][TS]
  for (int j = 0; j < k; ++j) {
    for (int m0 = 0; m0 < TS; m0 += MC) {
      __syncthreads();
      // columns m0 .. m0+MC-1 of tiles (i, j) and (k, j): contiguous in the tile layout
#pragma unroll
      for (int comp = 0; comp <= T; ++comp) {
        const double2* ga = (const double2*)(tile_ptr<T>(A, ntiles, c, comp, i, j) + m0 * TS);
        const double2* gb = (const double2*)(tile_ptr<T>(A, ntiles, c, comp, k, j) + m0 * TS);
        ((double2*)(As + comp * MC * TS))[t] = ga[t];
        ((double2*)(Bs + comp * MC * TS))[t] = gb[t];
      }
      __syncthreads();
#pragma unroll 4
      for (int m = 0; m < MC; ++m) {
        double2 a[T + 1], b[T + 1];
#pragma unroll
        for (int comp = 0; comp <= T; ++comp) {
          a[comp] = ((const double2*)(As + (comp * MC + m) * TS))[tr];
          b[comp] = ((const double2*)(Bs + (comp * MC + m) * TS))[tc];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            const double a0 = u ? a[0].y : a[0].x, b0 = v ? b[0].y : b[0].x;
            acc[u][v][0] = fma(-a0, b0, acc[u][v][0]);
#pragma unroll
            for (int comp = 1; comp <= T; ++comp)
              acc[u][v][comp] = fma(-a0, v ? b[comp].y : b[comp].x, fma(-(u ? a[comp].y : a[comp].x), b0, acc[u][v][comp]));
          }
      }
    }
  }
  if (i > k) {
#pragma unroll
    for (int comp = 0; comp <= T; ++comp) {
      double* tp = tile_ptr<T>(A, ntiles, c, comp, i, k);
      for (int v = 0; v < 2; ++v) {
        double2 o;
        o.x = acc[0][v][comp];
        o.y = acc[1][v][comp];
        ((double2*)(tp + (2 * tc + v) * TS))[tr] = o;
      }
    }
    return;
  }
  // diagonal tile: factor in LDS
  __syncthreads();
  double* Ds = sh;  // [T+1][col][row], column stride TSP
  for (int comp = 0; comp <= T; ++comp)
    for (int u = 0; u < 2; ++u)
      for (int v = 0; v < 2; ++v) DS(comp, 2 * tc + v, 2 * tr + u) = acc[u][v][comp];
  __syncthreads();
  for (int j = 0; j < TS; ++j) {
    if (t == 0) {
      const double l = sqrt(DS(0, j, j));
      DS(0, j, j) = l;
      for (int comp = 1; comp <= T; ++comp) DS(comp, j, j) = DS(comp, j, j) / (2.0 * l);
    }
    __syncthreads();
    const double l = DS(0, j, j);
    for (int r = j + 1 + t; r < TS; r += blockDim.x) {
      const double v = DS(0, j, r) / l;
      for (int comp = 1; comp <= T; ++comp) DS(comp, j, r) = (DS(comp, j, r) - v * DS(comp, j, j)) / l;
      DS(0, j, r) = v;
    }
    __syncthreads();
    const int m = TS - 1 - j;  // trailing lower triangle: pairs j < kk <= r < TS
    for (int e = t; e < m * (m + 1) / 2; e += blockDim.x) {
      const int row = (int)((sqrt(8.0 * e + 1.0) - 1.0) / 2.0);
      const int col = e - row * (row + 1) / 2;
      const int r = j + 1 + row, kk = j + 1 + col;
      const double lrj = DS(0, j, r), lkj = DS(0, j, kk);
      DS(0, kk, r) -= lrj * lkj;
      for (int comp = 1; comp <= T; ++comp) DS(comp, kk, r) -= DS(comp, j, r) * lkj + lrj * DS(comp, j, kk);
    }
    __syncthreads();
  }
  for (int e = t; e < (T + 1) * TT; e += blockDim.x) {
    const int comp = e / TT, col = (e / TS) % TS, row = e % TS;
    tile_ptr<T>(A, ntiles, c, comp, k, k)[col * TS + row] = row >= col ? DS(comp, col, row) : 0.0;
  }
  if (t == 0) {
    double s = 0.0;
    for (int r = 0; r < TS; ++r) s += log(DS(0, r, r));
    ldp[(size_t)c * nb + k] = s;
  }
  if (q && t < (T + 1) * TS) {
    const int comp = t / TS, r = t % TS;
    const float* z = q + (size_t)c * stride + T + k * TS;
    double s = 0.0;
    for (int cc = 0; cc <= r; ++cc) s += DS(comp, cc, r) * (double)z[cc];
    pf[((size_t)c * nb + k) * (T + 1) * N + (size_t)comp * N + k * TS + r] = s;
  }
}

template <int T>
__global__ __launch_bounds__(256) void k_gp_panel(
    const float* __restrict__ q, int stride, double* __restrict__ A, double* __restrict__ pf, int N, int k, int C,
    int G) {
  __shared__ double Ds[(T + 1) * TS * TSP];
  __shared__ double inv_d[TS];
  const int nb = N / TS, nt = nb - k - 1;
  const int w = xcd_item(G);
  if (w >= C * nt) return;
  const int c = w / nt, i = k + 1 + w % nt;
  const int ntiles = nb * (nb + 1) / 2;
  const int t = threadIdx.x;
  for (int e = t; e < (T + 1) * TT; e += blockDim.x) {
    const int comp = e / TT, col = (e / TS) % TS, row = e % TS;
    DS(comp, col, row) = tile_ptr<T>(A, ntiles, c, comp, k, k)[e % TT];
  }
  __syncthreads();
  if (t < TS) inv_d[t] = 1.0 / DS(0, t, t);
  const int r = t >> 3, g = t & 7;  // row r of the tile; columns g, g+8, g+16, g+24
  double x[4][T + 1];
#pragma unroll
  for (int comp = 0; comp <= T; ++comp) {
    const double* tp = tile_ptr<T>(A, ntiles, c, comp, i, k);
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s][comp] = tp[(g + 8 * s) * TS + r];
  }
  __syncthreads();
  // row r solves  L_r D^T = A_r  column by column; column jj lives in lane g = jj % 8
#pragma unroll
  for (int s0 = 0; s0 < 4; ++s0)
#pragma unroll 2
  for (int g0 = 0; g0 < 8; ++g0) {
    const int jj = 8 * s0 + g0;
    const double d = inv_d[jj];
    double v[T + 1];
    v[0] = x[s0][0] * d;
#pragma unroll
    for (int comp = 1; comp <= T; ++comp) v[comp] = fma(-v[0], DS(comp, jj, jj), x[s0][comp]) * d;
#pragma unroll
    for (int comp = 0; comp <= T; ++comp) v[comp] = __shfl(v[comp], g0, 8);
    if (g == g0) {
#pragma unroll
      for (int comp = 0; comp <= T; ++comp) x[s0][comp] = v[comp];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int mm = g + 8 * s;
      if (mm > jj) {
        const double dm = DS(0, jj, mm);
        x[s][0] = fma(-v[0], dm, x[s][0]);
#pragma unroll
        for (int comp = 1; comp <= T; ++comp) x[s][comp] = fma(-v[0], DS(comp, jj, mm), fma(-v[comp], dm, x[s][comp]));
      }
    }
  }
#pragma unroll
  for (int comp = 0; comp <= T; ++comp) {
    double* tp = tile_ptr<T>(A, ntiles, c, comp, i, k);
#pragma unroll
    for (int s = 0; s < 4; ++s) tp[(g + 8 * s) * TS + r] = x[s][comp];
  }
  if (q) {
    const float* z = q + (size_t)c * stride + T + k * TS;
    double zv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) zv[s] = (double)z[g + 8 * s];
#pragma unroll
    for (int comp = 0; comp <= T; ++comp) {
      double s = 0.0;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) s += x[s4][comp] * zv[s4];
      s += __shfl_xor(s, 1, 8);
      s += __shfl_xor(s, 2, 8);
      s += __shfl_xor(s, 4, 8);
      if (g == 0) pf[((size_t)c * nb + k) * (T + 1) * N + (size_t)comp * N + i * TS + r] = s;
    }
  }
}
#undef DS

// deterministic tree sum of v over the workgroup (256 threads) through LDS
__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <int T>
__global__ __launch_bounds__(256) void k_gp_resid(GpPlan p, const float* __restrict__ q,
                                                  const double* __restrict__ pf, const double* __restrict__ ldp,
                                                  double* __restrict__ resid, double* __restrict__ logdet,
                                                  float* __restrict__ grad, float* __restrict__ lj, int P) {
  constexpr int D = T - 1;
  __shared__ double red[256];
  const int N = p.N, nb = N / TS;
  const int c = blockIdx.x;
  const float* qc = q + (size_t)c * P;
  double ll = 0.0, zz = 0.0, gh[T];
  for (int kk = 0; kk < T; ++kk) gh[kk] = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int ib = i / TS;
    double f = 0.0, df[T];
    for (int kk = 0; kk < T; ++kk) df[kk] = 0.0;
    for (int kb = 0; kb <= ib; ++kb) {
      const double* pp = pf + ((size_t)c * nb + kb) * (T + 1) * N + i;
      f += pp[0];
      for (int kk = 0; kk < T; ++kk) df[kk] += pp[(size_t)(kk + 1) * N];
    }
    const double ef = exp(f);
    const double ri = p.y[i] - ef;
    resid[(size_t)c * N + i] = ri;
    ll += p.y[i] * f - ef;
    for (int kk = 0; kk < T; ++kk) gh[kk] += ri * df[kk];
    const double zi = (double)qc[T + i];
    zz += zi * zi;
  }
  ll = block_sum(ll, red);
  zz = block_sum(zz, red);
  for (int kk = 0; kk < T; ++kk) gh[kk] = block_sum(gh[kk], red);
  if (threadIdx.x == 0) {
    double ld = 0.0;
    for (int kb = 0; kb < nb; ++kb) ld += ldp[(size_t)c * nb + kb];
    // priors: log l_d ~ normal(ls_loc, ls_scale); a ~ halfnormal(amp_scale) with the
    // log transform's log-Jacobian u
    double lp = 0.0;
    const double inv_ls = 1.0 / p.ls_scale;
    for (int d = 0; d < D; ++d) {
      const double u = (double)qc[d];
      const double w = (u - p.ls_loc) * inv_ls;
      lp -= 0.5 * w * w;
      grad[(size_t)c * P + d] = (float)(gh[d] - w * inv_ls);
    }
    const double ua = (double)qc[D];
    const double a = exp(ua);
    const double wa = a / p.amp_scale;
    lp += -0.5 * wa * wa + ua;
    grad[(size_t)c * P + D] = (float)(gh[D] - wa * wa + 1.0);
    lj[c] = (float)(ll + lp - 0.5 * zz + p.total_const);
    logdet[c] = 2.0 * ld;
  }
}

// grad_z = L^T r - z, one workgroup per (chain, tile column jb): lane group g of
// column col reads rows 4g .. 4g+3 of every tile below
template <int T>
__global__ __launch_bounds__(256) void k_gp_ltr(const float* __restrict__ q, const double* __restrict__ A,
                                                const double* __restrict__ resid, float* __restrict__ grad, int N,
                                                int P, int C, int G) {
  const int nb = N / TS, ntiles = nb * (nb + 1) / 2;
  const int w = xcd_item(G);
  if (w >= C * nb) return;
  const int c = w / nb, jb = w % nb;
  const int t = threadIdx.x, g = t & 7, col = t >> 3;
  double s = 0.0;
  for (int ib = jb; ib < nb; ++ib) {
    const double* tp = A + ((size_t)c * (T + 1) * ntiles + tri(ib) + jb) * TT + col * TS + 4 * g;
    const double* rr = resid + (size_t)c * N + ib * TS + 4 * g;
    const double2 l01 = ((const double2*)tp)[0], l23 = ((const double2*)tp)[1];
    s += l01.x * rr[0] + l01.y * rr[1] + l23.x * rr[2] + l23.y * rr[3];
  }
  s += __shfl_xor(s, 1, 8);
  s += __shfl_xor(s, 2, 8);
  s += __shfl_xor(s, 4, 8);
  if (g == 0) {
    const size_t j = (size_t)c * P + T + jb * TS + col;
    grad[j] = (float)(s - (double)q[j]);
  }
}

template <int T>
hipError_t factor_t(const GpPlan& p, const float* q, int stride, const double* hyper_d, int C, GpScratch& s,
                    hipStream_t stream) {
  const int nb = p.N / TS;
  for (int k = 0; k < nb; ++k) {
    const int nt = nb - k;
    const int G = round8((long long)C * nt);
    hipLaunchKernelGGL(k_gp_column<T>, dim3(G), dim3(256), 0, stream, p, q, stride, hyper_d, s.A, s.pf, s.ldp, k, C,
                       G);
    if (nt > 1) {
      const int G2 = round8((long long)C * (nt - 1));
      hipLaunchKernelGGL(k_gp_panel<T>, dim3(G2), dim3(256), 0, stream, q, stride, s.A, s.pf, p.N, k, C, G2);
    }
  }
  return hipGetLastError();
}

template <int T>
hipError_t grad_t(const GpPlan& p, const float* q, float* grad, float* lj, int C, GpScratch& s, hipStream_t stream) {
  const int P = T + p.N;
  hipError_t e = factor_t<T>(p, q, P, nullptr, C, s, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gp_resid<T>, dim3(C), dim3(256), 0, stream, p, q, s.pf, s.ldp, s.r, s.logdet, grad, lj, P);
  const int G = round8((long long)C * (p.N / TS));
  hipLaunchKernelGGL(k_gp_ltr<T>, dim3(G), dim3(256), 0, stream, q, s.A, s.r, grad, p.N, P, C, G);
  return hipGetLastError();
}

}  // namespace

hipError_t gp_scratch_alloc(GpScratch& s, const GpPlan& p, int C) {
  const int T = p.D + 1, nb = p.N / TS;
  hipError_t e = hipMalloc(&s.A, (size_t)C * (T + 1) * gp_tiles(p.N) * TT * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s.pf, (size_t)C * nb * (T + 1) * p.N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s.ldp, (size_t)C * nb * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s.r, (size_t)C * p.N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s.logdet, (size_t)C * sizeof(double));
  if (e != hipSuccess) gp_scratch_free(s);
  return e;
}

void gp_scratch_free(GpScratch& s) {
  double* ptrs[] = {s.A, s.pf, s.ldp, s.r, s.logdet};
  for (double* x : ptrs)
    if (x) (void)hipFree(x);
  s = GpScratch{};
}

hipError_t launch_gp_factor(const GpPlan& p, const float* q, int stride, const double* hyper_d, int C,
                            GpScratch& s, hipStream_t stream) {
  if (p.N % TS != 0 || p.D < 1 || p.D + 1 > GP_MAX_T || C <= 0) return hipErrorInvalidValue;
  if (!hyper_d && !q) return hipErrorInvalidValue;
  return p.D == 1 ? factor_t<2>(p, q, stride, hyper_d, C, s, stream) : factor_t<3>(p, q, stride, hyper_d, C, s, stream);
}

hipError_t launch_gp_grad(const GpPlan& p, const float* q, float* grad, float* lj, int C, GpScratch& s,
                          hipStream_t stream) {
  if (p.N % TS != 0 || p.D < 1 || p.D + 1 > GP_MAX_T || C <= 0) return hipErrorInvalidValue;
  return p.D == 1 ? grad_t<2>(p, q, grad, lj, C, s, stream) : grad_t<3>(p, q, grad, lj, C, s, stream);
}

}  // namespace ut
