// ut_gp.hip -- latent-GP Poisson log density and gradient for all chains, with a
// batched binary64 Cholesky factorisation that carries the tangents of the kernel
// hyperparameters (reference model pieces: gp_cholesky and _gp_kernel_covariance,
// src/distributions/gaussian_process.jl:180-188,283-294; poisson
// src/device/math.jl:230-235; lognormal :93-97; halfnormal :120-124; log transform
// :1144-1159).
//
// Per chain the dual matrix (K, dK/dlog l_1, .., dK/dlog a) is factorised in place,
// right-looking, in 32 x 32 tiles, one launch per phase so the grid itself orders
// the phases:
//   k_gp_diag    factor the diagonal tile in LDS (one workgroup per chain)
//   k_gp_panel   L_ib = A_ib D^-T for every tile row below (one workgroup per tile)
//   k_gp_update  A_ib,jb -= L_ib L_jb^T for the trailing lower triangle of tiles
// with the tangent rules  dL_jj = dA_jj / (2 L_jj),  dL_ij = (dA_ij - L_ij dL_jj) / L_jj,
// and d(A - L M^T) = dA - dL M^T - L dM^T.  Then k_gp_finish forms f = L z and
// df = dL z, the Poisson terms, grad_z = L^T (y - e^f) - z, grad_hyper = r . df
// plus the priors, and the log-determinant 2 sum log L_ii.
#include <math.h>

#include "ut_gp.h"
#include "ut_math.h"

namespace ut {

namespace {

constexpr int TS = GP_TILE;
constexpr int TSP = TS + 1;  // LDS row stride (doubles)
constexpr int NC = GP_MAX_T + 1;

__device__ inline double* mat(double* A, int c, int comp, int T, int N) {
  return A + ((size_t)c * (T + 1) + comp) * (size_t)N * N;
}

// dual kernel matrix: value and tangents along each log lengthscale and log amplitude
__global__ void k_gp_build(GpPlan p, const float* __restrict__ q, int stride, const double* __restrict__ hyper_d,
                           double* __restrict__ A) {
  const int c = blockIdx.y;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = p.N, D = p.D, T = D + 1;
  if (idx >= (long long)N * N) return;
  const int i = (int)(idx / N), j = (int)(idx % N);
  double h[GP_MAX_T];
  for (int k = 0; k < T; ++k) h[k] = hyper_d ? hyper_d[(size_t)c * stride + k] : (double)q[(size_t)c * stride + k];
  double val = 0.0, tan[GP_MAX_T] = {0.0, 0.0, 0.0};
  if (j <= i) {
    double s = 0.0, terms[GP_MAX_T - 1] = {0.0, 0.0};
    for (int d = 0; d < D; ++d) {
      const double inv_l = exp(-h[d]);
      const double diff = (p.X[(size_t)i * D + d] - p.X[(size_t)j * D + d]) * inv_l;
      terms[d] = diff * diff;
      s += terms[d];
    }
    const double k = exp(2.0 * h[D]) * exp(-0.5 * s);
    for (int d = 0; d < D; ++d) tan[d] = k * terms[d];
    tan[D] = 2.0 * k;
    val = k + (i == j ? p.noise2 : 0.0);
  }
  mat(A, c, 0, T, N)[idx] = val;
  for (int t = 0; t < T; ++t) mat(A, c, t + 1, T, N)[idx] = tan[t];
}

__device__ void load_tile(double (*dst)[TS][TSP], double* A, int c, int T, int N, int ti, int tj) {
  for (int e = threadIdx.x; e < (T + 1) * TS * TS; e += blockDim.x) {
    const int comp = e / (TS * TS), r = (e / TS) % TS, cc = e % TS;
    dst[comp][r][cc] = mat(A, c, comp, T, N)[(size_t)(ti * TS + r) * N + tj * TS + cc];
  }
}
__device__ void store_tile(double (*src)[TS][TSP], double* A, int c, int T, int N, int ti, int tj) {
  for (int e = threadIdx.x; e < (T + 1) * TS * TS; e += blockDim.x) {
    const int comp = e / (TS * TS), r = (e / TS) % TS, cc = e % TS;
    mat(A, c, comp, T, N)[(size_t)(ti * TS + r) * N + tj * TS + cc] = src[comp][r][cc];
  }
}

__global__ __launch_bounds__(256) void k_gp_diag(double* __restrict__ A, int N, int T, int kb) {
  __shared__ double t[NC][TS][TSP];
  const int c = blockIdx.x;
  load_tile(t, A, c, T, N, kb, kb);
  __syncthreads();
  for (int j = 0; j < TS; ++j) {
    if (threadIdx.x == 0) {
      const double l = sqrt(t[0][j][j]);
      t[0][j][j] = l;
      for (int k = 1; k <= T; ++k) t[k][j][j] = t[k][j][j] / (2.0 * l);
    }
    __syncthreads();
    const double l = t[0][j][j];
    for (int i = j + 1 + threadIdx.x; i < TS; i += blockDim.x) {
      const double v = t[0][i][j] / l;
      for (int k = 1; k <= T; ++k) t[k][i][j] = (t[k][i][j] - v * t[k][j][j]) / l;
      t[0][i][j] = v;
    }
    __syncthreads();
    const int m = TS - 1 - j;  // trailing lower triangle: pairs j < kk <= i < TS
    for (int e = threadIdx.x; e < m * (m + 1) / 2; e += blockDim.x) {
      const int row = (int)((sqrt(8.0 * e + 1.0) - 1.0) / 2.0);
      const int col = e - row * (row + 1) / 2;
      const int i = j + 1 + row, kk = j + 1 + col;
      const double lij = t[0][i][j], lkj = t[0][kk][j];
      t[0][i][kk] -= lij * lkj;
      for (int k = 1; k <= T; ++k) t[k][i][kk] -= t[k][i][j] * lkj + lij * t[k][kk][j];
    }
    __syncthreads();
  }
  store_tile(t, A, c, T, N, kb, kb);
}

// one workgroup per (chain, tile row below the diagonal): row r of the tile is
// solved by thread r against the factored diagonal tile
__global__ __launch_bounds__(64) void k_gp_panel(double* __restrict__ A, int N, int T, int kb) {
  __shared__ double d[NC][TS][TSP];
  __shared__ double x[NC][TS][TSP];
  const int c = blockIdx.x;
  const int ib = kb + 1 + blockIdx.y;
  load_tile(d, A, c, T, N, kb, kb);
  load_tile(x, A, c, T, N, ib, kb);
  __syncthreads();
  const int r = threadIdx.x;
  if (r < TS) {
    for (int jj = 0; jj < TS; ++jj) {
      double v = x[0][r][jj];
      double dv[GP_MAX_T];
      for (int k = 1; k <= T; ++k) dv[k - 1] = x[k][r][jj];
      for (int m = 0; m < jj; ++m) {
        const double lrm = x[0][r][m], djm = d[0][jj][m];
        v -= lrm * djm;
        for (int k = 1; k <= T; ++k) dv[k - 1] -= x[k][r][m] * djm + lrm * d[k][jj][m];
      }
      const double djj = d[0][jj][jj];
      const double lv = v / djj;
      x[0][r][jj] = lv;
      for (int k = 1; k <= T; ++k) x[k][r][jj] = (dv[k - 1] - lv * d[k][jj][jj]) / djj;
    }
  }
  __syncthreads();
  store_tile(x, A, c, T, N, ib, kb);
}

// one workgroup per (chain, trailing tile pair ib >= jb > kb)
__global__ __launch_bounds__(256) void k_gp_update(double* __restrict__ A, int N, int T, int kb) {
  __shared__ double li[NC][TS][TSP];
  __shared__ double lj[NC][TS][TSP];
  const int c = blockIdx.x;
  const int e = blockIdx.y;
  const int row = (int)((sqrt(8.0 * e + 1.0) - 1.0) / 2.0);
  const int col = e - row * (row + 1) / 2;
  const int ib = kb + 1 + row, jb = kb + 1 + col;
  load_tile(li, A, c, T, N, ib, kb);
  load_tile(lj, A, c, T, N, jb, kb);
  __syncthreads();
  const int r = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 4;
  double acc[4][NC];
  for (int u = 0; u < 4; ++u)
    for (int k = 0; k < NC; ++k) acc[u][k] = 0.0;
  for (int m = 0; m < TS; ++m) {
    const double a0 = li[0][r][m];
    double ak[GP_MAX_T];
    for (int k = 1; k <= T; ++k) ak[k - 1] = li[k][r][m];
    for (int u = 0; u < 4; ++u) {
      const double b0 = lj[0][c0 + u][m];
      acc[u][0] += a0 * b0;
      for (int k = 1; k <= T; ++k) acc[u][k] += ak[k - 1] * b0 + a0 * lj[k][c0 + u][m];
    }
  }
  for (int k = 0; k <= T; ++k) {
    double* M = mat(A, c, k, T, N) + (size_t)(ib * TS + r) * N + jb * TS + c0;
    for (int u = 0; u < 4; ++u) M[u] -= acc[u][k];
  }
}

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

__global__ __launch_bounds__(256) void k_gp_finish(GpPlan p, const float* __restrict__ q, double* __restrict__ A,
                                                   double* __restrict__ logdet, float* __restrict__ grad,
                                                   float* __restrict__ lj, int P) {
  extern __shared__ double sh[];
  const int N = p.N, D = p.D, T = D + 1;
  const int c = blockIdx.x;
  double* z = sh;          // N
  double* r = sh + N;      // N
  double* red = sh + 2 * N;  // 256
  const float* qc = q + (size_t)c * P;
  for (int i = threadIdx.x; i < N; i += blockDim.x) z[i] = (double)qc[D + 1 + i];
  __syncthreads();
  const double* L = mat(A, c, 0, T, N);
  double ll = 0.0, zz = 0.0, ld = 0.0, gh[GP_MAX_T] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double f = 0.0, df[GP_MAX_T] = {0.0, 0.0, 0.0};
    const double* Li = L + (size_t)i * N;
    for (int j = 0; j <= i; ++j) f += Li[j] * z[j];
    for (int k = 0; k < T; ++k) {
      const double* dLi = mat(A, c, k + 1, T, N) + (size_t)i * N;
      for (int j = 0; j <= i; ++j) df[k] += dLi[j] * z[j];
    }
    const double ef = exp(f);
    const double ri = p.y[i] - ef;
    r[i] = ri;
    ll += p.y[i] * f - ef;
    for (int k = 0; k < T; ++k) gh[k] += ri * df[k];
    zz += z[i] * z[i];
    ld += log(Li[i]);
  }
  __syncthreads();
  // grad_z = L^T r - z (column access: consecutive threads read consecutive j)
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    double s = 0.0;
    for (int i = j; i < N; ++i) s += L[(size_t)i * N + j] * r[i];
    grad[(size_t)c * P + D + 1 + j] = (float)(s - z[j]);
  }
  ll = block_sum(ll, red);
  zz = block_sum(zz, red);
  ld = block_sum(ld, red);
  for (int k = 0; k < T; ++k) gh[k] = block_sum(gh[k], red);
  if (threadIdx.x == 0) {
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
    if (logdet) logdet[c] = 2.0 * ld;
  }
}

}  // namespace

hipError_t launch_gp_factor(const GpPlan& p, const float* q, int stride, const double* hyper_d, int C,
                            GpScratch& s, hipStream_t stream) {
  const int N = p.N, T = p.D + 1;
  if (N % TS != 0 || p.D < 1 || T > GP_MAX_T) return hipErrorInvalidValue;
  const long long nn = (long long)N * N;
  hipLaunchKernelGGL(k_gp_build, dim3((unsigned)((nn + 255) / 256), C), dim3(256), 0, stream, p, q, stride, hyper_d,
                     s.A);
  const int nbk = N / TS;
  for (int kb = 0; kb < nbk; ++kb) {
    hipLaunchKernelGGL(k_gp_diag, dim3(C), dim3(256), 0, stream, s.A, N, T, kb);
    const int m = nbk - kb - 1;
    if (m == 0) continue;
    hipLaunchKernelGGL(k_gp_panel, dim3(C, m), dim3(64), 0, stream, s.A, N, T, kb);
    hipLaunchKernelGGL(k_gp_update, dim3(C, m * (m + 1) / 2), dim3(256), 0, stream, s.A, N, T, kb);
  }
  return hipGetLastError();
}

hipError_t launch_gp_grad(const GpPlan& p, const float* q, float* grad, float* lj, int C, GpScratch& s,
                          hipStream_t stream) {
  const int P = p.D + 1 + p.N;
  hipError_t e = launch_gp_factor(p, q, P, nullptr, C, s, stream);
  if (e != hipSuccess) return e;
  const size_t shmem = (size_t)(2 * p.N + 256) * sizeof(double);
  hipLaunchKernelGGL(k_gp_finish, dim3(C), dim3(256), shmem, stream, p, q, s.A, s.logdet, grad, lj, P);
  return hipGetLastError();
}

}  // namespace ut
