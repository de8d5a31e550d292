// ut_glm.hip -- fused logistic-GLM log density and gradient over all chains on the
// gfx950 matrix cores (N rows, D = 256 covariates, C chains; the large logistic
// regression of the ChEES workload).
//
// For chain block b (128 chains) and row split z, one workgroup of four wavefronts
// walks its rows in blocks of 32; wavefront w owns chains 32w .. 32w + 31 of the
// block and, per row block,
//   stage 1  eta[32 x 32]  = X[32 x 256] * B[256 x 32]          (MFMA, K = 256)
//   epilogue r = y - logistic(eta) -> LDS; log-likelihood partials
//   stage 2  Gt[32 x 256] += R^T[32 x 32] * X[32 x 256]          (MFMA, K = 32)
// so eta and R never leave the chip.  X is the data and is bf16 exactly.  The
// positions B and the residuals R are fp32 values carried as bf16 hi + lo pairs
// (v = hi + lo to ~16 mantissa bits) and both stages multiply both parts, so eta,
// the log density and the gradient keep fp32-level accuracy with fp32
// accumulation.  Partials per split are reduced by k_glm_reduce, which adds the
// normal prior.
//
// Work mapping is XCD-aware: workgroup g handles split g % S and chain block g / S;
// workgroups are dispatched round-robin over the 8 XCDs, so with S a multiple of 8
// every XCD streams the same rows of X through its own L2 for all its chain blocks.
#include "ut_glm.h"
#include "ut_math.h"
#include "ut_mfma.h"

namespace ut {

namespace {
constexpr int GD = 256;          // covariates
constexpr int CB = 128;          // chains per workgroup
constexpr int RB = 32;           // rows per block
constexpr int BSTR = GD + 8;     // LDS row stride of the position tiles (bf16)
constexpr int RSTR = RB + 8;     // LDS row stride of the residual tiles (bf16)

__device__ inline float fast_logistic(float eta) {
  const float e = __builtin_amdgcn_exp2f(-eta * 1.4426950408889634f);
  return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ inline float fast_log1p_exp(float eta) {
  const float a = fabsf(eta);
  const float e = __builtin_amdgcn_exp2f(-a * 1.4426950408889634f);
  return fmaxf(eta, 0.0f) + __builtin_amdgcn_logf(1.0f + e) * 0.6931471805599453f;
}

typedef uint16_t bf16x4_t __attribute__((ext_vector_type(4)));

template <bool NEED_LJ>
__global__ __launch_bounds__(256) void k_glm_grad(const bf16_t* __restrict__ X, const bf16_t* __restrict__ Xt,
                                                  const float* __restrict__ y, const bf16_t* __restrict__ Bhi,
                                                  const bf16_t* __restrict__ Blo, int N, int C, int S,
                                                  int rows_per_split, float* __restrict__ Gpart,
                                                  double* __restrict__ LLpart) {
  __shared__ __attribute__((aligned(16))) bf16_t Bs_hi[CB * BSTR];
  __shared__ __attribute__((aligned(16))) bf16_t Bs_lo[CB * BSTR];
  __shared__ __attribute__((aligned(16))) bf16_t Rs_hi[CB * RSTR];
  __shared__ __attribute__((aligned(16))) bf16_t Rs_lo[CB * RSTR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, half = lane >> 5;
  const int split = blockIdx.x % S;
  const int cb = blockIdx.x / S;
  const long long c0 = (long long)cb * CB;
  const long long r_begin = (long long)split * rows_per_split;
  const long long r_end = min((long long)N, r_begin + rows_per_split);

  // positions of this chain block -> LDS (128 x 256 bf16, hi and lo)
  for (int c = tid; c < CB * (GD / 8); c += 256) {
    const int row = c / (GD / 8), col = (c % (GD / 8)) * 8;
    *(uint4*)(Bs_hi + row * BSTR + col) = *(const uint4*)(Bhi + (c0 + row) * GD + col);
    *(uint4*)(Bs_lo + row * BSTR + col) = *(const uint4*)(Blo + (c0 + row) * GD + col);
  }
  __syncthreads();

  const int cw = 32 * wave;  // this wavefront's chains within the block
  const bf16_t* bh_row = Bs_hi + (cw + l32) * BSTR + 8 * half;
  const bf16_t* bl_row = Bs_lo + (cw + l32) * BSTR + 8 * half;
  bf16_t* rh_row = Rs_hi + (cw + l32) * RSTR;
  bf16_t* rl_row = Rs_lo + (cw + l32) * RSTR;

  f32x16 g2[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) g2[j][r] = 0.0f;
  double ll = 0.0;

  // operand fragments of the current row block, loaded one block ahead
  bf16x8 xa[GD / 16];     // stage 1 A: X[rows][covariates]
  bf16x8 xb[8][RB / 16];  // stage 2 B: Xt[covariates][rows]
  auto load_x = [&](long long rb0) {
    const bf16_t* xrow = X + (rb0 + l32) * GD + 8 * half;
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) xa[t] = *(const bf16x8*)(xrow + 16 * t);
  };
  auto load_xt = [&](long long rb0) {
#pragma unroll
    for (int jn = 0; jn < 8; ++jn)
#pragma unroll
      for (int t = 0; t < RB / 16; ++t)
        xb[jn][t] = *(const bf16x8*)(Xt + (long long)(32 * jn + l32) * N + rb0 + 16 * t + 8 * half);
  };
  if (r_begin < r_end) {
    load_x(r_begin);
    load_xt(r_begin);
  }

  for (long long rb0 = r_begin; rb0 < r_end; rb0 += RB) {
    const bool more = rb0 + RB < r_end;
    // ---- stage 1: eta tile (rows x chains of this wavefront)
    f32x16 e1;
#pragma unroll
    for (int r = 0; r < 16; ++r) e1[r] = 0.0f;
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) {
      const bf16x8 bh = *(const bf16x8*)(bh_row + 16 * t);
      const bf16x8 bl = *(const bf16x8*)(bl_row + 16 * t);
      e1 = mfma32(xa[t], bh, e1);
      e1 = mfma32(xa[t], bl, e1);
    }
    if (more) load_x(rb0 + RB);
    // ---- epilogue: residual hi/lo to LDS (four consecutive rows per store),
    // log-likelihood partials
    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      bf16x4_t vh, vl;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * q + i;
        const int row = mfma32_row(lane, r);
        const float eta = e1[r];
        const float yv = y[rb0 + row];
        const float res = yv - fast_logistic(eta);
        const bf16_t h = f2bf(res);
        vh[i] = h;
        vl[i] = f2bf(res - bf2f(h));
        if (NEED_LJ) lsum += yv * eta - fast_log1p_exp(eta);
      }
      const int row0 = 8 * q + 4 * half;
      *(bf16x4_t*)(rh_row + row0) = vh;
      *(bf16x4_t*)(rl_row + row0) = vl;
    }
    if (NEED_LJ) ll += (double)lsum;
    __syncthreads();
    // ---- stage 2: Gt[32 chains x 256] += R^T X
#pragma unroll
    for (int t = 0; t < RB / 16; ++t) {
      const bf16x8 ah = *(const bf16x8*)(rh_row + 16 * t + 8 * half);
      const bf16x8 al = *(const bf16x8*)(rl_row + 16 * t + 8 * half);
#pragma unroll
      for (int jn = 0; jn < 8; ++jn) {
        g2[jn] = mfma32(ah, xb[jn][t], g2[jn]);
        g2[jn] = mfma32(al, xb[jn][t], g2[jn]);
      }
    }
    if (more) load_xt(rb0 + RB);
    __syncthreads();
  }
  float* G = Gpart + (long long)split * C * GD;
#pragma unroll
  for (int jn = 0; jn < 8; ++jn)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long long c = c0 + cw + mfma32_row(lane, r);
      G[c * GD + 32 * jn + l32] = g2[jn][r];
    }
  if (NEED_LJ) {
    // lanes l and l + 32 hold the same chain (different rows)
    const double other = __shfl_xor(ll, 32);
    if (lane < 32) LLpart[(long long)split * C + c0 + cw + lane] = ll + other;
  }
}

// positions (fp32, C x D) -> bf16 hi and lo parts
__global__ void k_split_bf16(const float* __restrict__ q, bf16_t* __restrict__ hi, bf16_t* __restrict__ lo,
                             long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = q[i];
  const bf16_t h = f2bf(v);
  hi[i] = h;
  lo[i] = f2bf(v - bf2f(h));
}

// gradient = sum over splits (fixed order) + prior; logjoint likewise in binary64
// (one wavefront per chain, four covariates per lane)
__global__ __launch_bounds__(64) void k_glm_reduce(const float* __restrict__ q, const float* __restrict__ Gpart,
                                                  const double* __restrict__ LLpart, int C, int S, float loc,
                                                  float inv_scale, double total_const, int need_lj,
                                                  float* __restrict__ grad, float* __restrict__ lj) {
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  double lp = 0.0;
#pragma unroll
  for (int k = 0; k < GD / 64; ++k) {
    const int d = 64 * k + lane;
    float s = 0.0f;
    for (int z = 0; z < S; ++z) s += Gpart[((long long)z * C + c) * GD + d];
    const float zq = (q[c * GD + d] - loc) * inv_scale;
    grad[c * GD + d] = s - zq * inv_scale;
    lp -= 0.5 * (double)zq * (double)zq;
  }
  if (need_lj) {
    for (int m = 32; m >= 1; m >>= 1) lp += __shfl_xor(lp, m);
    if (lane == 0) {
      double s = 0.0;
      for (int z = 0; z < S; ++z) s += LLpart[(long long)z * C + c];
      lj[c] = (float)((s + lp) + total_const);
    }
  }
}
}  // namespace

int glm_splits(int N, int C) {
  // enough workgroups to cover the 256 CUs, a multiple of 8 splits (one per XCD
  // when C is large), at least one row block per split
  const int blocks = C / CB;
  int S = (256 + blocks - 1) / blocks;
  S = ((S + 7) / 8) * 8;
  const int max_s = N / RB;
  if (S > max_s) S = max_s;
  return S < 1 ? 1 : S;
}

hipError_t launch_glm_grad(const GlmPlan& p, const float* q, float* grad, float* lj, int C, bool need_lj,
                           GlmScratch& s, hipStream_t stream) {
  if (C % CB != 0 || p.D != GD || p.N % RB != 0 || s.splits < 1) return hipErrorInvalidValue;
  const long long n = (long long)C * GD;
  hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, q, s.bhi, s.blo, n);
  const int S = s.splits;
  const int rows = ((p.N / S + RB - 1) / RB) * RB;
  const dim3 grid((unsigned)(S * (C / CB)));
  if (need_lj)
    hipLaunchKernelGGL(k_glm_grad<true>, grid, dim3(256), 0, stream, p.X, p.Xt, p.y, s.bhi, s.blo, p.N, C, S, rows,
                       s.gpart, s.llpart);
  else
    hipLaunchKernelGGL(k_glm_grad<false>, grid, dim3(256), 0, stream, p.X, p.Xt, p.y, s.bhi, s.blo, p.N, C, S,
                       rows, s.gpart, s.llpart);
  hipLaunchKernelGGL(k_glm_reduce, dim3(C), dim3(64), 0, stream, q, s.gpart, s.llpart, C, S, (float)p.loc,
                     (float)(1.0 / p.scale), p.total_const, need_lj ? 1 : 0, grad, lj);
  return hipGetLastError();
}

}  // namespace ut
