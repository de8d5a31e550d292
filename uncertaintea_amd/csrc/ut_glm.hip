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
// Latency: every wavefront runs alone on its SIMD (the position tiles fill the
// LDS), so the kernel hides global latency itself: the X fragments and outcomes of
// row block k + 1 are requested right after stage 1 of block k and the X^T
// fragments right after its stage 2, the position fragments are read from LDS one
// k-step ahead of the multiply, and the loop has no workgroup barrier (each
// wavefront's residual tile is its own).
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

typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> packed bf16 pair, round to nearest even (one v_cvt_pk_bf16_f32)
__device__ inline uint32_t pack_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2_hw));
}
// hi and lo bf16 pairs of two floats: v = hi + lo to ~16 mantissa bits
__device__ inline void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
  hi = pack_bf16(a, b);
  const float ha = __builtin_bit_cast(float, hi << 16);
  const float hb = __builtin_bit_cast(float, hi & 0xffff0000u);
  lo = pack_bf16(a - ha, b - hb);
}

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
  float4 yv[4];           // outcomes of rows 8 q + 4 half .. + 3 (the epilogue's rows)
  auto load_x = [&](long long rb0) {
    const bf16_t* xrow = X + (rb0 + l32) * GD + 8 * half;
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) xa[t] = *(const bf16x8*)(xrow + 16 * t);
#pragma unroll
    for (int q = 0; q < 4; ++q) yv[q] = *(const float4*)(y + rb0 + 8 * q + 4 * half);
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
    // the prefetch is unconditional (the last block re-reads itself), so the
    // compiler's in-order load counters stay exact across the loop
    const long long next = rb0 + RB < r_end ? rb0 + RB : rb0;
    // ---- stage 1: eta tile (rows x chains of this wavefront)
    f32x16 e1;
#pragma unroll
    for (int r = 0; r < 16; ++r) e1[r] = 0.0f;
    bf16x8 bh = *(const bf16x8*)(bh_row);
    bf16x8 bl = *(const bf16x8*)(bl_row);
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) {
      bf16x8 nh = bh, nl = bl;
      if (t + 1 < GD / 16) {
        nh = *(const bf16x8*)(bh_row + 16 * (t + 1));
        nl = *(const bf16x8*)(bl_row + 16 * (t + 1));
      }
      e1 = mfma32(xa[t], bh, e1);
      e1 = mfma32(xa[t], bl, e1);
      bh = nh;
      bl = nl;
    }
    float ycur[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ycur[4 * q + 0] = yv[q].x;
      ycur[4 * q + 1] = yv[q].y;
      ycur[4 * q + 2] = yv[q].z;
      ycur[4 * q + 3] = yv[q].w;
    }
    load_x(next);
    // ---- epilogue: residual hi/lo to LDS (four consecutive rows per store),
    // log-likelihood partials
    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float res[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * q + i;
        const float eta = e1[r];
        const float yr = ycur[r];
        res[i] = yr - fast_logistic(eta);
        if (NEED_LJ) lsum += yr * eta - fast_log1p_exp(eta);
      }
      uint2 vh, vl;
      split_pair(res[0], res[1], vh.x, vl.x);
      split_pair(res[2], res[3], vh.y, vl.y);
      const int row0 = 8 * q + 4 * half;
      *(uint2*)(rh_row + row0) = vh;
      *(uint2*)(rl_row + row0) = vl;
    }
    if (NEED_LJ) ll += (double)lsum;
    // the residual tile is private to this wavefront, whose LDS operations
    // complete in program order: a wavefront-scope fence orders them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
    load_xt(next);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// positions (fp32, C x D) -> bf16 hi and lo parts, two per thread (n even)
__global__ void k_split_bf16(const float* __restrict__ q, bf16_t* __restrict__ hi, bf16_t* __restrict__ lo,
                             long long n) {
  const long long i = 2 * ((long long)blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const float2 v = *(const float2*)(q + i);
  uint32_t h, l;
  split_pair(v.x, v.y, h, l);
  *(uint32_t*)(hi + i) = h;
  *(uint32_t*)(lo + i) = l;
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
  hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)((n / 2 + 255) / 256)), dim3(256), 0, stream, q, s.bhi, s.blo, n);
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
