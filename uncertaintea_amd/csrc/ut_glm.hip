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
// Data movement.  A wavefront's position fragments (hi and lo, 32 chains x 256)
// stay in registers for the whole kernel.  Each 32-row block of X is read from
// global memory once per workgroup (16 KB: four 16-byte loads per thread, issued
// about one block ahead) and written to LDS twice: row-major for stage 1 (A
// operand: 8 covariates of one row per lane) and covariate-major for stage 2 (B
// operand: 8 rows of one covariate per lane), so X is stored only once in HBM.
// The tiles are triple-buffered (blocks k, k + 1 and the one being staged): one
// workgroup barrier per row block.  The loop is software-pipelined: stage 1 of
// block k + 1 is issued interleaved with the epilogue of block k (the vector work
// runs while the matrix cores multiply), and the staging writes of block k + 2 are
// interleaved with stage 2 of block k.  LDS reads run four fragments ahead of their
// multiplies.  Row strides are padded and the covariate-major rows permuted within
// groups of eight (row_of_cov) so that the 16-byte reads of both stages and the
// 8-byte transposing writes are free of bank conflicts.
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
constexpr int XRS = GD + 8;      // LDS row stride of the row-major X tile (bf16)
constexpr int XCS = RB + 8;      // LDS row stride of the covariate-major X tile (bf16)
constexpr int RSTR = RB + 8;     // LDS row stride of the residual tiles (bf16)
constexpr int PF = 4;            // LDS fragments in flight ahead of the multiplies
constexpr int NBUF = 3;          // X tile buffers: block k (stage 2), k + 1 (stage 1), k + 2 (staging)

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

// LDS row of covariate j in the covariate-major tile: rows rotate within each group
// of eight by the group index, so that eight consecutive covariates (the stage-2
// reads) and covariates j, j + 8, ..., j + 56 (one transposing write) all fall on
// distinct bank groups.
__device__ inline int row_of_cov(int j) { return (j & ~7) | ((j + (j >> 3)) & 7); }

template <bool NEED_LJ>
__global__ __launch_bounds__(256) void k_glm_grad(const bf16_t* __restrict__ X, const float* __restrict__ y,
                                                  const bf16_t* __restrict__ Bhi, const bf16_t* __restrict__ Blo,
                                                  int N, int C, int S, int rows_per_split,
                                                  float* __restrict__ Gpart, double* __restrict__ LLpart) {
  __shared__ __attribute__((aligned(16))) bf16_t Xr[NBUF][RB * XRS];
  __shared__ __attribute__((aligned(16))) bf16_t Xc[NBUF][GD * XCS];
  __shared__ __attribute__((aligned(16))) bf16_t Rs_hi[CB * RSTR];
  __shared__ __attribute__((aligned(16))) bf16_t Rs_lo[CB * RSTR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, half = lane >> 5;
  const int split = blockIdx.x % S;
  const int cb = blockIdx.x / S;
  const long long c0 = (long long)cb * CB;
  const long long r_begin = (long long)split * rows_per_split;
  const long long r_end = min((long long)N, r_begin + rows_per_split);
  const int nb = r_begin < r_end ? (int)((r_end - r_begin) / RB) : 0;
  const int cw = 32 * wave;  // this wavefront's chains within the block
  // first row of block k; blocks past the end re-read the last one, so every load
  // and multiply of the pipeline is unconditional (their results are not used)
  auto blk = [&](int k) { return r_begin + (long long)min(k, nb - 1) * RB; };

  // staging role of this thread: rows 4 sq .. 4 sq + 3, covariates 8 sc .. 8 sc + 7
  // (32 consecutive lanes cover one row: coalesced loads, conflict-free LDS writes)
  const int sq = tid >> 5, sc = tid & 31;
  typedef uint4 Stage[4];
  auto load_block = [&](Stage& stg, long long rb0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) stg[i] = *(const uint4*)(X + (rb0 + 4 * sq + i) * GD + 8 * sc);
  };
  // part p (0..7) of writing a staged block into buffer `buf`: covariate 8 sc + p of
  // the covariate-major tile (rows 4 sq .. 4 sq + 3 as one 8-byte store) and, for
  // p < 4, row 4 sq + p of the row-major tile
  auto store_part = [&](const Stage& stg, int buf, int p) {
    if (p < 4) *(uint4*)(&Xr[buf][(4 * sq + p) * XRS + 8 * sc]) = stg[p];
    const int wd = p >> 1;
    const uint32_t a0 = ((const uint32_t*)&stg[0])[wd], a1 = ((const uint32_t*)&stg[1])[wd];
    const uint32_t a2 = ((const uint32_t*)&stg[2])[wd], a3 = ((const uint32_t*)&stg[3])[wd];
    uint2 v;
    if (p & 1) {
      v.x = (a0 >> 16) | (a1 & 0xffff0000u);
      v.y = (a2 >> 16) | (a3 & 0xffff0000u);
    } else {
      v.x = (a0 & 0xffffu) | (a1 << 16);
      v.y = (a2 & 0xffffu) | (a3 << 16);
    }
    *(uint2*)(&Xc[buf][row_of_cov(8 * sc + p) * XCS + 4 * sq]) = v;
  };

  // this wavefront's positions (hi and lo): B operand of stage 1, lane l holds chain
  // cw + l % 32, covariates 16 t + 8 (l / 32) .. + 7
  bf16x8 bh[GD / 16], bl[GD / 16];
  {
    const long long brow = (c0 + cw + l32) * GD + 8 * half;
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) {
      bh[t] = *(const bf16x8*)(Bhi + brow + 16 * t);
      bl[t] = *(const bf16x8*)(Blo + brow + 16 * t);
    }
  }
  bf16_t* rh_row = Rs_hi + (cw + l32) * RSTR;
  bf16_t* rl_row = Rs_lo + (cw + l32) * RSTR;

  f32x16 g2[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) g2[j][r] = 0.0f;
  double ll = 0.0;

  float4 yv[4];  // outcomes of rows 8 q + 4 half .. + 3 (the epilogue's rows), one block ahead
  auto load_y = [&](long long rb0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) yv[q] = *(const float4*)(y + rb0 + 8 * q + 4 * half);
  };
  f32x16 eh, el;  // stage-1 result (hi and lo parts) of the block whose epilogue is next

  // One pipeline step for block k: stage 1 of block k + 1 with the epilogue of block k,
  // then stage 2 of block k while the staged block k + 2 (`stg`, loaded during the
  // previous step) goes into the buffer last read in step k - 1; then the loads of
  // block k + 3 into `stg`.
  Stage stg;
  auto step = [&](int k) {
    const int b0 = k % NBUF, b1 = (k + 1) % NBUF, b2 = (k + 2) % NBUF;
    float ycur[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ycur[4 * q + 0] = yv[q].x;
      ycur[4 * q + 1] = yv[q].y;
      ycur[4 * q + 2] = yv[q].z;
      ycur[4 * q + 3] = yv[q].w;
    }
    load_y(blk(k + 1));
    // ---- stage 1 of block k + 1 interleaved with the epilogue of block k: each step
    // issues two multiplies and then computes one residual of block k while they run
    // (residuals hi/lo to LDS, four consecutive rows per store; log-likelihood
    // partials).  LDS reads run PF fragments ahead; the scheduling fences keep this
    // order.
    f32x16 nh, nl;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      nh[r] = 0.0f;
      nl[r] = 0.0f;
    }
    const bf16_t* xrow = &Xr[b1][l32 * XRS + 8 * half];
    bf16x8 xa[PF];
#pragma unroll
    for (int t = 0; t < PF; ++t) xa[t] = *(const bf16x8*)(xrow + 16 * t);
    float lsum = 0.0f;
    float res[4];
#pragma unroll
    for (int t = 0; t < GD / 16; ++t) {
      __builtin_amdgcn_sched_barrier(0);
      const bf16x8 a = xa[t % PF];
      if (t + PF < GD / 16) xa[t % PF] = *(const bf16x8*)(xrow + 16 * (t + PF));
      nh = mfma32(a, bh[t], nh);
      nl = mfma32(a, bl[t], nl);
      // residual of accumulator register t (row 8 (t / 4) + 4 half + t % 4):
      // u = e^-|eta| serves both the logistic and log(1 + e^eta)
      const float eta = eh[t] + el[t];
      const float yr = ycur[t];
      const float u = __builtin_amdgcn_exp2f(-fabsf(eta) * 1.4426950408889634f);
      const float pr = __builtin_amdgcn_rcpf(1.0f + u);
      res[t & 3] = yr - (eta >= 0.0f ? pr : u * pr);
      if (NEED_LJ) lsum += yr * eta - (fmaxf(eta, 0.0f) + __builtin_amdgcn_logf(1.0f + u) * 0.6931471805599453f);
      if ((t & 3) == 3) {
        uint2 vh, vl;
        split_pair(res[0], res[1], vh.x, vl.x);
        split_pair(res[2], res[3], vh.y, vl.y);
        const int row0 = 8 * (t >> 2) + 4 * half;
        *(uint2*)(rh_row + row0) = vh;
        *(uint2*)(rl_row + row0) = vl;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (NEED_LJ) ll += (double)lsum;
    // the residual tile is private to this wavefront, whose LDS operations complete
    // in program order: a wavefront-scope fence orders them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- stage 2 of block k: Gt[32 chains x 256] += R^T X (B operand: 8 rows of
    // covariate 32 jn + l % 32 from the covariate-major tile), with one part of the
    // staged block's LDS writes per group
    const bf16_t* xc = &Xc[b0][8 * half];
    auto xc_frag = [&](int i) {  // fragment i = 8 t + jn
      return *(const bf16x8*)(xc + row_of_cov(32 * (i & 7) + l32) * XCS + 16 * (i >> 3));
    };
    bf16x8 ah[RB / 16], al[RB / 16], xb[PF];
#pragma unroll
    for (int t = 0; t < RB / 16; ++t) {
      ah[t] = *(const bf16x8*)(rh_row + 16 * t + 8 * half);
      al[t] = *(const bf16x8*)(rl_row + 16 * t + 8 * half);
    }
#pragma unroll
    for (int i = 0; i < PF; ++i) xb[i] = xc_frag(i);
    // two covariate tiles per group: hi parts of both, then lo parts, so no multiply
    // waits on the one just issued
#pragma unroll
    for (int i = 0; i < 8 * (RB / 16); i += 2) {
      __builtin_amdgcn_sched_barrier(0);
      const bf16x8 x0 = xb[i % PF], x1 = xb[(i + 1) % PF];
      if (i + PF < 8 * (RB / 16)) {
        xb[i % PF] = xc_frag(i + PF);
        xb[(i + 1) % PF] = xc_frag(i + 1 + PF);
      }
      const int jn = i & 7, t = i >> 3;
      g2[jn] = mfma32(ah[t], x0, g2[jn]);
      g2[jn + 1] = mfma32(ah[t], x1, g2[jn + 1]);
      g2[jn] = mfma32(al[t], x0, g2[jn]);
      g2[jn + 1] = mfma32(al[t], x1, g2[jn + 1]);
      store_part(stg, b2, i >> 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    load_block(stg, blk(k + 3));
    __syncthreads();
    eh = nh;
    el = nl;
  };

  // an empty split (possible when the rows do not divide evenly) writes zero partials;
  // nb is uniform over the workgroup, so no part of it skips a barrier
  if (nb > 0) {
    load_block(stg, blk(0));
#pragma unroll
    for (int p = 0; p < 8; ++p) store_part(stg, 0, p);
    load_block(stg, blk(1));
#pragma unroll
    for (int p = 0; p < 8; ++p) store_part(stg, 1, p);
    load_block(stg, blk(2));
    load_y(blk(0));
    __syncthreads();
    // stage 1 of block 0
    {
      const bf16_t* xrow = &Xr[0][l32 * XRS + 8 * half];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        eh[r] = 0.0f;
        el[r] = 0.0f;
      }
#pragma unroll
      for (int t = 0; t < GD / 16; ++t) {
        const bf16x8 a = *(const bf16x8*)(xrow + 16 * t);
        eh = mfma32(a, bh[t], eh);
        el = mfma32(a, bl[t], el);
      }
    }
    for (int k = 0; k < nb; ++k) step(k);
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
    hipLaunchKernelGGL(k_glm_grad<true>, grid, dim3(256), 0, stream, p.X, p.y, s.bhi, s.blo, p.N, C, S, rows,
                       s.gpart, s.llpart);
  else
    hipLaunchKernelGGL(k_glm_grad<false>, grid, dim3(256), 0, stream, p.X, p.y, s.bhi, s.blo, p.N, C, S, rows,
                       s.gpart, s.llpart);
  hipLaunchKernelGGL(k_glm_reduce, dim3(C), dim3(64), 0, stream, q, s.gpart, s.llpart, C, S, (float)p.loc,
                     (float)(1.0 / p.scale), p.total_const, need_lj ? 1 : 0, grad, lj);
  return hipGetLastError();
}

}  // namespace ut
