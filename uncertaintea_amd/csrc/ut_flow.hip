// ut_flow.hip -- forward, inverse, both backward passes and Adam training of the
// affine-coupling flow preconditioner (structure and conventions in ut_flow.h).
//
// Rows are chains (sampling) or draws (training); every per-row quantity is
// chain-major like the samplers' state.  Per coupling layer the forward pass is
//   gather passive x -> bf16 [B x kp]          k_gather
//   h = tanh(X_p W1^T + b1) (fp32 and bf16)    MFMA GEMM (K = kp), bias and tanh in its epilogue
//   O = h W2^T                                 MFMA GEMM (K = H)
//   s, t from O + b2; x_a <- x_a e^s + t       k_coupling_fwd   (k_coupling_inv: inverse)
// and the backward pass mirrors it with the transposed bf16 weight copies
//   dO (coupling), dU = (dO W2) (1 - h^2) (GEMM + epilogue), dX_p = dU W1 added
//   into the passive coordinates' gradient by the GEMM's epilogue;
// training adds dW2 = dO^T h and dW1 = dU^T X_p (transposed operands through LDS
// tiles) and the bias column sums.
#include <math.h>

#include <cstdlib>
#include <cstring>

#include "ut_flow.h"
#include "ut_gemm.h"
#include "ut_math.h"
#include "ut_mfma.h"

namespace ut {

namespace {

constexpr int T256 = 256;
inline unsigned nblk(long long n) { return (unsigned)((n + T256 - 1) / T256); }
inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

__global__ void k_logdet_init(const float* __restrict__ ls, int P, float sign, float* __restrict__ logdet, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.0f;
  for (int i = 0; i < P; ++i) s += ls[i];
  logdet[b] = sign * s;
}

// x = mu + e^ls z
__global__ void k_base_fwd(const float* __restrict__ z, const float* __restrict__ prm, int P, float* __restrict__ x,
                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % P);
  x[i] = prm[c] + expf(prm[P + c]) * z[i];
}

// z = (x - mu) e^-ls
__global__ void k_base_inv(const float* __restrict__ x, const float* __restrict__ prm, int P, float* __restrict__ z,
                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % P);
  z[i] = (x[i] - prm[c]) * expf(-prm[P + c]);
}

// g <- g e^ls (forward base, sampling backward)
__global__ void k_base_fwd_bwd(float* __restrict__ g, const float* __restrict__ prm, int P, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  g[i] *= expf(prm[P + (int)(i % P)]);
}

// loss 0.5 |z|^2 + sum ls + ...: g_x0 = z e^-ls; parameter sums per column
__global__ void k_base_inv_bwd(const float* __restrict__ z, const float* __restrict__ prm, int P,
                               float* __restrict__ g, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  g[i] = z[i] * expf(-prm[P + (int)(i % P)]);
}
// base-parameter gradient terms per element: -z e^-ls (mu) and 1 - z^2 (log scale)
__global__ void k_base_param_terms(const float* __restrict__ z, const float* __restrict__ prm, int P,
                                   float* __restrict__ tmu, float* __restrict__ tls, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = z[i];
  tmu[i] = -v * expf(-prm[P + (int)(i % P)]);
  tls[i] = 1.0f - v * v;
}

// forward-direction log-scale gradient terms per element: g_x0 e^ls z + rw (the
// base's log-determinant sum ls per valid row)
__global__ void k_base_fwd_param_terms(const float* __restrict__ z, const float* __restrict__ g,
                                       const float* __restrict__ prm, int P, const float* __restrict__ rw,
                                       float* __restrict__ tls, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % P);
  tls[i] = g[i] * expf(prm[P + c]) * z[i] + rw[i / P];
}

__global__ void k_normal_fill(float* __restrict__ z, long long n, uint32_t seed, uint32_t step) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  z[i] = rand_normal_f(seed, step, 11u, (uint32_t)i);
}

__global__ void k_gather(const float* __restrict__ x, int P, const int* __restrict__ pidx, int np, int kp,
                         bf16_t* __restrict__ xp, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long b = i / kp;
  const int k = (int)(i % kp);
  xp[i] = k < np ? f2bf(x[b * P + pidx[k]]) : (bf16_t)0;
}

// Coupling kernels: one wavefront per row, lane j handles active coordinates
// j, j + 64, ...; the row's sum of s is a wavefront reduction.
__device__ inline float wave_sum_f(float v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// forward coupling: x_a <- x_a e^s + t, logdet += sum s
__global__ __launch_bounds__(256) void k_coupling_fwd(float* __restrict__ x, int P, const float* __restrict__ O,
                                                      int op, const float* __restrict__ b2,
                                                      const int* __restrict__ aidx, int na, float c,
                                                      float* __restrict__ s_out, float* __restrict__ xs_out,
                                                      float* __restrict__ logdet, int B) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  float sum = 0.0f;
  for (int j = lane; j < na; j += 64) {
    const float s = c * flow_tanh((O[b * op + j] + b2[j]) * (1.0f / c));
    const float t = O[b * op + na + j] + b2[na + j];
    const long long ia = b * P + aidx[j];
    const float xs = x[ia] * expf(s);
    x[ia] = xs + t;
    s_out[b * na + j] = s;
    xs_out[b * na + j] = xs;
    sum += s;
  }
  sum = wave_sum_f(sum);
  if (lane == 0) logdet[b] += sum;
}

// inverse coupling: x_a <- (x_a - t) e^-s, logdet_inv -= sum s
__global__ __launch_bounds__(256) void k_coupling_inv(float* __restrict__ x, int P, const float* __restrict__ O,
                                                      int op, const float* __restrict__ b2,
                                                      const int* __restrict__ aidx, int na, float c,
                                                      float* __restrict__ s_out, float* __restrict__ xo_out,
                                                      float* __restrict__ logdet, int B) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  float sum = 0.0f;
  for (int j = lane; j < na; j += 64) {
    const float s = c * flow_tanh((O[b * op + j] + b2[j]) * (1.0f / c));
    const float t = O[b * op + na + j] + b2[na + j];
    const long long ia = b * P + aidx[j];
    const float xo = (x[ia] - t) * expf(-s);
    x[ia] = xo;
    s_out[b * na + j] = s;
    xo_out[b * na + j] = xo;
    sum += s;
  }
  sum = wave_sum_f(sum);
  if (lane == 0) logdet[b] -= sum;
}

// backward of the forward coupling (objective includes + sum s); writes dO (and
// pads) and updates g on the active coordinates
__global__ __launch_bounds__(256) void k_coupling_fwd_bwd(float* __restrict__ g, int P,
                                                          const float* __restrict__ s_in,
                                                          const float* __restrict__ xs_in,
                                                          const int* __restrict__ aidx, int na, int op, float c,
                                                          bf16_t* __restrict__ gO, float* __restrict__ gOf, int B,
                                                          const float* __restrict__ rw) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  const float wl = rw ? rw[b] : 1.0f;  // weight of the row's log-determinant term
  for (int j = lane; j < na; j += 64) {
    const long long ia = b * P + aidx[j];
    const float ga = g[ia];
    const float s = s_in[b * na + j];
    const float gs = ga * xs_in[b * na + j] + wl;
    const float r = s * (1.0f / c);
    const float gos = gs * (1.0f - r * r);
    gO[b * op + j] = f2bf(gos);
    gO[b * op + na + j] = f2bf(ga);
    if (gOf) {
      gOf[b * op + j] = gos;
      gOf[b * op + na + j] = ga;
    }
    g[ia] = ga * expf(s);
  }
  for (int j = 2 * na + lane; j < op; j += 64) {
    gO[b * op + j] = 0;
    if (gOf) gOf[b * op + j] = 0.0f;
  }
}

// backward of the inverse coupling for the loss 0.5|z|^2 + sum s
__global__ __launch_bounds__(256) void k_coupling_inv_bwd(float* __restrict__ g, int P,
                                                          const float* __restrict__ s_in,
                                                          const float* __restrict__ xo_in,
                                                          const int* __restrict__ aidx, int na, int op, float c,
                                                          bf16_t* __restrict__ gO, float* __restrict__ gOf, int B) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  for (int j = lane; j < na; j += 64) {
    const long long ia = b * P + aidx[j];
    const float go = g[ia];
    const float s = s_in[b * na + j];
    const float e = expf(-s);
    const float gt = -go * e;
    const float gs = -go * xo_in[b * na + j] + 1.0f;
    const float r = s * (1.0f / c);
    const float gos = gs * (1.0f - r * r);
    gO[b * op + j] = f2bf(gos);
    gO[b * op + na + j] = f2bf(gt);
    gOf[b * op + j] = gos;
    gOf[b * op + na + j] = gt;
    g[ia] = go * e;
  }
  for (int j = 2 * na + lane; j < op; j += 64) {
    gO[b * op + j] = 0;
    gOf[b * op + j] = 0.0f;
  }
}

// ---- fused coupling layers (sampling path, small widths) ---------------------------
// One workgroup per 32 rows runs a whole layer with both weight matrices staged in LDS
// once (all of their loads in flight together): the forward gathers the passive inputs,
// multiplies by W1 with bias and tanh applied to the accumulators, multiplies by W2 and
// applies the coupling; the backward runs the coupling's backward, multiplies by W2 with
// the tanh backward applied to the accumulators, multiplies by W1 and adds into the
// passive coordinates.  Each accumulator sums k in the order of k_gemm_nt (32-wide
// steps, two MFMAs each), so results equal the per-step kernels bit for bit.  Serves
// H = 256 with passive inputs and conditioner outputs of width 128 (both matrices then
// fit in LDS: 2 x 256 x 136 bf16 + the row tiles).
constexpr int LR = 32;              // rows per workgroup
constexpr int LH = 256;             // hidden width served
constexpr int LW = 128;             // passive / output width served
constexpr int LHP = LH + 8, LWP = LW + 8;  // LDS row strides (bf16)
constexpr size_t LAYER_LDS = (size_t)(2 * LH * LWP + LR * LHP) * sizeof(bf16_t);

// acc[t] = A[32 x K] (LDS, row stride lda) * Bs[col_t .. col_t + 31][K]^T (LDS, row
// stride ldb) for the wavefront's tiles col_t = 32 wave + 128 t, t < ntw
template <int NT>
__device__ inline void lds_gemm(const bf16_t* As, int lda, const bf16_t* Bs, int ldb, int K, int ntw,
                                f32x16 (&acc)[NT]) {
  const int lane = threadIdx.x & 63, l32 = lane & 31, half = lane >> 5, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const bf16x8 a = *(const bf16x8*)(As + l32 * lda + k0 + 16 * ks + 8 * half);
#pragma unroll
      for (int t = 0; t < NT; ++t)
        if (t < ntw) {
          const bf16x8 b = *(const bf16x8*)(Bs + (32 * wave + 128 * t + l32) * ldb + k0 + 16 * ks + 8 * half);
          acc[t] = mfma32(a, b, acc[t]);
        }
    }
  }
}

// The layer kernels issue every global load they need before their first wait (the
// weights, the row tile's inputs, biases, the coupling's operands and the values the
// epilogues combine with), so a workgroup waits on memory once; with one workgroup per
// CU and no other work in flight, each further dependent round trip would add its full
// latency to the layer.  The conditioner output stays in LDS (over the first matrix,
// dead once the first GEMM is done).
constexpr int NSTAGE = LH * LW / 8 / 256;  // 16-byte weight chunks per thread and matrix

// The fused pair keeps the tanh activations h (LR x LH per workgroup) in the first
// GEMM's accumulator order, private to k_layer_fwd / k_layer_bwd: element r of
// tile t of lane `lane` in wavefront `wave` of block `blk` sits at
// ((((blk * 4 + wave) * 2 + t) * 4 + r / 4) * 64 + lane) * 4 + r % 4, so each thread
// moves its 32 values as eight coalesced 16-byte vectors.
__device__ inline long long hfrag(int blk, int wave, int t, int rq, int lane) {
  return ((((long long)blk * 4 + wave) * 2 + t) * 4 + rq) * 256 + lane * 4;
}
constexpr int OSP = LW + 4;                // fp32 row stride of the conditioner output in LDS

// staged weights travel as native 4 x 32-bit vectors (HIP's uint4 is a struct, whose
// copies through a plain pointer keep the staging arrays in scratch memory)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Pointers read from the layer descriptors are generic; accesses through them would
// compile to flat instructions, which also count against the LDS wait counter (every
// LDS wait would then wait for the next layer's weights in flight).  They are global
// memory: say so.
#define UT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline UT_GLOBAL T* glb(T* p) {
  return (UT_GLOBAL T*)p;
}

__device__ inline void load_weights(const bf16_t* __restrict__ s1, const bf16_t* __restrict__ s2,
                                    u32x4 (&v1)[NSTAGE], u32x4 (&v2)[NSTAGE]) {
#pragma unroll
  for (int i = 0; i < NSTAGE; ++i) {
    const int e = threadIdx.x + 256 * i;
    v1[i] = *glb((const u32x4*)(s1 + (long long)(e / (LW / 8)) * LW + (e % (LW / 8)) * 8));
    v2[i] = *glb((const u32x4*)(s2 + (long long)(e / (LH / 8)) * LH + (e % (LH / 8)) * 8));
  }
}

// first matrix (LH x LW) -> LDS row stride LWP, second (LW x LH) -> LHP
__device__ inline void store_weights(bf16_t* d1, bf16_t* d2, const u32x4 (&v1)[NSTAGE], const u32x4 (&v2)[NSTAGE]) {
#pragma unroll
  for (int i = 0; i < NSTAGE; ++i) {
    const int e = threadIdx.x + 256 * i;
    *(u32x4*)(d1 + (e / (LW / 8)) * LWP + (e % (LW / 8)) * 8) = v1[i];
    *(u32x4*)(d2 + (e / (LH / 8)) * LHP + (e % (LH / 8)) * 8) = v2[i];
  }
}

// `par` is the layer's parity: its passive coordinates are 2k + par and its active ones
// 2j + 1 - par (the alternating masks), so no index list is loaded before the gathers
// One layer of the fused forward for this workgroup's LR rows.  v1 / v2 hold the
// layer's weights (loaded, or still in flight); with `next` the next layer's weight
// loads are issued into them as soon as this layer's weights are in LDS, so they fly
// while this layer computes.
__device__ __forceinline__ void layer_fwd(float* __restrict__ x, int P, const FusedFwdLayer& ly,
                                          const FusedFwdLayer* next, float c, float* __restrict__ logdet,
                                          u32x4 (&v1)[NSTAGE], u32x4 (&v2)[NSTAGE], bf16_t* lsh) {
  bf16_t* W1s = lsh;                  // LH x LWP: W1 [H x kp]
  bf16_t* W2s = lsh + LH * LWP;       // LW x LHP: W2 [op x H]
  bf16_t* R = lsh + 2 * LH * LWP;     // LR x LHP: passive inputs (stride LWP), then h
  float* Os = (float*)lsh;            // LR x OSP fp32 conditioner output, over W1 after the first GEMM
  const long long r0 = (long long)blockIdx.x * LR;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l32 = lane & 31;
  constexpr int G = LR * LW / 256;    // passive inputs gathered per thread
  constexpr int RW = LR / 4;          // coupling rows per wavefront
  const int np = ly.np, na = ly.na, par = ly.par;
  const UT_GLOBAL float* b1 = glb(ly.b1);
  const UT_GLOBAL float* b2 = glb(ly.b2);
  UT_GLOBAL float* hout = glb(ly.h);
  UT_GLOBAL float* sout = glb(ly.s);
  UT_GLOBAL float* xsout = glb(ly.xs);
  const bool act = lane < na;
  // ---- every row load of the layer up front; the passive and active coordinates
  // are 2k + par and 2j + 1 - par (the alternating masks)
  const int kq = tid & (LW - 1);      // the passive column this thread gathers (fixed: 256 = 2 LW)
  const int pk = 2 * kq + par;
  const int ak = 2 * lane + 1 - par;
  float xv[G];
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const int r = (tid + 256 * i) / LW;
    xv[i] = kq < np ? x[(r0 + r) * P + pk] : 0.0f;
  }
  float b1v[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) b1v[t] = b1[32 * wave + 128 * t + l32];
  const float b2s = act ? b2[lane] : 0.0f, b2t = act ? b2[na + lane] : 0.0f;
  float xa[RW], ldv[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const long long b = r0 + wave * RW + rr;
    xa[rr] = act ? x[b * P + ak] : 0.0f;
    ldv[rr] = lane == 0 ? logdet[b] : 0.0f;
  }
  __builtin_amdgcn_sched_barrier(0);
  store_weights(W1s, W2s, v1, v2);
  if (next) load_weights(next->W1b, next->W2b, v1, v2);
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const int r = (tid + 256 * i) / LW;
    R[r * LWP + kq] = kq < np ? f2bf(xv[i]) : (bf16_t)0;
  }
  __syncthreads();
  f32x16 acc[2];
  lds_gemm<2>(R, LWP, W1s, LWP, LW, 2, acc);
  __syncthreads();  // every wavefront is done with the passive inputs and W1; R now takes h
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rq = 0; rq < 4; ++rq) {
      f32x4 hv;
#pragma unroll
      for (int r3 = 0; r3 < 4; ++r3) {
        const int r = 4 * rq + r3;
        const int row = mfma32_row(lane, r), col = 32 * wave + 128 * t + l32;
        const float v = flow_tanh(acc[t][r] + b1v[t]);
        hv[r3] = v;
        R[row * LHP + col] = f2bf(v);
      }
      *(UT_GLOBAL f32x4*)(hout + hfrag(blockIdx.x, wave, t, rq, lane)) = hv;
    }
  __syncthreads();
  f32x16 acc2[1];
  lds_gemm<1>(R, LHP, W2s, LHP, LH, 1, acc2);
#pragma unroll
  for (int r = 0; r < 16; ++r) Os[mfma32_row(lane, r) * OSP + 32 * wave + l32] = acc2[0][r];
  __syncthreads();
  // coupling, as k_coupling_fwd (na <= 64: lane j owns active coordinate j), eight rows
  // per wavefront; the rows' log-determinant sums are reduced together (each in the
  // butterfly order of wave_sum_f) so their shuffle latencies overlap
  float sum[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wave * RW + rr;
    const long long b = r0 + row;
    sum[rr] = 0.0f;
    if (act) {
      const float sv = c * flow_tanh((Os[row * OSP + lane] + b2s) * (1.0f / c));
      const float tv = Os[row * OSP + na + lane] + b2t;
      const float xs = xa[rr] * expf(sv);
      x[b * P + ak] = xs + tv;
      sout[b * na + lane] = sv;
      xsout[b * na + lane] = xs;
      sum[rr] += sv;
    }
  }
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) sum[rr] += __shfl_xor(sum[rr], m);
  if (lane == 0)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) logdet[r0 + wave * RW + rr] = ldv[rr] + sum[rr];
}

// One layer of the fused backward (the coupling's backward, dO W2 with the tanh
// backward, dU W1 added into the passive gradients); weights as in layer_fwd.
__device__ __forceinline__ void layer_bwd(float* __restrict__ g, int P, const FusedBwdLayer& ly,
                                          const FusedBwdLayer* next, float c, u32x4 (&v1)[NSTAGE],
                                          u32x4 (&v2)[NSTAGE], bf16_t* lsh) {
  bf16_t* W2Ts = lsh;                 // LH x LWP: W2^T [H x op]
  bf16_t* W1Ts = lsh + LH * LWP;      // LW x LHP: W1^T [kp x H]
  bf16_t* R = lsh + 2 * LH * LWP;     // LR x LHP: dO (stride LWP), then dU
  const long long r0 = (long long)blockIdx.x * LR;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l32 = lane & 31;
  constexpr int RW = LR / 4;
  const int np = ly.np, na = ly.na, par = ly.par;
  const bool act = lane < na;
  const int pcol = 32 * wave + l32;   // the passive output column of the second GEMM
  // ---- every row load up front: the coupling backward's operands, the activations
  // of the tanh backward and the passive gradients the result adds to (the coupling
  // writes only active coordinates)
  const int ak = 2 * lane + 1 - par;
  const int pk = 2 * pcol + par;
  const UT_GLOBAL float* sin_ = glb(ly.s);
  const UT_GLOBAL float* xsin = glb(ly.xs);
  const UT_GLOBAL float* hin = glb(ly.h);
  float gv[RW], sv[RW], xv[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const long long b = r0 + wave * RW + rr;
    gv[rr] = act ? g[b * P + ak] : 0.0f;
    sv[rr] = act ? sin_[b * na + lane] : 0.0f;
    xv[rr] = act ? xsin[b * na + lane] : 0.0f;
  }
  f32x4 hq[2][4];
  float gp[16];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rq = 0; rq < 4; ++rq) hq[t][rq] = *(const UT_GLOBAL f32x4*)(hin + hfrag(blockIdx.x, wave, t, rq, lane));
#pragma unroll
  for (int r = 0; r < 16; ++r) gp[r] = pcol < np ? g[(r0 + mfma32_row(lane, r)) * P + pk] : 0.0f;
  __builtin_amdgcn_sched_barrier(0);
  store_weights(W2Ts, W1Ts, v1, v2);
  if (next) load_weights(next->W2Tb, next->W1Tb, v1, v2);
  // coupling backward, as k_coupling_fwd_bwd (na <= 64: lane j owns active coordinate j),
  // into LDS, eight rows per wavefront
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wave * RW + rr;
    if (act) {
      const float ga = gv[rr];
      const float gs = ga * xv[rr] + 1.0f;
      const float r = sv[rr] * (1.0f / c);
      const float gos = gs * (1.0f - r * r);
      R[row * LWP + lane] = f2bf(gos);
      R[row * LWP + na + lane] = f2bf(ga);
      g[(r0 + row) * P + ak] = ga * expf(sv[rr]);
    }
    for (int j = 2 * na + lane; j < LW; j += 64) R[row * LWP + j] = (bf16_t)0;
  }
  __syncthreads();
  f32x16 acc[2];
  lds_gemm<2>(R, LWP, W2Ts, LWP, LW, 2, acc);
  __syncthreads();  // every wavefront is done with dO; R now takes dU
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = mfma32_row(lane, r), col = 32 * wave + 128 * t + l32;
      const float hv = hq[t][r / 4][r % 4];
      R[row * LHP + col] = f2bf(acc[t][r] * (1.0f - hv * hv));
    }
  __syncthreads();
  f32x16 acc2[1];
  lds_gemm<1>(R, LHP, W1Ts, LHP, LH, 1, acc2);
#pragma unroll
  for (int r = 0; r < 16; ++r)
    if (pcol < np) g[(r0 + mfma32_row(lane, r)) * P + pk] = gp[r] + acc2[0][r];
}

// The kernels run one workgroup per CU (the weights fill the LDS), so one wavefront
// per SIMD: the whole register file is available to keep every staged load in flight.
// With the default occupancy target the compiler sinks each weight load to its LDS
// store to save registers, and the 32 loads then run one round trip at a time.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_layer_fwd(
    float* __restrict__ x, int P, const FusedFwdLayer* __restrict__ layers, int l, float c,
    float* __restrict__ logdet) {
  extern __shared__ __attribute__((aligned(16))) bf16_t lsh[];
  const FusedFwdLayer ly = layers[l];
  u32x4 v1[NSTAGE], v2[NSTAGE];
  load_weights(ly.W1b, ly.W2b, v1, v2);
  layer_fwd(x, P, ly, nullptr, c, logdet, v1, v2, lsh);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_layer_bwd(
    float* __restrict__ g, int P, const FusedBwdLayer* __restrict__ layers, int l, float c) {
  extern __shared__ __attribute__((aligned(16))) bf16_t lsh[];
  const FusedBwdLayer ly = layers[l];
  u32x4 v1[NSTAGE], v2[NSTAGE];
  load_weights(ly.W2Tb, ly.W1Tb, v1, v2);
  layer_bwd(g, P, ly, nullptr, c, v1, v2, lsh);
}

// ---- all-layer kernels: rows resident in LDS, weights in registers -------------------
// The sampling path runs the whole flow in one launch per direction.  A workgroup keeps
// its rows (fp32, row stride P) in LDS from the first layer to the last, so a layer
// reads and writes no row through global memory, and each wavefront holds the B operand
// of exactly the MFMA tiles it computes in registers, loaded straight from global (L2),
// so no weight passes through LDS.  Each matrix's fragments for the next layer are
// requested as soon as this layer's GEMM with that matrix is done (half a layer ahead,
// one register set per matrix); the LDS operands of each step are loaded as one batch
// before its arithmetic, so a wavefront waits on LDS once per batch.  The arithmetic, its
// order and the activations written for the backward are those of layer_fwd / layer_bwd
// (and so of the per-step kernels), bit for bit.

// Workgroup barrier for LDS traffic only: this wavefront's LDS operations complete, then
// s_barrier.  __syncthreads() also makes global memory coherent across the workgroup,
// which waits for every global load and store in flight (the next layer's weights, the
// activations just written); the resident kernels exchange data between wavefronts only
// through LDS, so their global traffic keeps flying across the barriers.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// v_i <- sum over the wavefront in the xor-butterfly order (partners i ^ 32, i ^ 16, ...,
// i ^ 1) for RW values at once, every lane ending with the same bits.  The 32 and 16
// steps exchange through ds_bpermute; from there on a lane's value depends only on i mod
// 16 (then mod 8, ...), so a rotation by 8 (then 4) inside the 16-lane DPP row reads the
// same value as lane i ^ 8 (i ^ 4), and quad permutations give i ^ 2 and i ^ 1 exactly.
template <int CTRL>
__device__ inline float flow_dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int N>
__device__ inline void xor_sum_rows(float (&v)[N]) {
#pragma unroll
  for (int m = 32; m >= 16; m >>= 1)
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += __shfl_xor(v[i], m);
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += flow_dpp<0x128>(v[i]);  // row_ror:8
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += flow_dpp<0x124>(v[i]);  // row_ror:4
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += flow_dpp<0x4E>(v[i]);   // quad_perm [2,3,0,1]
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += flow_dpp<0xB1>(v[i]);   // quad_perm [1,0,3,2]
}

constexpr int KF1 = LW / 16;  // 16-deep k-steps of the first GEMM (K = LW)
constexpr int KF2 = LH / 16;  // and of the second (K = LH)
// A workgroup serves RL = 32 or 16 rows.  The MFMA tiles stay 32 rows tall either way
// (so every accumulator sums k exactly as the per-step kernels do); with RL = 16 the
// A-operand rows 16..31 hold whatever the LDS holds and the result rows they produce are
// never read (row i of a product depends on row i of A only), and each thread keeps the
// first NR = RL / 2 accumulator registers of a tile, the ones holding rows < RL.  Sixteen
// rows per workgroup put 4096 rows on 256 workgroups, one per CU.
constexpr int FLOW_MAX_FUSED_P = LW + LW / 2;  // the fused layers' widths bound P (np <= 128, na <= 64)
template <int RL>
inline size_t res_lds_bytes(int P) {
  return (size_t)RL * P * sizeof(float) + (size_t)LR * LHP * sizeof(bf16_t) + (size_t)RL * OSP * sizeof(float);
}
// the fused kernels' activation fragment order (hfrag) for RL-row workgroups
template <int RL>
__device__ inline long long hfrag_rl(int blk, int wave, int t, int rq, int lane) {
  return ((((long long)blk * 4 + wave) * 2 + t) * (RL / 8) + rq) * 256 + lane * 4;
}

// B-operand fragments for this wavefront, 16-byte K-contiguous pieces (row l32 of the
// tile, k = 16 kk + 8 half .. + 7) of a K-contiguous weight matrix: the first GEMM's two
// 32-column tiles 32 wave + 128 t (K = LW, row stride LW) and the second GEMM's one tile
// 32 wave (K = LH, row stride LH)
// The matrices are read from copies in fragment order (written by k_weights_bf16), where the
// pieces of one load instruction -- one (wavefront, tile, k-step) for all 64 lanes -- are
// 1 KiB contiguous: fully coalesced loads.
__device__ inline long long frag_index1(int wave, int t, int kk, int lane) {
  return ((((long long)wave * 2 + t) * KF1 + kk) * 64 + lane) * 8;
}
__device__ inline long long frag_index2(int wave, int kk, int lane) {
  return (((long long)wave * KF2 + kk) * 64 + lane) * 8;
}
__device__ inline void frag_load1(const bf16_t* m, bf16x8 (&f)[2][KF1]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const UT_GLOBAL bf16_t* g = glb(m);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int kk = 0; kk < KF1; ++kk) f[t][kk] = *(const UT_GLOBAL bf16x8*)(g + frag_index1(wave, t, kk, lane));
}
__device__ inline void frag_load2(const bf16_t* m, bf16x8 (&f)[1][KF2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const UT_GLOBAL bf16_t* g = glb(m);
#pragma unroll
  for (int kk = 0; kk < KF2; ++kk) f[0][kk] = *(const UT_GLOBAL bf16x8*)(g + frag_index2(wave, kk, lane));
}

// acc[t] = A[32 x 16 KS] (LDS, row stride lda) * (tile t of the wavefront)^T, k in the
// order of lds_gemm; the A fragments are read in batches of 8
template <int NT, int KS>
__device__ inline void reg_gemm(const bf16_t* As, int lda, const bf16x8 (&B)[NT][KS], f32x16 (&acc)[NT]) {
  const int lane = threadIdx.x & 63, l32 = lane & 31, half = lane >> 5;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
#pragma unroll
  for (int k0 = 0; k0 < KS; k0 += 8) {
    bf16x8 a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = *(const bf16x8*)(As + l32 * lda + 16 * (k0 + j) + 8 * half);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = mfma32(a[j], B[t][k0 + j], acc[t]);
  }
}

// the forward layer's biases for this thread
struct FwdBias {
  float b1[2], b2s, b2t;
};
__device__ inline void fbias_load(const FusedFwdLayer& ly, FwdBias& b) {
  const int lane = threadIdx.x & 63, l32 = lane & 31, wave = threadIdx.x >> 6;
  const UT_GLOBAL float* b1 = glb(ly.b1);
  const UT_GLOBAL float* b2 = glb(ly.b2);
#pragma unroll
  for (int t = 0; t < 2; ++t) b.b1[t] = b1[32 * wave + 128 * t + l32];
  const bool act = lane < ly.na;
  b.b2s = act ? b2[lane] : 0.0f;
  b.b2t = act ? b2[ly.na + lane] : 0.0f;
}

// One forward layer on the resident rows Xs (RL x P); ldv[rr] is the log-determinant of
// row wave * RW + rr.  W1 / W2 / bs hold this layer's operands on entry and, when `next`
// is given, the next layer's on exit (requested as each is consumed).  With `gathered`
// R already holds the layer's bf16 passive inputs: the masks alternate, so the next
// layer's passive coordinates are this layer's active ones (np of the next layer = na),
// and the coupling writes them into R as it produces them.
template <int RL>
__device__ __forceinline__ void layer_fwd_res(float* Xs, int P, const FusedFwdLayer& ly, const FusedFwdLayer& nx,
                                              bool next, bool gathered, bf16x8 (&W1)[2][KF1], bf16x8 (&W2)[1][KF2],
                                              FwdBias& bs, float c, float (&ldv)[RL / 4], bf16_t* R, float* Os,
                                              long long r0) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l32 = lane & 31;
  constexpr int G = RL * LW / 256;
  constexpr int RW = RL / 4;
  constexpr int NR = RL / 2;
  const int np = ly.np, na = ly.na, par = ly.par;
  UT_GLOBAL float* hout = glb(ly.h);
  UT_GLOBAL float* sout = glb(ly.s);
  UT_GLOBAL float* xsout = glb(ly.xs);
  const bool act = lane < na;
  const int kq = tid & (LW - 1);
  const int pk = 2 * kq + par;
  const int ak = 2 * lane + 1 - par;
  if (!gathered) {
    float xv[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int r = (tid + 256 * i) / LW;
      xv[i] = Xs[r * P + min(pk, P - 1)];  // read in bounds for every lane, selected below
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int r = (tid + 256 * i) / LW;
      R[r * LWP + kq] = kq < np ? f2bf(xv[i]) : (bf16_t)0;
    }
    lds_barrier();
  }
  f32x16 acc[2];
  reg_gemm<2, KF1>(R, LWP, W1, acc);
  const FwdBias cb = bs;
  __builtin_amdgcn_sched_barrier(0);
  if (next) {
    frag_load1(nx.W1f, W1);
    fbias_load(nx, bs);
  }
  lds_barrier();  // every wavefront is done with the passive inputs; R now takes h
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rq = 0; rq < NR / 4; ++rq) {
      f32x4 hv;
#pragma unroll
      for (int r3 = 0; r3 < 4; ++r3) {
        const int r = 4 * rq + r3;
        const int row = mfma32_row(lane, r), col = 32 * wave + 128 * t + l32;
        const float v = flow_tanh(acc[t][r] + cb.b1[t]);
        hv[r3] = v;
        R[row * LHP + col] = f2bf(v);
      }
      *(UT_GLOBAL f32x4*)(hout + hfrag_rl<RL>(blockIdx.x, wave, t, rq, lane)) = hv;
    }
  lds_barrier();
  f32x16 acc2[1];
  reg_gemm<1, KF2>(R, LHP, W2, acc2);
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load2(nx.W2f, W2);
#pragma unroll
  for (int r = 0; r < NR; ++r) Os[mfma32_row(lane, r) * OSP + 32 * wave + l32] = acc2[0][r];
  lds_barrier();
  float os[RW], ot[RW], xa[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wave * RW + rr;
    os[rr] = Os[row * OSP + lane];  // lane + na < OSP: in bounds for every lane
    ot[rr] = Os[row * OSP + na + lane];
    xa[rr] = Xs[row * P + min(ak, P - 1)];
  }
  float sum[RW], sv[RW], xs[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    sv[rr] = c * flow_tanh((os[rr] + cb.b2s) * (1.0f / c));
    xs[rr] = xa[rr] * expf(sv[rr]);
    sum[rr] = 0.0f + (act ? sv[rr] : 0.0f);
  }
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wave * RW + rr;
    const long long b = r0 + row;
    const float xn = xs[rr] + (ot[rr] + cb.b2t);
    if (act) {
      Xs[row * P + ak] = xn;
      sout[b * na + lane] = sv[rr];
      xsout[b * na + lane] = xs[rr];
    }
    if (next) {  // the next layer's passive input j = lane (zero beyond its np = na)
      R[row * LWP + lane] = act ? f2bf(xn) : (bf16_t)0;
      R[row * LWP + 64 + lane] = (bf16_t)0;
    }
  }
  xor_sum_rows<RW>(sum);
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) ldv[rr] = ldv[rr] + sum[rr];
  lds_barrier();  // the active coordinates are written before the next layer gathers
}

// The whole forward in one launch: the elementwise base (x = mu + e^ls z and the
// log-determinant start sum ls, in the order of k_base_fwd / k_logdet_init), then all L
// coupling layers on the resident rows; the rows go to x (and q when given), the
// log-determinants to logdet (and logdet_out when given).
// LT > 0 fixes the layer count at compile time (L == LT) and unrolls the layers, so the
// waits for the operands requested half a layer ahead are counted exactly; LT == 0 runs
// any L in a loop, whose back edge has to wait for every request in flight.
template <int RL, int LT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RL == 8 ? 2 : 1, RL == 8 ? 2 : 1))) void k_flow_fwd_all(
    const float* __restrict__ z, const float* __restrict__ prm, float* __restrict__ x, float* __restrict__ q, int P,
    const FusedFwdLayer* __restrict__ layers, int L_, float c, float* __restrict__ logdet,
    float* __restrict__ logdet_out) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lsf[];
  float* Xs = lsf;
  bf16_t* R = (bf16_t*)(lsf + RL * P);
  float* Os = (float*)(R + LR * LHP);
  constexpr int RW = RL / 4;
  bf16x8 W1[2][KF1], W2[1][KF2];
  FwdBias bs;
  frag_load1(layers[0].W1f, W1);
  frag_load2(layers[0].W2f, W2);
  fbias_load(layers[0], bs);
  const long long r0 = (long long)blockIdx.x * RL;
  const int tid = threadIdx.x;
  for (int e0 = tid; e0 < RL * P; e0 += 4 * 256) {
    float zv[4], mv[4], lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 256 * k;
      const bool in = e < RL * P;
      const int cidx = in ? e % P : 0;
      zv[k] = in ? z[r0 * P + e] : 0.0f;
      mv[k] = prm[cidx];
      lv[k] = prm[P + cidx];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 256 * k;
      if (e < RL * P) Xs[e] = mv[k] + expf(lv[k]) * zv[k];
    }
  }
  float ldv[RW];
  {
    // the log-determinant start (every lane computes it; lane 0's is stored), sequential
    // sum in the order of k_logdet_init, loads eight at a time
    float sl = 0.0f;
    int i = 0;
    for (; i + 8 <= P; i += 8) {
      float t[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) t[k] = prm[P + i + k];
#pragma unroll
      for (int k = 0; k < 8; ++k) sl += t[k];
    }
    for (; i < P; ++i) sl += prm[P + i];
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) ldv[rr] = sl;
  }
  lds_barrier();  // the base rows are in LDS
#pragma unroll
  for (int l = 0; l < (LT > 0 ? LT : L); ++l) {
    // both descriptors through scalar loads: the next layer's pointers and widths are
    // then uniform registers, so requesting its operands waits on no vector memory
    const FusedFwdLayer ly = layers[l];
    const FusedFwdLayer nx = layers[l + 1 < L ? l + 1 : l];
    layer_fwd_res<RL>(Xs, P, ly, nx, l + 1 < L, l > 0, W1, W2, bs, c, ldv, R, Os, r0);
  }
  for (int e = tid; e < RL * P; e += 256) {
    const float v = Xs[e];
    x[r0 * P + e] = v;
    if (q) q[r0 * P + e] = v;
  }
  if ((tid & 63) == 0)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const long long b = r0 + (tid >> 6) * RW + rr;
      logdet[b] = ldv[rr];
      if (logdet_out) logdet_out[b] = ldv[rr];
    }
}

// What one backward layer reads besides its weights, for this thread: the scales and
// scaled inputs of its rows, and the tanh activations of its accumulator positions
template <int RL>
struct BwdRows {
  float s[RL / 4], xs[RL / 4];
};
template <int RL>
__device__ inline void brows_load(const FusedBwdLayer& ly, long long r0, BwdRows<RL>& a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int RW = RL / 4;
  const bool act = lane < ly.na;
  const UT_GLOBAL float* sin_ = glb(ly.s);
  const UT_GLOBAL float* xsin = glb(ly.xs);
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const long long b = r0 + wave * RW + rr;
    a.s[rr] = act ? sin_[b * ly.na + lane] : 0.0f;
    a.xs[rr] = act ? xsin[b * ly.na + lane] : 0.0f;
  }
}
template <int RL>
__device__ inline void bh_load(const FusedBwdLayer& ly, f32x4 (&h)[2][RL / 8]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const UT_GLOBAL float* hin = glb(ly.h);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rq = 0; rq < RL / 8; ++rq)
      h[t][rq] = *(const UT_GLOBAL f32x4*)(hin + hfrag_rl<RL>(blockIdx.x, wave, t, rq, lane));
}

// One backward layer on the resident gradient rows Gs (RL x P); the operands are
// replaced by the next (lower) layer's as they are consumed, as in layer_fwd_res.
template <int RL>
__device__ __forceinline__ void layer_bwd_res(float* Gs, int P, const FusedBwdLayer& ly, const FusedBwdLayer& nx,
                                              bool next, bf16x8 (&W2T)[2][KF1], bf16x8 (&W1T)[1][KF2],
                                              BwdRows<RL>& av, f32x4 (&h)[2][RL / 8], float c, bf16_t* R,
                                              long long r0) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l32 = lane & 31;
  constexpr int RW = RL / 4;
  constexpr int NR = RL / 2;
  const int np = ly.np, na = ly.na, par = ly.par;
  const bool act = lane < na;
  const int pcol = 32 * wave + l32;
  const int ak = 2 * lane + 1 - par;
  const int pk = 2 * pcol + par;
  float gv[RW];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) gv[rr] = Gs[(wave * RW + rr) * P + min(ak, P - 1)];  // selected by act below
  const BwdRows<RL> cr = av;
  __builtin_amdgcn_sched_barrier(0);
  if (next) brows_load<RL>(nx, r0, av);
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wave * RW + rr;
    if (act) {
      const float ga = gv[rr];
      const float gs = ga * cr.xs[rr] + 1.0f;
      const float r = cr.s[rr] * (1.0f / c);
      const float gos = gs * (1.0f - r * r);
      R[row * LWP + lane] = f2bf(gos);
      R[row * LWP + na + lane] = f2bf(ga);
      Gs[row * P + ak] = ga * expf(cr.s[rr]);
    }
    for (int j = 2 * na + lane; j < LW; j += 64) R[row * LWP + j] = (bf16_t)0;
  }
  lds_barrier();
  f32x16 acc[2];
  reg_gemm<2, KF1>(R, LWP, W2T, acc);
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load1(nx.W2Tf, W2T);
  lds_barrier();  // every wavefront is done with dO; R now takes dU
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int row = mfma32_row(lane, r), col = 32 * wave + 128 * t + l32;
      const float hv = h[t][r / 4][r % 4];
      R[row * LHP + col] = f2bf(acc[t][r] * (1.0f - hv * hv));
    }
  __builtin_amdgcn_sched_barrier(0);
  if (next) bh_load<RL>(nx, h);
  lds_barrier();
  f32x16 acc2[1];
  reg_gemm<1, KF2>(R, LHP, W1T, acc2);
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load2(nx.W1Tf, W1T);
  {
    float gp[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) gp[r] = Gs[mfma32_row(lane, r) * P + min(pk, P - 1)];
    if (pcol < np)
#pragma unroll
      for (int r = 0; r < NR; ++r) Gs[mfma32_row(lane, r) * P + pk] = gp[r] + acc2[0][r];
  }
  lds_barrier();  // the passive gradients are written; R is free
}

// The whole backward in one launch: the gradient rows g_q enter LDS once, all L coupling
// layers run in reverse on them, then the base's g_z = g e^ls (as k_base_fwd_bwd) is
// written straight to the output (g_z may alias g_q: a workgroup reads its rows before
// it writes them).
template <int RL, int LT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RL == 8 ? 2 : 1, RL == 8 ? 2 : 1))) void k_flow_bwd_all(
    const float* g_q, const float* __restrict__ prm, float* gz, int P, const FusedBwdLayer* __restrict__ layers,
    int L_, float c) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lsf[];
  float* Gs = lsf;
  bf16_t* R = (bf16_t*)(lsf + RL * P);
  const long long r0 = (long long)blockIdx.x * RL;
  const int tid = threadIdx.x;
  bf16x8 W2T[2][KF1], W1T[1][KF2];
  BwdRows<RL> av;
  f32x4 h[2][RL / 8];
  frag_load1(layers[L - 1].W2Tf, W2T);
  frag_load2(layers[L - 1].W1Tf, W1T);
  brows_load<RL>(layers[L - 1], r0, av);
  bh_load<RL>(layers[L - 1], h);
  for (int e = tid; e < RL * P; e += 256) Gs[e] = g_q[r0 * P + e];
  lds_barrier();
#pragma unroll
  for (int l = (LT > 0 ? LT : L) - 1; l >= 0; --l) {
    const FusedBwdLayer ly = layers[l];
    const FusedBwdLayer nx = layers[l > 0 ? l - 1 : l];
    layer_bwd_res<RL>(Gs, P, ly, nx, l > 0, W2T, W1T, av, h, c, R, r0);
  }
  for (int e0 = tid; e0 < RL * P; e0 += 4 * 256) {
    float lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 256 * k;
      lv[k] = prm[P + (e < RL * P ? e % P : 0)];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 256 * k;
      if (e < RL * P) gz[r0 * P + e] = Gs[e] * expf(lv[k]);
    }
  }
}

// ---- the 8-wavefront variant (16 rows, 512 threads) ----------------------------------
// Two wavefronts per SIMD hide each other's LDS, MFMA and memory latencies.  Wavefront w
// owns first-GEMM tile w (columns 32 w .. 32 w + 31 of H = 256: tile (w % 4, w / 4) of the
// 4-wavefront kernels, so the same fragment-order copies serve) and, for w < 4, the second
// GEMM's tile w; the coupling runs two rows per wavefront.  Same arithmetic, same order.
constexpr int W8 = 8;
constexpr int RL8 = 16;
constexpr int RW8 = RL8 / W8;  // coupling rows per wavefront
constexpr int NR8 = RL8 / 2;   // accumulator registers holding rows < 16

// the first GEMM's tile of wavefront w in fragment order (frag_index1's layout)
__device__ inline void frag_load1_w8(const bf16_t* m, bf16x8 (&f)[1][KF1]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const UT_GLOBAL bf16_t* g = glb(m);
#pragma unroll
  for (int kk = 0; kk < KF1; ++kk) f[0][kk] = *(const UT_GLOBAL bf16x8*)(g + frag_index1(w & 3, w >> 2, kk, lane));
}
__device__ inline void frag_load2_w8(const bf16_t* m, bf16x8 (&f)[1][KF2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (w >= 4) return;
  const UT_GLOBAL bf16_t* g = glb(m);
#pragma unroll
  for (int kk = 0; kk < KF2; ++kk) f[0][kk] = *(const UT_GLOBAL bf16x8*)(g + frag_index2(w, kk, lane));
}
// activations of accumulator register r = 4 rq + r3 (rq < 2) of wavefront w, lane
__device__ inline long long hfrag8(int blk, int w, int rq, int lane) {
  return ((((long long)blk * W8 + w) * 2 + rq) * 64 + lane) * 4;
}

struct FwdBias8 {
  float b1, b2s, b2t;
};
__device__ inline void fbias_load8(const FusedFwdLayer& ly, FwdBias8& b) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  b.b1 = glb(ly.b1)[32 * w + (lane & 31)];
  const bool act = lane < ly.na;
  b.b2s = act ? glb(ly.b2)[lane] : 0.0f;
  b.b2t = act ? glb(ly.b2)[ly.na + lane] : 0.0f;
}

__device__ __forceinline__ void layer_fwd_w8(float* Xs, int P, const FusedFwdLayer& ly, const FusedFwdLayer& nx,
                                             bool next, bool gathered, bf16x8 (&W1)[1][KF1], bf16x8 (&W2)[1][KF2],
                                             FwdBias8& bs, float c, float (&ldv)[RW8], bf16_t* R, float* Os,
                                             long long r0) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, l32 = lane & 31;
  const int np = ly.np, na = ly.na, par = ly.par;
  UT_GLOBAL float* hout = glb(ly.h);
  UT_GLOBAL float* sout = glb(ly.s);
  UT_GLOBAL float* xsout = glb(ly.xs);
  const bool act = lane < na;
  const int ak = 2 * lane + 1 - par;
  if (!gathered) {
    constexpr int G = RL8 * LW / 512;
    const int kq = tid & (LW - 1);
    const int pk = 2 * kq + par;
    float xv[G];
#pragma unroll
    for (int i = 0; i < G; ++i) xv[i] = Xs[((tid + 512 * i) / LW) * P + min(pk, P - 1)];
#pragma unroll
    for (int i = 0; i < G; ++i) R[((tid + 512 * i) / LW) * LWP + kq] = kq < np ? f2bf(xv[i]) : (bf16_t)0;
    lds_barrier();
  }
  f32x16 acc[1];
  reg_gemm<1, KF1>(R, LWP, W1, acc);
  const FwdBias8 cb = bs;
  __builtin_amdgcn_sched_barrier(0);
  if (next) {
    frag_load1_w8(nx.W1f, W1);
    fbias_load8(nx, bs);
  }
  lds_barrier();  // every wavefront is done with the passive inputs; R now takes h
#pragma unroll
  for (int rq = 0; rq < NR8 / 4; ++rq) {
    f32x4 hv;
#pragma unroll
    for (int r3 = 0; r3 < 4; ++r3) {
      const int r = 4 * rq + r3;
      const float v = flow_tanh(acc[0][r] + cb.b1);
      hv[r3] = v;
      R[mfma32_row(lane, r) * LHP + 32 * w + l32] = f2bf(v);
    }
    *(UT_GLOBAL f32x4*)(hout + hfrag8(blockIdx.x, w, rq, lane)) = hv;
  }
  lds_barrier();
  if (w < 4) {
    f32x16 acc2[1];
    reg_gemm<1, KF2>(R, LHP, W2, acc2);
#pragma unroll
    for (int r = 0; r < NR8; ++r) Os[mfma32_row(lane, r) * OSP + 32 * w + l32] = acc2[0][r];
  }
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load2_w8(nx.W2f, W2);
  lds_barrier();
  float os[RW8], ot[RW8], xa[RW8];
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) {
    const int row = w * RW8 + rr;
    os[rr] = Os[row * OSP + lane];
    ot[rr] = Os[row * OSP + na + lane];
    xa[rr] = Xs[row * P + min(ak, P - 1)];
  }
  float sum[RW8];
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) {
    const int row = w * RW8 + rr;
    const long long b = r0 + row;
    const float sv = c * flow_tanh((os[rr] + cb.b2s) * (1.0f / c));
    const float xs = xa[rr] * expf(sv);
    const float xn = xs + (ot[rr] + cb.b2t);
    sum[rr] = 0.0f + (act ? sv : 0.0f);
    if (act) {
      Xs[row * P + ak] = xn;
      sout[b * na + lane] = sv;
      xsout[b * na + lane] = xs;
    }
    if (next) {
      R[row * LWP + lane] = act ? f2bf(xn) : (bf16_t)0;
      R[row * LWP + 64 + lane] = (bf16_t)0;
    }
  }
  xor_sum_rows<RW8>(sum);
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) ldv[rr] = ldv[rr] + sum[rr];
  lds_barrier();
}

template <int LT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_flow_fwd_w8(
    const float* __restrict__ z, const float* __restrict__ prm, float* __restrict__ x, float* __restrict__ q, int P,
    const FusedFwdLayer* __restrict__ layers, int L_, float c, float* __restrict__ logdet,
    float* __restrict__ logdet_out) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lsf[];
  float* Xs = lsf;
  bf16_t* R = (bf16_t*)(lsf + RL8 * P);
  float* Os = (float*)(R + LR * LHP);
  bf16x8 W1[1][KF1], W2[1][KF2];
  FwdBias8 bs;
  frag_load1_w8(layers[0].W1f, W1);
  frag_load2_w8(layers[0].W2f, W2);
  fbias_load8(layers[0], bs);
  const long long r0 = (long long)blockIdx.x * RL8;
  const int tid = threadIdx.x;
  for (int e0 = tid; e0 < RL8 * P; e0 += 4 * 512) {
    float zv[4], mv[4], lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 512 * k;
      const bool in = e < RL8 * P;
      const int cidx = in ? e % P : 0;
      zv[k] = in ? z[r0 * P + e] : 0.0f;
      mv[k] = prm[cidx];
      lv[k] = prm[P + cidx];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 512 * k;
      if (e < RL8 * P) Xs[e] = mv[k] + expf(lv[k]) * zv[k];
    }
  }
  float ldv[RW8];
  {
    float sl = 0.0f;
    int i = 0;
    for (; i + 8 <= P; i += 8) {
      float t[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) t[k] = prm[P + i + k];
#pragma unroll
      for (int k = 0; k < 8; ++k) sl += t[k];
    }
    for (; i < P; ++i) sl += prm[P + i];
#pragma unroll
    for (int rr = 0; rr < RW8; ++rr) ldv[rr] = sl;
  }
  lds_barrier();
#pragma unroll
  for (int l = 0; l < (LT > 0 ? LT : L); ++l) {
    const FusedFwdLayer ly = layers[l];
    const FusedFwdLayer nx = layers[l + 1 < L ? l + 1 : l];
    layer_fwd_w8(Xs, P, ly, nx, l + 1 < L, l > 0, W1, W2, bs, c, ldv, R, Os, r0);
  }
  for (int e = tid; e < RL8 * P; e += 512) {
    const float v = Xs[e];
    x[r0 * P + e] = v;
    if (q) q[r0 * P + e] = v;
  }
  if ((tid & 63) == 0)
#pragma unroll
    for (int rr = 0; rr < RW8; ++rr) {
      const long long b = r0 + (tid >> 6) * RW8 + rr;
      logdet[b] = ldv[rr];
      if (logdet_out) logdet_out[b] = ldv[rr];
    }
}

struct BwdRows8 {
  float s[RW8], xs[RW8];
};
__device__ inline void brows_load8(const FusedBwdLayer& ly, long long r0, BwdRows8& a) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool act = lane < ly.na;
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) {
    const long long b = r0 + w * RW8 + rr;
    a.s[rr] = act ? glb(ly.s)[b * ly.na + lane] : 0.0f;
    a.xs[rr] = act ? glb(ly.xs)[b * ly.na + lane] : 0.0f;
  }
}
__device__ inline void bh_load8(const FusedBwdLayer& ly, f32x4 (&h)[NR8 / 4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int rq = 0; rq < NR8 / 4; ++rq) h[rq] = *(const UT_GLOBAL f32x4*)(glb(ly.h) + hfrag8(blockIdx.x, w, rq, lane));
}

__device__ __forceinline__ void layer_bwd_w8(float* Gs, int P, const FusedBwdLayer& ly, const FusedBwdLayer& nx,
                                             bool next, bf16x8 (&W2T)[1][KF1], bf16x8 (&W1T)[1][KF2], BwdRows8& av,
                                             f32x4 (&h)[NR8 / 4], float c, bf16_t* R, long long r0) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, l32 = lane & 31;
  const int np = ly.np, na = ly.na, par = ly.par;
  const bool act = lane < na;
  const int ak = 2 * lane + 1 - par;
  float gv[RW8];
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) gv[rr] = Gs[(w * RW8 + rr) * P + min(ak, P - 1)];
  const BwdRows8 cr = av;
  __builtin_amdgcn_sched_barrier(0);
  if (next) brows_load8(nx, r0, av);
#pragma unroll
  for (int rr = 0; rr < RW8; ++rr) {
    const int row = w * RW8 + rr;
    if (act) {
      const float ga = gv[rr];
      const float gs = ga * cr.xs[rr] + 1.0f;
      const float r = cr.s[rr] * (1.0f / c);
      const float gos = gs * (1.0f - r * r);
      R[row * LWP + lane] = f2bf(gos);
      R[row * LWP + na + lane] = f2bf(ga);
      Gs[row * P + ak] = ga * expf(cr.s[rr]);
    }
    for (int j = 2 * na + lane; j < LW; j += 64) R[row * LWP + j] = (bf16_t)0;
  }
  lds_barrier();
  f32x16 acc[1];
  reg_gemm<1, KF1>(R, LWP, W2T, acc);
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load1_w8(nx.W2Tf, W2T);
  lds_barrier();  // every wavefront is done with dO; R now takes dU
#pragma unroll
  for (int r = 0; r < NR8; ++r) {
    const float hv = h[r / 4][r % 4];
    R[mfma32_row(lane, r) * LHP + 32 * w + l32] = f2bf(acc[0][r] * (1.0f - hv * hv));
  }
  __builtin_amdgcn_sched_barrier(0);
  if (next) bh_load8(nx, h);
  lds_barrier();
  if (w < 4) {
    f32x16 acc2[1];
    reg_gemm<1, KF2>(R, LHP, W1T, acc2);
    const int pcol = 32 * w + l32;
    const int pk = 2 * pcol + par;
    float gp[NR8];
#pragma unroll
    for (int r = 0; r < NR8; ++r) gp[r] = Gs[mfma32_row(lane, r) * P + min(pk, P - 1)];
    if (pcol < np)
#pragma unroll
      for (int r = 0; r < NR8; ++r) Gs[mfma32_row(lane, r) * P + pk] = gp[r] + acc2[0][r];
  }
  __builtin_amdgcn_sched_barrier(0);
  if (next) frag_load2_w8(nx.W1Tf, W1T);
  lds_barrier();
}

template <int LT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_flow_bwd_w8(
    const float* g_q, const float* __restrict__ prm, float* gz, int P, const FusedBwdLayer* __restrict__ layers,
    int L_, float c) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lsf[];
  float* Gs = lsf;
  bf16_t* R = (bf16_t*)(lsf + RL8 * P);
  const long long r0 = (long long)blockIdx.x * RL8;
  const int tid = threadIdx.x;
  bf16x8 W2T[1][KF1], W1T[1][KF2];
  BwdRows8 av;
  f32x4 h[NR8 / 4];
  frag_load1_w8(layers[L - 1].W2Tf, W2T);
  frag_load2_w8(layers[L - 1].W1Tf, W1T);
  brows_load8(layers[L - 1], r0, av);
  bh_load8(layers[L - 1], h);
  for (int e = tid; e < RL8 * P; e += 512) Gs[e] = g_q[r0 * P + e];
  lds_barrier();
#pragma unroll
  for (int l = (LT > 0 ? LT : L) - 1; l >= 0; --l) {
    const FusedBwdLayer ly = layers[l];
    const FusedBwdLayer nx = layers[l > 0 ? l - 1 : l];
    layer_bwd_w8(Gs, P, ly, nx, l > 0, W2T, W1T, av, h, c, R, r0);
  }
  for (int e0 = tid; e0 < RL8 * P; e0 += 4 * 512) {
    float lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 512 * k;
      lv[k] = prm[P + (e < RL8 * P ? e % P : 0)];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + 512 * k;
      if (e < RL8 * P) gz[r0 * P + e] = Gs[e] * expf(lv[k]);
    }
  }
}

// rows per workgroup of the all-layer kernels: 16 while 32 would leave CUs idle (fewer
// than 256 workgroups), else 32 (half the weight traffic per row)
static int resident_rows(int B) {
  static const int forced = [] {
    const char* e = std::getenv("UT_FLOW_ROWS");  // 8, 16 or 32 rows per workgroup (comparison runs)
    const int v = e ? std::atoi(e) : 0;
    return (v == 8 || v == 16 || v == 32) ? v : 0;
  }();
  if (forced) return forced;
  return B / 32 < 256 ? 16 : 32;
}

// the all-layer kernel instances: layer counts 2, 4, 6 and 8 unrolled, others in a loop
typedef void (*FwdAllFn)(const float*, const float*, float*, float*, int, const FusedFwdLayer*, int, float, float*,
                         float*);
typedef void (*BwdAllFn)(const float*, const float*, float*, int, const FusedBwdLayer*, int, float);
template <int RL>
static FwdAllFn fwd_all_rl(int L) {
  switch (L) {
    case 2: return k_flow_fwd_all<RL, 2>;
    case 4: return k_flow_fwd_all<RL, 4>;
    case 6: return k_flow_fwd_all<RL, 6>;
    case 8: return k_flow_fwd_all<RL, 8>;
    default: return k_flow_fwd_all<RL, 0>;
  }
}
template <int RL>
static BwdAllFn bwd_all_rl(int L) {
  switch (L) {
    case 2: return k_flow_bwd_all<RL, 2>;
    case 4: return k_flow_bwd_all<RL, 4>;
    case 6: return k_flow_bwd_all<RL, 6>;
    case 8: return k_flow_bwd_all<RL, 8>;
    default: return k_flow_bwd_all<RL, 0>;
  }
}
static FwdAllFn fwd_w8_kernel(int L) {
  switch (L) {
    case 2: return k_flow_fwd_w8<2>;
    case 4: return k_flow_fwd_w8<4>;
    case 6: return k_flow_fwd_w8<6>;
    case 8: return k_flow_fwd_w8<8>;
    default: return k_flow_fwd_w8<0>;
  }
}
static BwdAllFn bwd_w8_kernel(int L) {
  switch (L) {
    case 2: return k_flow_bwd_w8<2>;
    case 4: return k_flow_bwd_w8<4>;
    case 6: return k_flow_bwd_w8<6>;
    case 8: return k_flow_bwd_w8<8>;
    default: return k_flow_bwd_w8<0>;
  }
}
// 16-row workgroups run the 8-wavefront kernels unless UT_FLOW_WAVES=4 (comparison runs)
static bool resident_w8(int rl) {
  static const bool four = [] {
    const char* e = std::getenv("UT_FLOW_WAVES");
    return e && std::strcmp(e, "4") == 0;
  }();
  return rl == 16 && !four;
}
static FwdAllFn fwd_all_kernel(int rl, int L) {
  return rl == 8 ? fwd_all_rl<8>(L) : rl == 16 ? fwd_all_rl<16>(L) : fwd_all_rl<32>(L);
}
static BwdAllFn bwd_all_kernel(int rl, int L) {
  return rl == 8 ? bwd_all_rl<8>(L) : rl == 16 ? bwd_all_rl<16>(L) : bwd_all_rl<32>(L);
}
template <int RL>
static size_t res_lds(int P) { return res_lds_bytes<RL>(P); }
static size_t res_lds_rl(int rl, int P) { return rl == 8 ? res_lds<8>(P) : rl == 16 ? res_lds<16>(P) : res_lds<32>(P); }

// column sums in two passes (fixed order): COLSUM_CHUNKS row chunks, then chunks
constexpr int COLSUM_CHUNKS = 256;  // 256 workgroups in the first pass (16 rows each at 4096 rows)
__global__ void k_colsum_part(const float* __restrict__ M, int B, int n, float* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int per = (B + COLSUM_CHUNKS - 1) / COLSUM_CHUNKS;
  const int r0 = blockIdx.y * per, r1 = min(B, r0 + per);
  float s = 0.0f;
#pragma unroll 16
  for (int b = r0; b < r1; ++b) s += M[(long long)b * n + c];
  part[(long long)blockIdx.y * n + c] = s;
}
__global__ void k_colsum_final(const float* __restrict__ part, int n, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s = 0.0f;
#pragma unroll 16
  for (int k = 0; k < COLSUM_CHUNKS; ++k) s += part[(long long)k * n + c];
  out[c] = s;
}

// out (Cc x R) = in (R x Cc)^T, bf16, 32 x 32 tiles through LDS
__global__ void k_transpose_bf16(const bf16_t* __restrict__ in, int R, int Cc, bf16_t* __restrict__ out) {
  __shared__ bf16_t tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
  for (int k = ty; k < 32; k += 8) tile[k][tx] = in[(long long)(r0 + k) * Cc + c0 + tx];
  __syncthreads();
  for (int k = ty; k < 32; k += 8) out[(long long)(c0 + k) * R + r0 + tx] = tile[tx][k];
}

// position of element (r, c) of a 256 x 128 (order 1) or 128 x 256 (order 2) matrix in
// the all-layer kernels' fragment order (frag_index1 / frag_index2 plus the element)
__device__ inline long long frag_pos(int r, int c, int rows) {
  const int kk = c / 16, lane = (r & 31) + 32 * ((c & 15) >> 3), e = c & 7;
  if (rows == LH) return frag_index1((r & 127) >> 5, r >> 7, kk, lane) + e;
  return frag_index2(r >> 5, kk, lane) + e;
}

// bf16 copies of the R x Cc fp32 matrix W: row-major, transposed and, when given (the
// fused widths: one of R, Cc is 256 and the other 128), both in fragment order
__global__ void k_weights_bf16(const float* __restrict__ W, int R, int Cc, bf16_t* __restrict__ Wb,
                               bf16_t* __restrict__ WTb, bf16_t* __restrict__ Wf, bf16_t* __restrict__ WTf) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)R * Cc) return;
  const int r = (int)(i / Cc), c = (int)(i % Cc);
  const bf16_t v = f2bf(W[i]);
  Wb[i] = v;
  WTb[(long long)c * R + r] = v;
  if (Wf) {
    Wf[frag_pos(r, c, R)] = v;
    WTf[frag_pos(c, r, Cc)] = v;
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2, float scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * scale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_gather_rows(const float* __restrict__ src, long long n, int P, float* __restrict__ dst, int B,
                              uint32_t seed, uint32_t step) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)B * P) return;
  const int b = (int)(i / P), c = (int)(i % P);
  const double u = rand_uniform_d(seed, step, 9u, (uint32_t)b);
  long long r = (long long)(u * (double)n);
  if (r >= n) r = n - 1;
  dst[i] = src[r * P + c];
}

__global__ void k_add_logdet(float* __restrict__ lj, const float* __restrict__ logdet, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) lj[b] = lj[b] + logdet[b];
}
__global__ void k_sub_logdet(double* __restrict__ lj, const float* __restrict__ logdet, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) lj[b] = lj[b] - (double)logdet[b];
}

__global__ __launch_bounds__(256) void k_row_loss(const float* __restrict__ z, const float* __restrict__ logdet_inv,
                                                  int B, int P, double* __restrict__ out) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  double s = 0.0;
  for (int c = lane; c < P; c += 64) {
    const double v = z[b * P + c];
    s += v * v;
  }
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
  if (lane == 0) out[b] = 0.5 * s - (double)logdet_inv[b] + 0.5 * P * 1.8378770664093453;
}

// one wavefront per row: validity, gradient masking and the row's log-weight
__global__ __launch_bounds__(256) void k_elbo_rows(const float* __restrict__ z, const float* __restrict__ lj,
                                                   const float* __restrict__ logdet, float* __restrict__ g, int B,
                                                   int P, float* __restrict__ rw, double* __restrict__ logw) {
  const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  bool ok = true;
  double zz = 0.0;
  for (int c = lane; c < P; c += 64) {
    ok = ok && isfinite(g[b * P + c]);
    const double v = z[b * P + c];
    zz += v * v;
  }
  for (int m = 32; m >= 1; m >>= 1) zz += __shfl_xor(zz, m);
  const bool valid = __all(ok) && isfinite(lj[b]) && isfinite(logdet[b]);
  if (!valid)
    for (int c = lane; c < P; c += 64) g[b * P + c] = 0.0f;
  if (lane == 0) {
    rw[b] = valid ? 1.0f : 0.0f;
    logw[b] = valid ? (double)lj[b] + 0.5 * zz + 0.9189385332046727 * (double)P + (double)logdet[b] : 0.0;
  }
}

// block partials: part[blk] = sum of squares of g over a grid-stride slice
__global__ __launch_bounds__(256) void k_sumsq_part(const float* __restrict__ g, long long n,
                                                    double* __restrict__ part) {
  __shared__ double sh[256];
  double a = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double v = g[i];
    a += v * v;
  }
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// one block: valid rows, mean log-weight, the importance-weighted bound when group
// statistics are given, clipped update scale (fixed summation order).  stats =
// {divisor (valid rows, or groups used), objective, gradient norm of the mean,
// update scale, standard ELBO}
__global__ __launch_bounds__(256) void k_elbo_stats(const float* __restrict__ rw, const double* __restrict__ logw,
                                                    int B, const double* __restrict__ part, int nparts,
                                                    const double* __restrict__ gstat, int ngroups, double clip,
                                                    double* __restrict__ stats) {
  __shared__ double sv[256], sl[256], sq[256], su[256], sb[256];
  double v = 0.0, l = 0.0, q = 0.0, u = 0.0, bd = 0.0;
  for (int b = threadIdx.x; b < B; b += 256) {
    v += (double)rw[b];
    l += logw[b];
  }
  for (int i = threadIdx.x; i < nparts; i += 256) q += part[i];
  if (gstat)
    for (int g = threadIdx.x; g < ngroups; g += 256) {
      bd += gstat[2 * g];
      u += gstat[2 * g + 1];
    }
  sv[threadIdx.x] = v;
  sl[threadIdx.x] = l;
  sq[threadIdx.x] = q;
  su[threadIdx.x] = u;
  sb[threadIdx.x] = bd;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sv[threadIdx.x] += sv[threadIdx.x + w];
      sl[threadIdx.x] += sl[threadIdx.x + w];
      sq[threadIdx.x] += sq[threadIdx.x + w];
      su[threadIdx.x] += su[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nv = sv[0];
    const double dv = gstat ? su[0] : nv;  // IWAE: group weights already sum to one per group
    const double norm = dv > 0.0 ? sqrt(sq[0]) / dv : 0.0;
    const double cf = (isfinite(clip) && norm > clip) ? clip / norm : 1.0;
    const double elbo = nv > 0.0 ? sl[0] / nv : 0.0;
    stats[0] = dv;
    stats[1] = gstat ? (dv > 0.0 ? sb[0] / dv : 0.0) : elbo;
    stats[2] = norm;
    stats[3] = dv > 0.0 ? cf / dv : 0.0;
    stats[4] = elbo;
  }
}

// one wavefront per group of K particles (the reference's _iwae_weights_and_bound!,
// src/inference/vi/guides.jl:383-425): log-sum-exp of the valid log-weights,
// self-normalised weights wts (zero for invalid particles; divided by the groups
// used later, through the update scale) and gstat[2g] = lse - log(valid count),
// gstat[2g + 1] = 1 when the group has a valid particle
__global__ __launch_bounds__(256) void k_iwae_groups(const float* __restrict__ rw, const double* __restrict__ logw,
                                                     int K, int ngroups, float* __restrict__ wts,
                                                     double* __restrict__ gstat) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;
  const long long b0 = (long long)g * K;
  double m = -INFINITY, cnt = 0.0;
  for (int k = lane; k < K; k += 64)
    if (rw[b0 + k] != 0.0f) {
      m = fmax(m, logw[b0 + k]);
      cnt += 1.0;
    }
  for (int o = 32; o >= 1; o >>= 1) {
    m = fmax(m, __shfl_xor(m, o));
    cnt += __shfl_xor(cnt, o);
  }
  double se = 0.0;
  for (int k = lane; k < K; k += 64)
    if (rw[b0 + k] != 0.0f) se += exp(logw[b0 + k] - m);
  for (int o = 32; o >= 1; o >>= 1) se += __shfl_xor(se, o);
  const double lse = m + log(se);
  for (int k = lane; k < K; k += 64) wts[b0 + k] = rw[b0 + k] != 0.0f ? (float)exp(logw[b0 + k] - lse) : 0.0f;
  if (lane == 0) {
    gstat[2 * g] = cnt > 0.0 ? lse - log(cnt) : 0.0;
    gstat[2 * g + 1] = cnt > 0.0 ? 1.0 : 0.0;
  }
}

// g[b, :] *= wts[b] (the model-gradient half of the weighted accumulation)
__global__ void k_scale_rows(float* __restrict__ g, const float* __restrict__ wts, int P, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] *= wts[i / P];
}

// Adam ascent (the reference's _adam_ascent_step!, src/inference/vi.jl:205-227) on
// the gradient scaled by stats[3]
__global__ void k_adam_ascent(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                              float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float bc1,
                              float bc2, const double* __restrict__ stats, const ElboTrack* __restrict__ track) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(stats[0] > 0.0) || (track && track->failed)) return;
  const float gi = g[i] * (float)stats[3];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] += lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// one thread: record iteration `it`'s statistics, the failure flag and whether the
// objective is a new best (the host loop's bookkeeping, kept on the device so the
// fit runs without a per-iteration synchronisation)
__global__ void k_elbo_track(const double* __restrict__ stats, int it, double* __restrict__ hist,
                             ElboTrack* __restrict__ tr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  tr->take = 0;
  if (tr->failed) return;
  if (!(stats[0] > 0.0)) {
    tr->failed = 1;
    tr->fail_iter = it;
    return;
  }
  hist[3LL * it] = stats[1];
  hist[3LL * it + 1] = stats[4];
  hist[3LL * it + 2] = stats[2];
  if (stats[1] > tr->best) {
    tr->best = stats[1];
    tr->take = 1;
  }
}

__global__ void k_copy_if_best(float* __restrict__ dst, const float* __restrict__ src, long long n,
                               const ElboTrack* __restrict__ tr) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && tr->take) dst[i] = src[i];
}

// ---- full-rank Gaussian guide: q = mu + e^ls z + F z, F strictly lower
// (src/inference/vi/guides.jl:96-117).  One thread per element; sums in binary64.

// x[b, i] += sum_{j<i} F[i, j] z[b, j]  (x already holds mu + e^ls z)
__global__ __launch_bounds__(256) void k_factor_fwd(const float* __restrict__ z, const float* __restrict__ F, int P,
                                                    float* __restrict__ x, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const long long b = t / P;
  const int i = (int)(t - b * P);
  const float* zr = z + b * P;
  const float* Fr = F + (long long)i * P;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;  // four chains: the sum is latency-bound
  int j = 0;
  for (; j + 4 <= i; j += 4) {
    a0 += (double)Fr[j] * (double)zr[j];
    a1 += (double)Fr[j + 1] * (double)zr[j + 1];
    a2 += (double)Fr[j + 2] * (double)zr[j + 2];
    a3 += (double)Fr[j + 3] * (double)zr[j + 3];
  }
  for (; j < i; ++j) a0 += (double)Fr[j] * (double)zr[j];
  x[t] = (float)((double)x[t] + ((a0 + a1) + (a2 + a3)));
}

// out[b, j] = sum_{i>j} F[i, j] g[b, i]: the factor's part of (dq/dz)^T g
__global__ __launch_bounds__(256) void k_factor_vjp(const float* __restrict__ g, const float* __restrict__ F, int P,
                                                    float* __restrict__ out, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const long long b = t / P;
  const int j = (int)(t - b * P);
  const float* gr = g + b * P;
  double a0 = 0.0, a1 = 0.0;
  int i = j + 1;
  for (; i + 2 <= P; i += 2) {
    a0 += (double)F[(long long)i * P + j] * (double)gr[i];
    a1 += (double)F[(long long)(i + 1) * P + j] * (double)gr[i + 1];
  }
  for (; i < P; ++i) a0 += (double)F[(long long)i * P + j] * (double)gr[i];
  out[t] = (float)(a0 + a1);
}

__global__ __launch_bounds__(256) void k_add_to(float* __restrict__ x, const float* __restrict__ a, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] += a[t];
}

// z = L^-1 (x - mu), L = diag(e^ls) + F: forward substitution, one thread per row
__global__ __launch_bounds__(256) void k_factor_inv(const float* __restrict__ x, const float* __restrict__ prm, int P,
                                                    float* __restrict__ z, int B) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= B) return;
  const float* xr = x + (long long)b * P;
  float* zr = z + (long long)b * P;
  const float* F = prm + 2LL * P;
  for (int i = 0; i < P; ++i) {
    double a = (double)xr[i] - (double)prm[i];
    const float* Fr = F + (long long)i * P;
    for (int j = 0; j < i; ++j) a -= (double)Fr[j] * (double)zr[j];
    zr[i] = (float)(a * exp(-(double)prm[P + i]));
  }
}

// part[y][i, j] = sum over row chunk y of g[b, i] z[b, j] for j < i (0 elsewhere):
// the factor's ELBO gradient (src/inference/vi/guides.jl:164-199) before the row sum
__global__ __launch_bounds__(256) void k_factor_grad_part(const float* __restrict__ g, const float* __restrict__ z,
                                                          int B, int P, double* __restrict__ part) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long PP = (long long)P * P;
  if (t >= PP) return;
  const int i = (int)(t / P), j = (int)(t - (long long)i * P);
  const int per = (B + (int)gridDim.y - 1) / (int)gridDim.y;
  const int b0 = (int)blockIdx.y * per, b1 = min(B, b0 + per);
  double a0 = 0.0, a1 = 0.0;
  if (j < i) {
    int b = b0;
    for (; b + 2 <= b1; b += 2) {
      a0 += (double)g[(long long)b * P + i] * (double)z[(long long)b * P + j];
      a1 += (double)g[(long long)(b + 1) * P + i] * (double)z[(long long)(b + 1) * P + j];
    }
    for (; b < b1; ++b) a0 += (double)g[(long long)b * P + i] * (double)z[(long long)b * P + j];
  }
  part[(long long)blockIdx.y * PP + t] = a0 + a1;
}

// out[t] = sum over the splits of part[k][t]: 8 lanes per entry (32 entries per
// workgroup), each summing every 8th split, then a fixed xor tree over the 8 lanes
__global__ __launch_bounds__(256) void k_factor_grad_final(const double* __restrict__ part, int splits, long long PP,
                                                           float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 32 + (threadIdx.x >> 3);
  const int sub = threadIdx.x & 7;
  double a = 0.0;
  if (t < PP)
    for (int k = sub; k < splits; k += 8) a += part[(long long)k * PP + t];
  a += __shfl_xor(a, 1, 8);
  a += __shfl_xor(a, 2, 8);
  a += __shfl_xor(a, 4, 8);
  if (t < PP && sub == 0) out[t] = (float)a;
}

}  // namespace

// ------------------------------------------------------------------ host side
template <class T>
static hipError_t own(Flow& f, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e == hipSuccess) f.owned.push_back(*p);
  return e;
}

hipError_t flow_alloc(Flow& f, int P, int L, int H, bool factor) {
  if (factor && L != 0) return hipErrorInvalidValue;
  f.P = P;
  f.L = L;
  f.H = H;
  f.factor = factor;
  {
    // UT_FLOW_FUSED=0 keeps the per-step kernels on the sampling path (for comparison runs)
    const char* e = std::getenv("UT_FLOW_FUSED");
    f.fused = !(e && std::strcmp(e, "0") == 0);
    static bool lds_set = false;  // the layer kernels use more than 64 KB of dynamic LDS
    if (!lds_set) {
      (void)hipFuncSetAttribute((const void*)k_layer_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LAYER_LDS);
      (void)hipFuncSetAttribute((const void*)k_layer_bwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LAYER_LDS);
      for (int rl : {8, 16, 32})
        for (int l : {0, 2, 4, 6, 8}) {
          const int bytes = (int)res_lds_rl(rl, FLOW_MAX_FUSED_P);
          (void)hipFuncSetAttribute((const void*)fwd_all_kernel(rl, l), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    bytes);
          (void)hipFuncSetAttribute((const void*)bwd_all_kernel(rl, l), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    bytes);
        }
      lds_set = true;
    }
  }
  long long off = 2LL * P;
  f.layers.assign(L, FlowLayer());
  std::vector<std::vector<int>> pid(L), aid(L);
  for (int l = 0; l < L; ++l) {
    for (int i = 0; i < P; ++i) ((i + l) % 2 == 0 ? pid[l] : aid[l]).push_back(i);
    FlowLayer& ly = f.layers[l];
    ly.np = (int)pid[l].size();
    ly.na = (int)aid[l].size();
    const int kp = round_up(ly.np, FLOW_PAD), op = round_up(2 * ly.na, FLOW_PAD);
    ly.oW1 = off;
    off += (long long)H * kp;
    ly.ob1 = off;
    off += H;
    ly.oW2 = off;
    off += (long long)op * H;
    ly.ob2 = off;
    off += op;
  }
  if (factor) {
    f.ofac = off;
    off += (long long)P * P;
  }
  f.nparams = off;
  hipError_t e = own(f, &f.params, (size_t)off);
  if (factor) f.fsplits = (int)std::max(1LL, std::min((long long)FACTOR_SPLITS, (16LL << 20) / ((long long)P * P)));
  if (e == hipSuccess && factor) e = own(f, &f.fpart, (size_t)f.fsplits * P * P);
  if (e == hipSuccess && L > 0) e = own(f, &f.dfwd, (size_t)L);
  if (e == hipSuccess && L > 0) e = own(f, &f.dbwd, (size_t)L);
  if (e == hipSuccess && L > 0) e = own(f, &f.d16, (size_t)L);
  if (e == hipSuccess) e = own(f, &f.ldbase, (size_t)1);
  if (e == hipSuccess) e = own(f, &f.grads, (size_t)off);
  if (e == hipSuccess) e = own(f, &f.adam_m, (size_t)off);
  if (e == hipSuccess) e = own(f, &f.adam_v, (size_t)off);
  if (e == hipSuccess) e = hipMemset(f.adam_m, 0, (size_t)off * 4);
  if (e == hipSuccess) e = hipMemset(f.adam_v, 0, (size_t)off * 4);
  for (int l = 0; l < L && e == hipSuccess; ++l) {
    FlowLayer& ly = f.layers[l];
    const int kp = round_up(ly.np, FLOW_PAD), op = round_up(2 * ly.na, FLOW_PAD);
    e = own(f, &ly.pidx, (size_t)ly.np);
    if (e == hipSuccess) e = own(f, &ly.aidx, (size_t)std::max(ly.na, 1));
    if (e == hipSuccess) e = hipMemcpy(ly.pidx, pid[l].data(), ly.np * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && ly.na) e = hipMemcpy(ly.aidx, aid[l].data(), ly.na * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = own(f, &ly.W1b, (size_t)H * kp);
    if (e == hipSuccess) e = own(f, &ly.W1Tb, (size_t)H * kp);
    if (e == hipSuccess) e = own(f, &ly.W2b, (size_t)H * op);
    if (e == hipSuccess) e = own(f, &ly.W2Tb, (size_t)H * op);
    if (e == hipSuccess) e = own(f, &ly.W1f, (size_t)H * kp);
    if (e == hipSuccess) e = own(f, &ly.W2f, (size_t)H * op);
    if (e == hipSuccess) e = own(f, &ly.W2Tf, (size_t)H * op);
    if (e == hipSuccess) e = own(f, &ly.W1Tf, (size_t)H * kp);
  }
  f.m16 = flow16_eligible(f);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking);
  return e;
}

void flow_free(Flow& f) {
  for (void* p : f.owned) (void)hipFree(p);
  f.owned.clear();
  if (f.stream) (void)hipStreamDestroy(f.stream);
  f.stream = nullptr;
}

hipError_t flow_scratch(Flow& f, int B) {
  if (B <= f.sc.B) return hipSuccess;
  FlowScratch& s = f.sc;
  s = FlowScratch();
  s.B = B;
  f.gen += 1;
  int kpm = FLOW_PAD, opm = FLOW_PAD;
  for (const FlowLayer& ly : f.layers) {
    kpm = std::max(kpm, round_up(ly.np, FLOW_PAD));
    opm = std::max(opm, round_up(2 * ly.na, FLOW_PAD));
  }
  const size_t H = f.H;
  hipError_t e = hipSuccess;
  s.xp.assign(f.L, nullptr);
  s.h.assign(f.L, nullptr);
  s.hb.assign(f.L, nullptr);
  s.s.assign(f.L, nullptr);
  s.xa.assign(f.L, nullptr);
  s.fwd_fused.assign(f.L, 0);
  for (int l = 0; l < f.L && e == hipSuccess; ++l) {
    const FlowLayer& ly = f.layers[l];
    e = own(f, &s.xp[l], (size_t)B * round_up(ly.np, FLOW_PAD));
    if (e == hipSuccess) e = own(f, &s.h[l], (size_t)B * H);
    if (e == hipSuccess) e = own(f, &s.hb[l], (size_t)B * H);
    // row stride na on the per-step path, 64 in the 16x16x32-MFMA kernels (unconditional stores)
    if (e == hipSuccess) e = own(f, &s.s[l], (size_t)B * std::max(ly.na, 64));
    if (e == hipSuccess) e = own(f, &s.xa[l], (size_t)B * std::max(ly.na, 64));
  }
  const size_t wide = std::max<size_t>(H, (size_t)std::max(kpm, opm));
  if (e == hipSuccess) e = own(f, &s.O, (size_t)B * opm);
  if (e == hipSuccess) e = own(f, &s.gO, (size_t)B * opm);
  if (e == hipSuccess) e = own(f, &s.gOf, (size_t)B * opm);
  if (e == hipSuccess) e = own(f, &s.gU, (size_t)B * H);
  if (e == hipSuccess) e = own(f, &s.gUf, (size_t)B * H);
  if (e == hipSuccess) e = own(f, &s.tA, (size_t)B * wide);
  if (e == hipSuccess) e = own(f, &s.tB, (size_t)B * wide);
  if (e == hipSuccess) e = own(f, &s.x, (size_t)B * f.P);
  if (e == hipSuccess) e = own(f, &s.g, (size_t)B * f.P);
  if (e == hipSuccess) e = own(f, &s.aux, (size_t)B * f.P);
  if (e == hipSuccess) e = own(f, &s.logdet, (size_t)B);
  if (e == hipSuccess) e = own(f, &s.colsum, (size_t)COLSUM_CHUNKS * std::max<size_t>(wide, (size_t)f.P));
  if (e == hipSuccess) e = own(f, &s.wpart, (size_t)FLOW_WSPLITS * H * wide);
  if (e == hipSuccess && f.L > 0) {
    std::vector<FusedFwdLayer> fw(f.L);
    std::vector<FusedBwdLayer> bw(f.L);
    for (int l = 0; l < f.L; ++l) {
      const FlowLayer& ly = f.layers[l];
      fw[l] = {ly.W1b, f.params + ly.ob1, ly.W2b, f.params + ly.ob2, s.h[l], s.s[l], s.xa[l], ly.np, ly.na, l & 1,
               ly.W1f, ly.W2f};
      bw[l] = {ly.W2Tb, ly.W1Tb, s.h[l], s.s[l], s.xa[l], ly.np, ly.na, l & 1, ly.W2Tf, ly.W1Tf};
    }
    e = hipMemcpy(f.dfwd, fw.data(), fw.size() * sizeof(FusedFwdLayer), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f.dbwd, bw.data(), bw.size() * sizeof(FusedBwdLayer), hipMemcpyHostToDevice);
    std::vector<Flow16Layer> d16(f.L);
    for (int l = 0; l < f.L; ++l) {
      const FlowLayer& ly = f.layers[l];
      d16[l] = {ly.W1f, ly.W2f, ly.W2Tf, ly.W1Tf, f.params + ly.ob1, f.params + ly.ob2, s.s[l], s.xa[l], s.xp[l],
                ly.np, ly.na, l & 1};
    }
    if (e == hipSuccess) e = hipMemcpy(f.d16, d16.data(), d16.size() * sizeof(Flow16Layer), hipMemcpyHostToDevice);
  }
  return e;
}

hipError_t flow_sync_weights(Flow& f, hipStream_t st) {
  for (int l = 0; l < f.L; ++l) {
    FlowLayer& ly = f.layers[l];
    const int kp = round_up(ly.np, FLOW_PAD), op = round_up(2 * ly.na, FLOW_PAD);
    // the all-layer kernels' widths; the 16x16x32 kernels write their own fragment order
    const bool frag = f.H == LH && kp == LW && op == LW && !f.m16;
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk((long long)f.H * kp)), dim3(T256), 0, st, f.params + ly.oW1, f.H,
                       kp, ly.W1b, ly.W1Tb, frag ? ly.W1f : nullptr, frag ? ly.W1Tf : nullptr);
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk((long long)op * f.H)), dim3(T256), 0, st, f.params + ly.oW2, op,
                       f.H, ly.W2b, ly.W2Tb, frag ? ly.W2f : nullptr, frag ? ly.W2Tf : nullptr);
  }
  if (f.m16) return flow16_sync_weights(f, st);
  return hipGetLastError();
}

static void colsum(const float* M, int B, int n, float* out, float* part, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum_part, dim3(nblk(n), COLSUM_CHUNKS), dim3(T256), 0, st, M, B, n, part);
  hipLaunchKernelGGL(k_colsum_final, dim3(nblk(n)), dim3(T256), 0, st, part, n, out);
}

// the LDS-staged layer kernels serve H = 256 with passive and output widths of 128

static bool fused_layer(const Flow& f, int l) {
  const FlowLayer& ly = f.layers[l];
  return f.fused && f.H == LH && round_up(ly.np, FLOW_PAD) == LW && round_up(2 * ly.na, FLOW_PAD) == LW;
}

// conditioner of layer l on the current rows of sc.x: sc.O = h W2^T (bias added by the coupling)
static hipError_t conditioner(Flow& f, int l, int B, hipStream_t st) {
  FlowLayer& ly = f.layers[l];
  FlowScratch& s = f.sc;
  const int kp = round_up(ly.np, FLOW_PAD), op = round_up(2 * ly.na, FLOW_PAD), H = f.H;
  hipLaunchKernelGGL(k_gather, dim3(nblk((long long)B * kp)), dim3(T256), 0, st, s.x, f.P, ly.pidx, ly.np, kp,
                     s.xp[l], (long long)B * kp);
  // h = tanh(x_p W1^T + b1) straight from the GEMM's epilogue
  hipError_t e = launch_gemm_nt_bias_tanh(s.xp[l], kp, ly.W1b, kp, B, H, kp,
                                          GemmEpiBiasTanh{f.params + ly.ob1, s.h[l], s.hb[l], H}, st);
  if (e != hipSuccess) return e;
  return launch_gemm_nt(s.hb[l], H, ly.W2b, H, s.O, op, B, op, H, 1, 0, st);
}

// backward through the conditioner of layer l from sc.gO into sc.g (passive coordinates)
static hipError_t conditioner_bwd(Flow& f, int l, int B, bool weights, hipStream_t st) {
  FlowLayer& ly = f.layers[l];
  FlowScratch& s = f.sc;
  const int kp = round_up(ly.np, FLOW_PAD), op = round_up(2 * ly.na, FLOW_PAD), H = f.H;
  // gU = (gO W2) (1 - h^2) and g[passive] += gU W1, both in the GEMMs' epilogues
  hipError_t e = launch_gemm_nt_tanh_bwd(s.gO, op, ly.W2Tb, op, B, H, op,
                                         GemmEpiTanhBwd{s.h[l], s.gU, weights ? s.gUf : nullptr, H}, st);
  if (e != hipSuccess) return e;
  e = launch_gemm_nt_scatter_add(s.gU, H, ly.W1Tb, H, B, kp, H, GemmEpiScatterAdd{s.g, f.P, ly.pidx, ly.np}, st);
  if (e != hipSuccess) return e;
  if (!weights) return hipGetLastError();
  // dW2 (op x H) = gO^T h ; dW1 (H x kp) = gU^T X_p ; bias column sums
  hipLaunchKernelGGL(k_transpose_bf16, dim3(op / 32, B / 32), dim3(T256), 0, st, s.gO, B, op, s.tA);
  hipLaunchKernelGGL(k_transpose_bf16, dim3(H / 32, B / 32), dim3(T256), 0, st, s.hb[l], B, H, s.tB);
  const int splits = std::min(FLOW_WSPLITS, B / FLOW_PAD);  // K = B rows: split so the grid fills the CUs
  e = launch_gemm_nt_splitk(s.tA, B, s.tB, B, f.grads + ly.oW2, H, op, H, B, splits, s.wpart, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_transpose_bf16, dim3(H / 32, B / 32), dim3(T256), 0, st, s.gU, B, H, s.tA);
  hipLaunchKernelGGL(k_transpose_bf16, dim3(kp / 32, B / 32), dim3(T256), 0, st, s.xp[l], B, kp, s.tB);
  e = launch_gemm_nt_splitk(s.tA, B, s.tB, B, f.grads + ly.oW1, kp, H, kp, B, splits, s.wpart, st);
  if (e != hipSuccess) return e;
  colsum(s.gUf, B, H, f.grads + ly.ob1, s.colsum, st);
  colsum(s.gOf, B, op, f.grads + ly.ob2, s.colsum, st);
  return hipGetLastError();
}

static hipError_t forward_impl(Flow& f, const float* z, float* q, float* logdet, int B, bool allow_fused,
                               hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  const long long n = (long long)B * f.P;
  bool all = allow_fused && f.L > 0;
  for (int l = 0; l < f.L; ++l) {
    s.fwd_fused[l] = allow_fused && fused_layer(f, l);
    all = all && s.fwd_fused[l];
  }
  s.fwd16 = false;
  if (all) {
    // the rows go to s.x and the log-determinants to s.logdet (read by the model gradient
    // and the log-density) and also straight to q / logdet when given
    float* q2 = (q && q != s.x) ? q : nullptr;
    float* l2 = (logdet && logdet != s.logdet) ? logdet : nullptr;
    if (f.m16) {
      s.fwd16 = true;
      return flow16_forward(f, z, s.x, q2, s.logdet, l2, B, st);
    }
    const int rl = resident_rows(B);
    const bool w8 = resident_w8(rl);
    hipLaunchKernelGGL(w8 ? fwd_w8_kernel(f.L) : fwd_all_kernel(rl, f.L), dim3(B / rl), dim3(w8 ? 512 : T256),
                       res_lds_rl(rl, f.P), st, z, f.params, s.x, q2, f.P, f.dfwd, f.L, f.clamp, s.logdet, l2);
    return hipGetLastError();
  } else {
    hipLaunchKernelGGL(k_logdet_init, dim3(nblk(B)), dim3(T256), 0, st, f.params + f.P, f.P, 1.0f, s.logdet, B);
    hipLaunchKernelGGL(k_base_fwd, dim3(nblk(n)), dim3(T256), 0, st, z, f.params, f.P, s.x, n);
    if (f.factor)
      hipLaunchKernelGGL(k_factor_fwd, dim3(nblk(n)), dim3(T256), 0, st, z, f.params + f.ofac, f.P, s.x, n);
    for (int l = 0; l < f.L; ++l) {
      FlowLayer& ly = f.layers[l];
      if (s.fwd_fused[l]) {
        hipLaunchKernelGGL(k_layer_fwd, dim3(B / LR), dim3(T256), LAYER_LDS, st, s.x, f.P, f.dfwd, l, f.clamp,
                           s.logdet);
        continue;
      }
      hipError_t e = conditioner(f, l, B, st);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_coupling_fwd, dim3((B + 3) / 4), dim3(T256), 0, st, s.x, f.P, s.O,
                         round_up(2 * ly.na, FLOW_PAD), f.params + ly.ob2, ly.aidx, ly.na, f.clamp, s.s[l], s.xa[l],
                         s.logdet, B);
    }
  }
  if (q && q != s.x) {
    hipError_t e = hipMemcpyAsync(q, s.x, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  if (logdet && logdet != s.logdet) {
    hipError_t e = hipMemcpyAsync(logdet, s.logdet, (size_t)B * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t flow_forward(Flow& f, const float* z, float* q, float* logdet, int B, hipStream_t st) {
  return forward_impl(f, z, q, logdet, B, true, st);
}

hipError_t flow_forward_train(Flow& f, const float* z, int B, hipStream_t st) {
  return forward_impl(f, z, nullptr, nullptr, B, false, st);
}

hipError_t flow_forward_model(Flow& f, const float* z, const FlowRowModel& m, float* g, float* lj, int B,
                              hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  bool all = f.m16 && f.L > 0;
  for (int l = 0; l < f.L && all; ++l) all = fused_layer(f, l);
  if (!all) return hipErrorNotSupported;
  for (int l = 0; l < f.L; ++l) s.fwd_fused[l] = true;
  s.fwd16 = true;
  return flow16_forward_model(f, z, s.x, s.logdet, m, g, lj, B, st);
}

hipError_t flow_backward(Flow& f, const float* g_q, float* g_z, int B, hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  const long long n = (long long)B * f.P;
  if (s.fwd16) return flow16_backward(f, g_q, g_z, B, st);
  bool all = f.L > 0;
  for (int l = 0; l < f.L; ++l) all = all && s.fwd_fused[l];
  if (all) {
    // the gradient rows are read from g_q once; base backward fused in, written to g_z
    const int rl = resident_rows(B);
    const bool w8 = resident_w8(rl);
    hipLaunchKernelGGL(w8 ? bwd_w8_kernel(f.L) : bwd_all_kernel(rl, f.L), dim3(B / rl), dim3(w8 ? 512 : T256),
                       res_lds_rl(rl, f.P), st, g_q, f.params, g_z, f.P, f.dbwd, f.L, f.clamp);
    return hipGetLastError();
  }
  if (g_q != s.g) {
    hipError_t e = hipMemcpyAsync(s.g, g_q, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  {
    for (int l = f.L - 1; l >= 0; --l) {
      FlowLayer& ly = f.layers[l];
      const int op = round_up(2 * ly.na, FLOW_PAD);
      if (s.fwd_fused[l]) {
        hipLaunchKernelGGL(k_layer_bwd, dim3(B / LR), dim3(T256), LAYER_LDS, st, s.g, f.P, f.dbwd, l, f.clamp);
        continue;
      }
      hipLaunchKernelGGL(k_coupling_fwd_bwd, dim3((B + 3) / 4), dim3(T256), 0, st, s.g, f.P, s.s[l], s.xa[l],
                         ly.aidx, ly.na, op, f.clamp, s.gO, (float*)nullptr, B, (const float*)nullptr);
      hipError_t e = conditioner_bwd(f, l, B, false, st);
      if (e != hipSuccess) return e;
    }
  }
  if (f.factor)  // F^T g from the gradient before the base scales it
    hipLaunchKernelGGL(k_factor_vjp, dim3(nblk(n)), dim3(T256), 0, st, s.g, f.params + f.ofac, f.P, s.aux, n);
  hipLaunchKernelGGL(k_base_fwd_bwd, dim3(nblk(n)), dim3(T256), 0, st, s.g, f.params, f.P, n);
  if (f.factor) hipLaunchKernelGGL(k_add_to, dim3(nblk(n)), dim3(T256), 0, st, s.g, s.aux, n);
  if (g_z != s.g) {
    hipError_t e = hipMemcpyAsync(g_z, s.g, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t flow_inverse(Flow& f, const float* q, float* z, float* logdet_inv, int B, hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  s.fwd16 = false;  // the saved values below are the inverse's
  const long long n = (long long)B * f.P;
  if (q != s.x) {
    hipError_t e = hipMemcpyAsync(s.x, q, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_logdet_init, dim3(nblk(B)), dim3(T256), 0, st, f.params + f.P, f.P, -1.0f, s.logdet, B);
  for (int l = f.L - 1; l >= 0; --l) {
    FlowLayer& ly = f.layers[l];
    hipError_t e = conditioner(f, l, B, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_coupling_inv, dim3((B + 3) / 4), dim3(T256), 0, st, s.x, f.P, s.O,
                       round_up(2 * ly.na, FLOW_PAD), f.params + ly.ob2, ly.aidx, ly.na, f.clamp, s.s[l], s.xa[l],
                       s.logdet, B);
  }
  if (f.factor)
    hipLaunchKernelGGL(k_factor_inv, dim3(nblk(B)), dim3(T256), 0, st, s.x, f.params, f.P, z, B);
  else
    hipLaunchKernelGGL(k_base_inv, dim3(nblk(n)), dim3(T256), 0, st, s.x, f.params, f.P, z, n);
  if (logdet_inv && logdet_inv != s.logdet) {
    hipError_t e = hipMemcpyAsync(logdet_inv, s.logdet, (size_t)B * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t flow_loss_backward(Flow& f, const float* z, int B, hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  const long long n = (long long)B * f.P;
  hipLaunchKernelGGL(k_base_inv_bwd, dim3(nblk(n)), dim3(T256), 0, st, z, f.params, f.P, s.g, n);
  // mu and log-scale gradients: per-element terms summed over rows (s.x is free
  // once the inverse is done: the backward reads the per-layer activations)
  float* tmu = s.x;
  float* tls = s.aux;
  hipLaunchKernelGGL(k_base_param_terms, dim3(nblk(n)), dim3(T256), 0, st, z, f.params, f.P, tmu, tls, n);
  colsum(tmu, B, f.P, f.grads, s.colsum, st);
  colsum(tls, B, f.P, f.grads + f.P, s.colsum, st);
  for (int l = 0; l < f.L; ++l) {
    FlowLayer& ly = f.layers[l];
    const int op = round_up(2 * ly.na, FLOW_PAD);
    hipLaunchKernelGGL(k_coupling_inv_bwd, dim3((B + 3) / 4), dim3(T256), 0, st, s.g, f.P, s.s[l], s.xa[l], ly.aidx,
                       ly.na, op, f.clamp, s.gO, s.gOf, B);
    hipError_t e = conditioner_bwd(f, l, B, true, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t flow_adam(Flow& f, int B, double lr, double b1, double b2, double eps, hipStream_t st) {
  f.adam_t += 1;
  const double bc1 = 1.0 - pow(b1, (double)f.adam_t), bc2 = 1.0 - pow(b2, (double)f.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(f.nparams)), dim3(T256), 0, st, f.params, f.grads, f.adam_m, f.adam_v,
                     f.nparams, (float)lr, (float)b1, (float)b2, (float)eps, (float)bc1, (float)bc2,
                     1.0f / (float)B);
  return hipGetLastError();
}

hipError_t flow_normal_fill(float* z, int B, int P, uint32_t seed, uint32_t step, hipStream_t st) {
  const long long n = (long long)B * P;
  hipLaunchKernelGGL(k_normal_fill, dim3(nblk(n)), dim3(T256), 0, st, z, n, seed, step);
  return hipGetLastError();
}

hipError_t flow_elbo_rows(Flow& f, const float* z, const float* lj, int B, float* rw, double* logw,
                          hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_elbo_rows, dim3((B + 3) / 4), dim3(T256), 0, st, z, lj, f.sc.logdet, f.sc.g, B, f.P, rw,
                     logw);
  return hipGetLastError();
}

hipError_t flow_elbo_backward(Flow& f, const float* z, const float* rw, int B, hipStream_t st) {
  if (B % FLOW_PAD || B > f.sc.B) return hipErrorInvalidValue;
  FlowScratch& s = f.sc;
  const long long n = (long long)B * f.P;
  for (int l = f.L - 1; l >= 0; --l) {
    FlowLayer& ly = f.layers[l];
    const int op = round_up(2 * ly.na, FLOW_PAD);
    hipLaunchKernelGGL(k_coupling_fwd_bwd, dim3((B + 3) / 4), dim3(T256), 0, st, s.g, f.P, s.s[l], s.xa[l], ly.aidx,
                       ly.na, op, f.clamp, s.gO, s.gOf, B, rw);
    hipError_t e = conditioner_bwd(f, l, B, true, st);
    if (e != hipSuccess) return e;
  }
  // s.g is now d/dx0: mu's gradient is its column sum; log-scale terms g e^ls z + rw
  colsum(s.g, B, f.P, f.grads, s.colsum, st);
  hipLaunchKernelGGL(k_base_fwd_param_terms, dim3(nblk(n)), dim3(T256), 0, st, z, s.g, f.params, f.P, rw, s.aux, n);
  colsum(s.aux, B, f.P, f.grads + f.P, s.colsum, st);
  if (f.factor) {  // d/dF[i, j] = sum over rows of g_i z_j, j < i
    const long long PP = (long long)f.P * f.P;
    hipLaunchKernelGGL(k_factor_grad_part, dim3(nblk(PP), f.fsplits), dim3(T256), 0, st, s.g, z, B, f.P,
                       f.fpart);
    hipLaunchKernelGGL(k_factor_grad_final, dim3((unsigned)((PP + 31) / 32)), dim3(T256), 0, st, f.fpart,
                       f.fsplits, PP, f.grads + f.ofac);
  }
  return hipGetLastError();
}

hipError_t flow_elbo_stats(Flow& f, const float* rw, const double* logw, int B, const double* gstat, int ngroups,
                           double clip, double* part, double* stats, hipStream_t st) {
  constexpr int NPART = 256;
  hipLaunchKernelGGL(k_sumsq_part, dim3(NPART), dim3(T256), 0, st, f.grads, f.nparams, part);
  hipLaunchKernelGGL(k_elbo_stats, dim3(1), dim3(T256), 0, st, rw, logw, B, part, NPART, gstat, ngroups, clip,
                     stats);
  return hipGetLastError();
}

hipError_t flow_iwae_weights(Flow& f, const float* rw, const double* logw, int B, int K, float* wts, double* gstat,
                             hipStream_t st) {
  if (K < 1 || B % K) return hipErrorInvalidValue;
  const int ngroups = B / K;
  hipLaunchKernelGGL(k_iwae_groups, dim3((ngroups + 3) / 4), dim3(T256), 0, st, rw, logw, K, ngroups, wts, gstat);
  const long long n = (long long)B * f.P;
  hipLaunchKernelGGL(k_scale_rows, dim3(nblk(n)), dim3(T256), 0, st, f.sc.g, wts, f.P, n);
  return hipGetLastError();
}

hipError_t flow_adam_ascent(Flow& f, const double* stats, double lr, double b1, double b2, double eps,
                            hipStream_t st, const ElboTrack* track) {
  f.adam_t += 1;
  const double bc1 = 1.0 - pow(b1, (double)f.adam_t), bc2 = 1.0 - pow(b2, (double)f.adam_t);
  hipLaunchKernelGGL(k_adam_ascent, dim3(nblk(f.nparams)), dim3(T256), 0, st, f.params, f.grads, f.adam_m,
                     f.adam_v, f.nparams, (float)lr, (float)b1, (float)b2, (float)eps, (float)bc1, (float)bc2,
                     stats, track);
  return hipGetLastError();
}

hipError_t flow_elbo_track(Flow& f, const double* stats, int it, double* hist, ElboTrack* track, float* best_w,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_elbo_track, dim3(1), dim3(64), 0, st, stats, it, hist, track);
  hipLaunchKernelGGL(k_copy_if_best, dim3(nblk(f.nparams)), dim3(T256), 0, st, best_w, f.params, f.nparams,
                     (const ElboTrack*)track);
  return hipGetLastError();
}

hipError_t flow_gather_rows(const float* src, long long n, int P, float* dst, int B, uint32_t seed, uint32_t step,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)B * P)), dim3(T256), 0, st, src, n, P, dst, B, seed, step);
  return hipGetLastError();
}

hipError_t flow_add_logdet(float* lj, const float* logdet, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_add_logdet, dim3(nblk(B)), dim3(T256), 0, st, lj, logdet, B);
  return hipGetLastError();
}
hipError_t flow_sub_logdet(double* lj, const float* logdet, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_logdet, dim3(nblk(B)), dim3(T256), 0, st, lj, logdet, B);
  return hipGetLastError();
}

hipError_t flow_loss_value(const float* z, const float* logdet_inv, int B, int P, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_row_loss, dim3((B + 3) / 4), dim3(T256), 0, st, z, logdet_inv, B, P, out);
  return hipGetLastError();
}

}  // namespace ut
