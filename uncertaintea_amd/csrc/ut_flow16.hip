// ut_flow16.hip -- the sampling path's flow kernels on v_mfma_f32_16x16x32_bf16 for flows whose
// coupling layers have at most 64 passive and 64 active coordinates (P <= 128) and H = 256
// (the funnel's P = 101 preconditioner).
//
// One launch per direction carries a workgroup's 16 rows through every layer; eight
// wavefronts, two per SIMD.  Every conditioner GEMM is computed transposed (weights as the A
// operand, held in registers from fragment-order copies; the 16 rows as the N dimension), so a
// lane's accumulator holds four consecutive hidden units (or outputs) of one row:
//
//   forward, layer l    U^T[256 x 16] = W1 X_p^T       (wavefront w: hidden 32 w .. 32 w + 31)
//                       h = tanh(U + b1) -> LDS (bf16, 8-byte stores)
//                       O^T[128 x 16] = W2' h^T         (wavefront w: outputs 16 w .. 16 w + 15)
//                       coupling in registers: W2' interleaves W2's rows so that a lane holds
//                       (s_j, t_j, s_j+1, t_j+1) of one row, j = 8 w + 2 (lane / 16)
//   backward, layer l   coupling backward in registers -> dO (bf16) in the same interleaved order
//                       dh^T = W2'^T dO^T, with U recomputed from the stored bf16 passive
//                       inputs (exactly the forward's arithmetic), dU = dh (1 - h^2) -> LDS
//                       dX_p^T[64 x 16] = W1^T dU^T, added to the passive coordinates
//
// The forward stores per layer only what the backward cannot recompute cheaply: the bf16
// passive inputs (in the B-operand fragment order, 2 KB per workgroup), s and x_a e^s (fp32);
// the hidden activations never leave the chip.  Two workgroup barriers per forward layer, three
// per backward layer.  The padded passive width is 64 (two 32-deep k-steps), so no multiply runs
// on the zero padding of the 128-wide parameter layout.  The log-determinant of the base
// (sum of the log-scales) is summed once per weight update (k_ldbase16), each lane keeps its
// coordinates' share of the coupling log-determinants across layers, and the row sums are
// formed once at the end.
//
// Arithmetic: the same operations as the per-step kernels (bf16 operands, fp32 accumulation,
// flow_tanh, the clamp c tanh(s / c)), summed in the orders of this kernel's tiles, so results
// agree with the per-step path to fp32 rounding (and the occasional bf16 rounding flip of an
// operand), not bit for bit.
#include <algorithm>
#include <type_traits>

#include "ut_flow.h"
#include "ut_math.h"
#include "ut_mfma.h"
#include "ut_models.h"

namespace ut {

namespace {

#define UT_GLOBAL16 __attribute__((address_space(1)))

struct NoRowModel {};
template <typename T>
__device__ inline UT_GLOBAL16 T* g16(T* p) {
  return (UT_GLOBAL16 T*)p;
}
template <typename T>
__device__ inline const UT_GLOBAL16 T* g16(const T* p) {
  return (const UT_GLOBAL16 T*)p;
}

constexpr int R16 = 16;          // rows per workgroup
constexpr int NW16 = 8;          // wavefronts per workgroup
constexpr int THREADS16 = 64 * NW16;
constexpr int HID16 = 256;       // hidden width
constexpr int KP16 = 64;         // padded passive / active width
constexpr int XPS16 = KP16 + 8;  // LDS row stride (bf16) of the passive-input tile
constexpr int HS16 = HID16 + 8;  // of the hidden tile
constexpr int OS16 = 2 * KP16 + 8;  // of the conditioner-output gradient tile
// fragment-order weight copies, elements per layer
constexpr int W1F16 = NW16 * 2 * 2 * 512;   // [wave][tile][k-step][lane][8]
constexpr int W2F16 = NW16 * 8 * 512;       // [wave][k-step][lane][8]
constexpr int W2TF16 = NW16 * 2 * 4 * 512;  // [wave][tile][k-step][lane][8]
constexpr int W1TF16 = 4 * 8 * 512;         // [wave < 4][k-step][lane][8]

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// s_waitcnt lgkmcnt(0) + s_barrier: the kernels exchange data between wavefronts only through
// LDS, so global loads and stores keep flying across the barrier
__device__ __forceinline__ void lds_barrier16() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ inline uint32_t pack2(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}

// W2' row o (the interleaved output order s_0, t_0, s_1, t_1, ...) -> W2 row, or -1 past 2 na
__device__ inline int w2_row(int o, int na) {
  const int j = o >> 1;
  return j >= na ? -1 : ((o & 1) ? na + j : j);
}

// fragment-order bf16 copies of one layer's weights (W1: H x kp, W2: op x H, fp32 masters)
__global__ void k_frag16(const float* __restrict__ W1, int kp, const float* __restrict__ W2, int np, int na,
                         bf16_t* __restrict__ W1f, bf16_t* __restrict__ W2f, bf16_t* __restrict__ W2Tf,
                         bf16_t* __restrict__ W1Tf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = i & 7, lane = (i >> 3) & 63, lr = lane & 15, lg = lane >> 4;
  if (i < W1F16) {  // A operand of U^T: W1[hidden][k]
    const int ks = (i >> 9) & 1, t = (i >> 10) & 1, w = i >> 11;
    const int hid = 32 * w + 16 * t + lr, k = 32 * ks + 8 * lg + e;
    W1f[i] = k < np ? f2bf(W1[(long long)hid * kp + k]) : (bf16_t)0;
    return;
  }
  i -= W1F16;
  if (i < W2F16) {  // A operand of O^T: W2'[o][hidden]
    const int ks = (i >> 9) & 7, w = i >> 12;
    const int r = w2_row(16 * w + lr, na), hid = 32 * ks + 8 * lg + e;
    W2f[i] = r < 0 ? (bf16_t)0 : f2bf(W2[(long long)r * HID16 + hid]);
    return;
  }
  i -= W2F16;
  if (i < W2TF16) {  // A operand of dh^T: W2'^T[hidden][o]
    const int ks = (i >> 9) & 3, t = (i >> 11) & 1, w = i >> 12;
    const int hid = 32 * w + 16 * t + lr, r = w2_row(32 * ks + 8 * lg + e, na);
    W2Tf[i] = r < 0 ? (bf16_t)0 : f2bf(W2[(long long)r * HID16 + hid]);
    return;
  }
  i -= W2TF16;
  if (i < W1TF16) {  // A operand of dX_p^T: W1^T[k][hidden]
    const int ks = (i >> 9) & 7, w = i >> 12;
    const int k = 16 * w + lr, hid = 32 * ks + 8 * lg + e;
    W1Tf[i] = k < np ? f2bf(W1[(long long)hid * kp + k]) : (bf16_t)0;
  }
}

// sum of the base log-scales in index order (the order of k_logdet_init)
__global__ void k_ldbase16(const float* __restrict__ ls, int P, float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (int i = 0; i < P; ++i) s += ls[i];
  out[0] = s;
}

struct Fwd16Regs {
  bf16x8 a1[2][2];  // W1 tiles x k-steps
  bf16x8 a2[8];     // W2' k-steps
  float b1[2][4];
  float b2s[2], b2t[2];
};
// the first GEMM's weights and biases of layer `ly`
__device__ inline void fwd16_load1(const Flow16Layer& ly, Fwd16Regs& g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lg = lane >> 4;
  const UT_GLOBAL16 bf16_t* W1f = g16(ly.W1f);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
      g.a1[t][ks] = *(const UT_GLOBAL16 bf16x8*)(W1f + (((w * 2 + t) * 2 + ks) * 64 + lane) * 8);
  const UT_GLOBAL16 float* b1 = g16(ly.b1);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) g.b1[t][r] = b1[32 * w + 16 * t + 4 * lg + r];
}
// the second GEMM's weights and the coupling biases of this lane's coordinates (none on a
// wavefront whose outputs are all padding)
__device__ inline void fwd16_load2(const Flow16Layer& ly, Fwd16Regs& g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lg = lane >> 4;
  if (16 * __builtin_amdgcn_readfirstlane(w) >= 2 * ly.na) return;
  const UT_GLOBAL16 bf16_t* W2f = g16(ly.W2f);
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) g.a2[ks] = *(const UT_GLOBAL16 bf16x8*)(W2f + ((w * 8 + ks) * 64 + lane) * 8);
  const UT_GLOBAL16 float* b2 = g16(ly.b2);
  const int j0 = 8 * w + 2 * lg;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int j = min(j0 + e, ly.na - 1);  // the padded bias vector holds >= 2 na entries
    g.b2s[e] = b2[j];
    g.b2t[e] = b2[ly.na + j];
  }
}

// Where the forward keeps what the backward reads again: per layer the clamped scales s and
// the scaled active inputs x_a e^s of each lane's coordinates, and the passive-input
// fragments.  GlobalSaved: the layer's HBM buffers (row stride 64; the separate forward and
// backward kernels); LdsSaved: LDS of the fused gradient kernel, per layer 1024 of each in
// the order of the lanes that write and read them (coordinate j = 8 w + 2 lg + e of row lr
// at e 512 + 64 w + 16 lg + lr: consecutive lanes, consecutive banks), and 2 x 64 x 8 bf16
// fragments.  LdsSaved also keeps the hidden activations h (fp32, exactly the forward's
// values) of the layers from h_from on, in the order of the lanes that compute them, so the
// backward reads them instead of recomputing U and tanh (and those layers need no passive-input
// fragments); its s and x_a e^s hold `slots` floats per layer, 128 per wavefront with outputs.
struct GlobalSaved {
  __device__ void put(const Flow16Layer& ly, int, long long r0, int lr, int j, float sv, float xs) const {
    g16(ly.s)[(r0 + lr) * KP16 + j] = sv;
    g16(ly.xs)[(r0 + lr) * KP16 + j] = xs;
  }
  __device__ void put_frag(const Flow16Layer& ly, int, int ks, int lane, const bf16x8& v) const {
    *(UT_GLOBAL16 bf16x8*)(g16(ly.xpg) + (((long long)blockIdx.x * 2 + ks) * 64 + lane) * 8) = v;
  }
  __device__ void get(const Flow16Layer& ly, int, long long r0, int lr, int j, float& sv, float& xs) const {
    sv = g16(ly.s)[(r0 + lr) * KP16 + j];
    xs = g16(ly.xs)[(r0 + lr) * KP16 + j];
  }
  __device__ bf16x8 get_frag(const Flow16Layer& ly, int, int ks, int lane) const {
    return *(const UT_GLOBAL16 bf16x8*)(g16(ly.xpg) + (((long long)blockIdx.x * 2 + ks) * 64 + lane) * 8);
  }
  __device__ bool keeps_h(int) const { return false; }
  __device__ void put_h(int, int, int, int, const f32x4&) const {}
  __device__ f32x4 get_h(int, int, int, int) const { return (f32x4){0.0f, 0.0f, 0.0f, 0.0f}; }
};
constexpr int SAVED16_BF = 2 * 64 * 8;   // bf16 of the passive-input fragments per layer
constexpr int SAVED16_H = R16 * HID16;   // floats of h per kept layer
struct LdsSaved {
  float* s;
  float* xs;
  bf16_t* xp;  // layers below h_from
  float* h;    // layers from h_from on
  int h_from;
  int slots;
  __device__ int at(int l, int lr, int j) const {
    return l * slots + (j & 1) * (slots >> 1) + (j >> 3) * 64 + ((j >> 1) & 3) * 16 + lr;
  }
  __device__ void put(const Flow16Layer&, int l, long long, int lr, int j, float sv, float x) const {
    s[at(l, lr, j)] = sv;
    xs[at(l, lr, j)] = x;
  }
  __device__ void put_frag(const Flow16Layer&, int l, int ks, int lane, const bf16x8& v) const {
    *(bf16x8*)(xp + l * SAVED16_BF + (ks * 64 + lane) * 8) = v;
  }
  __device__ void get(const Flow16Layer&, int l, long long, int lr, int j, float& sv, float& x) const {
    sv = s[at(l, lr, j)];
    x = xs[at(l, lr, j)];
  }
  __device__ bf16x8 get_frag(const Flow16Layer&, int l, int ks, int lane) const {
    return *(const bf16x8*)(xp + l * SAVED16_BF + (ks * 64 + lane) * 8);
  }
  __device__ bool keeps_h(int l) const { return l >= h_from; }
  // hidden units 32 w + 16 t + 4 lg .. + 3 of row lr
  __device__ void put_h(int l, int w, int t, int lane, const f32x4& v) const {
    *(f32x4*)(h + (l - h_from) * SAVED16_H + ((w * 2 + t) * 64 + lane) * 4) = v;
  }
  __device__ f32x4 get_h(int l, int w, int t, int lane) const {
    return *(const f32x4*)(h + (l - h_from) * SAVED16_H + ((w * 2 + t) * 64 + lane) * 4);
  }
};

// The forward through every layer of one workgroup's 16 rows: the base x = mu + e^ls z into Xs,
// then per layer U^T = W1 X_p^T, h = tanh(U + b1) (Hs), O^T = W2' h^T and the coupling
// x_a <- x_a e^s + t; returns this lane's share of its row's coupling log-determinants.
template <int LT, class SAVE>
__device__ __forceinline__ float flow16_fwd_body(const float* __restrict__ z, const float* __restrict__ prm, int P,
                                                 const Flow16Layer* __restrict__ layers, int L, float c, float* Xs,
                                                 bf16_t* Xp, bf16_t* Hs, const SAVE& save) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const long long r0 = (long long)blockIdx.x * R16;
  const int j0 = 8 * w + 2 * lg;  // this lane's coupling coordinates j0, j0 + 1 (row lr)
  const int wu = __builtin_amdgcn_readfirstlane(w);  // the wavefront index as a scalar (uniform branches)
  // weights in two register sets: layer l computes with one while layer l + 1's are in flight
  Fwd16Regs ga, gb;
  fwd16_load1(layers[0], ga);
  fwd16_load2(layers[0], ga);
  if (L > 1) {
    fwd16_load1(layers[1], gb);
    fwd16_load2(layers[1], gb);
  }
  // base: x = mu + e^ls z
  for (int e0 = tid; e0 < R16 * P; e0 += 4 * THREADS16) {
    float zv[4], mv[4], lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = min(e0 + THREADS16 * k, R16 * P - 1);
      const int ci = e % P;
      zv[k] = z[r0 * P + e];
      mv[k] = prm[ci];
      lv[k] = prm[P + ci];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + THREADS16 * k;
      if (e < R16 * P) Xs[e] = mv[k] + expf(lv[k]) * zv[k];
    }
  }
  lds_barrier16();
  {  // layer 0's passive inputs
    const int np0 = layers[0].np, par0 = layers[0].par;
    for (int e = tid; e < R16 * KP16; e += THREADS16) {
      const int row = e / KP16, k = e % KP16;
      Xp[row * XPS16 + k] = k < np0 ? f2bf(Xs[row * P + 2 * k + par0]) : (bf16_t)0;
    }
  }
  lds_barrier16();
  float ldp = 0.0f;  // this lane's share of its row's coupling log-determinants
  // one layer with weights `g`; afterwards `g` receives layer l + 2's weights
  auto layer = [&](int l, Fwd16Regs& g) {
    const Flow16Layer ly = layers[l];
    const bool next2 = l + 2 < L;
    const Flow16Layer& nx = layers[next2 ? l + 2 : l];
    const int na = ly.na, par = ly.par;
    // ---- U^T = W1 X_p^T: B operand = 8 passive inputs of row lr.  Every wavefront holds the
    // same fragments; wavefront 0 saves them for the backward's recomputation of U
    bf16x8 xb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) xb[ks] = *(const bf16x8*)(Xp + lr * XPS16 + 32 * ks + 8 * lg);
    if (wu == 0 && !save.keeps_h(l)) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) save.put_frag(ly, l, ks, lane, xb[ks]);
    }
    f32x4 u[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      u[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) u[t] = mfma16(g.a1[t][ks], xb[ks], u[t]);
    }
    float cb1[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) cb1[t][r] = g.b1[t][r];
    __builtin_amdgcn_sched_barrier(0);
    if (next2) fwd16_load1(nx, g);  // a layer and a half ahead
    // ---- h = tanh(U + b1): hidden 32 w + 16 t + 4 lg .. + 3 of row lr
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      float hv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) hv[r] = flow_tanh(u[t][r] + cb1[t][r]);
      if (save.keeps_h(l)) save.put_h(l, w, t, lane, (f32x4){hv[0], hv[1], hv[2], hv[3]});
      uint2 pk;
      pk.x = pack2(hv[0], hv[1]);
      pk.y = pack2(hv[2], hv[3]);
      *(uint2*)(Hs + lr * HS16 + 32 * w + 16 * t + 4 * lg) = pk;
    }
    lds_barrier16();  // h complete; every wavefront is done with X_p
    // ---- O^T = W2' h^T (outputs 16 w .. 16 w + 15: coordinates j0, j0 + 1) and the coupling
    // x_a <- x_a e^s + t, on the wavefronts whose outputs exist
    float xn[2] = {0.0f, 0.0f};
    if (16 * wu < 2 * na) {
      f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k0 = 0; k0 < 8; k0 += 4) {  // B fragments four at a time (register pressure)
        bf16x8 hb[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) hb[ks] = *(const bf16x8*)(Hs + lr * HS16 + 32 * (k0 + ks) + 8 * lg);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) o = mfma16(g.a2[k0 + ks], hb[ks], o);
      }
      const float cb2s[2] = {g.b2s[0], g.b2s[1]}, cb2t[2] = {g.b2t[0], g.b2t[1]};
      __builtin_amdgcn_sched_barrier(0);
      if (next2) fwd16_load2(nx, g);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int j = j0 + e;
        if (j < na) {
          const int ak = 2 * j + 1 - par;
          const float sv = c * flow_tanh((o[2 * e] + cb2s[e]) * (1.0f / c));
          const float xs = Xs[lr * P + ak] * expf(sv);
          xn[e] = xs + (o[2 * e + 1] + cb2t[e]);
          Xs[lr * P + ak] = xn[e];
          ldp += sv;
          save.put(ly, l, r0, lr, j, sv, xs);  // for the backward
        }
      }
    } else if (next2) {
      fwd16_load2(nx, g);
    }
    // the next layer's passive inputs are this layer's active outputs (zero past na)
    *(uint32_t*)(Xp + lr * XPS16 + j0) = pack2(xn[0], xn[1]);
    lds_barrier16();
  };
  if (LT > 0) {
#pragma unroll
    for (int l = 0; l < LT; l += 2) {
      layer(l, ga);
      if (l + 1 < LT) layer(l + 1, gb);
    }
  } else {
    for (int l = 0; l < L; l += 2) {
      layer(l, ga);
      if (l + 1 < L) layer(l + 1, gb);
    }
  }
  return ldp;
}

// the row model on the 16 rows in Xs: one wavefront per row, the density kernel's loads,
// arithmetic and stores (gradient rows to `g` with row stride P, log densities to lj)
template <class RM>
__device__ __forceinline__ void flow16_row_model(const RM& rm, const float* Xs, int P, float* g, float* lj) {
  const int lane = threadIdx.x & 63;
  const int wu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  typename RM::template Lane<float> ld;
  rm.prepare(ld);
  for (int rr = wu; rr < R16; rr += NW16) {
    float qv[RM::K], gv[RM::K];
#pragma unroll
    for (int k = 0; k < RM::K; ++k) {
      const int a = rm.api(lane, k);
      qv[k] = a >= 0 ? Xs[rr * P + a] : 0.0f;
    }
    bool ok;
    const float val = rm.eval(ld, qv, gv, ok);
#pragma unroll
    for (int k = 0; k < RM::K; ++k) {
      const int a = rm.api(lane, k);
      if (a >= 0) g[rr * P + a] = gv[k];
    }
    if (lane == 0) lj[rr] = val;
  }
}

template <int LT, class RM>
__global__ __launch_bounds__(THREADS16) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_flow16_fwd(
    const float* __restrict__ z, const float* __restrict__ prm, const float* __restrict__ ldbase,
    float* __restrict__ x, float* __restrict__ q, int P, const Flow16Layer* __restrict__ layers, int L_, float c,
    float* __restrict__ logdet, float* __restrict__ logdet_out, RM rm, float* __restrict__ g_rm,
    float* __restrict__ lj_rm) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lds16[];
  float* Xs = lds16;                          // 16 x P fp32 rows
  bf16_t* Xp = (bf16_t*)(Xs + R16 * P);       // 16 x XPS16 passive inputs (bf16)
  bf16_t* Hs = Xp + R16 * XPS16;              // 16 x HS16 hidden activations (bf16)
  float* Red = (float*)(Hs + R16 * HS16);     // 8 x 16 log-determinant partials
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15;
  const long long r0 = (long long)blockIdx.x * R16;
  float ldp = flow16_fwd_body<LT>(z, prm, P, layers, L, c, Xs, Xp, Hs, GlobalSaved{});
  // log-determinant: the base's, plus the row sums of the lanes' shares
  ldp += __shfl_xor(ldp, 16);
  ldp += __shfl_xor(ldp, 32);
  if (lane < 16) Red[w * 16 + lr] = ldp;
  lds_barrier16();
  if (tid < R16) {
    float s = ldbase[0];
#pragma unroll
    for (int ww = 0; ww < NW16; ++ww) s += Red[ww * 16 + tid];
    logdet[r0 + tid] = s;
    if (logdet_out) logdet_out[r0 + tid] = s;
  }
  for (int e = tid; e < R16 * P; e += THREADS16) {
    const float v = Xs[e];
    x[r0 * P + e] = v;
    if (q) q[r0 * P + e] = v;
  }
  if constexpr (!std::is_same<RM, NoRowModel>::value) flow16_row_model(rm, Xs, P, g_rm + r0 * P, lj_rm + r0);
}

struct Bwd16Regs {
  bf16x8 a1[2][2];   // W1 (recompute of U)
  bf16x8 a2t[2][4];  // W2'^T tiles x k-steps
  bf16x8 a1t[8];     // W1^T k-steps (wavefront w & 3's tile)
  float b1[2][4];
};
// the weights of layer `ly`; W1 (for the recomputation of U) only where `need_a1`
__device__ inline void bwd16_load(const Flow16Layer& ly, Bwd16Regs& g, bool need_a1) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lg = lane >> 4;
  const UT_GLOBAL16 bf16_t* W1f = g16(ly.W1f);
  const UT_GLOBAL16 bf16_t* W2Tf = g16(ly.W2Tf);
  const UT_GLOBAL16 bf16_t* W1Tf = g16(ly.W1Tf);
  if (need_a1)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
        g.a1[t][ks] = *(const UT_GLOBAL16 bf16x8*)(W1f + (((w * 2 + t) * 2 + ks) * 64 + lane) * 8);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      g.a2t[t][ks] = *(const UT_GLOBAL16 bf16x8*)(W2Tf + (((w * 2 + t) * 4 + ks) * 64 + lane) * 8);
  }
  if (w < 4)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) g.a1t[ks] = *(const UT_GLOBAL16 bf16x8*)(W1Tf + ((w * 8 + ks) * 64 + lane) * 8);
  const UT_GLOBAL16 float* b1 = g16(ly.b1);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) g.b1[t][r] = b1[32 * w + 16 * t + 4 * lg + r];
}

// The backward's registers carried from its prologue into the layers: the weights of the
// last two layers and the last layer's saved values of this lane's coordinates.
struct Bwd16Carry {
  Bwd16Regs ga, gb;
  float sv[2], xsv[2];
  bf16x8 xb[2];
};
template <class SAVE>
__device__ __forceinline__ void flow16_bwd_prologue(const Flow16Layer* __restrict__ layers, int L, const SAVE& save,
                                                    Bwd16Carry& cy) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const long long r0 = (long long)blockIdx.x * R16;
  const int j0 = 8 * w + 2 * lg;
  bwd16_load(layers[L - 1], cy.ga, !save.keeps_h(L - 1));
  if (L > 1) bwd16_load(layers[L - 2], cy.gb, !save.keeps_h(L - 2));
#pragma unroll
  for (int e = 0; e < 2; ++e) save.get(layers[L - 1], L - 1, r0, lr, j0 + e, cy.sv[e], cy.xsv[e]);
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
    cy.xb[ks] = save.keeps_h(L - 1) ? (bf16x8){} : save.get_frag(layers[L - 1], L - 1, ks, lane);
}

// The backward through every layer (gradient rows in Gs, in place): coupling backward,
// dh^T = W2'^T dO^T with U recomputed from the saved passive inputs, dU = dh (1 - h^2),
// dX_p^T = W1^T dU^T.
template <int LT, class SAVE>
__device__ __forceinline__ void flow16_bwd_body(int P, const Flow16Layer* __restrict__ layers, int L, float c,
                                                float* Gs, bf16_t* dO, bf16_t* DU, const SAVE& save, Bwd16Carry& cy) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const long long r0 = (long long)blockIdx.x * R16;
  const int j0 = 8 * w + 2 * lg;
  const int wu = __builtin_amdgcn_readfirstlane(w);  // the wavefront index as a scalar (uniform branches)
  (void)tid;
  // one layer (l = L - 1 - li) with weights `g`; afterwards `g` receives layer l - 2's weights
  auto layer = [&](int li, Bwd16Regs& g) {
    const int l = L - 1 - li;
    const Flow16Layer ly = layers[l];
    const bool next = l > 0;
    const Flow16Layer& nx = layers[next ? l - 1 : l];
    const Flow16Layer& nx2 = layers[l > 1 ? l - 2 : l];
    const int np = ly.np, na = ly.na, par = ly.par;
    // ---- coupling backward: x_a <- x_a e^s + t, log-determinant + s, s = c tanh(s_raw / c)
    float d[4];  // (ds_raw, dt) of j0, j0 + 1, zero past na
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int j = j0 + e;
      const bool valid = j < na;
      const int ak = valid ? 2 * j + 1 - par : 0;
      const float ga = valid ? Gs[lr * P + ak] : 0.0f;
      const float r = cy.sv[e] * (1.0f / c);
      d[2 * e] = valid ? (ga * cy.xsv[e] + 1.0f) * (1.0f - r * r) : 0.0f;
      d[2 * e + 1] = ga;
      if (valid) Gs[lr * P + ak] = ga * expf(cy.sv[e]);
    }
    {
      uint2 pk;
      pk.x = pack2(d[0], d[1]);
      pk.y = pack2(d[2], d[3]);
      *(uint2*)(dO + lr * OS16 + 16 * w + 4 * lg) = pk;
    }
    float sv_n[2], xsv_n[2];
    if (next) {
#pragma unroll
      for (int e = 0; e < 2; ++e) save.get(nx, l - 1, r0, lr, j0 + e, sv_n[e], xsv_n[e]);
    }
    lds_barrier16();
    // ---- dh^T = W2'^T dO^T; U recomputed from the forward's passive inputs; dU = dh (1 - h^2)
    bf16x8 ob[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) ob[ks] = *(const bf16x8*)(dO + lr * OS16 + 32 * ks + 8 * lg);
    f32x4 dh[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      dh[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) dh[t] = mfma16(g.a2t[t][ks], ob[ks], dh[t]);
    }
    // h: the forward's values where kept, else recomputed with the forward's arithmetic
    f32x4 hk[2];
    if (save.keeps_h(l)) {
#pragma unroll
      for (int t = 0; t < 2; ++t) hk[t] = save.get_h(l, w, t, lane);
    } else {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        f32x4 u = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) u = mfma16(g.a1[t][ks], cy.xb[ks], u);
#pragma unroll
        for (int r = 0; r < 4; ++r) hk[t][r] = flow_tanh(u[r] + g.b1[t][r]);
      }
    }
    bf16x8 xb_n[2] = {};
    __builtin_amdgcn_sched_barrier(0);
    if (next && !save.keeps_h(l - 1)) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) xb_n[ks] = save.get_frag(nx, l - 1, ks, lane);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      float du[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float hv = hk[t][r];
        du[r] = dh[t][r] * (1.0f - hv * hv);
      }
      uint2 pk;
      pk.x = pack2(du[0], du[1]);
      pk.y = pack2(du[2], du[3]);
      *(uint2*)(DU + lr * HS16 + 32 * w + 16 * t + 4 * lg) = pk;
    }
    lds_barrier16();
    // ---- dX_p^T = W1^T dU^T: passive coordinates 16 w + 4 lg .. + 3 of row lr
    if (16 * wu < np) {
      f32x4 gp = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k0 = 0; k0 < 8; k0 += 4) {  // B fragments four at a time (register pressure)
        bf16x8 ub[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) ub[ks] = *(const bf16x8*)(DU + lr * HS16 + 32 * (k0 + ks) + 8 * lg);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) gp = mfma16(g.a1t[k0 + ks], ub[ks], gp);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 16 * w + 4 * lg + r;
        if (k < np) Gs[lr * P + 2 * k + par] += gp[r];
      }
    }
    if (l > 1) bwd16_load(nx2, g, !save.keeps_h(l - 2));
    lds_barrier16();
    if (next) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        cy.sv[e] = sv_n[e];
        cy.xsv[e] = xsv_n[e];
        cy.xb[e] = xb_n[e];
      }
    }
  };
  if (LT > 0) {
#pragma unroll
    for (int li = 0; li < LT; li += 2) {
      layer(li, cy.ga);
      if (li + 1 < LT) layer(li + 1, cy.gb);
    }
  } else {
    for (int li = 0; li < L; li += 2) {
      layer(li, cy.ga);
      if (li + 1 < L) layer(li + 1, cy.gb);
    }
  }
}

// base backward g_z = g_x e^ls of the workgroup's rows, and with STEP the lock-step
// leapfrog's full kick and drift (k_ls_step): rows whose gradient is finite move, a chain
// with a non-finite entry is marked invalid, invalid chains stay.  With STEP the gradient is
// used only by the kick, so g_z is not written (the trajectory's last gradient, without
// STEP, is the one the samplers read).
template <bool STEP>
__device__ __forceinline__ void flow16_bwd_base(const float* __restrict__ prm, int P, float* Gs, float* gz,
                                                const Flow16Step& stp) {
  const int tid = threadIdx.x;
  const long long r0 = (long long)blockIdx.x * R16;
  __shared__ int row_ok[R16];
  // with STEP, this thread's momenta, positions, inverse masses and chain flags (elements
  // tid + 512 k, P <= 128) are loaded first, so their latency overlaps the base backward; the
  // barriers below wait for LDS only
  constexpr int EPT = R16 * 128 / THREADS16;
  float pw[EPT], pq[EPT], pim[EPT];
  bool pok[EPT];
  if (STEP) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = min(tid + THREADS16 * k, R16 * P - 1);
      const int row = e / P, a = e - row * P;
      const long long cr = r0 + row;
      pw[k] = stp.wp[cr * P + a];
      pq[k] = stp.wq[cr * P + a];
      pim[k] = stp.im_f[a];
      pok[k] = stp.valid[cr] != 0;
    }
    if (tid < R16) row_ok[tid] = 1;
    lds_barrier16();
  }
  for (int e0 = tid; e0 < R16 * P; e0 += 4 * THREADS16) {
    float lv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) lv[k] = prm[P + min(e0 + THREADS16 * k, R16 * P - 1) % P];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + THREADS16 * k;
      if (e < R16 * P) {
        const float gv = Gs[e] * expf(lv[k]);
        if (STEP) {
          Gs[e] = gv;  // the kick's gradient, read back by this thread below
          if (!isfinite_f(gv)) row_ok[e / P] = 0;
        } else {
          gz[r0 * P + e] = gv;
        }
      }
    }
  }
  if (STEP) {
    lds_barrier16();
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = tid + THREADS16 * k;
      if (e < R16 * P) {
        const int row = e / P, a = e - row * P;
        const long long cr = r0 + row;
        if (row_ok[row] && pok[k]) {
          const float gv = Gs[e];
          const float pv = pw[k] + (1.0f * stp.eps) * gv;
          const float qv = pq[k] + (stp.eps * pim[k]) * pv;
          stp.wq[cr * P + a] = qv;
          stp.wp[cr * P + a] = pv;
        }
      }
    }
    if (tid < R16 && !row_ok[tid] && stp.valid[r0 + tid]) stp.valid[r0 + tid] = 0;
  }
}

template <int LT, bool STEP>
__global__ __launch_bounds__(THREADS16) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_flow16_bwd(
    const float* g_q, const float* __restrict__ prm, float* gz, int P, const Flow16Layer* __restrict__ layers, int L_,
    float c, Flow16Step stp) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lds16[];
  float* Gs = lds16;                      // 16 x P fp32 gradient rows
  bf16_t* dO = (bf16_t*)(Gs + R16 * P);   // 16 x OS16 conditioner-output gradient (interleaved order)
  bf16_t* DU = dO + R16 * OS16;           // 16 x HS16 hidden pre-activation gradient
  const long long r0 = (long long)blockIdx.x * R16;
  Bwd16Carry cy;
  flow16_bwd_prologue(layers, L, GlobalSaved{}, cy);
  for (int e = threadIdx.x; e < R16 * P; e += THREADS16) Gs[e] = g_q[r0 * P + e];
  lds_barrier16();
  flow16_bwd_body<LT>(P, layers, L, c, Gs, dO, DU, GlobalSaved{}, cy);
  flow16_bwd_base<STEP>(prm, P, Gs, gz, stp);
}

// The flow-wrapped gradient of a row model, in one kernel per 16-row tile: forward (saved
// values kept in LDS), the row model's gradient into the backward's gradient rows, backward
// and base backward; then with STEP the lock-step leapfrog step (lj_rm: the model's log
// density), without it the log density plus the flow's log-determinant (a trajectory's last
// step).  The arithmetic of k_flow16_fwd + the density kernel + k_flow16_bwd + k_ls_step (or
// k_add_logdet), so the results are theirs.
template <int LT, class RM, bool STEP>
__global__ __launch_bounds__(THREADS16) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_flow16_grad(
    const float* __restrict__ z, const float* __restrict__ prm, const float* __restrict__ ldbase, int P,
    const Flow16Layer* __restrict__ layers, int L_, float c, RM rm, float* __restrict__ lj_rm,
    float* __restrict__ gz, Flow16Step stp, int h_from, int slots) {
  const int L = LT > 0 ? LT : L_;
  extern __shared__ __attribute__((aligned(16))) float lds16[];
  // the forward's bf16 tiles and the backward's share their space (dead once the forward ends);
  // the gradient rows replace the rows in place (the row model reads each element and writes
  // its gradient from the same lane, and nothing reads the rows afterwards)
  float* Xs = lds16;                          // 16 x P fp32 rows, then their gradient rows
  float* Gs = Xs;
  bf16_t* Xp = (bf16_t*)(Xs + R16 * P);       // forward: 16 x XPS16 passive inputs (bf16)
  bf16_t* dO = Xp;                            // backward: 16 x OS16 (OS16 >= XPS16)
  bf16_t* Hs = Xp + R16 * OS16;               // forward: 16 x HS16 hidden activations (bf16)
  bf16_t* DU = Hs;                            // backward: 16 x HS16
  float* Red = (float*)(Hs + R16 * HS16);     // 8 x 16 log-determinant partials
  float* LJs = Red + NW16 * 16;               // 16 model log densities
  float* sav = LJs + R16;                     // L x (s, x_a e^s), fragments below h_from, kept h
  bf16_t* xpf = (bf16_t*)(sav + 2 * L * slots);
  const LdsSaved save{sav, sav + L * slots, xpf, (float*)(xpf + h_from * SAVED16_BF), h_from, slots};
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15;
  const long long r0 = (long long)blockIdx.x * R16;
  float ldp = flow16_fwd_body<LT>(z, prm, P, layers, L, c, Xs, Xp, Hs, save);
  Bwd16Carry cy;
  flow16_bwd_prologue(layers, L, save, cy);  // the backward's first weights fly during the model
  if (STEP) {
    flow16_row_model(rm, Xs, P, Gs, lj_rm + r0);
  } else {
    flow16_row_model(rm, Xs, P, Gs, LJs);
    // log-determinant as k_flow16_fwd forms it, added to the model's log density
    ldp += __shfl_xor(ldp, 16);
    ldp += __shfl_xor(ldp, 32);
    if (lane < 16) Red[w * 16 + lr] = ldp;
  }
  lds_barrier16();
  if (!STEP && tid < R16) {
    float s = ldbase[0];
#pragma unroll
    for (int ww = 0; ww < NW16; ++ww) s += Red[ww * 16 + tid];
    lj_rm[r0 + tid] = LJs[tid] + s;
  }
  flow16_bwd_body<LT>(P, layers, L, c, Gs, dO, DU, save, cy);
  flow16_bwd_base<STEP>(prm, P, Gs, gz, stp);
}

template <class RM>
using Fwd16Fn = void (*)(const float*, const float*, const float*, float*, float*, int, const Flow16Layer*, int, float,
                         float*, float*, RM, float*, float*);
typedef void (*Bwd16Fn)(const float*, const float*, float*, int, const Flow16Layer*, int, float, Flow16Step);
template <class RM>
Fwd16Fn<RM> fwd16_kernel(int L) {
  switch (L) {
    case 2: return k_flow16_fwd<2, RM>;
    case 4: return k_flow16_fwd<4, RM>;
    case 6: return k_flow16_fwd<6, RM>;
    case 8: return k_flow16_fwd<8, RM>;
    default: return k_flow16_fwd<0, RM>;
  }
}
template <bool STEP>
Bwd16Fn bwd16_kernel(int L) {
  switch (L) {
    case 2: return k_flow16_bwd<2, STEP>;
    case 4: return k_flow16_bwd<4, STEP>;
    case 6: return k_flow16_bwd<6, STEP>;
    case 8: return k_flow16_bwd<8, STEP>;
    default: return k_flow16_bwd<0, STEP>;
  }
}
template <class RM>
using Grad16Fn = void (*)(const float*, const float*, const float*, int, const Flow16Layer*, int, float, RM, float*,
                          float*, Flow16Step, int, int);
template <class RM, bool STEP>
Grad16Fn<RM> grad16_kernel(int L) {
  switch (L) {
    case 2: return k_flow16_grad<2, RM, STEP>;
    case 4: return k_flow16_grad<4, RM, STEP>;
    case 6: return k_flow16_grad<6, RM, STEP>;
    case 8: return k_flow16_grad<8, RM, STEP>;
    default: return k_flow16_grad<0, RM, STEP>;
  }
}
static_assert(OS16 >= XPS16, "the backward's output-gradient tile holds the forward's passive-input tile");
// the fused kernel's LDS with the hidden activations of the last nh layers kept and `slots`
// saved s (and x_a e^s) per layer
size_t grad16_lds(int P, int L, int nh, int slots) {
  return (size_t)R16 * P * 4 + (size_t)R16 * OS16 * 2 + (size_t)R16 * HS16 * 2 + (size_t)(NW16 * 16 + R16) * 4 +
         (size_t)L * 2 * slots * 4 + (size_t)(L - nh) * SAVED16_BF * 2 + (size_t)nh * SAVED16_H * 4;
}
size_t fwd16_lds(int P) { return (size_t)R16 * P * 4 + (size_t)R16 * XPS16 * 2 + (size_t)R16 * HS16 * 2 + NW16 * 16 * 4; }
size_t bwd16_lds(int P) { return (size_t)R16 * P * 4 + (size_t)R16 * OS16 * 2 + (size_t)R16 * HS16 * 2; }

}  // namespace

bool flow16_eligible(const Flow& f) {
  if (!f.fused || f.factor || f.L < 1 || f.H != HID16 || f.P > 2 * KP16) return false;
  for (const FlowLayer& ly : f.layers)
    if (ly.np > KP16 || ly.na > KP16 || ly.np < 1 || ly.na < 1) return false;
  return true;
}

hipError_t flow16_sync_weights(Flow& f, hipStream_t st) {
  for (const FlowLayer& ly : f.layers) {
    const int n = W1F16 + W2F16 + W2TF16 + W1TF16;
    hipLaunchKernelGGL(k_frag16, dim3((n + 255) / 256), dim3(256), 0, st, f.params + ly.oW1, FLOW_PAD,
                       f.params + ly.oW2, ly.np, ly.na, ly.W1f, ly.W2f, ly.W2Tf, ly.W1Tf);
  }
  hipLaunchKernelGGL(k_ldbase16, dim3(1), dim3(64), 0, st, f.params + f.P, f.P, f.ldbase);
  return hipGetLastError();
}

hipError_t flow16_forward(Flow& f, const float* z, float* x, float* q, float* logdet, float* logdet_out, int B,
                          hipStream_t st) {
  if (B % R16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fwd16_kernel<NoRowModel>(f.L), dim3(B / R16), dim3(THREADS16), fwd16_lds(f.P), st, z, f.params,
                     f.ldbase, x, q, f.P, f.d16, f.L, f.clamp, logdet, logdet_out, NoRowModel{}, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t flow16_forward_model(Flow& f, const float* z, float* x, float* logdet, const FlowRowModel& m, float* g,
                                float* lj, int B, hipStream_t st) {
  if (B % R16 || !g || !lj) return hipErrorInvalidValue;
  auto launch = [&](auto fm) {
    fm.dim = m.dim;
    fm.v_scale = m.v_scale;
    fm.total_const = m.total_const;
    using FM = decltype(fm);
    hipLaunchKernelGGL(fwd16_kernel<FM>(f.L), dim3(B / R16), dim3(THREADS16), fwd16_lds(f.P), st, z, f.params,
                       f.ldbase, x, nullptr, f.P, f.d16, f.L, f.clamp, logdet, nullptr, fm, g, lj);
    return hipGetLastError();
  };
  if (m.K == 1) return launch(FunnelModel<1>{});
  if (m.K == 2) return launch(FunnelModel<2>{});
  return hipErrorInvalidValue;
}

hipError_t flow16_backward(Flow& f, const float* g_q, float* g_z, int B, hipStream_t st) {
  if (B % R16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(bwd16_kernel<false>(f.L), dim3(B / R16), dim3(THREADS16), bwd16_lds(f.P), st, g_q, f.params, g_z,
                     f.P, f.d16, f.L, f.clamp, Flow16Step{});
  return hipGetLastError();
}

hipError_t flow16_grad_step(Flow& f, const float* z, const FlowRowModel& m, float* lj, float* g_z, int B,
                           const Flow16Step* stp, hipStream_t st) {
  if (B % R16 || !lj) return hipErrorInvalidValue;
  if (stp && (!stp->wq || !stp->wp || !stp->im_f || !stp->valid)) return hipErrorInvalidValue;
  constexpr size_t kLds = 160 * 1024;
  // saved s and x_a e^s: 128 per wavefront that has outputs (16 w < 2 na) in some layer
  int na_max = 0;
  for (const FlowLayer& fl : f.layers) na_max = std::max(na_max, fl.na);
  const int slots = 128 * ((2 * na_max + 15) / 16);
  if (grad16_lds(f.P, f.L, 0, slots) > kLds) return hipErrorNotSupported;
  // keep the hidden activations of as many of the last layers as LDS holds
  int nh = f.L;
  while (nh > 0 && grad16_lds(f.P, f.L, nh, slots) > kLds) --nh;
  const size_t lds = grad16_lds(f.P, f.L, nh, slots);
  auto launch = [&](auto fm, auto step) {
    fm.dim = m.dim;
    fm.v_scale = m.v_scale;
    fm.total_const = m.total_const;
    using FM = decltype(fm);
    const auto kern = grad16_kernel<FM, decltype(step)::value>(f.L);
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(B / R16), dim3(THREADS16), lds, st, z, f.params, f.ldbase, f.P, f.d16, f.L, f.clamp,
                       fm, lj, g_z, stp ? *stp : Flow16Step{}, f.L - nh, slots);
    return hipGetLastError();
  };
  using T = std::true_type;
  using F = std::false_type;
  if (m.K == 1) return stp ? launch(FunnelModel<1>{}, T{}) : launch(FunnelModel<1>{}, F{});
  if (m.K == 2) return stp ? launch(FunnelModel<2>{}, T{}) : launch(FunnelModel<2>{}, F{});
  return hipErrorInvalidValue;
}

hipError_t flow16_backward_step(Flow& f, const float* g_q, float* g_z, int B, const Flow16Step& stp,
                                hipStream_t st) {
  if (B % R16 || !stp.wq || !stp.wp || !stp.im_f || !stp.valid) return hipErrorInvalidValue;
  hipLaunchKernelGGL(bwd16_kernel<true>(f.L), dim3(B / R16), dim3(THREADS16), bwd16_lds(f.P), st, g_q, f.params, g_z,
                     f.P, f.d16, f.L, f.clamp, stp);
  return hipGetLastError();
}

}  // namespace ut
