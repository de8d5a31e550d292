// ut_mfma.h -- BF16 matrix-core building blocks for gfx950 (v_mfma_f32_32x32x16_bf16).
//
// Operand layouts of one 32x32x16 MFMA (per 64-lane wavefront):
//   A (32 x 16, M x K):  lane l holds A[l % 32][8 (l / 32) .. 8 (l / 32) + 7]
//   B (16 x 32, K x N):  lane l holds B[8 (l / 32) .. + 7][l % 32]
//   D (32 x 32, fp32):   lane l, register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// Both operands are therefore read as 8 consecutive K elements (16 bytes), so every
// GEMM here takes its operands "K-contiguous": A as M x K row-major and B as the
// transposed N x K row-major matrix.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16_t;  // storage type of bfloat16 values

__device__ inline f32x16 mfma32(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Round-to-nearest-even float -> bfloat16 (NaN kept quiet).  On the device this is
// gfx950's conversion instruction (v_cvt_pk_bf16_f32), which gives these bits for every
// one of the 2^32 inputs (NaN quieting included); the host restates it.
__host__ __device__ inline bf16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(bf16_t, (__bf16)f);
#endif
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (bf16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// tanh for the flow's conditioners and scale clamps (~17 instructions against ~60 for
// tanhf; within 2 ulp of tanh in float32 emulation plus the hardware exp2 / rcp):
// |x| < 0.55: odd Taylor series of tanh through x^15 (the x^17 term is < 5e-8
// relative there); otherwise 1 - 2 / (1 + e^(2|x|)), which is >= 0.5 so the
// subtraction loses nothing, with the sign of x restored.
__device__ inline float flow_tanh(float x) {
  const float a = fabsf(x);
  const float u = x * x;
  float p = -0.0014558344f;     // -929569 / 638512875
  p = fmaf(p, u, 0.003592128f);  // 21844 / 6081075
  p = fmaf(p, u, -0.008863236f); // -1382 / 155925
  p = fmaf(p, u, 0.021869488f);  // 62 / 2835
  p = fmaf(p, u, -0.053968254f); // -17 / 315
  p = fmaf(p, u, 0.13333334f);   // 2 / 15
  p = fmaf(p, u, -0.33333334f);  // -1 / 3
  const float small = fmaf(x * u, p, x);
  const float e = __builtin_amdgcn_exp2f(a * 2.8853900817779268f);  // e^(2a) = 2^(2a log2(e))
  const float big = fmaf(-2.0f, __builtin_amdgcn_rcpf(1.0f + e), 1.0f);
  return a < 0.55f ? small : copysignf(big, x);
}

// Row index in a 32x32 MFMA result held by lane `lane`, register `r`.
__device__ inline int mfma32_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

}  // namespace ut
