// ut_math.h -- scalar arithmetic shared bit-for-bit by the gfx950 kernels and the
// host (g++) build of the same functions.
//
// Everything here is written from published methods using only IEEE-754 basic
// operations (+ - * / sqrt, fma) and integer bit manipulation, so a value computed
// on the GPU and on the CPU from the same inputs is the same bit pattern, provided
// both sides compile with floating-point contraction disabled (-ffp-contract=off)
// and without fast-math.  That is what makes NUTS tree decisions reproducible
// between the device kernels and a CPU restatement driven by the same Philox
// coordinates.
//
// Methods:
//   * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
//     as 1, 2, 3" (SC'11).  Same (key, counter) conventions as the reference's
//     device RNG (src/device/rng.jl:175-243) and its uniform / Box-Muller
//     conversions (rng.jl:255-262, 320-343).
//   * exp: Cody & Waite (1980) additive range reduction x = k ln2 + r with a split
//     ln2, then the Taylor series of e^r evaluated by Horner's rule.
//   * log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = f/(2+f),
//     f = m-1, written as f - (f^2/2 - s (f^2/2 + R(s^2))) so the rounding error of
//     s only enters a small correction term.
//   * log1p: Goldberg (1991), "What every computer scientist should know about
//     floating-point arithmetic", Theorem 4: log1p(x) = log(u) * x / (u - 1),
//     u = 1 + x.
//   * cos: Cody-Waite reduction by pi/2 (three-part split) + Taylor series of sin
//     and cos on [-pi/4, pi/4].
// Accuracy targets are a few ulp; tests compare against libm/numpy.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define UT_HD __host__ __device__ inline
#else
#define UT_HD inline
#endif

namespace ut {

// Wavefront width of gfx950; the kernels map one chain to one wavefront.
constexpr int WAVE = 64;

UT_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
UT_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }
UT_HD uint64_t d2u(double x) { return __builtin_bit_cast(uint64_t, x); }
UT_HD double u2d(uint64_t x) { return __builtin_bit_cast(double, x); }

UT_HD float inf_f() { return u2f(0x7f800000u); }
UT_HD float nan_f() { return u2f(0x7fc00000u); }
UT_HD double inf_d() { return u2d(0x7ff0000000000000ull); }
UT_HD double nan_d() { return u2d(0x7ff8000000000000ull); }
UT_HD bool isfinite_f(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }
UT_HD bool isfinite_d(double x) { return (d2u(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }

UT_HD float fma_f(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
UT_HD double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }
UT_HD float sqrt_f(float x) { return __builtin_sqrtf(x); }
UT_HD double sqrt_d(double x) { return __builtin_sqrt(x); }

// ---------------------------------------------------------------- Philox4x32-10
struct philox_block { uint32_t w[4]; };

UT_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Ten rounds; the key is bumped by the Weyl increments before rounds 2..10.
UT_HD philox_block philox4x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1,
                              uint32_t c2, uint32_t c3) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c0, hi0, lo0);
    philox_mulhilo(M1, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  philox_block b;
  b.w[0] = c0; b.w[1] = c1; b.w[2] = c2; b.w[3] = c3;
  return b;
}

// [0, 1 - 2^-24]: top 24 bits of one word times 2^-24 (exact).
UT_HD float u01_f(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-08f; }
// [0, 1 - 2^-53]: top 53 bits of (w0 << 32 | w1) times 2^-53 (exact).
UT_HD double u01_d(uint32_t w0, uint32_t w1) {
  uint64_t bits = ((uint64_t)w0 << 32) | (uint64_t)w1;
  return (double)(bits >> 11) * 1.1102230246251565404236316680908203125e-16;
}

// Coordinate-indexed draws: key = (chain, iteration), counter = (stream, index, 0, 0).
UT_HD philox_block draw_block(uint32_t chain, uint32_t iteration, uint32_t stream,
                              uint32_t index) {
  return philox4x32(chain, iteration, stream, index, 0u, 0u);
}
UT_HD float rand_uniform_f(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t index) {
  return u01_f(draw_block(chain, iteration, stream, index).w[0]);
}
UT_HD double rand_uniform_d(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t index) {
  philox_block b = draw_block(chain, iteration, stream, index);
  return u01_d(b.w[0], b.w[1]);
}

// ----------------------------------------------------------------- binary32 math
UT_HD float ldexp_f(float p, int k) {
  // two exact-range scalings so subnormal results round once, in the last multiply
  int k1 = k / 2;
  int k2 = k - k1;
  float s1 = u2f((uint32_t)(k1 + 127) << 23);
  float s2 = u2f((uint32_t)(k2 + 127) << 23);
  return (p * s1) * s2;
}

UT_HD float exp_f(float x) {
  if (!(x == x)) return x;
  if (x > 88.72283935546875f) return inf_f();
  if (x < -103.97208404541015625f) return 0.0f;
  const float log2e = 1.44269502162933349609375f;
  const float shifter = 12582912.0f;  // 1.5 * 2^23: x + shifter - shifter rounds to integer
  float kf = (x * log2e + shifter) - shifter;
  const float ln2_hi = 0.693145751953125f;       // 16 significant bits: k*ln2_hi exact
  const float ln2_lo = 1.428606765330187e-06f;
  float r = fma_f(-kf, ln2_hi, x);
  r = fma_f(-kf, ln2_lo, r);
  // Taylor series of e^r, |r| <= ln2/2, through r^7/7!
  float p = 1.98412701138295233249664306640625e-04f;  // 1/5040
  p = fma_f(p, r, 1.388888922519981861114501953125e-03f);  // 1/720
  p = fma_f(p, r, 8.333333767950534820556640625e-03f);     // 1/120
  p = fma_f(p, r, 4.1666667163372039794921875e-02f);       // 1/24
  p = fma_f(p, r, 1.66666671633720397949218750e-01f);      // 1/6
  p = fma_f(p, r, 0.5f);
  p = fma_f(p, r, 1.0f);
  p = fma_f(p, r, 1.0f);
  return ldexp_f(p, (int)kf);
}

UT_HD float log_f(float x) {
  if (!(x == x)) return x;
  if (x < 0.0f) return nan_f();
  if (x == 0.0f) return -inf_f();
  if (x == inf_f()) return x;
  int e = 0;
  uint32_t ix = f2u(x);
  if (ix < 0x00800000u) {  // subnormal: scale into the normal range
    x = x * 8388608.0f;
    e = -23;
    ix = f2u(x);
  }
  e += (int)(ix >> 23) - 127;
  float m = u2f((ix & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421353816986083984375f) {
    m = m * 0.5f;
    e += 1;
  }
  float f = m - 1.0f;  // exact
  float s = f / (2.0f + f);
  float z = s * s;
  // R(z) = 2z/3 + 2z^2/5 + 2z^3/7 + 2z^4/9 + 2z^5/11
  float R = 0.181818187236785888671875f;                 // 2/11
  R = fma_f(R, z, 0.2222222238779067993164062f);         // 2/9
  R = fma_f(R, z, 0.2857142984867095947265625f);         // 2/7
  R = fma_f(R, z, 0.4000000059604644775390625f);         // 2/5
  R = fma_f(R, z, 0.666666686534881591796875f);          // 2/3
  R = R * z;
  float hfsq = 0.5f * f * f;
  float logm = f - (hfsq - s * (hfsq + R));
  float ef = (float)e;
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.428606765330187e-06f;
  return fma_f(ef, ln2_hi, fma_f(ef, ln2_lo, logm));
}

UT_HD float log1p_f(float x) {
  if (!(x == x)) return x;
  if (x == inf_f()) return x;
  float u = 1.0f + x;
  if (u == 1.0f) return x;
  return log_f(u) * (x / (u - 1.0f));
}

// cos on the Box-Muller argument range; valid for |x| < 2^15.
UT_HD float cos_f(float x) {
  if (!isfinite_f(x)) return nan_f();
  const float two_over_pi = 0.636619746685028076171875f;
  const float shifter = 12582912.0f;
  float kf = (x * two_over_pi + shifter) - shifter;
  const float p1 = 1.5703125f, p2 = 4.837512969970703125e-04f, p3 = 7.549790126404332e-08f;
  float r = fma_f(-kf, p1, x);
  r = fma_f(-kf, p2, r);
  r = fma_f(-kf, p3, r);
  float z = r * r;
  // sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9!
  float sp = 2.7557319224737950403e-06f;
  sp = fma_f(sp, z, -1.98412701138295233249664306640625e-04f);
  sp = fma_f(sp, z, 8.333333767950534820556640625e-03f);
  sp = fma_f(sp, z, -1.66666671633720397949218750e-01f);
  float sinr = fma_f(r * z, sp, r);
  // cos r = 1 - r^2/2 + r^4/4! - r^6/6! + r^8/8! - r^10/10!
  float cp = -2.7557319224737950403e-07f;
  cp = fma_f(cp, z, 2.4801587301587301587e-05f);
  cp = fma_f(cp, z, -1.388888922519981861114501953125e-03f);
  cp = fma_f(cp, z, 4.1666667163372039794921875e-02f);
  cp = fma_f(cp, z, -0.5f);
  float cosr = fma_f(cp, z, 1.0f);
  int q = ((int)kf) & 3;
  if (q == 0) return cosr;
  if (q == 1) return -sinr;
  if (q == 2) return -cosr;
  return sinr;
}

// N(0,1) from one Philox block: words 0 and 1 (reference rng.jl:332-337).
UT_HD float box_muller_f(philox_block b) {
  float u1 = u01_f(b.w[0]);
  float u2 = u01_f(b.w[1]);
  float radius = sqrt_f(-2.0f * log_f(1.0f - u1));
  return radius * cos_f(6.2831854820251465f * u2);
}
UT_HD float rand_normal_f(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t index) {
  return box_muller_f(draw_block(chain, iteration, stream, index));
}

// --------------------------------------------------------------- binary64 math
UT_HD double ldexp_d(double p, int k) {
  int k1 = k / 2;
  int k2 = k - k1;
  double s1 = u2d((uint64_t)(k1 + 1023) << 52);
  double s2 = u2d((uint64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

UT_HD double exp_d(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893383973096) return inf_d();
  if (x < -745.13321910194110842) return 0.0;
  const double log2e = 1.4426950408889634;
  const double shifter = 6755399441055744.0;  // 1.5 * 2^52
  double kf = (x * log2e + shifter) - shifter;
  const double ln2_hi = 0.6931471803691238;     // 32 significant bits: k*ln2_hi exact
  const double ln2_lo = 1.9082149292705877e-10;
  double r = fma_d(-kf, ln2_hi, x);
  r = fma_d(-kf, ln2_lo, r);
  // Taylor series through r^13/13!
  double p = 1.6059043836821614599e-10;       // 1/13!
  p = fma_d(p, r, 2.0876756987868098979e-09);  // 1/12!
  p = fma_d(p, r, 2.5052108385441718775e-08);  // 1/11!
  p = fma_d(p, r, 2.7557319223985890653e-07);  // 1/10!
  p = fma_d(p, r, 2.7557319223985890653e-06);  // 1/9!
  p = fma_d(p, r, 2.4801587301587301587e-05);  // 1/8!
  p = fma_d(p, r, 1.9841269841269841270e-04);  // 1/7!
  p = fma_d(p, r, 1.3888888888888888889e-03);  // 1/6!
  p = fma_d(p, r, 8.3333333333333333333e-03);  // 1/5!
  p = fma_d(p, r, 4.1666666666666666667e-02);  // 1/4!
  p = fma_d(p, r, 1.6666666666666666667e-01);  // 1/3!
  p = fma_d(p, r, 0.5);
  p = fma_d(p, r, 1.0);
  p = fma_d(p, r, 1.0);
  return ldexp_d(p, (int)kf);
}

UT_HD double log_d(double x) {
  if (!(x == x)) return x;
  if (x < 0.0) return nan_d();
  if (x == 0.0) return -inf_d();
  if (x == inf_d()) return x;
  int e = 0;
  uint64_t ix = d2u(x);
  if (ix < 0x0010000000000000ull) {
    x = x * 4503599627370496.0;  // 2^52
    e = -52;
    ix = d2u(x);
  }
  e += (int)(ix >> 52) - 1023;
  double m = u2d((ix & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  // R(z) = sum_{n=1..11} 2 z^n / (2n+1)
  double R = 2.0 / 23.0;
  R = fma_d(R, z, 2.0 / 21.0);
  R = fma_d(R, z, 2.0 / 19.0);
  R = fma_d(R, z, 2.0 / 17.0);
  R = fma_d(R, z, 2.0 / 15.0);
  R = fma_d(R, z, 2.0 / 13.0);
  R = fma_d(R, z, 2.0 / 11.0);
  R = fma_d(R, z, 2.0 / 9.0);
  R = fma_d(R, z, 2.0 / 7.0);
  R = fma_d(R, z, 2.0 / 5.0);
  R = fma_d(R, z, 2.0 / 3.0);
  R = R * z;
  double hfsq = 0.5 * f * f;
  double logm = f - (hfsq - s * (hfsq + R));
  double ef = (double)e;
  const double ln2_hi = 0.6931471803691238;
  const double ln2_lo = 1.9082149292705877e-10;
  return fma_d(ef, ln2_hi, fma_d(ef, ln2_lo, logm));
}

UT_HD double log1p_d(double x) {
  if (!(x == x)) return x;
  if (x == inf_d()) return x;
  double u = 1.0 + x;
  if (u == 1.0) return x;
  return log_d(u) * (x / (u - 1.0));
}

// x^y for x > 0 (dual-averaging schedule only).
UT_HD double pow_d(double x, double y) { return exp_d(y * log_d(x)); }

// log(e^x + e^y) = hi + log1p(e^(lo - hi)) with -inf handling (reference
// state_init.jl:284-293, persistent_nuts.jl:83-92), in two halves so that a
// caller can evaluate the exponential together with another one: lae_split gives
// hi and the exponent (-inf when either input is -inf, so the sum is the other
// input), lae_join adds log1p of the exponential.
UT_HD void lae_split(double x, double y, double& hi, double& arg) {
  const double ninf = -inf_d();
  if (x == ninf) { hi = y; arg = ninf; return; }
  if (y == ninf) { hi = x; arg = ninf; return; }
  hi = x > y ? x : y;
  arg = (x > y ? y : x) - hi;
}
UT_HD double lae_join(double hi, double e) { return hi + log1p_d(e); }
UT_HD double logaddexp_d(double x, double y) {
  double hi, arg;
  lae_split(x, y, hi, arg);
  return lae_join(hi, exp_d(arg));
}

// log(1 + e^eta) without overflow (reference math.jl:202-208).
UT_HD float log1p_exp_f(float eta) {
  return eta > 0.0f ? eta + log1p_f(exp_f(-eta)) : log1p_f(exp_f(eta));
}
UT_HD double log1p_exp_d(double eta) {
  return eta > 0.0 ? eta + log1p_d(exp_d(-eta)) : log1p_d(exp_d(eta));
}
// logistic(eta) = 1 / (1 + e^-eta), the reference's exp(-log1p_exp(-eta)) (math.jl:224),
// evaluated from t = e^-|eta| as 1 / (1 + t) for eta > 0 and t / (1 + t) otherwise: one
// exponential with the argument log1p_exp takes (a caller computing both shares it)
// and no cancellation for either sign.
UT_HD float logistic_f(float eta) {
  const float t = exp_f(eta > 0.0f ? -eta : eta);
  return eta > 0.0f ? 1.0f / (1.0f + t) : t / (1.0f + t);
}
UT_HD double logistic_d(double eta) {
  const double t = exp_d(eta > 0.0 ? -eta : eta);
  return eta > 0.0 ? 1.0 / (1.0 + t) : t / (1.0 + t);
}

}  // namespace ut
