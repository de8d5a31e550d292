// ut_plan.h -- arithmetic of the general static execution plan, shared by the
// device interpreter (GenericModel in ut_models.h) and the CPU oracle's twin of it.
//
// Everything here is a pure scalar function of its arguments compiled identically
// for host and device (UT_HD, no contraction): family log densities with their
// partial derivatives, the parameter transforms, one expression node forward and
// backward, log-gamma and digamma.  Summation orders across lanes are the
// interpreter's business; what is here must give the same bits on both sides.
//
// Families follow the reference's device logpdfs (src/device/math.jl:50-235): out
// of support scores -Inf with zero partials; the positive families with a log(x)
// term treat subnormal x as outside the support (`_device_positive_normal`,
// math.jl:64-66).  Transforms follow `_device_transform` (math.jl:1137-1159).
#pragma once
#include "../../include/uncertaintea_hip.h"
#include "ut_math.h"

namespace ut {

constexpr int PLAN_MAX_SLOTS = 32;
constexpr int PLAN_MAX_NODES = 24;   // nodes per expression
constexpr int PLAN_MAX_STEPS = 256;
constexpr int PLAN_MAX_PARAMS = 256;
constexpr int PLAN_MAX_LOOP_PARAMS = 128;  // interpreter, when a loop body reads parameters

// ---- precision dispatch -----------------------------------------------------------
template <class T> struct PM;
template <> struct PM<float> {
  UT_HD static float exp(float x) { return exp_f(x); }
  UT_HD static float log(float x) { return log_f(x); }
  UT_HD static float log1p(float x) { return log1p_f(x); }
  UT_HD static float sqrt(float x) { return sqrt_f(x); }
  UT_HD static float log1p_exp(float x) { return log1p_exp_f(x); }
  UT_HD static float logistic(float x) { return logistic_f(x); }
  UT_HD static float inf() { return inf_f(); }
  UT_HD static float nan() { return nan_f(); }
  UT_HD static bool finite(float x) { return isfinite_f(x); }
  UT_HD static float tiny() { return 1.17549435e-38f; }  // floatmin
};
template <> struct PM<double> {
  UT_HD static double exp(double x) { return exp_d(x); }
  UT_HD static double log(double x) { return log_d(x); }
  UT_HD static double log1p(double x) { return log1p_d(x); }
  UT_HD static double sqrt(double x) { return sqrt_d(x); }
  UT_HD static double log1p_exp(double x) { return log1p_exp_d(x); }
  UT_HD static double logistic(double x) { return logistic_d(x); }
  UT_HD static double inf() { return inf_d(); }
  UT_HD static double nan() { return nan_d(); }
  UT_HD static bool finite(double x) { return isfinite_d(x); }
  UT_HD static double tiny() { return 2.2250738585072014e-308; }
};

template <class T> UT_HD T t_abs(T x) { return x < T(0) ? -x : x; }
template <class T> UT_HD T t_sign(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

// tanh from the exponential: sign(x) (1 - e) / (1 + e), e = exp(-2|x|)
template <class T> UT_HD T t_tanh(T x) {
  const T e = PM<T>::exp(T(-2) * t_abs(x));
  const T r = (T(1) - e) / (T(1) + e);
  return x < T(0) ? -r : r;
}

// log Gamma(x) for x > 0 (NaN otherwise): recurrence up to x >= 12, then Stirling's
// series (Abramowitz & Stegun 6.1.41) through the x^-7 term.
template <class T> UT_HD T t_lgamma(T x) {
  if (!(x > T(0))) return PM<T>::nan();
  if (!PM<T>::finite(x)) return x;
  T shift = T(0);
  T prod = T(1);
  while (x < T(12)) {
    prod = prod * x;
    x = x + T(1);
    if (prod > T(1e30)) {
      shift = shift + PM<T>::log(prod);
      prod = T(1);
    }
  }
  shift = shift + PM<T>::log(prod);
  const T r = T(1) / x;
  const T r2 = r * r;
  const T series = r * (T(1.0 / 12.0) - r2 * (T(1.0 / 360.0) - r2 * (T(1.0 / 1260.0) - r2 * T(1.0 / 1680.0))));
  return ((x - T(0.5)) * PM<T>::log(x) - x + T(0.91893853320467274178)) + series - shift;
}

// digamma(x) for x > 0: recurrence up to x >= 12, then the asymptotic series
// (Abramowitz & Stegun 6.3.18) through the x^-8 term.
template <class T> UT_HD T t_digamma(T x) {
  if (!(x > T(0))) return PM<T>::nan();
  if (!PM<T>::finite(x)) return x;
  T acc = T(0);
  while (x < T(12)) {
    acc = acc + T(1) / x;
    x = x + T(1);
  }
  const T r = T(1) / x;
  const T r2 = r * r;
  const T series = r2 * (T(1.0 / 12.0) - r2 * (T(1.0 / 120.0) - r2 * (T(1.0 / 252.0) - r2 * T(1.0 / 240.0))));
  return ((PM<T>::log(x) - T(0.5) * r) - series) - acc;
}

// ---- transforms (reference _device_transform, math.jl:1137-1159) ------------------
template <class T> struct TransformOut {
  T c, lad, dc, dlad;  // value, log-abs-det, and their derivatives in u
};
template <class T> UT_HD TransformOut<T> plan_transform(int code, T u) {
  TransformOut<T> o;
  if (code == UT_TRANSFORM_LOG) {
    o.c = PM<T>::exp(u);
    o.lad = u;
    o.dc = o.c;
    o.dlad = T(1);
  } else if (code == UT_TRANSFORM_LOGIT) {
    o.c = T(1) / (T(1) + PM<T>::exp(-u));
    const T m = t_abs(u);
    const T em = PM<T>::exp(-m);
    o.lad = -m - T(2) * PM<T>::log1p(em);
    o.dc = o.c * (T(1) - o.c);
    o.dlad = -t_tanh(T(0.5) * u);
  } else {
    o.c = u;
    o.lad = T(0);
    o.dc = T(1);
    o.dlad = T(0);
  }
  return o;
}

// ---- families -------------------------------------------------------------------
template <class T> struct FamilyOut {
  T lp, dx, da[3];
};

template <class T> UT_HD FamilyOut<T> fam_outside() {
  FamilyOut<T> o;
  o.lp = -PM<T>::inf();
  o.dx = T(0);
  o.da[0] = o.da[1] = o.da[2] = T(0);
  return o;
}

constexpr double PLAN_HALF_LOG_2PI = 0.9189385332046727;
constexpr double PLAN_LOG_PI = 1.1447298858494002;
constexpr double PLAN_LOG_2 = 0.6931471805599453;

// exact non-negative integer (the count families' support; |x| < 2^52 holds for counts)
template <class T> UT_HD bool plan_count_ok(T x) {
  if (!(x >= T(0) && PM<T>::finite(x))) return false;
  return (T)(long long)x == x;
}

// log density of `x` and its partials in x and the arguments a[0..2]
template <class T> UT_HD FamilyOut<T> plan_family(int fam, T x, const T* a) {
  FamilyOut<T> o;
  o.dx = T(0);
  o.da[0] = o.da[1] = o.da[2] = T(0);
  switch (fam) {
    case UT_FAMILY_NORMAL: {  // (mu, sigma)
      const T s = a[1], z = (x - a[0]) / s;
      o.lp = (-PM<T>::log(s) - T(PLAN_HALF_LOG_2PI)) - z * z / T(2);
      o.dx = -z / s;
      o.da[0] = z / s;
      o.da[1] = (z * z - T(1)) / s;
      return o;
    }
    case UT_FAMILY_LOGNORMAL: {  // (mu, sigma) of log x
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T s = a[1], lx = PM<T>::log(x), z = (lx - a[0]) / s;
      o.lp = ((-PM<T>::log(s) - T(PLAN_HALF_LOG_2PI)) - z * z / T(2)) - lx;
      o.dx = (-z / s - T(1)) / x;
      o.da[0] = z / s;
      o.da[1] = (z * z - T(1)) / s;
      return o;
    }
    case UT_FAMILY_HALFNORMAL: {  // (sigma)
      if (!(x >= T(0))) return fam_outside<T>();
      const T s = a[0], z = x / s;
      o.lp = ((T(PLAN_LOG_2) - PM<T>::log(s)) - T(PLAN_HALF_LOG_2PI)) - z * z / T(2);
      o.dx = -z / s;
      o.da[0] = (z * z - T(1)) / s;
      return o;
    }
    case UT_FAMILY_HALFCAUCHY: {  // (scale)
      if (!(x >= T(0))) return fam_outside<T>();
      const T s = a[0], z = x / s, w = T(1) + z * z;
      o.lp = ((T(PLAN_LOG_2) - T(PLAN_LOG_PI)) - PM<T>::log(s)) - PM<T>::log1p(z * z);
      o.dx = -(T(2) * z) / (s * w);
      o.da[0] = (T(2) * z * z / w - T(1)) / s;
      return o;
    }
    case UT_FAMILY_EXPONENTIAL: {  // (rate)
      if (!(x >= T(0))) return fam_outside<T>();
      const T r = a[0];
      o.lp = PM<T>::log(r) - r * x;
      o.dx = -r;
      o.da[0] = T(1) / r - x;
      return o;
    }
    case UT_FAMILY_CAUCHY: {  // (mu, sigma)
      const T s = a[1], z = (x - a[0]) / s, w = T(1) + z * z;
      o.lp = (-T(PLAN_LOG_PI) - PM<T>::log(s)) - PM<T>::log1p(z * z);
      o.dx = -(T(2) * z) / (s * w);
      o.da[0] = -o.dx;
      o.da[1] = (T(2) * z * z / w - T(1)) / s;
      return o;
    }
    case UT_FAMILY_LAPLACE: {  // (loc, scale)
      const T b = a[1], d = x - a[0], ad = t_abs(d);
      o.lp = -PM<T>::log(T(2) * b) - ad / b;
      o.dx = -t_sign(d) / b;
      o.da[0] = t_sign(d) / b;
      o.da[1] = (ad / b - T(1)) / b;
      return o;
    }
    case UT_FAMILY_LOGISTIC: {  // (mu, scale)
      const T s = a[1], z = (x - a[0]) / s, az = t_abs(z);
      o.lp = (-PM<T>::log(s) - az) - T(2) * PM<T>::log1p(PM<T>::exp(-az));
      const T th = t_tanh(z / T(2));
      o.dx = -th / s;
      o.da[0] = th / s;
      o.da[1] = (z * th - T(1)) / s;
      return o;
    }
    case UT_FAMILY_STUDENTT: {  // (nu, mu, sigma)
      const T nu = a[0], s = a[2], z = (x - a[1]) / s, z2 = z * z, w = nu + z2;
      const T half = nu / T(2);
      o.lp = ((t_lgamma(half + T(0.5)) - t_lgamma(half)) - (PM<T>::log(nu) + T(PLAN_LOG_PI)) / T(2)) -
             PM<T>::log(s) - (nu + T(1)) * PM<T>::log1p(z2 / nu) / T(2);
      o.dx = -((nu + T(1)) * z) / (s * w);
      o.da[1] = -o.dx;
      o.da[2] = ((nu + T(1)) * z2 / w - T(1)) / s;
      o.da[0] = (((t_digamma(half + T(0.5)) - t_digamma(half)) / T(2) - T(1) / (T(2) * nu)) -
                 PM<T>::log1p(z2 / nu) / T(2)) + ((nu + T(1)) * z2) / (T(2) * nu * w);
      return o;
    }
    case UT_FAMILY_GAMMA: {  // (shape, rate)
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T k = a[0], r = a[1], lx = PM<T>::log(x), lr = PM<T>::log(r);
      o.lp = ((k * lr - t_lgamma(k)) + (k - T(1)) * lx) - r * x;
      o.dx = (k - T(1)) / x - r;
      o.da[0] = (lr - t_digamma(k)) + lx;
      o.da[1] = k / r - x;
      return o;
    }
    case UT_FAMILY_BERNOULLI: {  // (p)
      if (!(x == T(0) || x == T(1))) return fam_outside<T>();
      const T p = a[0];
      if (x > T(0.5)) {
        o.lp = PM<T>::log(p);
        o.da[0] = T(1) / p;
      } else {
        o.lp = PM<T>::log1p(-p);
        o.da[0] = -T(1) / (T(1) - p);
      }
      return o;
    }
    case UT_FAMILY_BERNOULLI_LOGIT: {  // (eta)
      if (!(x == T(0) || x == T(1))) return fam_outside<T>();
      const T eta = a[0];
      o.lp = x * eta - PM<T>::log1p_exp(eta);
      o.da[0] = x - PM<T>::logistic(eta);
      return o;
    }
    case UT_FAMILY_POISSON: {  // (lambda); counts are exact non-negative integers
      const T lam = a[0];
      const T k = x;
      if (!plan_count_ok(x)) return fam_outside<T>();
      o.lp = (k * PM<T>::log(lam) - lam) - t_lgamma(k + T(1));
      o.da[0] = k / lam - T(1);
      return o;
    }
    case UT_FAMILY_GUMBEL: {  // (mu, scale)
      const T s = a[1], z = (x - a[0]) / s, ez = PM<T>::exp(-z);
      o.lp = (-PM<T>::log(s) - z) - ez;
      o.dx = (ez - T(1)) / s;
      o.da[0] = -o.dx;
      o.da[1] = (z * (T(1) - ez) - T(1)) / s;
      return o;
    }
    case UT_FAMILY_FRECHET: {  // (shape, scale)
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T k = a[0], s = a[1], ls = PM<T>::log(s), lz = PM<T>::log(x) - ls, t = PM<T>::exp(-k * lz);
      o.lp = ((PM<T>::log(k) - ls) - (T(1) + k) * lz) - t;
      o.dx = (k * t - (T(1) + k)) / x;
      o.da[0] = T(1) / k - lz * (T(1) - t);
      o.da[1] = k * (T(1) - t) / s;
      return o;
    }
    case UT_FAMILY_RAYLEIGH: {  // (scale)
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T s = a[0], s2 = s * s;
      o.lp = (PM<T>::log(x) - T(2) * PM<T>::log(s)) - x * x / (T(2) * s2);
      o.dx = T(1) / x - x / s2;
      o.da[0] = (x * x / s2 - T(2)) / s;
      return o;
    }
    case UT_FAMILY_INVERSEGAUSSIAN: {  // (mu, lambda)
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T mu = a[0], lam = a[1], d = x - mu, mu2 = mu * mu;
      o.lp = ((T(0.5) * PM<T>::log(lam) - T(PLAN_HALF_LOG_2PI)) - T(1.5) * PM<T>::log(x)) -
             lam * d * d / (T(2) * mu2 * x);
      o.dx = -T(1.5) / x - lam * d * (x + mu) / (T(2) * mu2 * x * x);
      o.da[0] = lam * d / (mu2 * mu);
      o.da[1] = T(0.5) / lam - d * d / (T(2) * mu2 * x);
      return o;
    }
    case UT_FAMILY_BETA: {  // (alpha, beta)
      if (!(x > T(0) && x < T(1))) return fam_outside<T>();
      const T al = a[0], be = a[1], lx = PM<T>::log(x), l1x = PM<T>::log1p(-x);
      const T lbeta = (t_lgamma(al) + t_lgamma(be)) - t_lgamma(al + be);
      o.lp = ((al - T(1)) * lx + (be - T(1)) * l1x) - lbeta;
      o.dx = (al - T(1)) / x - (be - T(1)) / (T(1) - x);
      const T dab = t_digamma(al + be);
      o.da[0] = (lx - t_digamma(al)) + dab;
      o.da[1] = (l1x - t_digamma(be)) + dab;
      return o;
    }
    case UT_FAMILY_INVERSEGAMMA: {  // (shape, scale)
      if (!(x >= PM<T>::tiny())) return fam_outside<T>();
      const T k = a[0], s = a[1], lx = PM<T>::log(x), ls = PM<T>::log(s);
      o.lp = ((k * ls - t_lgamma(k)) - (k + T(1)) * lx) - s / x;
      o.dx = -(k + T(1)) / x + s / (x * x);
      o.da[0] = (ls - t_digamma(k)) - lx;
      o.da[1] = k / s - T(1) / x;
      return o;
    }
    case UT_FAMILY_WEIBULL: {  // (shape, scale)
      const T k = a[0], s = a[1], ls = PM<T>::log(s);
      if (!(x >= PM<T>::tiny())) {
        if (!(x == T(0))) return fam_outside<T>();
        // density at 0: +Inf below shape 1, 1/scale at shape 1, 0 above (math.jl:181-186)
        if (k < T(1)) {
          o.lp = PM<T>::inf();
        } else if (k == T(1)) {
          o.lp = -ls;
          o.da[1] = -T(1) / s;
        } else {
          return fam_outside<T>();
        }
        return o;
      }
      const T lx = PM<T>::log(x), lr = lx - ls, w = PM<T>::exp(k * lr);
      o.lp = ((PM<T>::log(k) + (k - T(1)) * lx) - k * ls) - w;
      o.dx = ((k - T(1)) - k * w) / x;
      o.da[0] = (T(1) / k + lr) - w * lr;
      o.da[1] = k * (w - T(1)) / s;
      return o;
    }
    case UT_FAMILY_BINOMIAL: {  // (trials, p); the count and the trials are exact integers
      const T n = a[0], p = a[1], k = x;
      if (!plan_count_ok(x) || !plan_count_ok(n) || !(k <= n)) return fam_outside<T>();
      o.lp = (t_lgamma(n + T(1)) - t_lgamma(k + T(1))) - t_lgamma((n - k) + T(1));
      if (k > T(0)) {
        o.lp = o.lp + k * PM<T>::log(p);
        o.da[1] = k / p;
      }
      if (k < n) {
        o.lp = o.lp + (n - k) * PM<T>::log1p(-p);
        o.da[1] = o.da[1] - (n - k) / (T(1) - p);
      }
      return o;  // no derivative in the integer trial count
    }
    case UT_FAMILY_GEOMETRIC: {  // (p): failures before the first success
      const T p = a[0], k = x;
      if (!plan_count_ok(x)) return fam_outside<T>();
      o.lp = PM<T>::log(p);
      o.da[0] = T(1) / p;
      if (k > T(0)) {
        o.lp = o.lp + k * PM<T>::log1p(-p);
        o.da[0] = o.da[0] - k / (T(1) - p);
      }
      return o;
    }
    case UT_FAMILY_NEGATIVEBINOMIAL: {  // (successes r, p)
      const T r = a[0], p = a[1], k = x;
      if (!plan_count_ok(x)) return fam_outside<T>();
      o.lp = (((t_lgamma(k + r) - t_lgamma(r)) - t_lgamma(k + T(1))) + r * PM<T>::log(p));
      o.da[0] = (t_digamma(k + r) - t_digamma(r)) + PM<T>::log(p);
      o.da[1] = r / p;
      if (k > T(0)) {
        o.lp = o.lp + k * PM<T>::log1p(-p);
        o.da[1] = o.da[1] - k / (T(1) - p);
      }
      return o;
    }
  }
  o.lp = PM<T>::nan();
  return o;
}

// number of family arguments
UT_HD int plan_family_arity(int fam) {
  switch (fam) {
    case UT_FAMILY_HALFNORMAL:
    case UT_FAMILY_HALFCAUCHY:
    case UT_FAMILY_EXPONENTIAL:
    case UT_FAMILY_BERNOULLI:
    case UT_FAMILY_BERNOULLI_LOGIT:
    case UT_FAMILY_POISSON:
    case UT_FAMILY_BERNOULLI_LOGIT_GLM:
    case UT_FAMILY_RAYLEIGH:
    case UT_FAMILY_GEOMETRIC:
      return 1;
    case UT_FAMILY_STUDENTT:
      return 3;
    default:
      return 2;
  }
}

// ---- expression nodes --------------------------------------------------------------
// Forward value of node `e` from its operand values va, vb (already computed).
template <class T> UT_HD T plan_node_forward(int op, T va, T vb) {
  switch (op) {
    case UT_OP_ADD: return va + vb;
    case UT_OP_SUB: return va - vb;
    case UT_OP_MUL: return va * vb;
    case UT_OP_DIV: return va / vb;
    case UT_OP_NEG: return -va;
    case UT_OP_EXP: return PM<T>::exp(va);
    case UT_OP_LOG: return PM<T>::log(va);
    case UT_OP_LOG1P: return PM<T>::log1p(va);
    case UT_OP_SQRT: return PM<T>::sqrt(va);
    case UT_OP_POW: return PM<T>::exp(vb * PM<T>::log(va));
    case UT_OP_ABS: return t_abs(va);
    case UT_OP_TANH: return t_tanh(va);
    case UT_OP_LOGISTIC: return PM<T>::logistic(va);
  }
  return PM<T>::nan();
}

// Local partials of an arithmetic node (value v) in its operands.
template <class T> UT_HD void plan_node_partials(int op, T va, T vb, T v, T& da, T& db) {
  da = T(0);
  db = T(0);
  switch (op) {
    case UT_OP_ADD: da = T(1); db = T(1); return;
    case UT_OP_SUB: da = T(1); db = T(-1); return;
    case UT_OP_MUL: da = vb; db = va; return;
    case UT_OP_DIV: da = T(1) / vb; db = -v / vb; return;
    case UT_OP_NEG: da = T(-1); return;
    case UT_OP_EXP: da = v; return;
    case UT_OP_LOG: da = T(1) / va; return;
    case UT_OP_LOG1P: da = T(1) / (T(1) + va); return;
    case UT_OP_SQRT: da = T(0.5) / v; return;
    case UT_OP_POW: da = vb * (v / va); db = v * PM<T>::log(va); return;
    case UT_OP_ABS: da = t_sign(va); return;
    case UT_OP_TANH: da = T(1) - v * v; return;
    case UT_OP_LOGISTIC: da = v * (T(1) - v); return;
  }
}

UT_HD bool plan_op_binary(int op) {
  return op == UT_OP_ADD || op == UT_OP_SUB || op == UT_OP_MUL || op == UT_OP_DIV || op == UT_OP_POW;
}
UT_HD bool plan_op_leaf(int op) {
  return op == UT_OP_LITERAL || op == UT_OP_SLOT || op == UT_OP_OBS || op == UT_OP_PARAM;
}

// Device-side view of a plan: the ABI arrays plus what the audit derived.
//   first_node[n]   first node of the expression whose root is n (its nodes are
//                   first_node[n] .. n)
//   cursor[i]       staged row where step i reads (a top-level observed choice or
//                   observed DIAG_NORMAL) or starts (a LOOP)
//   loop_info[4i..] for a LOOP step i: offset and count in `lists` of the slots
//                   its body reads but does not bind (outer), then of the slots
//                   it binds (local, the iterator included)
//   slot_local[s]   1 if slot s (0-based) is bound inside a loop body
struct PlanView {
  const ut_plan_step* steps = nullptr;
  const ut_plan_expr* exprs = nullptr;
  const int* first_node = nullptr;
  const float* obs_f = nullptr;
  const double* obs_d = nullptr;
  const int* trip = nullptr;
  const int* start = nullptr;
  const int* cursor = nullptr;
  const int* loop_info = nullptr;
  const int* lists = nullptr;
  const int* slot_local = nullptr;
  int num_steps = 0;
  int num_params = 0;
  int num_slots = 0;
  int max_nodes = 1;    // longest expression
  int loop_params = 0;  // num_params if a loop body reads parameters, else 0
};

template <class T> struct PObs;
template <> struct PObs<float> {
  UT_HD static float get(const PlanView& v, long long i) { return v.obs_f[i]; }
};
template <> struct PObs<double> {
  UT_HD static double get(const PlanView& v, long long i) { return v.obs_d[i]; }
};

// bytes of the interpreter's per-wavefront LDS region in precision T:
// slot values and adjoints (uniform), per-lane slot values and adjoints, per-lane
// expression node values and adjoints, per-lane parameter-gradient partials of the
// loop bodies, and the parameters themselves (uniform copy)
template <class T> UT_HD size_t plan_lds_bytes(const PlanView& v) {
  const size_t S = (size_t)v.num_slots, N = (size_t)v.max_nodes, PL = (size_t)v.loop_params;
  return sizeof(T) * (2 * S + 2 * S * 64 + 2 * N * 64 + PL * 64 + (size_t)v.num_params);
}

}  // namespace ut
