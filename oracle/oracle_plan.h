// oracle_plan.h -- test infrastructure: the CPU oracle of the general static
// execution plan (ut_plan_desc).  Included by oracle.cpp only.
//
// Two restatements:
//  * plan_ref_eval: binary64 in the reference's own semantics -- steps in order, a
//    loop's iterations in order with the observation cursor threaded through
//    (src/device/logjoint_kernel.jl:620-637, 716-738), the family log densities as
//    src/device/math.jl:88-235 writes them and the transforms of math.jl:1137-1159
//    -- differentiated by forward-mode dual numbers, as the reference's gradient
//    kernel does (src/device/gradient_kernel.jl:1124-1145, duals src/device/dual.jl).
//    It shares nothing with the device interpreter except the plan format, so the
//    device's hand-written reverse-mode partials (ut_plan.h) are checked against an
//    independent derivation.
//  * plan_twin_eval<T>: the device interpreter's exact arithmetic order (lanes,
//    butterfly reductions, reverse sweep), on the shared scalar arithmetic of
//    ut_plan.h, for bit-exact comparisons.
#pragma once
#include <cmath>
#include <memory>
#include <vector>

#include "../uncertaintea_amd/csrc/ut_plan_audit.h"

namespace plan_oracle {

using namespace ut;

struct PlanData {
  std::vector<ut_plan_step> steps;
  std::vector<ut_plan_expr> exprs;
  std::vector<double> observed;
  std::vector<float> observed_f;
  std::vector<int32_t> trips, starts;
  ut_plan_desc desc{};
  PlanAudit audit;
};

inline std::shared_ptr<PlanData> copy_plan(const ut_plan_desc& d) {
  auto p = std::make_shared<PlanData>();
  p->steps.assign(d.steps, d.steps + d.num_steps);
  p->exprs.assign(d.exprs, d.exprs + d.num_exprs);
  p->observed.assign(d.observed, d.observed + d.num_observed);
  p->observed_f.assign(p->observed.begin(), p->observed.end());
  p->trips.assign(d.trip_counts, d.trip_counts + d.num_loops);
  p->starts.assign(d.loop_starts, d.loop_starts + d.num_loops);
  p->desc = d;
  p->desc.steps = p->steps.data();
  p->desc.exprs = p->exprs.data();
  p->desc.observed = p->observed.data();
  p->desc.trip_counts = p->trips.data();
  p->desc.loop_starts = p->starts.data();
  return p;
}

// ---------------------------------------------------------------- binary64 duals
struct Du {
  double v = 0.0;
  std::vector<double> d;
};
inline Du cst(double v, int P) { return Du{v, std::vector<double>(P, 0.0)}; }
inline Du lift(const Du& a, double v, double da) {
  Du o{v, std::vector<double>(a.d.size())};
  for (size_t i = 0; i < a.d.size(); ++i) o.d[i] = da * a.d[i];
  return o;
}
inline Du lift2(const Du& a, const Du& b, double v, double da, double db) {
  Du o{v, std::vector<double>(a.d.size())};
  for (size_t i = 0; i < a.d.size(); ++i) o.d[i] = da * a.d[i] + db * b.d[i];
  return o;
}
inline Du operator+(const Du& a, const Du& b) { return lift2(a, b, a.v + b.v, 1.0, 1.0); }
inline Du operator-(const Du& a, const Du& b) { return lift2(a, b, a.v - b.v, 1.0, -1.0); }
inline Du operator*(const Du& a, const Du& b) { return lift2(a, b, a.v * b.v, b.v, a.v); }
inline Du operator/(const Du& a, const Du& b) { return lift2(a, b, a.v / b.v, 1.0 / b.v, -a.v / (b.v * b.v)); }
inline Du operator-(const Du& a) { return lift(a, -a.v, -1.0); }
inline Du operator+(const Du& a, double b) { return lift(a, a.v + b, 1.0); }
inline Du operator-(const Du& a, double b) { return lift(a, a.v - b, 1.0); }
inline Du operator-(double a, const Du& b) { return lift(b, a - b.v, -1.0); }
inline Du operator*(double a, const Du& b) { return lift(b, a * b.v, a); }
inline Du operator/(const Du& a, double b) { return lift(a, a.v / b, 1.0 / b); }
inline Du operator/(double a, const Du& b) { return lift(b, a / b.v, -a / (b.v * b.v)); }
inline Du dexp(const Du& a) { const double e = std::exp(a.v); return lift(a, e, e); }
inline Du dlog(const Du& a) { return lift(a, std::log(a.v), 1.0 / a.v); }
inline Du dlog1p(const Du& a) { return lift(a, std::log1p(a.v), 1.0 / (1.0 + a.v)); }
inline Du dsqrt(const Du& a) { const double s = std::sqrt(a.v); return lift(a, s, 0.5 / s); }
inline Du dabs(const Du& a) { return lift(a, std::fabs(a.v), a.v > 0 ? 1.0 : (a.v < 0 ? -1.0 : 0.0)); }
inline Du dtanh(const Du& a) { const double t = std::tanh(a.v); return lift(a, t, 1.0 - t * t); }
inline Du dlogistic(const Du& a) { const double s = 1.0 / (1.0 + std::exp(-a.v)); return lift(a, s, s * (1.0 - s)); }
inline Du dpow(const Du& a, const Du& b) { return dexp(b * dlog(a)); }
// log Gamma with the derivative by a central difference of std::lgamma
inline Du dlgamma(const Du& a) {
  const double h = 1e-5 * std::max(1.0, std::fabs(a.v));
  return lift(a, std::lgamma(a.v), (std::lgamma(a.v + h) - std::lgamma(a.v - h)) / (2.0 * h));
}
inline Du dlog1p_exp(const Du& eta) { return eta.v > 0 ? eta + dlog1p(dexp(-eta)) : dlog1p(dexp(eta)); }
inline Du neg_inf(int P) { return cst(-INFINITY, P); }

constexpr double C_HALF_LOG_2PI = 0.9189385332046727, C_LOG_PI = 1.1447298858494002, C_LOG_2 = 0.6931471805599453;
constexpr double DMIN = 2.2250738585072014e-308;  // floatmin(Float64)

// the reference's device log densities (src/device/math.jl:88-293)
inline Du ref_family(int fam, const Du& x, const Du* a, int P) {
  switch (fam) {
    case UT_FAMILY_NORMAL: {
      const Du z = (x - a[0]) / a[1];
      return ((-dlog(a[1])) - C_HALF_LOG_2PI) - (z * z) / 2.0;
    }
    case UT_FAMILY_LOGNORMAL: {
      if (!(x.v >= 2.2250738585072014e-308)) return neg_inf(P);
      const Du lx = dlog(x), z = (lx - a[0]) / a[1];
      return (((-dlog(a[1])) - C_HALF_LOG_2PI) - (z * z) / 2.0) - lx;
    }
    case UT_FAMILY_EXPONENTIAL:
      if (!(x.v >= 0)) return neg_inf(P);
      return dlog(a[0]) - a[0] * x;
    case UT_FAMILY_LAPLACE:
      return (-dlog(2.0 * a[1])) - dabs(x - a[0]) / a[1];
    case UT_FAMILY_CAUCHY: {
      const Du z = (x - a[0]) / a[1];
      return (cst(-C_LOG_PI, P) - dlog(a[1])) - dlog1p(z * z);
    }
    case UT_FAMILY_HALFNORMAL: {
      if (!(x.v >= 0)) return neg_inf(P);
      const Du z = x / a[0];
      return ((C_LOG_2 - dlog(a[0])) - C_HALF_LOG_2PI) - (z * z) / 2.0;
    }
    case UT_FAMILY_HALFCAUCHY: {
      if (!(x.v >= 0)) return neg_inf(P);
      const Du z = x / a[0];
      return ((cst(C_LOG_2 - C_LOG_PI, P) - dlog(a[0])) - dlog1p(z * z));
    }
    case UT_FAMILY_LOGISTIC: {
      const Du z = (x - a[0]) / a[1], az = dabs(z);
      return ((-dlog(a[1])) - az) - 2.0 * dlog1p(dexp(-az));
    }
    case UT_FAMILY_STUDENTT: {
      const Du nu = a[0], z = (x - a[1]) / a[2];
      return (((dlgamma((nu + 1.0) / 2.0) - dlgamma(nu / 2.0)) - (dlog(nu) + C_LOG_PI) / 2.0) - dlog(a[2])) -
             ((nu + 1.0) * dlog1p((z * z) / nu)) / 2.0;
    }
    case UT_FAMILY_GAMMA: {
      if (!(x.v >= 2.2250738585072014e-308)) return neg_inf(P);
      return ((a[0] * dlog(a[1]) - dlgamma(a[0])) + (a[0] - 1.0) * dlog(x)) - a[1] * x;
    }
    case UT_FAMILY_BERNOULLI:
      if (!(x.v == 0.0 || x.v == 1.0)) return neg_inf(P);
      return x.v > 0.5 ? dlog(a[0]) : dlog1p(-a[0]);
    case UT_FAMILY_BERNOULLI_LOGIT:
      if (!(x.v == 0.0 || x.v == 1.0)) return neg_inf(P);
      return x * a[0] - dlog1p_exp(a[0]);
    case UT_FAMILY_POISSON: {
      const double k = std::round(x.v);
      if (!(x.v >= 0 && x.v == k)) return neg_inf(P);
      return (k * dlog(a[0]) - a[0]) - cst(std::lgamma(k + 1.0), P);
    }
    case UT_FAMILY_GUMBEL: {
      const Du z = (x - a[0]) / a[1];
      return ((-dlog(a[1])) - z) - dexp(-z);
    }
    case UT_FAMILY_FRECHET: {
      if (!(x.v >= DMIN)) return neg_inf(P);
      const Du lz = dlog(x) - dlog(a[1]);
      return ((dlog(a[0]) - dlog(a[1])) - (a[0] + 1.0) * lz) - dexp(-(a[0] * lz));
    }
    case UT_FAMILY_RAYLEIGH:
      if (!(x.v >= DMIN)) return neg_inf(P);
      return (dlog(x) - 2.0 * dlog(a[0])) - (x * x) / (2.0 * (a[0] * a[0]));
    case UT_FAMILY_INVERSEGAUSSIAN: {
      if (!(x.v >= DMIN)) return neg_inf(P);
      const Du d = x - a[0];
      return ((0.5 * dlog(a[1]) - C_HALF_LOG_2PI) - 1.5 * dlog(x)) - (a[1] * (d * d)) / (2.0 * (a[0] * a[0]) * x);
    }
    case UT_FAMILY_BETA: {
      if (!(x.v > 0.0 && x.v < 1.0)) return neg_inf(P);
      const Du lb = (dlgamma(a[0]) + dlgamma(a[1])) - dlgamma(a[0] + a[1]);
      return ((a[0] - 1.0) * dlog(x) + (a[1] - 1.0) * dlog1p(-x)) - lb;
    }
    case UT_FAMILY_INVERSEGAMMA:
      if (!(x.v >= DMIN)) return neg_inf(P);
      return ((a[0] * dlog(a[1]) - dlgamma(a[0])) - (a[0] + 1.0) * dlog(x)) - a[1] / x;
    case UT_FAMILY_WEIBULL: {
      if (!(x.v >= DMIN)) {
        if (!(x.v == 0.0)) return neg_inf(P);
        if (a[0].v < 1.0) return cst(INFINITY, P);
        if (a[0].v == 1.0) return -dlog(a[1]);
        return neg_inf(P);
      }
      return ((dlog(a[0]) + (a[0] - 1.0) * dlog(x)) - a[0] * dlog(a[1])) - dexp(a[0] * (dlog(x) - dlog(a[1])));
    }
    case UT_FAMILY_BINOMIAL: {
      const double k = std::round(x.v), n = std::round(a[0].v);
      if (!(x.v >= 0 && x.v == k && a[0].v >= 0 && a[0].v == n && k <= n)) return neg_inf(P);
      Du out = cst(std::lgamma(n + 1.0) - std::lgamma(k + 1.0) - std::lgamma(n - k + 1.0), P);
      if (k > 0) out = out + k * dlog(a[1]);
      if (k < n) out = out + (n - k) * dlog1p(-a[1]);
      return out;
    }
    case UT_FAMILY_GEOMETRIC: {
      const double k = std::round(x.v);
      if (!(x.v >= 0 && x.v == k)) return neg_inf(P);
      return k > 0 ? dlog(a[0]) + k * dlog1p(-a[0]) : dlog(a[0]);
    }
    case UT_FAMILY_NEGATIVEBINOMIAL: {
      const double k = std::round(x.v);
      if (!(x.v >= 0 && x.v == k)) return neg_inf(P);
      Du out = ((dlgamma(a[0] + k) - dlgamma(a[0])) - cst(std::lgamma(k + 1.0), P)) + a[0] * dlog(a[1]);
      if (k > 0) out = out + k * dlog1p(-a[1]);
      return out;
    }
  }
  return cst(NAN, P);
}

struct RefCtx {
  const PlanData& p;
  int P;
  std::vector<Du> slots;
  const std::vector<Du>* params;
  long long cur = 0;  // the iteration's block (loop bodies)
  Du node(int n) const {
    const ut_plan_expr& e = p.desc.exprs[n];
    switch (e.op) {
      case UT_OP_LITERAL: return cst(e.value, P);
      case UT_OP_SLOT: return slots[e.index - 1];
      case UT_OP_OBS: return cst(p.observed[cur + e.index], P);
      case UT_OP_PARAM: {
        const long long off = e.a >= 0 ? (long long)std::llround(node(e.a).v) : 0;
        return (*params)[e.index - 1 + off];
      }
    }
    const Du a = node(e.a);
    switch (e.op) {
      case UT_OP_ADD: return a + node(e.b);
      case UT_OP_SUB: return a - node(e.b);
      case UT_OP_MUL: return a * node(e.b);
      case UT_OP_DIV: return a / node(e.b);
      case UT_OP_POW: return dpow(a, node(e.b));
      case UT_OP_NEG: return -a;
      case UT_OP_EXP: return dexp(a);
      case UT_OP_LOG: return dlog(a);
      case UT_OP_LOG1P: return dlog1p(a);
      case UT_OP_SQRT: return dsqrt(a);
      case UT_OP_ABS: return dabs(a);
      case UT_OP_TANH: return dtanh(a);
      case UT_OP_LOGISTIC: return dlogistic(a);
    }
    return cst(NAN, P);
  }
};

// binary64 logjoint and gradient (reference semantics, forward-mode duals)
inline double plan_ref_eval(const PlanData& p, const double* q, double* g) {
  const ut_plan_desc& d = p.desc;
  const int P = d.num_params;
  std::vector<Du> params(P);
  for (int r = 0; r < P; ++r) {
    params[r] = cst(q[r], P);
    params[r].d[r] = 1.0;
  }
  RefCtx cx{p, P, std::vector<Du>(d.num_slots > 0 ? d.num_slots : 1, cst(0.0, P)), &params, 0};
  Du total = cst(0.0, P);
  long long cursor = 0;
  auto choice = [&](const ut_plan_step& st, bool in_loop) {
    const int na = plan_family_arity(st.family);
    Du a[3] = {cst(0, P), cst(0, P), cst(0, P)};
    if (st.family != UT_FAMILY_BERNOULLI_LOGIT_GLM)
      for (int j = 0; j < na; ++j) a[j] = cx.node(st.args[j]);
    if (st.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
      Du eta = cx.node(st.args[0]);
      for (int k = 0; k < st.dimension; ++k)
        eta = eta + cst(p.observed[cx.cur + st.value_offset + k], P) * params[st.coef_source - 1 + k];
      const Du y = cst(p.observed[cx.cur + st.value_offset + st.dimension], P);
      total = total + ref_family(UT_FAMILY_BERNOULLI_LOGIT, y, &eta, P);
      return;
    }
    if (st.family == UT_FAMILY_DIAG_NORMAL) {
      for (int k = 0; k < st.dimension; ++k) {
        const Du x = st.value_source > 0 ? params[st.value_source - 1 + k] : cst(p.observed[cursor + k], P);
        total = total + ref_family(UT_FAMILY_NORMAL, x, a, P);
      }
      if (st.value_source <= 0) cursor += st.dimension;
      return;
    }
    if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
      const Du z = params[st.value_source - 1];
      const Du std_args[2] = {cst(0.0, P), cst(1.0, P)};
      total = total + ref_family(UT_FAMILY_NORMAL, z, std_args, P);
      if (st.binding_slot > 0) cx.slots[st.binding_slot - 1] = a[0] + a[1] * z;
      return;
    }
    Du x, lad = cst(0.0, P);
    if (st.value_source > 0) {
      const Du u = params[st.value_source - 1];
      if (st.transform == UT_TRANSFORM_LOG) {
        x = dexp(u);
        lad = u;
      } else if (st.transform == UT_TRANSFORM_LOGIT) {
        x = 1.0 / (dexp(-u) + 1.0);
        const Du m = dabs(u);
        lad = (-m) - 2.0 * dlog1p(dexp(-m));
      } else {
        x = u;
      }
    } else if (in_loop) {
      x = cst(p.observed[cx.cur + st.value_offset], P);
    } else {
      x = cst(p.observed[cursor++], P);
    }
    total = total + (ref_family(st.family, x, a, P) + lad);
    if (st.binding_slot > 0) cx.slots[st.binding_slot - 1] = x;
  };
  for (int si = 0; si < d.num_steps; ++si) {
    const ut_plan_step& st = d.steps[si];
    if (st.kind == UT_STEP_LOOP) {
      const int count = d.trip_counts[st.loop_id - 1], start = d.loop_starts[st.loop_id - 1];
      for (int t = 0; t < count; ++t) {
        cx.cur = cursor + (long long)t * st.stride;
        if (st.iterator_slot > 0) cx.slots[st.iterator_slot - 1] = cst((double)(start + t), P);
        for (int bi = si + 1; bi <= si + st.body_length; ++bi) {
          const ut_plan_step& b = d.steps[bi];
          if (b.kind == UT_STEP_DETERMINISTIC) cx.slots[b.binding_slot - 1] = cx.node(b.args[0]);
          else choice(b, true);
        }
      }
      cursor += (long long)count * st.stride;
      si += st.body_length;
      continue;
    }
    if (st.kind == UT_STEP_DETERMINISTIC) {
      cx.slots[st.binding_slot - 1] = cx.node(st.args[0]);
      continue;
    }
    choice(st, false);
  }
  for (int r = 0; r < P; ++r) g[r] = total.d[r];
  return total.v;
}

// ---------------------------------------------------------------- kernel-order twin
// The device interpreter (GenericModel) on 64 emulated lanes: vectors are K x 64
// (row i at lane i % 64, slot i / 64); reductions are the kernels' butterfly.
template <class T>
T bfly64(const T* in) {
  T v[64], w[64];
  for (int l = 0; l < 64; ++l) v[l] = in[l];
  const int pm[4] = {1, 2, 7, 15};
  for (int s = 0; s < 4; ++s) {
    for (int l = 0; l < 64; ++l) w[l] = v[l] + v[l ^ pm[s]];
    for (int l = 0; l < 64; ++l) v[l] = w[l];
  }
  return (v[0] + v[16]) + (v[32] + v[48]);
}

template <class T> struct TwinObs;
template <> struct TwinObs<float> {
  static float get(const PlanData& p, long long i) { return p.observed_f[i]; }
};
template <> struct TwinObs<double> {
  static double get(const PlanData& p, long long i) { return p.observed[i]; }
};

template <class T>
struct Twin {
  const PlanData& p;
  const ut_plan_desc& d;
  const PlanAudit& a;
  int P, S, K;
  std::vector<T> su, au, qu;
  std::vector<T> sl, al, ap;  // [S or P][64]
  T* g;                        // K x 64

  Twin(const PlanData& pp, int K_, T* g_)
      : p(pp), d(pp.desc), a(pp.audit), P(pp.desc.num_params), S(pp.desc.num_slots), K(K_), g(g_) {
    su.assign(S > 0 ? S : 1, T(0));
    au.assign(S > 0 ? S : 1, T(0));
    sl.assign((size_t)(S > 0 ? S : 1) * 64, T(0));
    al.assign((size_t)(S > 0 ? S : 1) * 64, T(0));
    ap.assign((size_t)(a.loop_params > 0 ? a.loop_params : 1) * 64, T(0));
  }
  T slot_read(int s, bool lp, int lane) const { return (lp && a.slot_local[s]) ? sl[(size_t)s * 64 + lane] : su[s]; }
  void add_row(int row, T v) { g[row] = g[row] + v; }

  T expr_forward(int root, bool lp, long long cur, int lane, T* nv) const {
    const int first = a.first_node[root];
    for (int n = first; n <= root; ++n) {
      const ut_plan_expr& e = d.exprs[n];
      T v;
      if (e.op == UT_OP_LITERAL) v = (T)e.value;
      else if (e.op == UT_OP_SLOT) v = slot_read(e.index - 1, lp, lane);
      else if (e.op == UT_OP_OBS) v = TwinObs<T>::get(p, cur + e.index);
      else if (e.op == UT_OP_PARAM) {
        const int off = e.a >= 0 ? (int)(nv[e.a - first] + T(0.5)) : 0;
        v = qu[e.index - 1 + off];
      } else {
        const T va = nv[e.a - first];
        const T vb = plan_op_binary(e.op) ? nv[e.b - first] : T(0);
        v = plan_node_forward<T>(e.op, va, vb);
      }
      nv[n - first] = v;
    }
    return nv[root - first];
  }
  void expr_backward(int root, T seed, bool lp, long long cur, int lane) {
    T nv[PLAN_MAX_NODES], na[PLAN_MAX_NODES];
    const int first = a.first_node[root];
    expr_forward(root, lp, cur, lane, nv);
    for (int n = first; n <= root; ++n) na[n - first] = T(0);
    na[root - first] = seed;
    for (int n = root; n >= first; --n) {
      const ut_plan_expr& e = d.exprs[n];
      const T adj = na[n - first];
      if (e.op == UT_OP_LITERAL || e.op == UT_OP_OBS) continue;
      if (e.op == UT_OP_SLOT) {
        const int s = e.index - 1;
        if (lp) al[(size_t)s * 64 + lane] = al[(size_t)s * 64 + lane] + adj;
        else au[s] = au[s] + adj;
        continue;
      }
      if (e.op == UT_OP_PARAM) {
        const int off = e.a >= 0 ? (int)(nv[e.a - first] + T(0.5)) : 0;
        const int row = e.index - 1 + off;
        if (lp) ap[(size_t)row * 64 + lane] = ap[(size_t)row * 64 + lane] + adj;
        else add_row(row, adj);
        continue;
      }
      const T va = nv[e.a - first];
      const bool bin = plan_op_binary(e.op);
      const T vb = bin ? nv[e.b - first] : T(0);
      const T v = nv[n - first];
      T da, db;
      plan_node_partials<T>(e.op, va, vb, v, da, db);
      na[e.a - first] = na[e.a - first] + adj * da;
      if (bin) na[e.b - first] = na[e.b - first] + adj * db;
    }
  }
  void args_forward(const ut_plan_step& st, bool lp, long long cur, int lane, T (&av)[3]) const {
    const int n = plan_family_arity(st.family);
    T nv[PLAN_MAX_NODES];
    for (int j = 0; j < 3; ++j) av[j] = j < n ? expr_forward(st.args[j], lp, cur, lane, nv) : T(0);
  }
  T glm_eta(const ut_plan_step& st, long long cur, int lane) const {
    T nv[PLAN_MAX_NODES];
    T eta = expr_forward(st.args[0], true, cur, lane, nv);
    for (int k = 0; k < st.dimension; ++k)
      eta = eta + qu[st.coef_source - 1 + k] * TwinObs<T>::get(p, cur + st.value_offset + k);
    return eta;
  }

  T run_loop(int si) {
    const ut_plan_step& lst = d.steps[si];
    const int count = d.trip_counts[lst.loop_id - 1], start = d.loop_starts[lst.loop_id - 1];
    const long long base = a.cursor[si];
    const int* info = &a.loop_info[4 * (size_t)si];
    const int b0 = si + 1, b1 = si + lst.body_length;
    for (int i = 0; i < info[1]; ++i)
      for (int l = 0; l < 64; ++l) al[(size_t)a.lists[info[0] + i] * 64 + l] = T(0);
    T lpl[64];
    for (int l = 0; l < 64; ++l) lpl[l] = T(0);
    for (int t0 = 0; t0 < count; t0 += 64)
      for (int lane = 0; lane < 64; ++lane) {
        const int t = t0 + lane;
        if (t >= count) continue;
        const long long cur = base + (long long)t * lst.stride;
        if (lst.iterator_slot > 0) sl[(size_t)(lst.iterator_slot - 1) * 64 + lane] = (T)(start + t);
        T nv[PLAN_MAX_NODES];
        for (int bi = b0; bi <= b1; ++bi) {
          const ut_plan_step& st = d.steps[bi];
          if (st.kind == UT_STEP_DETERMINISTIC) {
            sl[(size_t)(st.binding_slot - 1) * 64 + lane] = expr_forward(st.args[0], true, cur, lane, nv);
            continue;
          }
          if (st.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
            T eta = glm_eta(st, cur, lane);
            const T y = TwinObs<T>::get(p, cur + st.value_offset + st.dimension);
            lpl[lane] = lpl[lane] + plan_family<T>(UT_FAMILY_BERNOULLI_LOGIT, y, &eta).lp;
            continue;
          }
          T av[3];
          args_forward(st, true, cur, lane, av);
          const T x = TwinObs<T>::get(p, cur + st.value_offset);
          lpl[lane] = lpl[lane] + plan_family<T>(st.family, x, av).lp;
          if (st.binding_slot > 0) sl[(size_t)(st.binding_slot - 1) * 64 + lane] = x;
        }
        for (int i = 0; i < info[3]; ++i) al[(size_t)a.lists[info[2] + i] * 64 + lane] = T(0);
        for (int bi = b1; bi >= b0; --bi) {
          const ut_plan_step& st = d.steps[bi];
          if (st.kind == UT_STEP_DETERMINISTIC) {
            expr_backward(st.args[0], al[(size_t)(st.binding_slot - 1) * 64 + lane], true, cur, lane);
            continue;
          }
          if (st.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
            T eta = glm_eta(st, cur, lane);
            const T y = TwinObs<T>::get(p, cur + st.value_offset + st.dimension);
            const T r = plan_family<T>(UT_FAMILY_BERNOULLI_LOGIT, y, &eta).da[0];
            for (int k = 0; k < st.dimension; ++k) {
              const int row = st.coef_source - 1 + k;
              ap[(size_t)row * 64 + lane] =
                  ap[(size_t)row * 64 + lane] + r * TwinObs<T>::get(p, cur + st.value_offset + k);
            }
            expr_backward(st.args[0], r, true, cur, lane);
            continue;
          }
          T av[3];
          args_forward(st, true, cur, lane, av);
          const T x = TwinObs<T>::get(p, cur + st.value_offset);
          const FamilyOut<T> fo = plan_family<T>(st.family, x, av);
          const int n = plan_family_arity(st.family);
          for (int j = 0; j < n; ++j) expr_backward(st.args[j], fo.da[j], true, cur, lane);
        }
      }
    for (int i = 0; i < info[1]; ++i) {
      const int s = a.lists[info[0] + i];
      au[s] = au[s] + bfly64(&al[(size_t)s * 64]);
    }
    return bfly64(lpl);
  }

  T eval(const T* q) {
    qu.assign(q, q + P);
    for (int i = 0; i < K * 64; ++i) g[i] = T(0);
    T total = T(0);
    for (int si = 0; si < d.num_steps; ++si) {
      const ut_plan_step& st = d.steps[si];
      if (st.kind == UT_STEP_LOOP) {
        total = total + run_loop(si);
        si += st.body_length;
        continue;
      }
      T nv[PLAN_MAX_NODES];
      if (st.kind == UT_STEP_DETERMINISTIC) {
        su[st.binding_slot - 1] = expr_forward(st.args[0], false, 0, 0, nv);
        continue;
      }
      T av[3];
      args_forward(st, false, 0, 0, av);
      if (st.family == UT_FAMILY_DIAG_NORMAL) {
        T part[64];
        const T fa[3] = {av[0], av[1], T(0)};
        for (int l = 0; l < 64; ++l) {
          part[l] = T(0);
          if (st.value_source > 0) {
            for (int k = 0; k < K; ++k) {
              const int r = k * 64 + l;
              if (r >= st.value_source - 1 && r < st.value_source - 1 + st.dimension)
                part[l] = part[l] + plan_family<T>(UT_FAMILY_NORMAL, q[r], fa).lp;
            }
          } else {
            for (int dd = l; dd < st.dimension; dd += 64)
              part[l] = part[l] + plan_family<T>(UT_FAMILY_NORMAL, TwinObs<T>::get(p, a.cursor[si] + dd), fa).lp;
          }
        }
        total = total + bfly64(part);
        continue;
      }
      T x, lad = T(0);
      FamilyOut<T> fo;
      if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
        const T z = qu[st.value_source - 1];
        const T sa[3] = {T(0), T(1), T(0)};
        fo = plan_family<T>(UT_FAMILY_NORMAL, z, sa);
        x = av[0] + av[1] * z;
      } else {
        if (st.value_source > 0) {
          const TransformOut<T> tr = plan_transform<T>(st.transform, qu[st.value_source - 1]);
          x = tr.c;
          lad = tr.lad;
        } else {
          x = TwinObs<T>::get(p, a.cursor[si]);
        }
        fo = plan_family<T>(st.family, x, av);
      }
      total = total + (fo.lp + lad);
      if (st.binding_slot > 0) su[st.binding_slot - 1] = x;
    }
    for (int r = 0; r < a.loop_params; ++r) add_row(r, bfly64(&ap[(size_t)r * 64]));
    for (int si = d.num_steps - 1; si >= 0; --si) {
      const ut_plan_step& st = d.steps[si];
      if (st.kind == UT_STEP_LOOP || a.cursor[si] == -2) continue;
      if (st.kind == UT_STEP_DETERMINISTIC) {
        expr_backward(st.args[0], au[st.binding_slot - 1], false, 0, 0);
        continue;
      }
      T av[3];
      args_forward(st, false, 0, 0, av);
      const int n = plan_family_arity(st.family);
      T seed[3] = {T(0), T(0), T(0)};
      if (st.family == UT_FAMILY_DIAG_NORMAL) {
        T pmu[64], psg[64];
        const T fa[3] = {av[0], av[1], T(0)};
        for (int l = 0; l < 64; ++l) {
          pmu[l] = psg[l] = T(0);
          if (st.value_source > 0) {
            for (int k = 0; k < K; ++k) {
              const int r = k * 64 + l;
              if (r >= st.value_source - 1 && r < st.value_source - 1 + st.dimension) {
                const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_NORMAL, q[r], fa);
                g[r] = g[r] + fo.dx;
                pmu[l] = pmu[l] + fo.da[0];
                psg[l] = psg[l] + fo.da[1];
              }
            }
          } else {
            for (int dd = l; dd < st.dimension; dd += 64) {
              const FamilyOut<T> fo =
                  plan_family<T>(UT_FAMILY_NORMAL, TwinObs<T>::get(p, a.cursor[si] + dd), fa);
              pmu[l] = pmu[l] + fo.da[0];
              psg[l] = psg[l] + fo.da[1];
            }
          }
        }
        seed[0] = bfly64(pmu);
        seed[1] = bfly64(psg);
      } else if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
        const T z = qu[st.value_source - 1];
        const T sa[3] = {T(0), T(1), T(0)};
        const FamilyOut<T> fo = plan_family<T>(UT_FAMILY_NORMAL, z, sa);
        const T adj = st.binding_slot > 0 ? au[st.binding_slot - 1] : T(0);
        add_row(st.value_source - 1, fo.dx + adj * av[1]);
        seed[0] = adj;
        seed[1] = adj * z;
      } else if (st.value_source > 0) {
        const TransformOut<T> tr = plan_transform<T>(st.transform, qu[st.value_source - 1]);
        const FamilyOut<T> fo = plan_family<T>(st.family, tr.c, av);
        const T adj_x = st.binding_slot > 0 ? fo.dx + au[st.binding_slot - 1] : fo.dx;
        add_row(st.value_source - 1, adj_x * tr.dc + tr.dlad);
        for (int j = 0; j < 3; ++j) seed[j] = fo.da[j];
      } else {
        const FamilyOut<T> fo = plan_family<T>(st.family, TwinObs<T>::get(p, a.cursor[si]), av);
        for (int j = 0; j < 3; ++j) seed[j] = fo.da[j];
      }
      for (int j = 0; j < n; ++j) expr_backward(st.args[j], seed[j], false, 0, 0);
    }
    return total;
  }
};

// kernel-exact logjoint: q, g are K x 64 vectors (row i at index i)
template <class T>
T plan_twin_eval(const PlanData& p, int K, const T* q, T* g, bool& ok) {
  Twin<T> tw(p, K, g);
  const T total = tw.eval(q);
  ok = std::isfinite(total);
  for (int i = 0; i < K * 64 && ok; ++i) ok = std::isfinite(g[i]);
  return total;
}

}  // namespace plan_oracle
