// ut_plan_audit.h -- host-side checks and derived tables of a general static
// execution plan (ut_plan_desc), and its prior draw.  Shared by the C ABI (plan
// creation) and the CPU oracle (its restatement of the interpreter), so that both
// read a plan the same way.
//
// The checks restate the reference's lowering rules and read/write audit
// (src/device/plan/lowering.jl:1360-1478 and the audit it runs): loops are not
// nested and hold no latent choices; every slot is written by one step and read
// only after it; values bound inside a loop body are read only inside it; every
// parameter row belongs to exactly one latent choice; the staged rows are exactly
// those the steps consume.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "ut_plan.h"

namespace ut {

// families whose support is discrete: observed only
inline bool plan_family_discrete(int fam) {
  return fam == UT_FAMILY_BERNOULLI || fam == UT_FAMILY_BERNOULLI_LOGIT || fam == UT_FAMILY_POISSON ||
         fam == UT_FAMILY_BINOMIAL || fam == UT_FAMILY_GEOMETRIC || fam == UT_FAMILY_NEGATIVEBINOMIAL;
}

struct PlanAudit {
  std::vector<int> first_node;  // per node, the first node of the expression rooted there (-1 unused)
  std::vector<int> cursor;      // per step (see PlanView)
  std::vector<int> loop_info;   // 4 per step
  std::vector<int> lists;
  std::vector<int> slot_local;
  int max_nodes = 1;
  int loop_params = 0;
  bool has_loop_param_reads = false;
  std::string error;
};

namespace plan_detail {

inline bool valid_op(int op) { return op >= UT_OP_LITERAL && op <= UT_OP_LOGISTIC; }

// value of a data-only expression (LITERAL, OBS and arithmetic) at block `cur`
inline double data_eval(const ut_plan_desc& d, int root, long long cur) {
  struct Rec {
    const ut_plan_desc& d;
    long long cur;
    double go(int n) const {
      const ut_plan_expr& e = d.exprs[n];
      if (e.op == UT_OP_LITERAL) return e.value;
      if (e.op == UT_OP_OBS) return d.observed[cur + e.index];
      const double a = go(e.a);
      const double b = plan_op_binary(e.op) ? go(e.b) : 0.0;
      return plan_node_forward<double>(e.op, a, b);
    }
  } r{d, cur};
  return r.go(root);
}

}  // namespace plan_detail

// Returns UT_OK or an error code; `a.error` holds the reason.
inline int plan_audit(const ut_plan_desc* dp, PlanAudit& a) {
  auto bad = [&](int code, const std::string& m) {
    a.error = "ut_plan: " + m;
    return code;
  };
  if (!dp) return bad(UT_E_ARG, "null plan description");
  const ut_plan_desc& d = *dp;
  const int P = d.num_params, S = d.num_slots, NS = d.num_steps, NE = d.num_exprs;
  if (P < 1 || P > PLAN_MAX_PARAMS) return bad(UT_E_UNSUPPORTED, "1 <= num_params <= 256 parameters");
  if (S < 0 || S > PLAN_MAX_SLOTS) return bad(UT_E_UNSUPPORTED, "at most 32 slots");
  if (NS < 1 || NS > PLAN_MAX_STEPS || !d.steps) return bad(UT_E_UNSUPPORTED, "1 to 256 steps");
  if (NE < 0 || (NE > 0 && !d.exprs)) return bad(UT_E_ARG, "expression nodes missing");
  if (d.num_loops < 0 || (d.num_loops > 0 && (!d.trip_counts || !d.loop_starts)))
    return bad(UT_E_ARG, "trip_counts / loop_starts missing");
  if (d.num_observed < 0 || (d.num_observed > 0 && !d.observed)) return bad(UT_E_ARG, "observed rows missing");
  // ---- expression nodes
  for (int n = 0; n < NE; ++n) {
    const ut_plan_expr& e = d.exprs[n];
    if (!plan_detail::valid_op(e.op)) return bad(UT_E_UNSUPPORTED, "node " + std::to_string(n) + ": unknown op");
    if (e.op == UT_OP_SLOT && (e.index < 1 || e.index > S))
      return bad(UT_E_ARG, "node " + std::to_string(n) + ": slot out of range");
    if (e.op == UT_OP_PARAM) {
      if (e.index < 1 || e.index > P) return bad(UT_E_ARG, "node " + std::to_string(n) + ": parameter row out of range");
      if (e.a >= n) return bad(UT_E_ARG, "node " + std::to_string(n) + ": operands must precede their node");
    }
    if (e.op == UT_OP_OBS && e.index < 0) return bad(UT_E_ARG, "node " + std::to_string(n) + ": negative OBS row");
    if (!plan_op_leaf(e.op)) {
      if (e.a < 0 || e.a >= n || (plan_op_binary(e.op) && (e.b < 0 || e.b >= n)))
        return bad(UT_E_ARG, "node " + std::to_string(n) + ": operands must precede their node");
    }
  }
  a.first_node.assign(NE, -1);
  a.max_nodes = 1;
  // reach(n): nodes of the expression rooted at n; data(n): depends on staged data only
  std::vector<char> seen(NE, 0);
  auto collect = [&](int root, std::vector<int>& out) {
    out.clear();
    std::vector<int> stack{root};
    std::fill(seen.begin(), seen.end(), 0);
    while (!stack.empty()) {
      const int n = stack.back();
      stack.pop_back();
      if (seen[n]) continue;
      seen[n] = 1;
      out.push_back(n);
      const ut_plan_expr& e = d.exprs[n];
      if (e.op == UT_OP_PARAM) {
        if (e.a >= 0) stack.push_back(e.a);
      } else if (!plan_op_leaf(e.op)) {
        stack.push_back(e.a);
        if (plan_op_binary(e.op)) stack.push_back(e.b);
      }
    }
  };
  std::vector<int> nodes;
  // check one expression used by step `si` (in_loop: inside a loop body)
  int loop_stride = 0;  // stride of the loop whose body is being checked
  auto check_expr = [&](int root, int si, bool in_loop, std::vector<int>* slots_read, bool* reads_params) -> int {
    if (root < 0 || root >= NE)
      return bad(UT_E_ARG, "step " + std::to_string(si) + ": argument expression out of range");
    collect(root, nodes);
    int lo = root;
    for (int n : nodes) lo = n < lo ? n : lo;
    if ((int)nodes.size() != root - lo + 1)
      return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) +
                                       ": an expression's nodes must be contiguous (post-order, unshared)");
    if ((int)nodes.size() > PLAN_MAX_NODES)
      return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": more than 24 nodes in an expression");
    if (a.first_node[root] >= 0 && a.first_node[root] != lo)
      return bad(UT_E_ARG, "node " + std::to_string(root) + ": inconsistent expression range");
    a.first_node[root] = lo;
    a.max_nodes = (int)nodes.size() > a.max_nodes ? (int)nodes.size() : a.max_nodes;
    for (int n : nodes) {
      const ut_plan_expr& e = d.exprs[n];
      if (e.op == UT_OP_OBS && !in_loop)
        return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": OBS nodes are only valid in loop bodies");
      if (e.op == UT_OP_OBS && e.index >= loop_stride)
        return bad(UT_E_ARG, "step " + std::to_string(si) + ": OBS row outside the iteration's block");
      if (e.op == UT_OP_SLOT && slots_read) slots_read->push_back(e.index - 1);
      if (e.op == UT_OP_PARAM) {
        if (in_loop && reads_params) *reads_params = true;
        if (e.a >= 0 && !in_loop)
          return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": PARAM row offsets are only valid in loop bodies");
        if (e.a >= 0) {
          // the row offset depends on staged data only
          std::vector<int> sub;
          std::vector<int> st{e.a};
          while (!st.empty()) {
            const int m = st.back();
            st.pop_back();
            const ut_plan_expr& f = d.exprs[m];
            if (f.op == UT_OP_SLOT || f.op == UT_OP_PARAM)
              return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) +
                                               ": a PARAM row offset must depend on staged data only");
            if (!plan_op_leaf(f.op)) {
              st.push_back(f.a);
              if (plan_op_binary(f.op)) st.push_back(f.b);
            }
          }
        }
      }
    }
    return UT_OK;
  };
  // ---- steps
  a.cursor.assign(NS, -1);
  a.loop_info.assign(4 * (size_t)NS, 0);
  a.lists.clear();
  a.slot_local.assign(S, 0);
  std::vector<int> writer(S, -1);   // step that binds each slot
  std::vector<int> owner(P, -1);    // latent step owning each parameter row
  std::vector<int> loop_of(S, -1);  // loop step whose body binds the slot
  long long cursor = 0;
  auto bind = [&](int slot1, int si, int loop_si) -> int {
    if (slot1 == 0) return UT_OK;
    if (slot1 < 0 || slot1 > S) return bad(UT_E_ARG, "step " + std::to_string(si) + ": binding slot out of range");
    const int s = slot1 - 1;
    if (writer[s] >= 0)
      return bad(UT_E_UNSUPPORTED, "slot " + std::to_string(slot1) + " is written by two steps");
    writer[s] = si;
    if (loop_si >= 0) {
      a.slot_local[s] = 1;
      loop_of[s] = loop_si;
    }
    return UT_OK;
  };
  auto check_reads = [&](const std::vector<int>& reads, int si, int loop_si) -> int {
    for (int s : reads) {
      if (writer[s] < 0 || writer[s] > si)
        // within a body the iterator is written by the loop step itself (< si)
        return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + " reads slot " + std::to_string(s + 1) +
                                         " before it is written");
      if (a.slot_local[s] && loop_of[s] != loop_si)
        return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + " reads slot " + std::to_string(s + 1) +
                                         ", bound inside another loop body (loop-local values do not escape)");
    }
    return UT_OK;
  };
  auto latent_rows = [&](int si, int first, int count) -> int {
    if (first < 1 || first + count - 1 > P)
      return bad(UT_E_ARG, "step " + std::to_string(si) + ": latent rows out of range");
    for (int r = first - 1; r < first - 1 + count; ++r) {
      if (owner[r] >= 0) return bad(UT_E_ARG, "parameter row " + std::to_string(r + 1) + " belongs to two latents");
      owner[r] = si;
    }
    return UT_OK;
  };
  std::vector<int> reads;
  int rc;
  for (int si = 0; si < NS; ++si) {
    const ut_plan_step& st = d.steps[si];
    if (st.kind == UT_STEP_LOOP) {
      if (st.loop_id < 1 || st.loop_id > d.num_loops) return bad(UT_E_ARG, "step " + std::to_string(si) + ": loop_id");
      if (st.body_length < 0 || si + st.body_length > NS - 1)
        return bad(UT_E_ARG, "step " + std::to_string(si) + ": loop body runs past the step list");
      if (st.stride < 0) return bad(UT_E_ARG, "step " + std::to_string(si) + ": negative stride");
      const int count = d.trip_counts[st.loop_id - 1];
      if (count < 0) return bad(UT_E_ARG, "loop " + std::to_string(st.loop_id) + ": negative trip count");
      a.cursor[si] = (int)cursor;
      loop_stride = st.stride;
      if ((rc = bind(st.iterator_slot, si, si))) return rc;
      std::vector<int> outer, local;
      if (st.iterator_slot > 0) local.push_back(st.iterator_slot - 1);
      bool params_read = false;
      for (int bi = si + 1; bi <= si + st.body_length; ++bi) {
        const ut_plan_step& b = d.steps[bi];
        a.cursor[bi] = -2;
        if (b.kind == UT_STEP_LOOP) return bad(UT_E_UNSUPPORTED, "nested loops are not lowerable");
        reads.clear();
        if (b.kind == UT_STEP_DETERMINISTIC) {
          if ((rc = check_expr(b.args[0], bi, true, &reads, &params_read))) return rc;
          if ((rc = check_reads(reads, bi, si))) return rc;
          if (b.binding_slot < 1) return bad(UT_E_ARG, "step " + std::to_string(bi) + ": deterministic without slot");
          if ((rc = bind(b.binding_slot, bi, si))) return rc;
          local.push_back(b.binding_slot - 1);
        } else if (b.kind == UT_STEP_CHOICE) {
          if (b.value_source != 0)
            return bad(UT_E_UNSUPPORTED, "step " + std::to_string(bi) +
                                             ": latent choices inside loops are not lowerable");
          if (b.family == UT_FAMILY_DIAG_NORMAL || b.family == UT_FAMILY_NONCENTERED_NORMAL)
            return bad(UT_E_UNSUPPORTED, "step " + std::to_string(bi) + ": family not valid in a loop body");
          if (b.family < UT_FAMILY_NORMAL || b.family > UT_FAMILY_NEGATIVEBINOMIAL)
            return bad(UT_E_UNSUPPORTED, "step " + std::to_string(bi) + ": unknown family");
          const int na = plan_family_arity(b.family);
          for (int j = 0; j < na; ++j)
            if ((rc = check_expr(b.args[j], bi, true, &reads, &params_read))) return rc;
          if ((rc = check_reads(reads, bi, si))) return rc;
          int rows = 1;
          if (b.family == UT_FAMILY_BERNOULLI_LOGIT_GLM) {
            if (b.dimension < 1 || b.coef_source < 1 || b.coef_source + b.dimension - 1 > P)
              return bad(UT_E_ARG, "step " + std::to_string(bi) + ": GLM coefficient rows out of range");
            rows = b.dimension + 1;
            params_read = true;
          }
          if (b.value_offset < 0 || b.value_offset + rows > st.stride)
            return bad(UT_E_ARG, "step " + std::to_string(bi) + ": value rows outside the iteration's block");
          if (b.binding_slot > 0) {
            if ((rc = bind(b.binding_slot, bi, si))) return rc;
            local.push_back(b.binding_slot - 1);
          }
        } else {
          return bad(UT_E_ARG, "step " + std::to_string(bi) + ": unknown step kind");
        }
        for (int s : reads)
          if (!a.slot_local[s] || loop_of[s] != si) {
            bool have = false;
            for (int o : outer) have = have || o == s;
            if (!have) outer.push_back(s);
          }
      }
      // PARAM row offsets in range for every iteration
      for (int bi = si + 1; bi <= si + st.body_length; ++bi) {
        const ut_plan_step& b = d.steps[bi];
        const int na = b.kind == UT_STEP_DETERMINISTIC ? 1 : plan_family_arity(b.family);
        for (int j = 0; j < na; ++j) {
          collect(b.args[j], nodes);
          for (int n : nodes) {
            const ut_plan_expr& e = d.exprs[n];
            if (e.op != UT_OP_PARAM || e.a < 0) continue;
            for (int t = 0; t < count; ++t) {
              const long long cur = cursor + (long long)t * st.stride;
              if (cur + st.stride > d.num_observed) break;  // reported below
              const double off = plan_detail::data_eval(d, e.a, cur);
              const long long row = (long long)e.index - 1 + (long long)std::floor(off + 0.5);
              if (!(off == std::floor(off)) || row < 0 || row >= P)
                return bad(UT_E_ARG, "step " + std::to_string(bi) + ": PARAM row offset " + std::to_string(off) +
                                         " at iteration " + std::to_string(t) + " is not a valid row");
            }
          }
        }
      }
      a.has_loop_param_reads = a.has_loop_param_reads || params_read;
      a.loop_info[4 * si + 0] = (int)a.lists.size();
      a.loop_info[4 * si + 1] = (int)outer.size();
      a.lists.insert(a.lists.end(), outer.begin(), outer.end());
      a.loop_info[4 * si + 2] = (int)a.lists.size();
      a.loop_info[4 * si + 3] = (int)local.size();
      a.lists.insert(a.lists.end(), local.begin(), local.end());
      cursor += (long long)count * st.stride;
      si += st.body_length;
      continue;
    }
    reads.clear();
    if (st.kind == UT_STEP_DETERMINISTIC) {
      if ((rc = check_expr(st.args[0], si, false, &reads, nullptr))) return rc;
      if ((rc = check_reads(reads, si, -1))) return rc;
      if (st.binding_slot < 1) return bad(UT_E_ARG, "step " + std::to_string(si) + ": deterministic without slot");
      if ((rc = bind(st.binding_slot, si, -1))) return rc;
      continue;
    }
    if (st.kind != UT_STEP_CHOICE) return bad(UT_E_ARG, "step " + std::to_string(si) + ": unknown step kind");
    const int fam = st.family;
    if (fam < UT_FAMILY_NORMAL || fam > UT_FAMILY_NEGATIVEBINOMIAL)
      return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": unknown family");
    if (fam == UT_FAMILY_BERNOULLI_LOGIT_GLM)
      return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": the GLM step belongs in a loop body");
    if (st.transform < UT_TRANSFORM_IDENTITY || st.transform > UT_TRANSFORM_LOGIT)
      return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": unknown transform");
    const int na = plan_family_arity(fam);
    for (int j = 0; j < na; ++j)
      if ((rc = check_expr(st.args[j], si, false, &reads, nullptr))) return rc;
    if ((rc = check_reads(reads, si, -1))) return rc;
    const bool discrete = plan_family_discrete(fam);
    if (st.value_source > 0) {
      if (discrete) return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": discrete latents are not lowerable");
      if (fam == UT_FAMILY_DIAG_NORMAL) {
        if (st.dimension < 1) return bad(UT_E_ARG, "step " + std::to_string(si) + ": DIAG_NORMAL dimension");
        if (st.transform != UT_TRANSFORM_IDENTITY)
          return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": vector latents take the identity transform");
        if (st.binding_slot != 0)
          return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": vector bindings are not materialised");
        if ((rc = latent_rows(si, st.value_source, st.dimension))) return rc;
      } else {
        if (fam == UT_FAMILY_NONCENTERED_NORMAL && st.transform != UT_TRANSFORM_IDENTITY)
          return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": a non-centred latent is unconstrained");
        if ((rc = latent_rows(si, st.value_source, 1))) return rc;
      }
    } else {
      if (fam == UT_FAMILY_NONCENTERED_NORMAL)
        return bad(UT_E_UNSUPPORTED, "step " + std::to_string(si) + ": a non-centred choice must be latent");
      a.cursor[si] = (int)cursor;
      if (fam == UT_FAMILY_DIAG_NORMAL) {
        if (st.dimension < 0 || st.binding_slot != 0)
          return bad(UT_E_ARG, "step " + std::to_string(si) + ": observed DIAG_NORMAL");
        cursor += st.dimension;
      } else {
        cursor += 1;
      }
    }
    if ((rc = bind(st.binding_slot, si, -1))) return rc;
  }
  for (int r = 0; r < P; ++r)
    if (owner[r] < 0) return bad(UT_E_ARG, "parameter row " + std::to_string(r + 1) + " belongs to no latent choice");
  if (cursor != d.num_observed)
    return bad(UT_E_DIM, "the steps consume " + std::to_string(cursor) + " staged rows, " +
                             std::to_string(d.num_observed) + " were staged");
  if (cursor > 2147483647LL) return bad(UT_E_UNSUPPORTED, "more than 2^31 staged rows");
  a.loop_params = a.has_loop_param_reads ? P : 0;
  return UT_OK;
}

// ---- prior draw ----------------------------------------------------------------------
// One column of unconstrained initial positions: each latent drawn from its family
// at the arguments its step computes from the values drawn before it (ancestral
// order), then mapped through the inverse transform.  Normals and uniforms are
// Philox coordinates (chain, key, stream, index): the normal of row r is index r,
// a second normal P + r, a uniform 2P + r.  The Student-t and gamma families have
// no closed-form draw here and start Uniform(-2, 2) in the unconstrained space.
inline void plan_prior_draw(const ut_plan_desc& d, int strategy, uint32_t chain, uint32_t key, uint32_t stream,
                            double* col) {
  const int P = d.num_params;
  if (strategy == UT_INIT_UNIFORM) {
    for (int r = 0; r < P; ++r) col[r] = 4.0 * rand_uniform_d(chain, key, stream, (uint32_t)r) - 2.0;
    return;
  }
  auto z = [&](int idx) { return (double)rand_normal_f(chain, key, stream, (uint32_t)idx); };
  auto u = [&](int idx) { return rand_uniform_d(chain, key, stream, (uint32_t)(2 * P + idx)); };
  std::vector<double> slots(d.num_slots > 0 ? d.num_slots : 1, 0.0);
  struct Ev {
    const ut_plan_desc& d;
    const std::vector<double>& slots;
    const double* col;
    double go(int n) const {
      const ut_plan_expr& e = d.exprs[n];
      switch (e.op) {
        case UT_OP_LITERAL: return e.value;
        case UT_OP_SLOT: return slots[e.index - 1];
        case UT_OP_PARAM: return col[e.index - 1];
        case UT_OP_OBS: return 0.0;
      }
      const double a = go(e.a);
      const double b = plan_op_binary(e.op) ? go(e.b) : 0.0;
      return plan_node_forward<double>(e.op, a, b);
    }
  } ev{d, slots, col};
  for (int r = 0; r < P; ++r) col[r] = 0.0;
  long long cursor = 0;
  for (int si = 0; si < d.num_steps; ++si) {
    const ut_plan_step& st = d.steps[si];
    if (st.kind == UT_STEP_LOOP) {
      cursor += (long long)d.trip_counts[st.loop_id - 1] * st.stride;
      si += st.body_length;
      continue;
    }
    if (st.kind == UT_STEP_DETERMINISTIC) {
      slots[st.binding_slot - 1] = ev.go(st.args[0]);
      continue;
    }
    const int na = plan_family_arity(st.family);
    double a[3] = {0, 0, 0};
    for (int j = 0; j < na; ++j) a[j] = ev.go(st.args[j]);
    if (st.value_source <= 0) {
      const double x = d.observed[cursor];
      cursor += st.family == UT_FAMILY_DIAG_NORMAL ? st.dimension : 1;
      if (st.binding_slot > 0) slots[st.binding_slot - 1] = x;
      continue;
    }
    const int r = st.value_source - 1;
    if (st.family == UT_FAMILY_DIAG_NORMAL) {
      for (int k = 0; k < st.dimension; ++k) col[r + k] = a[0] + a[1] * z(r + k);
      continue;
    }
    double x = 0.0;
    bool unconstrained_uniform = false;
    switch (st.family) {
      case UT_FAMILY_NORMAL: x = a[0] + a[1] * z(r); break;
      case UT_FAMILY_NONCENTERED_NORMAL: x = z(r); break;
      case UT_FAMILY_LOGNORMAL: x = exp_d(a[0] + a[1] * z(r)); break;
      case UT_FAMILY_HALFNORMAL: x = a[0] * std::fabs(z(r)); break;
      case UT_FAMILY_HALFCAUCHY: x = a[0] * std::fabs(z(r) / z(P + r)); break;
      case UT_FAMILY_CAUCHY: x = a[0] + a[1] * (z(r) / z(P + r)); break;
      case UT_FAMILY_EXPONENTIAL: x = -log_d(1.0 - u(r)) / a[0]; break;
      case UT_FAMILY_LAPLACE: {
        const double v = u(r) - 0.5;
        x = a[0] - a[1] * (v < 0 ? -1.0 : 1.0) * log_d(1.0 - 2.0 * std::fabs(v));
        break;
      }
      case UT_FAMILY_LOGISTIC: {
        double v = u(r);
        v = v < 1e-300 ? 1e-300 : v;
        x = a[0] + a[1] * log_d(v / (1.0 - v));
        break;
      }
      case UT_FAMILY_GUMBEL: {
        double v = u(r);
        v = v < 1e-300 ? 1e-300 : v;
        x = a[0] - a[1] * log_d(-log_d(v));
        break;
      }
      case UT_FAMILY_FRECHET: {
        double v = u(r);
        v = v < 1e-300 ? 1e-300 : v;
        x = a[1] * exp_d(-log_d(-log_d(v)) / a[0]);
        break;
      }
      case UT_FAMILY_RAYLEIGH: x = a[0] * sqrt_d(-2.0 * log_d(1.0 - u(r))); break;
      case UT_FAMILY_WEIBULL: x = a[1] * exp_d(log_d(-log_d(1.0 - u(r))) / a[0]); break;
      default: unconstrained_uniform = true;
    }
    double uu;
    if (unconstrained_uniform) {
      uu = 4.0 * rand_uniform_d(chain, key, stream, (uint32_t)r) - 2.0;
      x = plan_transform<double>(st.transform, uu).c;
    } else if (st.family == UT_FAMILY_NONCENTERED_NORMAL) {
      uu = x;
      x = a[0] + a[1] * x;
    } else if (st.transform == UT_TRANSFORM_LOG) {
      uu = log_d(x);
    } else if (st.transform == UT_TRANSFORM_LOGIT) {
      uu = log_d(x / (1.0 - x));
    } else {
      uu = x;
    }
    col[r] = uu;
    if (st.binding_slot > 0) slots[st.binding_slot - 1] = x;
  }
}

}  // namespace ut
