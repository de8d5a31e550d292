// ut_api.hip -- the C ABI of libuncertaintea_hip (declared in include/uncertaintea_hip.h).
//
// Host responsibilities: staging a plan's data into HBM once, owning per-workspace
// device buffers and one HIP stream, drawing initial positions, and sequencing the
// sampler kernels: warmup runs as one launch range per mass-adaptation window
// (per-chain dual averaging happens inside the kernel), each window end replays
// the pooled running variance on the device and re-searches every chain's step
// size, and sampling runs in fixed launch ranges that write draws straight into
// device result buffers, copied out once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/uncertaintea_hip.h"
#include "ut_chees.h"
#include "ut_flow.h"
#include "ut_gemm.h"
#include "ut_glm.h"
#include "ut_gp.h"
#include "ut_kernels.h"
#include "ut_lockstep.h"
#include "ut_math.h"
#include "ut_meads.h"
#include "ut_mfma.h"
#include "ut_plan_audit.h"

using namespace ut;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  // unsupported models point at the lowering report, as the reference's
  // _device_unsupported_message does (src/device/workspace.jl:24-35)
  if (code == UT_E_UNSUPPORTED) g_last_error += "\nSee device_lowering_report(model) for details.";
  return code;
}

#define UT_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(UT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc((void**)p, count * sizeof(T));
}

template <class T>
hipError_t upload(T** p, const std::vector<T>& v) {
  hipError_t e = dalloc(p, v.size());
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

double now_seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int k_slots(int P) {
  int k = (P + WAVE - 1) / WAVE;
  if (k <= 1) return 1;
  if (k <= 2) return 2;
  if (k <= 4) return 4;
  return -1;
}

constexpr double HALF_LOG_2PI = 0.9189385332046728;
constexpr double LOG_PI = 1.1447298858494002;
constexpr double LOG_2 = 0.6931471805599453;

}  // namespace

// ---------------------------------------------------------------------------- plan
// Host copy of a general plan description (the arrays the description points at).
struct HostPlan {
  std::vector<ut_plan_step> steps;
  std::vector<ut_plan_expr> exprs;
  std::vector<double> observed;
  std::vector<int32_t> trips, starts;
  ut_plan_desc desc{};
  void copy(const ut_plan_desc& d) {
    steps.assign(d.steps, d.steps + d.num_steps);
    exprs.assign(d.exprs, d.exprs + d.num_exprs);
    observed.assign(d.observed, d.observed + d.num_observed);
    trips.assign(d.trip_counts, d.trip_counts + d.num_loops);
    starts.assign(d.loop_starts, d.loop_starts + d.num_loops);
    desc = d;
    desc.steps = steps.data();
    desc.exprs = exprs.data();
    desc.observed = observed.data();
    desc.trip_counts = trips.data();
    desc.loop_starts = starts.data();
  }
};

struct ut_plan {
  DevModelDesc dev;
  HostPlan host;  // general plans run by the interpreter
  // matrix-core GLM (MODEL_GLM_MFMA): evaluated for all chains at once
  GlmPlan glm;
  // latent GP Poisson (MODEL_GP): evaluated for all chains at once
  GpPlan gp;
  // host copies for initial draws and validation
  int P = 0;
  std::vector<void*> owned;  // device allocations
  ~ut_plan() {
    for (void* p : owned) (void)hipFree(p);
  }
  template <class T>
  int stage(const T** dst, const std::vector<T>& v) {
    T* p = nullptr;
    hipError_t e = upload(&p, v);
    if (e != hipSuccess) return fail(UT_E_HIP, std::string("plan staging: ") + hipGetErrorString(e));
    owned.push_back(p);
    *dst = p;
    return UT_OK;
  }
};

// A trained (or training) affine-coupling flow; attached to a workspace it turns
// the lock-step samplers into samplers of the flow's base space.
struct ut_flow {
  Flow f;
  ~ut_flow() { flow_free(f); }
};

struct ut_workspace {
  const ut_plan* plan = nullptr;
  // the plan's model description with this workspace's chain layout (dev.waves
  // wavefronts per chain in the density kernel and the persistent samplers)
  DevModelDesc dev;
  int waves_request = 0;  // 0: automatic, else 1, 2, 4, 8
  ut_flow* flow = nullptr;  // not owned
  int C = 0;
  int P = 0;
  int precision = UT_FLOAT32;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;  // result copy-out overlapping compute
  // density scratch
  void* d_params = nullptr;
  void* d_values = nullptr;
  void* d_grads = nullptr;
  // sampler state
  ChainState st;
  double* d_im = nullptr;
  double* d_summary = nullptr;
  GlmScratch glm;  // MODEL_GLM_MFMA only
  GpScratch gp;    // MODEL_GP only
  // flow-wrapped batched gradients captured as HIP graphs, one per buffer set
  struct GradGraph {
    const float* q;
    float* lj;
    float* grad;
    bool need_lj;
    bool fwd_only;  // the flow forward and the model gradient only (the backward launched apart)
    const void* flow;
    unsigned gen;
    hipGraphExec_t exec;
  };
  std::vector<GradGraph> graphs;
  bool use_graphs = true;
  bool fused_step = true;  // lock-step leapfrog kick and drift in the flow backward's epilogue
  bool fused_model = true;  // a row model's gradient in the flow forward's epilogue
  void clear_graphs() {
    for (GradGraph& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~ut_workspace() {
    clear_graphs();
    void* ptrs[] = {d_params, d_values, d_grads, st.pos, st.grad, st.lj,
                    st.step, st.dual, st.im_f, d_im, d_summary,
                    glm.bhi, glm.blo, glm.gpart, glm.llpart};
    for (void* p : ptrs) (void)hipFree(p);
    gp_scratch_free(gp);
    if (stream) (void)hipStreamDestroy(stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }
};

// ------------------------------------------------------- chain layout
// Wavefronts per chain.  One wavefront per chain fills the chip when there are
// thousands of chains; with few chains over many observations the logistic model
// splits each chain's rows over a workgroup of 2, 4 or 8 wavefronts instead.  The
// automatic choice, for N >= kChainWaveMinRows, takes the most wavefronts with at
// most kChainWaveBudget wavefronts in all (16 per compute unit), at least 256 rows
// per wavefront, and the NUTS checkpoint stacks of every wavefront within the
// 160 KB of LDS at the given tree depth (depth < 0: no NUTS kernel).  Measured on
// the density kernel (N = 8192 and 65536): 8 wavefronts per chain are 5.4-6x
// faster than one at 64 chains, 4 are 1.7x faster at 1024 chains, and at 4096
// chains one wavefront per chain is within 15 % of the best.  An explicit request
// is returned when the model can take it and it fits, else -1.
constexpr int kDefaultTreeDepth = 10;
constexpr long long kChainWaveBudget = 4096;
constexpr long long kChainWaveMinRows = 2048;
constexpr size_t kLdsBytes = 160 * 1024;

static bool chain_waves_fit(const DevModelDesc& d, int w, int depth) {
  DevModelDesc t = d;
  t.waves = w;
  const size_t xs = w > 1 ? (size_t)w * ((size_t)d.K * WAVE + 1) * sizeof(float) : 0;
  const size_t need = depth >= 0 ? nuts_lds_bytes(t, depth) : xs;
  return need <= kLdsBytes;
}

static int chain_waves_for(const DevModelDesc& d, int C, int request, int depth) {
  if (request == 1) return 1;
  if (request > 1) return d.kind == MODEL_LOGISTIC && chain_waves_fit(d, request, depth) ? request : -1;
  if (d.kind != MODEL_LOGISTIC || d.n < kChainWaveMinRows) return 1;
  int w = 8;
  while (w > 1 && ((long long)C * w > kChainWaveBudget || d.n < 256LL * w || !chain_waves_fit(d, w, depth)))
    w /= 2;
  return w;
}

// ------------------------------------------------------- allocation-free density
// Scratch of ut_logjoint_gradient_noalloc: one file-scope device array per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, looked up once per device with that
// device current, and only that address is used.
static __device__ __align__(256) unsigned char scratch[UT_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of the current device's copy of `scratch`
int noalloc_scratch(void** out) {
  int dev = -1;
  UT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(UT_E_HIP, "ut_logjoint_gradient_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(UT_E_HIP, std::string("ut_logjoint_gradient_noalloc: hipGetSymbolAddress(scratch): ") +
                                hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return UT_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

extern "C" {

const char* ut_last_error(void) { return g_last_error.c_str(); }
int32_t ut_abi_version(void) { return 2; }

int32_t ut_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n > 0 ? 1 : 0;
}

void ut_philox4x32(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]) {
  philox_block b = philox4x32(key[0], key[1], ctr[0], ctr[1], ctr[2], ctr[3]);
  for (int i = 0; i < 4; ++i) out[i] = b.w[i];
}

int ut_plan_logistic_glm(const float* x, const float* y, int64_t n, int32_t d, double prior_loc,
                         double prior_scale, ut_plan** out) {
  if (!out || !x || !y) return fail(UT_E_ARG, "ut_plan_logistic_glm: null pointer");
  *out = nullptr;
  if (n <= 0 || d <= 0) return fail(UT_E_DIM, "ut_plan_logistic_glm: n and d must be positive");
  if (k_slots(d) < 0)
    return fail(UT_E_UNSUPPORTED, "ut_plan_logistic_glm: d > 256 is not lowerable to the wavefront-per-chain kernels");
  if (!(prior_scale > 0)) return fail(UT_E_ARG, "ut_plan_logistic_glm: prior_scale must be > 0");
  if ((long long)n * d >= (1ll << 31)) return fail(UT_E_DIM, "ut_plan_logistic_glm: n*d too large");
  for (int64_t i = 0; i < n; ++i)
    if (!(y[i] == 0.0f || y[i] == 1.0f))
      return fail(UT_E_ARG, "ut_plan_logistic_glm: y must hold 0/1 outcomes");
  auto* plan = new ut_plan();
  // feature-major, zero-padded to the model instance's feature bound
  const int rows = logistic_feature_rows(d, k_slots(d));
  std::vector<float> xT_f((size_t)n * rows, 0.0f);
  std::vector<double> xT_d((size_t)n * rows, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < d; ++j) {
      xT_f[(size_t)j * n + i] = x[(size_t)i * d + j];
      xT_d[(size_t)j * n + i] = (double)x[(size_t)i * d + j];
    }
  std::vector<float> y_f(y, y + n);
  std::vector<double> y_d(y, y + n);
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_LOGISTIC;
  m.P = d;
  m.K = k_slots(d);
  m.n = n;
  m.D = d;
  m.loc = prior_loc;
  m.scale = prior_scale;
  m.total_const = (double)d * (-log_d(prior_scale) - HALF_LOG_2PI);
  int rc;
  if ((rc = plan->stage(&m.xT_f, xT_f)) || (rc = plan->stage(&m.xT_d, xT_d)) ||
      (rc = plan->stage(&m.yv_f, y_f)) || (rc = plan->stage(&m.yv_d, y_d))) {
    delete plan;
    return rc;
  }
  plan->P = d;
  *out = plan;
  return UT_OK;
}

int ut_plan_logistic_glm_bf16(const uint16_t* x_bf16, const float* y, int64_t n, int32_t d, double prior_loc,
                              double prior_scale, ut_plan** out) {
  if (!out || !x_bf16 || !y) return fail(UT_E_ARG, "ut_plan_logistic_glm_bf16: null pointer");
  *out = nullptr;
  if (n <= 0 || d <= 0) return fail(UT_E_DIM, "ut_plan_logistic_glm_bf16: n and d must be positive");
  if (d != 256 || n % 32 != 0)
    return fail(UT_E_UNSUPPORTED, "ut_plan_logistic_glm_bf16: the matrix-core GLM needs d == 256 and n a multiple of 32");
  if (n > (1ll << 30)) return fail(UT_E_DIM, "ut_plan_logistic_glm_bf16: n too large");
  if (!(prior_scale > 0)) return fail(UT_E_ARG, "ut_plan_logistic_glm_bf16: prior_scale must be > 0");
  for (int64_t i = 0; i < n; ++i)
    if (!(y[i] == 0.0f || y[i] == 1.0f))
      return fail(UT_E_ARG, "ut_plan_logistic_glm_bf16: y must hold 0/1 outcomes");
  auto* plan = new ut_plan();
  std::vector<uint16_t> X(x_bf16, x_bf16 + (size_t)n * d);
  std::vector<float> y_f(y, y + n);
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_GLM_MFMA;
  m.P = d;
  m.n = n;
  m.D = d;
  m.loc = prior_loc;
  m.scale = prior_scale;
  m.total_const = (double)d * (-log_d(prior_scale) - HALF_LOG_2PI);
  GlmPlan& g = plan->glm;
  g.N = (int)n;
  g.D = d;
  g.loc = prior_loc;
  g.scale = prior_scale;
  g.total_const = m.total_const;
  int rc;
  if ((rc = plan->stage(&g.X, X)) || (rc = plan->stage(&g.y, y_f))) {
    delete plan;
    return rc;
  }
  plan->P = d;
  *out = plan;
  return UT_OK;
}

int ut_plan_gp_poisson(const double* inputs, const int32_t* counts, int32_t n, int32_t d, double noise,
                       double ls_loc, double ls_scale, double amp_scale, ut_plan** out) {
  if (!out || !inputs || !counts) return fail(UT_E_ARG, "ut_plan_gp_poisson: null pointer");
  *out = nullptr;
  if (n <= 0 || n % GP_TILE != 0) return fail(UT_E_UNSUPPORTED, "ut_plan_gp_poisson: n must be a positive multiple of 32");
  if (d < 1 || d + 1 > GP_MAX_T) return fail(UT_E_UNSUPPORTED, "ut_plan_gp_poisson: 1 <= d <= 2 input dimensions");
  if (!(noise >= 0) || !(ls_scale > 0) || !(amp_scale > 0)) return fail(UT_E_ARG, "ut_plan_gp_poisson: invalid scale");
  for (int i = 0; i < n; ++i)
    if (counts[i] < 0) return fail(UT_E_ARG, "ut_plan_gp_poisson: counts must be >= 0");
  auto* plan = new ut_plan();
  std::vector<double> X(inputs, inputs + (size_t)n * d), y(n);
  double lg = 0.0;
  for (int i = 0; i < n; ++i) {
    y[i] = (double)counts[i];
    lg += std::lgamma(y[i] + 1.0);
  }
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_GP;
  m.P = d + 1 + n;
  m.D = d;
  m.n = n;
  m.loc = ls_loc;
  m.scale = ls_scale;
  m.v_scale = amp_scale;
  GpPlan& g = plan->gp;
  g.N = n;
  g.D = d;
  g.noise2 = noise * noise + 1e-8;
  g.ls_loc = ls_loc;
  g.ls_scale = ls_scale;
  g.amp_scale = amp_scale;
  g.total_const = -lg + (double)n * (-HALF_LOG_2PI) + (double)d * (-log_d(ls_scale) - HALF_LOG_2PI) +
                  (LOG_2 - HALF_LOG_2PI - log_d(amp_scale));
  int rc;
  if ((rc = plan->stage(&g.X, X)) || (rc = plan->stage(&g.y, y))) {
    delete plan;
    return rc;
  }
  plan->P = m.P;
  *out = plan;
  return UT_OK;
}

int ut_plan_hier_normal_nc(const float* y, const int32_t* group_offsets, const double* sigma,
                           int32_t num_groups, double mu_loc, double mu_scale, double tau_scale,
                           ut_plan** out) {
  if (!out || !y || !group_offsets || !sigma) return fail(UT_E_ARG, "ut_plan_hier_normal_nc: null pointer");
  *out = nullptr;
  if (num_groups <= 0 || num_groups > WAVE)
    return fail(UT_E_UNSUPPORTED, "ut_plan_hier_normal_nc: 1 <= num_groups <= 64 required");
  if (!(mu_scale > 0) || !(tau_scale > 0)) return fail(UT_E_ARG, "ut_plan_hier_normal_nc: scales must be > 0");
  int max_n = 0;
  for (int j = 0; j < num_groups; ++j) {
    int nj = group_offsets[j + 1] - group_offsets[j];
    if (nj < 0) return fail(UT_E_ARG, "ut_plan_hier_normal_nc: group_offsets must be non-decreasing");
    if (!(sigma[j] > 0)) return fail(UT_E_ARG, "ut_plan_hier_normal_nc: sigma_j must be > 0");
    max_n = nj > max_n ? nj : max_n;
  }
  auto* plan = new ut_plan();
  std::vector<float> y_f((size_t)(max_n > 0 ? max_n : 1) * WAVE, 0.0f);
  std::vector<double> y_d(y_f.size(), 0.0);
  std::vector<int> nobs(WAVE, 0);
  // per-lane constants [6][64]: c_j, hinv_j, inv_j, then the group's sufficient
  // statistics ybar_j, S_j = sum_i (y_ij - ybar_j)^2 and n_j (binary64, summed in
  // observation order, two passes)
  std::vector<float> lane_f(6 * WAVE, 0.0f);
  std::vector<double> lane_d(6 * WAVE, 0.0);
  for (int j = 0; j < num_groups; ++j) {
    int nj = group_offsets[j + 1] - group_offsets[j];
    nobs[j] = nj;
    double sy = 0.0;
    for (int i = 0; i < nj; ++i) {
      float v = y[group_offsets[j] + i];
      y_f[(size_t)i * WAVE + j] = v;
      y_d[(size_t)i * WAVE + j] = (double)v;
      sy = sy + (double)v;
    }
    const double ybar = nj > 0 ? sy / (double)nj : 0.0;
    double ss = 0.0;
    for (int i = 0; i < nj; ++i) {
      const double dv = (double)y[group_offsets[j] + i] - ybar;
      ss = ss + dv * dv;
    }
    const double s = sigma[j];
    const double cj = -(double)nj * log_d(s) - (double)nj * HALF_LOG_2PI;
    const double hinv = 1.0 / (2.0 * s * s);
    const double inv = 1.0 / (s * s);
    const double row[6] = {cj, hinv, inv, ybar, ss, (double)nj};
    for (int r = 0; r < 6; ++r) {
      lane_d[r * WAVE + j] = row[r];
      lane_f[r * WAVE + j] = (float)row[r];
    }
  }
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_HIER;
  m.P = num_groups + 2;
  m.K = 2;
  m.J = num_groups;
  m.max_n = max_n;
  m.mu_loc = mu_loc;
  m.mu_scale = mu_scale;
  m.tau_scale = tau_scale;
  m.inv_mu_scale = 1.0 / mu_scale;
  m.inv_tau_scale = 1.0 / tau_scale;
  m.total_const = (double)num_groups * (-HALF_LOG_2PI) + (-log_d(mu_scale) - HALF_LOG_2PI) +
                  (LOG_2 - LOG_PI - log_d(tau_scale));
  int rc;
  if ((rc = plan->stage(&m.y_f, y_f)) || (rc = plan->stage(&m.y_d, y_d)) ||
      (rc = plan->stage(&m.nobs, nobs)) || (rc = plan->stage(&m.lane_f, lane_f)) ||
      (rc = plan->stage(&m.lane_d, lane_d))) {
    delete plan;
    return rc;
  }
  plan->P = m.P;
  *out = plan;
  return UT_OK;
}

int ut_plan_funnel(int32_t dim, double v_scale, ut_plan** out) {
  if (!out) return fail(UT_E_ARG, "ut_plan_funnel: null pointer");
  *out = nullptr;
  if (dim <= 0) return fail(UT_E_DIM, "ut_plan_funnel: dim must be positive");
  if (k_slots(dim + 1) < 0) return fail(UT_E_UNSUPPORTED, "ut_plan_funnel: dim + 1 > 256 unsupported");
  if (!(v_scale > 0)) return fail(UT_E_ARG, "ut_plan_funnel: v_scale must be > 0");
  auto* plan = new ut_plan();
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_FUNNEL;
  m.P = dim + 1;
  m.K = k_slots(dim + 1);
  m.dim = dim;
  m.v_scale = v_scale;
  m.total_const = -(double)dim * HALF_LOG_2PI + (-log_d(v_scale) - HALF_LOG_2PI);
  plan->P = m.P;
  *out = plan;
  return UT_OK;
}

// ------------------------------------------------------------- general plans
static int plan_route(const ut_plan_desc& d, const PlanAudit& a, int* engine);

int ut_plan_validate(const ut_plan_desc* desc) {
  PlanAudit a;
  const int rc = plan_audit(desc, a);
  if (rc != UT_OK) return fail(rc, a.error);
  int engine = 0;
  return plan_route(*desc, a, &engine);
}

}  // extern "C"

namespace {
// Recognition of the plans whose density a fused kernel computes: the caller's
// description is matched node by node, and the data are pulled out of the staged
// rows.  A plan that does not match exactly runs on the interpreter.
struct PlanMatch {
  const ut_plan_desc& d;
  bool in(int i) const { return i >= 0 && i < d.num_exprs; }
  bool lit(int i, double* v = nullptr) const {
    if (!in(i) || d.exprs[i].op != UT_OP_LITERAL) return false;
    if (v) *v = d.exprs[i].value;
    return true;
  }
  bool lit_is(int i, double v) const {
    double x;
    return lit(i, &x) && x == v;
  }
  bool slot(int i, int s) const { return in(i) && d.exprs[i].op == UT_OP_SLOT && d.exprs[i].index == s; }
  bool obs(int i, int k) const { return in(i) && d.exprs[i].op == UT_OP_OBS && d.exprs[i].index == k; }
  bool op(int i, int o) const { return in(i) && d.exprs[i].op == o; }
  const ut_plan_expr& e(int i) const { return d.exprs[i]; }
  bool choice(const ut_plan_step& st, int fam, int vs, int tr) const {
    return st.kind == UT_STEP_CHOICE && st.family == fam && st.value_source == vs && st.transform == tr;
  }
};

// steps: normal(0, v_scale) latent v; diag normal of x_1..x_dim at (0, exp(v / 2))
int match_funnel(const ut_plan_desc& d, int* dim, double* v_scale) {
  const PlanMatch m{d};
  if (d.num_steps != 2 || d.num_observed != 0) return 0;
  const ut_plan_step &s0 = d.steps[0], &s1 = d.steps[1];
  if (!m.choice(s0, UT_FAMILY_NORMAL, 1, UT_TRANSFORM_IDENTITY) || s0.binding_slot < 1) return 0;
  if (!m.lit_is(s0.args[0], 0.0) || !m.lit(s0.args[1], v_scale) || !(*v_scale > 0)) return 0;
  if (!m.choice(s1, UT_FAMILY_DIAG_NORMAL, 2, UT_TRANSFORM_IDENTITY) || s1.dimension + 1 != d.num_params) return 0;
  if (!m.lit_is(s1.args[0], 0.0) || !m.op(s1.args[1], UT_OP_EXP)) return 0;
  const int x = m.e(s1.args[1]).a, b = s0.binding_slot;
  const bool half = (m.op(x, UT_OP_MUL) && ((m.slot(m.e(x).a, b) && m.lit_is(m.e(x).b, 0.5)) ||
                                            (m.lit_is(m.e(x).a, 0.5) && m.slot(m.e(x).b, b)))) ||
                    (m.op(x, UT_OP_DIV) && m.slot(m.e(x).a, b) && m.lit_is(m.e(x).b, 2.0));
  if (!half) return 0;
  *dim = s1.dimension;
  return 1;
}

// steps: diag normal of beta (loc, scale); loop over rows: bernoullilogit GLM of
// the staged covariates (intercept 0) and outcome
int match_logistic(const ut_plan_desc& d, std::vector<float>& x, std::vector<float>& y, double* loc, double* scale) {
  const PlanMatch m{d};
  if (d.num_steps != 3) return 0;
  const ut_plan_step &s0 = d.steps[0], &s1 = d.steps[1], &s2 = d.steps[2];
  const int D = d.num_params;
  if (!m.choice(s0, UT_FAMILY_DIAG_NORMAL, 1, UT_TRANSFORM_IDENTITY) || s0.dimension != D) return 0;
  if (!m.lit(s0.args[0], loc) || !m.lit(s0.args[1], scale) || !(*scale > 0)) return 0;
  if (s1.kind != UT_STEP_LOOP || s1.body_length != 1 || s1.stride != D + 1) return 0;
  if (!m.choice(s2, UT_FAMILY_BERNOULLI_LOGIT_GLM, 0, UT_TRANSFORM_IDENTITY) || s2.coef_source != 1 ||
      s2.dimension != D || s2.value_offset != 0 || !m.lit_is(s2.args[0], 0.0))
    return 0;
  const long long N = d.trip_counts[s1.loop_id - 1];
  if (N < 1) return 0;
  x.resize((size_t)N * D);
  y.resize((size_t)N);
  for (long long i = 0; i < N; ++i) {
    const double* row = d.observed + i * (D + 1);
    for (int k = 0; k < D; ++k) x[(size_t)i * D + k] = (float)row[k];
    if (!(row[D] == 0.0 || row[D] == 1.0)) return 0;
    y[(size_t)i] = (float)row[D];
  }
  return 1;
}

// steps: mu ~ normal(m, s); tau ~ halfcauchy(t) (log); theta_raw ~ diag normal(0, 1);
// loop over observations (group, sigma, y): y ~ normal(mu + tau theta_raw[group], sigma)
int match_hier(const ut_plan_desc& d, std::vector<float>& yv, std::vector<int32_t>& off, std::vector<double>& sig,
               double* mloc, double* mscale, double* tscale) {
  const PlanMatch m{d};
  if (d.num_steps != 5) return 0;
  const ut_plan_step &s0 = d.steps[0], &s1 = d.steps[1], &s2 = d.steps[2], &s3 = d.steps[3], &s4 = d.steps[4];
  const int J = d.num_params - 2;
  if (J < 1 || J > 64) return 0;
  if (!m.choice(s0, UT_FAMILY_NORMAL, 1, UT_TRANSFORM_IDENTITY) || s0.binding_slot < 1) return 0;
  if (!m.lit(s0.args[0], mloc) || !m.lit(s0.args[1], mscale) || !(*mscale > 0)) return 0;
  if (!m.choice(s1, UT_FAMILY_HALFCAUCHY, 2, UT_TRANSFORM_LOG) || s1.binding_slot < 1) return 0;
  if (!m.lit(s1.args[0], tscale) || !(*tscale > 0)) return 0;
  if (!m.choice(s2, UT_FAMILY_DIAG_NORMAL, 3, UT_TRANSFORM_IDENTITY) || s2.dimension != J) return 0;
  if (!m.lit_is(s2.args[0], 0.0) || !m.lit_is(s2.args[1], 1.0)) return 0;
  if (s3.kind != UT_STEP_LOOP || s3.body_length != 1 || s3.stride != 3) return 0;
  if (!m.choice(s4, UT_FAMILY_NORMAL, 0, UT_TRANSFORM_IDENTITY) || s4.value_offset != 2 || !m.obs(s4.args[1], 1))
    return 0;
  const int mu = s4.args[0], b1 = s0.binding_slot, b2 = s1.binding_slot;
  if (!m.op(mu, UT_OP_ADD) || !m.slot(m.e(mu).a, b1) || !m.op(m.e(mu).b, UT_OP_MUL)) return 0;
  const ut_plan_expr& mul = m.e(m.e(mu).b);
  int pnode = -1;
  if (m.slot(mul.a, b2) && m.op(mul.b, UT_OP_PARAM)) pnode = mul.b;
  else if (m.op(mul.a, UT_OP_PARAM) && m.slot(mul.b, b2)) pnode = mul.a;
  if (pnode < 0 || m.e(pnode).index != 3 || !m.obs(m.e(pnode).a, 0)) return 0;
  const long long N = d.trip_counts[s3.loop_id - 1];
  off.assign((size_t)J + 1, 0);
  sig.assign((size_t)J, 0.0);
  yv.resize((size_t)N);
  std::vector<int> seen(J, 0);
  int last = 0;
  for (long long i = 0; i < N; ++i) {
    const double* row = d.observed + 3 * i;
    const int j = (int)row[0];
    if ((double)j != row[0] || j < last || j >= J || !(row[1] > 0)) return 0;
    if (seen[j] && sig[j] != row[1]) return 0;
    seen[j] = 1;
    sig[j] = row[1];
    last = j;
    ++off[(size_t)j + 1];
    yv[(size_t)i] = (float)row[2];
  }
  for (int j = 0; j < J; ++j) {
    if (!seen[j]) sig[j] = 1.0;  // empty groups: any positive scale
    off[(size_t)j + 1] += off[(size_t)j];
  }
  return 1;
}

template <class T>
int stage_plan_array(ut_plan* plan, const T** dst, const std::vector<T>& v) {
  if (v.empty()) {
    *dst = nullptr;
    return UT_OK;
  }
  return plan->stage(dst, v);
}
}  // namespace

// The engine that serves a validated plan: a fused family's kernels when the plan
// matches one (and the flags allow it), else the interpreter within its limits.
static int plan_route(const ut_plan_desc& d, const PlanAudit& a, int* engine) {
  if (!(d.flags & UT_PLAN_GENERIC)) {
    int dim = 0;
    double v = 0, loc = 0, scale = 0, t = 0;
    std::vector<float> x, y;
    std::vector<int32_t> off;
    std::vector<double> sig;
    if (match_funnel(d, &dim, &v)) {
      *engine = UT_ENGINE_FUNNEL;
      return UT_OK;
    }
    if (match_logistic(d, x, y, &loc, &scale)) {
      if ((d.flags & UT_PLAN_MATRIX_CORES) && d.num_params == 256 && y.size() % 32 == 0) {
        *engine = UT_ENGINE_GLM_MATRIX_CORES;
        return UT_OK;
      }
      if (k_slots(d.num_params) > 0) {
        *engine = UT_ENGINE_LOGISTIC;
        return UT_OK;
      }
    }
    if (match_hier(d, y, off, sig, &loc, &scale, &t)) {
      *engine = UT_ENGINE_HIER;
      return UT_OK;
    }
  }
  *engine = UT_ENGINE_INTERPRETER;
  if (k_slots(d.num_params) < 0) return fail(UT_E_UNSUPPORTED, "ut_plan: the interpreter takes at most 256 parameters");
  if (a.has_loop_param_reads && d.num_params > PLAN_MAX_LOOP_PARAMS)
    return fail(UT_E_UNSUPPORTED, "ut_plan: loop bodies read parameters: the interpreter takes at most 128 parameters");
  return UT_OK;
}

extern "C" {

int ut_plan_create(const ut_plan_desc* desc, ut_plan** out) {
  if (!out) return fail(UT_E_ARG, "ut_plan_create: null pointer");
  *out = nullptr;
  PlanAudit a;
  int rc = plan_audit(desc, a);
  if (rc != UT_OK) return fail(rc, a.error);
  const ut_plan_desc& d = *desc;
  if (!(d.flags & UT_PLAN_GENERIC)) {
    int dim = 0;
    double v_scale = 0, loc = 0, scale = 0, tscale = 0;
    std::vector<float> x, y;
    std::vector<int32_t> off;
    std::vector<double> sig;
    if (match_funnel(d, &dim, &v_scale)) {
      rc = ut_plan_funnel(dim, v_scale, out);
      return rc;
    }
    if (match_logistic(d, x, y, &loc, &scale)) {
      const int D = d.num_params;
      const long long N = (long long)y.size();
      if ((d.flags & UT_PLAN_MATRIX_CORES) && D == 256 && N % 32 == 0) {
        std::vector<uint16_t> xb(x.size());
        for (size_t i = 0; i < x.size(); ++i) xb[i] = f2bf(x[i]);
        rc = ut_plan_logistic_glm_bf16(xb.data(), y.data(), N, D, loc, scale, out);
          return rc;
      }
      if (k_slots(D) > 0) {
        rc = ut_plan_logistic_glm(x.data(), y.data(), N, D, loc, scale, out);
          return rc;
      }
    }
    if (match_hier(d, y, off, sig, &loc, &scale, &tscale)) {
      rc = ut_plan_hier_normal_nc(y.data(), off.data(), sig.data(), (int32_t)sig.size(), loc, scale, tscale, out);
      return rc;
    }
  }
  // ---- the interpreter
  int engine = 0;
  if ((rc = plan_route(d, a, &engine))) return rc;
  const int K = k_slots(d.num_params);
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_plan_create: no HIP device visible");
  auto* plan = new ut_plan();
  plan->host.copy(d);
  plan->P = d.num_params;
  DevModelDesc& m = plan->dev;
  m.kind = MODEL_PLAN;
  m.P = d.num_params;
  m.K = K;
  m.host_plan = &plan->host.desc;
  PlanView& v = m.plan;
  v.num_steps = d.num_steps;
  v.num_params = d.num_params;
  v.num_slots = d.num_slots;
  v.max_nodes = a.max_nodes;
  v.loop_params = a.loop_params;
  std::vector<float> obs_f(plan->host.observed.begin(), plan->host.observed.end());
  std::vector<int> trip(plan->host.trips.begin(), plan->host.trips.end());
  std::vector<int> start(plan->host.starts.begin(), plan->host.starts.end());
  std::vector<int> slot_local(a.slot_local.begin(), a.slot_local.end());
  if (slot_local.empty()) slot_local.push_back(0);
  if ((rc = stage_plan_array(plan, &v.steps, plan->host.steps)) ||
      (rc = stage_plan_array(plan, &v.exprs, plan->host.exprs)) ||
      (rc = stage_plan_array(plan, &v.first_node, a.first_node)) ||
      (rc = stage_plan_array(plan, &v.obs_f, obs_f)) ||
      (rc = stage_plan_array(plan, &v.obs_d, plan->host.observed)) ||
      (rc = stage_plan_array(plan, &v.trip, trip)) || (rc = stage_plan_array(plan, &v.start, start)) ||
      (rc = stage_plan_array(plan, &v.cursor, a.cursor)) || (rc = stage_plan_array(plan, &v.loop_info, a.loop_info)) ||
      (rc = stage_plan_array(plan, &v.lists, a.lists)) || (rc = stage_plan_array(plan, &v.slot_local, slot_local))) {
    delete plan;
    return rc;
  }
  *out = plan;
  return UT_OK;
}

int ut_plan_engine(const ut_plan* plan, int32_t* engine) {
  if (!plan || !engine) return fail(UT_E_ARG, "ut_plan_engine: null pointer");
  switch (plan->dev.kind) {
    case MODEL_LOGISTIC: *engine = UT_ENGINE_LOGISTIC; break;
    case MODEL_HIER: *engine = UT_ENGINE_HIER; break;
    case MODEL_FUNNEL: *engine = UT_ENGINE_FUNNEL; break;
    case MODEL_GLM_MFMA: *engine = UT_ENGINE_GLM_MATRIX_CORES; break;
    case MODEL_GP: *engine = UT_ENGINE_GP; break;
    default: *engine = UT_ENGINE_INTERPRETER;
  }
  return UT_OK;
}

int ut_plan_destroy(ut_plan* plan) {
  delete plan;
  return UT_OK;
}

int ut_plan_num_params(const ut_plan* plan, int32_t* num_params) {
  if (!plan || !num_params) return fail(UT_E_ARG, "ut_plan_num_params: null pointer");
  *num_params = plan->P;
  return UT_OK;
}

int ut_workspace_create(const ut_plan* plan, int32_t num_chains, int32_t precision,
                        ut_workspace** out) {
  if (!plan || !out) return fail(UT_E_ARG, "ut_workspace_create: null pointer");
  *out = nullptr;
  if (num_chains <= 0) return fail(UT_E_DIM, "ut_workspace_create: num_chains must be positive");
  if (precision != UT_FLOAT32 && precision != UT_FLOAT64)
    return fail(UT_E_PRECISION, "ut_workspace_create: precision must be UT_FLOAT32 or UT_FLOAT64");
  const bool mfma = plan->dev.kind == MODEL_GLM_MFMA;
  const bool gp = plan->dev.kind == MODEL_GP;
  if ((mfma || gp) && precision != UT_FLOAT32)
    return fail(UT_E_PRECISION, "ut_workspace_create: the matrix-core GLM and GP plans evaluate in UT_FLOAT32");
  if (mfma && num_chains % 128 != 0)
    return fail(UT_E_DIM, "ut_workspace_create: the matrix-core GLM plan needs num_chains a multiple of 128");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_workspace_create: no HIP device visible");
  auto* ws = new ut_workspace();
  ws->plan = plan;
  ws->dev = plan->dev;
  ws->dev.waves = chain_waves_for(plan->dev, num_chains, 0, kDefaultTreeDepth);
  ws->C = num_chains;
  ws->P = plan->P;
  ws->precision = precision;
  {
    // UT_FLOW_GRAPHS=0 keeps the flow-wrapped gradient eager (for comparison runs)
    const char* g = std::getenv("UT_FLOW_GRAPHS");
    ws->use_graphs = !(g && std::strcmp(g, "0") == 0);
    // UT_FUSED_STEP=0 keeps k_ls_step a kernel of its own (for comparison runs)
    const char* fs = std::getenv("UT_FUSED_STEP");
    ws->fused_step = !(fs && std::strcmp(fs, "0") == 0);
    // UT_FUSED_MODEL=0 keeps the funnel's density kernel a launch of its own
    const char* fm = std::getenv("UT_FUSED_MODEL");
    ws->fused_model = !(fm && std::strcmp(fm, "0") == 0);
  }
  const size_t PC = (size_t)ws->P * num_chains;
  const size_t es = precision == UT_FLOAT32 ? 4 : 8;
  hipError_t e = hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ws->copy_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ws->d_params, PC * es);
  if (e == hipSuccess) e = hipMalloc(&ws->d_values, (size_t)num_chains * es);
  if (e == hipSuccess) e = hipMalloc(&ws->d_grads, PC * es);
  if (e == hipSuccess) e = dalloc(&ws->st.pos, PC);
  if (e == hipSuccess) e = dalloc(&ws->st.grad, PC);
  if (e == hipSuccess) e = dalloc(&ws->st.lj, (size_t)num_chains);
  if (e == hipSuccess) e = dalloc(&ws->st.step, (size_t)num_chains);
  if (e == hipSuccess) e = dalloc(&ws->st.dual, (size_t)num_chains * 5);
  if (e == hipSuccess) e = dalloc(&ws->st.im_f, (size_t)ws->P);
  if (e == hipSuccess) e = dalloc(&ws->d_im, (size_t)ws->P);
  if (e == hipSuccess) e = dalloc(&ws->d_summary, 8);
  if (mfma) {
    GlmScratch& g = ws->glm;
    g.splits = glm_splits(plan->glm.N, num_chains);
    if (e == hipSuccess) e = dalloc(&g.bhi, PC);
    if (e == hipSuccess) e = dalloc(&g.blo, PC);
    if (e == hipSuccess) e = dalloc(&g.gpart, PC * (size_t)g.splits);
    if (e == hipSuccess) e = dalloc(&g.llpart, (size_t)num_chains * g.splits);
  }
  if (gp) {
    if (e == hipSuccess) e = gp_scratch_alloc(ws->gp, plan->gp, num_chains);
  }
  if (e != hipSuccess) {
    delete ws;
    return fail(UT_E_HIP, std::string("ut_workspace_create: ") + hipGetErrorString(e));
  }
  *out = ws;
  return UT_OK;
}

int ut_workspace_destroy(ut_workspace* ws) {
  delete ws;
  return UT_OK;
}

// Logjoint (and gradient) of all chains at fp32 positions q (C x P, device), with
// the plan's gradient engine: the matrix-core GLM, or one wavefront per chain.
static hipError_t model_grad(ut_workspace* ws, const float* q, float* lj, float* grad, bool need_lj) {
  if (ws->dev.kind == MODEL_GLM_MFMA)
    return launch_glm_grad(ws->plan->glm, q, grad, lj, ws->C, need_lj, ws->glm, ws->stream);
  if (ws->dev.kind == MODEL_GP) return launch_gp_grad(ws->plan->gp, q, grad, lj, ws->C, ws->gp, ws->stream);
  return launch_density_f32(ws->dev, q, lj, grad, ws->C, ws->stream);
}

// The samplers' gradient engine.  With a flow attached, positions are base-space
// z: q = f(z), the model's logjoint and gradient at q, then
//   log p_z(z) = log p(q) + log |det df/dz|,  grad_z = (dq/dz)^T grad_q + grad_z logdet
// (the NeuTra reparameterisation).
// The flow forward and the model's gradient at its rows (into f.sc.g): Neal's funnel, whose
// terms are row-local, is evaluated in the 16-row forward kernel's epilogue (the density
// kernel's arithmetic); every other model by its own gradient engine after the forward.
static hipError_t flow_forward_grad(ut_workspace* ws, const float* q, float* lj, bool need_lj) {
  Flow& f = ws->flow->f;
  if (ws->fused_model && ws->dev.kind == MODEL_FUNNEL && f.m16 && (ws->dev.K == 1 || ws->dev.K == 2)) {
    FlowRowModel rm;
    rm.K = ws->dev.K;
    rm.dim = ws->dev.dim;
    rm.v_scale = ws->dev.v_scale;
    rm.total_const = ws->dev.total_const;
    const hipError_t e = flow_forward_model(f, q, rm, f.sc.g, lj, ws->C, ws->stream);
    if (e != hipErrorNotSupported) return e;
  }
  hipError_t e = flow_forward(f, q, nullptr, nullptr, ws->C, ws->stream);
  return e == hipSuccess ? model_grad(ws, f.sc.x, lj, f.sc.g, need_lj) : e;
}

static hipError_t flow_grad_eager(ut_workspace* ws, const float* q, float* lj, float* grad, bool need_lj) {
  Flow& f = ws->flow->f;
  hipError_t e = flow_forward_grad(ws, q, lj, need_lj);
  if (e == hipSuccess) e = flow_backward(f, f.sc.g, grad, ws->C, ws->stream);
  if (e == hipSuccess && need_lj) e = flow_add_logdet(lj, f.sc.logdet, ws->C, ws->stream);
  return e;
}

// With a flow attached, the ~40 launches of one batched gradient are host-bound, so
// the sequence is captured once per buffer set as a HIP graph (kernels and
// device-to-device copies only) and replayed.
static hipError_t batched_grad(ut_workspace* ws, const float* q, float* lj, float* grad, bool need_lj) {
  if (!ws->flow) return model_grad(ws, q, lj, grad, need_lj);
  if (!ws->use_graphs) return flow_grad_eager(ws, q, lj, grad, need_lj);
  const Flow& f = ws->flow->f;
  for (const ut_workspace::GradGraph& g : ws->graphs) {
    if (g.flow != &f || g.gen != f.gen) {
      ws->clear_graphs();
      break;
    }
  }
  for (const ut_workspace::GradGraph& g : ws->graphs)
    if (!g.fwd_only && g.q == q && g.lj == lj && g.grad == grad && g.need_lj == need_lj)
      return hipGraphLaunch(g.exec, ws->stream);
  hipError_t e = hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t ec = flow_grad_eager(ws, q, lj, grad, need_lj);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(ws->stream, &graph);
  if (ec != hipSuccess || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return ec != hipSuccess ? ec : e;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return e;
  ws->graphs.push_back({q, lj, grad, need_lj, false, &f, f.gen, exec});
  return hipGraphLaunch(exec, ws->stream);
}

// A trajectory's last gradient in the lock-step samplers (log density included): for Neal's
// funnel under a flow on the 16-row kernels one fused kernel (forward, model, backward and the
// log-determinant), else the batched gradient.
static hipError_t batched_grad_final(ut_workspace* ws, const LockstepState& ls) {
  if (ws->flow && ws->flow->f.m16 && ws->fused_model && ws->fused_step && ws->dev.kind == MODEL_FUNNEL &&
      (ws->dev.K == 1 || ws->dev.K == 2)) {
    FlowRowModel rm;
    rm.K = ws->dev.K;
    rm.dim = ws->dev.dim;
    rm.v_scale = ws->dev.v_scale;
    rm.total_const = ws->dev.total_const;
    const hipError_t e = flow16_grad_step(ws->flow->f, ls.wq, rm, ls.wlj, ls.wg, ws->C, nullptr, ws->stream);
    if (e != hipErrorNotSupported) return e;
  }
  return batched_grad(ws, ls.wq, ls.wlj, ls.wg, true);
}

// One leapfrog step of the lock-step samplers after the first: the batched gradient at
// ls.wq, then the full kick and drift (k_ls_step).  With a flow on the 16-row kernels the
// step runs in the flow backward's epilogue: the flow forward and the model gradient are
// replayed from a HIP graph, and the backward (with the step size as an argument) is
// launched after it; the arithmetic is k_ls_step's, so the draws are the same.
static hipError_t batched_grad_step(ut_workspace* ws, const LockstepState& ls, int P, int C, float eps) {
  if (!(ws->flow && ws->flow->f.m16 && ws->fused_step)) {
    hipError_t e = batched_grad(ws, ls.wq, ls.wlj, ls.wg, false);
    return e == hipSuccess ? launch_ls_step(ls, P, C, eps, ws->stream) : e;
  }
  Flow& f = ws->flow->f;
  if (ws->fused_model && ws->dev.kind == MODEL_FUNNEL && (ws->dev.K == 1 || ws->dev.K == 2)) {
    // the funnel: forward, model, backward and step in one kernel per 16-row tile
    FlowRowModel rm;
    rm.K = ws->dev.K;
    rm.dim = ws->dev.dim;
    rm.v_scale = ws->dev.v_scale;
    rm.total_const = ws->dev.total_const;
    Flow16Step stp;
    stp.wq = ls.wq;
    stp.wp = ls.wp;
    stp.im_f = ls.im_f;
    stp.valid = ls.valid;
    stp.eps = eps;
    const hipError_t e = flow16_grad_step(f, ls.wq, rm, ls.wlj, ls.wg, ws->C, &stp, ws->stream);
    if (e != hipErrorNotSupported) return e;
  }
  hipError_t e = hipSuccess;
  auto forward = [&]() -> hipError_t { return flow_forward_grad(ws, ls.wq, ls.wlj, false); };
  if (!ws->use_graphs) {
    e = forward();
  } else {
    for (const ut_workspace::GradGraph& g : ws->graphs) {
      if (g.flow != &f || g.gen != f.gen) {
        ws->clear_graphs();
        break;
      }
    }
    const ut_workspace::GradGraph* hit = nullptr;
    for (const ut_workspace::GradGraph& g : ws->graphs)
      if (g.fwd_only && g.q == ls.wq && g.lj == ls.wlj) hit = &g;
    if (hit) {
      e = hipGraphLaunch(hit->exec, ws->stream);
    } else {
      e = hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return e;
      const hipError_t ec = forward();
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(ws->stream, &graph);
      if (ec != hipSuccess || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return ec != hipSuccess ? ec : e;
      }
      hipGraphExec_t exec = nullptr;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return e;
      ws->graphs.push_back({ls.wq, ls.wlj, nullptr, false, true, &f, f.gen, exec});
      e = hipGraphLaunch(exec, ws->stream);
    }
  }
  if (e != hipSuccess) return e;
  if (!f.sc.fwd16) {  // the forward ran other kernels: their backward, then the step
    e = flow_backward(f, f.sc.g, ls.wg, ws->C, ws->stream);
    return e == hipSuccess ? launch_ls_step(ls, P, C, eps, ws->stream) : e;
  }
  Flow16Step stp;
  stp.wq = ls.wq;
  stp.wp = ls.wp;
  stp.im_f = ls.im_f;
  stp.valid = ls.valid;
  stp.eps = eps;
  return flow16_backward_step(f, f.sc.g, ls.wg, ws->C, stp, ws->stream);
}

static int density_call(ut_workspace* ws, const void* params, void* values, void* gradients) {
  if (!ws || !params || !values) return fail(UT_E_ARG, "density: null pointer");
  const size_t PC = (size_t)ws->P * ws->C;
  const size_t es = ws->precision == UT_FLOAT32 ? 4 : 8;
  UT_HIP(hipMemcpyAsync(ws->d_params, params, PC * es, hipMemcpyHostToDevice, ws->stream));
  hipError_t e;
  if (ws->dev.kind == MODEL_GLM_MFMA || ws->dev.kind == MODEL_GP)
    e = model_grad(ws, (const float*)ws->d_params, (float*)ws->d_values, (float*)ws->d_grads, true);
  else if (ws->precision == UT_FLOAT32)
    e = launch_density_f32(ws->dev, (const float*)ws->d_params, (float*)ws->d_values,
                           gradients ? (float*)ws->d_grads : nullptr, ws->C, ws->stream);
  else
    e = launch_density_f64(ws->dev, (const double*)ws->d_params, (double*)ws->d_values,
                           gradients ? (double*)ws->d_grads : nullptr, ws->C, ws->stream);
  UT_HIP(e);
  UT_HIP(hipMemcpyAsync(values, ws->d_values, (size_t)ws->C * es, hipMemcpyDeviceToHost, ws->stream));
  if (gradients)
    UT_HIP(hipMemcpyAsync(gradients, ws->d_grads, PC * es, hipMemcpyDeviceToHost, ws->stream));
  UT_HIP(hipStreamSynchronize(ws->stream));
  return UT_OK;
}

int ut_logjoint(ut_workspace* ws, const void* params, void* values) {
  return density_call(ws, params, values, nullptr);
}

int ut_logjoint_gradient(ut_workspace* ws, const void* params, void* values, void* gradients) {
  if (!gradients) return fail(UT_E_ARG, "ut_logjoint_gradient: null gradients");
  return density_call(ws, params, values, gradients);
}

int ut_gradient_kernel_time(ut_workspace* ws, const void* params, int32_t reps, int32_t need_logjoint,
                            double* us_per_call) {
  if (!ws || !params || !us_per_call) return fail(UT_E_ARG, "ut_gradient_kernel_time: null pointer");
  if (reps < 1) return fail(UT_E_ARG, "ut_gradient_kernel_time: reps must be >= 1");
  if (ws->precision != UT_FLOAT32) return fail(UT_E_PRECISION, "ut_gradient_kernel_time: float32 workspaces only");
  const size_t PC = (size_t)ws->P * ws->C;
  UT_HIP(hipMemcpyAsync(ws->d_params, params, PC * 4, hipMemcpyHostToDevice, ws->stream));
  const float* q = (const float*)ws->d_params;
  float* lj = (float*)ws->d_values;
  float* g = (float*)ws->d_grads;
  const bool nl = need_logjoint != 0;
  hipEvent_t ev[2];
  for (auto& e : ev) UT_HIP(hipEventCreate(&e));
  UT_HIP(batched_grad(ws, q, lj, g, nl));  // warm (and capture, when graphs are used)
  UT_HIP(hipEventRecord(ev[0], ws->stream));
  for (int r = 0; r < reps; ++r) UT_HIP(batched_grad(ws, q, lj, g, nl));
  UT_HIP(hipEventRecord(ev[1], ws->stream));
  UT_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.0f;
  UT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *us_per_call = 1e3 * ms / reps;
  return UT_OK;
}

int64_t ut_noalloc_scratch_bytes(const ut_plan* plan, int32_t num_chains) {
  if (!plan || num_chains <= 0) return -1;
  if (plan->dev.kind == MODEL_GP) return -1;
  if (plan->dev.kind != MODEL_GLM_MFMA) return 0;
  const size_t PC = (size_t)plan->P * (size_t)num_chains;
  const size_t S = (size_t)glm_splits(plan->glm.N, num_chains);
  return (int64_t)(2 * align256(PC * sizeof(uint16_t)) + align256(S * PC * sizeof(float)) +
                   align256(S * (size_t)num_chains * sizeof(double)));
}

int ut_logjoint_gradient_noalloc(const ut_plan* plan, int32_t num_chains, const float* params, float* values,
                                 float* gradients, void* stream) {
  if (!plan || !params || !values || !gradients)
    return fail(UT_E_ARG, "ut_logjoint_gradient_noalloc: null pointer");
  if (num_chains <= 0) return fail(UT_E_DIM, "ut_logjoint_gradient_noalloc: num_chains must be positive");
  if (plan->dev.kind == MODEL_GP)
    return fail(UT_E_UNSUPPORTED,
                "ut_logjoint_gradient_noalloc: the GP plan's factorisation scratch (about 4.5 MiB per chain) "
                "does not fit the fixed scratch; use ut_logjoint_gradient");
  hipStream_t s = (hipStream_t)stream;
  if (plan->dev.kind != MODEL_GLM_MFMA) {
    // wavefront-per-chain density kernels: no scratch beyond inputs and outputs
    UT_HIP(launch_density_f32(plan->dev, params, values, gradients, num_chains, s));
    return UT_OK;
  }
  if (num_chains % 128 != 0)
    return fail(UT_E_DIM, "ut_logjoint_gradient_noalloc: the matrix-core GLM plan needs num_chains a multiple of 128");
  const int64_t need = ut_noalloc_scratch_bytes(plan, num_chains);
  if (need > (int64_t)UT_NOALLOC_SCRATCH_BYTES)
    return fail(UT_E_DIM, "ut_logjoint_gradient_noalloc: " + std::to_string(num_chains) + " chains need " +
                              std::to_string(need) + " bytes of scratch, more than the fixed " +
                              std::to_string((long long)UT_NOALLOC_SCRATCH_BYTES) + " (UT_NOALLOC_SCRATCH_BYTES)");
  void* base = nullptr;
  const int rc = noalloc_scratch(&base);
  if (rc != UT_OK) return rc;
  // carve the scratch: bf16 hi and lo positions, per-split gradient and
  // log-likelihood partials.  Every byte used is written by the kernels before it
  // is read (each split stores its whole partial block), so nothing is cleared.
  const size_t PC = (size_t)plan->P * (size_t)num_chains;
  GlmScratch g;
  g.splits = glm_splits(plan->glm.N, num_chains);
  unsigned char* p = (unsigned char*)base;
  g.bhi = (uint16_t*)p;
  p += align256(PC * sizeof(uint16_t));
  g.blo = (uint16_t*)p;
  p += align256(PC * sizeof(uint16_t));
  g.gpart = (float*)p;
  p += align256((size_t)g.splits * PC * sizeof(float));
  g.llpart = (double*)p;
  UT_HIP(launch_glm_grad(plan->glm, params, gradients, values, num_chains, true, g, s));
  return UT_OK;
}

void ut_sampler_opts_default(ut_sampler_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_tree_depth = 10;
  o->num_leapfrog_steps = 10;
  o->step_size = 0.1;
  o->target_accept = 0.8;
  o->divergence_threshold = 1000.0;
  o->mass_matrix_regularization = 1e-3;
  o->mass_matrix_min_samples = 10;
  o->adapt_step_size = 1;
  o->adapt_mass_matrix = 1;
  o->find_reasonable_step_size = 0;
  o->init_strategy = UT_INIT_PRIOR;
  o->init_max_retries = 100;
  o->initial_params = nullptr;
  o->seed = 0;
  o->callback = nullptr;
  o->callback_user = nullptr;
  o->callback_every = 10;
}

int ut_host_alloc(int64_t bytes, void** out) {
  if (!out || bytes < 0) return fail(UT_E_ARG, "ut_host_alloc: invalid argument");
  *out = nullptr;
  if (bytes == 0) return UT_OK;
  hipError_t e = hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault);
  if (e != hipSuccess) return fail(UT_E_HIP, std::string("ut_host_alloc: ") + hipGetErrorString(e));
  return UT_OK;
}

int ut_host_free(void* p) {
  if (!p) return UT_OK;
  hipError_t e = hipHostFree(p);
  if (e != hipSuccess) return fail(UT_E_HIP, std::string("ut_host_free: ") + hipGetErrorString(e));
  return UT_OK;
}

int ut_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t first, int64_t n,
                   uint32_t* blocks, float* uniform_f32, float* normal_f32, double* uniform_f64) {
  if (n < 0) return fail(UT_E_ARG, "ut_philox_fill: n < 0");
  if (n == 0) return UT_OK;
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_philox_fill: no HIP device visible");
  uint32_t* db = nullptr;
  float *du = nullptr, *dn = nullptr;
  double* dd = nullptr;
  hipError_t e = hipSuccess;
  if (blocks) e = dalloc(&db, (size_t)n * 4);
  if (e == hipSuccess && uniform_f32) e = dalloc(&du, (size_t)n);
  if (e == hipSuccess && normal_f32) e = dalloc(&dn, (size_t)n);
  if (e == hipSuccess && uniform_f64) e = dalloc(&dd, (size_t)n);
  if (e == hipSuccess) e = launch_philox_fill(chain, iteration, stream, first, n, db, du, dn, dd, nullptr);
  if (e == hipSuccess && db) e = hipMemcpy(blocks, db, (size_t)n * 16, hipMemcpyDeviceToHost);
  if (e == hipSuccess && du) e = hipMemcpy(uniform_f32, du, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dn) e = hipMemcpy(normal_f32, dn, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dd) e = hipMemcpy(uniform_f64, dd, (size_t)n * 8, hipMemcpyDeviceToHost);
  (void)hipFree(db); (void)hipFree(du); (void)hipFree(dn); (void)hipFree(dd);
  if (e != hipSuccess) return fail(UT_E_HIP, std::string("ut_philox_fill: ") + hipGetErrorString(e));
  return UT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- sampler driver
namespace {

struct Schedule {
  int initial_buffer = 0;
  std::vector<int> ends;
  int terminal_buffer = 0;
};

// Stan-style warmup windows (reference nuts/adaptation.jl:103-138).
Schedule warmup_schedule(int n) {
  Schedule s;
  if (n <= 0) return s;
  if (n < 20) {
    int ib = std::min(5, std::max(n - 1, 0));
    s.initial_buffer = ib;
    if (ib == n) { s.initial_buffer = n; return s; }
    s.ends.push_back(n);
    return s;
  }
  int ib = std::min(std::max(5, n / 10), n);
  int tb = std::min(std::max(5, n / 10), std::max(n - ib, 0));
  int slow = n - ib - tb;
  if (slow <= 0) { s.initial_buffer = n; return s; }
  s.initial_buffer = ib;
  s.terminal_buffer = tb;
  int next = std::min(25, slow);
  int start = ib + 1;
  int remaining = slow;
  while (remaining > 0) {
    int size = remaining <= (3 * next) / 2 ? remaining : next;
    int end = start + size - 1;
    s.ends.push_back(end);
    remaining -= size;
    start = end + 1;
    next *= 2;
  }
  return s;
}

uint32_t run_seed_of(uint64_t seed) {
  return philox4x32((uint32_t)seed, (uint32_t)(seed >> 32), 0x5EEDu, 0u, 0u, 0u).w[0];
}

// One column's initial unconstrained position (reference hmc_support.jl:266-290):
// uniform(-2, 2) per coordinate, or a draw from the model prior mapped to the
// unconstrained space.  Draws are Philox coordinates (chain, run_seed + attempt,
// STREAM_INIT, index).
void draw_initial_column(const DevModelDesc& m, int strategy, uint32_t chain, uint32_t key,
                         double* col) {
  const int P = m.P;
  if (strategy == UT_INIT_UNIFORM) {
    for (int a = 0; a < P; ++a) col[a] = 4.0 * rand_uniform_d(chain, key, STREAM_INIT, (uint32_t)a) - 2.0;
    return;
  }
  auto z = [&](int idx) { return (double)rand_normal_f(chain, key, STREAM_INIT, (uint32_t)idx); };
  if (m.kind == MODEL_LOGISTIC || m.kind == MODEL_GLM_MFMA) {
    for (int a = 0; a < P; ++a) col[a] = m.loc + m.scale * z(a);
  } else if (m.kind == MODEL_HIER) {
    col[0] = m.mu_loc + m.mu_scale * z(0);
    const double tau = m.tau_scale * std::fabs(z(1) / z(P + 1));
    col[1] = log_d(tau);
    for (int j = 0; j < m.J; ++j) col[2 + j] = z(2 + j);
  } else if (m.kind == MODEL_GP) {
    for (int k = 0; k < m.D; ++k) col[k] = m.loc + m.scale * z(k);
    col[m.D] = log_d(m.v_scale * std::fabs(z(m.D)));
    for (int i = 0; i < (int)m.n; ++i) col[m.D + 1 + i] = z(m.D + 1 + i);
  } else if (m.kind == MODEL_PLAN) {
    plan_prior_draw(*m.host_plan, strategy, chain, key, STREAM_INIT, col);
  } else if (m.kind == MODEL_FUNNEL) {
    const double v = m.v_scale * z(0);
    col[0] = v;
    const double s = exp_d(v / 2.0);
    for (int i = 0; i < m.dim; ++i) col[1 + i] = s * z(1 + i);
  }
}

struct DeviceBuffers {
  std::vector<void*> ptrs;
  ~DeviceBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** p, size_t n) {
    hipError_t e = dalloc(p, n);
    if (e == hipSuccess && *p) ptrs.push_back(*p);
    return e;
  }
};

int validate_opts(const ut_workspace* ws, const ut_sampler_opts* o, const ut_chain_out* out,
                  bool nuts) {
  if (!ws || !o || !out) return fail(UT_E_ARG, "sampler: null pointer");
  if (ws->precision != UT_FLOAT32)
    return fail(UT_E_PRECISION, "sampler: the device samplers run in UT_FLOAT32 (binary64 energies)");
  if (!out->samples) return fail(UT_E_ARG, "sampler: out->samples is required");
  if (o->num_samples < 0 || o->num_warmup < 0) return fail(UT_E_ARG, "sampler: num_samples and num_warmup must be >= 0");
  if (!(o->step_size > 0) || !std::isfinite(o->step_size)) return fail(UT_E_ARG, "sampler: step_size must be finite and > 0");
  if (!(o->target_accept > 0 && o->target_accept < 1)) return fail(UT_E_ARG, "sampler: target_accept must lie in (0, 1)");
  if (!(o->divergence_threshold > 0)) return fail(UT_E_ARG, "sampler: divergence_threshold must be > 0");
  if (!(o->mass_matrix_regularization > 0)) return fail(UT_E_ARG, "sampler: mass_matrix_regularization must be > 0");
  if (o->mass_matrix_min_samples < 1) return fail(UT_E_ARG, "sampler: mass_matrix_min_samples must be >= 1");
  if (o->init_strategy != UT_INIT_PRIOR && o->init_strategy != UT_INIT_UNIFORM)
    return fail(UT_E_ARG, "sampler: init_strategy must be UT_INIT_PRIOR or UT_INIT_UNIFORM");
  if (o->init_max_retries < 0) return fail(UT_E_ARG, "sampler: init_max_retries must be >= 0");
  if (nuts) {
    if (o->max_tree_depth < 1 || o->max_tree_depth > 20) return fail(UT_E_ARG, "batched_nuts: max_tree_depth must lie in 1..20");
  } else {
    if (o->num_leapfrog_steps < 1) return fail(UT_E_ARG, "batched_hmc: num_leapfrog_steps must be >= 1");
  }
  if (ws->P > 1024) return fail(UT_E_UNSUPPORTED, "sampler: more than 1024 parameters");
  return UT_OK;
}

// Sampling iterations per launch: at least kChunk, and the sampling phase in at most
// kSampleLaunches launches.  A launch ends when its slowest chain does, and chains
// occasionally build much deeper NUTS trees than the rest, so short launches sum
// many per-launch maxima; longer ones average them out.  Each launch's draws are
// copied out while the next one runs.
constexpr int kChunk = 50;
constexpr int kSampleLaunches = 6;

// Initial positions, logjoint and gradient of every chain (reference
// _initial_batched_hmc_positions + the non-finite retry loop,
// api_batched_nuts.jl:184-200): columns whose logjoint or gradient is not finite
// are re-drawn with the next attempt key, up to init_max_retries times.
// UT_PROGRESS=n: the lock-step drivers wait for every n-th iteration to finish and
// report it on stderr (long runs stay visibly alive); unset or 0: silent.
int progress_every() {
  static const int n = [] {
    const char* e = std::getenv("UT_PROGRESS");
    return e ? std::max(0, std::atoi(e)) : 0;
  }();
  return n;
}

// Progress callback of the sampler options (reference _invoke_progress_callback,
// src/inference/api_single_chain.jl:12-30): due every `every` iterations of a phase
// and on its final iteration.  Divergences are counted on the host from the
// per-iteration flags the samplers already record (warmup: hist_flags, W x C;
// sampling: o_divergent, C x S), over the iterations since the last report.
struct Progress {
  ut_progress_fn fn = nullptr;
  void* user = nullptr;
  int every = 10;
  long long warm_divs = 0, samp_divs = 0;
  int warm_done = 0, samp_done = 0;
  std::vector<uint8_t> buf;
  bool due(int iteration, int total) const {
    return fn && iteration > 0 && ((every > 0 && iteration % every == 0) || iteration == total);
  }
  // next iteration in (from, to] at which a report is due, or `to`
  int next_stop(int from, int to, int total) const {
    if (!fn) return to;
    for (int i = from + 1; i < to; ++i)
      if (due(i, total)) return i;
    return to;
  }
  hipError_t warmup(const SamplerParams& sp, int C, int it, int W, double step, hipStream_t st) {
    if (!due(it, W)) return hipSuccess;
    const size_t n = (size_t)(it - warm_done) * C;
    buf.resize(n);
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && n && sp.hist_flags) {
      e = hipMemcpy(buf.data(), sp.hist_flags + (size_t)warm_done * C, n, hipMemcpyDeviceToHost);
      for (uint8_t f : buf) warm_divs += (f & FLAG_DIVERGENT) ? 1 : 0;
    }
    if (e != hipSuccess) return e;
    warm_done = it;
    fn(user, UT_PHASE_WARMUP, it, W, step, warm_divs);
    return hipSuccess;
  }
  hipError_t sample(const SamplerParams& sp, int C, int s, int S, double step, hipStream_t st) {
    if (!due(s, S)) return hipSuccess;
    const size_t w = (size_t)(s - samp_done);
    buf.resize(w * C);
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && w)
      e = hipMemcpy2D(buf.data(), w, sp.o_divergent + samp_done, (size_t)S, w, (size_t)C, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (uint8_t f : buf) samp_divs += f ? 1 : 0;
    samp_done = s;
    fn(user, UT_PHASE_SAMPLE, s, S, step, samp_divs);
    return hipSuccess;
  }
};

// mean over chains of the per-chain step sizes (the reported step of the
// per-chain samplers)
hipError_t mean_step(const double* d_step, int C, std::vector<double>& tmp, double* out) {
  tmp.resize(C);
  hipError_t e = hipMemcpy(tmp.data(), d_step, (size_t)C * 8, hipMemcpyDeviceToHost);
  double s = 0.0;
  for (double v : tmp) s += v;
  *out = C > 0 ? s / C : std::numeric_limits<double>::quiet_NaN();
  return e;
}

int init_chains(ut_workspace* ws, const ut_sampler_opts* o, uint32_t rs) {
  const int P = ws->P, C = ws->C;
  const DevModelDesc& m = ws->dev;
  hipStream_t stream = ws->stream;
  std::vector<double> pos((size_t)P * C);
  std::vector<float> pos_f((size_t)P * C);
  std::vector<float> lj_h(C), grad_h((size_t)P * C);
  if (o->initial_params) {
    std::memcpy(pos.data(), o->initial_params, pos.size() * sizeof(double));
  } else {
    for (int c = 0; c < C; ++c) draw_initial_column(m, o->init_strategy, (uint32_t)c, rs, &pos[(size_t)c * P]);
  }
  for (int attempt = 0;; ++attempt) {
    for (size_t i = 0; i < pos.size(); ++i) pos_f[i] = (float)pos[i];
    UT_HIP(hipMemcpyAsync(ws->st.pos, pos_f.data(), pos_f.size() * 4, hipMemcpyHostToDevice, stream));
    // with a flow the chains start at the base-space image of the drawn positions
    if (ws->flow) UT_HIP(flow_inverse(ws->flow->f, ws->st.pos, ws->st.pos, nullptr, C, stream));
    UT_HIP(batched_grad(ws, ws->st.pos, ws->st.lj, ws->st.grad, true));
    UT_HIP(hipMemcpyAsync(lj_h.data(), ws->st.lj, (size_t)C * 4, hipMemcpyDeviceToHost, stream));
    UT_HIP(hipMemcpyAsync(grad_h.data(), ws->st.grad, grad_h.size() * 4, hipMemcpyDeviceToHost, stream));
    UT_HIP(hipStreamSynchronize(stream));
    std::vector<int> bad;
    for (int c = 0; c < C; ++c) {
      bool ok = std::isfinite(lj_h[c]);
      for (int a = 0; a < P && ok; ++a) ok = std::isfinite(grad_h[(size_t)c * P + a]);
      if (!ok) bad.push_back(c);
    }
    if (bad.empty()) return UT_OK;
    if (o->initial_params || attempt >= o->init_max_retries) {
      char buf[256];
      std::snprintf(buf, sizeof(buf),
                    "initial batched parameters produced a non-finite unconstrained logjoint or gradient in %zu of %d chain(s)%s",
                    bad.size(), C, o->initial_params ? "" : " after the allowed re-draws");
      return fail(UT_E_NONFINITE_INIT, buf);
    }
    for (int c : bad)
      draw_initial_column(m, o->init_strategy, (uint32_t)c, rs + (uint32_t)(attempt + 1), &pos[(size_t)c * P]);
  }
}

// masked: NUTS with tree_strategy = masked -- every chain advances one leaf of its
// tree per step, the gradient of all chains from one batched call (any plan, a
// flow included); the adaptation is the persistent path's, step for step.
int run_sampler(ut_workspace* ws, const ut_sampler_opts* o, ut_chain_out* out, bool nuts, bool masked = false) {
  int rc = validate_opts(ws, o, out, nuts);
  if (rc) return rc;
  if (!masked && (ws->dev.kind == MODEL_GLM_MFMA || ws->dev.kind == MODEL_GP))
    return fail(UT_E_UNSUPPORTED,
                "this plan evaluates all chains at once; sample it with the lock-step samplers "
                "(ut_nuts_masked_run, ut_chees_run, ut_hmc_shared_run, ut_meads_run)");
  if (!masked && ws->flow)
    return fail(UT_E_UNSUPPORTED,
                "a flow preconditions the lock-step samplers (ut_nuts_masked_run, ut_chees_run, ut_hmc_shared_run); "
                "detach it first");
  if (masked && masked_slots(ws->dev) < 0)
    return fail(UT_E_UNSUPPORTED, "masked NUTS: more than 1024 parameters");
  {
    // the run's chain layout: every density evaluation of the run (initial points,
    // step search, transitions) takes the same wavefronts per chain
    const int depth = nuts && !masked ? o->max_tree_depth : -1;
    const int nw = chain_waves_for(ws->plan->dev, ws->C, ws->waves_request, depth);
    if (nw < 0)
      return fail(UT_E_ARG, "wavefronts per chain: " + std::to_string(ws->waves_request) +
                                " wavefronts' NUTS checkpoint stacks at max_tree_depth " +
                                std::to_string(o->max_tree_depth) + " exceed the 160 KB of LDS");
    ws->dev.waves = nw;
  }
  const double t_start = now_seconds();
  const int P = ws->P, C = ws->C, S = o->num_samples, W = o->num_warmup;
  const DevModelDesc& m = ws->dev;
  hipStream_t stream = ws->stream;
  const uint32_t rs = run_seed_of(o->seed);
  if ((rc = init_chains(ws, o, rs))) return rc;

  // ---- adaptation state
  std::vector<double> im_h(P, 1.0);
  std::vector<float> im_f(P, 1.0f);
  UT_HIP(hipMemcpyAsync(ws->d_im, im_h.data(), P * 8, hipMemcpyHostToDevice, stream));
  UT_HIP(hipMemcpyAsync(ws->st.im_f, im_f.data(), P * 4, hipMemcpyHostToDevice, stream));
  std::vector<double> step_h(C, o->step_size);
  UT_HIP(hipMemcpyAsync(ws->st.step, step_h.data(), (size_t)C * 8, hipMemcpyHostToDevice, stream));
  const int do_search = (o->find_reasonable_step_size || (W > 0 && o->adapt_step_size)) ? 1 : 0;

  DeviceBuffers buf;
  // ---- masked NUTS state and its lock-step drivers
  MnState ms;
  int* h_active = nullptr;
  struct PinnedInts {
    int* p = nullptr;
    ~PinnedInts() {
      if (p) (void)hipHostFree(p);
    }
  } pinned;
  const int mk = masked ? masked_slots(m) : 0;
  if (masked) {
    UT_HIP(buf.alloc(&ms.vec, (size_t)C * MN_NV * mk * WAVE));
    UT_HIP(buf.alloc(&ms.ck, (size_t)C * (o->max_tree_depth + 1) * 2 * mk * WAVE));
    UT_HIP(buf.alloc(&ms.ch, (size_t)C));
    UT_HIP(buf.alloc(&ms.active, 1));
    UT_HIP(buf.alloc(&ms.wq, (size_t)C * P));
    float *wg = nullptr, *wlj = nullptr;
    UT_HIP(buf.alloc(&wg, (size_t)C * P));
    UT_HIP(buf.alloc(&wlj, (size_t)C));
    ms.wg = wg;
    ms.wlj = wlj;
    UT_HIP(hipMemsetAsync(ms.vec, 0, (size_t)C * MN_NV * mk * WAVE * sizeof(float), stream));
    UT_HIP(hipMemsetAsync(ms.wq, 0, (size_t)C * P * sizeof(float), stream));
    UT_HIP(hipHostMalloc((void**)&pinned.p, 2 * sizeof(int), hipHostMallocDefault));
    h_active = pinned.p;
  }
  // lock-step loop: steps until no chain is running; the count of running chains
  // is read back one step behind, so the loop never waits on the device
  auto lockstep_until_idle = [&](auto&& step) -> int {
    hipEvent_t ev[2] = {nullptr, nullptr};
    UT_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    UT_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    struct EvRelease {
      hipEvent_t* e;
      ~EvRelease() {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
      }
    } release{ev};
    for (long long i = 0;; ++i) {
      UT_HIP(hipMemsetAsync(ms.active, 0, sizeof(int), stream));
      int r = step();
      if (r) return r;
      UT_HIP(hipMemcpyAsync(h_active + (i & 1), ms.active, sizeof(int), hipMemcpyDeviceToHost, stream));
      UT_HIP(hipEventRecord(ev[i & 1], stream));
      if (i > 0) {
        UT_HIP(hipEventSynchronize(ev[(i - 1) & 1]));
        if (h_active[(i - 1) & 1] == 0) break;
      }
    }
    return UT_OK;
  };
  auto search = [&](uint32_t itc, int do_it) -> int {
    if (!masked) {
      UT_HIP(launch_step_search(m, ws->st, C, itc, do_it, o->target_accept, stream));
      return UT_OK;
    }
    UT_HIP(hipMemsetAsync(ms.active, 0, sizeof(int), stream));
    UT_HIP(launch_mn_search_begin(m, ws->st, ms, C, P, itc, do_it, stream));
    if (!do_it) return UT_OK;
    UT_HIP(batched_grad(ws, ms.wq, (float*)ms.wlj, (float*)ms.wg, true));
    return lockstep_until_idle([&]() -> int {
      UT_HIP(launch_mn_search_step(m, ws->st, ms, C, P, itc, stream));
      UT_HIP(batched_grad(ws, ms.wq, (float*)ms.wlj, (float*)ms.wg, true));
      return UT_OK;
    });
  };
  if ((rc = search(rs, do_search))) return rc;
  SamplerParams sp;
  sp.run_seed = rs;
  sp.P = P;
  sp.C = C;
  sp.max_depth = o->max_tree_depth;
  sp.num_leapfrog = o->num_leapfrog_steps;
  sp.div_thr = o->divergence_threshold;
  sp.target_accept = o->target_accept;
  sp.adapt_step = o->adapt_step_size ? 1 : 0;
  sp.num_warmup = W;
  sp.num_samples = S;
  UT_HIP(buf.alloc(&sp.hist_pos, (size_t)W * C * P));
  UT_HIP(buf.alloc(&sp.hist_acc, (size_t)W * C));
  UT_HIP(buf.alloc(&sp.hist_flags, (size_t)W * C));
  UT_HIP(buf.alloc(&sp.draws, (size_t)S * C * P));
  UT_HIP(buf.alloc(&sp.o_lj, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_acc, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_energy, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_err, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_accepted, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_divergent, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.eval_counter, 1));
  UT_HIP(hipMemsetAsync(sp.eval_counter, 0, sizeof(unsigned long long), stream));
  UT_HIP(buf.alloc(&sp.o_depth, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_steps, (size_t)S * C));

  // a warmup window runs as one launch (nothing is copied out during warmup); masked
  // NUTS runs it in lock-step leaves
  auto masked_range = [&](int a, int b) -> int {
    UT_HIP(launch_mn_reset(ms, C, a, stream));
    return lockstep_until_idle([&]() -> int {
      UT_HIP(launch_mn_step(m, ws->st, sp, ms, b, stream));
      UT_HIP(batched_grad(ws, ms.wq, (float*)ms.wlj, (float*)ms.wg, true));
      return UT_OK;
    });
  };
  auto run_range = [&](int a, int b) -> int {
    if (masked) return masked_range(a, b);
    UT_HIP(nuts ? launch_nuts(m, ws->st, sp, a, b, stream) : launch_hmc(m, ws->st, sp, a, b, stream));
    return UT_OK;
  };

  // ---- warmup: one launch range per mass window, device replay at each window end
  const Schedule sch = warmup_schedule(W);
  std::vector<int> bounds;
  if (o->adapt_mass_matrix)
    for (int e : sch.ends) bounds.push_back(e);
  if (W > 0 && (bounds.empty() || bounds.back() != W)) bounds.push_back(W);
  int it = 1;
  int window = 0;
  Progress prog;
  prog.fn = o->callback;
  prog.user = o->callback_user;
  prog.every = o->callback_every;
  std::vector<double> step_tmp;
  auto report_step = [&](double* st) -> hipError_t {
    hipError_t e = hipStreamSynchronize(stream);
    return e == hipSuccess ? mean_step(ws->st.step, C, step_tmp, st) : e;
  };
  for (int b : bounds) {
    // with a progress callback the window's launch ends at every reported iteration
    for (int a = it; a < b;) {
      const int e = prog.next_stop(a - 1, b, W);
      if (e == b) break;
      if ((rc = run_range(a, e))) return rc;
      double st = 0.0;
      UT_HIP(report_step(&st));
      UT_HIP(prog.warmup(sp, C, e, W, st, stream));
      it = a = e + 1;
    }
    if ((rc = run_range(it, b))) return rc;
    if (prog.due(b, W)) {
      double st = 0.0;
      UT_HIP(report_step(&st));
      UT_HIP(prog.warmup(sp, C, b, W, st, stream));
    }
    const bool is_window_end = o->adapt_mass_matrix && window < (int)sch.ends.size() && sch.ends[window] == b;
    if (is_window_end) {
      const int start = window == 0 ? sch.initial_buffer + 1 : sch.ends[window - 1] + 1;
      const int wl = b - start + 1;
      UT_HIP(launch_mass_replay(sp.hist_pos, sp.hist_acc, sp.hist_flags, C, P, start, b, wl, nuts ? 0 : 1,
                                o->mass_matrix_regularization, o->mass_matrix_min_samples, ws->d_im,
                                ws->st.im_f, ws->d_summary, stream));
      if (o->adapt_step_size && b < W)
        if ((rc = search(rs + (uint32_t)b, 1))) return rc;
      ++window;
    }
    it = b + 1;
  }
  UT_HIP(hipStreamSynchronize(stream));
  const double t_warm = now_seconds();

  // ---- sampling: device time bracketed by events on the workspace stream; each
  // launch range's results are copied out on the copy stream while the next range
  // computes (strided 2-D copies: the outputs are chain-major).
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  UT_HIP(hipEventCreate(&ev0));
  UT_HIP(hipEventCreate(&ev1));
  std::vector<hipEvent_t> done;
  UT_HIP(hipEventRecord(ev0, stream));
  const int chunk = std::max(kChunk, (S + kSampleLaunches - 1) / kSampleLaunches);
  for (int s0 = 0, s1 = 0; s0 < S; s0 = s1) {
    s1 = prog.next_stop(s0, std::min(S, s0 + chunk), S);
    if ((rc = run_range(W + 1 + s0, W + s1))) return rc;
    hipEvent_t ev;
    UT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    done.push_back(ev);
    UT_HIP(hipEventRecord(ev, stream));
    UT_HIP(hipStreamWaitEvent(ws->copy_stream, ev, 0));
    auto copy2d = [&](void* dst, const void* src, size_t elem) -> hipError_t {
      if (!dst) return hipSuccess;
      const size_t pitch = (size_t)S * elem;
      return hipMemcpy2DAsync((char*)dst + (size_t)s0 * elem, pitch, (const char*)src + (size_t)s0 * elem,
                              pitch, (size_t)(s1 - s0) * elem, (size_t)C, hipMemcpyDeviceToHost,
                              ws->copy_stream);
    };
    // with a flow the draws (base space) and their logjoint are mapped after sampling
    if (!ws->flow) {
      UT_HIP(copy2d(out->samples, sp.draws, (size_t)P * 4));
      UT_HIP(copy2d(out->logjoint, sp.o_lj, 8));
    }
    UT_HIP(copy2d(out->accept_stat, sp.o_acc, 8));
    UT_HIP(copy2d(out->energy, sp.o_energy, 8));
    UT_HIP(copy2d(out->energy_error, sp.o_err, 8));
    UT_HIP(copy2d(out->accepted, sp.o_accepted, 1));
    UT_HIP(copy2d(out->divergent, sp.o_divergent, 1));
    UT_HIP(copy2d(out->tree_depth, sp.o_depth, 4));
    UT_HIP(copy2d(out->integration_steps, sp.o_steps, 4));
    if (prog.due(s1, S)) {
      double st = 0.0;
      UT_HIP(report_step(&st));
      UT_HIP(prog.sample(sp, C, s1, S, st, stream));
    }
  }
  UT_HIP(hipEventRecord(ev1, stream));
  if (ws->flow && S > 0) {
    // draws back to the model's space, C rows at a time; logjoint values lose the
    // log-determinant
    Flow& f = ws->flow->f;
    const size_t SC = (size_t)S * C;
    for (size_t r = 0; r < SC; r += (size_t)C) {
      float* rows = sp.draws + r * P;
      UT_HIP(flow_forward(f, rows, rows, nullptr, C, stream));
      UT_HIP(flow_sub_logdet(sp.o_lj + r, f.sc.logdet, C, stream));
    }
    UT_HIP(hipMemcpyAsync(out->samples, sp.draws, SC * P * 4, hipMemcpyDeviceToHost, stream));
    if (out->logjoint) UT_HIP(hipMemcpyAsync(out->logjoint, sp.o_lj, SC * 8, hipMemcpyDeviceToHost, stream));
  }
  unsigned long long evals = 0;
  UT_HIP(hipMemcpyAsync(&evals, sp.eval_counter, sizeof(evals), hipMemcpyDeviceToHost, stream));
  UT_HIP(hipStreamSynchronize(stream));
  UT_HIP(hipStreamSynchronize(ws->copy_stream));
  const double t_end = now_seconds();
  float kernel_ms = 0.0f;
  UT_HIP(hipEventElapsedTime(&kernel_ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  for (hipEvent_t ev : done) (void)hipEventDestroy(ev);
  out->sampling_kernel_seconds = 1e-3 * (double)kernel_ms;
  out->sampling_launches = (int32_t)done.size();
  if (out->step_size) UT_HIP(hipMemcpy(out->step_size, ws->st.step, (size_t)C * 8, hipMemcpyDeviceToHost));
  if (out->inverse_mass) UT_HIP(hipMemcpy(out->inverse_mass, ws->d_im, (size_t)P * 8, hipMemcpyDeviceToHost));
  out->warmup_seconds = t_warm - t_start;
  out->sampling_seconds = t_end - t_warm;
  out->sampling_gradient_evals = (int64_t)evals;
  return UT_OK;
}

// ---------------------------------------------------------------- ChEES driver
// Shared dual averaging on the host (reference _dual_averaging_state /
// _update_step_size! / _final_step_size, nuts/adaptation.jl:1-28), the same
// arithmetic as the kernels' per-chain dual_update.
struct HostDual {
  double mu = 0, log_step = 0, log_avg = 0, hbar = 0, it = 0, target = 0.8;
  void reset(double step, double tgt) {
    target = tgt;
    mu = log_d(10.0 * step);
    log_step = log_d(step);
    log_avg = log_step;
    hbar = 0.0;
    it = 0.0;
  }
  double update(double accept) {
    const double gamma = 0.05, t0 = 10.0, kappa = 0.75;
    it = it + 1.0;
    const double eta = 1.0 / (it + t0);
    hbar = (1.0 - eta) * hbar + eta * (target - accept);
    log_step = mu - sqrt_d(it) / gamma * hbar;
    const double eta_bar = pow_d(it, -kappa);
    log_avg = eta_bar * log_step + (1.0 - eta_bar) * log_avg;
    return exp_d(log_step);
  }
  double final_step() const { return exp_d(log_avg); }
};

struct PinnedStats {
  double* p = nullptr;
  ~PinnedStats() {
    if (p) (void)hipHostFree(p);
  }
};

int validate_chees(const ut_workspace* ws, const ut_chees_opts* co, const ut_chain_out* out) {
  if (!co) return fail(UT_E_ARG, "batched_chees: null options");
  int rc = validate_opts(ws, &co->base, out, false);
  if (rc) return rc;
  if (!(co->jitter_amount >= 0.0 && co->jitter_amount <= 1.0))
    return fail(UT_E_ARG, "batched_chees jitter_amount must be in [0, 1]");
  if (co->max_leapfrog_steps < 1) return fail(UT_E_ARG, "batched_chees max_leapfrog_steps must be >= 1");
  if (!(co->trajectory_decay_rate > 0.0 && co->trajectory_decay_rate <= 1.0))
    return fail(UT_E_ARG, "batched_chees trajectory_decay_rate must be in (0, 1]");
  if (!(co->trajectory_learning_rate > 0.0))
    return fail(UT_E_ARG, "batched_chees trajectory_learning_rate must be > 0");
  if (!(co->trajectory_adam_beta1 >= 0.0 && co->trajectory_adam_beta1 < 1.0))
    return fail(UT_E_ARG, "batched_chees trajectory_adam_beta1 must be in [0, 1)");
  if (!(co->trajectory_adam_beta2 >= 0.0 && co->trajectory_adam_beta2 < 1.0))
    return fail(UT_E_ARG, "batched_chees trajectory_adam_beta2 must be in [0, 1)");
  if (!(co->trajectory_adam_epsilon > 0.0))
    return fail(UT_E_ARG, "batched_chees trajectory_adam_epsilon must be > 0");
  if (ws->P > 256) return fail(UT_E_UNSUPPORTED, "batched_chees: more than 256 parameters");
  return UT_OK;
}

// Lock-step samplers: ChEES-HMC (reference batched_chees host loop,
// api_batched_chees.jl:435-700) and, with hmc = true, fixed-length HMC with shared
// adaptation (reference batched_hmc with per_chain_adaptation = false,
// api_batched_hmc.jl:184-330: L = num_leapfrog_steps, dual averaging to the mean
// acceptance with divergent chains counted as 0), both on the shared WarmupDriver
// (warmup.jl:262-327).  Every chain integrates the same number of leapfrog steps
// with the shared step size; the gradient between the trajectory kernels comes
// from the workspace's engine for all chains at once (through the flow when one is
// attached).  Warmup reads four binary64 scalars back per iteration (the ChEES
// numerator and denominator, the harmonic-mean and the mean acceptance) and runs
// the scalar state machines on the host; sampling launches without any host
// synchronisation.
int run_lockstep(ut_workspace* ws, const ut_chees_opts* co, ut_chain_out* out, double* trace, bool hmc) {
  int rc = validate_chees(ws, co, out);
  if (rc) return rc;
  if (ws->flow && (ws->C % FLOW_PAD != 0))
    return fail(UT_E_DIM, "a flow-preconditioned workspace needs num_chains a multiple of 128");
  const ut_sampler_opts* o = &co->base;
  const double t_start = now_seconds();
  const int P = ws->P, C = ws->C, S = o->num_samples, W = o->num_warmup;
  hipStream_t stream = ws->stream;
  const uint32_t rs = run_seed_of(o->seed);
  if ((rc = init_chains(ws, o, rs))) return rc;

  std::vector<double> im_h(P, 1.0);
  std::vector<float> im_f(P, 1.0f);
  UT_HIP(hipMemcpyAsync(ws->d_im, im_h.data(), P * 8, hipMemcpyHostToDevice, stream));
  UT_HIP(hipMemcpyAsync(ws->st.im_f, im_f.data(), P * 4, hipMemcpyHostToDevice, stream));

  DeviceBuffers buf;
  LockstepState ls;
  ls.pos = ws->st.pos;
  ls.grad = ws->st.grad;
  ls.lj = ws->st.lj;
  ls.im_f = ws->st.im_f;
  const size_t PC = (size_t)P * C;
  UT_HIP(buf.alloc(&ls.wq, PC));
  UT_HIP(buf.alloc(&ls.wp, PC));
  UT_HIP(buf.alloc(&ls.wg, PC));
  UT_HIP(buf.alloc(&ls.wlj, (size_t)C));
  UT_HIP(buf.alloc(&ls.h0, (size_t)C));
  UT_HIP(buf.alloc(&ls.valid, (size_t)C));
  UT_HIP(buf.alloc(&ls.initials, PC));
  LockstepRecord rec;
  UT_HIP(buf.alloc(&rec.accept_prob, (size_t)C));
  UT_HIP(buf.alloc(&rec.flags, (size_t)C));
  UT_HIP(buf.alloc(&rec.pfinite, (size_t)C));
  UT_HIP(buf.alloc(&rec.qfinite, (size_t)C));
  UT_HIP(buf.alloc(&rec.ifinite, (size_t)C));
  UT_HIP(buf.alloc(&rec.energy, (size_t)C));
  UT_HIP(buf.alloc(&rec.error, (size_t)C));
  UT_HIP(buf.alloc(&rec.imean, (size_t)P));
  UT_HIP(buf.alloc(&rec.pmean, (size_t)P));
  UT_HIP(buf.alloc(&rec.inv_sqrt, (size_t)P));
  UT_HIP(buf.alloc(&rec.term, (size_t)C));
  UT_HIP(buf.alloc(&rec.include, (size_t)C));
  double* d_stats = nullptr;
  UT_HIP(buf.alloc(&d_stats, 4));
  PinnedStats hs;
  UT_HIP(hipHostMalloc((void**)&hs.p, 4 * sizeof(double), hipHostMallocDefault));

  SamplerParams sp;
  sp.run_seed = rs;
  sp.P = P;
  sp.C = C;
  sp.num_warmup = W;
  sp.num_samples = S;
  UT_HIP(buf.alloc(&sp.hist_pos, (size_t)W * C * P));
  UT_HIP(buf.alloc(&sp.hist_acc, (size_t)W * C));
  UT_HIP(buf.alloc(&sp.hist_flags, (size_t)W * C));
  UT_HIP(buf.alloc(&sp.draws, (size_t)S * C * P));
  UT_HIP(buf.alloc(&sp.o_lj, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_acc, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_energy, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_err, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_accepted, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_divergent, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_depth, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_steps, (size_t)S * C));

  auto read_stats = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(hs.p, d_stats, 4 * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
  };

  // shared reasonable step size (reference _find_reasonable_batched_step_size,
  // nuts/adaptation.jl:303-375): one momentum draw, a one-step trajectory per trial,
  // doubling or halving until the mean acceptance (divergent chains count 0)
  // crosses 1/2, at most 20 doublings
  auto search = [&](double step, uint32_t itc, double* result) -> int {
    const double min_step = 1e-8, max_step = 1e3, target = 0.5;
    double direction = 0.0;
    for (int trial = 0; trial <= 20; ++trial) {
      const float eps = (float)step;
      UT_HIP(launch_ls_begin(ls, P, C, itc, STREAM_STEP_SEARCH, eps, 0, stream));
      UT_HIP(batched_grad(ws, ls.wq, ls.wlj, ls.wg, true));
      UT_HIP(launch_ls_end(ls, rec, P, C, itc, eps, o->divergence_threshold, 1, stream));
      UT_HIP(launch_chees_stats(ls, rec, ws->d_im, P, C, false, 1.0, d_stats, stream));
      UT_HIP(read_stats());
      const double mean_accept = hs.p[3];
      if (trial == 0) {
        direction = mean_accept > target ? 1.0 : -1.0;
      } else if ((direction > 0 && mean_accept <= target) || (direction < 0 && mean_accept >= target)) {
        break;
      }
      const double next = step * (direction > 0 ? 2.0 : 0.5);
      if (next < min_step || next > max_step) break;
      step = next;
    }
    *result = step < min_step ? min_step : (step > max_step ? max_step : step);
    return UT_OK;
  };

  double eps = o->step_size;
  if (o->find_reasonable_step_size || (W > 0 && o->adapt_step_size))
    if ((rc = search(eps, rs, &eps))) return rc;
  HostDual dual;
  dual.reset(eps, o->target_accept);
  const double T0 = (double)o->num_leapfrog_steps * eps;
  CheesTrajectory traj{T0, log_d(T0),
                       CheesAdam{co->trajectory_adam_beta1, co->trajectory_adam_beta2, co->trajectory_adam_epsilon,
                                 co->trajectory_learning_rate},
                       1, co->trajectory_decay_rate, co->max_leapfrog_steps};
  const Schedule sch = warmup_schedule(W);
  int window = 0;
  const bool adapt_traj = !hmc && W > 0 && co->adapt_trajectory_length;

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  UT_HIP(hipEventCreate(&ev0));
  UT_HIP(hipEventCreate(&ev1));
  double t_warm = t_start;
  long long evals = 0;
  int launches = 0;
  Progress prog;
  prog.fn = o->callback;
  prog.user = o->callback_user;
  prog.every = o->callback_every;
  for (int it = 1; it <= W + S; ++it) {
    if (progress_every() > 0 && it % progress_every() == 0) {
      UT_HIP(hipStreamSynchronize(stream));
      std::fprintf(stderr, "lock-step sampler: iteration %d of %d done\n", it - 1, W + S);
      std::fflush(stderr);
    }
    if (it == W + 1) {
      UT_HIP(hipStreamSynchronize(stream));
      t_warm = now_seconds();
      UT_HIP(hipEventRecord(ev0, stream));
    }
    const double jitter = chees_jitter(co->jitter_amount, it);
    const int L = hmc ? o->num_leapfrog_steps : chees_steps(jitter, traj.T, eps);
    const uint32_t itc = rs + (uint32_t)it;
    const bool warm = it <= W;
    const float eps_f = (float)eps;
    UT_HIP(launch_ls_begin(ls, P, C, itc, STREAM_MOMENTUM, eps_f, (warm && adapt_traj) ? 1 : 0, stream));
    for (int l = 1; l < L; ++l) UT_HIP(batched_grad_step(ws, ls, P, C, eps_f));
    UT_HIP(batched_grad_final(ws, ls));
    UT_HIP(launch_ls_end(ls, rec, P, C, itc, eps_f, o->divergence_threshold, 0, stream));
    UT_HIP(launch_ls_record(ls, rec, sp, it, L, stream));
    if (!warm) {
      evals += (long long)L * C;
      launches += 2 * L + 2;
      UT_HIP(prog.sample(sp, C, it - W, S, eps, stream));
      continue;
    }
    UT_HIP(launch_chees_stats(ls, rec, ws->d_im, P, C, adapt_traj, co->forward_velocity ? -1.0 : 1.0, d_stats,
                              stream));
    UT_HIP(read_stats());
    const double num = hs.p[0], den = hs.p[1], harmonic = hs.p[2], mean_accept = hs.p[3];
    if (o->adapt_step_size) eps = dual.update(hmc ? mean_accept : harmonic);
    if (o->adapt_mass_matrix && window < (int)sch.ends.size() && sch.ends[window] == it) {
      const int start = window == 0 ? sch.initial_buffer + 1 : sch.ends[window - 1] + 1;
      UT_HIP(launch_mass_replay(sp.hist_pos, sp.hist_acc, sp.hist_flags, C, P, start, it, it - start + 1, 1,
                                o->mass_matrix_regularization, o->mass_matrix_min_samples, ws->d_im,
                                ws->st.im_f, ws->d_summary, stream));
      ++window;
      if (o->adapt_step_size && it < W) {
        if ((rc = search(eps, rs + (uint32_t)it, &eps))) return rc;
        dual.reset(eps, o->target_accept);
      }
    }
    if (adapt_traj) traj.apply(num, den, jitter, eps);
    if (trace && !hmc) trace[it - 1] = traj.T;
    if (it == W && o->adapt_step_size) eps = dual.final_step();
    UT_HIP(prog.warmup(sp, C, it, W, eps, stream));
  }
  if (S == 0) {
    UT_HIP(hipStreamSynchronize(stream));
    t_warm = now_seconds();
    UT_HIP(hipEventRecord(ev0, stream));
  }
  UT_HIP(hipEventRecord(ev1, stream));
  const size_t SC = (size_t)S * C;
  if (ws->flow) {
    // draws back to the model's space, C rows at a time (rows are independent);
    // logjoint values lose the log-determinant
    Flow& f = ws->flow->f;
    for (size_t r = 0; r < SC; r += (size_t)C) {
      float* rows = sp.draws + r * P;
      UT_HIP(flow_forward(f, rows, rows, nullptr, C, stream));
      UT_HIP(flow_sub_logdet(sp.o_lj + r, f.sc.logdet, C, stream));
    }
  }
  auto copy = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
  };
  UT_HIP(copy(out->samples, sp.draws, SC * P * 4));
  UT_HIP(copy(out->logjoint, sp.o_lj, SC * 8));
  UT_HIP(copy(out->accept_stat, sp.o_acc, SC * 8));
  UT_HIP(copy(out->energy, sp.o_energy, SC * 8));
  UT_HIP(copy(out->energy_error, sp.o_err, SC * 8));
  UT_HIP(copy(out->accepted, sp.o_accepted, SC));
  UT_HIP(copy(out->divergent, sp.o_divergent, SC));
  UT_HIP(copy(out->tree_depth, sp.o_depth, SC * 4));
  UT_HIP(copy(out->integration_steps, sp.o_steps, SC * 4));
  if (out->inverse_mass) UT_HIP(copy(out->inverse_mass, ws->d_im, (size_t)P * 8));
  UT_HIP(hipStreamSynchronize(stream));
  const double t_end = now_seconds();
  float kernel_ms = 0.0f;
  UT_HIP(hipEventElapsedTime(&kernel_ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (out->step_size)
    for (int c = 0; c < C; ++c) out->step_size[c] = eps;
  out->warmup_seconds = t_warm - t_start;
  out->sampling_seconds = t_end - t_warm;
  out->sampling_gradient_evals = evals;
  out->sampling_kernel_seconds = 1e-3 * (double)kernel_ms;
  out->sampling_launches = launches;
  return UT_OK;
}

// ---------------------------------------------------------------- MEADS driver
// Reference batched_meads (api_batched_meads.jl:170-547): contiguous folds, the
// neighbour map k <- (k - 1) mod K, skipped fold (t - 1) mod K, initial persistent
// momentum and slice from Philox (iteration coordinate run_seed), then per
// iteration fold parameters -> refresh + half step -> gradient -> accept -> record.
int run_meads(ut_workspace* ws, const ut_meads_opts* mo, ut_chain_out* out, double* trace) {
  if (!ws || !mo || !out) return fail(UT_E_ARG, "batched_meads: null pointer");
  if (!out->samples) return fail(UT_E_ARG, "batched_meads: out->samples is required");
  if (ws->precision != UT_FLOAT32) return fail(UT_E_PRECISION, "batched_meads runs in UT_FLOAT32 (binary64 energies)");
  const int P = ws->P, C = ws->C, S = mo->num_samples, W = mo->num_warmup, K = mo->num_folds;
  if (K < 2) return fail(UT_E_ARG, "batched_meads num_folds must be >= 2");
  if (C < 2 * K) return fail(UT_E_ARG, "batched_meads needs num_chains >= 2 * num_folds");
  if (S <= 0) return fail(UT_E_ARG, "batched_meads num_samples must be > 0");
  if (W < 0) return fail(UT_E_ARG, "batched_meads num_warmup must be >= 0");
  if (mo->init_strategy != UT_INIT_PRIOR && mo->init_strategy != UT_INIT_UNIFORM)
    return fail(UT_E_ARG, "batched_meads init_strategy must be UT_INIT_PRIOR or UT_INIT_UNIFORM");
  if (mo->init_max_retries < 0) return fail(UT_E_ARG, "batched_meads init_max_retries must be >= 0");
  if (P > 1024) return fail(UT_E_UNSUPPORTED, "batched_meads: more than 1024 parameters");
  if (ws->flow && C % FLOW_PAD != 0)
    return fail(UT_E_DIM, "a flow-preconditioned workspace needs num_chains a multiple of 128");
  hipStream_t stream = ws->stream;
  const double t_start = now_seconds();
  const uint32_t rs = run_seed_of(mo->seed);
  ut_sampler_opts io;
  ut_sampler_opts_default(&io);
  io.init_strategy = mo->init_strategy;
  io.init_max_retries = mo->init_max_retries;
  io.initial_params = mo->initial_params;
  io.seed = mo->seed;
  int rc = init_chains(ws, &io, rs);
  if (rc) return rc;

  // folds: contiguous partition, the first C mod K folds one chain larger
  std::vector<int> start(K + 1, 0), fold_of(C);
  const int base = C / K, rem = C % K;
  for (int k = 0; k < K; ++k) start[k + 1] = start[k] + base + (k < rem ? 1 : 0);
  for (int k = 0; k < K; ++k)
    for (int c = start[k]; c < start[k + 1]; ++c) fold_of[c] = k;
  const int nmax = base + (rem ? 1 : 0);

  DeviceBuffers buf;
  MeadsState m;
  m.pos = ws->st.pos;
  m.grad = ws->st.grad;
  m.lj = ws->st.lj;
  const size_t PC = (size_t)P * C;
  UT_HIP(buf.alloc(&m.mom, PC));
  UT_HIP(buf.alloc(&m.slice, (size_t)C));
  UT_HIP(buf.alloc(&m.wq, PC));
  UT_HIP(buf.alloc(&m.wp, PC));
  UT_HIP(buf.alloc(&m.wg, PC));
  UT_HIP(buf.alloc(&m.wlj, (size_t)C));
  UT_HIP(buf.alloc(&m.refreshed, PC));
  UT_HIP(buf.alloc(&m.h0, (size_t)C));
  UT_HIP(buf.alloc(&m.valid, (size_t)C));
  UT_HIP(buf.alloc(&m.sigma, (size_t)K * P));
  UT_HIP(buf.alloc(&m.sigma_f, (size_t)K * P));
  UT_HIP(buf.alloc(&m.white, (size_t)2 * K * nmax * P));
  UT_HIP(buf.alloc(&m.gram_row, (size_t)2 * K * nmax));
  UT_HIP(buf.alloc(&m.gram, (size_t)2 * K * nmax * nmax));
  UT_HIP(buf.alloc(&m.fold, (size_t)4 * K));
  UT_HIP(buf.alloc(&m.fold_start, (size_t)K + 1));
  int* d_fold_of = nullptr;
  UT_HIP(buf.alloc(&d_fold_of, (size_t)C));
  UT_HIP(hipMemcpyAsync(m.fold_start, start.data(), (K + 1) * 4, hipMemcpyHostToDevice, stream));
  UT_HIP(hipMemcpyAsync(d_fold_of, fold_of.data(), (size_t)C * 4, hipMemcpyHostToDevice, stream));
  MeadsRecord rec;
  UT_HIP(buf.alloc(&rec.accept_prob, (size_t)C));
  UT_HIP(buf.alloc(&rec.flags, (size_t)C));
  UT_HIP(buf.alloc(&rec.energy_error, (size_t)C));
  const int T = W + S;
  double* d_trace = nullptr;
  UT_HIP(buf.alloc(&d_trace, (size_t)4 * K * T));

  SamplerParams sp;
  sp.P = P;
  sp.C = C;
  sp.num_warmup = W;
  sp.num_samples = S;
  UT_HIP(buf.alloc(&sp.draws, (size_t)S * C * P));
  UT_HIP(buf.alloc(&sp.o_lj, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_acc, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_energy, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_err, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_accepted, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_divergent, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_depth, (size_t)S * C));
  UT_HIP(buf.alloc(&sp.o_steps, (size_t)S * C));

  UT_HIP(launch_meads_init(m, P, C, rs, stream));
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  UT_HIP(hipEventCreate(&ev0));
  UT_HIP(hipEventCreate(&ev1));
  double t_warm = t_start;
  Progress prog;
  prog.fn = mo->callback;
  prog.user = mo->callback_user;
  prog.every = mo->callback_every;
  for (int it = 1; it <= T; ++it) {
    if (it == W + 1) {
      UT_HIP(hipStreamSynchronize(stream));
      t_warm = now_seconds();
      UT_HIP(hipEventRecord(ev0, stream));
    }
    if (progress_every() > 0 && it % progress_every() == 0) {
      UT_HIP(hipStreamSynchronize(stream));
      std::fprintf(stderr, "ut_meads_run: iteration %d of %d done\n", it - 1, T);
      std::fflush(stderr);
    }
    const int skip = (it - 1) % K;
    UT_HIP(launch_meads_fold_params(m, P, K, nmax, it, stream));
    UT_HIP(hipMemcpyAsync(d_trace + (size_t)4 * K * (it - 1), m.fold, (size_t)4 * K * 8, hipMemcpyDeviceToDevice,
                          stream));
    UT_HIP(launch_meads_begin(m, d_fold_of, P, C, rs + (uint32_t)it, stream));
    UT_HIP(batched_grad(ws, m.wq, m.wlj, m.wg, true));
    UT_HIP(launch_meads_end(m, rec, d_fold_of, P, C, skip, mo->divergence_threshold, stream));
    UT_HIP(launch_meads_record(m, rec, d_fold_of, sp, it, skip, stream));
    // MEADS has no single step size (one per fold) and records no warmup flags
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (it <= W) UT_HIP(prog.warmup(sp, C, it, W, nan, stream));
    else UT_HIP(prog.sample(sp, C, it - W, S, nan, stream));
  }
  UT_HIP(hipEventRecord(ev1, stream));
  const size_t SC = (size_t)S * C;
  if (ws->flow) {
    Flow& f = ws->flow->f;
    for (size_t r = 0; r < SC; r += (size_t)C) {
      float* rows = sp.draws + r * P;
      UT_HIP(flow_forward(f, rows, rows, nullptr, C, stream));
      UT_HIP(flow_sub_logdet(sp.o_lj + r, f.sc.logdet, C, stream));
    }
  }
  auto copy = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
  };
  UT_HIP(copy(out->samples, sp.draws, SC * P * 4));
  UT_HIP(copy(out->logjoint, sp.o_lj, SC * 8));
  UT_HIP(copy(out->accept_stat, sp.o_acc, SC * 8));
  UT_HIP(copy(out->energy, sp.o_energy, SC * 8));
  UT_HIP(copy(out->energy_error, sp.o_err, SC * 8));
  UT_HIP(copy(out->accepted, sp.o_accepted, SC));
  UT_HIP(copy(out->divergent, sp.o_divergent, SC));
  UT_HIP(copy(out->tree_depth, sp.o_depth, SC * 4));
  UT_HIP(copy(out->integration_steps, sp.o_steps, SC * 4));
  std::vector<double> tr((size_t)4 * K * T), sig((size_t)P);
  UT_HIP(copy(tr.data(), d_trace, tr.size() * 8));
  UT_HIP(copy(sig.data(), m.sigma, (size_t)P * 8));
  UT_HIP(hipStreamSynchronize(stream));
  const double t_end = now_seconds();
  float kernel_ms = 0.0f;
  UT_HIP(hipEventElapsedTime(&kernel_ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  double final_step = 0.0;
  for (int it = 0; it < T; ++it) {
    double ssum = 0.0, asum = 0.0, mx = 0.0;
    for (int k = 0; k < K; ++k) {
      ssum += tr[(size_t)4 * K * it + 4 * k + 0];
      asum += tr[(size_t)4 * K * it + 4 * k + 1];
      mx = std::max(mx, tr[(size_t)4 * K * it + 4 * k + 3]);
    }
    if (trace) {
      trace[3 * (size_t)it + 0] = ssum / K;
      trace[3 * (size_t)it + 1] = asum / K;
      trace[3 * (size_t)it + 2] = mx;
    }
    final_step = ssum / K;
  }
  if (out->step_size)
    for (int c = 0; c < C; ++c) out->step_size[c] = final_step;
  if (out->inverse_mass)
    for (int a = 0; a < P; ++a) out->inverse_mass[a] = sig[a] * sig[a];
  out->warmup_seconds = t_warm - t_start;
  out->sampling_seconds = t_end - t_warm;
  out->sampling_gradient_evals = (int64_t)S * C;
  out->sampling_kernel_seconds = 1e-3 * (double)kernel_ms;
  out->sampling_launches = S * 4;
  return UT_OK;
}

}  // namespace

extern "C" {

void ut_meads_opts_default(ut_meads_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->num_folds = 4;
  o->init_strategy = UT_INIT_PRIOR;
  o->init_max_retries = 100;
  o->divergence_threshold = 1000.0;
  o->callback_every = 10;
}

int ut_meads_run(ut_workspace* ws, const ut_meads_opts* opts, ut_chain_out* out, double* adapt_trace) {
  return run_meads(ws, opts, out, adapt_trace);
}

void ut_chees_opts_default(ut_chees_opts* co) {
  if (!co) return;
  std::memset(co, 0, sizeof(*co));
  ut_sampler_opts_default(&co->base);
  co->base.target_accept = 0.651;
  co->jitter_amount = 1.0;
  co->adapt_trajectory_length = 1;
  co->max_leapfrog_steps = 1000;
  co->trajectory_decay_rate = 0.5;
  co->trajectory_learning_rate = 0.5;
  co->trajectory_adam_beta1 = 0.0;
  co->trajectory_adam_beta2 = 0.95;
  co->trajectory_adam_epsilon = 1e-8;
  co->forward_velocity = 0;
}

int ut_chees_run(ut_workspace* ws, const ut_chees_opts* opts, ut_chain_out* out, double* trajectory_trace) {
  return run_lockstep(ws, opts, out, trajectory_trace, false);
}

int ut_hmc_shared_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out) {
  if (!opts) return fail(UT_E_ARG, "batched_hmc: null options");
  ut_chees_opts co;
  ut_chees_opts_default(&co);
  co.base = *opts;
  co.adapt_trajectory_length = 0;
  co.jitter_amount = 0.0;
  return run_lockstep(ws, &co, out, nullptr, true);
}

int ut_nuts_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out) {
  return run_sampler(ws, opts, out, true);
}

int ut_hmc_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out) {
  return run_sampler(ws, opts, out, false);
}

int ut_nuts_masked_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out) {
  if (ws && ws->flow && ws->C % 128 != 0)
    return fail(UT_E_DIM, "masked NUTS with a flow: num_chains must be a multiple of 128");
  return run_sampler(ws, opts, out, true, true);
}

}  // extern "C"

extern "C" {

int ut_gemm_bf16(const uint16_t* a, const uint16_t* bt, float* c, int32_t m, int32_t n, int32_t k,
                 int32_t splits) {
  if (!a || !bt || !c) return fail(UT_E_ARG, "ut_gemm_bf16: null pointer");
  if (m <= 0 || n <= 0 || k <= 0 || m % 128 || n % 128 || k % 32 || splits < 1)
    return fail(UT_E_DIM, "ut_gemm_bf16: m, n must be multiples of 128, k of 32, splits >= 1");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_gemm_bf16: no HIP device visible");
  DeviceBuffers buf;
  uint16_t *da = nullptr, *db = nullptr;
  float* dc = nullptr;
  const size_t mn = (size_t)m * n;
  UT_HIP(buf.alloc(&da, (size_t)m * k));
  UT_HIP(buf.alloc(&db, (size_t)n * k));
  UT_HIP(buf.alloc(&dc, mn * splits));
  UT_HIP(hipMemcpy(da, a, (size_t)m * k * 2, hipMemcpyHostToDevice));
  UT_HIP(hipMemcpy(db, bt, (size_t)n * k * 2, hipMemcpyHostToDevice));
  UT_HIP(launch_gemm_nt(da, k, db, k, dc, n, m, n, k, splits, (long long)mn, nullptr));
  std::vector<float> part(mn * splits);
  UT_HIP(hipMemcpy(part.data(), dc, part.size() * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < mn; ++i) {
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * mn + i];
    c[i] = s;
  }
  return UT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- flow preconditioner
namespace {

int flow_rows_check(const ut_flow* fl, int64_t n, const char* who) {
  if (!fl) return fail(UT_E_ARG, std::string(who) + ": null flow");
  if (n <= 0 || n % FLOW_PAD != 0)
    return fail(UT_E_DIM, std::string(who) + ": the row count must be a positive multiple of 128");
  if (n > (1 << 20)) return fail(UT_E_DIM, std::string(who) + ": at most 1048576 rows per call");
  return UT_OK;
}

}  // namespace

extern "C" {

int ut_flow_create(int32_t num_params, int32_t num_layers, int32_t hidden, double scale_clamp, uint64_t seed,
                   ut_flow** out) {
  if (!out) return fail(UT_E_ARG, "ut_flow_create: null pointer");
  *out = nullptr;
  if (num_params < 2) return fail(UT_E_DIM, "ut_flow_create: at least 2 parameters");
  if (num_layers < 0) return fail(UT_E_ARG, "ut_flow_create: num_layers must be >= 0");
  if (hidden < 128 || hidden % 128 != 0) return fail(UT_E_ARG, "ut_flow_create: hidden must be a positive multiple of 128");
  if (!(scale_clamp > 0)) return fail(UT_E_ARG, "ut_flow_create: scale_clamp must be > 0");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_flow_create: no HIP device visible");
  auto* fl = new ut_flow();
  Flow& f = fl->f;
  f.clamp = (float)scale_clamp;
  hipError_t e = flow_alloc(f, num_params, num_layers, hidden);
  if (e != hipSuccess) {
    delete fl;
    return fail(UT_E_HIP, std::string("ut_flow_create: ") + hipGetErrorString(e));
  }
  // identity start: W2 = 0, b = 0, mu = 0, ls = 0; W1 ~ N(0, 1/np) from Philox
  std::vector<float> prm((size_t)f.nparams, 0.0f);
  for (int l = 0; l < num_layers; ++l) {
    const FlowLayer& ly = f.layers[l];
    const int kp = (ly.np + FLOW_PAD - 1) / FLOW_PAD * FLOW_PAD;
    const float sd = 1.0f / sqrt_f((float)ly.np);
    for (int j = 0; j < hidden; ++j)
      for (int k = 0; k < ly.np; ++k)
        prm[(size_t)ly.oW1 + (size_t)j * kp + k] =
            sd * rand_normal_f((uint32_t)seed, (uint32_t)(seed >> 32) + (uint32_t)l, 10u, (uint32_t)(j * kp + k));
  }
  e = hipMemcpy(f.params, prm.data(), prm.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = flow_sync_weights(f, f.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(f.stream);
  if (e != hipSuccess) {
    delete fl;
    return fail(UT_E_HIP, std::string("ut_flow_create: ") + hipGetErrorString(e));
  }
  *out = fl;
  return UT_OK;
}

int ut_flow_create_fullrank(int32_t num_params, ut_flow** out) {
  if (!out) return fail(UT_E_ARG, "ut_flow_create_fullrank: null pointer");
  *out = nullptr;
  if (num_params < 2) return fail(UT_E_DIM, "ut_flow_create_fullrank: at least 2 parameters");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_flow_create_fullrank: no HIP device visible");
  auto* fl = new ut_flow();
  Flow& f = fl->f;
  // H only sizes unused scratch here (no coupling layers)
  hipError_t e = flow_alloc(f, num_params, 0, 128, true);
  // mu = 0, ls = 0 and a zero factor: the mean-field start (src/inference/vi.jl:459-460)
  if (e == hipSuccess) e = hipMemset(f.params, 0, (size_t)f.nparams * 4);
  if (e != hipSuccess) {
    delete fl;
    return fail(UT_E_HIP, std::string("ut_flow_create_fullrank: ") + hipGetErrorString(e));
  }
  *out = fl;
  return UT_OK;
}

int ut_flow_destroy(ut_flow* flow) {
  delete flow;
  return UT_OK;
}

int ut_flow_num_weights(const ut_flow* flow, int64_t* n) {
  if (!flow || !n) return fail(UT_E_ARG, "ut_flow_num_weights: null pointer");
  *n = flow->f.nparams;
  return UT_OK;
}

int ut_flow_layer_layout(const ut_flow* flow, int32_t layer, int32_t* np, int32_t* na, int64_t* offsets) {
  if (!flow || !np || !na || !offsets) return fail(UT_E_ARG, "ut_flow_layer_layout: null pointer");
  if (layer < 0 || layer >= flow->f.L) return fail(UT_E_ARG, "ut_flow_layer_layout: layer out of range");
  const FlowLayer& ly = flow->f.layers[layer];
  *np = ly.np;
  *na = ly.na;
  offsets[0] = ly.oW1;
  offsets[1] = ly.ob1;
  offsets[2] = ly.oW2;
  offsets[3] = ly.ob2;
  return UT_OK;
}

int ut_flow_get_weights(const ut_flow* flow, float* out) {
  if (!flow || !out) return fail(UT_E_ARG, "ut_flow_get_weights: null pointer");
  UT_HIP(hipMemcpy(out, flow->f.params, (size_t)flow->f.nparams * 4, hipMemcpyDeviceToHost));
  return UT_OK;
}

int ut_flow_set_weights(ut_flow* flow, const float* in) {
  if (!flow || !in) return fail(UT_E_ARG, "ut_flow_set_weights: null pointer");
  Flow& f = flow->f;
  UT_HIP(hipMemcpy(f.params, in, (size_t)f.nparams * 4, hipMemcpyHostToDevice));
  UT_HIP(flow_sync_weights(f, f.stream));
  UT_HIP(hipStreamSynchronize(f.stream));
  return UT_OK;
}

int ut_flow_set_base(ut_flow* flow, const double* mu, const double* log_scale) {
  if (!flow || !mu || !log_scale) return fail(UT_E_ARG, "ut_flow_set_base: null pointer");
  Flow& f = flow->f;
  std::vector<float> v(2 * (size_t)f.P);
  for (int i = 0; i < f.P; ++i) {
    v[i] = (float)mu[i];
    v[f.P + i] = (float)log_scale[i];
  }
  UT_HIP(hipMemcpy(f.params, v.data(), v.size() * 4, hipMemcpyHostToDevice));
  return UT_OK;
}

int ut_flow_train(ut_flow* flow, const float* draws, int64_t n, int32_t steps, int32_t batch, double learning_rate,
                  uint64_t seed, double* loss_trace) {
  if (!flow || !draws) return fail(UT_E_ARG, "ut_flow_train: null pointer");
  if (flow->f.factor) return fail(UT_E_ARG, "ut_flow_train: the full-rank guide is fitted by ut_flow_fit_elbo");
  if (n <= 0 || steps < 0) return fail(UT_E_ARG, "ut_flow_train: n must be > 0 and steps >= 0");
  int rc = flow_rows_check(flow, batch, "ut_flow_train (batch)");
  if (rc) return rc;
  if (!(learning_rate > 0)) return fail(UT_E_ARG, "ut_flow_train: learning_rate must be > 0");
  Flow& f = flow->f;
  hipStream_t st = f.stream;
  UT_HIP(flow_scratch(f, batch));
  DeviceBuffers buf;
  float *d_draws = nullptr, *d_batch = nullptr, *d_z = nullptr;
  double* d_rows = nullptr;
  UT_HIP(buf.alloc(&d_draws, (size_t)n * f.P));
  UT_HIP(buf.alloc(&d_batch, (size_t)batch * f.P));
  UT_HIP(buf.alloc(&d_z, (size_t)batch * f.P));
  UT_HIP(buf.alloc(&d_rows, (size_t)batch));
  UT_HIP(hipMemcpyAsync(d_draws, draws, (size_t)n * f.P * 4, hipMemcpyHostToDevice, st));
  std::vector<double> rows(batch);
  // minibatch rows are Philox draws keyed by both halves of the 64-bit seed
  const uint32_t key = run_seed_of(seed);
  for (int s = 0; s < steps; ++s) {
    UT_HIP(flow_gather_rows(d_draws, n, f.P, d_batch, batch, key, (uint32_t)s, st));
    UT_HIP(flow_inverse(f, d_batch, d_z, nullptr, batch, st));
    if (loss_trace) UT_HIP(flow_loss_value(d_z, f.sc.logdet, batch, f.P, d_rows, st));
    UT_HIP(flow_loss_backward(f, d_z, batch, st));
    UT_HIP(flow_adam(f, batch, learning_rate, 0.9, 0.999, 1e-8, st));
    UT_HIP(flow_sync_weights(f, st));
    if (loss_trace) {
      UT_HIP(hipMemcpyAsync(rows.data(), d_rows, (size_t)batch * 8, hipMemcpyDeviceToHost, st));
      UT_HIP(hipStreamSynchronize(st));
      double acc = 0.0;
      for (int b = 0; b < batch; ++b) acc += rows[b];
      loss_trace[s] = acc / batch;
    }
  }
  UT_HIP(hipStreamSynchronize(st));
  return UT_OK;
}

int ut_flow_forward(ut_flow* flow, const float* z, int64_t n, float* q, float* logdet) {
  int rc = flow_rows_check(flow, n, "ut_flow_forward");
  if (rc) return rc;
  if (!z || !q) return fail(UT_E_ARG, "ut_flow_forward: null pointer");
  Flow& f = flow->f;
  UT_HIP(flow_scratch(f, (int)n));
  DeviceBuffers buf;
  float *dz = nullptr, *dq = nullptr, *dl = nullptr;
  UT_HIP(buf.alloc(&dz, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dq, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dl, (size_t)n));
  UT_HIP(hipMemcpyAsync(dz, z, (size_t)n * f.P * 4, hipMemcpyHostToDevice, f.stream));
  UT_HIP(flow_forward(f, dz, dq, dl, (int)n, f.stream));
  UT_HIP(hipMemcpyAsync(q, dq, (size_t)n * f.P * 4, hipMemcpyDeviceToHost, f.stream));
  if (logdet) UT_HIP(hipMemcpyAsync(logdet, dl, (size_t)n * 4, hipMemcpyDeviceToHost, f.stream));
  UT_HIP(hipStreamSynchronize(f.stream));
  return UT_OK;
}

int ut_flow_kernel_times(ut_flow* flow, int64_t n, int32_t reps, double* forward_us, double* backward_us) {
  int rc = flow_rows_check(flow, n, "ut_flow_kernel_times");
  if (rc) return rc;
  if (reps < 1 || !forward_us || !backward_us) return fail(UT_E_ARG, "ut_flow_kernel_times: reps >= 1 and outputs");
  Flow& f = flow->f;
  UT_HIP(flow_scratch(f, (int)n));
  DeviceBuffers buf;
  float *dz = nullptr, *dg = nullptr;
  UT_HIP(buf.alloc(&dz, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dg, (size_t)n * f.P));
  UT_HIP(flow_normal_fill(dz, (int)n, f.P, 0x7f10u, 0u, f.stream));
  UT_HIP(flow_normal_fill(dg, (int)n, f.P, 0x7f10u, 1u, f.stream));
  hipEvent_t ev[3];
  for (auto& e : ev) UT_HIP(hipEventCreate(&e));
  UT_HIP(flow_forward(f, dz, nullptr, nullptr, (int)n, f.stream));  // warm
  UT_HIP(flow_backward(f, dg, dg, (int)n, f.stream));
  UT_HIP(hipEventRecord(ev[0], f.stream));
  for (int r = 0; r < reps; ++r) UT_HIP(flow_forward(f, dz, nullptr, nullptr, (int)n, f.stream));
  UT_HIP(hipEventRecord(ev[1], f.stream));
  // the backward reads the last forward's activations and writes g_z over its input
  for (int r = 0; r < reps; ++r) UT_HIP(flow_backward(f, dg, dg, (int)n, f.stream));
  UT_HIP(hipEventRecord(ev[2], f.stream));
  UT_HIP(hipEventSynchronize(ev[2]));
  float ms_f = 0.0f, ms_b = 0.0f;
  UT_HIP(hipEventElapsedTime(&ms_f, ev[0], ev[1]));
  UT_HIP(hipEventElapsedTime(&ms_b, ev[1], ev[2]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *forward_us = 1e3 * ms_f / reps;
  *backward_us = 1e3 * ms_b / reps;
  return UT_OK;
}

int ut_flow_inverse(ut_flow* flow, const float* q, int64_t n, float* z, float* logdet_inv) {
  int rc = flow_rows_check(flow, n, "ut_flow_inverse");
  if (rc) return rc;
  if (!z || !q) return fail(UT_E_ARG, "ut_flow_inverse: null pointer");
  Flow& f = flow->f;
  UT_HIP(flow_scratch(f, (int)n));
  DeviceBuffers buf;
  float *dz = nullptr, *dq = nullptr, *dl = nullptr;
  UT_HIP(buf.alloc(&dz, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dq, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dl, (size_t)n));
  UT_HIP(hipMemcpyAsync(dq, q, (size_t)n * f.P * 4, hipMemcpyHostToDevice, f.stream));
  UT_HIP(flow_inverse(f, dq, dz, dl, (int)n, f.stream));
  UT_HIP(hipMemcpyAsync(z, dz, (size_t)n * f.P * 4, hipMemcpyDeviceToHost, f.stream));
  if (logdet_inv) UT_HIP(hipMemcpyAsync(logdet_inv, dl, (size_t)n * 4, hipMemcpyDeviceToHost, f.stream));
  UT_HIP(hipStreamSynchronize(f.stream));
  return UT_OK;
}

int ut_flow_vjp(ut_flow* flow, const float* z, int64_t n, const float* g_q, float* g_z) {
  int rc = flow_rows_check(flow, n, "ut_flow_vjp");
  if (rc) return rc;
  if (!z || !g_q || !g_z) return fail(UT_E_ARG, "ut_flow_vjp: null pointer");
  Flow& f = flow->f;
  UT_HIP(flow_scratch(f, (int)n));
  DeviceBuffers buf;
  float *dz = nullptr, *dg = nullptr;
  UT_HIP(buf.alloc(&dz, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dg, (size_t)n * f.P));
  UT_HIP(hipMemcpyAsync(dz, z, (size_t)n * f.P * 4, hipMemcpyHostToDevice, f.stream));
  UT_HIP(hipMemcpyAsync(dg, g_q, (size_t)n * f.P * 4, hipMemcpyHostToDevice, f.stream));
  UT_HIP(flow_forward(f, dz, nullptr, nullptr, (int)n, f.stream));
  UT_HIP(flow_backward(f, dg, dg, (int)n, f.stream));
  UT_HIP(hipMemcpyAsync(g_z, dg, (size_t)n * f.P * 4, hipMemcpyDeviceToHost, f.stream));
  UT_HIP(hipStreamSynchronize(f.stream));
  return UT_OK;
}

int ut_flow_loss_grad(ut_flow* flow, const float* q, int64_t n, double* loss, float* grad) {
  int rc = flow_rows_check(flow, n, "ut_flow_loss_grad");
  if (rc) return rc;
  if (!q || !grad) return fail(UT_E_ARG, "ut_flow_loss_grad: null pointer");
  if (flow->f.factor) return fail(UT_E_ARG, "ut_flow_loss_grad: not defined for the full-rank guide");
  Flow& f = flow->f;
  UT_HIP(flow_scratch(f, (int)n));
  DeviceBuffers buf;
  float *dq = nullptr, *dz = nullptr;
  double* drows = nullptr;
  UT_HIP(buf.alloc(&dq, (size_t)n * f.P));
  UT_HIP(buf.alloc(&dz, (size_t)n * f.P));
  UT_HIP(buf.alloc(&drows, (size_t)n));
  UT_HIP(hipMemcpyAsync(dq, q, (size_t)n * f.P * 4, hipMemcpyHostToDevice, f.stream));
  UT_HIP(flow_inverse(f, dq, dz, nullptr, (int)n, f.stream));
  UT_HIP(flow_loss_value(dz, f.sc.logdet, (int)n, f.P, drows, f.stream));
  UT_HIP(flow_loss_backward(f, dz, (int)n, f.stream));
  std::vector<double> rows((size_t)n);
  std::vector<float> g((size_t)f.nparams);
  UT_HIP(hipMemcpyAsync(rows.data(), drows, (size_t)n * 8, hipMemcpyDeviceToHost, f.stream));
  UT_HIP(hipMemcpyAsync(g.data(), f.grads, g.size() * 4, hipMemcpyDeviceToHost, f.stream));
  UT_HIP(hipStreamSynchronize(f.stream));
  double acc = 0.0;
  for (double r : rows) acc += r;
  if (loss) *loss = acc / (double)n;
  for (size_t i = 0; i < g.size(); ++i) grad[i] = g[i] / (float)n;
  return UT_OK;
}

int ut_workspace_set_chain_waves(ut_workspace* ws, int32_t waves) {
  if (!ws) return fail(UT_E_ARG, "ut_workspace_set_chain_waves: null workspace");
  if (waves != 0 && waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return fail(UT_E_ARG, "ut_workspace_set_chain_waves: waves must be 0 (automatic), 1, 2, 4 or 8");
  if (waves > 1 && ws->plan->dev.kind != MODEL_LOGISTIC)
    return fail(UT_E_UNSUPPORTED, "ut_workspace_set_chain_waves: only the logistic plan splits a chain over wavefronts");
  const int nw = chain_waves_for(ws->plan->dev, ws->C, waves, kDefaultTreeDepth);
  if (nw < 0 && chain_waves_for(ws->plan->dev, ws->C, waves, -1) < 0)
    return fail(UT_E_ARG, "ut_workspace_set_chain_waves: the exchange area exceeds the 160 KB of LDS");
  ws->waves_request = waves;
  ws->dev.waves = nw > 0 ? nw : waves;
  return UT_OK;
}

int ut_workspace_chain_waves(const ut_workspace* ws, int32_t* waves) {
  if (!ws || !waves) return fail(UT_E_ARG, "ut_workspace_chain_waves: null pointer");
  *waves = ws->dev.waves;
  return UT_OK;
}

int ut_workspace_set_flow(ut_workspace* ws, ut_flow* flow) {
  if (!ws) return fail(UT_E_ARG, "ut_workspace_set_flow: null workspace");
  if (flow) {
    if (flow->f.P != ws->P) return fail(UT_E_DIM, "ut_workspace_set_flow: flow and workspace parameter counts differ");
    if (ws->C % FLOW_PAD != 0 || ws->precision != UT_FLOAT32)
      return fail(UT_E_DIM, "ut_workspace_set_flow: the workspace needs UT_FLOAT32 and num_chains a multiple of 128");
    UT_HIP(flow_scratch(flow->f, ws->C));
  }
  ws->clear_graphs();
  ws->flow = flow;
  return UT_OK;
}

}  // extern "C"

namespace {

int elbo_check(ut_workspace* ws, ut_flow* flow, const char* who) {
  if (!ws || !flow) return fail(UT_E_ARG, std::string(who) + ": null workspace or flow");
  if (flow->f.P != ws->P) return fail(UT_E_DIM, std::string(who) + ": flow and workspace parameter counts differ");
  if (ws->C % FLOW_PAD != 0 || ws->precision != UT_FLOAT32)
    return fail(UT_E_DIM, std::string(who) + ": the workspace needs UT_FLOAT32 and num_chains a multiple of 128");
  return UT_OK;
}

// Device buffers of one ELBO evaluation over the workspace's C particles.
struct ElboBuffers {
  DeviceBuffers buf;
  float *z = nullptr, *lj = nullptr, *rw = nullptr, *wts = nullptr;
  double *logw = nullptr, *part = nullptr, *stats = nullptr, *gstat = nullptr;
  hipError_t alloc(int B, int P) {
    hipError_t e = buf.alloc(&z, (size_t)B * P);
    if (e == hipSuccess) e = buf.alloc(&lj, (size_t)B);
    if (e == hipSuccess) e = buf.alloc(&rw, (size_t)B);
    if (e == hipSuccess) e = buf.alloc(&wts, (size_t)B);
    if (e == hipSuccess) e = buf.alloc(&logw, (size_t)B);
    if (e == hipSuccess) e = buf.alloc(&part, 256);
    if (e == hipSuccess) e = buf.alloc(&stats, 5);
    if (e == hipSuccess) e = buf.alloc(&gstat, 2 * (size_t)B);
    return e;
  }
};

// particles z (in b.z) -> flow -> model logjoint and gradient -> per-particle
// validity and log-weights -> (importance weights of groups of K > 1) -> summed
// weight gradient (f.grads) -> b.stats
hipError_t elbo_eval(ut_workspace* ws, Flow& f, ElboBuffers& b, double clip, int K) {
  const int B = ws->C;
  hipStream_t st = ws->stream;
  hipError_t e = flow_forward_train(f, b.z, B, st);
  if (e == hipSuccess) e = model_grad(ws, f.sc.x, b.lj, f.sc.g, true);
  if (e == hipSuccess) e = flow_elbo_rows(f, b.z, b.lj, B, b.rw, b.logw, st);
  if (K > 1) {
    if (e == hipSuccess) e = flow_iwae_weights(f, b.rw, b.logw, B, K, b.wts, b.gstat, st);
    if (e == hipSuccess) e = flow_elbo_backward(f, b.z, b.wts, B, st);
    if (e == hipSuccess) e = flow_elbo_stats(f, b.rw, b.logw, B, b.gstat, B / K, clip, b.part, b.stats, st);
  } else {
    if (e == hipSuccess) e = flow_elbo_backward(f, b.z, b.rw, B, st);
    if (e == hipSuccess) e = flow_elbo_stats(f, b.rw, b.logw, B, nullptr, 0, clip, b.part, b.stats, st);
  }
  return e;
}

int iwae_check(int B, int32_t K, const char* who) {
  if (K < 1 || B % K != 0)
    return fail(UT_E_ARG, std::string(who) + ": iwae_group must be >= 1 and divide num_chains (1: the standard ELBO)");
  return UT_OK;
}

}  // namespace

extern "C" {

int ut_flow_elbo_grad(ut_workspace* ws, ut_flow* flow, const float* z, int32_t iwae_group, double* elbo, float* grad,
                      float* g_q, int32_t* valid, float* row_weights) {
  int rc = elbo_check(ws, flow, "ut_flow_elbo_grad");
  if (rc) return rc;
  if (!z || !grad) return fail(UT_E_ARG, "ut_flow_elbo_grad: null pointer");
  rc = iwae_check(ws->C, iwae_group, "ut_flow_elbo_grad");
  if (rc) return rc;
  Flow& f = flow->f;
  const int B = ws->C, P = ws->P;
  hipStream_t st = ws->stream;
  UT_HIP(hipStreamSynchronize(f.stream));
  UT_HIP(flow_scratch(f, B));
  ElboBuffers b;
  UT_HIP(b.alloc(B, P));
  UT_HIP(hipMemcpyAsync(b.z, z, (size_t)B * P * 4, hipMemcpyHostToDevice, st));
  UT_HIP(elbo_eval(ws, f, b, std::numeric_limits<double>::infinity(), iwae_group));
  double h[5];
  std::vector<float> rw((size_t)B), wt((size_t)B);
  if (iwae_group > 1) UT_HIP(hipMemcpyAsync(wt.data(), b.wts, (size_t)B * 4, hipMemcpyDeviceToHost, st));
  if (g_q) {  // the model gradient at f(z) (zero on invalid rows) before the backward pass consumed it
    UT_HIP(hipMemcpyAsync(b.z, z, (size_t)B * P * 4, hipMemcpyHostToDevice, st));
    UT_HIP(flow_forward_train(f, b.z, B, st));
    UT_HIP(model_grad(ws, f.sc.x, b.lj, f.sc.g, true));
    UT_HIP(flow_elbo_rows(f, b.z, b.lj, B, b.rw, b.logw, st));
    UT_HIP(hipMemcpyAsync(g_q, f.sc.g, (size_t)B * P * 4, hipMemcpyDeviceToHost, st));
  }
  UT_HIP(hipMemcpyAsync(h, b.stats, sizeof(h), hipMemcpyDeviceToHost, st));
  UT_HIP(hipMemcpyAsync(rw.data(), b.rw, (size_t)B * 4, hipMemcpyDeviceToHost, st));
  UT_HIP(hipMemcpyAsync(grad, f.grads, (size_t)f.nparams * 4, hipMemcpyDeviceToHost, st));
  UT_HIP(hipStreamSynchronize(st));
  if (valid)
    for (int i = 0; i < B; ++i) valid[i] = rw[i] != 0.0f;
  if (elbo) *elbo = h[1];
  const float sc = h[0] > 0 ? (float)(1.0 / h[0]) : 0.0f;
  for (long long i = 0; i < f.nparams; ++i) grad[i] *= sc;
  if (row_weights)
    for (int i = 0; i < B; ++i) row_weights[i] = iwae_group > 1 ? wt[i] * sc : rw[i] * sc;
  return UT_OK;
}

}  // extern "C"

namespace {

// Pointwise log-likelihood of the observed sites (the reference's
// pointwise_loglikelihood, src/evaluator_pointwise.jl:342, result_interface.jl:384-410):
// one thread per (draw, observation), binary64, draws in unconstrained space.
// y_i ~ bernoullilogit(x_i . beta): y eta - log(1 + e^eta)
__global__ __launch_bounds__(256) void k_ll_logistic(const double* __restrict__ draws, long long S, int D,
                                                     const double* __restrict__ xT, const double* __restrict__ yv,
                                                     long long n, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= S * n) return;
  const long long s = t / n, i = t - s * n;
  const double* b = draws + s * D;
  double eta = 0.0;
  for (int d = 0; d < D; ++d) eta += xT[(long long)d * n + i] * b[d];
  const double softplus = eta > 0.0 ? eta + log1p(exp(-eta)) : log1p(exp(eta));
  out[t] = yv[i] * eta - softplus;
}

// the same for the bf16 plan: X is the plan's bf16 matrix (row-major, X[i D + d]),
// widened exactly, as the matrix-core gradient evaluates it; 16-byte row chunks
__global__ __launch_bounds__(256) void k_ll_glm_bf16(const double* __restrict__ draws, long long S, int D,
                                                     const uint16_t* __restrict__ X, const float* __restrict__ yv,
                                                     long long n, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= S * n) return;
  const long long s = t / n, i = t - s * n;
  const double* b = draws + s * D;
  const uint4* xr = (const uint4*)(X + i * D);
  double eta = 0.0;
  for (int d8 = 0; d8 < D / 8; ++d8) {
    const uint4 v = xr[d8];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      eta += (double)__uint_as_float(w[k] << 16) * b[8 * d8 + 2 * k];
      eta += (double)__uint_as_float(w[k] & 0xffff0000u) * b[8 * d8 + 2 * k + 1];
    }
  }
  const double softplus = eta > 0.0 ? eta + log1p(exp(-eta)) : log1p(exp(eta));
  out[t] = (double)yv[i] * eta - softplus;
}

// latent GP: y_i ~ poisson(e^{f_i}) with f = L z summed from the factorisation's
// per-tile-column partial products (value component, tile columns kb <= i's tile, in
// the order the gradient kernel sums them)
__global__ __launch_bounds__(256) void k_ll_gp(const double* __restrict__ pf, int nb, int T, int N,
                                               const double* __restrict__ y, long long S, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= S * N) return;
  const long long s = t / N;
  const int i = (int)(t - s * N), ib = i / GP_TILE;
  double f = 0.0;
  for (int kb = 0; kb <= ib; ++kb) f += pf[((size_t)s * nb + kb) * (T + 1) * N + i];
  out[t] = y[i] * f - exp(f) - lgamma(y[i] + 1.0);
}

// y ~ normal(mu + e^{log tau} theta_raw_j, sigma_j) for observation k = (group grp[k],
// row idx[k] of the plan's lane-padded y); inv_var[j] = 1 / sigma_j^2
__global__ __launch_bounds__(256) void k_ll_hier(const double* __restrict__ draws, long long S, int J,
                                                 const double* __restrict__ ypad, const double* __restrict__ inv_var,
                                                 const int* __restrict__ grp, const int* __restrict__ idx,
                                                 long long n, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= S * n) return;
  const long long s = t / n, k = t - s * n;
  const int j = grp[k];
  const double* th = draws + s * (J + 2);
  const double theta = th[0] + exp(th[1]) * th[2 + j];
  const double r = ypad[(long long)idx[k] * WAVE + j] - theta;
  const double iv = inv_var[j];
  out[t] = -HALF_LOG_2PI + 0.5 * log(iv) - 0.5 * r * r * iv;
}

// observation count of a plan: logistic n, hierarchical sum n_j, funnel 0
int ll_count(const ut_plan* plan, std::vector<int>* nobs_out, int64_t* n) {
  const DevModelDesc& m = plan->dev;
  switch (m.kind) {
    case MODEL_LOGISTIC: *n = m.n; return UT_OK;
    case MODEL_GLM_MFMA: *n = plan->glm.N; return UT_OK;
    case MODEL_GP: *n = plan->gp.N; return UT_OK;
    case MODEL_FUNNEL: *n = 0; return UT_OK;
    case MODEL_HIER: {
      std::vector<int> nobs(WAVE, 0);
      UT_HIP(hipMemcpy(nobs.data(), m.nobs, WAVE * sizeof(int), hipMemcpyDeviceToHost));
      int64_t total = 0;
      for (int j = 0; j < m.J; ++j) total += nobs[j];
      if (nobs_out) *nobs_out = nobs;
      *n = total;
      return UT_OK;
    }
    default:
      return fail(UT_E_UNSUPPORTED, "ut_pointwise_loglik: the pointwise log-likelihood is built for the "
                                    "logistic (float and bf16), hierarchical, funnel and GP plans");
  }
}

}  // namespace

extern "C" {

int ut_pointwise_loglik_size(const ut_plan* plan, int64_t* num_observations) {
  if (!plan || !num_observations) return fail(UT_E_ARG, "ut_pointwise_loglik_size: null pointer");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_pointwise_loglik_size: no HIP device visible");
  return ll_count(plan, nullptr, num_observations);
}

int ut_pointwise_loglik(const ut_plan* plan, const double* draws, int64_t num_draws, double* out) {
  if (!plan || (num_draws > 0 && (!draws || !out))) return fail(UT_E_ARG, "ut_pointwise_loglik: null pointer");
  if (num_draws < 0) return fail(UT_E_DIM, "ut_pointwise_loglik: num_draws must be >= 0");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_pointwise_loglik: no HIP device visible");
  const DevModelDesc& m = plan->dev;
  std::vector<int> nobs;
  int64_t n = 0;
  int rc = ll_count(plan, &nobs, &n);
  if (rc) return rc;
  if (n == 0 || num_draws == 0) return UT_OK;
  const int P = m.P;
  DeviceBuffers buf;
  if (m.kind == MODEL_GP) {  // factorise K per draw (binary32 positions, as the model's evaluator takes them)
    const GpPlan& g = plan->gp;
    const int T = g.D + 1, nb = g.N / GP_TILE;
    const size_t per_draw = ((size_t)(T + 1) * gp_tiles(g.N) * GP_TILE * GP_TILE + (size_t)nb * (T + 1) * g.N) * 8;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>({num_draws, 128, (int64_t)((size_t(1) << 30) / per_draw)}));
    GpScratch gs;
    UT_HIP(gp_scratch_alloc(gs, g, (int)chunk));
    struct Release {
      GpScratch& s;
      ~Release() { gp_scratch_free(s); }
    } release{gs};
    float* dq = nullptr;
    double* dout = nullptr;
    UT_HIP(buf.alloc(&dq, (size_t)chunk * P));
    UT_HIP(buf.alloc(&dout, (size_t)chunk * n));
    std::vector<float> hq((size_t)chunk * P);
    for (int64_t s0 = 0; s0 < num_draws; s0 += chunk) {
      const int64_t S = std::min(chunk, num_draws - s0);
      for (size_t k = 0; k < (size_t)S * P; ++k) hq[k] = (float)draws[(size_t)s0 * P + k];
      UT_HIP(hipMemcpy(dq, hq.data(), (size_t)S * P * 4, hipMemcpyHostToDevice));
      UT_HIP(launch_gp_factor(g, dq, P, nullptr, (int)S, gs, 0));
      hipLaunchKernelGGL(k_ll_gp, dim3((unsigned)((S * n + 255) / 256)), dim3(256), 0, 0, gs.pf, nb, T, g.N, g.y,
                         (long long)S, dout);
      UT_HIP(hipGetLastError());
      UT_HIP(hipMemcpy(out + s0 * n, dout, (size_t)S * n * 8, hipMemcpyDeviceToHost));
    }
    return UT_OK;
  }
  int *grp = nullptr, *idx = nullptr;
  if (m.kind == MODEL_HIER) {  // observation order: group by group, rows in data order
    std::vector<int> hg((size_t)n), hi((size_t)n);
    size_t k = 0;
    for (int j = 0; j < m.J; ++j)
      for (int i = 0; i < nobs[j]; ++i, ++k) {
        hg[k] = j;
        hi[k] = i;
      }
    UT_HIP(buf.alloc(&grp, (size_t)n));
    UT_HIP(buf.alloc(&idx, (size_t)n));
    UT_HIP(hipMemcpy(grp, hg.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    UT_HIP(hipMemcpy(idx, hi.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  }
  // draws in chunks so the device output stays within 512 MiB
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(num_draws, (int64_t(1) << 26) / n));
  double *dd = nullptr, *dout = nullptr;
  UT_HIP(buf.alloc(&dd, (size_t)chunk * P));
  UT_HIP(buf.alloc(&dout, (size_t)chunk * n));
  for (int64_t s0 = 0; s0 < num_draws; s0 += chunk) {
    const int64_t S = std::min(chunk, num_draws - s0);
    UT_HIP(hipMemcpy(dd, draws + s0 * P, (size_t)S * P * 8, hipMemcpyHostToDevice));
    const unsigned blocks = (unsigned)((S * n + 255) / 256);
    if (m.kind == MODEL_GLM_MFMA)
      hipLaunchKernelGGL(k_ll_glm_bf16, dim3(blocks), dim3(256), 0, 0, dd, (long long)S, plan->glm.D, plan->glm.X,
                         plan->glm.y, (long long)n, dout);
    else if (m.kind == MODEL_LOGISTIC)
      hipLaunchKernelGGL(k_ll_logistic, dim3(blocks), dim3(256), 0, 0, dd, (long long)S, m.D, m.xT_d, m.yv_d,
                         (long long)n, dout);
    else
      hipLaunchKernelGGL(k_ll_hier, dim3(blocks), dim3(256), 0, 0, dd, (long long)S, m.J, m.y_d, m.lane_d + 2 * WAVE,
                         grp, idx, (long long)n, dout);
    UT_HIP(hipGetLastError());
    UT_HIP(hipMemcpy(out + s0 * n, dout, (size_t)S * n * 8, hipMemcpyDeviceToHost));
  }
  return UT_OK;
}

int ut_initial_params(const ut_plan* plan, int32_t strategy, uint64_t seed, int32_t num_chains, double* out) {
  if (!plan || !out) return fail(UT_E_ARG, "ut_initial_params: null pointer");
  if (num_chains <= 0) return fail(UT_E_DIM, "ut_initial_params: num_chains must be positive");
  if (strategy != UT_INIT_PRIOR && strategy != UT_INIT_UNIFORM)
    return fail(UT_E_ARG, "ut_initial_params: strategy must be UT_INIT_PRIOR or UT_INIT_UNIFORM");
  const int P = plan->dev.P;
  for (int c = 0; c < num_chains; ++c)
    draw_initial_column(plan->dev, strategy, (uint32_t)c, run_seed_of(seed), out + (size_t)c * P);
  return UT_OK;
}

int ut_flow_fit_elbo(ut_workspace* ws, ut_flow* flow, int32_t num_iterations, double learning_rate, double beta1,
                     double beta2, double adam_epsilon, double gradient_clip, int32_t iwae_group, uint64_t seed,
                     double* elbo_history, double* standard_elbo_history, double* grad_norm_history,
                     float* best_weights, double* best_elbo) {
  int rc = elbo_check(ws, flow, "ut_flow_fit_elbo");
  if (rc) return rc;
  rc = iwae_check(ws->C, iwae_group, "ut_flow_fit_elbo");
  if (rc) return rc;
  if (num_iterations < 0) return fail(UT_E_ARG, "ut_flow_fit_elbo: num_iterations must be >= 0");
  if (!(learning_rate > 0)) return fail(UT_E_ARG, "ut_flow_fit_elbo: learning_rate must be > 0");
  if (!(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1))
    return fail(UT_E_ARG, "ut_flow_fit_elbo: requires 0 <= beta1 < 1 and 0 <= beta2 < 1");
  if (!(adam_epsilon > 0)) return fail(UT_E_ARG, "ut_flow_fit_elbo: adam_epsilon must be > 0");
  if (!(gradient_clip > 0)) return fail(UT_E_ARG, "ut_flow_fit_elbo: gradient_clip must be > 0 (inf: no clipping)");
  Flow& f = flow->f;
  const int B = ws->C, P = ws->P;
  hipStream_t st = ws->stream;
  UT_HIP(hipStreamSynchronize(f.stream));  // the flow's own calls run on its stream
  UT_HIP(flow_scratch(f, B));
  ElboBuffers b;
  UT_HIP(b.alloc(B, P));
  DeviceBuffers buf;
  float* best = nullptr;
  UT_HIP(buf.alloc(&best, (size_t)f.nparams));
  UT_HIP(hipMemcpyAsync(best, f.params, (size_t)f.nparams * 4, hipMemcpyDeviceToDevice, st));
  // fresh Adam moments per fit, as each reference ADVI run starts its own
  UT_HIP(hipMemsetAsync(f.adam_m, 0, (size_t)f.nparams * 4, st));
  UT_HIP(hipMemsetAsync(f.adam_v, 0, (size_t)f.nparams * 4, st));
  f.adam_t = 0;
  // The bookkeeping (histories, best weights, the stop at an iteration without a valid
  // particle) runs on the device, so the iterations queue without a host round trip;
  // after a failed iteration the weights and the best copy stay as they were.
  double* hist = nullptr;
  ElboTrack* track = nullptr;
  UT_HIP(buf.alloc(&hist, 3 * (size_t)std::max(num_iterations, 1)));
  UT_HIP(buf.alloc(&track, 1));
  ElboTrack t0{-std::numeric_limits<double>::infinity(), 0, -1, 0, 0};
  UT_HIP(hipMemcpyAsync(track, &t0, sizeof(t0), hipMemcpyHostToDevice, st));
  // particles are Philox normals keyed by both halves of the 64-bit seed
  const uint32_t key = run_seed_of(seed);
  for (int it = 0; it < num_iterations; ++it) {
    UT_HIP(flow_normal_fill(b.z, B, P, key, (uint32_t)it, st));
    UT_HIP(elbo_eval(ws, f, b, gradient_clip, iwae_group));
    UT_HIP(flow_elbo_track(f, b.stats, it, hist, track, best, st));
    UT_HIP(flow_adam_ascent(f, b.stats, learning_rate, beta1, beta2, adam_epsilon, st, track));
    UT_HIP(flow_sync_weights(f, st));
  }
  ElboTrack tr;
  std::vector<double> h(3 * (size_t)std::max(num_iterations, 1));
  UT_HIP(hipMemcpyAsync(&tr, track, sizeof(tr), hipMemcpyDeviceToHost, st));
  if (num_iterations > 0) UT_HIP(hipMemcpyAsync(h.data(), hist, h.size() * 8, hipMemcpyDeviceToHost, st));
  if (best_weights) UT_HIP(hipMemcpyAsync(best_weights, best, (size_t)f.nparams * 4, hipMemcpyDeviceToHost, st));
  UT_HIP(hipStreamSynchronize(st));
  if (tr.failed)
    return fail(UT_E_ARG, std::string("ut_flow_fit_elbo: only non-finite ") +
                              (iwae_group > 1 ? "importance weights" : "logjoint values or gradients") +
                              " at iteration " + std::to_string(tr.fail_iter));
  for (int it = 0; it < num_iterations; ++it) {
    if (elbo_history) elbo_history[it] = h[3 * (size_t)it];
    if (standard_elbo_history) standard_elbo_history[it] = h[3 * (size_t)it + 1];
    const double gn = h[3 * (size_t)it + 2];
    if (grad_norm_history) grad_norm_history[it] = std::isfinite(gradient_clip) ? std::min(gn, gradient_clip) : gn;
  }
  if (best_elbo) *best_elbo = tr.best;
  return UT_OK;
}

}  // extern "C"

extern "C" {

int ut_gp_factor(const ut_plan* plan, const double* hyper, int32_t num_chains, double* L, double* dL,
                 double* logdet) {
  if (!plan || !hyper) return fail(UT_E_ARG, "ut_gp_factor: null pointer");
  if (plan->dev.kind != MODEL_GP) return fail(UT_E_ARG, "ut_gp_factor: not a GP plan");
  if (num_chains <= 0) return fail(UT_E_DIM, "ut_gp_factor: num_chains must be positive");
  if (!ut_device_available()) return fail(UT_E_HIP, "ut_gp_factor: no HIP device visible");
  const GpPlan& g = plan->gp;
  const int T = g.D + 1;
  const size_t nn = (size_t)g.N * g.N;
  const int nb = g.N / GP_TILE;
  DeviceBuffers buf;
  GpScratch s;
  double* dh = nullptr;
  UT_HIP(gp_scratch_alloc(s, g, num_chains));
  struct Release {
    GpScratch& s;
    ~Release() { gp_scratch_free(s); }
  } release{s};
  UT_HIP(buf.alloc(&dh, (size_t)num_chains * T));
  UT_HIP(hipMemcpy(dh, hyper, (size_t)num_chains * T * 8, hipMemcpyHostToDevice));
  UT_HIP(launch_gp_factor(g, nullptr, T, dh, num_chains, s, nullptr));
  std::vector<double> A((size_t)num_chains * (T + 1) * gp_tiles(g.N) * GP_TILE * GP_TILE);
  UT_HIP(hipMemcpy(A.data(), s.A, A.size() * 8, hipMemcpyDeviceToHost));
  // tile-major, column-major tiles -> row-major lower triangles
  for (int c = 0; c < num_chains; ++c) {
    double ld = 0.0;
    for (int ib = 0; ib < nb; ++ib) {
      const double* dt = &A[gp_tile_offset(g.N, T, c, 0, ib, ib)];
      for (int r = 0; r < GP_TILE; ++r) ld += std::log(dt[r * GP_TILE + r]);
    }
    if (logdet) logdet[c] = 2.0 * ld;
    for (int comp = 0; comp <= T; ++comp) {
      double* dst = comp == 0 ? (L ? L + (size_t)c * nn : nullptr) : (dL ? dL + ((size_t)c * T + comp - 1) * nn : nullptr);
      if (!dst) continue;
      std::fill(dst, dst + nn, 0.0);
      for (int ib = 0; ib < nb; ++ib)
        for (int jb = 0; jb <= ib; ++jb) {
          const double* tp = &A[gp_tile_offset(g.N, T, c, comp, ib, jb)];
          for (int col = 0; col < GP_TILE; ++col)
            for (int row = 0; row < GP_TILE; ++row)
              dst[(size_t)(ib * GP_TILE + row) * g.N + jb * GP_TILE + col] = tp[col * GP_TILE + row];
        }
    }
  }
  return UT_OK;
}

}  // extern "C"

extern "C" {

int ut_gradient_step_kernel_time(ut_workspace* ws, const void* params, int32_t reps, double* us_per_call) {
  if (!ws || !params || !us_per_call) return fail(UT_E_ARG, "ut_gradient_step_kernel_time: null pointer");
  if (reps < 1) return fail(UT_E_ARG, "ut_gradient_step_kernel_time: reps must be >= 1");
  if (ws->precision != UT_FLOAT32)
    return fail(UT_E_PRECISION, "ut_gradient_step_kernel_time: float32 workspaces only");
  const int P = ws->P, C = ws->C;
  const size_t PC = (size_t)P * C;
  DeviceBuffers buf;
  LockstepState ls;
  float* im = nullptr;
  UT_HIP(buf.alloc(&ls.wq, PC));
  UT_HIP(buf.alloc(&ls.wp, PC));
  UT_HIP(buf.alloc(&ls.wg, PC));
  UT_HIP(buf.alloc(&ls.wlj, (size_t)C));
  UT_HIP(buf.alloc(&ls.valid, (size_t)C));
  UT_HIP(buf.alloc(&im, (size_t)P));
  ls.im_f = im;
  const std::vector<float> ones((size_t)P, 1.0f);
  UT_HIP(hipMemcpyAsync(ls.wq, params, PC * 4, hipMemcpyHostToDevice, ws->stream));
  UT_HIP(hipMemcpyAsync(im, ones.data(), (size_t)P * 4, hipMemcpyHostToDevice, ws->stream));
  UT_HIP(hipMemsetAsync(ls.wp, 0, PC * 4, ws->stream));
  UT_HIP(hipMemsetAsync(ls.valid, 1, (size_t)C, ws->stream));
  hipEvent_t ev[2];
  for (auto& e : ev) UT_HIP(hipEventCreate(&e));
  // step size 0: the positions stay, so every call does the same work
  UT_HIP(batched_grad_step(ws, ls, P, C, 0.0f));  // warm (and capture, when graphs are used)
  UT_HIP(hipEventRecord(ev[0], ws->stream));
  for (int r = 0; r < reps; ++r) UT_HIP(batched_grad_step(ws, ls, P, C, 0.0f));
  UT_HIP(hipEventRecord(ev[1], ws->stream));
  UT_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.0f;
  UT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  UT_HIP(hipStreamSynchronize(ws->stream));
  *us_per_call = 1e3 * ms / reps;
  return UT_OK;
}

}  // extern "C"
