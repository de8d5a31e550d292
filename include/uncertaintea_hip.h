/* uncertaintea_hip.h -- C ABI of libuncertaintea_hip, the MI355X (gfx950) device
 * backend of UncertainTea's batched density and batched MCMC path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an
 * int status (UT_OK on success).  On failure ut_last_error() returns a message for
 * the calling thread.  Matrices follow the reference's column-major Julia layout:
 * a "P x C" parameter matrix stores chain c's P unconstrained parameters
 * contiguously at offset c*P.  All calls are synchronous: they return after the
 * device work they launched has finished and results are in the caller's buffers.
 *
 * Reference interfaces replaced (paths are relative to the reference repository):
 *   plans       <- DeviceExecutionPlan lowering + observation staging
 *                  (src/device/plan/steps.jl:532-539, src/device/plan/lowering.jl:1360-1478,
 *                   src/device/staging.jl:499-550)
 *   workspace   <- DeviceBatchedWorkspace (src/device/workspace.jl:71-162)
 *   density     <- device_batched_logjoint[!] / device_batched_logjoint_gradient[!]
 *                  (src/device/api.jl:34-124)
 *   samplers    <- batched_hmc / batched_nuts(tree_strategy=:persistent) device branches
 *                  (src/inference/api_batched_hmc.jl:16-75, src/device/hmc_kernels.jl:306-723;
 *                   src/inference/api_batched_nuts.jl:22-144,761-980,
 *                   src/device/persistent_nuts.jl:226-786); batched_chees
 *                  (src/inference/api_batched_chees.jl:277-700, device variant
 *                   src/device/chees_kernels.jl)
 *   RNG         <- philox4x32 / device_rand_* (src/device/rng.jl:175-343)
 */
#ifndef UNCERTAINTEA_HIP_H
#define UNCERTAINTEA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------
 * UT_E_UNSUPPORTED   model/feature not lowerable (reference: ArgumentError naming
 *                    device_lowering_report, src/device/workspace.jl:24-35)
 * UT_E_DIM           wrong parameter count or chain count (reference:
 *                    _signature_length_error / DimensionMismatch, src/device/api.jl:99-109)
 * UT_E_PRECISION     unsupported precision (reference: src/device/workspace.jl:12-22)
 * UT_E_HIP           a HIP runtime call failed (message carries hipGetErrorString)
 * UT_E_NONFINITE_INIT initial positions non-finite after init_max_retries re-draws
 *                    (reference: src/inference/api_batched_nuts.jl:184-200)
 * UT_E_ARG           invalid argument value (reference: ArgumentError from the
 *                    _validate_batched_* helpers)                                   */
enum {
  UT_OK = 0,
  UT_E_UNSUPPORTED = 1,
  UT_E_DIM = 2,
  UT_E_PRECISION = 3,
  UT_E_HIP = 4,
  UT_E_NONFINITE_INIT = 5,
  UT_E_ARG = 6
};

/* precision codes (reference: the `precision::Type` kwarg) */
enum { UT_FLOAT32 = 0, UT_FLOAT64 = 1 };

/* init strategies (reference: init_strategy = :prior | :uniform,
 * src/inference/hmc_support.jl:266-290) */
enum { UT_INIT_PRIOR = 0, UT_INIT_UNIFORM = 1 };

typedef struct ut_plan ut_plan;
typedef struct ut_workspace ut_workspace;
typedef struct ut_flow ut_flow;

/* Thread-local message for the last failing call on this thread. */
const char* ut_last_error(void);
/* ABI version: bumps on any incompatible change (2: progress callbacks in the
 * sampler option structs). */
int32_t ut_abi_version(void);
/* 1 when a HIP device is visible, 0 otherwise (no error). */
int32_t ut_device_available(void);

/* ---- plans: static model + staged observations ---------------------------------
 * A plan is the device analogue of the reference's lowered DeviceExecutionPlan
 * plus its staged observation stream: the model structure is fixed at creation
 * and the data are copied to HBM once, immutable for the plan's lifetime.       */

/* Bayesian logistic regression without intercept:
 *   beta_d ~ normal(prior_loc, prior_scale), d = 1..D
 *   y_i ~ bernoullilogit(x_i . beta), i = 1..N
 * x is N x D row-major; y holds 0/1.  P = D.
 * Reference: GLM step lowering src/device/plan/lowering.jl:664-675, analytic GLM
 * gradient src/device/gradient_kernel.jl:578-657, bernoullilogit
 * src/device/math.jl:202-219.                                                   */
int ut_plan_logistic_glm(const float* x, const float* y, int64_t n, int32_t d,
                         double prior_loc, double prior_scale, ut_plan** out);

/* Non-centred hierarchical normal (eight-schools form):
 *   mu ~ normal(mu_loc, mu_scale); tau ~ halfcauchy(tau_scale) (log transform);
 *   theta_raw_j ~ normal(0, 1); theta_j = mu + tau * theta_raw_j;
 *   y_ij ~ normal(theta_j, sigma_j).
 * Observations of group j are y[group_offsets[j] .. group_offsets[j+1]).
 * Parameter order: (mu, log tau, theta_raw_1..J); P = J + 2.  J <= 64.
 * Reference families: normal src/device/math.jl:88-91, halfcauchy :126-130,
 * log transform :1144-1159.                                                     */
/* The same model for the matrix cores: x is N x D row-major bfloat16 bit patterns
 * (the design matrix is bf16 exactly), D == 256, N a multiple of 32.  Density and
 * gradient of all chains run as one fused two-stage MFMA contraction (eta = X B,
 * then X^T r) with fp32 accumulation; positions and residuals enter the matrix
 * cores as bf16 hi + lo pairs.  Workspaces of this plan need num_chains a multiple
 * of 128 and UT_FLOAT32; it is sampled by ut_chees_run (all chains in lock step).
 * Reference: the same GLM lowering and gradient as ut_plan_logistic_glm
 * (lowering.jl:664-675, gradient_kernel.jl:578-657).                            */
int ut_plan_logistic_glm_bf16(const uint16_t* x_bf16, const float* y, int64_t n, int32_t d,
                              double prior_loc, double prior_scale, ut_plan** out);

int ut_plan_hier_normal_nc(const float* y, const int32_t* group_offsets,
                           const double* sigma, int32_t num_groups, double mu_loc,
                           double mu_scale, double tau_scale, ut_plan** out);

/* Latent Gaussian process with Poisson counts (direct latent-function form of
 * gp_cholesky, src/distributions/gaussian_process.jl:180-188,283-294):
 *   log l_d ~ normal(ls_loc, ls_scale) (l_d ~ lognormal), d = 1..D (D <= 2)
 *   a ~ halfnormal(amp_scale) (log transform)
 *   z ~ normal(0, 1)^N; f = L z with L L^T = K,
 *   K_ij = a^2 exp(-0.5 sum_d (x_di - x_dj)^2 / l_d^2) + (noise^2 + 1e-8) delta_ij
 *   y_i ~ poisson(exp(f_i))          (poisson src/device/math.jl:230-235)
 * inputs is N x D row-major, counts N non-negative.  Parameter order
 * (log l_1..log l_D, log a, z_1..z_N); P = D + 1 + N; N a multiple of 32.  The
 * density and gradient of all chains at once: binary64 kernel matrices and a
 * batched Cholesky that carries the D + 1 hyperparameter tangents.  Workspaces of
 * this plan are UT_FLOAT32 and are sampled by the lock-step samplers.           */
int ut_plan_gp_poisson(const double* inputs, const int32_t* counts, int32_t n, int32_t d, double noise,
                       double ls_loc, double ls_scale, double amp_scale, ut_plan** out);
/* Verification entry point: the device factorisation of K at the log
 * hyperparameters hyper (C x (D+1)): L (C x N x N, lower), its tangents dL
 * (C x (D+1) x N x N) along each log hyperparameter, and log det K = 2 sum log L_ii.
 * Any output may be NULL.  No reference counterpart (reference: LinearAlgebra
 * cholesky with ForwardDiff duals, gaussian_process.jl:283-294).                */
int ut_gp_factor(const ut_plan* plan, const double* hyper, int32_t num_chains, double* L, double* dL,
                 double* logdet);

/* Neal's funnel: v ~ normal(0, v_scale); x_i ~ normal(0, exp(v/2)), i = 1..dim.
 * Parameter order: (v, x_1..x_dim); P = dim + 1.                                */
int ut_plan_funnel(int32_t dim, double v_scale, ut_plan** out);

/* ---- general static execution plans -------------------------------------------
 * The flat, serialised form of the reference's isbits DeviceExecutionPlan
 * (src/device/plan/steps.jl:1-539, lowered by src/device/plan/lowering.jl:1360-1478)
 * with its staged observation stream (src/device/staging.jl:499-550): a step list
 * in the reference's pre-order, expression nodes, and the observed rows, trip
 * counts and loop starts.  Indices into params and slots are 1-based as in the
 * reference.
 *
 * Steps.  UT_STEP_CHOICE scores a family at argument expressions args[0..2]; a
 * latent choice (value_source > 0) reads unconstrained row value_source through
 * `transform` and adds the transform's log-abs-det; an observed choice
 * (value_source == 0) reads the next staged row at top level, or row value_offset
 * of the iteration's block inside a loop; binding_slot > 0 stores the
 * (constrained) value.  UT_FAMILY_NONCENTERED_NORMAL scores its latent against
 * N(0, 1) and binds mu + sigma z.  UT_FAMILY_DIAG_NORMAL (top level) scores
 * `dimension` rows against one normal(args[0], args[1]) (the reference's
 * DeviceDiagNormalChoiceStepDyn; observed: `dimension` staged rows).
 * UT_FAMILY_BERNOULLI_LOGIT_GLM (loop bodies) scores the staged outcome at
 * value_offset + dimension against bernoullilogit(args[0] + sum_d
 * coef[d] x_d), coef[d] = row coef_source + d, x_d = staged row value_offset + d
 * (DeviceBernoulliLogitGLMChoiceStepDyn).  UT_STEP_DETERMINISTIC binds args[0].
 * UT_STEP_LOOP runs its body (the next body_length steps) trip_counts[loop_id]
 * times, iterator_slot = loop_starts[loop_id] + t, each iteration reading a block
 * of `stride` staged rows.  As in the reference: loops are not nested and hold no
 * latent choices; values bound inside a body are read only inside it.
 *
 * Expression nodes (post-order: operands precede their node; an expression is
 * named by its root node).  UT_OP_LITERAL `value`; UT_OP_SLOT slot `index`;
 * UT_OP_PARAM unconstrained row index + (a >= 0 ? round(node a) : 0), where node a
 * depends on staged data only (a vector latent indexed by data, e.g. theta[g[i]]);
 * UT_OP_OBS row `index` of the iteration's block (loop bodies only: covariates,
 * group indices, known scales); arithmetic ops on nodes a (and b).
 *
 * Families and transforms follow src/device/math.jl:50-293, 1137-1159: out of
 * support scores -Inf, never an error; the positive families with a log(x) or 1/x
 * term (lognormal, gamma, frechet, rayleigh, inversegaussian, inversegamma,
 * weibull) score x below the smallest normal number of the workspace precision
 * as outside the support (math.jl:53-66).  The count families (poisson,
 * binomial, geometric, negativebinomial, and bernoulli's 0/1) are observed only;
 * binomial's trial count carries no derivative.  Gradients are reverse mode through the
 * plan (per step and per loop iteration), fp32 or fp64 as the workspace.
 *
 * flags: UT_PLAN_GENERIC always runs the plan interpreter; otherwise the plans of
 * the fused families (iid-normal logistic GLM, non-centred hierarchical normal,
 * Neal's funnel) run their dedicated kernels, bit-identical to the ut_plan_*
 * constructors; UT_PLAN_MATRIX_CORES lets the GLM with D = 256 take the bf16
 * matrix-core plan (X rounded to bf16).  Limits: P <= 256 (P <= 128 when a loop
 * body reads parameters), 32 slots, 24 nodes per expression, 256 steps.        */
enum { UT_STEP_CHOICE = 1, UT_STEP_DETERMINISTIC = 2, UT_STEP_LOOP = 3 };
enum {
  UT_FAMILY_NORMAL = 1,
  UT_FAMILY_NONCENTERED_NORMAL = 2,
  UT_FAMILY_LOGNORMAL = 3,
  UT_FAMILY_HALFNORMAL = 4,
  UT_FAMILY_HALFCAUCHY = 5,
  UT_FAMILY_EXPONENTIAL = 6,
  UT_FAMILY_CAUCHY = 7,
  UT_FAMILY_LAPLACE = 8,
  UT_FAMILY_LOGISTIC = 9,
  UT_FAMILY_STUDENTT = 10, /* (nu, mu, sigma) */
  UT_FAMILY_GAMMA = 11,    /* (shape, rate) */
  UT_FAMILY_BERNOULLI = 12,
  UT_FAMILY_BERNOULLI_LOGIT = 13,
  UT_FAMILY_POISSON = 14,
  UT_FAMILY_DIAG_NORMAL = 15,
  UT_FAMILY_BERNOULLI_LOGIT_GLM = 16,
  UT_FAMILY_GUMBEL = 17,          /* (mu, scale) */
  UT_FAMILY_FRECHET = 18,         /* (shape, scale) */
  UT_FAMILY_RAYLEIGH = 19,        /* (scale) */
  UT_FAMILY_INVERSEGAUSSIAN = 20, /* (mu, lambda) */
  UT_FAMILY_BETA = 21,            /* (alpha, beta) */
  UT_FAMILY_INVERSEGAMMA = 22,    /* (shape, scale) */
  UT_FAMILY_WEIBULL = 23,         /* (shape, scale) */
  UT_FAMILY_BINOMIAL = 24,        /* (trials, p), observed only */
  UT_FAMILY_GEOMETRIC = 25,       /* (p), observed only */
  UT_FAMILY_NEGATIVEBINOMIAL = 26 /* (successes, p), observed only */
};
enum { UT_TRANSFORM_IDENTITY = 0, UT_TRANSFORM_LOG = 1, UT_TRANSFORM_LOGIT = 2 };
enum {
  UT_OP_LITERAL = 1, UT_OP_SLOT = 2, UT_OP_PARAM = 3, UT_OP_OBS = 4,
  UT_OP_ADD = 5, UT_OP_SUB = 6, UT_OP_MUL = 7, UT_OP_DIV = 8, UT_OP_NEG = 9,
  UT_OP_EXP = 10, UT_OP_LOG = 11, UT_OP_LOG1P = 12, UT_OP_SQRT = 13, UT_OP_POW = 14,
  UT_OP_ABS = 15, UT_OP_TANH = 16, UT_OP_LOGISTIC = 17
};
enum { UT_PLAN_GENERIC = 1, UT_PLAN_MATRIX_CORES = 2 };
/* which engine serves a plan (ut_plan_engine) */
enum { UT_ENGINE_INTERPRETER = 0, UT_ENGINE_LOGISTIC = 1, UT_ENGINE_HIER = 2, UT_ENGINE_FUNNEL = 3,
       UT_ENGINE_GLM_MATRIX_CORES = 4, UT_ENGINE_GP = 5 };

typedef struct ut_plan_expr {
  int32_t op;
  int32_t a, b;   /* operand node indices (-1 unused) */
  int32_t index;  /* SLOT: slot; PARAM: base row; OBS: row within the iteration's block */
  double value;   /* LITERAL */
} ut_plan_expr;

typedef struct ut_plan_step {
  int32_t kind;
  int32_t family;
  int32_t value_source;  /* latent: first unconstrained row; observed: 0 */
  int32_t transform;
  int32_t binding_slot;  /* 0: none */
  int32_t args[3];       /* root nodes of the family arguments / the deterministic expression; -1 unused */
  int32_t dimension;     /* DIAG_NORMAL rows; BERNOULLI_LOGIT_GLM coefficients */
  int32_t coef_source;   /* BERNOULLI_LOGIT_GLM first coefficient row */
  int32_t value_offset;  /* observed choice in a loop body: row within the iteration's block */
  int32_t loop_id;       /* LOOP: 1-based */
  int32_t iterator_slot; /* LOOP: 0 none */
  int32_t body_length;   /* LOOP */
  int32_t stride;        /* LOOP: staged rows per iteration */
} ut_plan_step;

typedef struct ut_plan_desc {
  int32_t num_params;
  int32_t num_slots;
  int32_t num_loops;
  int32_t num_steps;
  const ut_plan_step* steps;
  int32_t num_exprs;
  const ut_plan_expr* exprs;
  int64_t num_observed;     /* staged rows, one column shared by all chains */
  const double* observed;
  const int32_t* trip_counts;  /* num_loops */
  const int32_t* loop_starts;  /* num_loops */
  int32_t flags;
} ut_plan_desc;

/* Host-only audit of a plan description (no device touched): UT_OK, or
 * UT_E_UNSUPPORTED / UT_E_ARG / UT_E_DIM with the reason in ut_last_error(), the
 * checks of the reference's lowering and read/write audit.                     */
int ut_plan_validate(const ut_plan_desc* desc);
int ut_plan_create(const ut_plan_desc* desc, ut_plan** out);
int ut_plan_engine(const ut_plan* plan, int32_t* engine);

int ut_plan_destroy(ut_plan* plan);
int ut_plan_num_params(const ut_plan* plan, int32_t* num_params);

/* ---- workspace -----------------------------------------------------------------
 * Owns the device buffers for num_chains chains of one plan, one HIP stream and
 * the sampler state.  Not thread-safe: one caller at a time per workspace.      */
int ut_workspace_create(const ut_plan* plan, int32_t num_chains, int32_t precision,
                        ut_workspace** out);
int ut_workspace_destroy(ut_workspace* ws);
/* Chain layout of the density kernel and the persistent samplers: wavefronts per
 * chain.  0 (the default) chooses automatically: with few chains over many rows the
 * logistic plan splits each chain's observations over a workgroup of 2, 4 or 8
 * wavefronts (the observation-tiled gradient inside the persistent kernel that the
 * reference defers, docs/persistent-nuts.md:406-414; its tiling rule for the
 * batched gradient is src/device/workspace.jl:123-132); 1, 2, 4, 8 force a layout.
 * The per-wavefront totals are added in wavefront order, so each layout is its own
 * deterministic summation order.  waves > 1 on a plan other than the logistic one
 * is UT_E_UNSUPPORTED.  A NUTS run whose checkpoint stacks (per wavefront) would
 * exceed the LDS at its max_tree_depth fails with UT_E_ARG for an explicit layout
 * and takes fewer wavefronts for the automatic one.  ut_workspace_chain_waves
 * reports the layout in effect (after a sampler run: that run's).  No reference
 * counterpart as an entry point.                                               */
int ut_workspace_set_chain_waves(ut_workspace* ws, int32_t waves);
int ut_workspace_chain_waves(const ut_workspace* ws, int32_t* waves);

/* ---- density -------------------------------------------------------------------
 * params: P x C unconstrained parameters in the workspace precision (float or
 * double per the precision code).  Outputs the unconstrained logjoint (log density
 * plus transform log-abs-det) and its gradient, in the same precision.
 * Reference: device_batched_logjoint!/device_batched_logjoint_gradient!
 * (src/device/api.jl:34-51, 98-124).                                            */
int ut_logjoint(ut_workspace* ws, const void* params, void* values);
int ut_logjoint_gradient(ut_workspace* ws, const void* params, void* values,
                         void* gradients);
/* Device time of the samplers' batched gradient: copies params (P x C, float32) to
 * the workspace, then times `reps` back-to-back calls of the gradient engine the
 * samplers use (the plan's kernels, plus the flow when one is attached; with the
 * log density only when need_logjoint != 0) with HIP events on the workspace's
 * stream, and writes the mean microseconds per call.  A measurement utility with
 * no counterpart in the reference (the benchmark's roofline of the gradient).    */
int ut_gradient_kernel_time(ut_workspace* ws, const void* params, int32_t reps, int32_t need_logjoint,
                            double* us_per_call);
/* Device time of one leapfrog step of the lock-step samplers after a trajectory's first: the
 * batched gradient at params (P x C, float32) and the full kick and drift, with step size 0
 * so that every call does the same work; `reps` back-to-back calls timed with HIP events on
 * the workspace's stream (with a flow on the 16-row kernels and the funnel this is one
 * fused kernel per call).  A measurement utility with no counterpart in the reference.     */
int ut_gradient_step_kernel_time(ut_workspace* ws, const void* params, int32_t reps, double* us_per_call);

/* Allocation-free density and gradient.  No counterpart in the reference: this is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It computes what
 * ut_logjoint_gradient computes for a UT_FLOAT32 workspace of `plan` with
 * num_chains chains, with the same kernels, but takes DEVICE pointers and calls no
 * hipMalloc/hipFree and no allocating library routine: params (P x C,
 * column-major, fp32), values (C) and gradients (P x C) are the caller's device
 * buffers, and the matrix-core GLM plan's scratch (bf16 position planes and
 * per-split partials) is a file-scope device array of UT_NOALLOC_SCRATCH_BYTES
 * that each GPU holds until the process ends.  The work is enqueued on `stream`
 * (a hipStream_t, NULL for the null stream) and the call returns without waiting:
 * results are ready once that stream has reached this point.  Every call on one
 * GPU shares that GPU's scratch, so calls that overlap on one GPU (on different
 * streams, or from several threads) race; the scratch keeps the last call's
 * contents, and nothing in it needs clearing (the kernels write every byte they
 * read).  ut_noalloc_scratch_bytes returns the scratch a call needs (0 for the
 * wavefront-per-chain plans, which use none; -1 for the GP plan, which is
 * rejected with UT_E_UNSUPPORTED because its factorisation scratch is about
 * 4.5 MiB per chain); a call needing more than UT_NOALLOC_SCRATCH_BYTES fails with
 * UT_E_DIM (at D = 256: 4096 chains need 36.3 MiB).                            */
#define UT_NOALLOC_SCRATCH_BYTES (48u << 20)
int64_t ut_noalloc_scratch_bytes(const ut_plan* plan, int32_t num_chains);
int ut_logjoint_gradient_noalloc(const ut_plan* plan, int32_t num_chains, const float* params, float* values,
                                 float* gradients, void* stream);

/* ---- samplers ------------------------------------------------------------------
 * Options mirror the keyword arguments of batched_nuts / batched_hmc
 * (src/inference/api_batched_nuts.jl:22-51, api_batched_hmc.jl:16-39).          */

/* Progress callback (reference `callback` / `callback_every` keywords and
 * _invoke_progress_callback, src/inference/api_single_chain.jl:1-30): called on the
 * calling thread with phase UT_PHASE_WARMUP or UT_PHASE_SAMPLE, the 1-based
 * iteration within the phase, the phase's total, the current step size (the mean
 * over chains where each chain adapts its own; NaN where meaningless) and the
 * cumulative divergent transitions of the phase, every callback_every iterations
 * (<= 0: only at the end of each phase) and on each phase's final iteration.
 * NULL (the default) adds no work.  Draws do not depend on whether a callback is
 * set; each call costs one synchronisation (the persistent samplers then end a
 * launch at every reported iteration).                                          */
enum { UT_PHASE_WARMUP = 0, UT_PHASE_SAMPLE = 1 };
typedef void (*ut_progress_fn)(void* user, int32_t phase, int32_t iteration, int32_t total, double step_size,
                               int64_t divergences);

typedef struct ut_sampler_opts {
  int32_t num_samples;
  int32_t num_warmup;
  int32_t max_tree_depth;         /* NUTS (default 10) */
  int32_t num_leapfrog_steps;     /* HMC (default 10) */
  double step_size;               /* initial step size (NUTS/HMC default 0.1) */
  double target_accept;           /* 0.8 */
  double divergence_threshold;    /* 1000 */
  double mass_matrix_regularization; /* 1e-3 */
  int32_t mass_matrix_min_samples;   /* 10 */
  int32_t adapt_step_size;        /* bool, default 1 */
  int32_t adapt_mass_matrix;      /* bool, default 1 */
  int32_t find_reasonable_step_size; /* bool, default 0 */
  int32_t init_strategy;          /* UT_INIT_PRIOR / UT_INIT_UNIFORM */
  int32_t init_max_retries;       /* 100 */
  const double* initial_params;   /* P x C or NULL */
  uint64_t seed;                  /* run seed; draws are a pure function of it */
  ut_progress_fn callback;        /* NULL: no progress reports */
  void* callback_user;            /* passed back to callback */
  int32_t callback_every;         /* 10 */
} ut_sampler_opts;

/* Fills opts with the reference defaults (num_samples/num_warmup zero). */
void ut_sampler_opts_default(ut_sampler_opts* opts);

/* Per-chain results, laid out like the reference's HMCChain fields stacked over
 * chains (src/inference/data_types.jl:55-80).  samples is P x S x C in the
 * workspace precision (float for UT_FLOAT32); S x C matrices are column-major.
 * Any pointer except samples may be NULL to skip that output.                   */
typedef struct ut_chain_out {
  void* samples;               /* P x S x C unconstrained draws */
  double* logjoint;            /* S x C */
  double* accept_stat;         /* S x C */
  double* energy;              /* S x C */
  double* energy_error;        /* S x C */
  uint8_t* accepted;           /* S x C */
  uint8_t* divergent;          /* S x C */
  int32_t* tree_depth;         /* S x C (NUTS) */
  int32_t* integration_steps;  /* S x C */
  double* step_size;           /* C: final adapted step per chain */
  double* inverse_mass;        /* P: final shared diagonal inverse mass */
  double warmup_seconds;       /* wall time of initialisation + warmup */
  double sampling_seconds;     /* wall time of sampling incl. copy-out */
  int64_t sampling_gradient_evals; /* leapfrog gradient evaluations while sampling */
  double sampling_kernel_seconds;  /* device time of the sampling kernels (HIP events) */
  int32_t sampling_launches;       /* number of sampling kernel launches */
} ut_chain_out;

/* Persistent-kernel NUTS: every chain builds its whole multinomial NUTS tree in
 * one kernel; warmup uses per-chain dual averaging and a pooled diagonal mass
 * (reference _batched_nuts_device_per_chain!, api_batched_nuts.jl:761-980).    */
int ut_nuts_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out);

/* NUTS with tree_strategy = masked (reference src/device/nuts_kernels.jl:61-580,
 * api_batched_nuts.jl:44-47, 95-144): the same multinomial NUTS transition and the
 * same per-chain adaptation as ut_nuts_run, but every chain advances one leaf of its
 * own tree per step and the gradient of all chains comes from one batched call
 * (each leaf's gradient launch for all chains, chains that end an iteration start
 * their next one at once, chains past a launch range are masked).  It therefore
 * runs every plan, including the matrix-core GLM, the GP and workspaces with a flow
 * attached (NUTS in the flow's base space, draws returned mapped through it).  For
 * the wavefront-per-chain plans its draws equal ut_nuts_run's bit for bit.  The
 * per-chain step search runs in lock step the same way.                      */
int ut_nuts_masked_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out);

/* Fixed-length HMC with per-chain step and pooled diagonal mass
 * (reference _run_device_batched_hmc, src/device/hmc_kernels.jl:306-723).      */
int ut_hmc_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out);

/* ChEES-HMC: jittered lock-step HMC with a shared step size (dual averaging to
 * the harmonic-mean acceptance), pooled diagonal mass, and warmup adaptation of
 * the trajectory time T by Adam ascent on the cross-chain ChEES criterion
 * (reference batched_chees, src/inference/api_batched_chees.jl:277-700; its
 * keyword arguments map to the fields below, defaults from the reference).
 * base.num_leapfrog_steps sets the initial T = num_leapfrog_steps * step.
 * trajectory_trace, if not NULL, receives T after each warmup iteration
 * (num_warmup entries; reference `_trajectory_trace`).                          */
typedef struct ut_chees_opts {
  ut_sampler_opts base;              /* target_accept default 0.651 */
  double jitter_amount;              /* 1.0, in [0, 1] */
  int32_t adapt_trajectory_length;   /* bool, default 1 */
  int32_t max_leapfrog_steps;        /* 1000 */
  double trajectory_decay_rate;      /* 0.5 */
  double trajectory_learning_rate;   /* 0.5 */
  double trajectory_adam_beta1;      /* 0.0 */
  double trajectory_adam_beta2;      /* 0.95 */
  double trajectory_adam_epsilon;    /* 1e-8 */
  /* 0 (reference): the criterion's end velocity is the proposal momentum as the
   * reference passes it, i.e. after the final momentum flip; the Adam ascent on
   * that estimate then shortens T towards one step.  1: the velocity in the
   * direction of integration (the ChEES paper's d q'/dT), under which the same
   * ascent lengthens T where longer trajectories decorrelate more.  Not a
   * reference option; default 0 keeps parity.                                */
  int32_t forward_velocity;
} ut_chees_opts;

void ut_chees_opts_default(ut_chees_opts* opts);
int ut_chees_run(ut_workspace* ws, const ut_chees_opts* opts, ut_chain_out* out,
                 double* trajectory_trace);

/* Fixed-length HMC with SHARED adaptation, all chains in lock step: one step
 * size dual-averaged to the mean acceptance (divergent chains count 0), pooled
 * diagonal mass, shared reasonable-step search (reference batched_hmc with
 * per_chain_adaptation = false, src/inference/api_batched_hmc.jl:184-330).  The
 * gradient comes from the workspace's engine for all chains at once, so it runs
 * the matrix-core GLM plan and flow-preconditioned workspaces.                  */
int ut_hmc_shared_run(ut_workspace* ws, const ut_sampler_opts* opts, ut_chain_out* out);

/* MEADS: generalized HMC with one leapfrog step per iteration, persistent
 * partially refreshed momentum and a persistent-slice Metropolis test; step size,
 * damping, slice drift and a diagonal preconditioner come every iteration from the
 * neighbouring fold's ensemble statistics (K complementary folds, a rotating
 * skipped fold) -- reference batched_meads, src/inference/api_batched_meads.jl:
 * 170-547.  No warmup adaptation phase: num_warmup only discards draws.  The fold
 * reductions (Gram trace-ratio eigenvalue estimates) run on the device in
 * binary64.  adapt_trace, if not NULL, receives per iteration the mean step, mean
 * damping and largest standard deviation over folds (3 x (num_warmup +
 * num_samples), the reference `_adapt_trace`).  out->step_size gets the final
 * mean step for every chain; out->inverse_mass the last preconditioner of fold 0. */
typedef struct ut_meads_opts {
  int32_t num_samples;
  int32_t num_warmup;
  int32_t num_folds;              /* 4 */
  int32_t init_strategy;          /* UT_INIT_PRIOR */
  int32_t init_max_retries;       /* 100 */
  double divergence_threshold;    /* 1000 */
  const double* initial_params;   /* P x C or NULL */
  uint64_t seed;
  ut_progress_fn callback;        /* as in ut_sampler_opts */
  void* callback_user;
  int32_t callback_every;         /* 10 */
} ut_meads_opts;

void ut_meads_opts_default(ut_meads_opts* opts);
int ut_meads_run(ut_workspace* ws, const ut_meads_opts* opts, ut_chain_out* out, double* adapt_trace);

/* ---- flow preconditioner ----------------------------------------------------------
 * An affine-coupling normalising flow q = f(z) (the reference's FlowGuide
 * structure, src/inference/vi/guides.jl:1-21,282-369: elementwise affine base
 * x0 = mu + exp(log_scale) z, then num_layers coupling layers with the alternating
 * parity masks, active x <- x exp(s) + t), with MLP conditioners
 * [s_raw, t] = W2 tanh(W1 x_passive + b1) + b2 of width `hidden` and
 * s = c tanh(s_raw / c) (c = scale_clamp).  num_layers = 0 leaves the base alone:
 * the reference's mean-field Gaussian guide.  The reference trains linear
 * conditioners host-side for ADVI; here the network is trained on the GPU, by the
 * ELBO (ut_flow_fit_elbo) or by maximum likelihood on draws (ut_flow_train), and
 * can then precondition HMC (NeuTra): attached to a workspace, the lock-step samplers run in z space on
 * log p(f(z)) + log|det df/dz| and return draws mapped back through f.
 * Conditioner GEMMs run on the matrix cores with bf16 operands and fp32
 * accumulation; weights are fp32.  Row counts of the host-array calls below must
 * be multiples of 128.                                                           */
int ut_flow_create(int32_t num_params, int32_t num_layers, int32_t hidden, double scale_clamp,
                   uint64_t seed, ut_flow** out);
/* The full-rank Gaussian guide as a flow without coupling layers:
 * q = mu + exp(log_scale) z + F z with F strictly lower triangular (replaces the
 * reference's _draw_fullrank_particles! / _accumulate_fullrank_gradients!,
 * src/inference/vi/guides.jl:96-117,164-199, host-only there).  Weights: mu[P],
 * log_scale[P], F[P x P] row-major (entries on and above the diagonal are ignored
 * and get zero gradient); created at zero, the reference's starting point
 * (src/inference/vi.jl:459-460).  Supports ut_flow_forward / inverse / vjp,
 * ut_flow_elbo_grad, ut_flow_fit_elbo and attaching to a workspace (dense-metric
 * HMC in z space); ut_flow_train and ut_flow_loss_grad return UT_E_ARG.          */
int ut_flow_create_fullrank(int32_t num_params, ut_flow** out);
int ut_flow_destroy(ut_flow* flow);
/* flat fp32 weight vector: mu[P], log_scale[P], then per layer W1[hidden x kp],
 * b1[hidden], W2[op x hidden], b2[op] with kp = np rounded up to 128 and
 * op = 2 na rounded up to 128 (padding entries are zero); offsets per layer:
 * {W1, b1, W2, b2}                                                               */
int ut_flow_num_weights(const ut_flow* flow, int64_t* n);
int ut_flow_layer_layout(const ut_flow* flow, int32_t layer, int32_t* num_passive, int32_t* num_active,
                         int64_t* offsets);
int ut_flow_get_weights(const ut_flow* flow, float* out);
int ut_flow_set_weights(ut_flow* flow, const float* in);
int ut_flow_set_base(ut_flow* flow, const double* mu, const double* log_scale);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch` rows drawn
 * with replacement (Philox, `seed`) from draws (n x P row-major host array);
 * loss_trace[step] = minibatch mean negative log-likelihood, or NULL.            */
int ut_flow_train(ut_flow* flow, const float* draws, int64_t n, int32_t steps, int32_t batch,
                  double learning_rate, uint64_t seed, double* loss_trace);
/* q = f(z) with log|det df/dz| per row; z = f^-1(q) with log|det df^-1/dq|      */
int ut_flow_forward(ut_flow* flow, const float* z, int64_t n, float* q, float* logdet);
int ut_flow_inverse(ut_flow* flow, const float* q, int64_t n, float* z, float* logdet_inv);
/* g_z = (dq/dz)^T g_q + d log|det df/dz| / dz (the sampler's gradient path)       */
int ut_flow_vjp(ut_flow* flow, const float* z, int64_t n, const float* g_q, float* g_z);
/* Device time of the sampling path's flow kernels: `reps` back-to-back forward
 * launches, then `reps` backward launches, over n device-resident rows of N(0, I)
 * draws, timed with HIP events on the flow's stream; writes the mean microseconds
 * per launch of each direction.  A measurement utility with no counterpart in the
 * reference (the benchmark's roofline of the flow kernels).                        */
int ut_flow_kernel_times(ut_flow* flow, int64_t n, int32_t reps, double* forward_us, double* backward_us);
/* mean over rows of 0.5|f^-1(q)|^2 - log|det df^-1/dq| + P/2 log 2 pi and its
 * gradient over the flat weights (the training objective)                       */
int ut_flow_loss_grad(ut_flow* flow, const float* q, int64_t n, double* loss, float* grad);
/* attach (or detach with NULL) a flow: the lock-step samplers then precondition
 * with it.  The workspace must be UT_FLOAT32 with num_chains a multiple of 128.   */
int ut_workspace_set_flow(ut_workspace* ws, ut_flow* flow);
/* Reparameterised-ELBO fit of the flow against the workspace's model: the
 * reference's flow ADVI (batched_advi(...; guide = :flow, elbo = :standard),
 * src/inference/vi.jl:700-870, Adam ascent src/inference/vi.jl:205-227, flat-norm
 * clip src/inference/vi/guides.jl:492-502) on the device.  Each iteration draws
 * num_chains particles z ~ N(0, I) (Philox, key = (seed, iteration)), maps them
 * through the flow, scores them with the workspace's gradient engine (the plain
 * model, whether or not a flow is attached) and takes one Adam ascent step
 * (beta1, beta2, adam_epsilon; the reference's defaults 0.9, 0.999, 1e-8; fresh
 * moments per call) on the mean over the valid
 * particles (finite logjoint, log-determinant and gradient) of the gradient of
 * g . f(z) + log|det df/dz| with g = grad log p(f(z)) held fixed.  iwae_group = K > 1
 * (dividing num_chains) optimises the importance-weighted bound instead
 * (elbo = :iwae, src/inference/vi/guides.jl:373-425): per group of K particles the
 * log-mean-exp of log w = log p - log q, and the gradient weights each particle by
 * its within-group softmax weight, averaged over the groups with a valid particle.
 * elbo_history[i] = the optimised objective, standard_elbo_history[i] = the sampled
 * standard ELBO (mean log w over the valid particles), grad_norm_history[i] = the
 * gradient norm, clipped at gradient_clip
 * (INFINITY: no clipping); either may be NULL.  The flow keeps the last weights;
 * best_weights (ut_flow_num_weights floats, or NULL) receives the weights of the
 * best ELBO seen and best_elbo that value.  UT_E_ARG when an iteration has no
 * valid particle (reference: ArgumentError "encountered only non-finite ..."): the
 * iterations run without a host round trip, so the error is returned after the
 * loop, names the first such iteration, and the weights stay as they were before it
 * (as if the loop had stopped there).  Works for every guide: coupling flows, the
 * mean-field base (num_layers = 0) and ut_flow_create_fullrank's full-rank guide.
 * The workspace must be UT_FLOAT32 with num_chains a multiple of 128.            */
int ut_flow_fit_elbo(ut_workspace* ws, ut_flow* flow, int32_t num_iterations, double learning_rate, double beta1,
                     double beta2, double adam_epsilon, double gradient_clip, int32_t iwae_group, uint64_t seed,
                     double* elbo_history, double* standard_elbo_history, double* grad_norm_history,
                     float* best_weights, double* best_elbo);
/* Pointwise log-likelihood of the observed sites for draws in unconstrained space:
 * the reference's pointwise_loglikelihood (src/inference/result_interface.jl:384-410,
 * src/evaluator_pointwise.jl:342), the log_likelihood group of its four-argument
 * to_arviz_dict (src/inference/export.jl:276-338).  out is num_draws x
 * num_observations row-major (binary64); draws num_draws x P row-major.  Logistic
 * plans (float and bf16, the latter with its bf16 X): y_i ~ bernoullilogit(x_i . beta),
 * observations in data order.
 * Hierarchical plan: y ~ normal(mu + tau theta_raw_j, sigma_j), observations group by
 * group in the order of group_offsets.  GP plan: y_i ~ poisson(exp(f_i)), f = L z
 * from the binary64 factorisation at binary32 positions (as the model's evaluator
 * takes them).  Funnel plan: no observations (size 0).                           */
int ut_pointwise_loglik_size(const ut_plan* plan, int64_t* num_observations);
int ut_pointwise_loglik(const ut_plan* plan, const double* draws, int64_t num_draws, double* out);
/* num_chains initial positions (num_chains x P, chain-major, unconstrained) as the
 * samplers draw them for `seed` before their non-finite re-draw loop: prior draws
 * (UT_INIT_PRIOR) or Uniform(-2, 2) (UT_INIT_UNIFORM) from Philox; host-only.
 * Reference: _initial_hmc_position / _initial_batched_hmc_positions
 * (src/inference/hmc_support.jl:114-200, 266-290), used by batched_advi for its starting
 * location (src/inference/vi.jl:83-93).                                          */
int ut_initial_params(const ut_plan* plan, int32_t strategy, uint64_t seed, int32_t num_chains, double* out);
/* One evaluation of that objective at given particles z (num_chains x P host
 * rows): the sampled ELBO, the unclipped mean weight gradient over the valid
 * particles (ut_flow_num_weights floats), and optionally the model gradient at
 * f(z) that was held fixed (num_chains x P, zero on invalid particles) and the
 * validity flags (num_chains), and the particles' weights in the gradient
 * (num_chains: 1 / valid count, or the importance weights over the groups used).
 * The analogue of ut_flow_loss_grad for the ELBO.                                 */
int ut_flow_elbo_grad(ut_workspace* ws, ut_flow* flow, const float* z, int32_t iwae_group, double* elbo, float* grad,
                      float* g_q, int32_t* valid, float* row_weights);

/* ---- pinned host memory ----------------------------------------------------------
 * Page-locked host buffers for sampler outputs: the samplers copy each launch
 * range's draws out while the next range computes, at full link bandwidth when the
 * destination is pinned.  No reference counterpart (the reference returns Julia
 * arrays); used by the host layer to allocate ut_chain_out buffers.              */
int ut_host_alloc(int64_t bytes, void** out);
int ut_host_free(void* ptr);

/* ---- matrix cores ----------------------------------------------------------------
 * C[m x n] = A[m x k] * B[k x n] on the gfx950 MFMA units, bfloat16 inputs (A row-major,
 * B given transposed as n x k row-major), fp32 accumulation, split-K over `splits`.
 * Host pointers; m, n multiples of 128, k of 32.  The same kernel computes the
 * dense GLM contractions and the flow layers; exposed for verification.  No
 * reference counterpart.                                                          */
int ut_gemm_bf16(const uint16_t* a, const uint16_t* bt, float* c, int32_t m, int32_t n, int32_t k,
                 int32_t splits);

/* ---- RNG -----------------------------------------------------------------------
 * Host evaluation of one Philox4x32-10 block (reference philox4x32, rng.jl:219). */
void ut_philox4x32(const uint32_t key[2], const uint32_t ctr[4], uint32_t out[4]);

/* Device fill for the determinism checks: for i in [0, n):
 *   blocks[4i..4i+3] = philox4x32((chain, iteration), (stream, first + i, 0, 0))
 *   uniform_f32[i]  = device_rand_uniform(Float32, ...) of the same coordinate
 *   normal_f32[i]   = device_rand_normal(Float32, ...) of the same coordinate
 *   uniform_f64[i]  = device_rand_uniform(Float64, ...) of the same coordinate
 * Output arrays are host pointers; any may be NULL.                             */
int ut_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream, uint32_t first,
                   int64_t n, uint32_t* blocks, float* uniform_f32, float* normal_f32,
                   double* uniform_f64);

#ifdef __cplusplus
}
#endif

#endif /* UNCERTAINTEA_HIP_H */
