"""Static models lowered to device plans.

The reference lowers a `@tea (static)` model to an isbits DeviceExecutionPlan and
stages its observations (src/device/plan/lowering.jl:1360-1478,
src/device/staging.jl:499-550).  Here each supported model family is a class that
holds its data and creates the matching `ut_plan` through the C ABI; the plan
keeps the data resident in HBM for its lifetime.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib as L


class TeaModel:
    """A static model with a fixed parameter layout (unconstrained order)."""

    name = "model"

    def __init__(self):
        self._plan = None

    # layout ---------------------------------------------------------------
    @property
    def num_params(self) -> int:
        raise NotImplementedError

    @property
    def param_names(self) -> list[str]:
        raise NotImplementedError

    def constrain(self, unconstrained: np.ndarray) -> np.ndarray:
        """Map unconstrained draws (..., P) to constrained values (..., P)."""
        return np.asarray(unconstrained, dtype=np.float64)

    # conditioning -----------------------------------------------------------
    # The reference conditions a model at workspace creation on its arguments and a
    # choicemap of observed values (src/device/api.jl:34-88, staged once by
    # src/device/staging.jl:499-550).  Here a model object carries its data, so
    # conditioning returns a model of the same family with the arguments and
    # observations replaced; `args` and `constraints` name what each family takes.
    arg_names: tuple = ()
    observed_names: tuple = ()

    def condition(self, args=(), constraints=None) -> "TeaModel":
        """The model conditioned on `args` (a tuple in the order of `arg_names`;
        () keeps the model's own) and `constraints` (a mapping from the names in
        `observed_names` to observed values; None or {} keeps the model's own)."""
        args = tuple(args) if args is not None else ()
        cons = dict(constraints or {})
        if not args and not cons:
            return self
        if args and len(args) != len(self.arg_names):
            raise ValueError(f"DimensionMismatch: {self.name} takes {len(self.arg_names)} model arguments "
                             f"{self.arg_names}, got {len(args)}")
        unknown = sorted(set(cons) - set(self.observed_names))
        if unknown:
            raise ValueError(f"model {self.name} cannot be lowered to the device logjoint path with constraints on "
                             f"{unknown}: only the observed choices {list(self.observed_names)} can be constrained "
                             f"(latent choices stay parameters).\nSee device_lowering_report(model) for details.")
        return self._conditioned(args, cons)

    def _conditioned(self, args: tuple, constraints: dict) -> "TeaModel":
        raise ValueError(f"model {self.name} takes no arguments or observations.\n"
                         "See device_lowering_report(model) for details.")

    def lowering_report(self) -> dict:
        """How this model reaches the device (see device_lowering_report)."""
        return {"model": self.name, "lowerable": True, "plan": type(self).__name__, "num_params": self.num_params,
                "arguments": list(self.arg_names), "observed": list(self.observed_names), "issues": []}

    # device plan ------------------------------------------------------------
    def _create_plan(self, out: ct.c_void_p) -> int:
        raise NotImplementedError

    @property
    def plan(self) -> ct.c_void_p:
        if self._plan is None:
            h = ct.c_void_p()
            L.check(self._create_plan(ct.byref(h)))
            self._plan = h
            self._destroy = L.lib().ut_plan_destroy  # kept: module globals may be gone at exit
        return self._plan

    def __del__(self):
        if getattr(self, "_plan", None) is not None:
            self._destroy(self._plan)
            self._plan = None


class LogisticRegression(TeaModel):
    """beta_d ~ normal(prior_loc, prior_scale); y_i ~ bernoullilogit(x_i . beta)."""

    name = "logistic_regression"

    def __init__(self, x: np.ndarray, y: np.ndarray, prior_loc: float = 0.0, prior_scale: float = 1.0):
        super().__init__()
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        if self.x.ndim != 2 or self.x.shape[0] != self.y.shape[0]:
            raise ValueError("DimensionMismatch: x must be N x D with N = len(y)")
        self.prior_loc = float(prior_loc)
        self.prior_scale = float(prior_scale)

    arg_names = ("x",)
    observed_names = ("y",)

    @property
    def num_params(self):
        return self.x.shape[1]

    @property
    def param_names(self):
        return [f"beta[{d + 1}]" for d in range(self.num_params)]

    def _conditioned(self, args, constraints):
        x = args[0] if args else (self.x_bf16 if hasattr(self, "x_bf16") else self.x)
        return type(self)(x, constraints.get("y", self.y), self.prior_loc, self.prior_scale)

    def _create_plan(self, out):
        n, d = self.x.shape
        return L.lib().ut_plan_logistic_glm(L.ptr(self.x), L.ptr(self.y), n, d, self.prior_loc,
                                            self.prior_scale, out)


class LogisticRegressionBF16(LogisticRegression):
    """The logistic regression with a bfloat16 design matrix, evaluated for all
    chains at once on the matrix cores (D = 256, N a multiple of 32; chains in
    multiples of 128).  `x` is rounded to bfloat16 (round to nearest even) once, and
    that rounded matrix is the model's data: `x_model` holds it as float32."""

    name = "logistic_regression_bf16"

    def __init__(self, x: np.ndarray, y: np.ndarray, prior_loc: float = 0.0, prior_scale: float = 1.0):
        xb = np.asarray(x)
        self.x_bf16 = np.ascontiguousarray(xb if xb.dtype == np.uint16 else L.to_bf16(xb), dtype=np.uint16)
        super().__init__(L.from_bf16(self.x_bf16), y, prior_loc, prior_scale)

    @property
    def x_model(self) -> np.ndarray:
        return self.x

    def _create_plan(self, out):
        n, d = self.x_bf16.shape
        return L.lib().ut_plan_logistic_glm_bf16(L.ptr(self.x_bf16), L.ptr(self.y), n, d, self.prior_loc,
                                                 self.prior_scale, out)


class HierarchicalNormalNC(TeaModel):
    """Non-centred hierarchical normal.

    mu ~ normal(mu_loc, mu_scale); tau ~ halfcauchy(tau_scale);
    theta_raw_j ~ normal(0, 1); theta_j = mu + tau theta_raw_j; y_ij ~ normal(theta_j, sigma_j).
    Unconstrained layout: (mu, log tau, theta_raw_1..J).
    """

    name = "hierarchical_normal_nc"

    def __init__(self, y: np.ndarray, group_offsets: np.ndarray, sigma: np.ndarray, mu_loc: float = 0.0,
                 mu_scale: float = 5.0, tau_scale: float = 5.0):
        super().__init__()
        self.y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        self.offsets = np.ascontiguousarray(group_offsets, dtype=np.int32)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        if self.offsets.shape[0] != self.sigma.shape[0] + 1 or self.offsets[-1] != self.y.shape[0]:
            raise ValueError("DimensionMismatch: group_offsets must have J+1 entries ending at len(y)")
        self.mu_loc, self.mu_scale, self.tau_scale = float(mu_loc), float(mu_scale), float(tau_scale)

    arg_names = ("group_offsets", "sigma")
    observed_names = ("y",)

    def _conditioned(self, args, constraints):
        off, sig = args if args else (self.offsets, self.sigma)
        return HierarchicalNormalNC(constraints.get("y", self.y), off, sig, self.mu_loc, self.mu_scale,
                                    self.tau_scale)

    @property
    def num_groups(self):
        return self.sigma.shape[0]

    @property
    def num_params(self):
        return self.num_groups + 2

    @property
    def param_names(self):
        return ["mu", "tau"] + [f"theta_raw[{j + 1}]" for j in range(self.num_groups)]

    def constrain(self, u):
        c = np.array(u, dtype=np.float64, copy=True)
        c[..., 1] = np.exp(c[..., 1])
        return c

    def _create_plan(self, out):
        return L.lib().ut_plan_hier_normal_nc(L.ptr(self.y), L.ptr(self.offsets), L.ptr(self.sigma),
                                              self.num_groups, self.mu_loc, self.mu_scale, self.tau_scale, out)


class NealsFunnel(TeaModel):
    """v ~ normal(0, v_scale); x_i ~ normal(0, exp(v/2)), i = 1..dim.  Layout (v, x_1..x_dim)."""

    name = "neals_funnel"

    def __init__(self, dim: int = 100, v_scale: float = 3.0):
        super().__init__()
        self.dim = int(dim)
        self.v_scale = float(v_scale)

    @property
    def num_params(self):
        return self.dim + 1

    @property
    def param_names(self):
        return ["v"] + [f"x[{i + 1}]" for i in range(self.dim)]

    def _create_plan(self, out):
        return L.lib().ut_plan_funnel(self.dim, self.v_scale, out)


class GaussianProcessPoisson(TeaModel):
    """Latent GP with Poisson counts (direct latent-function form of gp_cholesky):
    log l_d ~ normal(ls_loc, ls_scale); a ~ halfnormal(amp_scale); z ~ normal(0, 1)^N;
    f = chol(K) z; y_i ~ poisson(exp(f_i)), K the ARD squared-exponential kernel plus
    (noise^2 + 1e-8) I.  Unconstrained layout (log l_1..log l_D, log a, z_1..z_N)."""

    name = "gp_poisson"

    def __init__(self, inputs: np.ndarray, counts: np.ndarray, noise: float = 1e-2, ls_loc: float = 0.0,
                 ls_scale: float = 0.5, amp_scale: float = 1.0):
        super().__init__()
        self.inputs = np.ascontiguousarray(inputs, dtype=np.float64)
        if self.inputs.ndim == 1:
            self.inputs = self.inputs[:, None]
        self.counts = np.ascontiguousarray(counts, dtype=np.int32).reshape(-1)
        if self.inputs.shape[0] != self.counts.shape[0]:
            raise ValueError("DimensionMismatch: inputs must be N x D with N = len(counts)")
        self.noise, self.ls_loc, self.ls_scale, self.amp_scale = float(noise), float(ls_loc), float(ls_scale), \
            float(amp_scale)

    arg_names = ("inputs",)
    observed_names = ("y",)

    def _conditioned(self, args, constraints):
        inputs = args[0] if args else self.inputs
        return GaussianProcessPoisson(inputs, constraints.get("y", self.counts), self.noise, self.ls_loc,
                                      self.ls_scale, self.amp_scale)

    @property
    def num_points(self):
        return self.inputs.shape[0]

    @property
    def input_dim(self):
        return self.inputs.shape[1]

    @property
    def num_params(self):
        return self.input_dim + 1 + self.num_points

    @property
    def param_names(self):
        return [f"lengthscale[{d + 1}]" for d in range(self.input_dim)] + ["amplitude"] + \
            [f"z[{i + 1}]" for i in range(self.num_points)]

    def constrain(self, u):
        c = np.array(u, dtype=np.float64, copy=True)
        c[..., :self.input_dim + 1] = np.exp(c[..., :self.input_dim + 1])
        return c

    def _create_plan(self, out):
        n, d = self.inputs.shape
        return L.lib().ut_plan_gp_poisson(L.ptr(self.inputs), L.ptr(self.counts), n, d, self.noise, self.ls_loc,
                                          self.ls_scale, self.amp_scale, out)

    def factor(self, log_hyper):
        """Device factorisation at log hyperparameters (C x (D+1)): (L, dL, logdet)."""
        h = np.ascontiguousarray(np.atleast_2d(log_hyper), dtype=np.float64)
        C, N, T = h.shape[0], self.num_points, self.input_dim + 1
        Lm = np.empty((C, N, N))
        dL = np.empty((C, T, N, N))
        ld = np.empty(C)
        L.check(L.lib().ut_gp_factor(self.plan, L.ptr(h), C, L.ptr(Lm), L.ptr(dL), L.ptr(ld)))
        return Lm, dL, ld
