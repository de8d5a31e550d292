"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
    L = ct.CDLL(str(LIB_PATH))
    vp, i32, i64, u32, dbl = ct.c_void_p, ct.c_int32, ct.c_int64, ct.c_uint32, ct.c_double
    sig = {
        "orc_philox4x32": (None, [vp, vp, vp]),
        "orc_philox_fill": (None, [u32, u32, u32, u32, i64, vp, vp, vp, vp]),
        "orc_math_f32": (None, [ct.c_int, vp, vp, i64]),
        "orc_math_f64": (None, [ct.c_int, vp, vp, i64]),
        "orc_model_logistic": (vp, [vp, vp, i64, i32, dbl, dbl]),
        "orc_model_set_tiles": (i32, [vp, i32]),
        "orc_model_hier": (vp, [vp, vp, vp, i32, dbl, dbl, dbl]),
        "orc_model_funnel": (vp, [i32, dbl]),
        "orc_model_gp": (vp, [vp, vp, i32, i32, dbl, dbl, dbl, dbl]),
        "orc_gp_eval": (None, [vp, vp, i32, vp, vp, vp]),
        "orc_step_search": (None, [vp, i32, vp, vp, vp, u32, vp]),
        "orc_nuts_subtree": (None, [vp, vp, vp, vp, dbl, i32, i32, vp, vp, vp, vp, vp]),
        "orc_model_free": (None, [vp]),
        "orc_model_plan": (vp, [vp]),
        "orc_plan_error": (ct.c_char_p, []),
        "orc_plan_prior": (None, [vp, i32, u32, u32, vp]),
        "orc_set_threads": (None, [i32]),
        "orc_model_num_params": (i32, [vp]),
        "orc_logjoint_grad_f64": (None, [vp, vp, i32, vp, vp]),
        "orc_logjoint_grad_twin": (None, [vp, i32, vp, i32, vp, vp]),
        "orc_sampler_run": (ct.c_int, [vp, i32, i32, vp, i32, i32, vp, vp, i32]),
        "orc_warmup_schedule": (i32, [i32, vp, vp, vp, i32]),
        "orc_nuts_sample_from": (i64, [vp, i32, i32, i32, vp, vp, vp, i32, i32, ct.c_uint64, i32]),
        "orc_dual_averaging": (None, [dbl, dbl, vp, i32, vp, vp]),
        "orc_running_variance": (dbl, [vp, vp, vp, i32, i32, i32, i32, i32, dbl, vp]),
        "orc_chees_run": (ct.c_int, [vp, i32, vp, i32, i32, vp, vp, vp, i32]),
        "orc_meads_run": (ct.c_int, [vp, i32, vp, i32, i32, vp, vp, vp]),
        "orc_meads_max_eig": (dbl, [vp, i32, i32]),
        "orc_meads_slice_drift": (dbl, [dbl, dbl]),
        "orc_flow_forward": (None, [i32, i32, i32, dbl, i32, vp, vp, i64, vp, vp, vp]),
        "orc_flow_inverse": (None, [i32, i32, i32, dbl, i32, vp, vp, i64, vp, vp]),
        "orc_flow_loss_grad": (dbl, [i32, i32, i32, dbl, i32, vp, vp, i64, vp]),
        "orc_flow_num_weights": (i64, [i32, i32, i32]),
        "orc_flow_elbo_grad": (None, [i32, i32, i32, dbl, i32, vp, vp, vp, vp, vp, i64, vp]),
        "orc_fullrank_forward": (None, [i32, vp, vp, i64, vp, vp]),
        "orc_fullrank_inverse": (None, [i32, vp, vp, i64, vp, vp]),
        "orc_fullrank_vjp": (None, [i32, vp, vp, i64, vp]),
        "orc_fullrank_elbo_grad": (None, [i32, vp, vp, vp, vp, vp, i64, vp]),
        "orc_chees_halton": (None, [i32, vp]),
        "orc_chees_steps": (i32, [dbl, dbl, dbl]),
        "orc_chees_trajectory": (None, [dbl, dbl, dbl, dbl, dbl, dbl, i32, vp, vp, vp, vp, i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle entry points that take no threads argument."""
    _load().orc_set_threads(int(n))


def _p(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


def philox4x32(key, ctr):
    k = np.ascontiguousarray(key, dtype=np.uint32)
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _load().orc_philox4x32(_p(k), _p(c), _p(out))
    return out


def philox_fill(chain, iteration, stream, first, n):
    blocks = np.zeros((n, 4), dtype=np.uint32)
    uf = np.zeros(n, dtype=np.float32)
    nf = np.zeros(n, dtype=np.float32)
    ud = np.zeros(n, dtype=np.float64)
    _load().orc_philox_fill(chain, iteration, stream, first, n, _p(blocks), _p(uf), _p(nf), _p(ud))
    return blocks, uf, nf, ud


MATH = {"exp": 0, "log": 1, "log1p": 2, "cos": 3, "log1p_exp": 4, "logistic": 5, "pow_m075": 6}


def math_f32(fn: str, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    _load().orc_math_f32(MATH[fn], _p(x), _p(y), x.size)
    return y


def math_f64(fn: str, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    _load().orc_math_f64(MATH[fn], _p(x), _p(y), x.size)
    return y


class OracleModel:
    """Oracle twin of a uncertaintea_amd model (same data, same layout)."""

    def __init__(self, model, waves_per_chain: int = 1):
        L = _load()
        self.model = model
        kind = type(model).__name__
        if kind in ("LogisticRegression", "LogisticRegressionBF16"):
            self._keep = (model.x, model.y)
            self.h = L.orc_model_logistic(_p(model.x), _p(model.y), model.x.shape[0], model.x.shape[1],
                                          model.prior_loc, model.prior_scale)
        elif kind == "HierarchicalNormalNC":
            self._keep = (model.y, model.offsets, model.sigma)
            self.h = L.orc_model_hier(_p(model.y), _p(model.offsets), _p(model.sigma), model.num_groups,
                                      model.mu_loc, model.mu_scale, model.tau_scale)
        elif kind == "NealsFunnel":
            self.h = L.orc_model_funnel(model.dim, model.v_scale)
        elif kind == "PlanModel":
            import ctypes as _ct
            self._keep = (model,)
            self.h = L.orc_model_plan(_ct.byref(model.desc))
            if not self.h:
                raise ValueError(L.orc_plan_error().decode())
        elif kind == "GaussianProcessPoisson":
            self._keep = (model.inputs, model.counts)
            self.h = L.orc_model_gp(_p(model.inputs), _p(model.counts), model.num_points, model.input_dim,
                                    model.noise, model.ls_loc, model.ls_scale, model.amp_scale)
        else:
            raise TypeError(kind)
        self.P = L.orc_model_num_params(self.h)
        if waves_per_chain != 1 and L.orc_model_set_tiles(self.h, int(waves_per_chain)) != 0:
            raise ValueError(f"waves_per_chain {waves_per_chain}: logistic models with 1, 2, 4 or 8 only")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_model_free(self.h)
            self.h = None

    def logjoint_grad_f64(self, params):
        """Reference-formula binary64 values/gradients for P x C params."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).T)
        C = p.shape[0]
        lj = np.empty(C)
        g = np.empty((C, self.P))
        _load().orc_logjoint_grad_f64(self.h, _p(p), C, _p(lj), _p(g))
        return lj, g.T

    def prior_column(self, strategy: int, chain: int, key: int) -> np.ndarray:
        """A plan model's initial column as the samplers draw it (UT_INIT_PRIOR or
        UT_INIT_UNIFORM) for the Philox coordinates (chain, key)."""
        col = np.empty(self.P)
        _load().orc_plan_prior(self.h, int(strategy), int(chain), int(key), _p(col))
        return col

    def gp_eval(self, params):
        """GP model: binary64 (logjoint, gradient P x C, log det K) for P x C params."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).T)
        C = p.shape[0]
        lj, ld = np.empty(C), np.empty(C)
        g = np.empty((C, self.P))
        _load().orc_gp_eval(self.h, _p(p), C, _p(lj), _p(g), _p(ld))
        return lj, g.T, ld

    def step_search(self, params, inverse_mass, anchors, itc):
        """binary64 per-chain reasonable step-size search from P x C params with per-chain
        anchors at Philox iteration coordinate itc."""
        q = np.ascontiguousarray(np.asarray(params, dtype=np.float64).T)
        C = q.shape[0]
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(anchors, dtype=np.float64), (C,)))
        im = np.ascontiguousarray(inverse_mass, dtype=np.float64)
        out = np.empty(C)
        _load().orc_step_search(self.h, C, _p(q), _p(im), _p(a), int(itc) & 0xFFFFFFFF, _p(out))
        return out

    def nuts_subtree(self, q, p, inv_mass, step, direction, depth):
        """binary64 NUTS subtree of 2**depth leaves from (q, p) with the checkpoint U-turn test.
        Returns (steps, turning, divergent, traj_q, traj_p) with the full leapfrog trajectory
        traj_* of shape (2**depth, P) in leaf order."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        im = np.ascontiguousarray(inv_mass, dtype=np.float64)
        n = 1 << int(depth)
        tq, tp = np.empty((n, self.P)), np.empty((n, self.P))
        st, tu, dv = ct.c_int32(), ct.c_int32(), ct.c_int32()
        _load().orc_nuts_subtree(self.h, _p(q), _p(p), _p(im), float(step), int(direction), int(depth),
                                 ct.byref(st), ct.byref(tu), ct.byref(dv), _p(tq), _p(tp))
        return st.value, bool(tu.value), bool(dv.value), tq, tp

    def logjoint_grad_twin(self, params, precision=np.float32):
        """Kernel-exact values/gradients (float32 or float64) for P x C params."""
        dt = np.dtype(precision)
        p = np.ascontiguousarray(np.asarray(params, dtype=dt).T)
        C = p.shape[0]
        lj = np.empty(C, dtype=dt)
        g = np.empty((C, self.P), dtype=dt)
        _load().orc_logjoint_grad_twin(self.h, 0 if dt == np.float32 else 1, _p(p), C, _p(lj), _p(g))
        return lj, g.T

    def run(self, kind: str, mode: str, num_chains: int, *, num_samples: int, num_warmup: int = 0,
            step_size: float = 0.1, max_tree_depth: int = 10, num_leapfrog_steps: int = 10,
            initial_params=None, init_strategy: str = "prior", target_accept=None, seed: int = 0,
            adapt_step_size=True, adapt_mass_matrix=True, find_reasonable_step_size=False, threads: int = 0,
            stop_after: int = -1, divergence_threshold=1000.0, chees=None, num_folds: int = 4):
        """Run the oracle sampler.  kind 'nuts'|'hmc'|'hmc_shared'|'chees'|'meads'; mode 'twin' (float32,
        kernel-exact) or 'f64' (binary64 reference formulas).  `chees` holds the
        ChEES keyword arguments (reference names).  Returns a dict."""
        from uncertaintea_amd import _lib as UL  # struct layouts only
        if target_accept is None:
            target_accept = 0.651 if kind == "chees" else 0.8  # the reference's defaults
        lockstep = kind in ("chees", "hmc_shared")
        o = UL.SamplerOpts()
        o.num_samples, o.num_warmup = num_samples, num_warmup
        o.max_tree_depth, o.num_leapfrog_steps = max_tree_depth, num_leapfrog_steps
        o.step_size, o.target_accept, o.divergence_threshold = step_size, target_accept, divergence_threshold
        o.mass_matrix_regularization, o.mass_matrix_min_samples = 1e-3, 10
        o.adapt_step_size, o.adapt_mass_matrix = int(adapt_step_size), int(adapt_mass_matrix)
        o.find_reasonable_step_size = int(find_reasonable_step_size)
        o.init_strategy = 0 if init_strategy == "prior" else 1
        o.init_max_retries = 100
        o.seed = seed
        C, P, S = num_chains, self.P, num_samples
        init = None
        if initial_params is not None:
            init = np.ascontiguousarray(np.asarray(initial_params, dtype=np.float64).T)
            o.initial_params = init.ctypes.data_as(ct.POINTER(ct.c_double))
        dt = np.float32 if mode == "twin" else np.float64
        samples = np.zeros((C, S, P), dtype=dt)
        res = {k: np.zeros((C, S)) for k in ("logjoint", "accept_stat", "energy", "energy_error")}
        acc = np.zeros((C, S), dtype=np.uint8)
        div = np.zeros((C, S), dtype=np.uint8)
        depth = np.zeros((C, S), dtype=np.int32)
        steps = np.zeros((C, S), dtype=np.int32)
        step = np.zeros(C)
        im = np.zeros(P)
        out = UL.ChainOut()
        out.samples = samples.ctypes.data_as(ct.c_void_p)
        for k in res:
            setattr(out, k, res[k].ctypes.data_as(ct.POINTER(ct.c_double)))
        out.accepted = acc.ctypes.data_as(ct.POINTER(ct.c_uint8))
        out.divergent = div.ctypes.data_as(ct.POINTER(ct.c_uint8))
        out.tree_depth = depth.ctypes.data_as(ct.POINTER(ct.c_int32))
        out.integration_steps = steps.ctypes.data_as(ct.POINTER(ct.c_int32))
        out.step_size = step.ctypes.data_as(ct.POINTER(ct.c_double))
        out.inverse_mass = im.ctypes.data_as(ct.POINTER(ct.c_double))
        trace = None
        if kind == "meads":
            mo = UL.MeadsOpts()
            mo.num_samples, mo.num_warmup = num_samples, num_warmup
            mo.num_folds = int(num_folds)
            mo.init_strategy = o.init_strategy
            mo.init_max_retries = 100
            mo.divergence_threshold = divergence_threshold
            mo.initial_params = o.initial_params
            mo.seed = seed
            trace = np.zeros((num_warmup + num_samples, 3))
            rc = _load().orc_meads_run(self.h, 0 if mode == "twin" else 1, ct.byref(mo), C, threads, _p(samples),
                                       ct.byref(out), _p(trace))
        elif lockstep:
            co = UL.CheesOpts()
            co.base = o  # ChEES defaults restated from the reference's keyword defaults
            kw = dict(jitter_amount=1.0, adapt_trajectory_length=True, max_leapfrog_steps=1000,
                      trajectory_decay_rate=0.5, trajectory_learning_rate=0.5, trajectory_adam_beta1=0.0,
                      trajectory_adam_beta2=0.95, trajectory_adam_epsilon=1e-8, forward_velocity=False)
            kw.update(chees or {})
            for k, v in kw.items():
                setattr(co, k, int(v) if isinstance(getattr(co, k), int) else float(v))
            trace = np.zeros(num_warmup)
            rc = _load().orc_chees_run(self.h, 0 if mode == "twin" else 1, ct.byref(co), C, threads, _p(samples),
                                       ct.byref(out), _p(trace), int(kind == "hmc_shared"))
        else:
            rc = _load().orc_sampler_run(self.h, 0 if kind == "nuts" else 1, 0 if mode == "twin" else 1,
                                         ct.byref(o), C, threads, _p(samples), ct.byref(out), stop_after)
        if rc != 0:
            raise RuntimeError(f"oracle sampler failed with status {rc}")
        res.update(samples=samples, accepted=acc.astype(bool), divergent=div.astype(bool), tree_depth=depth,
                   integration_steps=steps, step_size=step, inverse_mass=im, trajectory_trace=trace,
                   warmup_seconds=out.warmup_seconds, sampling_seconds=out.sampling_seconds)
        return res


def nuts_sample_from(om: OracleModel, mode: str, positions, steps, inverse_mass, first_iteration: int,
                     iters: int, seed: int = 0, max_depth: int = 10, threads: int = 0):
    """Run `iters` NUTS sampling iterations (no adaptation) from `positions` (P x C,
    updated in place); returns the number of leapfrog gradient evaluations."""
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).T)
    st = np.ascontiguousarray(steps, dtype=np.float64)
    im = np.ascontiguousarray(inverse_mass, dtype=np.float64)
    C = pos.shape[0]
    n = _load().orc_nuts_sample_from(om.h, 0 if mode == "twin" else 1, C, threads, _p(pos), _p(st), _p(im),
                                     first_iteration, iters, seed, max_depth)
    return int(n), pos.T


def warmup_schedule(n: int):
    ib = ct.c_int32()
    tb = ct.c_int32()
    ends = np.zeros(64, dtype=np.int32)
    k = _load().orc_warmup_schedule(n, ct.byref(ib), ct.byref(tb), _p(ends), 64)
    return ib.value, [int(e) for e in ends[:k]], tb.value


def dual_averaging(step: float, target: float, accepts):
    a = np.ascontiguousarray(accepts, dtype=np.float64)
    steps = np.empty_like(a)
    fin = ct.c_double()
    _load().orc_dual_averaging(step, target, _p(a), a.size, _p(steps), ct.byref(fin))
    return steps, fin.value


def running_variance(hist, acc, flags, window_length, use_accepted=False, reg=1e-3):
    h = np.ascontiguousarray(hist, dtype=np.float64)
    it, C, P = h.shape
    a = np.ascontiguousarray(acc, dtype=np.float64)
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    im = np.empty(P)
    eff = _load().orc_running_variance(_p(h), _p(a), _p(f), it, C, P, window_length, int(use_accepted), reg, _p(im))
    return im, eff


def chees_halton(n: int):
    out = np.zeros(n)
    _load().orc_chees_halton(n, _p(out))
    return out


def chees_steps(jitter: float, T: float, eps: float) -> int:
    return int(_load().orc_chees_steps(jitter, T, eps))


def chees_trajectory(T0, num, den, jitter, eps, lr=0.5, beta1=0.0, beta2=0.95, adam_eps=1e-8, decay=0.5,
                     max_steps=1000):
    """Trajectory times after successive ChEES updates from the given reduced terms."""
    num, den, jitter, eps = (np.ascontiguousarray(v, dtype=np.float64) for v in (num, den, jitter, eps))
    out = np.zeros(num.size)
    _load().orc_chees_trajectory(T0, lr, beta1, beta2, adam_eps, decay, max_steps, _p(num), _p(den), _p(jitter),
                                 _p(eps), num.size, _p(out))
    return out


class OracleFlow:
    """Binary64 restatement of the device flow with the same flat weights.  With
    bf16=True the matrix-core operands are rounded to bfloat16 as on the device."""

    def __init__(self, P: int, L: int, H: int, clamp: float, weights, bf16: bool = False):
        self.P, self.L, self.H, self.clamp, self.bf16 = int(P), int(L), int(H), float(clamp), bool(bf16)
        self.w = np.ascontiguousarray(weights, dtype=np.float32)
        n = int(_load().orc_flow_num_weights(self.P, self.L, self.H))
        if self.w.shape != (n,):
            raise ValueError(f"expected {n} weights, got {self.w.shape}")

    @classmethod
    def of(cls, flow, bf16: bool = False):
        return cls(flow.num_params, flow.num_layers, flow.hidden, flow.scale_clamp, flow.weights, bf16)

    @staticmethod
    def num_weights(P, L, H):
        return int(_load().orc_flow_num_weights(P, L, H))

    def forward(self, z, g_q=None):
        z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(-1, self.P))
        n = z.shape[0]
        q = np.empty_like(z)
        ld = np.empty(n)
        g = None if g_q is None else np.ascontiguousarray(np.asarray(g_q, dtype=np.float64).reshape(-1, self.P)).copy()
        _load().orc_flow_forward(self.P, self.L, self.H, self.clamp, int(self.bf16), _p(self.w), _p(z), n, _p(q),
                                 _p(ld), _p(g))
        return (q, ld) if g is None else (q, ld, g)

    def inverse(self, q):
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1, self.P))
        z = np.empty_like(q)
        ld = np.empty(q.shape[0])
        _load().orc_flow_inverse(self.P, self.L, self.H, self.clamp, int(self.bf16), _p(self.w), _p(q),
                                 q.shape[0], _p(z), _p(ld))
        return z, ld

    def loss_grad(self, q):
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(-1, self.P))
        g = np.empty(self.w.shape[0])
        loss = _load().orc_flow_loss_grad(self.P, self.L, self.H, self.clamp, int(self.bf16), _p(self.w), _p(q),
                                          q.shape[0], _p(g))
        return float(loss), g

    def elbo_grad(self, z, g_q, valid=None, row_weights=None):
        """Mean over valid rows of the weight gradient of g_q . f(z) + log|det df/dz|
        with g_q held fixed (the flow ADVI gradient, src/inference/vi.jl:807-826);
        with row_weights, the weighted sum over rows instead (the IWAE gradient)."""
        z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(-1, self.P))
        g_q = np.ascontiguousarray(np.asarray(g_q, dtype=np.float64).reshape(-1, self.P))
        v = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.int32))
        w = None if row_weights is None else np.ascontiguousarray(np.asarray(row_weights, dtype=np.float64))
        g = np.empty(self.w.shape[0])
        _load().orc_flow_elbo_grad(self.P, self.L, self.H, self.clamp, int(self.bf16), _p(self.w), _p(z), _p(g_q),
                                   _p(v), _p(w), z.shape[0], _p(g))
        return g


class OracleFullRank:
    """Binary64 restatement of the full-rank Gaussian guide with the device's flat
    weights: mu[P], log_scale[P], F[P x P] row-major (strict lower triangle used)."""

    def __init__(self, P: int, weights):
        self.P = int(P)
        self.w = np.ascontiguousarray(weights, dtype=np.float32)
        if self.w.shape != (2 * self.P + self.P * self.P,):
            raise ValueError(f"expected {2 * self.P + self.P * self.P} weights, got {self.w.shape}")

    @classmethod
    def of(cls, guide):
        return cls(guide.num_params, guide.weights)

    def _rows(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, self.P))

    def forward(self, z):
        z = self._rows(z)
        q = np.empty_like(z)
        ld = np.empty(z.shape[0])
        _load().orc_fullrank_forward(self.P, _p(self.w), _p(z), z.shape[0], _p(q), _p(ld))
        return q, ld

    def inverse(self, q):
        q = self._rows(q)
        z = np.empty_like(q)
        ld = np.empty(q.shape[0])
        _load().orc_fullrank_inverse(self.P, _p(self.w), _p(q), q.shape[0], _p(z), _p(ld))
        return z, ld

    def vjp(self, g_q):
        g = self._rows(g_q)
        out = np.empty_like(g)
        _load().orc_fullrank_vjp(self.P, _p(self.w), _p(g), g.shape[0], _p(out))
        return out

    def elbo_grad(self, z, g_q, valid=None, row_weights=None):
        """As OracleFlow.elbo_grad for the full-rank guide (src/inference/vi/guides.jl:
        164-199 and, with row_weights, 459-489)."""
        z = self._rows(z)
        g_q = self._rows(g_q)
        v = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.int32))
        w = None if row_weights is None else np.ascontiguousarray(np.asarray(row_weights, dtype=np.float64))
        g = np.empty(self.w.shape[0])
        _load().orc_fullrank_elbo_grad(self.P, _p(self.w), _p(z), _p(g_q), _p(v), _p(w), z.shape[0], _p(g))
        return g


def iwae_weights_and_bound(logw, valid, group_size):
    """The reference's _iwae_weights_and_bound! (src/inference/vi/guides.jl:383-425):
    per group of `group_size` particles the log-mean-exp of the valid log-weights,
    self-normalised weights within the group, all divided by the groups used.
    Returns (weights, bound)."""
    logw = np.asarray(logw, dtype=np.float64)
    valid = np.asarray(valid, dtype=bool)
    n = logw.size
    w = np.zeros(n)
    bound, used = 0.0, 0
    for g0 in range(0, n - n % group_size, group_size):
        idx = np.arange(g0, g0 + group_size)[valid[g0:g0 + group_size]]
        if idx.size == 0:
            continue
        m = logw[idx].max()
        lse = m + np.log(np.sum(np.exp(logw[idx] - m)))
        bound += lse - np.log(idx.size)
        w[idx] = np.exp(logw[idx] - lse)
        used += 1
    if used == 0:
        raise ValueError("batched_advi encountered only non-finite importance weights")
    return w / used, bound / used


def meads_slice_drift(u: float, delta: float) -> float:
    """The oracle's MEADS persistent-slice drift of slice value u by delta."""
    return float(_load().orc_meads_slice_drift(float(u), float(delta)))


def meads_max_eig(columns):
    """Trace-ratio largest-eigenvalue estimate of a P x N ensemble (columns = members)."""
    x = np.ascontiguousarray(np.asarray(columns, dtype=np.float64).T)
    n, P = x.shape
    return float(_load().orc_meads_max_eig(_p(x), n, P))
