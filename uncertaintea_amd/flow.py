"""Affine-coupling flow preconditioner (the C ABI's ut_flow).

Structure follows the reference's FlowGuide (src/inference/vi/guides.jl:1-21,
282-369): an elementwise affine base x0 = mu + exp(log_scale) z and coupling
layers with alternating parity masks, active x <- x exp(s) + t.  The reference's
conditioners are linear and trained host-side for ADVI; here they are MLPs of width
`hidden` trained on the GPU, either by the reference's objective (`fit_elbo`: the
reparameterised ELBO of batched_advi(...; guide=:flow), src/inference/vi.jl:700-870)
or by maximum likelihood on posterior draws (`fit`), and the trained flow
preconditions the lock-step HMC / ChEES / MEADS samplers (NeuTra):
`batched_hmc(model, ..., flow=flow)`.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib as L


def _pad_rows(a: np.ndarray, mult: int = 128) -> tuple[np.ndarray, int]:
    n = a.shape[0]
    m = -(-n // mult) * mult
    if m == n:
        return np.ascontiguousarray(a, dtype=np.float32), n
    out = np.zeros((m, a.shape[1]), dtype=np.float32)
    out[:n] = a
    return out, n


class AffineCouplingFlow:
    """q = f(z) on R^P with `num_layers` coupling layers and MLP width `hidden`."""

    def __init__(self, num_params: int, num_layers: int = 6, hidden: int = 256, scale_clamp: float = 3.0,
                 seed: int = 0):
        self.num_params = int(num_params)
        self.num_layers = int(num_layers)
        self.hidden = int(hidden)
        self.scale_clamp = float(scale_clamp)
        h = ct.c_void_p()
        L.check(L.lib().ut_flow_create(self.num_params, self.num_layers, self.hidden, self.scale_clamp,
                                       int(seed) & 0xFFFFFFFFFFFFFFFF, ct.byref(h)))
        self._adopt(h)

    def _adopt(self, h):
        self._h = h
        self._destroy = L.lib().ut_flow_destroy  # kept: module globals may be gone at exit
        n = ct.c_int64()
        L.check(L.lib().ut_flow_num_weights(h, ct.byref(n)))
        self.num_weights = int(n.value)
        self.loss_trace: np.ndarray | None = None
        self.elbo_trace: np.ndarray | None = None

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            self._destroy(self._h)
            self._h = None

    def layout(self):
        """Per layer: (num_passive, num_active, {W1, b1, W2, b2} offsets)."""
        out = []
        for layer in range(self.num_layers):
            npv, nav = ct.c_int32(), ct.c_int32()
            off = np.zeros(4, dtype=np.int64)
            L.check(L.lib().ut_flow_layer_layout(self._h, layer, ct.byref(npv), ct.byref(nav), L.ptr(off)))
            out.append((int(npv.value), int(nav.value), dict(zip(("W1", "b1", "W2", "b2"), off.tolist()))))
        return out

    @property
    def weights(self) -> np.ndarray:
        w = np.empty(self.num_weights, dtype=np.float32)
        L.check(L.lib().ut_flow_get_weights(self._h, L.ptr(w)))
        return w

    @weights.setter
    def weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.shape != (self.num_weights,):
            raise ValueError(f"DimensionMismatch: expected {self.num_weights} weights")
        L.check(L.lib().ut_flow_set_weights(self._h, L.ptr(w)))

    def set_base(self, mu, log_scale):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        ls = np.ascontiguousarray(log_scale, dtype=np.float64)
        if mu.shape != (self.num_params,) or ls.shape != (self.num_params,):
            raise ValueError("DimensionMismatch: mu and log_scale must have num_params entries")
        L.check(L.lib().ut_flow_set_base(self._h, L.ptr(mu), L.ptr(ls)))

    def fit(self, draws, steps: int = 2000, batch: int = 4096, learning_rate: float = 3e-4, seed: int = 0,
            init_base: bool = True) -> np.ndarray:
        """Maximum-likelihood training on `draws` (n x P, unconstrained space) with
        Adam on the GPU; returns (and keeps) the per-step minibatch loss."""
        d = np.ascontiguousarray(np.asarray(draws, dtype=np.float32).reshape(-1, self.num_params))
        if not np.all(np.isfinite(d)):
            d = d[np.all(np.isfinite(d), axis=1)]
        if init_base:
            mu = d.mean(axis=0, dtype=np.float64)
            sd = d.std(axis=0, dtype=np.float64)
            self.set_base(mu, np.log(np.maximum(sd, 1e-6)))
        trace = np.zeros(int(steps), dtype=np.float64)
        L.check(L.lib().ut_flow_train(self._h, L.ptr(d), d.shape[0], int(steps), int(batch), float(learning_rate),
                                      int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(trace)))
        self.loss_trace = trace
        return trace

    def fit_elbo(self, model, num_particles: int = 4096, num_iterations: int = 1000, learning_rate: float = 1e-3,
                 gradient_clip: float = float("inf"), seed: int = 0, keep_best: bool = False, beta1: float = 0.9,
                 beta2: float = 0.999, adam_epsilon: float = 1e-8, iwae_group: int = 1) -> dict:
        """Reparameterised-ELBO fit against `model` on the GPU, the reference's flow
        ADVI (src/inference/vi.jl:700-870): `num_iterations` Adam ascent steps, each
        on `num_particles` fresh particles (a multiple of 128); invalid particles
        (non-finite logjoint or gradient) carry no weight.  iwae_group = K > 1
        optimises the importance-weighted bound over groups of K particles.  Returns
        elbo_history (the optimised objective), standard_elbo_history,
        grad_norm_history, best_elbo and best_weights; the flow keeps its last
        weights unless keep_best."""
        from .device import DeviceBatchedWorkspace
        if model.num_params != self.num_params:
            raise ValueError("DimensionMismatch: the flow and the model have different parameter counts")
        ws = DeviceBatchedWorkspace(model, int(num_particles), np.float32)
        n = int(num_iterations)
        elbo = np.zeros(n)
        std_elbo = np.zeros(n)
        gnorm = np.zeros(n)
        best = np.empty(self.num_weights, dtype=np.float32)
        best_elbo = ct.c_double()
        L.check(L.lib().ut_flow_fit_elbo(ws.handle, self._h, n, float(learning_rate), float(beta1), float(beta2),
                                         float(adam_epsilon), float(gradient_clip), int(iwae_group),
                                         int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(elbo), L.ptr(std_elbo), L.ptr(gnorm),
                                         L.ptr(best), ct.byref(best_elbo)))
        if keep_best and n > 0:
            self.weights = best
        self.elbo_trace = elbo
        return {"elbo_history": elbo, "standard_elbo_history": std_elbo, "grad_norm_history": gnorm,
                "best_elbo": float(best_elbo.value), "best_weights": best}

    def elbo_grad(self, model, z, iwae_group: int = 1, return_weights: bool = False):
        """One ELBO (or, with iwae_group > 1, importance-weighted bound) evaluation at
        particles z (rows, a multiple of 128): (objective, mean weight gradient, model
        gradient at f(z), validity flags[, the particles' gradient weights])."""
        from .device import DeviceBatchedWorkspace
        zz = np.ascontiguousarray(np.asarray(z, dtype=np.float32).reshape(-1, self.num_params))
        ws = DeviceBatchedWorkspace(model, zz.shape[0], np.float32)
        elbo = ct.c_double()
        g = np.empty(self.num_weights, dtype=np.float32)
        gq = np.empty_like(zz)
        valid = np.empty(zz.shape[0], dtype=np.int32)
        rw = np.empty(zz.shape[0], dtype=np.float32)
        L.check(L.lib().ut_flow_elbo_grad(ws.handle, self._h, L.ptr(zz), int(iwae_group), ct.byref(elbo), L.ptr(g),
                                          L.ptr(gq), L.ptr(valid), L.ptr(rw)))
        out = (float(elbo.value), g, gq, valid.astype(bool))
        return out + (rw,) if return_weights else out

    def forward(self, z):
        """(q, log|det dq/dz|) for rows of z."""
        zp, n = _pad_rows(np.asarray(z).reshape(-1, self.num_params))
        q = np.empty_like(zp)
        ld = np.empty(zp.shape[0], dtype=np.float32)
        L.check(L.lib().ut_flow_forward(self._h, L.ptr(zp), zp.shape[0], L.ptr(q), L.ptr(ld)))
        return q[:n], ld[:n]

    def inverse(self, q):
        """(z, log|det dz/dq|) for rows of q."""
        qp, n = _pad_rows(np.asarray(q).reshape(-1, self.num_params))
        z = np.empty_like(qp)
        ld = np.empty(qp.shape[0], dtype=np.float32)
        L.check(L.lib().ut_flow_inverse(self._h, L.ptr(qp), qp.shape[0], L.ptr(z), L.ptr(ld)))
        return z[:n], ld[:n]

    def vjp(self, z, g_q):
        """(dq/dz)^T g_q + d log|det dq/dz| / dz for rows of z."""
        zp, n = _pad_rows(np.asarray(z).reshape(-1, self.num_params))
        gp, _ = _pad_rows(np.asarray(g_q).reshape(-1, self.num_params))
        gz = np.empty_like(zp)
        L.check(L.lib().ut_flow_vjp(self._h, L.ptr(zp), zp.shape[0], L.ptr(gp), L.ptr(gz)))
        return gz[:n]

    def kernel_times(self, rows: int, reps: int = 100):
        """Mean device microseconds per launch of the sampling path's flow forward and
        backward over `rows` device-resident rows (HIP events on the flow's stream)."""
        f, b = ct.c_double(), ct.c_double()
        L.check(L.lib().ut_flow_kernel_times(self._h, int(rows), int(reps), ct.byref(f), ct.byref(b)))
        return f.value, b.value

    def loss_grad(self, q):
        """(mean negative log-likelihood, its gradient over the flat weights); rows
        must be a multiple of 128."""
        qq = np.ascontiguousarray(np.asarray(q, dtype=np.float32).reshape(-1, self.num_params))
        g = np.empty(self.num_weights, dtype=np.float32)
        loss = ct.c_double()
        L.check(L.lib().ut_flow_loss_grad(self._h, L.ptr(qq), qq.shape[0], ct.byref(loss), L.ptr(g)))
        return float(loss.value), g


class FullRankGaussian(AffineCouplingFlow):
    """The full-rank Gaussian guide q = mu + exp(log_scale) z + F z, F strictly
    lower triangular (the reference's :fullrank guide, src/inference/vi/guides.jl:
    96-117,164-199), on the flow interface: the same weights, ELBO fit (`fit_elbo`),
    sampling (`forward`), `inverse` and `vjp`, and it preconditions the lock-step
    samplers like a flow (HMC in z space with the dense metric L L^T).  Weights:
    mu[P], log_scale[P], F[P x P] row-major, created at zero."""

    def __init__(self, num_params: int):
        self.num_params = int(num_params)
        self.num_layers = 0
        self.hidden = 128
        self.scale_clamp = 3.0
        h = ct.c_void_p()
        L.check(L.lib().ut_flow_create_fullrank(self.num_params, ct.byref(h)))
        self._adopt(h)

    def factor_of(self, weights=None) -> np.ndarray:
        """The strict lower triangle of the Cholesky factor (the reference's
        ADVIResult.scale_factor, src/inference/vi.jl:18-21), float64."""
        w = self.weights if weights is None else np.asarray(weights)
        P = self.num_params
        return np.tril(np.asarray(w[2 * P:], dtype=np.float64).reshape(P, P), -1)

    def cholesky(self, weights=None) -> np.ndarray:
        """L = diag(exp(log_scale)) + F, so the guide's covariance is L L^T."""
        w = self.weights if weights is None else np.asarray(weights)
        P = self.num_params
        return self.factor_of(w) + np.diag(np.exp(np.asarray(w[P:2 * P], dtype=np.float64)))

    def fit(self, *args, **kwargs):
        raise ValueError("FullRankGaussian is fitted by fit_elbo (the full-rank guide has no maximum-likelihood fit)")
