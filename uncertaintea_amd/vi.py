"""Variational inference on the device: batched_advi with the mean-field, the
full-rank and the flow guides (reference: src/inference/vi.jl,
src/inference/vi/guides.jl).

All three guides run through the C ABI's ut_flow_fit_elbo: the mean-field Gaussian
guide is the flow with no coupling layers (its elementwise affine base
x = mu + exp(log_scale) z), the full-rank guide adds a strictly lower factor F
(x = mu + exp(log_scale) z + F z, ut_flow_create_fullrank), the flow guide adds
`num_flow_layers` affine-coupling layers with MLP conditioners (linear in the
reference).  Each iteration draws the
particles on the device, scores them with the model's batched gradient engine,
drops invalid particles (non-finite logjoint or gradient; src/inference/vi.jl:
760-771), forms the reparameterised standard-ELBO gradient with the model
gradient held fixed, clips its flat norm and takes one Adam ascent step
(src/inference/vi.jl:205-227, 807-832).  The particles come from Philox on the
device, so results agree with the reference statistically, not draw for draw (as
the reference's own device ADVI path states, src/inference/vi.jl:367-368).
elbo="iwae" optimises the importance-weighted bound over groups of
`iwae_samples` particles (src/inference/vi/guides.jl:373-425).  Not built: the
low-rank guide.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .flow import AffineCouplingFlow, FullRankGaussian
from .models import TeaModel

_GUIDES = ("meanfield", "fullrank", "flow")


@dataclass
class ADVIResult:
    """Fields of the reference's ADVIResult (src/inference/vi.jl:3-30) that the
    device guides produce; `flow` / `best_flow` hold the guide (for the mean-field
    guide a flow without coupling layers, for the full-rank guide a
    FullRankGaussian) with the last / best weights; `scale_factor` /
    `best_scale_factor` are the full-rank guide's strict lower Cholesky triangle
    (None for the other guides)."""
    model: TeaModel
    location: np.ndarray
    log_scale: np.ndarray
    best_location: np.ndarray
    best_log_scale: np.ndarray
    elbo_history: np.ndarray
    gradient_norm_history: np.ndarray
    best_elbo: float
    num_particles: int
    learning_rate: float
    guide: str
    flow: AffineCouplingFlow
    best_flow: AffineCouplingFlow
    gradient_backend: str = "device"
    elbo: str = "standard"
    iwae_samples: int = 1
    standard_elbo_history: np.ndarray = field(default=None)
    scale_factor: np.ndarray | None = None
    best_scale_factor: np.ndarray | None = None

    def __post_init__(self):
        if self.standard_elbo_history is None:
            self.standard_elbo_history = self.elbo_history


def _initial_location(model: TeaModel, initial_params, seed: int) -> np.ndarray:
    """The reference's _resolve_unconstrained_point (src/inference/vi.jl:83-108): a
    prior draw when no point is given, else the point in unconstrained space."""
    P = model.num_params
    if initial_params is None:
        out = np.empty(P, dtype=np.float64)
        L.check(L.lib().ut_initial_params(model.plan, L.UT_INIT_PRIOR, int(seed) & 0xFFFFFFFFFFFFFFFF, 1, L.ptr(out)))
        return out
    p = np.asarray(initial_params, dtype=np.float64).reshape(-1)
    if p.shape != (P,):
        raise ValueError(f"DimensionMismatch: expected initial_params to have length {P} in unconstrained space, "
                         f"got {p.size}")
    return p


def _scale_vector(initial_log_scale, P: int) -> np.ndarray:
    """src/inference/vi.jl:111-119: a scalar fills every coordinate."""
    if np.isscalar(initial_log_scale):
        return np.full(P, float(initial_log_scale))
    v = np.asarray(initial_log_scale, dtype=np.float64).reshape(-1)
    if v.shape != (P,):
        raise ValueError(f"DimensionMismatch: expected initial_log_scale to have length {P}, got {v.size}")
    return v


def batched_advi(model: TeaModel, num_iterations: int, num_particles: int = 128, learning_rate: float = 0.05,
                 initial_params=None, initial_log_scale=-1.0, guide: str = "meanfield", num_flow_layers: int = 4,
                 flow_hidden: int = 256, elbo: str = "standard", iwae_samples: int = 0, beta1: float = 0.9,
                 beta2: float = 0.999, adam_epsilon: float = 1e-8, gradient_clip: float = float("inf"),
                 seed: int = 0) -> ADVIResult:
    """batched_advi(model; num_iterations, num_particles, learning_rate, guide, elbo,
    iwae_samples, ...) (src/inference/vi.jl:300-440) on the device.  `num_particles` must be a multiple
    of 128 (the particles are the rows of one batched gradient; the reference's
    default is 32).  Argument checks and messages follow the reference's."""
    if guide not in _GUIDES:
        if guide == "lowrank":
            raise ValueError("batched_advi guide=lowrank is CPU-only in the reference and not built here; the "
                             "device path supports meanfield, fullrank and flow")
        raise ValueError(f"batched_advi guide must be :meanfield, :fullrank, :lowrank, or :flow, got {guide}")
    if elbo not in ("standard", "iwae"):
        raise ValueError(f"batched_advi elbo must be :standard or :iwae, got :{elbo}")
    if num_iterations <= 0:
        raise ValueError("batched_advi requires num_iterations > 0")
    if num_particles <= 0:
        raise ValueError("batched_advi requires num_particles > 0")
    if num_particles % 128 != 0:
        raise ValueError("DimensionMismatch: the device batched_advi needs num_particles a multiple of 128")
    group = (int(iwae_samples) if iwae_samples > 0 else int(num_particles)) if elbo == "iwae" else 1
    if elbo == "iwae":
        if not 1 <= group <= num_particles:
            raise ValueError("batched_advi requires 1 <= iwae_samples <= num_particles")
        if num_particles % group != 0:
            raise ValueError("batched_advi requires iwae_samples to divide num_particles evenly")
    if not learning_rate > 0:
        raise ValueError("batched_advi requires learning_rate > 0")
    if not 0 <= beta1 < 1:
        raise ValueError("batched_advi requires 0 <= beta1 < 1")
    if not 0 <= beta2 < 1:
        raise ValueError("batched_advi requires 0 <= beta2 < 1")
    if not adam_epsilon > 0:
        raise ValueError("batched_advi requires adam_epsilon > 0")
    if not gradient_clip > 0:
        raise ValueError("batched_advi requires gradient_clip > 0")
    P = model.num_params
    location = _initial_location(model, initial_params, seed)
    log_scale = _scale_vector(initial_log_scale, P)
    layers = int(num_flow_layers) if guide == "flow" else 0

    def make():
        return FullRankGaussian(P) if guide == "fullrank" else AffineCouplingFlow(P, layers, flow_hidden, seed=seed)

    flow = make()
    flow.set_base(location, log_scale)
    out = flow.fit_elbo(model, num_particles=num_particles, num_iterations=num_iterations,
                        learning_rate=learning_rate, gradient_clip=gradient_clip, seed=seed, beta1=beta1,
                        beta2=beta2, adam_epsilon=adam_epsilon, iwae_group=group)
    w = flow.weights
    best = out["best_weights"]
    best_flow = make()
    best_flow.weights = best
    factors = (flow.factor_of(w), flow.factor_of(best)) if guide == "fullrank" else (None, None)
    return ADVIResult(model=model, location=w[:P].astype(np.float64), log_scale=w[P:2 * P].astype(np.float64),
                      best_location=best[:P].astype(np.float64), best_log_scale=best[P:2 * P].astype(np.float64),
                      elbo_history=out["elbo_history"], gradient_norm_history=out["grad_norm_history"],
                      best_elbo=out["best_elbo"], num_particles=int(num_particles),
                      learning_rate=float(learning_rate), guide=guide, flow=flow, best_flow=best_flow, elbo=elbo,
                      iwae_samples=group, standard_elbo_history=out["standard_elbo_history"],
                      scale_factor=factors[0], best_scale_factor=factors[1])


def variational_samples(result: ADVIResult, num_samples: int, space: str = "constrained", use_best: bool = True,
                        seed: int = 0) -> np.ndarray:
    """P x num_samples draws from the fitted guide (src/inference/vi.jl:251-300),
    pushed through the guide on the device."""
    if num_samples <= 0:
        raise ValueError("variational_samples requires num_samples > 0")
    if space not in ("constrained", "unconstrained"):
        raise ValueError("variational space must be :constrained or :unconstrained")
    flow = result.best_flow if use_best else result.flow
    z = np.random.default_rng(seed).standard_normal((num_samples, result.model.num_params)).astype(np.float32)
    q, _ = flow.forward(z)
    q = q.astype(np.float64)
    if space == "constrained":
        q = result.model.constrain(q)
    return np.ascontiguousarray(q.T)


def variational_mean(result: ADVIResult, space: str = "constrained", use_best: bool = True) -> np.ndarray:
    """src/inference/vi.jl:229-249: the location (mapped to the constrained space)
    for the mean-field guide, a fixed-seed 20000-draw estimate for the flow."""
    if space not in ("constrained", "unconstrained"):
        raise ValueError("variational space must be :constrained or :unconstrained")
    if result.guide == "flow":
        return variational_samples(result, 20000, space=space, use_best=use_best).mean(axis=1)
    loc = result.best_location if use_best else result.location
    return loc.copy() if space == "unconstrained" else result.model.constrain(loc)


def variational_covariance(result: ADVIResult, use_best: bool = True) -> np.ndarray:
    """src/inference/vi.jl:129-150: diagonal exp(2 log_scale) for the mean-field
    guide, L L^T for the full-rank guide, a fixed-seed 20000-draw estimate for the
    flow (unconstrained space)."""
    if result.guide == "flow":
        s = variational_samples(result, 20000, space="unconstrained", use_best=use_best)
        return np.cov(s)
    ls = result.best_log_scale if use_best else result.log_scale
    if result.guide == "fullrank":
        chol = (result.best_scale_factor if use_best else result.scale_factor) + np.diag(np.exp(ls))
        return chol @ chol.T
    return np.diag(np.exp(2.0 * ls))
