"""Batched samplers on the gfx950 backend: `batched_nuts`, `batched_hmc` and
`batched_chees`.

Signatures and defaults follow the reference (src/inference/api_batched_nuts.jl:22-51,
src/inference/api_batched_hmc.jl:16-39, src/inference/api_batched_chees.jl:277-306).  The device path is always the pooled-mass /
per-chain-step adaptation the reference uses on devices
(api_batched_nuts.jl:761-980, src/device/hmc_kernels.jl:306-723); all randomness is
drawn on the device from Philox coordinates derived from `seed`, so a run is a pure
function of its inputs.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

from . import _lib as L
from . import diagnostics as D
from .device import DeviceBatchedWorkspace
from .models import TeaModel


class ProgressInfo(NamedTuple):
    """What a progress callback receives (the reference's NamedTuple,
    src/inference/api_single_chain.jl:1-11)."""

    phase: str          # "warmup" or "sample"
    iteration: int      # 1-based within the phase
    total: int          # iterations in the phase
    step_size: float    # mean over chains of the per-chain steps; NaN for MEADS
    divergences: int    # cumulative divergent transitions of the phase (warmup MEADS: 0, untracked)


class _Callback:
    """Holds the C function pointer for one sampler call and re-raises the first
    exception the Python callback raised once the call has returned."""

    def __init__(self, fn):
        self.fn = fn
        self.error = None

        def trampoline(user, phase, iteration, total, step_size, divergences):
            if self.error is not None:
                return
            try:
                self.fn(ProgressInfo("warmup" if phase == L.UT_PHASE_WARMUP else "sample", int(iteration),
                                     int(total), float(step_size), int(divergences)))
            except BaseException as e:  # noqa: BLE001 -- re-raised after the C call returns
                self.error = e

        self.c_fn = L.PROGRESS_FN(trampoline)

    def install(self, opts, every: int) -> None:
        opts.callback = ct.cast(self.c_fn, ct.c_void_p)
        opts.callback_every = int(every)

    def reraise(self) -> None:
        if self.error is not None:
            raise self.error


@dataclass
class HMCChains:
    """Stacked per-chain results (reference HMCChain fields, data_types.jl:55-97).

    Arrays are chain-major: `unconstrained_samples[c, s, p]`, `logjoint_values[c, s]`.
    """

    sampler: str
    model: TeaModel
    unconstrained_samples: np.ndarray
    logjoint_values: np.ndarray
    acceptance_stats: np.ndarray
    energies: np.ndarray
    energy_errors: np.ndarray
    accepted: np.ndarray
    divergent: np.ndarray
    tree_depths: np.ndarray
    integration_steps: np.ndarray
    step_size: np.ndarray
    mass_matrix: np.ndarray
    max_tree_depth: int
    num_leapfrog_steps: int
    target_accept: float
    warmup_seconds: float
    sampling_seconds: float
    sampling_gradient_evals: int
    sampling_kernel_seconds: float = 0.0
    sampling_launches: int = 0
    trajectory_trace: np.ndarray | None = None
    _constrained: np.ndarray | None = field(default=None, repr=False)

    @property
    def num_chains(self) -> int:
        return self.unconstrained_samples.shape[0]

    @property
    def num_samples(self) -> int:
        return self.unconstrained_samples.shape[1]

    @property
    def constrained_samples(self) -> np.ndarray:
        if self._constrained is None:
            self._constrained = self.model.constrain(self.unconstrained_samples)
        return self._constrained

    def draws(self, space: str = "constrained") -> np.ndarray:
        if space == "constrained":
            return self.constrained_samples
        if space == "unconstrained":
            return self.unconstrained_samples
        raise ValueError(f"space must be 'constrained' or 'unconstrained', got {space!r}")

    def rhat(self, space: str = "constrained", method: str = "rank") -> np.ndarray:
        return D.rhat(self.draws(space), method=method)

    def ess(self, space: str = "constrained") -> np.ndarray:
        return D.ess(self.draws(space))

    def ess_bulk(self, space: str = "constrained") -> np.ndarray:
        return D.ess_bulk(self.draws(space))

    def divergence_rate(self) -> float:
        return float(self.divergent.mean()) if self.divergent.size else 0.0


def _opts(num_samples, num_warmup, step_size, target_accept, divergence_threshold, mass_matrix_regularization,
          mass_matrix_min_samples, adapt_step_size, adapt_mass_matrix, find_reasonable_step_size, init_strategy,
          init_max_retries, seed, max_tree_depth=10, num_leapfrog_steps=10) -> L.SamplerOpts:
    o = L.SamplerOpts()
    L.lib().ut_sampler_opts_default(ct.byref(o))
    if init_strategy not in ("prior", "uniform"):
        raise ValueError(f"init_strategy must be 'prior' or 'uniform', got {init_strategy!r}")
    o.num_samples = int(num_samples)
    o.num_warmup = int(num_warmup)
    o.max_tree_depth = int(max_tree_depth)
    o.num_leapfrog_steps = int(num_leapfrog_steps)
    o.step_size = float(step_size)
    o.target_accept = float(target_accept)
    o.divergence_threshold = float(divergence_threshold)
    o.mass_matrix_regularization = float(mass_matrix_regularization)
    o.mass_matrix_min_samples = int(mass_matrix_min_samples)
    o.adapt_step_size = int(bool(adapt_step_size))
    o.adapt_mass_matrix = int(bool(adapt_mass_matrix))
    o.find_reasonable_step_size = int(bool(find_reasonable_step_size))
    o.init_strategy = L.UT_INIT_PRIOR if init_strategy == "prior" else L.UT_INIT_UNIFORM
    o.init_max_retries = int(init_max_retries)
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return o


def _run(kind: str, model: TeaModel, num_chains: int, opts, initial_params,
         max_tree_depth: int, num_leapfrog_steps: int, target_accept: float, flow=None, callback=None,
         callback_every: int = 10, waves_per_chain: int = 0) -> HMCChains:
    if num_chains <= 0:
        raise ValueError("num_chains must be positive")
    cb = None
    if callback is not None:
        if not callable(callback):
            raise ValueError("callback must be callable (it receives a ProgressInfo)")
        cb = _Callback(callback)
        cb.install(opts.base if kind == "chees" else opts, callback_every)
    ws = DeviceBatchedWorkspace(model, num_chains, np.float32, waves_per_chain=waves_per_chain)
    if flow is not None:
        if flow.num_params != model.num_params:
            raise ValueError("DimensionMismatch: the flow and the model have different parameter counts")
        L.check(L.lib().ut_workspace_set_flow(ws.handle, flow.handle))
    base = opts.base if kind == "chees" else opts
    P, C, S = model.num_params, int(num_chains), int(base.num_samples)
    init = None
    if initial_params is not None:
        init = np.asarray(initial_params, dtype=np.float64)
        if init.shape == (P,):
            init = np.repeat(init[:, None], C, axis=1)
        if init.shape != (P, C):
            raise ValueError(f"initial_params must be P x C = {P} x {C}")
        init = np.ascontiguousarray(init.T)  # chain-major
        base.initial_params = L.dptr(init, ct.c_double)
    pe = L.pinned_empty  # page-locked: the sampler copies results out while it computes
    samples = pe((C, S, P), np.float32)
    lj, acc, en, err = pe((C, S), np.float64), pe((C, S), np.float64), pe((C, S), np.float64), pe((C, S), np.float64)
    accepted, div = pe((C, S), np.uint8), pe((C, S), np.uint8)
    depth, steps = pe((C, S), np.int32), pe((C, S), np.int32)
    step = np.empty(C); im = np.empty(P)
    out = L.ChainOut()
    out.samples = samples.ctypes.data_as(ct.c_void_p)
    out.logjoint = L.dptr(lj, ct.c_double)
    out.accept_stat = L.dptr(acc, ct.c_double)
    out.energy = L.dptr(en, ct.c_double)
    out.energy_error = L.dptr(err, ct.c_double)
    out.accepted = L.dptr(accepted, ct.c_uint8)
    out.divergent = L.dptr(div, ct.c_uint8)
    out.tree_depth = L.dptr(depth, ct.c_int32)
    out.integration_steps = L.dptr(steps, ct.c_int32)
    out.step_size = L.dptr(step, ct.c_double)
    out.inverse_mass = L.dptr(im, ct.c_double)
    trace = None
    if kind == "meads":
        trace = np.zeros((int(base.num_warmup) + int(base.num_samples), 3), dtype=np.float64)
        L.check(L.lib().ut_meads_run(ws.handle, ct.byref(opts), ct.byref(out), L.ptr(trace)))
    elif kind == "chees":
        trace = np.zeros(int(base.num_warmup), dtype=np.float64)
        L.check(L.lib().ut_chees_run(ws.handle, ct.byref(opts), ct.byref(out), L.ptr(trace)))
    else:
        fn = {"nuts": L.lib().ut_nuts_run, "nuts_masked": L.lib().ut_nuts_masked_run, "hmc": L.lib().ut_hmc_run,
              "hmc_shared": L.lib().ut_hmc_shared_run}[kind]
        L.check(fn(ws.handle, ct.byref(opts), ct.byref(out)))
    if cb is not None:
        cb.reraise()
    kind = {"hmc_shared": "hmc", "nuts_masked": "nuts"}.get(kind, kind)
    return HMCChains(kind, model, samples, lj, acc, en, err, accepted.view(bool), div.view(bool), depth,
                     steps, step, im, max_tree_depth, num_leapfrog_steps, target_accept, out.warmup_seconds,
                     out.sampling_seconds, int(out.sampling_gradient_evals), out.sampling_kernel_seconds,
                     int(out.sampling_launches), trace)


def batched_nuts(model: TeaModel, args=(), constraints=None, *, num_chains: int, num_samples: int,
                 num_warmup: int = 0, step_size: float = 0.1, max_tree_depth: int = 10, initial_params=None,
                 init_strategy: str = "prior", init_max_retries: int = 100, target_accept: float = 0.8,
                 adapt_step_size: bool = True, adapt_mass_matrix: bool = True, per_chain_adaptation=None,
                 find_reasonable_step_size: bool = False, divergence_threshold: float = 1000.0,
                 mass_matrix_regularization: float = 1e-3, mass_matrix_min_samples: int = 10,
                 tree_strategy: str = "persistent", backend: str = "hip", precision=np.float32,
                 seed: int = 0, callback=None, callback_every: int = 10, flow=None,
                 waves_per_chain: int = 0) -> HMCChains:
    """Many NUTS chains.  tree_strategy='persistent': each chain builds its whole tree
    in one kernel (wavefront-per-chain plans).  'masked': every chain advances one
    leaf per step with the gradient of all chains in one batched call -- any plan,
    the matrix-core GLM and the GP included, and a `flow` preconditioner (NUTS in its
    base space); the same draws as 'persistent' where both run.  waves_per_chain: 0
    chooses the chain layout automatically (few chains over many logistic rows: a
    workgroup of 2, 4 or 8 wavefronts per chain), 1/2/4/8 forces it."""
    model = model.condition(args, constraints)
    if tree_strategy not in ("persistent", "masked"):
        raise ValueError(f"batched_nuts tree_strategy must be 'persistent' or 'masked', got {tree_strategy!r}")
    if flow is not None and tree_strategy != "masked":
        raise ValueError("a flow preconditions NUTS with tree_strategy='masked' (leaves in lock step)")
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    if per_chain_adaptation is False:
        raise ValueError("the device path adapts a per-chain step with a pooled mass (per_chain_adaptation=True)")
    if np.dtype(precision) != np.float32:
        raise ValueError("the device samplers run in float32 with float64 energies")
    o = _opts(num_samples, num_warmup, step_size, target_accept, divergence_threshold, mass_matrix_regularization,
              mass_matrix_min_samples, adapt_step_size, adapt_mass_matrix, find_reasonable_step_size, init_strategy,
              init_max_retries, seed, max_tree_depth=max_tree_depth)
    return _run("nuts_masked" if tree_strategy == "masked" else "nuts", model, num_chains, o, initial_params,
                max_tree_depth, 0, target_accept, flow=flow, callback=callback, callback_every=callback_every,
                waves_per_chain=waves_per_chain)


def batched_hmc(model: TeaModel, args=(), constraints=None, *, num_chains: int, num_samples: int,
                num_warmup: int = 0, step_size: float = 0.1, num_leapfrog_steps: int = 10, initial_params=None,
                init_strategy: str = "prior", init_max_retries: int = 100, target_accept: float = 0.8,
                adapt_step_size: bool = True, adapt_mass_matrix: bool = True, per_chain_adaptation=None,
                find_reasonable_step_size: bool = False, divergence_threshold: float = 1000.0,
                mass_matrix_regularization: float = 1e-3, mass_matrix_min_samples: int = 10,
                backend: str = "hip", precision=np.float32, seed: int = 0, flow=None, callback=None,
                callback_every: int = 10, waves_per_chain: int = 0) -> HMCChains:
    """Many fixed-length HMC chains, device-resident (momentum, accept and adaptation on the GPU).

    per_chain_adaptation (default True, as the reference): each chain adapts its own
    step size inside the persistent kernel.  False: the reference's shared mode -- all
    chains in lock step with one dual-averaged step size.  `flow` (an
    AffineCouplingFlow) preconditions the shared mode: chains run in the flow's base
    space and draws are returned mapped through the flow.  waves_per_chain: the chain
    layout of the per-chain mode, as for batched_nuts."""
    model = model.condition(args, constraints)
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    if np.dtype(precision) != np.float32:
        raise ValueError("the device samplers run in float32 with float64 energies")
    if flow is not None and per_chain_adaptation:
        raise ValueError("a flow preconditions the shared-adaptation (lock-step) mode; pass per_chain_adaptation=False")
    shared = per_chain_adaptation is False or flow is not None
    o = _opts(num_samples, num_warmup, step_size, target_accept, divergence_threshold, mass_matrix_regularization,
              mass_matrix_min_samples, adapt_step_size, adapt_mass_matrix, find_reasonable_step_size, init_strategy,
              init_max_retries, seed, num_leapfrog_steps=num_leapfrog_steps)
    return _run("hmc_shared" if shared else "hmc", model, num_chains, o, initial_params, 0, num_leapfrog_steps,
                target_accept, flow=flow, callback=callback, callback_every=callback_every,
                waves_per_chain=waves_per_chain)


def batched_chees(model: TeaModel, args=(), constraints=None, *, num_chains: int, num_samples: int,
                  num_warmup: int = 0, step_size: float = 0.1, num_leapfrog_steps: int = 10,
                  jitter_amount: float = 1.0, adapt_trajectory_length: bool = True, max_leapfrog_steps: int = 1000,
                  trajectory_decay_rate: float = 0.5, trajectory_learning_rate: float = 0.5,
                  trajectory_adam_beta1: float = 0.0, trajectory_adam_beta2: float = 0.95,
                  trajectory_adam_epsilon: float = 1e-8, initial_params=None, init_strategy: str = "prior",
                  init_max_retries: int = 100, target_accept: float = 0.651, adapt_step_size: bool = True,
                  adapt_mass_matrix: bool = True, find_reasonable_step_size: bool = False,
                  divergence_threshold: float = 1000.0, mass_matrix_regularization: float = 1e-3,
                  mass_matrix_min_samples: int = 10, backend: str = "hip", precision=np.float32,
                  seed: int = 0, forward_velocity: bool = False, flow=None, callback=None,
                  callback_every: int = 10) -> HMCChains:
    """ChEES-HMC: every chain integrates the same Halton-jittered number of leapfrog
    steps with a shared step size; warmup adapts the step (dual averaging to the
    harmonic-mean acceptance), a pooled diagonal mass and the trajectory time T
    (Adam ascent on the cross-chain ChEES criterion).  `trajectory_trace` holds T
    after each warmup iteration.

    `forward_velocity` (not a reference option, default off for parity): the
    reference feeds the criterion the proposal momentum after its final flip, so
    its Adam ascent shortens T; True uses the velocity in the direction of
    integration, under which T grows where longer trajectories pay off."""
    model = model.condition(args, constraints)
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    if np.dtype(precision) != np.float32:
        raise ValueError("the device samplers run in float32 with float64 energies")
    co = L.CheesOpts()
    L.lib().ut_chees_opts_default(ct.byref(co))
    co.base = _opts(num_samples, num_warmup, step_size, target_accept, divergence_threshold,
                    mass_matrix_regularization, mass_matrix_min_samples, adapt_step_size, adapt_mass_matrix,
                    find_reasonable_step_size, init_strategy, init_max_retries, seed,
                    num_leapfrog_steps=num_leapfrog_steps)
    co.jitter_amount = float(jitter_amount)
    co.adapt_trajectory_length = int(bool(adapt_trajectory_length))
    co.max_leapfrog_steps = int(max_leapfrog_steps)
    co.trajectory_decay_rate = float(trajectory_decay_rate)
    co.trajectory_learning_rate = float(trajectory_learning_rate)
    co.trajectory_adam_beta1 = float(trajectory_adam_beta1)
    co.trajectory_adam_beta2 = float(trajectory_adam_beta2)
    co.trajectory_adam_epsilon = float(trajectory_adam_epsilon)
    co.forward_velocity = int(bool(forward_velocity))
    return _run("chees", model, num_chains, co, initial_params, 0, num_leapfrog_steps, target_accept, flow=flow,
                callback=callback, callback_every=callback_every)


def batched_meads(model: TeaModel, args=(), constraints=None, *, num_chains: int, num_samples: int,
                  num_warmup: int = 0, num_folds: int = 4, initial_params=None, init_strategy: str = "prior",
                  init_max_retries: int = 100, divergence_threshold: float = 1000.0, backend: str = "hip",
                  precision=np.float32, seed: int = 0, flow=None, callback=None,
                  callback_every: int = 10) -> HMCChains:
    """MEADS: generalized HMC, one leapfrog step per iteration with persistent
    momentum and a persistent-slice accept; step, damping, drift and a diagonal
    preconditioner from the neighbouring fold every iteration (no adaptation phase,
    `num_warmup` only discards draws).  `trajectory_trace` holds per iteration
    (mean step, mean damping, largest std) -- the reference's `_adapt_trace`.
    The reference runs MEADS on the host only; here the fold reductions and the
    transitions run on the device."""
    model = model.condition(args, constraints)
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    if np.dtype(precision) != np.float32:
        raise ValueError("the device samplers run in float32 with float64 energies")
    if init_strategy not in ("prior", "uniform"):
        raise ValueError(f"batched_meads init_strategy must be 'prior' or 'uniform', got {init_strategy!r}")
    mo = L.MeadsOpts()
    L.lib().ut_meads_opts_default(ct.byref(mo))
    mo.num_samples = int(num_samples)
    mo.num_warmup = int(num_warmup)
    mo.num_folds = int(num_folds)
    mo.init_strategy = L.UT_INIT_PRIOR if init_strategy == "prior" else L.UT_INIT_UNIFORM
    mo.init_max_retries = int(init_max_retries)
    mo.divergence_threshold = float(divergence_threshold)
    mo.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return _run("meads", model, num_chains, mo, initial_params, 0, 1, float("nan"), flow=flow, callback=callback,
                callback_every=callback_every)
