"""Result interchange: ArviZ-style dictionaries and the cross-PPL run files
(reference: src/inference/export.jl, bench/crossppl/julia/run.jl:255-285).

The device samplers return chain-major arrays (`HMCChains.unconstrained_samples[c, s, p]`);
these helpers lay them out the way the reference's exporters do, so downstream
tools (ArviZ's `from_dict`, the cross-PPL analysis that reads `draws` from an NPZ
file and the run's settings from a JSON sidecar) consume device runs directly.
"""
from __future__ import annotations

import ctypes as ct
import json
import os
import re

import numpy as np

from . import _lib as L
from .inference import HMCChains

_LAYOUTS = ("draw_chain", "chain_draw")


def parameter_names(chains: HMCChains, space: str = "constrained") -> list[str]:
    """Display names of the parameters, one per column (export.jl:42-51)."""
    if space not in ("constrained", "unconstrained"):
        raise ValueError(f"space must be 'constrained' or 'unconstrained', got {space!r}")
    return list(chains.model.param_names)


def posterior_array(chains: HMCChains, space: str = "constrained") -> np.ndarray:
    """(num_samples, num_chains, num_params) draws, float64 (export.jl:53-70)."""
    d = chains.draws(space)
    return np.ascontiguousarray(np.transpose(np.asarray(d, dtype=np.float64), (1, 0, 2)))


def _layout(m: np.ndarray, layout: str) -> np.ndarray:
    """Matrices are built (num_samples, num_chains[, k]); chain_draw swaps the first
    two axes (export.jl:84-91)."""
    return np.ascontiguousarray(np.swapaxes(m, 0, 1)) if layout == "chain_draw" else m


def _per_draw(chains: HMCChains, a) -> np.ndarray:
    """A per-chain or per-draw statistic as (num_samples, num_chains)."""
    C, S = chains.num_chains, chains.num_samples
    a = np.asarray(a)
    if a.ndim == 0:
        a = np.full((C, S), a)
    elif a.shape == (C,):
        a = np.repeat(a[:, None], S, axis=1)
    return np.ascontiguousarray(a.reshape(C, S).T)


def pointwise_loglikelihood(chains: HMCChains) -> np.ndarray:
    """(num_chains * num_samples, num_observations) log-likelihood of every observed
    site at every draw, rows chain-major (chain 1's draws, then chain 2's, ...): the
    reference's pointwise_loglikelihood(model, args, constraints, chains)
    (src/inference/result_interface.jl:384-410), computed on the device
    (ut_pointwise_loglik).  Observations: the logistic model's y in data order, the
    hierarchical model's y group by group; the funnel has none."""
    model = chains.model
    n = ct.c_int64()
    L.check(L.lib().ut_pointwise_loglik_size(model.plan, ct.byref(n)))
    d = np.ascontiguousarray(np.asarray(chains.unconstrained_samples, dtype=np.float64).reshape(-1, model.num_params))
    out = np.empty((d.shape[0], int(n.value)), dtype=np.float64)
    if out.size:
        L.check(L.lib().ut_pointwise_loglik(model.plan, L.ptr(d), d.shape[0], L.ptr(out)))
    return out


def to_arviz_dict(chains: HMCChains, layout: str = "draw_chain", flatten_vectors: bool = True,
                  log_likelihood: bool = False) -> dict:
    """ArviZ-style nested dictionary (export.jl:187-272): "posterior" (one array per
    constrained parameter, or per vector variable with a trailing component axis
    when flatten_vectors is False, with "coords" / "dims" in the from_dict
    convention), "sample_stats" (diverging, energy, tree_depth, acceptance_rate,
    lp, step_size, n_steps) and "attrs".  layout "chain_draw" gives the
    (chain, draw) order Python ArviZ expects.  log_likelihood=True adds the
    "log_likelihood" group of the reference's four-argument form (export.jl:276-338):
    one "y[k]" matrix per observation, or with flatten_vectors=False one "y" array
    with a trailing observation axis (dims "y_dim_0", coords 1..N)."""
    if layout not in _LAYOUTS:
        raise ValueError(f"to_arviz_dict layout must be :draw_chain or :chain_draw, got :{layout}")
    if chains.num_chains == 0:
        raise ValueError("to_arviz_dict requires at least one chain")
    draws = posterior_array(chains, "constrained")
    names = parameter_names(chains)
    posterior: dict = {}
    coords: dict = {}
    dims: dict = {}
    if flatten_vectors:
        for p, name in enumerate(names):
            posterior[name] = _layout(draws[:, :, p], layout)
    else:
        groups: dict[str, list[int]] = {}
        for p, name in enumerate(names):
            m = re.fullmatch(r"(.+)\[(\d+)\]", name)
            groups.setdefault(m.group(1) if m else name, []).append(p)
        for name, idx in groups.items():
            if len(idx) == 1 and names[idx[0]] == name:
                posterior[name] = _layout(draws[:, :, idx[0]], layout)
            else:
                posterior[name] = _layout(draws[:, :, idx], layout)
                dim = f"{name}_dim_0"
                dims[name] = [dim]
                coords[dim] = list(range(1, len(idx) + 1))
    stats = {
        "diverging": _per_draw(chains, chains.divergent).astype(bool),
        "energy": _per_draw(chains, chains.energies).astype(np.float64),
        "tree_depth": _per_draw(chains, chains.tree_depths).astype(np.float64),
        "acceptance_rate": _per_draw(chains, chains.acceptance_stats).astype(np.float64),
        "lp": _per_draw(chains, chains.logjoint_values).astype(np.float64),
        "step_size": _per_draw(chains, chains.step_size).astype(np.float64),
        "n_steps": _per_draw(chains, chains.integration_steps).astype(np.int64),
    }
    out = {"posterior": posterior,
           "sample_stats": {k: _layout(v, layout) for k, v in stats.items()},
           "attrs": {"layout": layout, "inference_library": "UncertainTea",
                     "backend": "uncertaintea_amd (MI355X HIP)", "sampler": chains.sampler}}
    if log_likelihood:
        ll = pointwise_loglikelihood(chains)
        C, S = chains.num_chains, chains.num_samples
        cube = np.ascontiguousarray(np.transpose(ll.reshape(C, S, -1), (1, 0, 2)))  # (S, C, N)
        group: dict = {}
        if flatten_vectors:
            for k in range(cube.shape[2]):
                group[f"y[{k + 1}]"] = _layout(cube[:, :, k], layout)
        elif cube.shape[2] > 0:
            group["y"] = _layout(cube, layout)
            dims["y"] = ["y_dim_0"]
            coords["y_dim_0"] = list(range(1, cube.shape[2] + 1))
        out["log_likelihood"] = group
    if not flatten_vectors:
        out["coords"] = coords
        out["dims"] = dims
    return out


def write_crossppl_run(chains: HMCChains, directory: str, model_name: str, variant: str, *, num_warmup: int,
                       seed: int, rep: int = 1, ttfx_s: float | None = None, sampler: dict | None = None) -> str:
    """One run in the cross-PPL layout (bench/crossppl/julia/run.jl:255-285):
    `<model>__uncertaintea-<variant>__chains<K>__seed<S>__rep<R>.npz` holding
    "draws" (chains, draws, params) float64, and a .json sidecar with the
    parameter names, sizes, seed, timings and sampler settings.  Returns the stem's
    path without extension."""
    os.makedirs(directory, exist_ok=True)
    stem = os.path.join(directory, f"{model_name}__uncertaintea-{variant}__chains{chains.num_chains}__seed{seed}"
                                   f"__rep{rep}")
    draws = np.ascontiguousarray(np.asarray(chains.constrained_samples, dtype=np.float64))
    np.savez(stem + ".npz", draws=draws)
    total = chains.warmup_seconds + chains.sampling_seconds
    meta = {
        "model": model_name, "framework": "uncertaintea", "variant": variant,
        "params": parameter_names(chains), "num_chains": chains.num_chains, "num_samples": chains.num_samples,
        "num_warmup": int(num_warmup), "seed": int(seed), "rep": int(rep),
        "timings": {"ttfx_s": ttfx_s, "warmup_s": chains.warmup_seconds, "total_s": total,
                    "sampling_s": chains.sampling_seconds},
        "sampler": sampler if sampler is not None else {"kind": chains.sampler},
    }
    with open(stem + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    return stem
