"""`UncertainTea.Device` API on the gfx950 backend.

Mirrors the reference's device density entry points (src/device/api.jl:34-124):
parameters are the P x C unconstrained matrix (column c = chain c), the output is
the unconstrained logjoint (log density plus transform log-abs-det) per chain, and
its gradient as a P x C matrix.  Every call is synchronous.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib as L
from .models import TeaModel

_PREC = {np.dtype(np.float32): L.UT_FLOAT32, np.dtype(np.float64): L.UT_FLOAT64}


def _precision(precision) -> np.dtype:
    dt = np.dtype(precision)
    if dt not in _PREC:
        raise ValueError(f"precision must be float32 or float64, got {dt}")
    return dt


class DeviceBatchedWorkspace:
    """Device buffers + one HIP stream for `num_chains` chains of one model
    (reference DeviceBatchedWorkspace, src/device/workspace.jl:71-162)."""

    def __init__(self, model: TeaModel, num_chains: int, precision=np.float32, waves_per_chain: int = 0):
        self.model = model
        self.num_chains = int(num_chains)
        self.dtype = _precision(precision)
        self.num_params = model.num_params
        h = ct.c_void_p()
        L.check(L.lib().ut_workspace_create(model.plan, self.num_chains, _PREC[self.dtype], ct.byref(h)))
        self._h = h
        self._destroy = L.lib().ut_workspace_destroy  # kept: module globals may be gone at exit
        if waves_per_chain:
            self.set_waves_per_chain(waves_per_chain)

    def set_waves_per_chain(self, waves: int) -> None:
        """Wavefronts per chain of the density kernel and the persistent samplers:
        0 automatic, 1, or 2/4/8 (logistic plans: one chain's rows split over a
        workgroup, each layout its own deterministic summation order)."""
        L.check(L.lib().ut_workspace_set_chain_waves(self._h, int(waves)))

    @property
    def waves_per_chain(self) -> int:
        w = ct.c_int32()
        L.check(L.lib().ut_workspace_chain_waves(self._h, ct.byref(w)))
        return w.value

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            self._destroy(self._h)
            self._h = None

    def gradient_kernel_time(self, params, reps: int = 20, need_logjoint: bool = False) -> float:
        """Mean device microseconds per batched gradient of all chains at `params`
        (P x C) over `reps` back-to-back calls (HIP events on the workspace stream)."""
        if self.dtype != np.float32:
            raise ValueError("gradient_kernel_time: float32 workspaces only")
        p = self._params(params)
        us = ct.c_double()
        L.check(L.lib().ut_gradient_kernel_time(self._h, L.ptr(p), int(reps), int(bool(need_logjoint)),
                                                ct.byref(us)))
        return us.value

    def gradient_step_kernel_time(self, params, reps: int = 20) -> float:
        """Mean device microseconds per leapfrog step of the lock-step samplers after a
        trajectory's first (batched gradient at `params` plus the kick and drift, step size
        0), over `reps` back-to-back calls (HIP events on the workspace stream)."""
        if self.dtype != np.float32:
            raise ValueError("gradient_step_kernel_time: float32 workspaces only")
        p = self._params(params)
        us = ct.c_double()
        L.check(L.lib().ut_gradient_step_kernel_time(self._h, L.ptr(p), int(reps), ct.byref(us)))
        return us.value

    def _params(self, params) -> np.ndarray:
        a = np.asarray(params)
        if a.ndim != 2:
            raise ValueError("params must be a P x C matrix")
        if a.shape[0] != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters per chain, got {a.shape[0]}")
        if a.shape[1] != self.num_chains:
            raise ValueError(f"DimensionMismatch: workspace has {self.num_chains} chains, got {a.shape[1]}")
        # column-major P x C == row-major C x P: chain c's parameters are contiguous
        return np.ascontiguousarray(a.T, dtype=self.dtype)


def device_batched_logjoint_inplace(ws: DeviceBatchedWorkspace, params) -> np.ndarray:
    p = ws._params(params)
    out = np.empty(ws.num_chains, dtype=ws.dtype)
    L.check(L.lib().ut_logjoint(ws.handle, L.ptr(p), L.ptr(out)))
    return out


def device_batched_logjoint_gradient_inplace(ws: DeviceBatchedWorkspace, params):
    p = ws._params(params)
    values = np.empty(ws.num_chains, dtype=ws.dtype)
    grads = np.empty((ws.num_chains, ws.num_params), dtype=ws.dtype)
    L.check(L.lib().ut_logjoint_gradient(ws.handle, L.ptr(p), L.ptr(values), L.ptr(grads)))
    return values, grads.T


def device_batched_logjoint(model: TeaModel, params, args=(), constraints=None, *, backend="hip",
                            precision=np.float32) -> np.ndarray:
    """Per-chain unconstrained logjoint of `params` (P x C), the model conditioned on
    `args` / `constraints` (TeaModel.condition)."""
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    model = model.condition(args, constraints)
    params = np.asarray(params)
    ws = DeviceBatchedWorkspace(model, params.shape[1], precision)
    return device_batched_logjoint_inplace(ws, params)


def device_batched_logjoint_gradient(model: TeaModel, params, args=(), constraints=None, *, backend="hip",
                                     precision=np.float32):
    """(values[C], gradients[P x C]) of the unconstrained logjoint, the model
    conditioned on `args` / `constraints` (TeaModel.condition)."""
    if backend != "hip":
        raise ValueError(f"backend must be 'hip', got {backend!r}")
    model = model.condition(args, constraints)
    params = np.asarray(params)
    ws = DeviceBatchedWorkspace(model, params.shape[1], precision)
    return device_batched_logjoint_gradient_inplace(ws, params)


def device_lowering_report(model) -> dict:
    """Whether and how `model` reaches the device (reference device_lowering_report,
    the report every unsupported-model error points at): the plan it lowers to, its
    parameter count, the arguments and observed choices it can be conditioned on,
    the samplers that run it, and the issues that keep it off the device (empty
    when lowerable).  Host-only: no device is touched."""
    rep = model.lowering_report()
    kind = rep.get("plan")
    batched = kind in ("LogisticRegressionBF16", "GaussianProcessPoisson") or rep.get("batched_only", False)
    rep["samplers"] = (["batched_chees", "batched_hmc(per_chain_adaptation=False)", "batched_meads"] if batched else
                       ["batched_nuts", "batched_hmc", "batched_chees", "batched_meads"]) if rep["lowerable"] else []
    return rep
