"""Convergence diagnostics: split R-hat, rank-normalised R-hat and split ESS.

Restates the reference's estimators (src/inference/diagnostics.jl:499-651):
split chains into halves, within/between variances, rank-normal transform with
fractional ranks (r - 3/8)/(S + 1/4) of the ordinal (stable-sort) ranks,
paired-autocorrelation (Geyer initial positive + monotone sequence) ESS capped at
the total draw count.  Bulk ESS is the split ESS of the rank-normalised draws
(Vehtari et al. 2021), the statistic the reference's cross-PPL harness gates on
(bench/crossppl/python/analyze.py:61-66).  Autocovariances are computed with an
FFT, which equals the reference's direct lag sum up to rounding.

The arithmetic runs in torch (binary64) on the GPU when one is visible, else on the
CPU, so diagnostics of a million-draw run take seconds.  Functions accept numpy
arrays shaped (chains, samples) for one parameter or (chains, samples, params) for
many, and return numpy arrays.  Every statistic is per parameter, so large arrays are
processed in parameter chunks of at most 2^27 draws (the rank transform needs several
binary64 copies of its input).
"""
from __future__ import annotations

import functools

import numpy as np
import torch

_CHUNK_DRAWS = 1 << 27


def _per_param_chunks(fn):
    @functools.wraps(fn)
    def wrapper(draws, *args, **kwargs):
        a = np.asarray(draws)
        if a.ndim < 3 or a.size <= _CHUNK_DRAWS:
            return fn(a, *args, **kwargs)
        step = max(1, _CHUNK_DRAWS // (a.shape[0] * a.shape[1]))
        outs = [fn(a[:, :, p0:p0 + step], *args, **kwargs) for p0 in range(0, a.shape[2], step)]
        if isinstance(outs[0], tuple):
            return tuple(np.concatenate([o[i] for o in outs]) for i in range(len(outs[0])))
        return np.concatenate(outs)
    return wrapper


def _device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def _t(draws) -> torch.Tensor:
    a = torch.as_tensor(np.ascontiguousarray(draws))
    return a.to(device=_device(), dtype=torch.float64)


def _split(d: torch.Tensor) -> torch.Tensor:
    """(C, S, ...) -> (2C, S//2, ...): first and second half of every chain."""
    half = d.shape[1] // 2
    out = torch.stack([d[:, :half], d[:, half:2 * half]], dim=1)
    return out.reshape((2 * d.shape[0], half) + tuple(d.shape[2:]))


def _stats(d: torch.Tensor):
    m, n = d.shape[0], d.shape[1]
    means = d.mean(dim=1)
    var = d.var(dim=1, unbiased=True) if n > 1 else torch.zeros_like(means)
    W = var.mean(dim=0)
    B = n * means.var(dim=0, unbiased=True) if m > 1 else torch.zeros_like(W)
    var_plus = (n - 1) / n * W + B / n
    return means, W, var_plus


def _rhat_of(sp: torch.Tensor) -> torch.Tensor:
    _, W, vp = _stats(sp)
    r = torch.sqrt(torch.clamp(vp / W, min=1.0))
    nan = torch.full_like(r, float("nan"))
    inf = torch.full_like(r, float("inf"))
    return torch.where(W == 0, torch.where(vp == 0, nan, inf), r)


def _rank_normal(sp: torch.Tensor) -> torch.Tensor:
    shp = sp.shape
    flat = sp.reshape(shp[0] * shp[1], -1)
    S = flat.shape[0]
    order = torch.argsort(flat, dim=0, stable=True)
    ranks = torch.empty_like(flat)
    pos = torch.arange(1, S + 1, device=flat.device, dtype=flat.dtype).unsqueeze(1).expand_as(flat)
    ranks.scatter_(0, order, pos)
    z = np.sqrt(2.0) * torch.special.erfinv(2.0 * ((ranks - 0.375) / (S + 0.25)) - 1.0)
    return z.reshape(shp)


def _ess_of(sp: torch.Tensor) -> torch.Tensor:
    m, n = sp.shape[0], sp.shape[1]
    _, W, vp = _stats(sp)
    total = float(m * n)
    x = sp - sp.mean(dim=1, keepdim=True)
    nfft = 1 << int(np.ceil(np.log2(2 * n)))
    f = torch.fft.rfft(x, n=nfft, dim=1)
    acov = torch.fft.irfft(f * torch.conj(f), n=nfft, dim=1)[:, :n] / n
    acov = acov.mean(dim=0)  # (n, ...)
    rho = 1.0 - (W.unsqueeze(0) - acov) / vp.unsqueeze(0)
    rho[0] = 1.0
    rho = torch.clamp(rho, max=1.0)
    if n % 2 == 1:
        rho = torch.cat([rho, torch.zeros_like(rho[:1])], dim=0)
    pairs = rho[0::2] + rho[1::2]
    keep = torch.cumprod((pairs > 0).to(torch.int64), dim=0).bool()
    mono = torch.cummin(torch.where(keep, pairs, torch.full_like(pairs, float("inf"))), dim=0).values
    s = torch.where(keep, mono, torch.zeros_like(mono)).sum(dim=0)
    tau = torch.clamp(-1.0 + 2.0 * s, min=1.0)
    ess = torch.clamp(total / tau, max=total)
    ess = torch.where(vp <= 0, torch.zeros_like(ess), ess)
    return torch.where((W == 0) & (vp == 0), torch.full_like(ess, float("nan")), ess)


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def split_chains(draws) -> np.ndarray:
    return _np(_split(_t(draws)))


@_per_param_chunks
def rhat(draws, method: str = "rank") -> np.ndarray:
    """Per-parameter R-hat of (C, S[, P]) draws; method 'split' or 'rank'."""
    if method not in ("split", "rank"):
        raise ValueError(f"rhat method must be 'split' or 'rank', got {method!r}")
    sp = _split(_t(draws))
    if method == "split":
        return _np(_rhat_of(sp))
    bulk = _rhat_of(_rank_normal(sp))
    flat = sp.reshape(sp.shape[0] * sp.shape[1], -1)
    # reference median: mean of the two middle order statistics for even counts
    srt = torch.sort(flat, dim=0).values
    S = flat.shape[0]
    med = 0.5 * (srt[(S - 1) // 2] + srt[S // 2])
    folded = _rhat_of(_rank_normal(torch.abs(sp - med.reshape(sp.shape[2:]))))
    return _np(torch.maximum(bulk, folded))


@_per_param_chunks
def ess(draws) -> np.ndarray:
    """Split ESS of the raw draws (the reference's `ess`)."""
    return _np(_ess_of(_split(_t(draws))))


@_per_param_chunks
def ess_bulk(draws) -> np.ndarray:
    """Bulk ESS: split ESS of the rank-normalised split draws."""
    return _np(_ess_of(_rank_normal(_split(_t(draws)))))


@_per_param_chunks
def rhat_and_ess_bulk(draws):
    """(rank-normalised R-hat, bulk ESS) sharing one transfer and one rank transform."""
    sp = _split(_t(draws))
    z = _rank_normal(sp)
    bulk_r = _rhat_of(z)
    ess_b = _ess_of(z)
    del z
    flat = sp.reshape(sp.shape[0] * sp.shape[1], -1)
    srt = torch.sort(flat, dim=0).values
    S = flat.shape[0]
    med = 0.5 * (srt[(S - 1) // 2] + srt[S // 2])
    del srt
    folded = _rhat_of(_rank_normal(torch.abs(sp - med.reshape(sp.shape[2:]))))
    return _np(torch.maximum(bulk_r, folded)), _np(ess_b)


@_per_param_chunks
def ess_tail(draws) -> np.ndarray:
    """Tail ESS: min over the 5% and 95% quantile indicator ESS."""
    sp = _split(_t(draws))
    flat = sp.reshape(sp.shape[0] * sp.shape[1], -1)
    out = None
    srt = torch.sort(flat, dim=0).values if flat.shape[0] > 16_000_000 else None
    for q in (0.05, 0.95):
        if srt is None:
            qv = torch.quantile(flat, q, dim=0)
        else:  # torch.quantile's size limit: the same linear interpolation on sorted draws
            pos = q * (flat.shape[0] - 1)
            lo = int(np.floor(pos))
            qv = srt[lo] + (pos - lo) * (srt[min(lo + 1, flat.shape[0] - 1)] - srt[lo])
        e = _ess_of((sp <= qv.reshape(sp.shape[2:])).to(torch.float64))
        out = e if out is None else torch.minimum(out, e)
    return _np(out)


@_per_param_chunks
def mcse_mean(draws) -> np.ndarray:
    d = _t(draws)
    sd = d.reshape((-1,) + tuple(d.shape[2:])).std(dim=0, unbiased=True)
    return _np(sd) / np.sqrt(ess(draws))
