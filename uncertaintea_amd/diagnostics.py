"""Convergence diagnostics: split R-hat, rank-normalised R-hat and split ESS.

Restates the reference's estimators (src/inference/diagnostics.jl:499-651):
split chains into halves, within/between variances, rank-normal transform with
fractional ranks (r - 3/8)/(S + 1/4), paired-autocorrelation (Geyer initial
positive + monotone sequence) ESS capped at the total draw count.  Bulk ESS is the
split ESS of the rank-normalised draws (Vehtari et al. 2021), the statistic the
reference's cross-PPL harness gates on (bench/crossppl/python/analyze.py:61-66).
Autocovariances are computed with an FFT, which equals the reference's direct
lag sum up to rounding.

All functions take draws shaped (chains, samples) for one parameter, or
(chains, samples, params) for many.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfinv


def split_chains(draws: np.ndarray) -> np.ndarray:
    """(C, S, ...) -> (2C, S//2, ...): first and second half of every chain."""
    d = np.asarray(draws, dtype=np.float64)
    half = d.shape[1] // 2
    a = d[:, :half]
    b = d[:, half:2 * half]
    out = np.empty((2 * d.shape[0], half) + d.shape[2:], dtype=np.float64)
    out[0::2] = a
    out[1::2] = b
    return out


def _stats(d: np.ndarray):
    m, n = d.shape[0], d.shape[1]
    means = d.mean(axis=1)
    var = d.var(axis=1, ddof=1) if n > 1 else np.zeros_like(means)
    W = var.mean(axis=0)
    B = n * means.var(axis=0, ddof=1) if m > 1 else np.zeros_like(W)
    var_plus = (n - 1) / n * W + B / n
    return means, W, var_plus


def split_rhat_of(split: np.ndarray) -> np.ndarray:
    """Classical R-hat of already-split draws (2C, n, ...)."""
    _, W, vp = _stats(split)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.sqrt(np.maximum(vp / W, 1.0))
    r = np.where(W == 0, np.where(vp == 0, np.nan, np.inf), r)
    return r


def rank_normal(split: np.ndarray) -> np.ndarray:
    """Pooled fractional-rank inverse-normal transform (ordinal ranks)."""
    shp = split.shape
    flat = split.reshape(shp[0] * shp[1], -1)
    S = flat.shape[0]
    order = np.argsort(flat, axis=0, kind="stable")
    ranks = np.empty_like(order, dtype=np.float64)
    np.put_along_axis(ranks, order, np.arange(1, S + 1, dtype=np.float64)[:, None], axis=0)
    z = np.sqrt(2.0) * erfinv(2.0 * ((ranks - 0.375) / (S + 0.25)) - 1.0)
    return z.reshape(shp)


def _autocov_fft(split: np.ndarray) -> np.ndarray:
    """acov[c, lag, ...] = sum_i (x_i - m)(x_{i+lag} - m) / n per split chain."""
    n = split.shape[1]
    x = split - split.mean(axis=1, keepdims=True)
    nfft = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(x, n=nfft, axis=1)
    ac = np.fft.irfft(f * np.conj(f), n=nfft, axis=1)[:, :n]
    return ac / n


def split_ess_of(split: np.ndarray) -> np.ndarray:
    """ESS of already-split draws (2C, n, ...), the reference's `_split_ess`."""
    m, n = split.shape[0], split.shape[1]
    _, W, vp = _stats(split)
    total = float(m * n)
    acov = _autocov_fft(split).mean(axis=0)  # (n, ...)
    with np.errstate(divide="ignore", invalid="ignore"):
        rho = 1.0 - (W[None] - acov) / vp[None]
    rho[0] = 1.0
    rho = np.minimum(rho, 1.0)
    npairs = (n + 1) // 2
    if n % 2 == 1:
        rho = np.concatenate([rho, np.zeros((1,) + rho.shape[1:])], axis=0)
    pairs = rho[0:2 * npairs:2] + rho[1:2 * npairs:2]
    if n % 2 == 1:
        pairs[-1] = rho[2 * npairs - 2]  # the final lag stands alone
    # initial positive sequence: stop at the first non-positive pair sum
    pos = pairs > 0
    keep = np.cumprod(pos, axis=0).astype(bool)
    # monotone: running minimum over the kept prefix
    mono = np.minimum.accumulate(np.where(keep, pairs, np.inf), axis=0)
    s = np.where(keep, mono, 0.0).sum(axis=0)
    tau = np.maximum(-1.0 + 2.0 * s, 1.0)
    ess = np.minimum(total, total / tau)
    ess = np.where((W == 0) & (vp == 0), np.nan, np.where(vp <= 0, 0.0, ess))
    return ess


def rhat(draws: np.ndarray, method: str = "rank") -> np.ndarray:
    """Per-parameter R-hat of (C, S[, P]) draws; method 'split' or 'rank'."""
    if method not in ("split", "rank"):
        raise ValueError(f"rhat method must be 'split' or 'rank', got {method!r}")
    sp = split_chains(draws)
    if method == "split":
        return split_rhat_of(sp)
    bulk = split_rhat_of(rank_normal(sp))
    flat = sp.reshape(sp.shape[0] * sp.shape[1], -1)
    med = np.median(flat, axis=0).reshape(sp.shape[2:])
    folded = split_rhat_of(rank_normal(np.abs(sp - med)))
    return np.maximum(bulk, folded)


def ess(draws: np.ndarray) -> np.ndarray:
    """Split ESS of the raw draws (the reference's `ess`)."""
    return split_ess_of(split_chains(draws))


def ess_bulk(draws: np.ndarray) -> np.ndarray:
    """Bulk ESS: split ESS of the rank-normalised split draws."""
    return split_ess_of(rank_normal(split_chains(draws)))


def ess_tail(draws: np.ndarray) -> np.ndarray:
    """Tail ESS: min over the 5% and 95% quantile indicator ESS."""
    sp = split_chains(draws)
    flat = sp.reshape(sp.shape[0] * sp.shape[1], -1)
    out = None
    for q in (0.05, 0.95):
        qv = np.quantile(flat, q, axis=0).reshape(sp.shape[2:])
        e = split_ess_of((sp <= qv).astype(np.float64))
        out = e if out is None else np.minimum(out, e)
    return out


def mcse_mean(draws: np.ndarray) -> np.ndarray:
    d = np.asarray(draws, dtype=np.float64)
    sd = d.reshape(-1, *d.shape[2:]).std(axis=0, ddof=1)
    return sd / np.sqrt(ess(d))
