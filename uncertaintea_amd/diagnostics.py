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


# --------------------------------------------------------------------------------------
# The cross-PPL harness's statistics (bench/crossppl/python/analyze.py:51-97), which call
# ArviZ 0.23.4 (the version bench/crossppl/python/uv.lock pins).  Restated from ArviZ's
# published estimators (Vehtari et al. 2021, arviz/stats/diagnostics.py), which differ
# from the reference's diagnostics.jl in four places: ties get average ranks; a split
# drops the middle draw of an odd-length chain; the autocorrelation sum stops at lag
# n - 3, adds the even term of the first non-positive pair, and is not clipped to 1; and
# ESS may exceed the draw count (tau >= 1 / log10(C S) instead of tau >= 1).

def _arviz_split(d: torch.Tensor) -> torch.Tensor:
    half = d.shape[1] // 2
    return torch.cat([d[:, :half], d[:, d.shape[1] - half:]], dim=0)


def _average_ranks(flat: torch.Tensor) -> torch.Tensor:
    """Ranks 1..S of each column with ties given their average rank (scipy 'average')."""
    S = flat.shape[0]
    srt, order = torch.sort(flat, dim=0, stable=True)
    out = torch.empty_like(flat)
    pos = torch.arange(1, S + 1, device=flat.device, dtype=flat.dtype)
    for j in range(flat.shape[1]):
        col = srt[:, j]
        new = torch.ones(S, dtype=torch.bool, device=flat.device)
        new[1:] = col[1:] != col[:-1]
        gid = torch.cumsum(new.to(torch.int64), dim=0) - 1
        ng = int(gid[-1].item()) + 1
        first = torch.zeros(ng, dtype=flat.dtype, device=flat.device).scatter_reduce(
            0, gid, pos, reduce="amin", include_self=False)
        last = torch.zeros(ng, dtype=flat.dtype, device=flat.device).scatter_reduce(
            0, gid, pos, reduce="amax", include_self=False)
        out[order[:, j], j] = (0.5 * (first + last))[gid]
    return out


def _arviz_z_scale(sp: torch.Tensor) -> torch.Tensor:
    shp = sp.shape
    flat = sp.reshape(shp[0] * shp[1], -1)
    S = flat.shape[0]
    r = (_average_ranks(flat) - 0.375) / (S + 0.25)
    return (np.sqrt(2.0) * torch.special.erfinv(2.0 * r - 1.0)).reshape(shp)


def _arviz_ess(sp: torch.Tensor) -> np.ndarray:
    """ArviZ `_ess` of (chains, draws, params) columns."""
    m, n = sp.shape[0], sp.shape[1]
    total = m * n
    flat = sp.reshape(m, n, -1)
    P = flat.shape[2]
    rng = (flat.amax(dim=(0, 1)) - flat.amin(dim=(0, 1)))
    x = flat - flat.mean(dim=1, keepdim=True)
    nfft = 1 << int(np.ceil(np.log2(2 * n)))
    f = torch.fft.rfft(x, n=nfft, dim=1)
    acov = (torch.fft.irfft(f * torch.conj(f), n=nfft, dim=1)[:, :n] / n).mean(dim=0)  # (n, P)
    mean_var = acov[0] * n / (n - 1.0)
    var_plus = mean_var * (n - 1.0) / n
    if m > 1:
        var_plus = var_plus + flat.mean(dim=1).var(dim=0, unbiased=True)
    rho_all = _np(1.0 - (mean_var.unsqueeze(0) - acov) / var_plus.unsqueeze(0))
    rng = _np(rng)
    out = np.empty(P)
    for j in range(P):
        if not np.isfinite(rng[j]):
            out[j] = np.nan
            continue
        if rng[j] < np.finfo(float).resolution:
            out[j] = float(total)
            continue
        rho = np.zeros(n)
        rho_all_j = rho_all[:, j]
        even, odd = 1.0, rho_all_j[1] if n > 1 else 0.0
        rho[0] = even
        if n > 1:
            rho[1] = odd
        t = 1
        while t < n - 3 and (even + odd) > 0.0:
            even, odd = rho_all_j[t + 1], rho_all_j[t + 2]
            if even + odd >= 0:
                rho[t + 1], rho[t + 2] = even, odd
            t += 2
        max_t = t - 2
        if even > 0:
            rho[max_t + 1] = even
        t = 1
        while t <= max_t - 2:
            if rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]:
                rho[t + 1] = (rho[t - 1] + rho[t]) / 2.0
                rho[t + 2] = rho[t + 1]
            t += 2
        tau = -1.0 + 2.0 * np.sum(rho[:max_t + 1]) + np.sum(rho[max_t + 1:max_t + 2])
        tau = max(tau, 1.0 / np.log10(total))
        out[j] = np.nan if np.isnan(rho).any() else total / tau
    return out


def _arviz_cols(draws) -> torch.Tensor:
    d = _t(draws)
    return d.unsqueeze(-1) if d.ndim == 2 else d


def _squeeze(draws, v):
    return v[0] if np.asarray(draws).ndim == 2 else v


@_per_param_chunks
def arviz_ess_bulk(draws) -> np.ndarray:
    """ArviZ bulk ESS (az.ess(method='bulk')) per parameter of (C, S[, P]) draws."""
    return _squeeze(draws, _arviz_ess(_arviz_z_scale(_arviz_split(_arviz_cols(draws)))))


@_per_param_chunks
def arviz_ess_mean(draws) -> np.ndarray:
    return _squeeze(draws, _arviz_ess(_arviz_split(_arviz_cols(draws))))


def _quantile_cols(flat: torch.Tensor, q: float) -> torch.Tensor:
    # numpy's default (linear) quantile, column-wise
    srt = torch.sort(flat, dim=0).values
    pos = q * (flat.shape[0] - 1)
    lo = int(np.floor(pos))
    hi = min(lo + 1, flat.shape[0] - 1)
    return srt[lo] + (pos - lo) * (srt[hi] - srt[lo])


@_per_param_chunks
def arviz_ess_quantile(draws, prob: float) -> np.ndarray:
    d = _arviz_cols(draws)
    qv = _quantile_cols(d.reshape(-1, d.shape[2]), prob)
    return _squeeze(draws, _arviz_ess(_arviz_split((d <= qv).to(torch.float64))))


def arviz_ess_tail(draws) -> np.ndarray:
    """ArviZ tail ESS (az.ess(method='tail')): min of the 5 % and 95 % quantile-indicator ESS."""
    return np.minimum(arviz_ess_quantile(draws, 0.05), arviz_ess_quantile(draws, 0.95))


def _arviz_rhat_cols(sp: torch.Tensor) -> torch.Tensor:
    n = sp.shape[1]
    chain_mean = sp.mean(dim=1)
    between = n * chain_mean.var(dim=0, unbiased=True)
    within = sp.var(dim=1, unbiased=True).mean(dim=0)
    return torch.sqrt((between / within + n - 1) / n)


@_per_param_chunks
def arviz_rhat(draws) -> np.ndarray:
    """ArviZ rank R-hat (az.rhat(method='rank')): max of the bulk and folded split R-hat."""
    d = _arviz_cols(draws)
    bulk = _arviz_rhat_cols(_arviz_z_scale(_arviz_split(d)))
    flat = d.reshape(-1, d.shape[2])
    srt = torch.sort(flat, dim=0).values
    S = flat.shape[0]
    med = 0.5 * (srt[(S - 1) // 2] + srt[S // 2])
    tail = _arviz_rhat_cols(_arviz_z_scale(_arviz_split(torch.abs(d - med))))
    return _squeeze(draws, _np(torch.maximum(bulk, tail)))


@_per_param_chunks
def arviz_mcse_mean(draws) -> np.ndarray:
    d = _arviz_cols(draws)
    sd = _np(d.reshape(-1, d.shape[2]).std(dim=0, unbiased=True))
    return _squeeze(draws, sd / np.sqrt(_arviz_ess(_arviz_split(d))))


@_per_param_chunks
def arviz_mcse_quantile(draws, prob: float) -> np.ndarray:
    """ArviZ quantile MCSE: half the width of the order-statistic interval whose Beta
    posterior (ESS of the quantile indicator) covers +-1 sd around `prob`."""
    from scipy import stats
    d = _arviz_cols(draws)
    ess_q = np.atleast_1d(arviz_ess_quantile(draws, prob))
    srt = _np(torch.sort(d.reshape(-1, d.shape[2]), dim=0).values)
    size = srt.shape[0]
    out = np.empty(srt.shape[1])
    for j in range(srt.shape[1]):
        a = stats.beta.ppf([0.1586553, 0.8413447], ess_q[j] * prob + 1, ess_q[j] * (1 - prob) + 1)
        th1 = srt[int(np.floor(np.nanmax((a[0] * size, 0)))), j]
        th2 = srt[int(np.ceil(np.nanmin((a[1] * size, size - 1)))), j]
        out[j] = (th2 - th1) / 2
    return _squeeze(draws, out)


@_per_param_chunks
def _harness_arrays(draws):
    d = _arviz_cols(draws)
    P = d.shape[2]
    flat = d.reshape(-1, P)
    size = flat.shape[0]
    srt = torch.sort(flat, dim=0).values

    def quant(q):
        pos = q * (size - 1)
        lo = int(np.floor(pos))
        return srt[lo] + (pos - lo) * (srt[min(lo + 1, size - 1)] - srt[lo])
    q05, q95 = quant(0.05), quant(0.95)
    med = 0.5 * (srt[(size - 1) // 2] + srt[size // 2])
    sp = _arviz_split(d)
    z = _arviz_z_scale(sp)
    ess_bulk = _arviz_ess(z)
    r_bulk = _arviz_rhat_cols(z)
    del z
    r_tail = _arviz_rhat_cols(_arviz_z_scale(_arviz_split(torch.abs(d - med))))
    ess_mean = _arviz_ess(sp)
    ess_q = [_arviz_ess(_arviz_split((d <= q).to(torch.float64))) for q in (q05, q95)]
    sd = _np(flat.std(dim=0, unbiased=True))
    srt_np = _np(srt)
    from scipy import stats
    mcse_q = []
    for prob, e in zip((0.05, 0.95), ess_q):
        m = np.empty(P)
        for j in range(P):
            a = stats.beta.ppf([0.1586553, 0.8413447], e[j] * prob + 1, e[j] * (1 - prob) + 1)
            th1 = srt_np[int(np.floor(np.nanmax((a[0] * size, 0)))), j]
            th2 = srt_np[int(np.ceil(np.nanmin((a[1] * size, size - 1)))), j]
            m[j] = (th2 - th1) / 2
        mcse_q.append(m)
    return (_np(flat.mean(dim=0)), sd, _np(q05), _np(q95), _np(torch.maximum(r_bulk, r_tail)), ess_bulk,
            np.minimum(ess_q[0], ess_q[1]), sd / np.sqrt(ess_mean), mcse_q[0], mcse_q[1])


def harness_stats(draws) -> dict:
    """analyze.py:51-72 per parameter of (C, S, P) draws: mean, sd, 5 % / 95 % quantiles,
    rank R-hat, bulk / tail ESS and the MCSE of the mean and of both quantiles (one sort
    and one rank transform shared per parameter chunk)."""
    keys = ("mean", "sd", "q05", "q95", "rhat", "ess_bulk", "ess_tail", "mcse_mean", "mcse_q05", "mcse_q95")
    return dict(zip(keys, _harness_arrays(np.asarray(draws))))


def harness_gate(stats: dict, ref: dict | None, rhat_max: float = 1.01, z_max: float = 4.0):
    """analyze.py:76-97: every rank R-hat below `rhat_max`, and (against a reference's
    statistics) mean, q05 and q95 within `z_max` combined MCSE.  A reference with zero MCSE
    (exact values) gives z against the run's own MCSE.  Returns (passed, failures, max z)."""
    failures = []
    for p in np.nonzero(~(np.asarray(stats["rhat"]) < rhat_max))[0]:
        failures.append(f"param {p}: rhat={float(stats['rhat'][p]):.4f}")
    zmax = 0.0
    if ref is not None:
        for key, mkey in (("mean", "mcse_mean"), ("q05", "mcse_q05"), ("q95", "mcse_q95")):
            denom = np.sqrt(np.asarray(stats[mkey]) ** 2 + np.asarray(ref.get(mkey, 0.0)) ** 2)
            z = np.where(denom > 0, np.abs(np.asarray(stats[key]) - np.asarray(ref[key])) / np.where(denom > 0, denom, 1),
                         0.0)
            zmax = max(zmax, float(np.nanmax(z)))
            for p in np.nonzero(z >= z_max)[0]:
                failures.append(f"param {p}: {key} z={float(z[p]):.1f}")
    return not failures, failures, zmax
