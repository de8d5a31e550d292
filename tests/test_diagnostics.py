"""Diagnostics against a direct restatement of the reference loops
(src/inference/diagnostics.jl:513-613) and against textbook values; the cross-PPL
harness's ArviZ 0.23.4 statistics against direct lag sums, scipy's average ranks,
AR(1) closed forms (including super-efficient antithetic chains) and the asymptotic
quantile MCSE.  ArviZ itself is not importable here, so agreement with its own code
is pinned only through these properties."""
import math

import numpy as np
import pytest
import torch
from scipy.special import erfinv

from uncertaintea_amd import diagnostics as D


def _direct_split_ess(draws):
    """Lag-by-lag pair loop exactly as the reference's `_split_ess`."""
    m, n = draws.shape
    means = draws.mean(axis=1)
    W = draws.var(axis=1, ddof=1).mean()
    B = n * means.var(ddof=1)
    vp = (n - 1) / n * W + B / n
    pair_sums = []
    for start in range(0, n, 2):
        s = 0.0
        for lag in range(start, min(start + 1, n - 1) + 1):
            ac = np.mean([np.sum((draws[c, :n - lag] - means[c]) * (draws[c, lag:] - means[c])) / n for c in range(m)])
            rho = 1.0 if lag == 0 else 1 - (W - ac) / vp
            s += min(rho, 1.0)
        if s <= 0:
            break
        pair_sums.append(s)
    for i in range(1, len(pair_sums)):
        pair_sums[i] = min(pair_sums[i], pair_sums[i - 1])
    tau = max(-1 + 2 * sum(pair_sums), 1.0)
    return min(m * n, m * n / tau)


def _split(d):
    h = d.shape[1] // 2
    out = np.empty((2 * d.shape[0], h))
    out[0::2] = d[:, :h]
    out[1::2] = d[:, h:2 * h]
    return out


def _ar1(rng, chains, n, rho):
    x = np.empty((chains, n))
    x[:, 0] = rng.normal(size=chains) / math.sqrt(1 - rho ** 2)
    for t in range(1, n):
        x[:, t] = rho * x[:, t - 1] + rng.normal(size=chains)
    return x


def test_ess_matches_direct_loop():
    rng = np.random.default_rng(0)
    for n in (40, 41):
        d = _ar1(rng, 4, n, 0.6)
        assert abs(D.ess(d) - _direct_split_ess(_split(d))) < 1e-8 * n


def test_bulk_ess_and_rank_rhat_match_direct():
    rng = np.random.default_rng(1)
    d = _ar1(rng, 3, 60, 0.3) ** 3
    sp = _split(d)
    flat = sp.ravel()
    ranks = np.empty(flat.size)
    ranks[np.argsort(flat, kind="stable")] = np.arange(1, flat.size + 1)
    z = (np.sqrt(2) * erfinv(2 * ((ranks - 0.375) / (flat.size + 0.25)) - 1)).reshape(sp.shape)
    assert abs(D.ess_bulk(d) - _direct_split_ess(z)) < 1e-8 * d.size

    def split_rhat(x):
        n = x.shape[1]
        W = x.var(axis=1, ddof=1).mean()
        B = n * x.mean(axis=1).var(ddof=1)
        return math.sqrt(max(((n - 1) / n * W + B / n) / W, 1.0))
    med = np.median(flat)
    f = np.abs(sp - med).ravel()
    fr = np.empty(f.size)
    fr[np.argsort(f, kind="stable")] = np.arange(1, f.size + 1)
    fz = (np.sqrt(2) * erfinv(2 * ((fr - 0.375) / (f.size + 0.25)) - 1)).reshape(sp.shape)
    assert abs(D.rhat(d) - max(split_rhat(z), split_rhat(fz))) < 1e-12


def test_textbook_values():
    rng = np.random.default_rng(2)
    iid = rng.normal(size=(8, 2000))
    assert abs(D.ess(iid) / iid.size - 1) < 0.1
    assert abs(D.rhat(iid, method="split") - 1) < 0.01
    ar = _ar1(rng, 8, 4000, 0.5)  # ESS/N -> (1 - rho)/(1 + rho) = 1/3
    assert abs(D.ess(ar) / ar.size - 1 / 3) < 0.05
    shifted = iid + np.arange(8)[:, None]  # chains disagree: R-hat far above 1
    assert D.rhat(shifted) > 1.5
    # vectorised over parameters equals per-parameter calls
    many = np.stack([iid, ar[:, :2000]], axis=2)
    assert np.allclose(D.ess_bulk(many), [D.ess_bulk(iid), D.ess_bulk(ar[:, :2000])])
    r, e = D.rhat_and_ess_bulk(many)
    assert np.allclose(r, D.rhat(many)) and np.allclose(e, D.ess_bulk(many))


def test_parameter_chunks_agree_with_whole_array(monkeypatch):
    """Large draw arrays are processed in parameter chunks; every statistic is per
    parameter, so chunking changes nothing beyond reduction-order rounding."""
    import uncertaintea_amd.diagnostics as DD
    rng = np.random.default_rng(3)
    x = rng.normal(size=(8, 400, 50)).astype(np.float32)
    whole = (*D.rhat_and_ess_bulk(x), D.mcse_mean(x), D.ess_tail(x))
    monkeypatch.setattr(DD, "_CHUNK_DRAWS", 1 << 13)
    chunked = (*D.rhat_and_ess_bulk(x), D.mcse_mean(x), D.ess_tail(x))
    for a, b in zip(whole, chunked):
        assert a.shape == b.shape == (50,)
        np.testing.assert_allclose(a, b, rtol=1e-12)


# ------------------------------------------------ the harness's (ArviZ 0.23.4) statistics
def _ar1_series(rho, C, S, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=(C, S))
    a = np.empty((C, S))
    a[:, 0] = e[:, 0] / np.sqrt(1 - rho * rho)  # stationary start
    for t in range(1, S):
        a[:, t] = rho * a[:, t - 1] + e[:, t]
    return a


def _arviz_ess_direct(x):
    """ArviZ's _ess written out with the direct lag-sum autocovariance (no FFT)."""
    m, n = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    acov = np.array([[np.dot(xc[c, :n - k], xc[c, k:]) / n for k in range(n)] for c in range(m)])
    mean_var = acov[:, 0].mean() * n / (n - 1)
    var_plus = mean_var * (n - 1) / n + np.var(x.mean(axis=1), ddof=1)
    r = lambda k: 1.0 - (mean_var - acov[:, k].mean()) / var_plus  # noqa: E731
    rho = np.zeros(n)
    even, odd = 1.0, r(1)
    rho[0], rho[1] = even, odd
    t = 1
    while t < n - 3 and even + odd > 0:
        even, odd = r(t + 1), r(t + 2)
        if even + odd >= 0:
            rho[t + 1], rho[t + 2] = even, odd
        t += 2
    max_t = t - 2
    if even > 0:
        rho[max_t + 1] = even
    t = 1
    while t <= max_t - 2:
        if rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]:
            rho[t + 1] = rho[t + 2] = (rho[t - 1] + rho[t]) / 2
        t += 2
    tau = max(-1 + 2 * rho[:max_t + 1].sum() + rho[max_t + 1:max_t + 2].sum(), 1 / np.log10(m * n))
    return m * n / tau


def test_arviz_ess_fft_equals_direct_lag_sums():
    x = _ar1_series(0.6, 4, 301, 3)
    half = x.shape[1] // 2
    sp = np.concatenate([x[:, :half], x[:, -half:]])  # ArviZ split: the middle draw is dropped
    assert D.arviz_ess_mean(x) == pytest.approx(_arviz_ess_direct(sp), rel=1e-10)


def test_arviz_ess_textbook_values():
    """iid: C S; stationary AR(1): C S (1 - rho) / (1 + rho), including the antithetic case
    where ESS exceeds the draw count (ArviZ caps at C S log10(C S), diagnostics.jl at C S)."""
    iid = np.random.default_rng(1).normal(size=(8, 2000, 2))
    assert np.allclose(D.arviz_ess_bulk(iid), 16000, rtol=0.06)
    for rho in (0.5, -0.5):
        a = _ar1_series(rho, 8, 20000, 4)
        want = a.size * (1 - rho) / (1 + rho)
        assert D.arviz_ess_bulk(a) == pytest.approx(want, rel=0.08)
        assert D.arviz_ess_mean(a) == pytest.approx(want, rel=0.08)
    anti = _ar1_series(-0.5, 8, 20000, 4)
    assert D.arviz_ess_bulk(anti) > 2.5 * anti.size and D.ess_bulk(anti) == anti.size
    const = np.ones((4, 100))
    assert D.arviz_ess_bulk(const) == 400.0


def test_arviz_ranks_average_ties_and_rhat():
    from scipy import stats
    v = np.random.default_rng(2).integers(0, 6, size=(3, 40)).astype(float)
    flat = torch.as_tensor(v.reshape(-1, 1))
    assert np.array_equal(D._average_ranks(flat).numpy()[:, 0], stats.rankdata(v.ravel(), method="average"))
    iid = np.random.default_rng(5).normal(size=(8, 1000, 3))
    assert np.all(D.arviz_rhat(iid) < 1.01)
    shifted = iid.copy()
    shifted[0] += 2.0
    assert np.all(D.arviz_rhat(shifted) > 1.1)


def test_arviz_mcse_quantile_asymptote_and_harness_gate():
    """For iid normal draws the quantile MCSE approaches sqrt(p (1 - p) / n) / phi(z_p); the
    gate passes the exact law and fails a 10-MCSE shift of the mean or of a quantile."""
    from scipy import stats
    x = np.random.default_rng(6).normal(size=(8, 50000, 2))
    n = 8 * 50000
    for p in (0.05, 0.95):
        want = np.sqrt(p * (1 - p) / n) / stats.norm.pdf(stats.norm.ppf(p))
        assert np.allclose(D.arviz_mcse_quantile(x, p), want, rtol=0.12)
    st = D.harness_stats(x)
    exact = {"mean": np.zeros(2), "q05": np.full(2, stats.norm.ppf(0.05)), "q95": np.full(2, stats.norm.ppf(0.95))}
    ok, fails, zmax = D.harness_gate(st, exact)
    assert ok and zmax < 4.0, fails
    bad = dict(exact, mean=exact["mean"] + 10 * st["mcse_mean"])
    assert not D.harness_gate(st, bad)[0]
    bad = dict(exact, q95=exact["q95"] - 10 * st["mcse_q95"])
    assert not D.harness_gate(st, bad)[0]


def test_harness_stats_share_of_work_equals_the_single_statistics(monkeypatch):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(8, 1001, 5)).cumsum(axis=1) * 0.1 + rng.normal(size=(8, 1001, 5))
    monkeypatch.setattr(D, "_CHUNK_DRAWS", 8 * 1001 * 2)  # parameter chunks of 2
    st = D.harness_stats(x)
    for key, want in (("ess_bulk", D.arviz_ess_bulk(x)), ("rhat", D.arviz_rhat(x)), ("ess_tail", D.arviz_ess_tail(x)),
                      ("mcse_mean", D.arviz_mcse_mean(x)), ("mcse_q05", D.arviz_mcse_quantile(x, 0.05)),
                      ("mcse_q95", D.arviz_mcse_quantile(x, 0.95)),
                      ("q05", np.quantile(x.reshape(-1, 5), 0.05, axis=0))):
        assert np.allclose(st[key], want, rtol=1e-12), key
