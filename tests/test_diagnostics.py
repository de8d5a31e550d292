"""Diagnostics against a direct restatement of the reference loops
(src/inference/diagnostics.jl:513-613) and against textbook values."""
import math

import numpy as np
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
