"""CPU checks that pin the oracle: known-answer vectors, library routines and
closed forms.  A plausible slip (wrong sign, dropped term, swapped index) in the
oracle fails at least one of these."""
import math

import numpy as np
import pytest
import torch

from oracle import binding as B
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S


def test_philox_kat(golden):
    for key, ctr, expected in golden["philox_kat"]:
        assert B.philox4x32(key, ctr).tolist() == expected


def test_philox_bijection_and_numpy_twin():
    blocks, uf, nf, ud = B.philox_fill(3, 9, 0, 0, 5000)
    assert len({tuple(b) for b in blocks.tolist()}) == 5000
    w = S.philox4x32(3, 9, 0, np.arange(5000, dtype=np.uint32), 0, 0)
    assert np.array_equal(np.stack(w, axis=1), blocks)
    # exact conversions (reference rng.jl:255-262)
    assert np.array_equal(uf, ((blocks[:, 0] >> 8).astype(np.float64) * 2.0 ** -24).astype(np.float32))
    bits = (blocks[:, 0].astype(np.uint64) << np.uint64(32)) | blocks[:, 1].astype(np.uint64)
    assert np.array_equal(ud, (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)


def test_philox_normal_and_uniform_distribution():
    _, uf, nf, ud = B.philox_fill(0, 0, 1, 0, 400_000)
    assert uf.min() >= 0 and uf.max() < 1 and ud.max() < 1
    assert abs(ud.mean() - 0.5) < 2e-3 and abs(ud.var() - 1 / 12) < 2e-3
    z = nf.astype(np.float64)
    assert np.all(np.isfinite(z))
    assert abs(z.mean()) < 6e-3 and abs(z.var() - 1) < 1e-2
    assert abs(np.mean(np.abs(z) <= 1) - 0.6827) < 5e-3
    # Box-Muller: float32 radius*cos of the first two uniforms, against libm binary64
    blocks, _, nf2, _ = B.philox_fill(7, 1, 0, 0, 20000)
    u1 = (blocks[:, 0] >> 8) * 2.0 ** -24
    u2 = (blocks[:, 1] >> 8) * 2.0 ** -24
    ref = np.sqrt(-2 * np.log(1 - u1)) * np.cos(np.float32(6.2831855) * u2)
    assert np.max(np.abs(nf2 - ref)) < 2e-5 * np.maximum(1, np.abs(ref)).max()


@pytest.mark.parametrize("fn,lo,hi,ref,tol", [
    ("exp", -87, 88, np.exp, 2e-7), ("log1p", -0.999, 1e4, np.log1p, 3e-7),
    ("cos", 0, 6.3, np.cos, 2e-7), ("log1p_exp", -80, 80, lambda x: np.logaddexp(0, x), 3e-7),
    ("logistic", -80, 80, lambda x: 1 / (1 + np.exp(-x)), 1e-6)])
def test_math_f32_accuracy(fn, lo, hi, ref, tol):
    x = np.linspace(lo, hi, 200_001).astype(np.float32)
    y = B.math_f32(fn, x).astype(np.float64)
    r = ref(x.astype(np.float64))
    if fn == "cos":
        assert np.max(np.abs(y - r)) < tol
    else:
        assert np.max(np.abs(y - r) / np.maximum(np.abs(r), 1e-30)) < tol


def test_math_f32_log_special_values():
    x = np.exp(np.linspace(-103, 88, 200_001)).astype(np.float32)
    x = np.concatenate([x, np.array([1e-45, 1e-40, 1.0, np.inf, 0.0, -1.0], dtype=np.float32)])
    y = B.math_f32("log", x)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.log(x.astype(np.float64))
    fin = np.isfinite(r)
    assert np.max(np.abs(y[fin] - r[fin])) < 1e-5
    assert y[-3] == np.inf and y[-2] == -np.inf and np.isnan(y[-1])
    assert B.math_f32("log", np.array([1.0], np.float32))[0] == 0.0


@pytest.mark.parametrize("fn,x,ref", [
    ("exp", np.linspace(-708, 709, 100_001), np.exp),
    ("log", np.exp(np.linspace(-740, 700, 100_001)), np.log),
    ("log1p", np.linspace(-0.9999, 1e3, 100_001), np.log1p),
    ("log1p_exp", np.linspace(-700, 700, 100_001), lambda x: np.logaddexp(0, x)),
    ("pow_m075", np.arange(1, 100_001, dtype=np.float64), lambda x: x ** -0.75)])
def test_math_f64_accuracy(fn, x, ref):
    y = B.math_f64(fn, x)
    r = ref(x)
    ok = r != 0
    tol = 1e-14 if fn == "pow_m075" else 2e-15
    assert np.max(np.abs(y[ok] - r[ok]) / np.abs(r[ok])) < tol


def test_warmup_schedule_kat(golden):
    for n, ib, ends, tb in golden["warmup_schedule"]:
        assert B.warmup_schedule(n) == (ib, ends, tb)
    # windows tile the slow phase; the doubling pattern of the reference
    ib, ends, tb = B.warmup_schedule(1000)
    assert (ib, tb) == (100, 100) and ends[-1] == 900 and ends[0] == 125
    assert B.warmup_schedule(0) == (0, [], 0)


def test_dual_averaging_properties():
    # accept == target keeps hbar = 0, so log step = mu = log(10 step0)
    steps, fin = B.dual_averaging(0.1, 0.8, np.full(50, 0.8))
    assert np.allclose(steps, 1.0, rtol=1e-14) and abs(fin - 1.0) < 1e-13
    # always accepting pushes the step up, never accepting pushes it down
    up, _ = B.dual_averaging(0.1, 0.8, np.ones(30))
    down, _ = B.dual_averaging(0.1, 0.8, np.zeros(30))
    assert up[-1] > 1.0 > down[-1]
    # first update, by hand from Hoffman & Gelman (2014) eq. 6 with t0=10, gamma=0.05
    hbar = (1 / 11) * (0.8 - 0.3)
    assert abs(down[0] if False else B.dual_averaging(0.1, 0.8, [0.3])[0][0]
               - math.exp(math.log(1.0) - 1 / 0.05 * hbar)) < 1e-12


def test_running_variance_matches_weighted_variance():
    # unit weights, no clipping in effect (well inside 8 sd): shrunk sample variance
    rng = np.random.default_rng(1)
    scale = np.array([1.0, 2.0, 0.5])
    hist = rng.normal(size=(50, 4, 3)) * scale
    # a wide first iteration keeps every later draw inside the 5-8 sd clip bound
    hist[0] = np.array([3.0, -3.0, 3.0, -3.0])[:, None] * scale
    acc = np.ones((50, 4))
    flags = np.full((50, 4), 2, np.uint8)  # accepted => weight 1 for HMC
    im, eff = B.running_variance(hist, acc, flags, window_length=50, use_accepted=True)
    flat = hist.reshape(-1, 3)
    n = flat.shape[0]
    assert abs(eff - n) < 1e-9
    var = flat.var(axis=0, ddof=1)
    shrink = n / (n + 5.0)
    assert np.allclose(im, shrink * var + (1 - shrink), rtol=1e-12)
    # divergent draws carry zero weight
    flags2 = flags.copy()
    flags2[:, 0] |= 1
    im2, eff2 = B.running_variance(hist, acc, flags2, window_length=50, use_accepted=True)
    assert abs(eff2 - 150) < 1e-9


def _model_from(kind, g):
    if kind == "logistic":
        return M.LogisticRegression(np.array(g["x"], np.float32), np.array(g["y"], np.float32), g["prior_loc"],
                                    g["prior_scale"])
    if kind == "hier":
        return M.HierarchicalNormalNC(np.array(g["y"], np.float32), np.array(g["offsets"], np.int32),
                                      np.array(g["sigma"]), g["mu_loc"], g["mu_scale"], g["tau_scale"])
    return M.NealsFunnel(g["dim"], g["v_scale"])


@pytest.mark.parametrize("kind", ["logistic", "hier", "funnel"])
def test_oracle_gradients_against_torch_golden(golden, kind):
    g = golden[kind]
    om = B.OracleModel(_model_from(kind, g))
    params = np.array([c["params"] for c in g["cases"]]).T
    lj_ref = np.array([c["logjoint"] for c in g["cases"]])
    g_ref = np.array([c["gradient"] for c in g["cases"]]).T
    lj, gr = om.logjoint_grad_f64(params)
    assert np.allclose(lj, lj_ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(gr, g_ref, rtol=1e-11, atol=1e-11)
    # kernel-order binary64 twin agrees to rounding, binary32 twin to float tolerance
    lj2, gr2 = om.logjoint_grad_twin(params, np.float64)
    assert np.allclose(lj2, lj_ref, rtol=1e-12, atol=1e-11)
    assert np.allclose(gr2, g_ref, rtol=1e-11, atol=1e-11)
    lj3, gr3 = om.logjoint_grad_twin(params.astype(np.float32), np.float32)
    assert np.allclose(lj3, lj_ref, rtol=1e-5, atol=1e-4)
    assert np.max(np.abs(gr3 - g_ref)) <= 1e-5 * max(1.0, np.max(np.abs(g_ref)))


def test_oracle_gradient_finite_differences():
    x, y, _ = S.logistic_data(200, 6)
    om = B.OracleModel(M.LogisticRegression(x, y))
    q = np.random.default_rng(2).normal(size=(6, 1)) * 0.3
    _, g = om.logjoint_grad_f64(q)
    for d in range(6):
        e = np.zeros((6, 1)); e[d] = 1e-6
        fd = (om.logjoint_grad_f64(q + e)[0] - om.logjoint_grad_f64(q - e)[0]) / 2e-6
        assert abs(fd[0] - g[d, 0]) < 1e-5 * max(1, abs(g[d, 0]))


def test_oracle_nuts_funnel_moments():
    # v ~ N(0, 1) marginally; E[x_i^2] = E[e^v] = e^{1/2}
    om = B.OracleModel(M.NealsFunnel(dim=2, v_scale=1.0))
    r = om.run("nuts", "f64", 64, num_samples=400, num_warmup=300, seed=5, threads=8)
    s = r["samples"]
    v = s[..., 0].ravel()
    assert abs(v.mean()) < 0.08 and abs(v.var() - 1.0) < 0.12
    assert abs(np.mean(s[..., 1] ** 2) - math.exp(0.5)) < 0.25
    assert r["divergent"].mean() < 0.02


def test_oracle_hmc_logistic_posterior_mean_near_truth():
    x, y, beta = S.logistic_data(1000, 5)
    om = B.OracleModel(M.LogisticRegression(x, y))
    r = om.run("hmc", "f64", 16, num_samples=300, num_warmup=300, num_leapfrog_steps=8, step_size=0.05,
               seed=3, threads=8)
    post = r["samples"].reshape(-1, 5).mean(axis=0)
    # posterior mean within a few posterior sd of the generating beta for N=1000
    sd = r["samples"].reshape(-1, 5).std(axis=0)
    assert np.all(np.abs(post - beta) < 4 * sd + 0.05)
    assert r["accepted"].mean() > 0.5


def test_twin_and_f64_nuts_agree_statistically():
    yy, off, sig, _ = S.hierarchical_data(16, 8)
    om = B.OracleModel(M.HierarchicalNormalNC(yy, off, sig))
    a = om.run("nuts", "twin", 32, num_samples=300, num_warmup=300, seed=11, threads=8)
    b = om.run("nuts", "f64", 32, num_samples=300, num_warmup=300, seed=12, threads=8)
    ma, mb = a["samples"][..., 0].mean(), b["samples"][..., 0].mean()
    sa = a["samples"][..., 0].std()
    assert abs(ma - mb) < 0.15 * sa + 0.05
    # same seed, different thread counts: identical draws
    c = om.run("nuts", "twin", 32, num_samples=50, num_warmup=50, seed=11, threads=2)
    d = om.run("nuts", "twin", 32, num_samples=50, num_warmup=50, seed=11, threads=7)
    assert np.array_equal(c["samples"], d["samples"])
    e = om.run("nuts", "twin", 32, num_samples=50, num_warmup=50, seed=13, threads=7)
    assert not np.array_equal(c["samples"], e["samples"])
