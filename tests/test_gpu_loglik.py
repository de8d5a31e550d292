"""GPU: the pointwise log-likelihood (ut_pointwise_loglik) and the ArviZ
log_likelihood group.

Tolerance: binary64 on the device against scipy's normal logpdf and stable
log-sigmoid (scipy.special.log_expit) of every observation (|d| <= 1e-12 (1 + |ref|)); the sum over observations plus the priors
(scipy.stats, with the log-transform Jacobian of tau) must equal the oracle's
binary64 logjoint at every draw (1e-10 (1 + |ref|)), which pins the model's
decomposition and the observation order at once."""
import numpy as np
import pytest
from scipy import stats
from scipy.special import log_expit

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


class _Draws:
    """The fields of HMCChains that pointwise_loglikelihood reads."""

    def __init__(self, model, d, C):
        self.model = model
        self.unconstrained_samples = d.reshape(C, -1, model.num_params)


def _close(a, ref):
    return float(np.max(np.abs(a - ref) / (1.0 + np.abs(ref))))


def test_logistic_pointwise_against_scipy_and_logjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, 6)).astype(np.float32)
    y = (rng.random(300) < 0.4).astype(np.float32)
    model = ut.LogisticRegression(x, y)
    d = rng.normal(size=(2 * 150, 6)) * 1.5
    ll = ut.pointwise_loglikelihood(_Draws(model, d, 2))
    eta = d @ x.astype(np.float64).T
    # scipy's stable log-sigmoid (bernoulli.logpmf(y, expit(eta)) takes log 1 - p of a
    # rounded p and loses ~1e-9 where |eta| is large)
    ref = y[None, :] * log_expit(eta) + (1 - y[None, :]) * log_expit(-eta)
    lj, _ = B.OracleModel(model).logjoint_grad_f64(d.T)
    e = _close(ll, ref)
    e_sum = _close(ll.sum(axis=1) + stats.norm.logpdf(d).sum(axis=1), lj)
    print(f"logistic: vs scipy {e:.1e}; sum + prior vs oracle logjoint {e_sum:.1e}", flush=True)
    assert ll.shape == (300, 300) and e <= 1e-12 and e_sum <= 1e-10


def test_bf16_glm_pointwise_against_scipy_and_logjoint():
    """The matrix-core GLM plan (D = 256): the log-likelihood uses its bf16 X, which the
    model holds widened as x_model."""
    rng = np.random.default_rng(4)
    x = (rng.normal(size=(256, 256)) / 16).astype(np.float32)
    y = (rng.random(256) < 0.5).astype(np.float32)
    model = ut.LogisticRegressionBF16(x, y)
    d = rng.normal(size=(2 * 64, 256))
    ll = ut.pointwise_loglikelihood(_Draws(model, d, 2))
    eta = d @ model.x_model.astype(np.float64).T
    ref = y[None, :] * log_expit(eta) + (1 - y[None, :]) * log_expit(-eta)
    lj, _ = B.OracleModel(model).logjoint_grad_f64(d.T)
    e = _close(ll, ref)
    e_sum = _close(ll.sum(axis=1) + stats.norm.logpdf(d).sum(axis=1), lj)
    print(f"bf16 glm: vs scipy {e:.1e}; sum + prior vs oracle logjoint {e_sum:.1e}", flush=True)
    assert ll.shape == (128, 256) and e <= 1e-12 and e_sum <= 1e-10


def test_hierarchical_pointwise_against_scipy_and_logjoint():
    rng = np.random.default_rng(2)
    sizes = [3, 5, 1, 7, 4, 2, 6, 3]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    y = rng.normal(size=off[-1]).astype(np.float32)
    sig = rng.uniform(0.5, 2.0, size=8)
    model = ut.HierarchicalNormalNC(y, off, sig, mu_loc=0.3, mu_scale=5.0, tau_scale=2.0)
    P = model.num_params
    d = rng.normal(size=(4 * 50, P)) * 0.7
    ll = ut.pointwise_loglikelihood(_Draws(model, d, 4))
    grp = np.repeat(np.arange(8), sizes)
    theta = d[:, [0]] + np.exp(d[:, [1]]) * d[:, 2:][:, grp]
    ref = stats.norm.logpdf(y.astype(np.float64)[None, :], theta, sig[grp][None, :])
    prior = (stats.norm.logpdf(d[:, 0], 0.3, 5.0) + stats.halfcauchy.logpdf(np.exp(d[:, 1]), scale=2.0) + d[:, 1]
             + stats.norm.logpdf(d[:, 2:]).sum(axis=1))
    lj, _ = B.OracleModel(model).logjoint_grad_f64(d.T)
    e = _close(ll, ref)
    e_sum = _close(ll.sum(axis=1) + prior, lj)
    print(f"hierarchical: vs scipy {e:.1e}; sum + prior vs oracle logjoint {e_sum:.1e}", flush=True)
    assert ll.shape == (200, off[-1]) and e <= 1e-12 and e_sum <= 1e-10


def test_funnel_has_no_observations():
    model = ut.NealsFunnel(4, 3.0)
    ll = ut.pointwise_loglikelihood(_Draws(model, np.zeros((10, 5)), 2))
    assert ll.shape == (10, 0)


def test_gp_pointwise_against_scipy_and_logjoint():
    """The latent GP (N = 96, three tiles): f = chol(K) z by numpy, poisson logpmf by
    scipy, at the binary32-rounded draws the device evaluates."""
    X, yc, _ = S.gp_poisson_data(96, 2, seed=3)
    gp = ut.GaussianProcessPoisson(X, yc)
    P = gp.num_params
    rng = np.random.default_rng(5)
    d = np.concatenate([rng.normal(size=(40, 3)) * 0.3, rng.normal(size=(40, P - 3))], axis=1)
    d = d.astype(np.float32).astype(np.float64)
    ll = ut.pointwise_loglikelihood(_Draws(gp, d, 2))
    ref = np.array([stats.poisson.logpmf(yc, np.exp(np.linalg.cholesky(
        S.gp_kernel(X, np.exp(r[:2]), np.exp(r[2]), 1e-2)) @ r[3:])) for r in d])
    prior = (stats.norm.logpdf(d[:, :2], 0.0, 0.5).sum(axis=1) + stats.halfnorm.logpdf(np.exp(d[:, 2]), scale=1.0)
             + d[:, 2] + stats.norm.logpdf(d[:, 3:]).sum(axis=1))
    lj = B.OracleModel(gp).gp_eval(d.T)[0]
    e, e_sum = _close(ll, ref), _close(ll.sum(axis=1) + prior, lj)
    print(f"gp: vs scipy/numpy {e:.1e}; sum + prior vs oracle logjoint {e_sum:.1e}", flush=True)
    assert ll.shape == (40, 96) and e <= 1e-10 and e_sum <= 1e-10


def test_arviz_log_likelihood_group():
    """to_arviz_dict(..., log_likelihood=True) from a short NUTS run: per-observation
    (draw, chain) matrices in both layouts and the grouped form with its dims/coords,
    all equal to pointwise_loglikelihood's chain-major rows."""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, 3)).astype(np.float32)
    y = (rng.random(40) < 0.5).astype(np.float32)
    model = ut.LogisticRegression(x, y)
    ch = ut.batched_nuts(model, num_chains=64, num_samples=30, num_warmup=30, seed=4)
    ll = ut.pointwise_loglikelihood(ch)
    assert ll.shape == (64 * 30, 40)
    d = ut.to_arviz_dict(ch, log_likelihood=True)
    g = d["log_likelihood"]
    assert sorted(g) == sorted(f"y[{k + 1}]" for k in range(40))
    assert np.array_equal(g["y[7]"], ll[:, 6].reshape(64, 30).T)
    dc = ut.to_arviz_dict(ch, layout="chain_draw", log_likelihood=True)
    assert np.array_equal(dc["log_likelihood"]["y[7]"], ll[:, 6].reshape(64, 30))
    dg = ut.to_arviz_dict(ch, flatten_vectors=False, log_likelihood=True)
    assert dg["log_likelihood"]["y"].shape == (30, 64, 40)
    assert dg["dims"]["y"] == ["y_dim_0"] and dg["coords"]["y_dim_0"] == list(range(1, 41))
    assert np.array_equal(dg["log_likelihood"]["y"][:, :, 6], g["y[7]"])
    assert "log_likelihood" not in ut.to_arviz_dict(ch)
