"""GPU: the latent-GP Poisson model.  The batched binary64 Cholesky against
numpy.linalg.cholesky of the same kernel matrix (|dL| <= 1e-11 max|L|, L L^T = K,
log det against numpy's slogdet), its hyperparameter tangents against central
finite differences of numpy's factor (relative 1e-6); logjoint and gradient of all
chains against the oracle's binary64 evaluation (which matches torch autograd) at
|d lj| <= 1e-6 |lj| and |d grad| <= 1e-5 max(1, |grad|_inf); MEADS and shared HMC on
the GP agree on the posterior within MCSE.  At the configuration's own size (N = 512,
P = 515): the factor of 8 chains against numpy, the density of 16 chains against the
oracle, and a short MEADS run (the fold reductions at P = 515) against the oracle's
MEADS on the same starting state, draws and adaptation trace to 1e-4."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [32, 64, 96])
def test_gp_cholesky_against_numpy(n):
    say("gp cholesky", n)
    X, y, _ = S.gp_poisson_data(n, 2, seed=3)
    model = M.GaussianProcessPoisson(X, y, noise=1e-2)
    hyper = np.array([[np.log(0.3), np.log(0.6), 0.0], [np.log(0.8), np.log(0.2), np.log(1.7)],
                      [np.log(0.5), np.log(0.5), np.log(0.4)]])
    Ld, dLd, ld = model.factor(hyper)
    for c, h in enumerate(hyper):
        K = S.gp_kernel(X, np.exp(h[:2]), np.exp(h[2]), 1e-2)
        Ln = np.linalg.cholesky(K)
        err = np.max(np.abs(Ld[c] - Ln)) / np.max(np.abs(Ln))
        assert err <= 1e-11, err
        assert np.max(np.abs(Ld[c] @ Ld[c].T - K)) <= 1e-11 * np.max(np.abs(K))
        assert ld[c] == pytest.approx(np.linalg.slogdet(K)[1], rel=1e-11, abs=1e-9)
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-6
            Kp = S.gp_kernel(X, np.exp((h + e)[:2]), np.exp((h + e)[2]), 1e-2)
            Km = S.gp_kernel(X, np.exp((h - e)[:2]), np.exp((h - e)[2]), 1e-2)
            fd = (np.linalg.cholesky(Kp) - np.linalg.cholesky(Km)) / 2e-6
            rel = np.max(np.abs(dLd[c, k] - fd)) / np.max(np.abs(fd))
            assert rel <= 1e-6, (c, k, rel)
    say(f"  max rel L error {err:.2e}")


def test_gp_one_input_dimension():
    """D = 1 (two hyperparameters, two tangents: the second kernel instance) against
    numpy's factor, finite differences of it, and the oracle's density."""
    say("gp d=1")
    n = 64
    X, y, _ = S.gp_poisson_data(n, 1, seed=7, lengthscales=(0.25,))
    model = M.GaussianProcessPoisson(X, y, noise=1e-2)
    hyper = np.array([[np.log(0.25), 0.0], [np.log(0.7), np.log(1.5)]])
    Ld, dLd, ld = model.factor(hyper)
    assert dLd.shape[1] == 2
    for c, h in enumerate(hyper):
        K = S.gp_kernel(X, np.exp(h[:1]), np.exp(h[1]), 1e-2)
        Ln = np.linalg.cholesky(K)
        assert np.max(np.abs(Ld[c] - Ln)) <= 1e-11 * np.max(np.abs(Ln))
        assert ld[c] == pytest.approx(np.linalg.slogdet(K)[1], rel=1e-11, abs=1e-9)
        for k in range(2):
            e = np.zeros(2)
            e[k] = 1e-6
            Kp = S.gp_kernel(X, np.exp((h + e)[:1]), np.exp((h + e)[1]), 1e-2)
            Km = S.gp_kernel(X, np.exp((h - e)[:1]), np.exp((h - e)[1]), 1e-2)
            fd = (np.linalg.cholesky(Kp) - np.linalg.cholesky(Km)) / 2e-6
            assert np.max(np.abs(dLd[c, k] - fd)) <= 1e-6 * np.max(np.abs(fd)), (c, k)
    om = B.OracleModel(model)
    rng = np.random.default_rng(8)
    C = 32
    params = np.concatenate([rng.normal(size=(2, C)) * 0.4, rng.normal(size=(n, C))], axis=0).astype(np.float32)
    v, g = ut.device_batched_logjoint_gradient(model, params)
    rv, rg, _ = om.gp_eval(params.astype(np.float64))
    assert np.max(np.abs(v - rv) / np.abs(rv)) <= 1e-6
    assert np.max(np.abs(g - rg)) <= 1e-5 * max(1.0, np.max(np.abs(rg)))


def test_gp_density_against_oracle():
    say("gp density")
    n = 128
    X, y, _ = S.gp_poisson_data(n, 2, seed=4)
    model = M.GaussianProcessPoisson(X, y, noise=1e-2)
    om = B.OracleModel(model)
    rng = np.random.default_rng(6)
    C = 64
    params = np.concatenate([rng.normal(size=(3, C)) * 0.4, rng.normal(size=(n, C))], axis=0).astype(np.float32)
    v, g = ut.device_batched_logjoint_gradient(model, params)
    rv, rg, _ = om.gp_eval(params.astype(np.float64))
    dlj = np.max(np.abs(v - rv) / np.abs(rv))
    dg = np.max(np.abs(g - rg)) / max(1.0, np.max(np.abs(rg)))
    say(f"  logjoint rel err {dlj:.2e}, gradient err {dg:.2e}")
    assert dlj <= 1e-6 and dg <= 1e-5


def test_gp_meads_matches_hmc():
    say("gp meads vs shared hmc")
    n = 64
    X, y, _ = S.gp_poisson_data(n, 2, seed=5)
    model = M.GaussianProcessPoisson(X, y, noise=1e-2)
    me = ut.batched_meads(model, num_chains=256, num_samples=1500, num_warmup=1500, seed=9)
    hm = ut.batched_hmc(model, num_chains=256, num_samples=600, num_warmup=400, num_leapfrog_steps=10,
                        per_chain_adaptation=False, seed=10)
    a = me.constrained_samples[:, :, :3]
    b = hm.constrained_samples[:, :, :3]
    se = np.sqrt(a.var(axis=(0, 1)) / np.asarray(D.ess_bulk(a)) + b.var(axis=(0, 1)) / np.asarray(D.ess_bulk(b)))
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / se
    say(f"  meads hyper means {a.mean(axis=(0, 1))}, hmc {b.mean(axis=(0, 1))}, z {z}, "
        f"meads rhat {float(np.max(np.asarray(D.rhat(me.unconstrained_samples)))):.3f}")
    assert z.max() < 5.0


def _cfg5():
    X, y, _ = S.gp_poisson_data(512, 2)
    return M.GaussianProcessPoisson(X, y, noise=1e-2), X


def test_gp_cfg5_full_size_factor_and_density():
    say("gp cfg5 N=512 factor + density")
    model, X = _cfg5()
    rng = np.random.default_rng(12)
    hyper = np.column_stack([rng.normal(np.log(0.4), 0.3, size=8), rng.normal(np.log(0.5), 0.3, size=8),
                             rng.normal(0.0, 0.3, size=8)])
    Ld, _, ld = model.factor(hyper)
    worst = 0.0
    for c, h in enumerate(hyper):
        K = S.gp_kernel(X, np.exp(h[:2]), np.exp(h[2]), 1e-2)
        Ln = np.linalg.cholesky(K)
        worst = max(worst, np.max(np.abs(Ld[c] - Ln)) / np.max(np.abs(Ln)))
        assert ld[c] == pytest.approx(np.linalg.slogdet(K)[1], rel=1e-10, abs=1e-8)
    say(f"  max rel L error {worst:.2e}")
    assert worst <= 1e-10
    om = B.OracleModel(model)
    C = 16
    params = np.concatenate([rng.normal(size=(3, C)) * 0.3, rng.normal(size=(512, C))], axis=0).astype(np.float32)
    v, g = ut.device_batched_logjoint_gradient(model, params)
    rv, rg, _ = om.gp_eval(params.astype(np.float64))
    dlj = np.max(np.abs(v - rv) / np.abs(rv))
    dg = np.max(np.abs(g - rg)) / max(1.0, np.max(np.abs(rg)))
    say(f"  logjoint rel err {dlj:.2e}, gradient err {dg:.2e}")
    assert dlj <= 1e-6 and dg <= 1e-5


def test_gp_cfg5_meads_short_run_against_oracle():
    """MEADS at P = 515 for 6 iterations from a common start: the device (blocked binary64
    Cholesky, fp32 elsewhere) and the oracle's MEADS (binary64 density rounded to binary32)
    make the same accept decisions, and draws and the per-iteration trace agree to 1e-4."""
    say("gp cfg5 meads short")
    model, _ = _cfg5()
    C, K = 16, 4
    rng = np.random.default_rng(13)
    init = np.concatenate([rng.normal(size=(3, C)) * 0.2 + np.array([[np.log(0.3)], [np.log(0.6)], [0.0]]),
                           rng.normal(size=(512, C)) * 0.5], axis=0)
    dev = ut.batched_meads(model, num_chains=C, num_samples=3, num_warmup=3, num_folds=K, seed=7,
                           initial_params=init.astype(np.float32))
    orc = B.OracleModel(model).run("meads", "twin", C, num_samples=3, num_warmup=3, num_folds=K, seed=7,
                                   threads=16, initial_params=init.astype(np.float32).astype(np.float64))
    assert np.array_equal(dev.accepted, orc["accepted"])
    d = np.max(np.abs(dev.unconstrained_samples.astype(np.float64) - orc["samples"].astype(np.float64)))
    t = np.max(np.abs(dev.trajectory_trace - orc["trajectory_trace"]) / np.maximum(1e-12, np.abs(orc["trajectory_trace"])))
    say(f"  max draw diff {d:.2e}, trace rel diff {t:.2e}, accepted {dev.accepted.mean():.2f}")
    assert d <= 1e-4 and t <= 1e-4
