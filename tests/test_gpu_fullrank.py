"""GPU: the full-rank Gaussian guide (ut_flow_create_fullrank).

Tolerance: fp32 device arithmetic with binary64 sums against the oracle's binary64
guide fed the same float32 weights: positions and vector-Jacobian products within
|d| <= 1e-5 (1 + |ref|) elementwise, the inverse (a triangular solve) by the same
bound on its backward error and by 1e-6 cond(L) on its forward error; the ELBO and IWAE weight gradients,
fed the device's model gradient and weights, within 1e-5 normwise, and exactly zero
on and above the factor's diagonal.  Statistical: on a logistic regression with
strongly correlated features the full-rank fit reaches a higher ELBO than the
mean-field one and its covariance L L^T matches NUTS's posterior correlations,
and HMC preconditioned by the fitted guide (chains in its base space) reproduces
NUTS's posterior means."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _guide(P, seed, fscale=0.1):
    g = ut.FullRankGaussian(P)
    rng = np.random.default_rng(seed)
    w = np.empty(g.num_weights, dtype=np.float32)
    w[:P] = rng.normal(size=P) * 0.3
    w[P:2 * P] = rng.normal(size=P) * 0.2 - 1.0
    w[2 * P:] = rng.normal(size=P * P) * fscale  # entries on and above the diagonal must be ignored
    g.weights = w
    return g


def _correlated_logistic(n=2000, d=8, seed=41):
    rng = np.random.default_rng(seed)
    base = rng.normal(size=(n, 1))
    x = (0.8 * base + 0.6 * rng.normal(size=(n, d))).astype(np.float32)  # pairwise feature correlation ~0.64
    beta = rng.normal(size=d) * 0.5
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-(x.astype(np.float64) @ beta)))).astype(np.float32)
    return M.LogisticRegression(x, y)


def _close(a, ref, tol=1e-5):
    return float(np.max(np.abs(a - ref) / (1.0 + np.abs(ref))))


@pytest.mark.parametrize("P", [7, 66, 130])
def test_fullrank_kernels_against_oracle(P):
    say(f"full-rank guide kernels vs oracle, P={P}")
    g = _guide(P, P)
    o = B.OracleFullRank.of(g)
    z = np.random.default_rng(P + 1).normal(size=(256, P)).astype(np.float32)
    q, ld = g.forward(z)
    rq, rld = o.forward(z)
    zi, ldi = g.inverse(rq.astype(np.float32))
    rzi, rldi = o.inverse(rq.astype(np.float32))
    gq = np.random.default_rng(P + 2).normal(size=(256, P)).astype(np.float32)
    gz = g.vjp(z, gq)
    rgz = o.vjp(gq)
    # the inverse is a triangular solve: judged by its backward error (the oracle maps
    # the device's z back onto q), its forward error by the factor's condition number
    back = _close(o.forward(zi)[0], rq.astype(np.float32).astype(np.float64))
    cond = np.linalg.cond(g.cholesky())
    e = (_close(q, rq), _close(ld, rld), back, _close(ldi, rldi), _close(gz, rgz))
    say("  forward %.1e logdet %.1e inverse (backward error) %.1e logdet_inv %.1e vjp %.1e" % e
        + f"; inverse forward error {_close(zi, rzi):.1e} at cond(L) {cond:.0f}")
    assert max(e) <= 1e-5
    assert _close(zi, rzi) <= 1e-6 * cond
    with pytest.raises(ValueError, match="full-rank"):
        g.loss_grad(rq.astype(np.float32))


def test_fullrank_elbo_and_iwae_gradients_against_oracle():
    say("full-rank ELBO / IWAE gradients vs oracle")
    model = _correlated_logistic(1000, 12)
    P = model.num_params
    g = _guide(P, 3, fscale=0.05)
    o = B.OracleFullRank.of(g)
    z = np.random.default_rng(4).normal(size=(256, P)).astype(np.float32)
    elbo, grad, gq, valid = g.elbo_grad(model, z)
    assert valid.all() and grad.shape == (2 * P + P * P,)
    rg = o.elbo_grad(z, gq)
    e_g = np.linalg.norm(grad - rg) / np.linalg.norm(rg)
    Fg = grad[2 * P:].reshape(P, P)
    q, ld = o.forward(z)
    lj, _ = B.OracleModel(model).logjoint_grad_f64(np.ascontiguousarray(q.T))
    logw = lj + 0.5 * np.sum(z.astype(np.float64) ** 2, axis=1) + 0.5 * P * np.log(2 * np.pi) + ld
    say(f"  elbo {elbo:.6f} vs {logw.mean():.6f}; gradient vs oracle {e_g:.2e}")
    assert e_g <= 1e-5
    assert np.all(np.triu(Fg) == 0.0) and np.any(np.tril(Fg, -1) != 0.0)
    assert abs(elbo - logw.mean()) <= 1e-5 * (1 + abs(logw.mean()))
    bound, gi, gqi, valid, rw = g.elbo_grad(model, z, iwae_group=8, return_weights=True)
    w_ref, b_ref = B.iwae_weights_and_bound(logw, valid, 8)
    rgi = o.elbo_grad(z, gqi, row_weights=rw.astype(np.float64))
    e_i = np.linalg.norm(gi - rgi) / np.linalg.norm(rgi)
    say(f"  iwae bound {bound:.6f} vs {b_ref:.6f}; weights max rel {np.max(np.abs(rw - w_ref)) / np.max(w_ref):.2e}; "
        f"gradient vs oracle {e_i:.2e}")
    assert abs(bound - b_ref) <= 1e-4 * (1 + abs(b_ref))
    assert e_i <= 1e-5
    assert bound >= elbo - 1e-4


def test_fullrank_large_p_gradient():
    """P = 643 (the latent GP with N = 640 inputs): the factor gradient's row splits
    drop from 64 to 40 so the binary64 partials stay within 128 MiB; forward and ELBO
    gradient still match the oracle."""
    say("full-rank guide at P=643")
    from uncertaintea_amd import synthetic as S
    rng = np.random.default_rng(51)
    X, y, _ = S.gp_poisson_data(640, 2, seed=53)
    model = ut.GaussianProcessPoisson(X, y)
    P = model.num_params
    assert P == 643
    g = _guide(P, 52, fscale=0.01)
    o = B.OracleFullRank.of(g)
    z = rng.normal(size=(128, P)).astype(np.float32)
    q, _ = g.forward(z)
    rq, _ = o.forward(z)
    elbo, grad, gq, valid = g.elbo_grad(model, z)
    rg = o.elbo_grad(z, gq)
    e_q, e_g = _close(q, rq), np.linalg.norm(grad - rg) / np.linalg.norm(rg)
    say(f"  forward {e_q:.1e}; gradient vs oracle {e_g:.2e}")
    assert valid.all() and e_q <= 1e-5 and e_g <= 1e-5
    assert np.all(np.triu(grad[2 * P:].reshape(P, P)) == 0.0)


def test_fullrank_invalid_particles_and_iwae_fit():
    """A particle pushed to v = -1e4 on the funnel overflows the logjoint: it is flagged
    invalid, carries no weight, and the gradient over the valid rows matches the oracle;
    batched_advi(guide="fullrank", elbo="iwae") keeps the bound above the standard ELBO
    of the same particles on every iteration and raises it from a too-narrow start."""
    say("full-rank guide: invalid particles, IWAE fit")
    model = M.NealsFunnel(10, 3.0)
    P = model.num_params
    g = _guide(P, 9, fscale=0.05)
    w = g.weights
    w[:P] = 0.0
    g.weights = w
    z = np.random.default_rng(10).normal(size=(256, P)).astype(np.float32)
    z[0, 0] = -1e4
    elbo, grad, gq, valid = g.elbo_grad(model, z)
    assert not valid[0] and valid[1:].all() and np.all(gq[0] == 0.0)
    rg = B.OracleFullRank.of(g).elbo_grad(z, gq, valid.astype(np.int32))
    e_g = np.linalg.norm(grad - rg) / np.linalg.norm(rg)
    say(f"  valid {valid.sum()} of {valid.size}; gradient vs oracle {e_g:.2e}")
    assert e_g <= 1e-5
    r = ut.batched_advi(model, num_iterations=300, num_particles=512, learning_rate=5e-3, guide="fullrank",
                        elbo="iwae", iwae_samples=16, initial_params=np.zeros(P), initial_log_scale=-1.5, seed=11)
    h, st = r.elbo_history, r.standard_elbo_history
    say(f"  iwae bound {h[:20].mean():.3f} -> {h[-50:].mean():.3f}; standard elbo {st[:20].mean():.3f} -> "
        f"{st[-50:].mean():.3f}")
    # from log-scale -1.5 the bound starts near -9.5 (log Z = 0); at log-scale 0 it is -1.2
    assert np.all(h >= st - 1e-9) and h[-50:].mean() > h[:20].mean() + 1.0
    assert r.guide == "fullrank" and r.iwae_samples == 16 and r.scale_factor.shape == (P, P)


def test_fullrank_advi_correlated_logistic():
    say("full-rank vs mean-field ADVI on a correlated logistic regression")
    model = _correlated_logistic()
    P = model.num_params
    res = {}
    for guide in ("meanfield", "fullrank"):
        res[guide] = ut.batched_advi(model, num_iterations=2000, num_particles=512, learning_rate=0.02, guide=guide,
                                     seed=5)
        say(f"  {guide}: best elbo {res[guide].best_elbo:.3f}")
    fr = res["fullrank"]
    assert fr.best_elbo > res["meanfield"].best_elbo + 1.0
    assert fr.best_scale_factor.shape == (P, P) and np.all(np.triu(fr.best_scale_factor) == 0.0)
    nuts = ut.batched_nuts(model, num_chains=128, num_samples=400, num_warmup=400, seed=6)
    d = nuts.unconstrained_samples.reshape(-1, P).astype(np.float64)
    cov_n = np.cov(d.T)
    cov_f = ut.variational_covariance(fr)
    sd_n, sd_f = np.sqrt(np.diag(cov_n)), np.sqrt(np.diag(cov_f))
    corr_n, corr_f = cov_n / np.outer(sd_n, sd_n), cov_f / np.outer(sd_f, sd_f)
    e_corr = np.max(np.abs(corr_f - corr_n))
    e_loc = np.max(np.abs(fr.best_location - d.mean(axis=0)) / sd_n)
    say(f"  posterior corr min {np.min(corr_n):.3f}; |corr_fr - corr_nuts| max {e_corr:.3f}; sd ratio "
        f"[{np.min(sd_f / sd_n):.3f}, {np.max(sd_f / sd_n):.3f}]; max |loc - mean| / sd {e_loc:.3f}")
    assert e_corr < 0.1 and e_loc < 0.3
    assert np.min(sd_f / sd_n) > 0.8 and np.max(sd_f / sd_n) < 1.2
    s = ut.variational_samples(fr, 16384, space="unconstrained")
    assert np.max(np.abs(np.cov(s) - cov_f)) < 0.1 * np.max(np.abs(cov_f))
    # the fitted guide preconditions shared-adaptation HMC (chains in its base space,
    # near-standard: eps settles near 0.8, so 2 steps give T ~ pi/2; 4 would give T ~ pi,
    # where a standard Gaussian's trajectories come back to their start)
    h = ut.batched_hmc(model, num_chains=256, num_samples=300, num_warmup=200, num_leapfrog_steps=2,
                       per_chain_adaptation=False, flow=fr.best_flow, seed=7)
    dh = h.unconstrained_samples.astype(np.float64)
    rh = float(np.max(np.asarray(D.rhat(h.unconstrained_samples))))
    e_m = np.max(np.abs(dh.reshape(-1, P).mean(axis=0) - d.mean(axis=0)) / sd_n)
    say(f"  preconditioned hmc: eps {float(h.step_size[0]):.3f}, rhat {rh:.4f}, max |mean - nuts| / sd {e_m:.3f}")
    assert rh < 1.02 and e_m < 0.1


def test_fullrank_graph_replay_matches_eager(monkeypatch):
    """Attached to a workspace, the full-rank guide's gradient (forward with F z, model
    gradient, L^T g) is captured as a HIP graph and replayed at every leapfrog step:
    whole shared-HMC and ChEES runs must equal the eager launches bit for bit."""
    say("full-rank guide: graph replay vs eager")
    model = M.NealsFunnel(10, 3.0)
    g = _guide(model.num_params, 23, fscale=0.1)
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("UT_FLOW_GRAPHS", mode)
        h = ut.batched_hmc(model, num_chains=128, num_samples=20, num_warmup=30, num_leapfrog_steps=6,
                           per_chain_adaptation=False, flow=g, seed=24)
        c = ut.batched_chees(model, num_chains=128, num_samples=20, num_warmup=30, flow=g, seed=25)
        runs[mode] = (h, c)
    for a, b in zip(runs["0"], runs["1"]):
        assert np.array_equal(a.unconstrained_samples.view(np.uint32), b.unconstrained_samples.view(np.uint32))
        assert np.array_equal(a.logjoint_values, b.logjoint_values)
        assert np.array_equal(a.step_size, b.step_size)
