"""GPU: batched_advi on the device (mean-field and flow guides).

Tolerance: the mean-field guide's ELBO weight gradient (the flow without coupling
layers) against the oracle's binary64 flow fed the device's model gradient, within
1e-5 normwise (no bf16 operands on this path: fp32 arithmetic only).  Statistical:
on the 25-feature logistic regression (N = 1000) the mean-field fit's location
sits within half a posterior sd of NUTS's posterior means and its scales within
[0.6, 1.1] of the posterior sds (the mean-field optimum under-states the sd by the
posterior's correlations, which are weak here); the flow guide's draws match the
funnel better than the mean-field guide (a higher ELBO; v's spread is printed)."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import _lib as L
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_meanfield_elbo_gradient_against_oracle():
    say("mean-field ELBO gradient vs oracle")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    P = model.num_params
    flow = ut.AffineCouplingFlow(P, 0, 256, seed=1)
    rng = np.random.default_rng(2)
    flow.set_base(rng.normal(size=P) * 0.3, rng.normal(size=P) * 0.2 - 1.0)
    z = rng.normal(size=(256, P)).astype(np.float32)
    elbo, g, gq, valid = flow.elbo_grad(model, z)
    assert valid.all() and g.shape == (2 * P,)
    rg = B.OracleFlow.of(flow).elbo_grad(z, gq)
    e = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    q, ld = B.OracleFlow.of(flow).forward(z)
    lj, _ = B.OracleModel(model).logjoint_grad_f64(np.ascontiguousarray(q.T))
    ref = np.mean(lj + 0.5 * np.sum(z.astype(np.float64) ** 2, axis=1) + 0.5 * P * np.log(2 * np.pi) + ld)
    say(f"  weight gradient vs oracle {e:.2e}; elbo {elbo:.6f} vs {ref:.6f}")
    assert e <= 1e-5
    assert abs(elbo - ref) <= 1e-5 * (1 + abs(ref))


def test_initial_params_hierarchical_prior():
    """The hierarchical plan (staged on the device): mu ~ N(0, 5), log tau with
    tau ~ halfcauchy(5) (median 5), theta_raw ~ N(0, 1)."""
    y = np.zeros(8, dtype=np.float32)
    h = ut.HierarchicalNormalNC(y, np.array([0, 4, 8], dtype=np.int32), np.ones(2))
    b = np.empty((40000, h.num_params))
    L.check(L.lib().ut_initial_params(h.plan, L.UT_INIT_PRIOR, 3, 40000, L.ptr(b)))
    assert abs(np.median(np.exp(b[:, 1])) / 5.0 - 1.0) < 0.05
    assert abs(b[:, 0].std() / 5.0 - 1.0) < 0.03 and abs(b[:, 2:].std() - 1.0) < 0.02


def test_meanfield_advi_logistic_against_nuts():
    say("mean-field ADVI on the logistic regression")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    res = ut.batched_advi(model, num_iterations=1500, num_particles=256, learning_rate=0.02, seed=3)
    h = res.elbo_history
    nuts = ut.batched_nuts(model, num_chains=128, num_samples=500, num_warmup=500, seed=4)
    d = nuts.unconstrained_samples.reshape(-1, model.num_params)
    m, sd = d.mean(axis=0), d.std(axis=0)
    loc, sc = res.best_location, np.exp(res.best_log_scale)
    say(f"  elbo {h[:20].mean():.2f} -> {h[-100:].mean():.2f} (best {res.best_elbo:.2f}); max |loc - mean| / sd "
        f"{np.max(np.abs(loc - m) / sd):.3f}; scale / sd in [{np.min(sc / sd):.3f}, {np.max(sc / sd):.3f}]")
    assert h[-100:].mean() > h[:20].mean()
    assert np.max(np.abs(loc - m) / sd) < 0.5
    assert np.min(sc / sd) > 0.6 and np.max(sc / sd) < 1.1
    mean = ut.variational_mean(res, space="unconstrained")
    assert np.array_equal(mean, res.best_location)
    cov = ut.variational_covariance(res)
    assert np.allclose(np.diag(cov), sc ** 2)
    s = ut.variational_samples(res, 4096, space="unconstrained")
    assert s.shape == (model.num_params, 4096)
    assert np.max(np.abs(s.mean(axis=1) - loc) / sc) < 5 / np.sqrt(4096) * 1.5


def test_flow_advi_funnel_spread():
    say("flow vs mean-field ADVI on the funnel")
    model = M.NealsFunnel(10, 3.0)
    out = {}
    for guide in ("meanfield", "flow"):
        r = ut.batched_advi(model, num_iterations=1500, num_particles=512, learning_rate=2e-3, guide=guide,
                            num_flow_layers=6, initial_params=np.zeros(11), initial_log_scale=0.0, seed=5)
        s = ut.variational_samples(r, 16384, space="unconstrained")
        out[guide] = (r.best_elbo, s[0].std())
        say(f"  {guide}: best elbo {r.best_elbo:.3f}, v sd {s[0].std():.3f} (exact 3)")
    # the flow family contains the mean-field one (zero conditioners): a better bound
    assert out["flow"][0] > out["meanfield"][0] + 0.5


def test_iwae_gradient_against_oracle():
    """Importance-weighted bound (groups of 8) for a mean-field guide on the funnel:
    the device's particle weights and bound against the reference's reduction
    applied to binary64 log-weights, and its gradient against the oracle's weighted
    gradient fed the device's model gradient and weights."""
    say("IWAE gradient vs oracle")
    model = M.NealsFunnel(10, 3.0)
    P = model.num_params
    flow = ut.AffineCouplingFlow(P, 0, 256, seed=6)
    rng = np.random.default_rng(7)
    flow.set_base(rng.normal(size=P) * 0.3, rng.normal(size=P) * 0.2 - 0.5)
    z = rng.normal(size=(256, P)).astype(np.float32)
    bound, g, gq, valid, rw = flow.elbo_grad(model, z, iwae_group=8, return_weights=True)
    assert valid.all()
    of = B.OracleFlow.of(flow)
    q, ld = of.forward(z)
    lj, _ = B.OracleModel(model).logjoint_grad_f64(np.ascontiguousarray(q.T))
    logw = lj + 0.5 * np.sum(z.astype(np.float64) ** 2, axis=1) + 0.5 * P * np.log(2 * np.pi) + ld
    w_ref, b_ref = B.iwae_weights_and_bound(logw, valid, 8)
    e_w = np.max(np.abs(rw - w_ref)) / np.max(w_ref)
    rg = of.elbo_grad(z, gq, row_weights=rw.astype(np.float64))
    e_g = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    say(f"  bound {bound:.6f} vs {b_ref:.6f}; weights max rel {e_w:.2e}; gradient vs oracle {e_g:.2e}")
    assert abs(bound - b_ref) <= 1e-4 * (1 + abs(b_ref))
    assert e_w <= 1e-3
    assert e_g <= 1e-5
    std, _, _, _ = flow.elbo_grad(model, z)
    assert bound >= std - 1e-4  # Jensen, on the same particles


def test_iwae_fit_bound_above_elbo():
    """batched_advi(elbo="iwae"): on every iteration the bound is at least the standard
    ELBO of the same particles, and the fit raises it: the bound of the fitted guide,
    evaluated on 16384 common particles (1024 groups of 16), exceeds the starting
    guide's on the same particles by more than four standard errors of the
    difference (the per-iteration histories, 512 particles each, are too noisy for a
    first-versus-last comparison)."""
    model = M.NealsFunnel(10, 3.0)
    P = model.num_params
    # start away from the optimum (location 1.5, scales e^-1): the bound's gradient
    # then carries a clear signal (from N(0, I) the 16-particle estimator's gradient
    # noise swamps the small remaining gain, so 400 iterations need not raise it)
    r = ut.batched_advi(model, num_iterations=400, num_particles=512, learning_rate=5e-3, elbo="iwae",
                        iwae_samples=16, initial_params=np.full(P, 1.5), initial_log_scale=-1.0, seed=8)
    h, s = r.elbo_history, r.standard_elbo_history
    say(f"  iwae bound {h[:20].mean():.3f} -> {h[-50:].mean():.3f}; standard elbo {s[:20].mean():.3f} -> "
        f"{s[-50:].mean():.3f}")
    assert np.all(h >= s - 1e-9)
    assert r.elbo == "iwae" and r.iwae_samples == 16
    z = np.random.default_rng(9).standard_normal((16384, P)).astype(np.float32)
    start = ut.AffineCouplingFlow(P, 0, 256, seed=1)
    start.set_base(np.full(P, 1.5), np.full(P, -1.0))
    lw = []
    for g in (start, r.flow):
        # the device objective is the mean over groups of log mean exp(log w); the
        # per-group values come from the binary64 oracle on the same particles
        obj, _, _, valid = g.elbo_grad(model, z, iwae_group=16)
        assert valid.all()
        q, ld = B.OracleFlow.of(g).forward(z)
        lj, _ = B.OracleModel(model).logjoint_grad_f64(np.ascontiguousarray(q.T))
        logw = lj + 0.5 * np.sum(z.astype(np.float64) ** 2, axis=1) + 0.5 * P * np.log(2 * np.pi) + ld
        m = logw.reshape(-1, 16)
        mx = m.max(axis=1, keepdims=True)
        lw.append((mx[:, 0] + np.log(np.mean(np.exp(m - mx), axis=1))))
        assert abs(obj - lw[-1].mean()) <= 1e-4 * (1 + abs(obj)), (obj, lw[-1].mean())
    d = lw[1] - lw[0]
    se = d.std(ddof=1) / np.sqrt(d.size)
    say(f"  bound on common particles: start {lw[0].mean():.4f}, fitted {lw[1].mean():.4f} (diff {d.mean():.4f} "
        f"+- {se:.4f})")
    assert d.mean() > 4 * se


def test_fit_without_valid_particles_stops_and_keeps_weights():
    """Every particle of a guide centred at v = -1e4 overflows the funnel's logjoint:
    ut_flow_fit_elbo reports the first such iteration (as the reference throws there,
    src/inference/vi.jl:762-764) and leaves the weights as they were, although the
    iterations queue on the device without a host check in between."""
    model = M.NealsFunnel(10, 3.0)
    P = model.num_params
    flow = ut.AffineCouplingFlow(P, 0, 256, seed=12)
    mu = np.zeros(P)
    mu[0] = -1e4
    flow.set_base(mu, np.full(P, -1.0))
    w0 = flow.weights.copy()
    with pytest.raises(ValueError, match="at iteration 0"):
        flow.fit_elbo(model, num_particles=256, num_iterations=20, learning_rate=1e-2, seed=13)
    assert np.array_equal(flow.weights, w0)


def test_seeds_differing_in_the_upper_half_give_different_streams():
    """Particles (ut_flow_fit_elbo) and minibatches (ut_flow_train) are keyed by both
    32-bit halves of the 64-bit seed: seeds that differ only in the upper half give
    different histories, and the same seed repeats its history exactly."""
    say("64-bit seeds: ELBO particles and training minibatches")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    P = model.num_params
    hist = []
    for seed in (7, 7 + (1 << 32), 7):
        flow = ut.AffineCouplingFlow(P, 0, 256, seed=1)
        hist.append(flow.fit_elbo(model, num_particles=128, num_iterations=5, learning_rate=1e-2,
                                  seed=seed)["elbo_history"])
    assert not np.array_equal(hist[0], hist[1])
    assert np.array_equal(hist[0], hist[2])
    draws = np.random.default_rng(3).normal(size=(4096, P)).astype(np.float32)
    loss = []
    for seed in (9, 9 + (1 << 32), 9):
        flow = ut.AffineCouplingFlow(P, 2, 256, seed=1)
        loss.append(flow.fit(draws, steps=4, batch=256, seed=seed))
    assert not np.array_equal(loss[0], loss[1])
    assert np.array_equal(loss[0], loss[2])
