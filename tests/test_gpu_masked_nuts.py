"""GPU: NUTS with tree_strategy='masked' (ut_nuts_masked_run).

Bit-exact: for the wavefront-per-chain plans (hierarchical cfg-2 data, logistic
cfg 1, funnel, an interpreted plan) the masked sampler -- one leaf per step for all
chains, gradients from the batched density kernel, the lock-step step search --
reproduces the persistent kernel's runs bit for bit: draws, tree depths, steps,
divergences, accept statistics, energies, step sizes and adapted mass.  That pins
the state-machine restatement of the transition against the persistent kernel,
which the oracle pins.  The conjugate-Gaussian gate and the two-parameter deep-tree
check of the reference's persistent-NUTS tests run on both strategies.

Statistical (the batched-only plans, where no kernel-order twin exists): masked
NUTS on the matrix-core GLM agrees with lock-step ChEES on the same data, and on the
GP with shared-adaptation HMC, posterior means within 5 combined Monte Carlo
standard errors (z printed); the cfg-3 shape itself (N = 131072, D = 256, 4096
chains) runs to finite draws.  With a flow attached (funnel), masked NUTS in the
flow's base space matches the funnel's exact law within 5 MCSE."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
import plan_models as PMS  # noqa: E402

import uncertaintea_amd as ut  # noqa: E402
from uncertaintea_amd import diagnostics as D  # noqa: E402
from uncertaintea_amd import models as M  # noqa: E402
from uncertaintea_amd import synthetic as S  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model(name):
    if name == "hier_cfg2":
        y, off, sig, _ = S.hierarchical_data(64, 32)
        return M.HierarchicalNormalNC(y, off, sig)
    if name == "logistic_cfg1":
        x, y, _ = S.logistic_data(1000, 25)
        return M.LogisticRegression(x, y)
    if name == "funnel_d10":
        return M.NealsFunnel(10, 3.0)
    return PMS.eight_schools_centered()[0]


@pytest.mark.parametrize("name,C,W,Sn", [("hier_cfg2", 64, 150, 50), ("logistic_cfg1", 64, 80, 30),
                                         ("funnel_d10", 32, 120, 40), ("eight_schools", 64, 100, 40)])
def test_masked_equals_persistent_bitexact(name, C, W, Sn):
    say("masked vs persistent", name, C, W, Sn)
    model = _model(name)
    a = ut.batched_nuts(model, num_chains=C, num_samples=Sn, num_warmup=W, seed=17)
    b = ut.batched_nuts(model, num_chains=C, num_samples=Sn, num_warmup=W, seed=17, tree_strategy="masked")
    same = np.all(a.unconstrained_samples.view(np.uint32) == b.unconstrained_samples.view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[0]
        pytest.fail(f"{name}: first draw mismatch at chain {c} sample {s}")
    for f in ("tree_depths", "integration_steps", "divergent", "accepted", "step_size", "mass_matrix",
              "logjoint_values", "acceptance_stats", "energies", "energy_errors"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), (name, f)
    assert a.sampling_gradient_evals == b.sampling_gradient_evals
    say(f"  mean depth {a.tree_depths.mean():.2f}")


def _conjugate(y=0.3):
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 1.0, name="mu")
    b.observe("normal", mu, 1.0, value=y, name="y")
    return b.build("dpnuts_conjugate_gauss")


def _two_param(y=0.7):
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 1.0, name="mu")
    ls = b.latent("normal", 0.0, 0.5, name="log_sigma")
    b.observe("normal", mu, ut.plan.exp(ls), value=y, name="y")
    return b.build("dpnuts_two_param")


def test_conjugate_gaussian_gate_persistent_and_masked():
    """device_persistent_nuts.jl:57-72 on the device: 64 chains, 500 warmup + 1000 draws of
    mu ~ N(0, 1), y ~ N(mu, 1) at y = 0.3 -- posterior mean 0.15 and sd sqrt(0.5) within 0.03,
    R-hat < 1.01, no divergences -- for the persistent kernel and the masked sampler, which
    agree bit for bit."""
    pm = _conjugate()
    a = ut.batched_nuts(pm, num_chains=64, num_samples=1000, num_warmup=500, seed=511)
    b = ut.batched_nuts(pm, num_chains=64, num_samples=1000, num_warmup=500, seed=511, tree_strategy="masked")
    assert np.array_equal(a.unconstrained_samples.view(np.uint32), b.unconstrained_samples.view(np.uint32))
    d = a.unconstrained_samples[..., 0].astype(np.float64)
    say(f"  conjugate mean {d.mean():.4f} sd {d.std(ddof=1):.4f} rhat {float(np.max(D.rhat(a.unconstrained_samples))):.4f}")
    assert np.all(np.isfinite(d))
    assert abs(d.mean() - 0.15) < 0.03 and abs(d.std(ddof=1) - np.sqrt(0.5)) < 0.03
    assert float(np.max(np.asarray(D.rhat(a.unconstrained_samples)))) < 1.01
    assert not a.divergent.any()


def test_two_parameter_model_deep_trees_device():
    """device_persistent_nuts.jl:99-120 on the device: depth >= 3 reached, R-hat < 1.01,
    divergence rate < 1 %, persistent and masked bit-identical."""
    pm = _two_param()
    a = ut.batched_nuts(pm, num_chains=64, num_samples=1000, num_warmup=500, seed=2)
    b = ut.batched_nuts(pm, num_chains=64, num_samples=1000, num_warmup=500, seed=2, tree_strategy="masked")
    assert np.array_equal(a.unconstrained_samples.view(np.uint32), b.unconstrained_samples.view(np.uint32))
    assert np.array_equal(a.tree_depths, b.tree_depths)
    say(f"  two-param max depth {a.tree_depths.max()} div {a.divergence_rate():.4f}")
    assert np.all(np.isfinite(a.unconstrained_samples))
    assert float(np.max(np.asarray(D.rhat(a.unconstrained_samples)))) < 1.01
    assert a.divergence_rate() < 0.01
    assert a.tree_depths.max() >= 3


def _zmeans(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / np.sqrt(np.asarray(D.mcse_mean(a)) ** 2 +
                                                                    np.asarray(D.mcse_mean(b)) ** 2)
    return float(np.nanmax(z))


def test_masked_nuts_matrix_core_glm_against_chees():
    x, y, _ = S.logistic_data(4096, 256, seed=31, cfg=3, x_scale=1.0 / 16.0)
    model = M.LogisticRegressionBF16(x, y)
    n = ut.batched_nuts(model, num_chains=256, num_samples=200, num_warmup=200, seed=3, tree_strategy="masked")
    c = ut.batched_chees(model, num_chains=256, num_samples=200, num_warmup=200, seed=4)
    z = _zmeans(n.constrained_samples, c.constrained_samples)
    rh = float(np.nanmax(D.rhat(n.constrained_samples)))
    say(f"  masked NUTS vs ChEES (bf16 GLM N=4096 D=256): max |z| {z:.2f}, rhat {rh:.4f}, depth "
        f"{n.tree_depths.mean():.2f}, divergences {int(n.divergent.sum())}")
    assert z < 5.0 and rh < 1.05 and not n.divergent.any()


def test_masked_nuts_runs_at_the_large_glm_shape():
    x, y, _ = S.logistic_data(131072, 256, seed=20261203, cfg=3, x_scale=1.0 / 16.0)
    model = M.LogisticRegressionBF16(x, y)
    r = ut.batched_nuts(model, num_chains=4096, num_samples=5, num_warmup=10, seed=5, tree_strategy="masked",
                        max_tree_depth=6)
    say(f"  N=131072 D=256 4096 chains: depth {r.tree_depths.mean():.2f}, step {np.median(r.step_size):.4f}, "
        f"sampling {r.sampling_seconds:.2f}s, gradient evals {r.sampling_gradient_evals}")
    assert np.isfinite(r.unconstrained_samples).all() and r.sampling_gradient_evals > 0


def test_masked_nuts_gp_against_shared_hmc():
    X, yc, _ = S.gp_poisson_data(32, 2, seed=3)
    model = M.GaussianProcessPoisson(X, yc)
    init = np.zeros((model.num_params, 128))
    init[2] = np.log(0.674)
    n = ut.batched_nuts(model, num_chains=128, num_samples=150, num_warmup=150, seed=6, tree_strategy="masked",
                        initial_params=init)
    h = ut.batched_hmc(model, num_chains=128, num_samples=300, num_warmup=300, seed=7, per_chain_adaptation=False,
                       initial_params=init, num_leapfrog_steps=12)
    z = _zmeans(n.constrained_samples, h.constrained_samples)
    say(f"  masked NUTS vs shared HMC (GP N=32): max |z| {z:.2f}, depth {n.tree_depths.mean():.2f}")
    assert z < 5.0


def test_masked_nuts_with_a_flow_on_the_funnel():
    model = M.NealsFunnel(10, 3.0)
    flow = ut.AffineCouplingFlow(model.num_params, 4, 256, seed=2)
    flow.fit_elbo(model, num_particles=1024, num_iterations=3000, learning_rate=1e-3, seed=3)
    r = ut.batched_nuts(model, num_chains=256, num_samples=400, num_warmup=200, seed=8, tree_strategy="masked",
                        flow=flow)
    d = np.asarray(r.constrained_samples, np.float64)
    u = d.copy()
    u[..., 1:] *= np.exp(-0.5 * d[..., :1])
    zm = np.abs(u.mean(axis=(0, 1))) / np.asarray(D.mcse_mean(u))
    zv = np.abs((u ** 2).mean(axis=(0, 1)) - np.concatenate([[9.0], np.ones(10)])) / np.asarray(D.mcse_mean(u ** 2))
    say(f"  flow NUTS funnel: v sd {d[..., 0].std():.3f}; max |z| mean {zm.max():.2f} var {zv.max():.2f}")
    assert zm.max() < 5.0 and zv.max() < 5.0
