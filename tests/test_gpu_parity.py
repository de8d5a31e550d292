"""GPU parity: the gfx950 kernels against the CPU oracle on identical inputs.

Bit-exact: Philox blocks and conversions, the binary32 and binary64 density
kernels against the kernel-order twin, and whole sampler runs (draws, NUTS tree
depths, divergences, step sizes, adapted mass).  Tolerance: binary32 densities
against the binary64 reference formulas (|dlogjoint| <= 1e-5 |logjoint| + 1e-4,
|dgrad| <= 1e-5 max(1, |grad|_inf))."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_visible():
    say("device check")
    assert ut.device_available()


@pytest.mark.parametrize("coord", [(0, 0, 0, 0), (5, 123456, 2, 17), (0xFFFFFFFF, 0x80000000, 6, 4000000000)])
def test_philox_device_bitexact(coord):
    say("philox", coord)
    chain, it, stream, first = coord
    d = ut.philox_fill(chain, it, stream, first, 20000)
    h = B.philox_fill(chain, it, stream, first, 20000)
    for a, b in zip(d, h):
        assert np.array_equal(a.view(np.uint32) if a.dtype != np.float64 else a.view(np.uint64),
                              b.view(np.uint32) if b.dtype != np.float64 else b.view(np.uint64))


def _models():
    x, y, _ = S.logistic_data(1000, 25)
    yy, off, sig, _ = S.hierarchical_data(64, 32)
    ragged = np.array([0, 3, 3, 8, 20, 21], np.int32)
    return {
        "logistic_cfg1": M.LogisticRegression(x, y),
        "logistic_d70": M.LogisticRegression(*S.logistic_data(300, 70)[:2]),
        "hier_cfg2": M.HierarchicalNormalNC(yy, off, sig),
        "hier_ragged": M.HierarchicalNormalNC(np.linspace(-2, 2, 21).astype(np.float32), ragged,
                                              np.array([0.5, 1.0, 1.5, 2.0, 0.7])),
        "funnel_cfg4": M.NealsFunnel(100, 3.0),
        "funnel_d5": M.NealsFunnel(5, 3.0),
    }


@pytest.mark.parametrize("name", list(_models().keys()))
def test_density_parity(name):
    say("density", name)
    model = _models()[name]
    om = B.OracleModel(model)
    rng = np.random.default_rng(7)
    P, C = model.num_params, 96
    params = rng.normal(size=(P, C)) * 0.5
    # binary32: bit-exact vs the twin, tolerance vs binary64 reference formulas
    v32, g32 = ut.device_batched_logjoint_gradient(model, params.astype(np.float32), precision=np.float32)
    tv, tg = om.logjoint_grad_twin(params.astype(np.float32), np.float32)
    assert np.array_equal(v32.view(np.uint32), tv.view(np.uint32))
    assert np.array_equal(np.ascontiguousarray(g32).view(np.uint32), np.ascontiguousarray(tg).view(np.uint32))
    rv, rg = om.logjoint_grad_f64(params.astype(np.float32).astype(np.float64))
    assert np.all(np.abs(v32 - rv) <= 1e-5 * np.abs(rv) + 1e-4)
    assert np.max(np.abs(g32 - rg)) <= 1e-5 * max(1.0, np.max(np.abs(rg)))
    # binary64: bit-exact vs the twin, 1e-12 vs reference formulas
    v64, g64 = ut.device_batched_logjoint_gradient(model, params, precision=np.float64)
    tv64, tg64 = om.logjoint_grad_twin(params, np.float64)
    assert np.array_equal(v64, tv64) and np.array_equal(g64, tg64)
    rv, rg = om.logjoint_grad_f64(params)
    assert np.allclose(v64, rv, rtol=1e-12, atol=1e-10)
    assert np.max(np.abs(g64 - rg)) <= 1e-11 * max(1.0, np.max(np.abs(rg)))
    # logjoint-only entry point agrees with the gradient entry point
    assert np.array_equal(ut.device_batched_logjoint(model, params.astype(np.float32)), v32)


def _compare_runs(dev, orc, label):
    s_d = dev.unconstrained_samples
    s_o = orc["samples"]
    same = np.all(s_d.view(np.uint32) == s_o.view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[np.lexsort(np.argwhere(~same).T[::-1])][0]
        pytest.fail(f"{label}: first draw mismatch at chain {c} sample {s}; "
                    f"step dev {dev.step_size[c]!r} orc {orc['step_size'][c]!r}")
    assert np.array_equal(dev.tree_depths, orc["tree_depth"]), label
    assert np.array_equal(dev.integration_steps, orc["integration_steps"]), label
    assert np.array_equal(dev.divergent, orc["divergent"]), label
    assert np.array_equal(dev.accepted, orc["accepted"]), label
    assert np.array_equal(dev.step_size, orc["step_size"]), label
    assert np.array_equal(dev.mass_matrix, orc["inverse_mass"]), label
    assert np.array_equal(dev.logjoint_values, orc["logjoint"]), label
    assert np.array_equal(dev.acceptance_stats, orc["accept_stat"]), label
    assert np.array_equal(dev.energies, orc["energy"]), label


@pytest.mark.parametrize("name,C,W,Sn", [("hier_cfg2", 64, 60, 30), ("hier_ragged", 32, 40, 20),
                                         ("funnel_d5", 32, 40, 20)])
def test_nuts_bitexact_against_twin(name, C, W, Sn):
    say("nuts bitexact", name, C, W, Sn)
    model = _models()[name]
    dev = ut.batched_nuts(model, num_chains=C, num_samples=Sn, num_warmup=W, seed=42)
    orc = B.OracleModel(model).run("nuts", "twin", C, num_samples=Sn, num_warmup=W, seed=42, threads=8)
    _compare_runs(dev, orc, name)
    assert dev.tree_depths.max() >= 2


@pytest.mark.parametrize("name,C,W,Sn,L", [("logistic_cfg1", 64, 60, 30, 8), ("funnel_cfg4", 64, 40, 20, 10)])
def test_hmc_bitexact_against_twin(name, C, W, Sn, L):
    say("hmc bitexact", name, C, W, Sn, L)
    model = _models()[name]
    dev = ut.batched_hmc(model, num_chains=C, num_samples=Sn, num_warmup=W, num_leapfrog_steps=L,
                         step_size=0.05, seed=7)
    orc = B.OracleModel(model).run("hmc", "twin", C, num_samples=Sn, num_warmup=W, num_leapfrog_steps=L,
                                   step_size=0.05, seed=7, threads=8)
    _compare_runs(dev, orc, name)


def test_nuts_explicit_init_and_uniform_init():
    say("nuts init variants")
    model = _models()["hier_cfg2"]
    init = np.zeros((model.num_params, 16))
    dev = ut.batched_nuts(model, num_chains=16, num_samples=5, num_warmup=10, initial_params=init, seed=3)
    orc = B.OracleModel(model).run("nuts", "twin", 16, num_samples=5, num_warmup=10, initial_params=init, seed=3)
    _compare_runs(dev, orc, "explicit init")
    dev = ut.batched_nuts(model, num_chains=16, num_samples=5, num_warmup=10, init_strategy="uniform", seed=4)
    orc = B.OracleModel(model).run("nuts", "twin", 16, num_samples=5, num_warmup=10, init_strategy="uniform", seed=4)
    _compare_runs(dev, orc, "uniform init")


def test_nuts_seed_determinism():
    say("nuts determinism")
    model = _models()["hier_ragged"]
    a = ut.batched_nuts(model, num_chains=8, num_samples=10, num_warmup=10, seed=1)
    b = ut.batched_nuts(model, num_chains=8, num_samples=10, num_warmup=10, seed=1)
    c = ut.batched_nuts(model, num_chains=8, num_samples=10, num_warmup=10, seed=2)
    assert np.array_equal(a.unconstrained_samples, b.unconstrained_samples)
    assert not np.array_equal(a.unconstrained_samples, c.unconstrained_samples)


def test_nonfinite_initial_params_error():
    say("nonfinite init")
    model = _models()["funnel_d5"]
    init = np.full((6, 4), np.nan)
    with pytest.raises(ut.NonFiniteInitError):
        ut.batched_nuts(model, num_chains=4, num_samples=2, num_warmup=0, initial_params=init)


def test_dimension_errors():
    say("dimension errors")
    model = _models()["funnel_d5"]
    ws = ut.DeviceBatchedWorkspace(model, 4)
    with pytest.raises(ValueError):
        ut.device_batched_logjoint_inplace(ws, np.zeros((5, 4), np.float32))
    with pytest.raises(ValueError):
        ut.device_batched_logjoint_inplace(ws, np.zeros((6, 3), np.float32))


def test_hier_cfg2_posterior_statistics():
    """Full 1024-chain config-2 run against the binary64 oracle (different seed):
    posterior means within 4 combined MCSE, R-hat < 1.01, no divergences.  The
    non-centred posterior couples tau with every theta_raw, so mu and tau mix slowly
    (bulk ESS ~ 0.05 per draw); 2000 draws per chain bring R-hat under 1.01."""
    say("hier cfg2 statistics")
    model = _models()["hier_cfg2"]
    dev = ut.batched_nuts(model, num_chains=1024, num_samples=2000, num_warmup=1000, seed=2026)
    d = dev.constrained_samples
    rh = ut.diagnostics.rhat(d)
    assert np.nanmax(rh) < 1.01
    assert dev.divergence_rate() < 0.01
    orc = B.OracleModel(model).run("nuts", "f64", 64, num_samples=400, num_warmup=400, seed=77, threads=8)
    o = model.constrain(orc["samples"])
    for p in (0, 1, 2, 40):
        m1, m2 = d[..., p].mean(), o[..., p].mean()
        se = np.hypot(ut.diagnostics.mcse_mean(d[..., p]), ut.diagnostics.mcse_mean(o[..., p]))
        assert abs(m1 - m2) < 4 * se + 1e-3, (p, m1, m2, se)
