"""GPU: general static execution plans (ut_plan_create) through the C ABI.

* Plans of the fused families (the configuration models re-expressed as plan
  descriptions: cfg 1 logistic N=1000 D=25, cfg 2 hierarchical 64 x 32, cfg 4 funnel
  D=100, cfg 3's bf16 GLM at N=4096 D=256 with matrix cores allowed) are routed to
  their dedicated kernels: the same engine and bit-identical densities and gradients
  as the ut_plan_* constructors.
* The same descriptions forced onto the plan interpreter, and non-configuration
  models (centred eight schools, linear regression, two models with every family,
  transform and op, a logistic regression with an intercept at N = 8192): density
  and gradient bit-exact against the oracle's kernel-order twin (binary32 and
  binary64), and within 1e-5 (binary32) of the oracle's binary64 reference-semantics
  evaluator; the dedicated kernels and the interpreter agree within 1e-5.
* Samplers on interpreted plans: persistent NUTS (eight schools; the N = 8192 GLM,
  the observation-strided in-kernel gradient), per-chain HMC and lock-step ChEES runs
  bit-exact against the twin: draws, tree depths, divergences, step sizes, adapted
  mass."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent))
import plan_models as PMS  # noqa: E402

import uncertaintea_amd as ut  # noqa: E402
from oracle import binding as B  # noqa: E402
from uncertaintea_amd import models as M  # noqa: E402
from uncertaintea_amd import synthetic as S  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))) /
                 max(1e-300, float(np.max(np.abs(np.asarray(b, np.float64))))))


def _params(P, C, seed, scale=0.4):
    return (np.random.default_rng(seed).normal(size=(P, C)) * scale).astype(np.float32)


@pytest.mark.parametrize("which", ["logistic_cfg1", "hier_cfg2", "funnel_cfg4", "glm_bf16_cfg3"])
def test_fused_families_are_routed_to_their_kernels(which):
    say("plan routing", which)
    if which == "logistic_cfg1":
        pm, (x, y) = PMS.logistic_plan(1000, 25)
        dm, eng = M.LogisticRegression(x, y), "logistic_glm"
    elif which == "hier_cfg2":
        pm, dm = PMS.hier_plan(64, 32, seed=S.DEFAULT_SEED)
        eng = "hierarchical_normal_nc"
    elif which == "funnel_cfg4":
        pm, dm = PMS.funnel_plan(100)
        eng = "funnel"
    else:
        pm, (x, y) = PMS.logistic_plan(4096, 256, cfg=3, x_scale=1.0 / 16.0, matrix_cores=True)
        dm, eng = M.LogisticRegressionBF16(x, y), "logistic_glm_matrix_cores"
    assert pm.engine == eng
    C = 256 if which == "glm_bf16_cfg3" else 64
    q = _params(pm.num_params, C, 1)
    v1, g1 = ut.device_batched_logjoint_gradient(pm, q)
    v2, g2 = ut.device_batched_logjoint_gradient(dm, q)
    assert _bits_equal(v1, v2) and _bits_equal(g1, g2)
    if which != "glm_bf16_cfg3":
        gm = pm.with_flags(generic=True)
        assert gm.engine == "interpreter"
        v3, g3 = ut.device_batched_logjoint_gradient(gm, q)
        say(f"  interpreter vs fused: logjoint {_rel(v3, v1):.2e}, gradient {_rel(g3, g1):.2e}")
        assert _rel(v3, v1) <= 1e-5 and _rel(g3, g1) <= 1e-5


def _interpreted(name):
    if name == "hier_generic":
        return PMS.hier_plan(64, 32, seed=S.DEFAULT_SEED, generic=True)[0]
    if name == "logistic_generic":
        return PMS.logistic_plan(1000, 25, generic=True)[0]
    if name == "funnel_generic":
        return PMS.funnel_plan(100, generic=True)[0]
    if name == "glm_large":
        return PMS.glm_large()[0]
    return getattr(PMS, name)()[0]


@pytest.mark.parametrize("name", ["hier_generic", "logistic_generic", "funnel_generic", "eight_schools_centered",
                                  "linear_regression", "kitchen_sink", "kitchen_sink_more", "glm_large"])
def test_interpreter_density_against_twin_and_reference(name):
    say("plan interpreter density", name)
    pm = _interpreted(name)
    assert pm.engine == "interpreter"
    om = B.OracleModel(pm)
    q = _params(pm.num_params, 96, 2)
    v, g = ut.device_batched_logjoint_gradient(pm, q)
    tv, tg = om.logjoint_grad_twin(q)
    assert _bits_equal(v, tv), name
    assert _bits_equal(g, tg), name
    rv, rg = om.logjoint_grad_f64(q.astype(np.float64))
    say(f"  vs binary64 reference: logjoint {_rel(v, rv):.2e}, gradient {_rel(g, rg):.2e}")
    assert _rel(v, rv) <= 1e-5 and _rel(g, rg) <= 1e-5
    q64 = q.astype(np.float64)
    v64, g64 = ut.device_batched_logjoint_gradient(pm, q64, precision=np.float64)
    tv64, tg64 = om.logjoint_grad_twin(q64, precision=np.float64)
    assert _bits_equal(v64, tv64) and _bits_equal(g64, tg64)


def _compare_runs(dev, orc, label, depth=True):
    same = np.all(dev.unconstrained_samples.view(np.uint32) == orc["samples"].view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[0]
        pytest.fail(f"{label}: first draw mismatch at chain {c} sample {s}")
    if depth:
        assert np.array_equal(dev.tree_depths, orc["tree_depth"]), label
    assert np.array_equal(dev.divergent, orc["divergent"]), label
    assert np.array_equal(dev.step_size, orc["step_size"]), label
    assert np.array_equal(dev.mass_matrix, orc["inverse_mass"]), label


@pytest.mark.parametrize("name,C,W,Sn", [("eight_schools_centered", 64, 150, 60), ("glm_large", 32, 40, 12)])
def test_nuts_on_interpreted_plans_bitexact(name, C, W, Sn):
    say("plan NUTS", name, C, W, Sn)
    pm = _interpreted(name)
    dev = ut.batched_nuts(pm, num_chains=C, num_samples=Sn, num_warmup=W, seed=13)
    orc = B.OracleModel(pm).run("nuts", "twin", C, num_samples=Sn, num_warmup=W, seed=13, threads=8)
    _compare_runs(dev, orc, name)
    say(f"  mean tree depth {dev.tree_depths.mean():.2f}, divergences {int(dev.divergent.sum())}")


def test_hmc_and_chees_on_an_interpreted_plan_bitexact():
    pm = _interpreted("linear_regression")
    say("plan HMC / ChEES linear regression")
    dev = ut.batched_hmc(pm, num_chains=64, num_samples=30, num_warmup=60, num_leapfrog_steps=6, seed=21)
    orc = B.OracleModel(pm).run("hmc", "twin", 64, num_samples=30, num_warmup=60, num_leapfrog_steps=6, seed=21,
                                threads=8)
    _compare_runs(dev, orc, "hmc", depth=False)
    dev = ut.batched_chees(pm, num_chains=64, num_samples=20, num_warmup=60, seed=22)
    orc = B.OracleModel(pm).run("chees", "twin", 64, num_samples=20, num_warmup=60, seed=22, threads=8)
    _compare_runs(dev, orc, "chees", depth=False)
    assert np.array_equal(dev.trajectory_trace, orc["trajectory_trace"])


def test_eight_schools_posterior_nuts_interpreter():
    """The centred eight-schools posterior under the interpreter: mean of mu and
    the school effects near their well-known values (mu ~ 4.4, theta_1 ~ 6)."""
    pm = _interpreted("eight_schools_centered")
    r = ut.batched_nuts(pm, num_chains=256, num_samples=500, num_warmup=500, seed=3, target_accept=0.95)
    c = r.constrained_samples
    say(f"  mu {c[:, :, 0].mean():.2f} tau median {np.median(c[:, :, 1]):.2f} theta1 {c[:, :, 2].mean():.2f} "
        f"div {r.divergence_rate():.4f}")
    assert 3.0 < c[:, :, 0].mean() < 5.8
    assert 4.5 < c[:, :, 2].mean() < 7.5


def test_float32_saturation_edges_on_the_device():
    """The binary32 saturation edges (device_float32_saturation.jl:78-160) on the device: the
    guarded positive families score -Inf with a finite gradient in the subnormal band of
    exp(theta) and past underflow, finite just outside it; exponential stays finite; the beta
    latent's logistic cliff rejects past theta = 18. Bit-exact against the twin at every point."""
    thetas = np.array([-120.0, -100.0, -90.0, -87.0, -80.0, -10.0, 0.5], dtype=np.float32)
    fams = {"gamma": (2.0, 1.0), "lognormal": (0.0, 1.0), "weibull": (2.0, 1.5), "inversegamma": (3.0, 2.0),
            "frechet": (2.0, 1.5), "rayleigh": (1.5,), "inversegaussian": (1.0, 2.0), "exponential": (1.5,)}
    for fam, args in fams.items():
        say("saturation", fam)
        b = ut.PlanBuilder()
        b.latent(fam, *args, transform="log", name="x")
        pm = b.build(f"dfs_{fam}")
        q = thetas.reshape(1, -1)
        v, g = ut.device_batched_logjoint_gradient(pm, q)
        tv, tg = B.OracleModel(pm).logjoint_grad_twin(q)
        assert _bits_equal(v, tv) and _bits_equal(g, tg), fam
        assert not np.isnan(g).any()
        if fam == "exponential":
            assert np.isfinite(v).all()
        else:
            # theta = -80 is finite where the density has no exp(-k log x) or 1/x blow-up
            # (device_float32_saturation.jl:90); frechet's exp(-k log z) overflows binary32 there
            first_finite = 4 if fam in ("gamma", "lognormal", "weibull", "rayleigh") else 5
            assert np.all(v[:3] == -np.inf) and np.isfinite(v[first_finite:]).all(), (fam, v)
    b = ut.PlanBuilder()
    p = b.latent("beta", 2.0, 2.0, transform="logit", name="p")
    b.observe("bernoulli", p, value=1.0, name="y")
    pm = b.build("dfs_beta_bern")
    q = np.array([[-1.5, 0.3, 2.0, 13.0, 16.0, 18.0, 20.0, 30.0]], dtype=np.float32)
    v, g = ut.device_batched_logjoint_gradient(pm, q)
    tv, tg = B.OracleModel(pm).logjoint_grad_twin(q)
    assert _bits_equal(v, tv) and _bits_equal(g, tg)
    assert np.isfinite(v[:5]).all() and np.all(v[5:] == -np.inf) and not np.isnan(g).any()
