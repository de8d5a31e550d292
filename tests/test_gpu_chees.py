"""GPU: lock-step ChEES and the matrix-core GLM.

Bit-exact: whole ChEES runs (draws, step counts, acceptance, divergences, step
size, adapted mass, trajectory-time trace) of the wavefront-per-chain models
against the oracle's binary32 twin.  Tolerance: the matrix-core GLM logjoint and
gradient against the binary64 reference formulas on the same bf16 design matrix,
errors printed and bounded: logjoint |d| <= 1e-6 |lj|; gradient normwise
|d| <= 1e-5 |g| and componentwise |d_d| <= 2e-6 (1 + sum_i |x_id|).  Measured on
MI355X: 2.5e-7, 1.7e-6 and 3.7e-7.  The same bounds hold at configuration 3's own
shape (N = 131072, D = 256, 4096 chains through the split-K layout), checked on a
128-chain slice from both ends of the batch against the binary64 oracle."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import _lib as L
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _models():
    x, y, _ = S.logistic_data(1000, 25)
    yy, off, sig, _ = S.hierarchical_data(8, 8, seed=3)
    return {
        "logistic_cfg1": M.LogisticRegression(x, y),
        "hier_small": M.HierarchicalNormalNC(yy, off, sig),
        "funnel_d10": M.NealsFunnel(10, 3.0),
    }


def _compare(dev, orc, label):
    s_d = dev.unconstrained_samples
    same = np.all(s_d.view(np.uint32) == orc["samples"].view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[0]
        pytest.fail(f"{label}: first draw mismatch at chain {c} sample {s}")
    assert np.array_equal(dev.trajectory_trace, orc["trajectory_trace"]), label
    assert np.array_equal(dev.integration_steps, orc["integration_steps"]), label
    assert np.array_equal(dev.divergent, orc["divergent"]), label
    assert np.array_equal(dev.accepted, orc["accepted"]), label
    assert np.array_equal(dev.step_size, orc["step_size"]), label
    assert np.array_equal(dev.mass_matrix, orc["inverse_mass"]), label
    assert np.array_equal(dev.logjoint_values, orc["logjoint"]), label
    assert np.array_equal(dev.acceptance_stats, orc["accept_stat"]), label
    assert np.array_equal(dev.energies, orc["energy"]), label
    assert np.array_equal(dev.energy_errors, orc["energy_error"]), label


@pytest.mark.parametrize("name,C,W,Sn,fwd", [("logistic_cfg1", 64, 120, 30, True), ("hier_small", 64, 80, 30, False),
                                             ("funnel_d10", 32, 60, 20, False)])
def test_chees_bitexact_against_twin(name, C, W, Sn, fwd):
    say("chees bitexact", name, C, W, Sn, fwd)
    model = _models()[name]
    dev = ut.batched_chees(model, num_chains=C, num_samples=Sn, num_warmup=W, seed=11, forward_velocity=fwd)
    orc = B.OracleModel(model).run("chees", "twin", C, num_samples=Sn, num_warmup=W, seed=11, threads=8,
                                   chees={"forward_velocity": fwd})
    _compare(dev, orc, name)
    assert dev.integration_steps.max() >= 1


def _glm(n, seed=20261203):
    x, y, _ = S.logistic_data(n, 256, seed=seed, cfg=3, x_scale=1.0 / 16.0)
    return M.LogisticRegressionBF16(x, y)


@pytest.mark.parametrize("n,C", [(2048, 128), (4096, 256)])
def test_glm_mfma_density_accuracy(n, C):
    say("glm mfma density", n, C)
    model = _glm(n)
    om = B.OracleModel(model)
    rng = np.random.default_rng(5)
    params = rng.normal(size=(256, C)).astype(np.float32)
    v, g = ut.device_batched_logjoint_gradient(model, params)
    rv, rg = om.logjoint_grad_f64(params.astype(np.float64))
    lj_rel = np.abs(v - rv) / np.abs(rv)
    scale = np.abs(model.x).sum(axis=0)[:, None] + 1.0  # sum_i |x_id| bounds |d grad_d / d r|
    g_err = np.abs(g - rg) / scale
    say(f"  logjoint rel err max {lj_rel.max():.3e}; gradient err / sum|x| max {g_err.max():.3e}; "
        f"normwise {np.linalg.norm(g - rg) / np.linalg.norm(rg):.3e}")
    assert lj_rel.max() <= 1e-6
    assert g_err.max() <= 2e-6
    assert np.linalg.norm(g - rg) <= 1e-5 * np.linalg.norm(rg)
    # logjoint-only call agrees
    assert np.array_equal(ut.device_batched_logjoint(model, params), v)


def test_glm_mfma_density_cfg3_full_shape():
    say("glm mfma density cfg3 N=131072 D=256 C=4096")
    model = _glm(131072)
    om = B.OracleModel(model)
    params = np.random.default_rng(9).normal(size=(256, 4096)).astype(np.float32)
    v, g = ut.device_batched_logjoint_gradient(model, params)
    cols = np.r_[0:64, 4032:4096]
    rv, rg = om.logjoint_grad_f64(params[:, cols].astype(np.float64))
    lj_rel = np.abs(v[cols] - rv) / np.abs(rv)
    scale = np.abs(model.x).sum(axis=0)[:, None] + 1.0
    g_err = np.abs(g[:, cols] - rg) / scale
    nw = np.linalg.norm(g[:, cols] - rg) / np.linalg.norm(rg)
    say(f"  logjoint rel err max {lj_rel.max():.3e}; gradient err / sum|x| max {g_err.max():.3e}; normwise {nw:.3e}")
    assert np.all(np.isfinite(v)) and np.all(np.isfinite(g))
    assert lj_rel.max() <= 1e-6 and g_err.max() <= 2e-6 and nw <= 1e-5


def test_glm_mfma_argument_errors():
    say("glm mfma errors")
    model = _glm(2048)
    with pytest.raises(ValueError):
        ut.DeviceBatchedWorkspace(model, 100)  # chains not a multiple of 128
    with pytest.raises(ValueError):
        ut.batched_nuts(model, num_chains=128, num_samples=1)
    bad = M.LogisticRegressionBF16(np.zeros((100, 256), np.float32), np.zeros(100, np.float32))
    with pytest.raises(ValueError):
        _ = bad.plan


def test_glm_mfma_chees_posterior():
    """ChEES on the matrix-core GLM against NUTS on the same bf16 data through
    the wavefront-per-chain model: posterior means agree within MCSE."""
    say("glm mfma chees posterior")
    n = 2048
    mf = _glm(n)
    wave = M.LogisticRegression(mf.x, mf.y)
    ch = ut.batched_chees(mf, num_chains=256, num_samples=300, num_warmup=300, seed=3, forward_velocity=True)
    nu = ut.batched_nuts(wave, num_chains=256, num_samples=200, num_warmup=300, seed=4)
    a, b = ch.unconstrained_samples.astype(np.float64), nu.unconstrained_samples.astype(np.float64)
    rh = float(np.max(np.asarray(D.rhat(a))))
    se = np.sqrt(a.var(axis=(0, 1)) / np.asarray(D.ess_bulk(a)) + b.var(axis=(0, 1)) / np.asarray(D.ess_bulk(b)))
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / se
    say(f"  chees rhat {rh:.4f} max z {z.max():.2f} divergent {ch.divergence_rate():.4f} "
        f"T {ch.trajectory_trace[-1]:.3f} eps {ch.step_size[0]:.4f}")
    assert rh < 1.05
    assert z.max() < 5.0
