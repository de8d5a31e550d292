"""GPU: MEADS against the oracle's binary32 twin, bit for bit (draws, acceptance,
divergences, energies, skipped-fold records, the per-iteration adaptation trace),
and its posterior against NUTS on the same model."""
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


def _models():
    x, y, _ = S.logistic_data(1000, 25)
    yy, off, sig, _ = S.hierarchical_data(8, 8, seed=3)
    return {
        "logistic_cfg1": M.LogisticRegression(x, y),
        "hier_small": M.HierarchicalNormalNC(yy, off, sig),
        "funnel_d10": M.NealsFunnel(10, 3.0),
    }


@pytest.mark.parametrize("name,C,K,W,Sn", [("logistic_cfg1", 64, 4, 30, 40), ("hier_small", 64, 4, 20, 40),
                                           ("funnel_d10", 66, 3, 10, 30)])
def test_meads_bitexact_against_twin(name, C, K, W, Sn):
    say("meads bitexact", name, C, K, W, Sn)
    model = _models()[name]
    dev = ut.batched_meads(model, num_chains=C, num_samples=Sn, num_warmup=W, num_folds=K, seed=5)
    orc = B.OracleModel(model).run("meads", "twin", C, num_samples=Sn, num_warmup=W, num_folds=K, seed=5, threads=8)
    same = np.all(dev.unconstrained_samples.view(np.uint32) == orc["samples"].view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[0]
        pytest.fail(f"{name}: first draw mismatch at chain {c} sample {s}")
    assert np.array_equal(dev.trajectory_trace, orc["trajectory_trace"]), name
    for k_dev, k_orc in (("accepted", "accepted"), ("divergent", "divergent"), ("logjoint_values", "logjoint"),
                         ("acceptance_stats", "accept_stat"), ("energies", "energy"),
                         ("energy_errors", "energy_error"), ("integration_steps", "integration_steps"),
                         ("step_size", "step_size"), ("mass_matrix", "inverse_mass")):
        assert np.array_equal(getattr(dev, k_dev), orc[k_orc]), (name, k_dev)


def test_meads_posterior_matches_nuts():
    say("meads posterior")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    # MEADS has no adaptation phase: discard the transient from the prior draws
    me = ut.batched_meads(model, num_chains=256, num_samples=1500, num_warmup=1500, seed=7)
    nu = ut.batched_nuts(model, num_chains=256, num_samples=300, num_warmup=300, seed=8)
    a, b = me.unconstrained_samples.astype(np.float64), nu.unconstrained_samples.astype(np.float64)
    rh = float(np.max(np.asarray(D.rhat(a))))
    se = np.sqrt(a.var(axis=(0, 1)) / np.asarray(D.ess_bulk(a)) + b.var(axis=(0, 1)) / np.asarray(D.ess_bulk(b)))
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / se
    say(f"  meads rhat {rh:.4f} max z {z.max():.2f} step {me.step_size[0]:.4f} div {me.divergence_rate():.4f}")
    assert rh < 1.02 and z.max() < 5.0
