"""Result interchange (no GPU): the ArviZ-style dictionary's groups, keys, axis
layouts and vector grouping (src/inference/export.jl:187-272), and the cross-PPL
NPZ + JSON sidecar files read back the way the cross-PPL analysis reads them
(bench/crossppl/python/analyze.py:40-44)."""
import json

import numpy as np
import pytest

import uncertaintea_amd as ut
from uncertaintea_amd.inference import HMCChains


def _chains(C=3, S=5, J=2):
    y = np.zeros(4 * J, dtype=np.float32)
    model = ut.HierarchicalNormalNC(y, np.arange(0, 4 * J + 1, 4, dtype=np.int32), np.ones(J))
    P = model.num_params
    rng = np.random.default_rng(0)
    u = rng.normal(size=(C, S, P)).astype(np.float32)
    return HMCChains(sampler="nuts", model=model, unconstrained_samples=u,
                     logjoint_values=rng.normal(size=(C, S)), acceptance_stats=rng.uniform(size=(C, S)),
                     energies=rng.normal(size=(C, S)), energy_errors=rng.normal(size=(C, S)),
                     accepted=np.ones((C, S), bool), divergent=rng.uniform(size=(C, S)) < 0.2,
                     tree_depths=rng.integers(1, 5, size=(C, S)), integration_steps=rng.integers(1, 31, size=(C, S)),
                     step_size=np.array([0.1, 0.2, 0.3][:C]), mass_matrix=np.ones((C, P)), max_tree_depth=10,
                     num_leapfrog_steps=0, target_accept=0.8, warmup_seconds=1.5, sampling_seconds=2.5,
                     sampling_gradient_evals=100)


def test_arviz_dict_flat_and_layouts():
    ch = _chains()
    d = ut.to_arviz_dict(ch)
    assert set(d) == {"posterior", "sample_stats", "attrs"}
    assert list(d["posterior"]) == ["mu", "tau", "theta_raw[1]", "theta_raw[2]"]
    # draw_chain: (num_samples, num_chains); tau is constrained (exp of log tau)
    assert d["posterior"]["tau"].shape == (5, 3)
    assert np.allclose(d["posterior"]["tau"], np.exp(ch.unconstrained_samples[:, :, 1].astype(np.float64)).T)
    assert set(d["sample_stats"]) == {"diverging", "energy", "tree_depth", "acceptance_rate", "lp", "step_size",
                                      "n_steps"}
    assert np.array_equal(d["sample_stats"]["step_size"][:, 2], np.full(5, 0.3))
    assert np.array_equal(d["sample_stats"]["n_steps"], ch.integration_steps.T)
    assert d["attrs"]["layout"] == "draw_chain"
    e = ut.to_arviz_dict(ch, layout="chain_draw")
    for k in d["posterior"]:
        assert np.array_equal(e["posterior"][k], d["posterior"][k].T)
    assert np.array_equal(e["sample_stats"]["diverging"], ch.divergent)
    with pytest.raises(ValueError, match="layout"):
        ut.to_arviz_dict(ch, layout="draws")


def test_arviz_dict_grouped_vectors():
    ch = _chains()
    d = ut.to_arviz_dict(ch, layout="chain_draw", flatten_vectors=False)
    assert list(d["posterior"]) == ["mu", "tau", "theta_raw"]
    assert d["posterior"]["theta_raw"].shape == (3, 5, 2)
    assert np.array_equal(d["posterior"]["theta_raw"], ch.unconstrained_samples[:, :, 2:].astype(np.float64))
    assert d["dims"] == {"theta_raw": ["theta_raw_dim_0"]}
    assert d["coords"] == {"theta_raw_dim_0": [1, 2]}


def test_posterior_array_and_crossppl_files(tmp_path):
    ch = _chains()
    pa = ut.posterior_array(ch)
    assert pa.shape == (5, 3, 4) and pa.dtype == np.float64
    stem = ut.write_crossppl_run(ch, str(tmp_path), "hier", "hip-nuts", num_warmup=7, seed=11, rep=2)
    assert stem.endswith("hier__uncertaintea-hip-nuts__chains3__seed11__rep2")
    meta = json.load(open(stem + ".json"))
    draws = np.load(stem + ".npz")["draws"]
    assert draws.shape == (3, 5, 4) and np.array_equal(draws, np.transpose(pa, (1, 0, 2)))
    assert meta["params"] == ut.parameter_names(ch) and meta["num_warmup"] == 7 and meta["seed"] == 11
    assert meta["timings"]["sampling_s"] == 2.5 and meta["timings"]["total_s"] == 4.0
    assert meta["framework"] == "uncertaintea" and meta["sampler"] == {"kind": "nuts"}
