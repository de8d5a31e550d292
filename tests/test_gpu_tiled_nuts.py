"""GPU parity of the workgroup-per-chain layout: one chain's logistic rows split over
the W wavefronts of its workgroup (W = 2, 4, 8), every wavefront running the same
transition on the same values.

Bit-exact against the oracle twin of the same summation order: the density kernel
in binary32 and binary64, and whole persistent NUTS and HMC runs (draws, tree depths,
step sizes, adapted mass, energies) on logistic regression with N = 8192 rows.  Also:
the automatic layout choice, the errors of invalid layouts, and the run time of the
few-chains / large-N case against the one-wavefront layout."""
import sys
import time

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S
from uncertaintea_amd.device import DeviceBatchedWorkspace, device_batched_logjoint_gradient_inplace

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _logistic(n, d, seed=20261204):
    x, y, _ = S.logistic_data(n, d, seed=seed)
    return M.LogisticRegression(x, y)


@pytest.mark.parametrize("W,n,d", [(2, 8192, 16), (4, 8192, 16), (8, 8192, 16), (4, 4100, 70), (8, 3001, 130)])
def test_tiled_density_bitexact(W, n, d):
    say("tiled density", W, n, d)
    model = _logistic(n, d)
    om = B.OracleModel(model, waves_per_chain=W)
    rng = np.random.default_rng(W * 1000 + d)
    C = 24
    params = rng.normal(size=(d, C)) * 0.3
    for dt in (np.float32, np.float64):
        ws = DeviceBatchedWorkspace(model, C, dt, waves_per_chain=W)
        assert ws.waves_per_chain == W
        p = params.astype(dt)
        v, g = device_batched_logjoint_gradient_inplace(ws, p)
        tv, tg = om.logjoint_grad_twin(p, dt)
        assert np.array_equal(v, tv), dt
        assert np.array_equal(np.ascontiguousarray(g), np.ascontiguousarray(tg)), dt
    rv, rg = om.logjoint_grad_f64(params.astype(np.float32).astype(np.float64))
    v32 = device_batched_logjoint_gradient_inplace(DeviceBatchedWorkspace(model, C, np.float32, waves_per_chain=W),
                                                   params.astype(np.float32))[0]
    assert np.all(np.abs(v32 - rv) <= 1e-5 * np.abs(rv) + 1e-4)


def _compare(dev, orc, label):
    s_d, s_o = dev.unconstrained_samples, orc["samples"]
    same = np.all(s_d.view(np.uint32) == s_o.view(np.uint32), axis=2)
    if not same.all():
        c, s = np.argwhere(~same)[0]
        pytest.fail(f"{label}: first draw mismatch at chain {c} sample {s}")
    for a, b in [("tree_depths", "tree_depth"), ("integration_steps", "integration_steps"),
                 ("divergent", "divergent"), ("accepted", "accepted"), ("step_size", "step_size"),
                 ("mass_matrix", "inverse_mass"), ("logjoint_values", "logjoint"),
                 ("acceptance_stats", "accept_stat"), ("energies", "energy")]:
        assert np.array_equal(getattr(dev, a), orc[b]), (label, a)


@pytest.mark.parametrize("W", [4, 8])
def test_tiled_nuts_bitexact_against_twin(W):
    say("tiled nuts", W)
    model = _logistic(8192, 16)
    C, Wu, Sn = 16, 40, 20
    dev = ut.batched_nuts(model, num_chains=C, num_samples=Sn, num_warmup=Wu, seed=21, waves_per_chain=W)
    orc = B.OracleModel(model, waves_per_chain=W).run("nuts", "twin", C, num_samples=Sn, num_warmup=Wu, seed=21,
                                                       threads=8)
    _compare(dev, orc, f"nuts W={W}")
    assert dev.tree_depths.max() >= 2


def test_tiled_nuts_two_slots_bitexact():
    say("tiled nuts K=2")
    model = _logistic(4100, 70)
    dev = ut.batched_nuts(model, num_chains=8, num_samples=10, num_warmup=30, seed=22, waves_per_chain=4)
    orc = B.OracleModel(model, waves_per_chain=4).run("nuts", "twin", 8, num_samples=10, num_warmup=30, seed=22,
                                                       threads=8)
    _compare(dev, orc, "nuts K=2 W=4")


def test_tiled_hmc_cfg1_shape_bitexact():
    """The reference's CPU-runnable case (N = 1000, D = 25, batched_hmc with 64 chains,
    8 leapfrog steps, step 0.05) with 8 wavefronts per chain, shortened to 100 + 50."""
    say("tiled hmc cfg1 shape")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    kw = dict(num_samples=50, num_warmup=100, num_leapfrog_steps=8, step_size=0.05, seed=25)
    dev = ut.batched_hmc(model, num_chains=64, waves_per_chain=8, **kw)
    orc = B.OracleModel(model, waves_per_chain=8).run("hmc", "twin", 64, threads=8, **kw)
    _compare(dev, orc, "hmc cfg1 W=8")


def test_tiled_hmc_bitexact_against_twin():
    say("tiled hmc")
    model = _logistic(8192, 16)
    dev = ut.batched_hmc(model, num_chains=16, num_samples=20, num_warmup=30, num_leapfrog_steps=6, step_size=0.02,
                         seed=23, waves_per_chain=2)
    orc = B.OracleModel(model, waves_per_chain=2).run("hmc", "twin", 16, num_samples=20, num_warmup=30,
                                                       num_leapfrog_steps=6, step_size=0.02, seed=23, threads=8)
    _compare(dev, orc, "hmc W=2")


def test_chain_layout_choice_and_errors():
    say("layout choice")
    big = _logistic(8192, 16)
    assert DeviceBatchedWorkspace(big, 64).waves_per_chain == 8
    assert DeviceBatchedWorkspace(big, 512).waves_per_chain == 8
    assert DeviceBatchedWorkspace(big, 1024).waves_per_chain == 4
    assert DeviceBatchedWorkspace(big, 4096).waves_per_chain == 1
    # below 2048 rows the automatic layout stays at one wavefront per chain
    assert DeviceBatchedWorkspace(_logistic(1000, 25), 16).waves_per_chain == 1
    assert DeviceBatchedWorkspace(M.NealsFunnel(100, 3.0), 16).waves_per_chain == 1
    ws = DeviceBatchedWorkspace(big, 64)
    with pytest.raises(ValueError):
        ws.set_waves_per_chain(3)
    with pytest.raises(ValueError):
        DeviceBatchedWorkspace(M.NealsFunnel(100, 3.0), 16, waves_per_chain=2)
    ws.set_waves_per_chain(1)
    assert ws.waves_per_chain == 1
    # D = 256 (four slots per lane): eight wavefronts' checkpoint stacks exceed the LDS
    wide = _logistic(4096, 256)
    with pytest.raises(ValueError):
        ut.batched_nuts(wide, num_chains=4, num_samples=2, num_warmup=2, seed=1, waves_per_chain=8)


def test_tiled_nuts_few_chains_speed():
    """64 chains over 16384 rows: the automatic layout (8 wavefronts per chain)
    against one wavefront per chain, same seed; both are checked for agreement of
    their posterior means (different summation orders, so not bit-equal)."""
    say("tiled speed")
    model = _logistic(16384, 32)
    kw = dict(num_chains=64, num_samples=100, num_warmup=100, seed=24)
    runs = {}
    for W in (1, 0):
        t0 = time.perf_counter()
        r = ut.batched_nuts(model, waves_per_chain=W, **kw)
        runs[W] = (r, time.perf_counter() - t0)
    a, b = runs[1][0], runs[0][0]
    say(f"  one wavefront: {runs[1][1]:.2f} s (sampling {a.sampling_seconds:.2f} s); automatic: "
        f"{runs[0][1]:.2f} s (sampling {b.sampling_seconds:.2f} s); speedup "
        f"{a.sampling_seconds / b.sampling_seconds:.2f}x")
    ma, mb = a.unconstrained_samples.mean(axis=(0, 1)), b.unconstrained_samples.mean(axis=(0, 1))
    sd = a.unconstrained_samples.std(axis=(0, 1))
    assert np.max(np.abs(ma - mb) / sd) < 0.5
    assert b.sampling_seconds < a.sampling_seconds
