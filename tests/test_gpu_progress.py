"""GPU: the samplers' progress callback (reference callback / callback_every,
src/inference/api_single_chain.jl:1-30).

Each sampler reports every `callback_every` iterations of a phase and on the phase's
last iteration, with the cumulative divergences of the phase; the draws of a run with
a callback are bit-identical to those without one (the persistent samplers then end
a launch at every reported iteration, which must not change the chains); an
exception raised by the callback surfaces after the call."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model():
    y, off, sig, _ = S.hierarchical_data(8, 8, seed=3)
    return ut.HierarchicalNormalNC(y, off, sig)


def _expected(every, total):
    its = [i for i in range(1, total + 1) if i % every == 0]
    return its if (its and its[-1] == total) else its + [total]


@pytest.mark.parametrize("sampler", ["nuts", "hmc", "chees", "hmc_shared", "meads"])
def test_progress_callback_schedule_and_identical_draws(sampler):
    say("progress callback", sampler)
    model = _model()
    W, Sn, every, C = 30, 25, 7, 64
    run = {"nuts": ut.batched_nuts, "hmc": ut.batched_hmc, "chees": ut.batched_chees,
           "hmc_shared": lambda *a, **k: ut.batched_hmc(*a, per_chain_adaptation=False, **k),
           "meads": ut.batched_meads}[sampler]
    kw = dict(num_chains=C, num_samples=Sn, num_warmup=W, seed=5)
    if sampler != "meads":
        kw["step_size"] = 0.6  # large enough for some divergent transitions early in warmup
    calls = []
    res = run(model, callback=calls.append, callback_every=every, **kw)
    ref = run(model, **kw)
    assert np.array_equal(res.unconstrained_samples.view(np.uint32), ref.unconstrained_samples.view(np.uint32))
    assert np.array_equal(res.divergent, ref.divergent) and np.array_equal(res.step_size, ref.step_size)
    warm = [c for c in calls if c.phase == "warmup"]
    samp = [c for c in calls if c.phase == "sample"]
    assert [c.iteration for c in warm] == _expected(every, W)
    assert [c.iteration for c in samp] == _expected(every, Sn)
    assert all(c.total == W for c in warm) and all(c.total == Sn for c in samp)
    for seq in (warm, samp):
        d = [c.divergences for c in seq]
        assert d == sorted(d) and d[0] >= 0
    assert samp[-1].divergences == int(res.divergent.sum())
    if sampler == "meads":
        assert all(np.isnan(c.step_size) for c in calls) and all(c.divergences == 0 for c in warm)
    else:
        assert all(np.isfinite(c.step_size) and c.step_size > 0 for c in calls)
        assert abs(samp[-1].step_size - float(np.mean(res.step_size))) <= 1e-12 * samp[-1].step_size
    say("  warmup divergences", warm[-1].divergences, "sampling", samp[-1].divergences)


def test_progress_callback_exception_surfaces():
    model = _model()

    def boom(info):
        if info.iteration >= 10:
            raise KeyError("stop here")

    with pytest.raises(KeyError, match="stop here"):
        ut.batched_nuts(model, num_chains=64, num_samples=20, num_warmup=20, seed=1, callback=boom, callback_every=5)
