"""Host side of the device VI path (no GPU): the samplers' initial-position draws
that batched_advi starts from (the reference's _initial_hmc_position, prior draws
in unconstrained space), and batched_advi's argument checks, which fail before any
device work (messages follow src/inference/vi.jl:300-390)."""
import numpy as np
import pytest

import uncertaintea_amd as ut
from uncertaintea_amd import _lib as L


def _initial(model, strategy, seed, n):
    out = np.empty((n, model.num_params))
    L.check(L.lib().ut_initial_params(model.plan, strategy, seed, n, L.ptr(out)))
    return out


def test_initial_params_follow_the_prior():
    """Funnel prior draws: v ~ N(0, 3) and x_i e^(-v/2) ~ N(0, 1), moments within
    5 standard errors at 40000 draws; seeded and deterministic."""
    m = ut.NealsFunnel(10, 3.0)
    a = _initial(m, L.UT_INIT_PRIOR, 7, 40000)
    v = a[:, 0]
    u = a[:, 1:] * np.exp(-0.5 * v[:, None])
    n = a.shape[0]
    assert abs(v.mean()) < 5 * 3.0 / np.sqrt(n) and abs(v.var() - 9.0) < 5 * 9.0 * np.sqrt(2.0 / n)
    assert np.all(np.abs(u.mean(axis=0)) < 5 / np.sqrt(n)) and np.all(np.abs(u.var(axis=0) - 1) < 5 * np.sqrt(2 / n))
    assert np.array_equal(a, _initial(m, L.UT_INIT_PRIOR, 7, 40000))
    assert not np.array_equal(a[:10], _initial(m, L.UT_INIT_PRIOR, 8, 10))
    # chain c's draw does not depend on how many chains are drawn
    assert np.array_equal(a[:5], _initial(m, L.UT_INIT_PRIOR, 7, 5))


def test_initial_params_uniform():
    m = ut.NealsFunnel(4, 3.0)
    a = _initial(m, L.UT_INIT_UNIFORM, 3, 20000)
    assert a.min() >= -2.0 and a.max() < 2.0 and abs(a.mean()) < 0.02


def test_initial_params_argument_errors():
    m = ut.NealsFunnel(4, 3.0)
    out = np.empty((2, m.num_params))
    with pytest.raises(ValueError):
        L.check(L.lib().ut_initial_params(m.plan, 7, 0, 2, L.ptr(out)))
    with pytest.raises(ValueError, match="DimensionMismatch"):
        L.check(L.lib().ut_initial_params(m.plan, L.UT_INIT_PRIOR, 0, 0, L.ptr(out)))


@pytest.mark.parametrize("kw,msg", [
    ({"guide": "lowrank"}, "CPU-only"),
    ({"guide": "bogus"}, "guide must be"),
    ({"elbo": "bogus"}, "elbo must be"),
    ({"elbo": "iwae", "iwae_samples": 100}, "divide num_particles evenly"),
    ({"elbo": "iwae", "iwae_samples": 256}, "1 <= iwae_samples"),
    ({"num_iterations": 0}, "num_iterations > 0"),
    ({"num_particles": 100}, "multiple of 128"),
    ({"learning_rate": 0.0}, "learning_rate > 0"),
    ({"beta1": 1.0}, "beta1"),
    ({"gradient_clip": 0.0}, "gradient_clip > 0"),
    ({"initial_params": np.zeros(3)}, "DimensionMismatch"),
])
def test_batched_advi_argument_errors(kw, msg):
    args = dict(num_iterations=10)
    args.update(kw)
    with pytest.raises(ValueError, match=msg):
        ut.batched_advi(ut.NealsFunnel(4, 3.0), **args)
