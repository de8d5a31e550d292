"""Pins of the oracle's workgroup-per-chain logistic order (W wavefronts per chain,
rows 64 (w + W k) + l on wavefront w, lane l; wavefront totals added in order).

- Row coverage: with beta = 0 every residual is y - 1/2 and, for small-integer
  covariates, every partial sum of the gradient is exact in binary32, so each W must
  give the closed-form gradient bit for bit; a row dropped or scored twice by the
  tiling changes it.
- The log density of that point is -N log 2 + prior: checked to binary32 rounding.
- Random data: every W agrees with the binary64 reference formulas to the binary32
  tolerance of the density parity tests, and W = 1 is the one-wavefront order."""
import math

import numpy as np
import pytest

from oracle import binding as B
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S


def _int_logistic(n, d, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.integers(-3, 4, size=(n, d)).astype(np.float32)
    y = (rng.random(n) < 0.5).astype(np.float32)
    return M.LogisticRegression(x, y, 0.0, 2.0), x, y


@pytest.mark.parametrize("W", [1, 2, 4, 8])
@pytest.mark.parametrize("n,d", [(8192, 16), (5000, 70), (37, 3)])
def test_tiled_row_coverage_exact(W, n, d):
    model, x, y = _int_logistic(n, d)
    om = B.OracleModel(model, waves_per_chain=W)
    params = np.zeros((d, 3), np.float32)
    v, g = om.logjoint_grad_twin(params, np.float32)
    # d/dbeta = X^T (y - 1/2) - (beta - loc) / scale^2, beta = 0, loc = 0: every term an
    # exact multiple of 1/2 far below 2^24
    exact = x.astype(np.float64).T @ (y.astype(np.float64) - 0.5)
    assert np.array_equal(g.astype(np.float64), np.repeat(exact[:, None], 3, axis=1))
    # log density: sum_i (y_i * 0 - log(1 + e^0)) + prior at zero
    ref = -n * math.log(2.0) + d * (-0.5 * math.log(2 * math.pi) - math.log(2.0))
    assert np.allclose(v, ref, rtol=2e-6 * max(1.0, math.log2(n)), atol=1e-3)


@pytest.mark.parametrize("W", [2, 4, 8])
def test_tiled_twin_against_binary64(W):
    x, y, _ = S.logistic_data(8192, 16)
    model = M.LogisticRegression(x, y)
    rng = np.random.default_rng(11)
    params = (rng.normal(size=(16, 8)) * 0.3).astype(np.float32)
    om = B.OracleModel(model, waves_per_chain=W)
    v, g = om.logjoint_grad_twin(params, np.float32)
    rv, rg = om.logjoint_grad_f64(params.astype(np.float64))
    assert np.all(np.abs(v - rv) <= 1e-5 * np.abs(rv) + 1e-4)
    assert np.max(np.abs(g - rg)) <= 1e-5 * max(1.0, np.max(np.abs(rg)))
    # binary64 twin of the same order: 1e-12 relative
    v64, g64 = om.logjoint_grad_twin(params.astype(np.float64), np.float64)
    assert np.allclose(v64, rv, rtol=1e-12)
    assert np.max(np.abs(g64 - rg)) <= 1e-11 * max(1.0, np.max(np.abs(rg)))
    # a different summation order: the binary32 results differ from W = 1 somewhere
    v1, g1 = B.OracleModel(model).logjoint_grad_twin(params, np.float32)
    assert not (np.array_equal(v, v1) and np.array_equal(g, g1))


def test_tiled_w1_is_the_wavefront_order():
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    params = (np.random.default_rng(2).normal(size=(25, 6)) * 0.5).astype(np.float32)
    a = B.OracleModel(model).logjoint_grad_twin(params, np.float32)
    b = B.OracleModel(model, waves_per_chain=1).logjoint_grad_twin(params, np.float32)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_tiled_twin_rejects_other_layouts():
    x, y, _ = S.logistic_data(100, 3)
    with pytest.raises(ValueError):
        B.OracleModel(M.LogisticRegression(x, y), waves_per_chain=3)
    with pytest.raises(ValueError):
        B.OracleModel(M.NealsFunnel(5, 3.0), waves_per_chain=2)
