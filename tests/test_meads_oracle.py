"""The oracle's MEADS restatement (CPU): the reference's known answers for the
trace-ratio eigenvalue estimate and the slice drift
(test/uncertaintea/core/batched_meads.jl:32-57), the estimate against numpy's
eigenvalues where the ratio is exact, the drift as a rotation of [-1, 1), and posterior recovery on a logistic
regression against the oracle's NUTS."""
import numpy as np
import pytest

from oracle import binding as B
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S


def test_max_eig_known_answers():
    assert B.meads_max_eig(np.array([[1.0], [2.0]])) == 0.0  # one member: no information
    v = np.array([3.0, 4.0])
    assert B.meads_max_eig(np.stack([v, v, v, v], axis=1)) == pytest.approx(25.0, abs=1e-10)
    e1 = np.array([1.0, 0.0, 0.0])
    big = np.stack([2 * e1, -2 * e1, 2 * e1, -2 * e1], axis=1)
    assert B.meads_max_eig(big) == pytest.approx(4.0, abs=1e-10)


def test_max_eig_is_the_trace_ratio():
    # brute force: the estimator is sum_{a != b} (x_a . x_b)^2 / (n (n - 1)) divided by
    # sum_a |x_a|^2 / n; for an ensemble of orthogonal equal-norm members the
    # off-diagonal part vanishes
    rng = np.random.default_rng(4)
    X = rng.normal(size=(6, 9))  # P x N
    n = X.shape[1]
    G = X.T @ X
    off = (G ** 2).sum() - (np.diag(G) ** 2).sum()
    expect = (off / (n * (n - 1))) / (np.trace(G) / n)
    assert B.meads_max_eig(X) == pytest.approx(expect, rel=1e-12)
    Q, _ = np.linalg.qr(rng.normal(size=(8, 8)))
    assert B.meads_max_eig(Q[:, :5] * 3.0) == pytest.approx(0.0, abs=1e-12)


def test_slice_drift_known_answers():
    # batched_meads.jl:48-55 against the oracle's drift function itself
    assert B.meads_slice_drift(0.9, 0.2) == pytest.approx(-0.9, abs=1e-12)
    assert B.meads_slice_drift(-0.9, 0.0) == pytest.approx(-0.9, abs=1e-12)
    for u in (-0.99, -0.5, 0.0, 0.4, 0.999):
        for d in (0.0, 0.1, 0.5, 0.9):
            assert -1.0 <= B.meads_slice_drift(u, d) < 1.0


def test_slice_drift_is_a_rotation_of_the_circle():
    # mod 2 arithmetic: drifting by d1 then d2 equals drifting by d1 + d2 (mod 2),
    # drifting by 2 is the identity and the map is a translation on [-1, 1)
    rng = np.random.default_rng(8)
    for _ in range(200):
        u = rng.uniform(-1.0, 1.0)
        d1, d2 = rng.uniform(0.0, 1.0, size=2)
        two = B.meads_slice_drift(B.meads_slice_drift(u, d1), d2)
        one = B.meads_slice_drift(u, d1 + d2)
        assert min(abs(two - one), 2.0 - abs(two - one)) < 1e-12
        assert B.meads_slice_drift(u, 2.0) == pytest.approx(u, abs=1e-12)
        t = B.meads_slice_drift(u, d1)
        assert min(abs((t - u) % 2.0 - d1), 2.0 - abs((t - u) % 2.0 - d1)) < 1e-12


def test_meads_posterior_matches_nuts():
    x, y, _ = S.logistic_data(1000, 25)
    om = B.OracleModel(M.LogisticRegression(x, y))
    me = om.run("meads", "f64", 64, num_samples=1500, num_warmup=1500, seed=3, threads=8)
    nu = om.run("nuts", "f64", 64, num_samples=300, num_warmup=300, seed=6, threads=8)
    a, b = me["samples"], nu["samples"]
    assert float(np.max(np.asarray(D.rhat(a)))) < 1.05
    se = np.sqrt(a.var(axis=(0, 1)) / np.asarray(D.ess_bulk(a)) + b.var(axis=(0, 1)) / np.asarray(D.ess_bulk(b)))
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / se
    assert z.max() < 5.0
    tr = me["trajectory_trace"]
    assert np.all(np.isfinite(tr)) and np.all(tr[:, 0] > 0) and np.all(tr[:, 0] <= 1.0)
    assert np.all((tr[:, 1] >= 0) & (tr[:, 1] <= 1))
    # skipped folds record trivially kept transitions
    steps = me["integration_steps"]
    assert set(np.unique(steps)) <= {0, 1} and (steps == 0).mean() == pytest.approx(0.25, abs=0.01)
