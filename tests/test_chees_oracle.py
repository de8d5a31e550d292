"""The oracle's ChEES restatement (CPU): known answers from the reference's own
tests (test/uncertaintea/core/batched_chees_adaptation.jl:27-71), closed forms of
the Adam / clip / moving-average update, and a posterior check against the
oracle's NUTS on the same model."""
import math

import numpy as np
import pytest

from oracle import binding as B
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S


def test_halton_base2_sequence():
    # 1/2, 1/4, 3/4, 1/8, 5/8, 3/8, 7/8, 1/16 (reference comment, api_batched_chees.jl:28-29)
    assert np.array_equal(B.chees_halton(8), [0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875, 0.0625])
    # radical inverse: bits of the index mirrored behind the binary point
    h = B.chees_halton(1000)
    for i in (5, 37, 512, 999):
        bits = bin(i)[2:][::-1]
        assert h[i - 1] == sum(int(b) * 2.0 ** -(k + 1) for k, b in enumerate(bits))


def test_leapfrog_step_counts():
    # jitter 0 => L = num_leapfrog_steps exactly at T = L * eps (the integer snap)
    for L in (1, 3, 10, 12, 37, 100):
        for eps in (0.1, 0.037, 1.6, 0.9):
            assert B.chees_steps(1.0, L * eps, eps) == L
    assert B.chees_steps(0.5, 10 * 0.1, 0.1) == 5
    assert B.chees_steps(0.31, 10 * 0.1, 0.1) == 4  # ceil(3.1)
    assert B.chees_steps(1e-6, 10 * 0.1, 0.1) == 1  # clamped


def _adam_closed_form(T0, grads_terms, lr=0.5, b1=0.0, b2=0.95, aeps=1e-8, decay=0.5, max_steps=1000):
    """Independent restatement: Adam ascent on log T, clip 0.35, moving average
    with weight step^-decay, clamp to [eps, max_steps * eps]."""
    T, ma, m, v = T0, math.log(T0), 0.0, 0.0
    out = []
    for t, (num, den, jit, eps) in enumerate(grads_terms, start=1):
        g = jit * T * num / den if den > 0 else 0.0
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        upd = lr * (m / (1 - b1 ** t)) / (math.sqrt(v / (1 - b2 ** t)) + aeps)
        upd = min(max(upd, -0.35), 0.35)
        w = t ** -decay
        ma = (1 - w) * ma + w * (math.log(T) + upd)
        T = min(max(math.exp(ma), eps), max_steps * eps)
        out.append(T)
    return np.array(out)


def test_trajectory_update_closed_form():
    # first update: b1 = 0 makes mhat = g and vhat = g^2, so the step is
    # lr * sign(g) = +-0.5, clipped to +-0.35; weight 1^-0.5 = 1 moves T fully
    T = B.chees_trajectory(1.0, [2.0], [1.0], [1.0], [0.1])
    assert T[0] == pytest.approx(math.exp(0.35), rel=1e-12)
    T = B.chees_trajectory(1.0, [-2.0], [1.0], [1.0], [0.1])
    assert T[0] == pytest.approx(math.exp(-0.35), rel=1e-12)
    # zero denominator: no gradient, no move
    assert B.chees_trajectory(2.0, [5.0], [0.0], [1.0], [0.1])[0] == pytest.approx(2.0, rel=1e-15)
    # a longer sequence with mixed signs, small gradients (unclipped) and the clamp
    rng = np.random.default_rng(3)
    n = 60
    num = rng.normal(size=n) * np.where(np.arange(n) < 30, 1e-3, 1.0)
    den = np.abs(rng.normal(size=n)) + 0.1
    jit = rng.uniform(0.5, 1.0, size=n)
    eps = np.full(n, 0.05)
    terms = list(zip(num, den, jit, eps))
    got = B.chees_trajectory(0.4, num, den, jit, eps)
    assert np.allclose(got, _adam_closed_form(0.4, terms), rtol=1e-12)
    # clamp at the step size and at max_steps * step
    assert B.chees_trajectory(0.06, [-1.0] * 5, [1.0] * 5, [1.0] * 5, [0.05] * 5)[-1] == 0.05
    assert B.chees_trajectory(1.0, [1.0] * 40, [1.0] * 40, [1.0] * 40, [0.05] * 40, max_steps=30)[-1] == \
        pytest.approx(1.5, rel=1e-15)


def test_nonfinite_update_reverts():
    # a NaN numerator makes the update non-finite: log T keeps its value, so the
    # first (weight 1) moving-average step leaves T unchanged
    T = B.chees_trajectory(1.3, [float("nan")], [1.0], [1.0], [0.1])
    assert T[0] == pytest.approx(1.3, rel=1e-15)


@pytest.fixture(scope="module")
def logistic():
    x, y, _ = S.logistic_data(1000, 25)
    return B.OracleModel(M.LogisticRegression(x, y))


def test_chees_posterior_matches_nuts(logistic):
    """Binary64 ChEES (forward velocity, where T settles at several steps) and
    binary64 NUTS on the same logistic regression agree in posterior mean within
    a few MCSE."""
    ch = logistic.run("chees", "f64", 64, num_samples=400, num_warmup=400, seed=5, threads=8,
                      chees={"forward_velocity": True})
    nu = logistic.run("nuts", "f64", 64, num_samples=300, num_warmup=300, seed=6, threads=8)
    a, b = ch["samples"], nu["samples"]
    assert float(np.max(np.asarray(D.rhat(a)))) < 1.02
    se = np.sqrt(a.var(axis=(0, 1)) / np.asarray(D.ess_bulk(a)) + b.var(axis=(0, 1)) / np.asarray(D.ess_bulk(b)))
    z = np.abs(a.mean(axis=(0, 1)) - b.mean(axis=(0, 1))) / se
    assert z.max() < 5.0
    assert ch["trajectory_trace"][-1] > 4 * ch["step_size"][0]


def test_reference_velocity_shortens_trajectory(logistic):
    """The reference's criterion uses the flipped end momentum; its ascent drives
    T down to about one step on this target, while the forward velocity lengthens
    it.  Both runs start from T = 10 steps."""
    ref = logistic.run("chees", "f64", 64, num_samples=10, num_warmup=400, seed=5, threads=8)
    fwd = logistic.run("chees", "f64", 64, num_samples=10, num_warmup=400, seed=5, threads=8,
                       chees={"forward_velocity": True})
    assert ref["trajectory_trace"][-1] < 1.5 * ref["step_size"][0]
    assert fwd["trajectory_trace"][-1] > 5.0 * fwd["step_size"][0]


def test_chees_adaptation_off_freezes_trajectory(logistic):
    r1 = logistic.run("chees", "twin", 32, num_samples=20, num_warmup=50, seed=9, threads=4,
                      chees={"adapt_trajectory_length": False})
    r2 = logistic.run("chees", "twin", 32, num_samples=20, num_warmup=50, seed=9, threads=4,
                      chees={"adapt_trajectory_length": False})
    assert len(np.unique(np.round(r1["trajectory_trace"], 10))) == 1
    assert np.array_equal(r1["samples"], r2["samples"])
