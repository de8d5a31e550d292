"""The oracle's binary64 flow (CPU), pinned independently of its own formulas:
identity at initialisation, f^-1(f(z)) = z, log-determinants against a brute-force
finite-difference Jacobian and numpy's slogdet, the sampler's vector-Jacobian
product and the training gradient against central finite differences."""
import numpy as np
import pytest

from oracle import binding as B


def _weights(P, L, H, seed, scale=0.3, base=True):
    """Random flat weights in the device layout (padding entries zero)."""
    rng = np.random.default_rng(seed)
    n = B.OracleFlow.num_weights(P, L, H)
    w = np.zeros(n, dtype=np.float32)
    if base:
        w[:P] = rng.normal(size=P) * 0.5
        w[P:2 * P] = rng.normal(size=P) * 0.3
    off = 2 * P
    for l in range(L):
        npass = sum(1 for i in range(P) if (i + l) % 2 == 0)
        nact = P - npass
        kp = -(-npass // 128) * 128
        op = -(-2 * nact // 128) * 128
        W1 = np.zeros((H, kp), np.float32)
        W1[:, :npass] = rng.normal(size=(H, npass)) / np.sqrt(npass)
        b1 = rng.normal(size=H) * 0.1
        W2 = np.zeros((op, H), np.float32)
        W2[:2 * nact] = rng.normal(size=(2 * nact, H)) * scale / np.sqrt(H)
        b2 = np.zeros(op, np.float32)
        b2[:2 * nact] = rng.normal(size=2 * nact) * 0.1
        for arr in (W1.ravel(), b1, W2.ravel(), b2):
            w[off:off + arr.size] = arr
            off += arr.size
    assert off == n
    return w


def test_identity_at_zero_conditioner():
    P, L, H = 7, 4, 8
    w = _weights(P, L, H, 1)
    off = 2 * P
    for l in range(L):  # zero every W2 and b2: s = t = 0
        npass = sum(1 for i in range(P) if (i + l) % 2 == 0)
        kp, op = 128, 128
        off += H * kp + H
        w[off:off + op * H + op] = 0.0
        off += op * H + op
    f = B.OracleFlow(P, L, H, 5.0, w)
    z = np.random.default_rng(2).normal(size=(5, P))
    q, ld = f.forward(z)
    assert np.allclose(q, w[:P] + np.exp(w[P:2 * P].astype(np.float64)) * z, rtol=1e-14, atol=1e-14)
    assert np.allclose(ld, np.sum(w[P:2 * P].astype(np.float64)), rtol=1e-14)


@pytest.mark.parametrize("P,L", [(5, 3), (8, 6), (11, 2)])
def test_inverse_and_logdet(P, L):
    H = 8
    f = B.OracleFlow(P, L, H, 5.0, _weights(P, L, H, 3 + P))
    z = np.random.default_rng(P).normal(size=(6, P))
    q, ld = f.forward(z)
    z2, ldi = f.inverse(q)
    assert np.allclose(z2, z, atol=1e-11)
    assert np.allclose(ldi, -ld, atol=1e-11)
    # brute force: central-difference Jacobian, numpy slogdet
    h = 1e-6
    for r in range(3):
        J = np.empty((P, P))
        for i in range(P):
            e = np.zeros(P)
            e[i] = h
            J[:, i] = (f.forward(z[r] + e)[0][0] - f.forward(z[r] - e)[0][0]) / (2 * h)
        sign, lad = np.linalg.slogdet(J)
        assert sign != 0
        assert abs(lad - ld[r]) < 1e-7 * max(1.0, abs(lad))


def test_vjp_matches_finite_differences():
    P, L, H = 9, 4, 8
    f = B.OracleFlow(P, L, H, 5.0, _weights(P, L, H, 11))
    rng = np.random.default_rng(12)
    z = rng.normal(size=(3, P))
    gq = rng.normal(size=(3, P))
    _, _, gz = f.forward(z, gq)
    h = 1e-6
    for r in range(3):
        def J(zz):
            q, ld = f.forward(zz)
            return float(gq[r] @ q[0] + ld[0])
        fd = np.array([(J(z[r] + h * np.eye(P)[i]) - J(z[r] - h * np.eye(P)[i])) / (2 * h) for i in range(P)])
        assert np.allclose(gz[r], fd, rtol=1e-6, atol=1e-7)


def test_loss_gradient_matches_finite_differences():
    P, L, H = 6, 3, 8
    w = _weights(P, L, H, 21)
    q = np.random.default_rng(22).normal(size=(16, P)) * 1.5
    loss, g = B.OracleFlow(P, L, H, 5.0, w).loss_grad(q)
    # the loss is the mean negative log-likelihood: closed form at the base
    z, ldi = B.OracleFlow(P, L, H, 5.0, w).inverse(q)
    assert loss == pytest.approx(np.mean(0.5 * np.sum(z * z, axis=1) - ldi + 0.5 * P * np.log(2 * np.pi)), rel=1e-12)
    rng = np.random.default_rng(23)
    nz = np.flatnonzero(w)  # perturb real (non-padding) weights, in float32 steps
    for i in rng.choice(nz, size=25, replace=False):
        hstep = np.float32(2.0 ** -13) * max(np.float32(1.0), abs(w[i]))
        wp, wm = w.copy(), w.copy()
        wp[i] += hstep
        wm[i] -= hstep
        lp, _ = B.OracleFlow(P, L, H, 5.0, wp).loss_grad(q)
        lm, _ = B.OracleFlow(P, L, H, 5.0, wm).loss_grad(q)
        fd = (lp - lm) / (float(wp[i]) - float(wm[i]))
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-7), i


def test_elbo_gradient_matches_finite_differences():
    """Weight gradient of the reparameterised ELBO (g_q held fixed), over valid rows
    only, against central differences of mean(g_q . f(z) + logdet) on those rows."""
    P, L, H = 6, 3, 8
    w = _weights(P, L, H, 31)
    rng = np.random.default_rng(32)
    z = rng.normal(size=(16, P))
    gq = rng.normal(size=(16, P))
    valid = (np.arange(16) % 5 != 3).astype(np.int32)
    g = B.OracleFlow(P, L, H, 5.0, w).elbo_grad(z, gq, valid)
    # dropping the invalid rows is the same as passing only the valid ones
    keep = valid.astype(bool)
    assert np.allclose(g, B.OracleFlow(P, L, H, 5.0, w).elbo_grad(z[keep], gq[keep]), rtol=1e-12, atol=1e-14)

    def J(ww):
        q, ld = B.OracleFlow(P, L, H, 5.0, ww).forward(z[keep])
        return float(np.mean(np.sum(gq[keep] * q, axis=1) + ld))

    nz = np.flatnonzero(w)
    # every base parameter, and a sample of the conditioner weights
    picks = np.concatenate([np.arange(2 * P), rng.choice(nz[nz >= 2 * P], size=25, replace=False)])
    for i in picks:
        hstep = np.float32(2.0 ** -13) * max(np.float32(1.0), abs(w[i]))
        wp, wm = w.copy(), w.copy()
        wp[i] += hstep
        wm[i] -= hstep
        fd = (J(wp) - J(wm)) / (float(wp[i]) - float(wm[i]))
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-7), i


def test_iwae_weights_and_bound():
    """The restated IWAE reduction: group size 1 is the standard ELBO with uniform
    weights; the bound is at least the standard ELBO (Jensen) and at most the
    largest log-weight; weights sum to one and vanish on invalid particles; a
    group without a valid particle is skipped; equal log-weights give the log-weight."""
    rng = np.random.default_rng(41)
    logw = rng.normal(size=64) * 3.0
    valid = np.ones(64, bool)
    valid[[5, 17, 18]] = False
    w1, b1 = B.iwae_weights_and_bound(logw, valid, 1)
    assert b1 == pytest.approx(logw[valid].mean(), rel=1e-14)
    assert np.allclose(w1[valid], 1.0 / valid.sum()) and np.all(w1[~valid] == 0)
    for K in (2, 8, 64):
        w, b = B.iwae_weights_and_bound(logw, valid, K)
        assert b >= b1 - 1e-12 and b <= logw[valid].max()
        assert w.sum() == pytest.approx(1.0, rel=1e-14) and np.all(w[~valid] == 0)
    v2 = valid.copy()
    v2[:8] = False
    w, b = B.iwae_weights_and_bound(logw, v2, 8)
    assert np.all(w[:8] == 0) and w.sum() == pytest.approx(1.0, rel=1e-14)
    w, b = B.iwae_weights_and_bound(np.full(16, -3.5), np.ones(16, bool), 4)
    assert b == pytest.approx(-3.5, rel=1e-14) and np.allclose(w, 1 / 16)


def test_weighted_elbo_gradient_matches_finite_differences():
    """With row weights the oracle's gradient is that of sum_r w_r (g_r . f(z_r) +
    logdet_r), weights held fixed (the IWAE gradient's form)."""
    P, L, H = 5, 2, 8
    w = _weights(P, L, H, 51)
    rng = np.random.default_rng(52)
    z = rng.normal(size=(12, P))
    gq = rng.normal(size=(12, P))
    rw = rng.uniform(size=12)
    g = B.OracleFlow(P, L, H, 5.0, w).elbo_grad(z, gq, row_weights=rw)

    def J(ww):
        q, ld = B.OracleFlow(P, L, H, 5.0, ww).forward(z)
        return float(np.sum(rw * (np.sum(gq * q, axis=1) + ld)))

    nz = np.flatnonzero(w)
    for i in np.concatenate([np.arange(2 * P), rng.choice(nz[nz >= 2 * P], size=15, replace=False)]):
        hstep = np.float32(2.0 ** -13) * max(np.float32(1.0), abs(w[i]))
        wp, wm = w.copy(), w.copy()
        wp[i] += hstep
        wm[i] -= hstep
        fd = (J(wp) - J(wm)) / (float(wp[i]) - float(wm[i]))
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-7), i


def test_bf16_operands_stay_close():
    P, L, H = 12, 4, 8
    w = _weights(P, L, H, 31)
    z = np.random.default_rng(32).normal(size=(8, P))
    q64, ld64 = B.OracleFlow(P, L, H, 5.0, w).forward(z)
    q16, ld16 = B.OracleFlow(P, L, H, 5.0, w, bf16=True).forward(z)
    assert np.max(np.abs(q16 - q64)) < 0.05 * max(1.0, np.max(np.abs(q64)))
    assert not np.array_equal(q16, q64)
