"""The oracle's full-rank Gaussian guide (orc_fullrank_*), pinned without retyping
its loops: against the matrix forms q = mu + L z, z = L^-1 (q - mu), L^T g with
L = diag(exp(log_scale)) + tril(F, -1) (entries of F on and above the diagonal are
set and must be ignored); the ELBO weight gradient against central finite
differences of its objective; and, for a Gaussian target N(m, S), against what the
mathematics fixes: with particles whose sample mean is 0 and sample covariance
exactly I the Monte Carlo ELBO equals the closed form, so its gradient is the
closed-form gradient and vanishes at the optimum (m, chol(S))."""
import numpy as np
import pytest
from scipy.linalg import solve_triangular

from oracle import binding as B


def _weights(P, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = np.empty(2 * P + P * P, dtype=np.float32)
    w[:P] = rng.normal(size=P)
    w[P:2 * P] = rng.normal(size=P) * 0.3
    w[2 * P:] = rng.normal(size=P * P) * scale  # diagonal and upper entries too: ignored
    return w


def _chol(w, P):
    w = w.astype(np.float64)
    return np.tril(w[2 * P:].reshape(P, P), -1) + np.diag(np.exp(w[P:2 * P]))


@pytest.mark.parametrize("P", [2, 7, 33])
def test_forward_inverse_vjp_match_matrix_forms(P):
    w = _weights(P, P)
    o = B.OracleFullRank(P, w)
    Lc = _chol(w, P)
    z = np.random.default_rng(P + 1).normal(size=(50, P))
    q, ld = o.forward(z)
    assert np.allclose(q, w[:P].astype(np.float64) + z @ Lc.T, rtol=1e-12, atol=1e-12)
    assert np.allclose(ld, np.linalg.slogdet(Lc)[1], rtol=1e-12, atol=1e-12)
    z2, ldi = o.inverse(q)
    assert np.allclose(z2, z, rtol=1e-10, atol=1e-10)
    assert np.allclose(z2, solve_triangular(Lc, (q - w[:P].astype(np.float64)).T, lower=True).T, atol=1e-10)
    assert np.allclose(ldi, -ld, rtol=1e-12, atol=1e-12)
    g = np.random.default_rng(P + 2).normal(size=(50, P))
    assert np.allclose(o.vjp(g), g @ Lc, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("weighted", [False, True])
def test_elbo_gradient_matches_finite_differences(weighted):
    P, n = 6, 40
    w = _weights(P, 21)
    rng = np.random.default_rng(22)
    z = rng.normal(size=(n, P))
    g_q = rng.normal(size=(n, P))
    valid = (rng.random(n) > 0.2).astype(np.int32)
    rw = rng.random(n) * valid if weighted else None

    def objective(ww):
        q, ld = B.OracleFullRank(P, ww).forward(z)
        per = np.sum(g_q * q, axis=1) + ld
        return float(np.sum(rw * per)) if weighted else float(per[valid != 0].mean())

    grad = B.OracleFullRank(P, w).elbo_grad(z, g_q, valid=None if weighted else valid, row_weights=rw)
    w64 = w.astype(np.float64)
    h = 1e-3
    for k in range(w.size):
        wp, wm = w64.copy(), w64.copy()
        wp[k] += h
        wm[k] -= h
        # the oracle reads float32 weights: perturb in float64 and difference the float32-exact values
        fp, fm = objective(wp.astype(np.float32)), objective(wm.astype(np.float32))
        step = float(np.float32(wp[k])) - float(np.float32(wm[k]))
        fd = (fp - fm) / step
        assert abs(fd - grad[k]) <= 2e-4 * (1 + abs(grad[k])), (k, fd, grad[k])
    F_grad = grad[2 * P:].reshape(P, P)
    assert np.all(np.triu(F_grad) == 0.0)


def test_gaussian_target_closed_form_and_optimum():
    P = 5
    rng = np.random.default_rng(31)
    A = rng.normal(size=(P, P))
    S = A @ A.T + P * np.eye(P)
    Si = np.linalg.inv(S)
    m = rng.normal(size=P)
    # 2P particles +-sqrt(P) e_k: sample mean 0, sample second moment exactly I
    z = np.concatenate([np.sqrt(P) * np.eye(P), -np.sqrt(P) * np.eye(P)])

    def grads(Lc, mu):
        w = np.concatenate([mu, np.log(np.diag(Lc)), np.tril(Lc, -1).reshape(-1)]).astype(np.float32)
        Lc = _chol(w, P)
        mu = w[:P].astype(np.float64)
        o = B.OracleFullRank(P, w)
        q, _ = o.forward(z)
        g_q = -(q - m) @ Si  # gradient of log N(q; m, S)
        return o.elbo_grad(z, g_q), Lc, mu

    g, Lc, mu = grads(np.linalg.cholesky(S), m)
    assert np.max(np.abs(g)) < 1e-5  # stationary at (m, chol S) up to float32 weights
    L0 = np.linalg.cholesky(S) + np.tril(rng.normal(size=(P, P)) * 0.3)
    L0[np.diag_indices(P)] = np.abs(np.diag(L0)) + 0.5
    g, Lc, mu = grads(L0, m + rng.normal(size=P))
    # closed form: ELBO = -1/2 tr(S^-1 L L^T) - 1/2 (mu - m)^T S^-1 (mu - m) + sum log L_ii + const
    dL = -Si @ Lc
    assert np.allclose(g[:P], -Si @ (mu - m), rtol=1e-10, atol=1e-10)
    assert np.allclose(g[P:2 * P], np.diag(dL) * np.diag(Lc) + 1.0, rtol=1e-10, atol=1e-10)
    assert np.allclose(g[2 * P:].reshape(P, P), np.tril(dL, -1), rtol=1e-10, atol=1e-10)
