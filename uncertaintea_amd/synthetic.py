"""Synthetic inputs of the benchmark configurations, drawn from Philox4x32-10.

Every value is a pure function of (seed, config id, stream, index), so the device
library, the CPU oracle and these generators agree on the data bit for bit.
Normals are binary64 Box-Muller of one Philox block (words (0,1) and (2,3) as two
53-bit uniforms), as in the reference's Float64 device_rand_normal
(src/device/rng.jl:338-343).
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20261203

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(k0, k1, c0, c1, c2, c3):
    """Vectorised Philox4x32-10 over numpy uint32 arrays (broadcasting)."""
    k0 = np.asarray(k0, dtype=np.uint32).copy()
    k1 = np.asarray(k1, dtype=np.uint32).copy()
    c = [np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3)]
    shape = np.broadcast_shapes(k0.shape, k1.shape, *(x.shape for x in c))
    k0 = np.broadcast_to(k0, shape).copy()
    k1 = np.broadcast_to(k1, shape).copy()
    c0, c1, c2, c3 = (np.broadcast_to(x, shape).copy() for x in c)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r > 0:
                k0 = k0 + _W0
                k1 = k1 + _W1
            p0 = c0.astype(np.uint64) * _M0
            p1 = c2.astype(np.uint64) * _M1
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def _u53(w0, w1):
    bits = (w0.astype(np.uint64) << np.uint64(32)) | w1.astype(np.uint64)
    return (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniforms(seed: int, cfg: int, stream: int, n: int) -> np.ndarray:
    """n binary64 uniforms in [0, 1) at coordinates (seed, cfg, stream, 0..n-1)."""
    idx = np.arange(n, dtype=np.uint32)
    w = philox4x32(np.uint32(seed & 0xFFFFFFFF), np.uint32(cfg), np.uint32(stream), idx, 0, 0)
    return _u53(w[0], w[1])


def normals(seed: int, cfg: int, stream: int, n: int) -> np.ndarray:
    """n binary64 standard normals (Box-Muller, one Philox block each)."""
    idx = np.arange(n, dtype=np.uint32)
    w = philox4x32(np.uint32(seed & 0xFFFFFFFF), np.uint32(cfg), np.uint32(stream), idx, 0, 0)
    u1 = _u53(w[0], w[1])
    u2 = _u53(w[2], w[3])
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)


def logistic_data(n: int, d: int, seed: int = DEFAULT_SEED, cfg: int = 1, x_scale: float = 1.0):
    """X ~ N(0, x_scale^2) (float32), beta* ~ N(0, 1), y ~ Bernoulli(sigmoid(X beta*))."""
    x = (x_scale * normals(seed, cfg, 0, n * d)).astype(np.float32).reshape(n, d)
    beta = normals(seed, cfg, 1, d)
    eta = x.astype(np.float64) @ beta
    u = uniforms(seed, cfg, 2, n)
    y = (u < 1.0 / (1.0 + np.exp(-eta))).astype(np.float32)
    return x, y, beta


def hierarchical_data(num_groups: int = 64, per_group: int = 32, seed: int = DEFAULT_SEED, cfg: int = 2,
                      mu: float = 0.0, tau: float = 1.0):
    """sigma_j ~ U(0.5, 2), theta_j = mu + tau z_j, y_ij ~ N(theta_j, sigma_j)."""
    sigma = 0.5 + 1.5 * uniforms(seed, cfg, 0, num_groups)
    theta = mu + tau * normals(seed, cfg, 1, num_groups)
    eps = normals(seed, cfg, 2, num_groups * per_group).reshape(num_groups, per_group)
    y = (theta[:, None] + sigma[:, None] * eps).astype(np.float32).reshape(-1)
    offsets = (np.arange(num_groups + 1) * per_group).astype(np.int32)
    return y, offsets, sigma, theta


def gp_kernel(inputs: np.ndarray, lengthscales, amplitude: float, noise: float) -> np.ndarray:
    """ARD squared-exponential kernel matrix plus (noise^2 + 1e-8) on the diagonal."""
    x = np.asarray(inputs, dtype=np.float64) / np.asarray(lengthscales, dtype=np.float64)
    sq = ((x[:, None, :] - x[None, :, :]) ** 2).sum(axis=2)
    return amplitude ** 2 * np.exp(-0.5 * sq) + (noise ** 2 + 1e-8) * np.eye(x.shape[0])


def gp_poisson_data(n: int = 512, d: int = 2, seed: int = DEFAULT_SEED, cfg: int = 5, lengthscales=(0.3, 0.6),
                    amplitude: float = 1.0, noise: float = 1e-2):
    """inputs ~ U([0,1]^d), z* ~ N(0, I), f* = chol(K*) z*, y ~ Poisson(exp(f*)) by
    inversion of Philox uniforms.  Returns (inputs, counts, f*)."""
    inputs = uniforms(seed, cfg, 0, n * d).reshape(n, d)
    z = normals(seed, cfg, 1, n)
    f = np.linalg.cholesky(gp_kernel(inputs, lengthscales[:d], amplitude, noise)) @ z
    u = uniforms(seed, cfg, 2, n)
    counts = np.zeros(n, dtype=np.int32)
    for i in range(n):
        lam = np.exp(f[i])
        k, p = 0, np.exp(-lam)
        cdf = p
        while u[i] > cdf and k < 10000:
            k += 1
            p *= lam / k
            cdf += p
        counts[i] = k
    return inputs, counts, f
