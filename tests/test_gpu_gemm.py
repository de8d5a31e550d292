"""Matrix-core GEMM (bf16 x bf16 -> fp32) against numpy on the same bf16 values.
Tolerance: fp32 accumulation of K products, |err| <= 1e-5 * K^0.5 * max|C| scale.
Small grids run the 64 x 64 workgroup tiles, grids of >= 256 128-wide tiles the
128 x 128 tiles."""
import sys

import numpy as np
import pytest

from uncertaintea_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k,splits", [(128, 128, 32, 1), (256, 384, 96, 1), (384, 128, 4096, 4),
                                          (128, 256, 512, 3), (2048, 2048, 64, 1), (1024, 1024, 96, 5)])
def test_gemm_matches_numpy(m, n, k, splits):
    print("gemm", m, n, k, splits, flush=True)
    rng = np.random.default_rng(m + n + k)
    a = L.to_bf16(rng.normal(size=(m, k)).astype(np.float32))
    bt = L.to_bf16(rng.normal(size=(n, k)).astype(np.float32))
    c = L.gemm_bf16(a, bt, splits)
    ref = L.from_bf16(a).astype(np.float64) @ L.from_bf16(bt).astype(np.float64).T
    err = np.max(np.abs(c - ref))
    assert err <= 2e-6 * np.sqrt(k) * max(1.0, np.max(np.abs(ref))), err


def test_gemm_layout_one_hot():
    """A single nonzero in A and B lands at exactly one output position."""
    print("gemm one-hot", flush=True)
    m = n = 128
    k = 32
    for (i, kk, j) in [(0, 0, 0), (5, 17, 70), (127, 31, 127), (33, 8, 96)]:
        a = np.zeros((m, k), np.float32)
        b = np.zeros((n, k), np.float32)
        a[i, kk] = 2.0
        b[j, kk] = 3.0
        c = L.gemm_bf16(L.to_bf16(a), L.to_bf16(b))
        nz = np.argwhere(c != 0)
        assert nz.tolist() == [[i, j]] and c[i, j] == 6.0
