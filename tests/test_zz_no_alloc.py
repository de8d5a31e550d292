"""GPU: the allocation-free density entry point against ut_logjoint_gradient.

ut_logjoint_gradient_noalloc runs the same kernels on caller-owned device buffers
with its scratch in a fixed device array; its values and gradients must equal the
workspace entry point's bit for bit, on two calls in a row (the second with fewer
chains, so it reuses the scratch the first left behind), for the matrix-core GLM
plan (which uses the scratch) and for a wavefront-per-chain plan (which uses none).
Device buffers come from torch (plumbing only); the calls go through the C ABI."""
import ctypes as ct
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from uncertaintea_amd import _lib as L
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(model, params):
    import torch
    P, C = params.shape
    dev = torch.device("cuda:0")
    q = torch.from_numpy(np.ascontiguousarray(params.T, dtype=np.float32)).to(dev)  # chain-major = P x C col-major
    vals = torch.full((C,), float("nan"), dtype=torch.float32, device=dev)
    grads = torch.full((C, P), float("nan"), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    rc = L.lib().ut_logjoint_gradient_noalloc(model.plan, C, ct.c_void_p(q.data_ptr()), ct.c_void_p(vals.data_ptr()),
                                              ct.c_void_p(grads.data_ptr()), None)
    L.check(rc)
    torch.cuda.synchronize()
    return vals.cpu().numpy(), grads.cpu().numpy().T


def _check(model, Cs, seed):
    rng = np.random.default_rng(seed)
    for C in Cs:
        say("noalloc", model.name, "chains", C, "scratch bytes", L.lib().ut_noalloc_scratch_bytes(model.plan, C))
        params = (rng.standard_normal((model.num_params, C)) * 0.3).astype(np.float32)
        v_ref, g_ref = ut.device_batched_logjoint_gradient(model, params)
        v, g = _noalloc(model, params)
        assert np.array_equal(v.view(np.uint32), v_ref.view(np.uint32)), (model.name, C)
        assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), (model.name, C)


def test_noalloc_glm_mfma_matches_workspace():
    x, y, _ = S.logistic_data(4096, 256, seed=20261203, cfg=3, x_scale=1.0 / 16.0)
    _check(M.LogisticRegressionBF16(x, y), (512, 256), seed=1)


def test_noalloc_wavefront_plan_matches_workspace():
    yy, off, sig, _ = S.hierarchical_data(64, 32, seed=3)
    _check(M.HierarchicalNormalNC(yy, off, sig), (96, 40), seed=2)


def test_noalloc_rejects_oversize_and_gp():
    x, y, _ = S.logistic_data(1024, 256, seed=5, cfg=3, x_scale=1.0 / 16.0)
    model = M.LogisticRegressionBF16(x, y)
    C = 1 << 15
    need = L.lib().ut_noalloc_scratch_bytes(model.plan, C)
    say("noalloc oversize", C, need)
    assert need > (48 << 20)
    rc = L.lib().ut_logjoint_gradient_noalloc(model.plan, C, ct.c_void_p(16), ct.c_void_p(16), ct.c_void_p(16), None)
    assert rc == L.UT_E_DIM
    assert b"UT_NOALLOC_SCRATCH_BYTES" in L.lib().ut_last_error()
    inputs, counts, _ = S.gp_poisson_data(32, 2, seed=4)
    gp = M.GaussianProcessPoisson(inputs, counts)
    rc = L.lib().ut_logjoint_gradient_noalloc(gp.plan, 8, ct.c_void_p(16), ct.c_void_p(16), ct.c_void_p(16), None)
    assert rc == L.UT_E_UNSUPPORTED
