"""The oracle's latent-GP Poisson model (CPU) against torch's binary64 autograd
through torch.linalg.cholesky (an independent library implementation of the
factorisation and its derivative), and its Cholesky against numpy."""
import numpy as np
import pytest
import torch

from oracle import binding as B
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S


def _torch_logjoint(q, X, y, noise, ls_loc=0.0, ls_scale=0.5, amp=1.0):
    D = X.shape[1]
    q = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    Xt = torch.tensor(X, dtype=torch.float64)
    yt = torch.tensor(y, dtype=torch.float64)
    ls = torch.exp(q[:D])
    a = torch.exp(q[D])
    z = q[D + 1:]
    xs = Xt / ls
    sq = ((xs[:, None, :] - xs[None, :, :]) ** 2).sum(-1)
    K = a * a * torch.exp(-0.5 * sq) + (noise ** 2 + 1e-8) * torch.eye(X.shape[0], dtype=torch.float64)
    L = torch.linalg.cholesky(K)
    f = L @ z
    ll = (yt * f - torch.exp(f) - torch.lgamma(yt + 1)).sum()
    lp = (-0.5 * ((q[:D] - ls_loc) / ls_scale) ** 2 - np.log(ls_scale) - 0.5 * np.log(2 * np.pi)).sum()
    lp = lp + np.log(2) - 0.5 * np.log(2 * np.pi) - np.log(amp) - 0.5 * (a / amp) ** 2 + q[D]
    lp = lp + (-0.5 * z * z - 0.5 * np.log(2 * np.pi)).sum()
    total = ll + lp
    total.backward()
    return total.item(), q.grad.numpy(), float(2 * torch.log(torch.diagonal(L.detach())).sum())


@pytest.mark.parametrize("n,d", [(32, 2), (64, 2), (32, 1)])
def test_gp_logjoint_and_gradient_match_torch(n, d):
    X, y, _ = S.gp_poisson_data(n, d, seed=11)
    model = M.GaussianProcessPoisson(X, y, noise=1e-2)
    om = B.OracleModel(model)
    rng = np.random.default_rng(n + d)
    params = np.concatenate([rng.normal(size=(d + 1, 3)) * 0.4, rng.normal(size=(n, 3))], axis=0)
    lj, g, ld = om.gp_eval(params)
    for c in range(3):
        tl, tg, tld = _torch_logjoint(params[:, c], X, y, 1e-2)
        assert lj[c] == pytest.approx(tl, rel=1e-11, abs=1e-9)
        assert np.allclose(g[:, c], tg, rtol=1e-8, atol=1e-8)
        assert ld[c] == pytest.approx(tld, rel=1e-10, abs=1e-9)
