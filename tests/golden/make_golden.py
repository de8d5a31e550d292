"""Writes the golden vectors under tests/golden/ (run from the repository root).

Gradients come from torch binary64 autograd of the models' log densities written
directly from their definitions (independent of the oracle's analytic gradients);
Philox known-answer vectors are the Random123 philox4x32 R=10 entries the
reference's own test holds (test/uncertaintea/core/device_rng.jl:77-87); the
warmup-schedule vectors are the reference's (nuts_scheduler_and_backend.jl:609-618).
"""
import json
import math
from pathlib import Path

import numpy as np
import torch

torch.set_default_dtype(torch.float64)
OUT = Path(__file__).resolve().parent
LOG_2PI_HALF = 0.5 * math.log(2 * math.pi)


def normal_lpdf(x, mu, sigma):
    return -torch.log(torch.as_tensor(sigma)) - LOG_2PI_HALF - 0.5 * ((x - mu) / sigma) ** 2


def grad_of(f, q):
    q = torch.tensor(q, requires_grad=True)
    v = f(q)
    (g,) = torch.autograd.grad(v, q)
    return float(v), g.numpy()


def main():
    rng = np.random.default_rng(20261203)
    gold = {"philox_kat": [
        [[0x00000000, 0x00000000], [0, 0, 0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]],
        [[0xffffffff, 0xffffffff], [0xffffffff] * 4, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]],
        [[0xa4093822, 0x299f31d0], [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]]],
        "warmup_schedule": [[12, 5, [12], 0], [150, 15, [40, 90, 135], 15]]}

    # logistic regression, N=64, D=8
    x = rng.normal(size=(64, 8)).astype(np.float32)
    y = (rng.uniform(size=64) < 0.5).astype(np.float32)
    X = torch.tensor(x.astype(np.float64))
    Y = torch.tensor(y.astype(np.float64))
    cases = []
    for _ in range(4):
        q = rng.normal(size=8) * 0.7

        def f(b):
            eta = X @ b
            return normal_lpdf(b, 0.0, 1.0).sum() + (Y * eta - torch.nn.functional.softplus(eta)).sum()
        v, g = grad_of(f, q)
        cases.append({"params": q.tolist(), "logjoint": v, "gradient": g.tolist()})
    gold["logistic"] = {"x": x.tolist(), "y": y.tolist(), "prior_loc": 0.0, "prior_scale": 1.0, "cases": cases}

    # hierarchical non-centred normal, J=8 groups of ragged size
    sizes = [3, 1, 5, 4, 2, 6, 0, 3]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    hy = rng.normal(size=offsets[-1]).astype(np.float32) * 2.0
    sigma = rng.uniform(0.5, 2.0, size=8)
    HY = torch.tensor(hy.astype(np.float64))
    cases = []
    for _ in range(4):
        q = rng.normal(size=10) * 0.8

        def f(u):
            mu, ltau, r = u[0], u[1], u[2:]
            tau = torch.exp(ltau)
            lp = normal_lpdf(mu, 0.0, 5.0)
            lp = lp + math.log(2 / math.pi) - math.log(5.0) - torch.log1p((tau / 5.0) ** 2) + ltau
            lp = lp + normal_lpdf(r, 0.0, 1.0).sum()
            theta = mu + tau * r
            for j in range(8):
                obs = HY[offsets[j]:offsets[j + 1]]
                lp = lp + normal_lpdf(obs, theta[j], float(sigma[j])).sum()
            return lp
        v, g = grad_of(f, q)
        cases.append({"params": q.tolist(), "logjoint": v, "gradient": g.tolist()})
    gold["hier"] = {"y": hy.tolist(), "offsets": offsets.tolist(), "sigma": sigma.tolist(), "mu_loc": 0.0,
                    "mu_scale": 5.0, "tau_scale": 5.0, "cases": cases}

    # funnel, dim 5
    cases = []
    for _ in range(4):
        q = rng.normal(size=6)

        def f(u):
            v, xs = u[0], u[1:]
            return normal_lpdf(v, 0.0, 3.0) + normal_lpdf(xs, 0.0, torch.exp(v / 2)).sum()
        v, g = grad_of(f, q)
        cases.append({"params": q.tolist(), "logjoint": v, "gradient": g.tolist()})
    gold["funnel"] = {"dim": 5, "v_scale": 3.0, "cases": cases}

    (OUT / "golden.json").write_text(json.dumps(gold, indent=1))


if __name__ == "__main__":
    main()
