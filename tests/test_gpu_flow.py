"""GPU: the affine-coupling flow preconditioner and the shared-adaptation HMC.

Tolerance: device flow forward / inverse / vector-Jacobian product / training
gradient against the oracle's binary64 flow evaluated with the same bf16-rounded
matrix-core operands.  The device's fp32 intermediates differ from binary64 by
~1e-7, which now and then moves an operand across a bf16 rounding boundary; such
flips change that operand by 2^-9 relative and propagate through later layers.
So a row whose operands saw no flip agrees to the arithmetic (|d| <= 1e-5 (1 + |ref|)), which
at least 90 % of the rows must do for positions (their row maximum) and log-determinants,
and every value within 2e-3 (the flips);
gradients normwise (|d|_2 <= 2e-3 |ref|_2).  Errors against the unrounded
binary64 flow are printed.  Bit-exact: shared-adaptation lock-step HMC against
the oracle's binary32 twin.  Statistical: a flow trained on exact funnel draws maps
held-out draws to a near-standard normal, and flow-preconditioned HMC on the
funnel matches the exact moments of v where plain HMC on the same budget mixes
far worse."""
import sys

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B
from uncertaintea_amd import diagnostics as D
from uncertaintea_amd import models as M
from uncertaintea_amd import synthetic as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _random_weights(flow, seed, scale=0.5):
    """Non-trivial weights in the flow's flat layout (padding entries stay zero)."""
    rng = np.random.default_rng(seed)
    w = flow.weights.copy()
    P, H = flow.num_params, flow.hidden
    w[:P] = rng.normal(size=P) * 0.5
    w[P:2 * P] = rng.normal(size=P) * 0.3
    for npass, nact, off in flow.layout():
        kp = -(-npass // 128) * 128
        op = -(-2 * nact // 128) * 128
        W2 = np.zeros((op, H), np.float32)
        W2[:2 * nact] = rng.normal(size=(2 * nact, H)) * scale / np.sqrt(H)
        w[off["W2"]:off["W2"] + op * H] = W2.ravel()
        b2 = np.zeros(op, np.float32)
        b2[:2 * nact] = rng.normal(size=2 * nact) * 0.1
        w[off["b2"]:off["b2"] + op] = b2
        w[off["b1"]:off["b1"] + H] = rng.normal(size=H) * 0.1
    return w


def _funnel_draws(n, dim, seed=1):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) * 3.0
    x = rng.normal(size=(n, dim)) * np.exp(v / 2.0)[:, None]
    return np.concatenate([v[:, None], x], axis=1).astype(np.float32)


@pytest.mark.parametrize("P", [101, 2, 128, 129])
def test_flow_kernels_against_oracle(P):
    """P = 101, 2 and 128 run the sampling path's 16x16x32-MFMA all-layer kernels (every layer
    with at most 64 passive and active coordinates; 128 is the widest), P = 129 the older
    all-layer kernels; the inverse and the training gradient run the per-step kernels."""
    say("flow kernels vs oracle", P)
    Lyr, H = 6, 256
    flow = ut.AffineCouplingFlow(P, Lyr, H, scale_clamp=5.0, seed=3)
    flow.weights = _random_weights(flow, 4, scale=0.2)
    o16 = B.OracleFlow.of(flow, bf16=True)
    o64 = B.OracleFlow.of(flow, bf16=False)
    rng = np.random.default_rng(5)
    z = rng.normal(size=(256, P)).astype(np.float32)
    q, ld = flow.forward(z)
    rq, rld = o16.forward(z)
    fq, fld = o64.forward(z)
    d_q = np.abs(q - rq) / (1 + np.abs(rq))
    d_ld = np.abs(ld - rld) / (1 + np.abs(rld))
    say(f"  forward: vs bf16-operand oracle q p99 {np.quantile(d_q, 0.99):.2e} max {d_q.max():.2e}, logdet p99 "
        f"{np.quantile(d_ld, 0.99):.2e} max {d_ld.max():.2e}; vs binary64 q max "
        f"{np.max(np.abs(q - fq) / (1 + np.abs(fq))):.2e} logdet max {np.max(np.abs(ld - fld)):.2e}")
    # a row's values depend on ~2000 bf16-rounded operands of that row, so a few percent of
    # the rows carry a flip (up to ~1e-3 on the row's later coordinates and its
    # log-determinant): at least 90 % of the rows must agree to the arithmetic (1e-5), all
    # of them to 2e-3
    assert np.median(d_q) <= 1e-6 and np.quantile(d_q.max(axis=1), 0.9) <= 1e-5 and d_q.max() <= 2e-3
    assert np.quantile(d_ld, 0.9) <= 1e-5 and d_ld.max() <= 2e-3
    zi, ldi = flow.inverse(q)
    rz, rldi = o16.inverse(q)
    d_z = np.abs(zi - rz) / (1 + np.abs(rz))
    d_ldi = np.abs(ldi - rldi) / (1 + np.abs(rldi))
    say(f"  inverse: z p99 {np.quantile(d_z, 0.99):.2e} max {d_z.max():.2e}, logdet max {d_ldi.max():.2e}; "
        f"round trip |z - f^-1(f(z))| max {np.max(np.abs(zi - z)):.2e}")
    assert np.quantile(d_z, 0.99) <= 1e-5 and d_z.max() <= 2e-3 and d_ldi.max() <= 2e-3
    gq = rng.normal(size=(256, P)).astype(np.float32)
    gz = flow.vjp(z, gq)
    _, _, rgz = o16.forward(z, gq)
    e_g = np.linalg.norm(gz - rgz) / np.linalg.norm(rgz)
    _, _, fgz = o64.forward(z, gq)
    say(f"  vjp: vs bf16-operand oracle {e_g:.2e}; vs binary64 {np.linalg.norm(gz - fgz) / np.linalg.norm(fgz):.2e}")
    assert e_g <= 2e-3
    qd = (rng.normal(size=(256, P)) * 1.3).astype(np.float32)
    loss, g = flow.loss_grad(qd)
    rloss, rg = o16.loss_grad(qd)
    floss, fg = o64.loss_grad(qd)
    e_w = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    say(f"  loss {loss:.6f} vs {rloss:.6f} (binary64 {floss:.6f}); weight gradient vs bf16-operand oracle {e_w:.2e}, "
        f"vs binary64 {np.linalg.norm(g - fg) / np.linalg.norm(fg):.2e}")
    assert abs(loss - rloss) <= 1e-4 * abs(rloss)
    assert e_w <= 2e-3


def test_elbo_gradient_against_oracle():
    """The device's ELBO weight gradient (model gradient held fixed) against the
    oracle's binary64 flow with bf16-rounded operands, fed the device's model
    gradient; the model gradient and the ELBO against the binary64 model at the
    oracle's f(z).  Row 0 is pushed to v = -1e4, where the funnel's logjoint
    overflows: it must be flagged invalid and carry no weight."""
    say("flow ELBO gradient vs oracle")
    dim = 20
    P = dim + 1
    model = M.NealsFunnel(dim, 3.0)
    flow = ut.AffineCouplingFlow(P, 6, 256, scale_clamp=5.0, seed=13)
    w = _random_weights(flow, 14, scale=0.2)
    w[:P] = 0.0  # base mu 0, log-scale small: f(z) stays in the funnel's finite range
    w[P:2 * P] = np.random.default_rng(15).normal(size=P).astype(np.float32) * 0.1
    flow.weights = w
    z = np.random.default_rng(16).normal(size=(256, P)).astype(np.float32)
    z[0, 0] = -1e4
    elbo, g, gq, valid = flow.elbo_grad(model, z)
    say(f"  valid {valid.sum()} of {valid.size}; elbo {elbo:.4f}")
    assert not valid[0] and valid[1:].all()
    assert np.all(gq[0] == 0.0) and np.all(np.isfinite(g))
    o16 = B.OracleFlow.of(flow, bf16=True)
    rg = o16.elbo_grad(z, gq, valid.astype(np.int32))
    e_w = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    say(f"  weight gradient vs bf16-operand oracle {e_w:.2e}")
    assert e_w <= 2e-3
    q, ld = o16.forward(z[1:])
    lj, grad = B.OracleModel(model).logjoint_grad_f64(np.ascontiguousarray(q.T))
    e_q = np.linalg.norm(gq[1:] - grad.T) / np.linalg.norm(grad)
    ref = np.mean(lj + 0.5 * np.sum(z[1:].astype(np.float64) ** 2, axis=1) + 0.5 * P * np.log(2 * np.pi) + ld)
    say(f"  model gradient vs binary64 {e_q:.2e}; elbo {elbo:.6f} vs {ref:.6f}")
    assert e_q <= 2e-3
    assert abs(elbo - ref) <= 1e-4 * (1 + abs(ref))


def test_elbo_fit_learns_funnel():
    """The reference's flow ADVI objective, fitted on the device from the identity
    flow: the sampled ELBO rises towards log Z = 0 (the funnel's density is
    normalised, so ELBO <= 0 up to sampling noise) and the pushed-forward v spreads
    towards its N(0, 3) marginal."""
    say("flow ELBO fit on the funnel")
    dim = 10
    model = M.NealsFunnel(dim, 3.0)
    flow = ut.AffineCouplingFlow(dim + 1, 6, 256, seed=17)
    out = flow.fit_elbo(model, num_particles=1024, num_iterations=600, learning_rate=1e-3, seed=18)
    h = out["elbo_history"]
    first, last = h[:20].mean(), h[-50:].mean()
    q, _ = flow.forward(np.random.default_rng(19).normal(size=(16384, dim + 1)).astype(np.float32))
    say(f"  elbo {first:.3f} -> {last:.3f} (best {out['best_elbo']:.3f}); pushed-forward v sd {q[:, 0].std():.3f} "
        f"(exact 3); grad norm {out['grad_norm_history'][:5].mean():.2f} -> {out['grad_norm_history'][-50:].mean():.2f}")
    assert np.all(np.isfinite(h))
    assert last > first + 1.0 and last < 0.1
    assert out["best_elbo"] >= h.max() - 1e-12
    assert q[:, 0].std() > 1.5


def test_flow_training_learns_funnel():
    say("flow training on exact funnel draws")
    dim = 10
    draws = _funnel_draws(65536, dim)
    flow = ut.AffineCouplingFlow(dim + 1, 6, 256, seed=7)
    trace = flow.fit(draws, steps=1500, batch=1024, learning_rate=3e-4, seed=8)
    entropy = 0.5 * np.log(2 * np.pi * np.e * 9.0) + dim * 0.5 * np.log(2 * np.pi * np.e)
    first, last = trace[:20].mean(), trace[-100:].mean()
    say(f"  loss {first:.3f} -> {last:.3f} (true entropy {entropy:.3f})")
    assert np.all(np.isfinite(trace))
    assert last < entropy + 0.5 * (first - entropy)
    # maximum likelihood makes the inverse map of held-out exact draws standard normal
    zi, _ = flow.inverse(_funnel_draws(16384, dim, seed=2))
    m, sd = zi.mean(axis=0), zi.std(axis=0)
    z = np.random.default_rng(9).normal(size=(16384, dim + 1)).astype(np.float32)
    q, _ = flow.forward(z)
    say(f"  inverse of held-out draws: |mean| max {np.abs(m).max():.3f}, sd range [{sd.min():.3f}, {sd.max():.3f}]; "
        f"pushed-forward v: mean {q[:, 0].mean():.3f} sd {q[:, 0].std():.3f} (exact 0, 3)")
    assert np.abs(m).max() < 1.0 and sd.min() > 0.5 and sd.max() < 2.0


def test_hmc_shared_bitexact_against_twin():
    say("shared-adaptation HMC bitexact")
    x, y, _ = S.logistic_data(1000, 25)
    model = M.LogisticRegression(x, y)
    dev = ut.batched_hmc(model, num_chains=64, num_samples=30, num_warmup=80, num_leapfrog_steps=8, step_size=0.05,
                         per_chain_adaptation=False, seed=13)
    orc = B.OracleModel(model).run("hmc_shared", "twin", 64, num_samples=30, num_warmup=80, num_leapfrog_steps=8,
                                   step_size=0.05, seed=13, threads=8)
    assert np.array_equal(dev.unconstrained_samples.view(np.uint32), orc["samples"].view(np.uint32))
    for k_dev, k_orc in (("step_size", "step_size"), ("mass_matrix", "inverse_mass"), ("accepted", "accepted"),
                         ("divergent", "divergent"), ("logjoint_values", "logjoint"),
                         ("acceptance_stats", "accept_stat"), ("integration_steps", "integration_steps")):
        assert np.array_equal(getattr(dev, k_dev), orc[k_orc]), k_dev


@pytest.mark.parametrize("P", [101, 2, 300, 129])
def test_fused_layers_match_unfused(monkeypatch, P):
    """The sampling path's all-layer kernels against the per-step kernels (UT_FLOW_FUSED=0).
    P = 101 and 2 run the 16x16x32-MFMA kernels, whose tiles sum in another order: forward
    values, log-determinants and the sampler's VJP agree to the flow tolerance (fp32 rounding
    and the occasional bf16 flip of an operand: 99th percentile <= 1e-5 (1 + |ref|), maximum
    2e-3; VJP normwise 2e-3), and an HMC run on the funnel with each path samples the same law
    (v's mean and sd within 5 combined MCSE).  P = 129 mixes per-layer LDS-staged kernels (even
    layers) with per-step ones (odd layers, 2 na = 130) and P = 300 exercises the per-step
    fallback: those match bit for bit."""
    say("fused vs unfused flow layers", P)
    flows = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("UT_FLOW_FUSED", mode)
        flows[mode] = ut.AffineCouplingFlow(P, 6, 256, seed=5)
    w = _random_weights(flows["0"], 6, scale=0.3)
    for f in flows.values():
        f.weights = w
    rng = np.random.default_rng(7)
    z = rng.normal(size=(256, P)).astype(np.float32)
    gq = rng.normal(size=(256, P)).astype(np.float32)
    q0, l0 = flows["0"].forward(z)
    q1, l1 = flows["1"].forward(z)
    g0 = flows["0"].vjp(z, gq)
    g1 = flows["1"].vjp(z, gq)
    if P in (300, 129):
        assert np.array_equal(q0.view(np.uint32), q1.view(np.uint32))
        assert np.array_equal(l0.view(np.uint32), l1.view(np.uint32))
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32))
        return
    d_q = np.abs(q1 - q0) / (1 + np.abs(q0))
    d_l = np.abs(l1 - l0) / (1 + np.abs(l0))
    e_g = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    say(f"  q p99 {np.quantile(d_q, 0.99):.2e} max {d_q.max():.2e}; logdet max {d_l.max():.2e}; vjp {e_g:.2e}")
    assert np.quantile(d_q, 0.99) <= 1e-5 and d_q.max() <= 2e-3
    assert np.quantile(d_l, 0.99) <= 1e-5 and d_l.max() <= 2e-3
    assert e_g <= 2e-3
    if P == 101:
        model = M.NealsFunnel(100, 3.0)
        runs = [ut.batched_hmc(model, num_chains=256, num_samples=300, num_warmup=200, num_leapfrog_steps=5,
                               per_chain_adaptation=False, flow=f, seed=23) for f in (flows["0"], flows["1"])]
        v = [r.unconstrained_samples[:, :, :1].astype(np.float64) for r in runs]
        for stat in (lambda a: a, lambda a: (a - a.mean()) ** 2):
            x0, x1 = stat(v[0]), stat(v[1])
            z_ = abs(x0.mean() - x1.mean()) / np.sqrt(float(D.mcse_mean(x0)[0]) ** 2 + float(D.mcse_mean(x1)[0]) ** 2)
            say(f"  HMC v statistic, per-step vs all-layer kernels: z {z_:.2f}")
            assert z_ < 5.0


def test_flow_graph_replay_matches_eager(monkeypatch):
    """The flow-wrapped batched gradient is captured once per buffer set as a HIP graph
    and replayed at every leapfrog step, with the samplers' own kernels in between:
    whole shared-HMC and ChEES runs must equal the eager launches bit for bit."""
    say("flow graph replay vs eager")
    model = M.NealsFunnel(10, 3.0)
    flow = ut.AffineCouplingFlow(model.num_params, 4, 256, seed=3)
    flow.weights = _random_weights(flow, 4, scale=0.3)
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("UT_FLOW_GRAPHS", mode)
        h = ut.batched_hmc(model, num_chains=128, num_samples=20, num_warmup=30, num_leapfrog_steps=6,
                           per_chain_adaptation=False, flow=flow, seed=21)
        c = ut.batched_chees(model, num_chains=128, num_samples=20, num_warmup=30, flow=flow, seed=22)
        runs[mode] = (h, c)
    for a, b in zip(runs["0"], runs["1"]):
        assert np.array_equal(a.unconstrained_samples.view(np.uint32), b.unconstrained_samples.view(np.uint32))
        assert np.array_equal(a.logjoint_values, b.logjoint_values)
        assert np.array_equal(a.step_size, b.step_size)
        assert np.array_equal(a.integration_steps, b.integration_steps)


@pytest.mark.parametrize("dim,layers", [(10, 4), (100, 6), (127, 8)])
def test_fused_leapfrog_step_matches_separate_kernel(monkeypatch, dim, layers):
    """With a flow on the 16-row kernels the lock-step samplers run each leapfrog step of
    Neal's funnel as one kernel (forward, model, backward, kick and drift; a trajectory's
    last gradient with the log-determinant instead of the step); with UT_FUSED_STEP=0 and
    UT_FUSED_MODEL=0 they run the separate forward, density, backward and step kernels.
    Same arithmetic: whole shared-HMC and ChEES runs must agree bit for bit.  The fused
    kernel keeps the hidden activations of as many of the last layers as LDS holds and
    recomputes the others: all of them at dim 10 and at the benchmark's dim 100, four of
    eight at dim 127 (P = 128, the widest)."""
    say("fused leapfrog step and row model", dim, layers)
    model = M.NealsFunnel(dim, 3.0)
    flow = ut.AffineCouplingFlow(model.num_params, layers, 256, seed=5)
    flow.weights = _random_weights(flow, layers, scale=0.2)
    runs = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("UT_FUSED_STEP", mode)
        monkeypatch.setenv("UT_FUSED_MODEL", mode)
        h = ut.batched_hmc(model, num_chains=256, num_samples=20, num_warmup=30, num_leapfrog_steps=6,
                           per_chain_adaptation=False, flow=flow, seed=23)
        c = ut.batched_chees(model, num_chains=256, num_samples=20, num_warmup=30, flow=flow, seed=24)
        runs[mode] = (h, c)
    for a, b in zip(runs["0"], runs["1"]):
        assert np.array_equal(a.unconstrained_samples.view(np.uint32), b.unconstrained_samples.view(np.uint32))
        assert np.array_equal(a.logjoint_values, b.logjoint_values)
        assert np.array_equal(a.step_size, b.step_size)
        assert np.array_equal(a.integration_steps, b.integration_steps)
        assert np.array_equal(a.divergent, b.divergent)


def test_flow_hmc_funnel_gains_over_plain():
    """NeuTra on Neal's funnel (dim 10): the same shared-adaptation HMC budget with
    and without a flow trained on exact funnel draws."""
    say("flow hmc on the funnel")
    dim, C, W, Sn = 10, 512, 300, 400
    model = M.NealsFunnel(dim, 3.0)
    flow = ut.AffineCouplingFlow(dim + 1, 6, 256, seed=17)
    flow.fit(_funnel_draws(65536, dim, seed=18), steps=1500, batch=1024, learning_rate=3e-4, seed=19)
    res = {}
    for name, fl in (("plain", None), ("flow", flow)):
        r = ut.batched_hmc(model, num_chains=C, num_samples=Sn, num_warmup=W, num_leapfrog_steps=10,
                           per_chain_adaptation=False, flow=fl, seed=21)
        v = r.unconstrained_samples[:, :, 0].astype(np.float64)
        ess_v = float(np.asarray(D.ess_bulk(r.unconstrained_samples[:, :, :1]))[0])
        res[name] = (v.mean(), v.std(), ess_v, float(np.max(np.asarray(D.rhat(r.unconstrained_samples)))),
                     r.divergence_rate(), float(r.step_size[0]))
        say(f"  {name}: v mean {res[name][0]:.3f} sd {res[name][1]:.3f} ess(v) {ess_v:.0f} rhat {res[name][3]:.3f} "
            f"div {res[name][4]:.4f} eps {res[name][5]:.4f}")
    m, sd, ess, rh, div, _ = res["flow"]
    assert abs(m) < 4 * 3.0 / np.sqrt(ess)
    assert abs(sd - 3.0) < 0.25
    # what the network gains: more effective draws of v, fewer divergences, better R-hat
    assert ess > 2.0 * res["plain"][2]
    assert div < res["plain"][4] and rh < res["plain"][3]


def test_flow_trained_on_the_samplers_draws_gains_over_plain_cfg4():
    """What the network gains when trained the NeuTra way, on the sampler's own draws,
    at the benchmark's shape (Neal's funnel D = 100, 4096 chains, L = 10): plain HMC
    with per-chain adaptation runs first; the flow is fitted by maximum likelihood to
    those draws, then re-fitted to a short flow run's draws (iterative NeuTra); the
    flow-preconditioned shared-adaptation HMC is compared with the plain run on bulk
    ESS of v per gradient evaluation, R-hat and the spread of v (exactly 3)."""
    say("flow trained on the plain run's draws vs plain, funnel D=100")
    dim, C, L = 100, 4096, 10
    model = M.NealsFunnel(dim, 3.0)
    plain = ut.batched_hmc(model, num_chains=C, num_samples=1000, num_warmup=1000, num_leapfrog_steps=L, seed=31)
    flow = ut.AffineCouplingFlow(dim + 1, 6, 256, seed=32)
    flow.fit(plain.unconstrained_samples.reshape(-1, dim + 1), steps=2000, batch=4096, seed=33)
    mid = ut.batched_hmc(model, num_chains=C, num_samples=500, num_warmup=300, num_leapfrog_steps=L,
                         per_chain_adaptation=False, flow=flow, seed=34, target_accept=0.9)
    flow.fit(mid.unconstrained_samples.reshape(-1, dim + 1), steps=2000, batch=4096, seed=35, init_base=False)
    fr = ut.batched_hmc(model, num_chains=C, num_samples=1000, num_warmup=500, num_leapfrog_steps=L,
                        per_chain_adaptation=False, flow=flow, seed=36, target_accept=0.9)
    out = {}
    for name, r in (("plain", plain), ("flow", fr)):
        ess_v = float(np.asarray(D.ess_bulk(r.constrained_samples[:, :, :1]))[0])
        rh = float(np.max(np.asarray(D.rhat(r.constrained_samples))))
        sd_v = float(r.constrained_samples[:, :, 0].std())
        out[name] = (ess_v / r.sampling_gradient_evals, rh, sd_v, ess_v / r.sampling_seconds)
        say(f"  {name}: ess(v) {ess_v:.0f} per gradient {out[name][0]:.2e} ({out[name][3]:.0f}/s) rhat {rh:.4f} "
            f"sd(v) {sd_v:.3f}")
    assert out["flow"][0] > 3.0 * out["plain"][0]
    assert out["flow"][1] < out["plain"][1]
    assert abs(out["flow"][2] - 3.0) < abs(out["plain"][2] - 3.0)


def test_elbo_flow_hmc_matches_funnel_law():
    """A flow fitted by the ELBO alone (no draws) preconditions shared-adaptation HMC
    on the funnel (dim 10) so that the draws follow the exact law: v ~ N(0, 3) and
    u_i = x_i e^(-v/2) ~ N(0, 1), means and variances within 5 Monte Carlo standard
    errors.  R-hat is bounded loosely (1.05): at 1000 draws x 1024 chains it measures
    the per-chain effective draws (~20-40 here), not the law."""
    say("ELBO-fitted flow hmc on the funnel")
    dim, C = 10, 1024
    model = M.NealsFunnel(dim, 3.0)
    flow = ut.AffineCouplingFlow(dim + 1, 6, 256, seed=23)
    flow.fit_elbo(model, num_particles=1024, num_iterations=3000, learning_rate=1e-3, seed=24)
    flow.fit_elbo(model, num_particles=1024, num_iterations=1000, learning_rate=2e-4, seed=25)
    r = ut.batched_hmc(model, num_chains=C, num_samples=1000, num_warmup=300, num_leapfrog_steps=10,
                       per_chain_adaptation=False, flow=flow, seed=26, target_accept=0.97)
    a = r.constrained_samples.astype(np.float64)
    u = a.copy()
    u[..., 1:] *= np.exp(-0.5 * a[..., :1])
    exact_var = np.concatenate([[9.0], np.ones(dim)])
    zm = np.abs(u.mean(axis=(0, 1))) / np.asarray(D.mcse_mean(u))
    dev = u ** 2
    zv = np.abs(dev.mean(axis=(0, 1)) - exact_var) / np.asarray(D.mcse_mean(dev))
    rh = float(np.max(np.asarray(D.rhat(r.constrained_samples))))
    say(f"  v mean {u[..., 0].mean():.3f} sd {u[..., 0].std():.3f}; max |z| mean {zm.max():.2f} var {zv.max():.2f}; "
        f"rhat {rh:.4f}; step {float(r.step_size[0]):.3f}; div {r.divergence_rate():.4f}")
    assert zm.max() < 5.0 and zv.max() < 5.0
    assert rh < 1.05
