"""Benchmark: min bulk-ESS/s at R-hat < 1.01 of batched NUTS on one MI355X.

Default workload (`--workload hier_nuts`, the line the driver records):

Workload: the non-centred hierarchical normal (J = 64 groups x 32 observations,
sigma_j ~ U(0.5, 2)), batched_nuts with 1024 chains, max tree depth 10, binary32
state with binary64 energies; W warmup + K sampling iterations.  One "step" is one
NUTS iteration of all 1024 chains.  `value` is min over parameters of bulk ESS
(rank-normalised, constrained space) divided by the sampling wall time (inputs and
chain state resident in HBM; the time includes copying the draws back to the host).

Also reported: gradient evaluations/s; the roofline of the dominant kernel (the
persistent NUTS kernel, FP32 VALU bound) from HIP-event device time; and the CPU
baseline -- the oracle's binary64 NUTS on the host cores, timed on a bounded number
of sampling iterations from the GPU run's adapted state, giving CPU s/iteration and
a projected ESS/s (GPU ESS per iteration / CPU s per iteration).

`--workload glm_chees`: the large logistic regression (N = 131072 rows, D = 256,
X ~ N(0, 1/D) stored in bf16, 4096 chains) under batched_chees with the fused
matrix-core gradient; roofline against the bf16 MFMA peak with the algorithmic
4 N D flops per chain-gradient, CPU baseline = the binary64 oracle gradient on the
host cores (projected: CPU s per batch gradient x mean leapfrog steps).

`--workload funnel_flow`: Neal's funnel D = 100 (P = 101), 4096 chains, batched_hmc
L = 10: plain (per-chain adaptation, persistent kernel) and flow-preconditioned
(a 6-layer, width-256 affine-coupling flow trained on the GPU on the plain run's
2000 draws per chain, then shared-adaptation lock-step HMC in the flow's base
space).  `value` is the flow run's min bulk-ESS/s; the plain run is reported beside
it.  Roofline: the flow's conditioner GEMMs against the bf16 MFMA peak.

`--workload advi`: batched_advi on the same hierarchical model (P = 66), 1024
particles per iteration, `--guide` meanfield / fullrank (default) / flow (4 layers,
width 256), standard ELBO, lr 0.01; `value` is Adam iterations/s of the whole call
(one iteration = draw, guide forward, batched model gradient over the particles,
ELBO reduction, guide backward, clip, Adam).  CPU baseline: the same iteration in
numpy around the oracle's binary64 gradient (full-rank guide), bounded in time.

`--workload gp_meads`: the latent GP with Poisson counts (N = 512 inputs in
[0,1]^2, ARD squared-exponential kernel, P = 515), batched_meads with 1024 chains,
binary64 batched Cholesky with hyperparameter tangents; without and with a flow
trained on the plain run's draws.  Roofline: the factorisation's binary64 flops
(7 N^3 / 3 per chain-gradient with three tangents) against the FP64 peak.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "min bulk-ESS/sec at R̂<1.01, plain vs flow-preconditioned HMC, 1 MI355X"
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
# Algorithmic work per NUTS leaf of the J=64 x 32 model: one chain-gradient
# (6 N + 10 J + 20 = 12,948 flops) plus kick/drift/U-turn/energy (12 P = 792 flops).
FLOPS_PER_GRADIENT = 6 * 2048 + 10 * 64 + 20
FLOPS_PER_LEAF_EXTRA = 12 * 66


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def moments_vs(D, a, b) -> dict:
    """Posterior means and variances of draws `a` against `b` (chains x draws x params)
    in units of their combined Monte Carlo standard errors: z = (m_a - m_b) / sqrt(se_a^2
    + se_b^2) for the mean and for the variance (the latter from the MCSE of the mean of
    the squared deviations).  max|z| below 4 over a few hundred parameters is what
    agreement within MCSE looks like."""
    def mom(x):
        m = x.mean(axis=(0, 1))
        dev = (x - m) ** 2
        return m, np.asarray(D.mcse_mean(x)), dev.mean(axis=(0, 1)), np.asarray(D.mcse_mean(dev))
    ma, sa, va, sva = mom(np.asarray(a, dtype=np.float64))
    mb, sb, vb, svb = mom(np.asarray(b, dtype=np.float64))
    zm = np.abs(ma - mb) / np.sqrt(sa ** 2 + sb ** 2)
    zv = np.abs(va - vb) / np.sqrt(sva ** 2 + svb ** 2)
    return {"max_abs_z_mean": float(np.nanmax(zm)), "max_abs_z_var": float(np.nanmax(zv)),
            "within_mcse": bool(np.nanmax(zm) < 4.0 and np.nanmax(zv) < 4.0)}


def moments_vs_exact(D, a, mean, var) -> dict:
    """Means and variances of draws `a` (chains x draws x params) against known exact
    values, in units of the draws' Monte Carlo standard errors."""
    x = np.asarray(a, dtype=np.float64)
    m = x.mean(axis=(0, 1))
    dev = (x - np.asarray(mean)) ** 2
    zm = np.abs(m - mean) / np.asarray(D.mcse_mean(x))
    zv = np.abs(dev.mean(axis=(0, 1)) - var) / np.asarray(D.mcse_mean(dev))
    return {"max_abs_z_mean": float(np.nanmax(zm)), "max_abs_z_var": float(np.nanmax(zv)),
            "within_mcse": bool(np.nanmax(zm) < 4.0 and np.nanmax(zv) < 4.0)}


def _funnel_standardised(draws):
    """(v, x_1 e^(-v/2), ...) per draw: exactly N(0, 9) x N(0, 1)^dim under the funnel."""
    a = np.asarray(draws, dtype=np.float64)
    out = a.copy()
    out[..., 1:] *= np.exp(-0.5 * a[..., :1])
    return out


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("hier_nuts", "glm_chees", "funnel_flow", "gp_meads", "advi"),
                    default="hier_nuts")
    ap.add_argument("--guide", choices=("meanfield", "fullrank", "flow"), default="fullrank",
                    help="advi: the variational guide")
    ap.add_argument("--steps", type=int, default=None, help="sampling iterations")
    ap.add_argument("--warmup", type=int, default=None, help="warmup iterations")
    ap.add_argument("--chains", type=int, default=None)
    ap.add_argument("--flow", action="store_true",
                    help="glm_chees: also train a flow on the plain run's draws and run ChEES with it")
    ap.add_argument("--forward-velocity", action="store_true",
                    help="glm_chees: ChEES criterion with the forward end velocity (not the reference's sign)")
    ap.add_argument("--flow-rounds", type=int, default=3,
                    help="funnel_flow / gp_meads: train the flow on the plain run's draws, then re-train it "
                         "on each flow run's draws this many times in all (iterative NeuTra)")
    ap.add_argument("--flow-target-accept", type=float, default=None,
                    help="funnel_flow: target acceptance of the flow runs' step-size adaptation (default 0.97 "
                         "after an ELBO fit: with L fixed at 10 the smaller step keeps the trajectory near a "
                         "quarter period of the near-Gaussian base; 0.9 after maximum likelihood, where 0.8 "
                         "gives ~1 %% divergent transitions in the neck)")
    ap.add_argument("--flow-fit", choices=("elbo", "ml"), default="elbo",
                    help="funnel_flow: fit the flow by the reparameterised ELBO against the model (the "
                         "reference's flow ADVI objective; no draws needed) or by maximum likelihood on the "
                         "plain run's draws with --flow-rounds re-training rounds")
    ap.add_argument("--elbo-schedule", default="30000:1e-3,10000:3e-4,10000:1e-4",
                    help="funnel_flow --flow-fit elbo: stages iterations:learning_rate, 4096 particles each")
    ap.add_argument("--final-steps", type=int, default=None,
                    help="funnel_flow / gp_meads: sampling iterations of the reported flow run (funnel default: "
                         "2 x --steps after an ELBO fit, --steps after maximum likelihood; GP default: --steps)")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "repeats": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(chains shard across GPUs without communication; see the design notes)", file=sys.stderr)
        return 2

    if args.workload == "glm_chees":
        return bench_glm_chees(args)
    if args.workload == "funnel_flow":
        return bench_funnel_flow(args)
    if args.workload == "gp_meads":
        return bench_gp_meads(args)
    if args.workload == "advi":
        return bench_advi(args)
    args.steps = 3000 if args.steps is None else args.steps
    args.warmup = 1000 if args.warmup is None else args.warmup
    args.chains = 1024 if args.chains is None else args.chains

    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D
    from uncertaintea_amd import synthetic as S

    y, off, sig, _ = S.hierarchical_data(64, 32, seed=args.seed)
    model = ut.HierarchicalNormalNC(y, off, sig, mu_loc=0.0, mu_scale=5.0, tau_scale=5.0)
    log(f"bench: batched_nuts hierarchical J=64x32 chains={args.chains} warmup={args.warmup} steps={args.steps}")
    t0 = time.perf_counter()
    res = ut.batched_nuts(model, num_chains=args.chains, num_samples=args.steps, num_warmup=args.warmup,
                          max_tree_depth=10, seed=args.seed)
    wall = time.perf_counter() - t0
    log(f"bench: sampler done warmup {res.warmup_seconds:.3f}s sampling {res.sampling_seconds:.3f}s wall {wall:.3f}s")
    draws = res.constrained_samples
    rh, eb = D.rhat_and_ess_bulk(draws)
    min_ess = float(np.nanmin(eb))
    rhat_max = float(np.nanmax(rh))
    value = min_ess / res.sampling_seconds
    log(f"bench: rhat max {rhat_max:.4f} min bulk ESS {min_ess:.1f} -> {value:.1f} ESS/s")

    evals = res.sampling_gradient_evals
    flops = evals * (FLOPS_PER_GRADIENT + FLOPS_PER_LEAF_EXTRA)
    ksec = res.sampling_kernel_seconds
    launches = max(res.sampling_launches, 1)
    achieved = flops / ksec / 1e12 if ksec > 0 else 0.0
    traffic = None
    pmc = ROOT / "profiles" / "k_nuts_hier_cfg2_pmc.json"
    if pmc.exists():  # HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE runs
        traffic = json.loads(pmc.read_text()).get("traffic_bytes_per_launch")
    roofline = {"bound": "mfma", "compute": "fp32 vector ALU (no HBM traffic in the tree)", "achieved": achieved,
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "kernel": "k_nuts<HierModel>", "launches": res.sampling_launches,
                "avg_launch_ms": 1e3 * ksec / launches, "flops_per_launch": flops / launches}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        cores = cpu_cores()
        pos = res.unconstrained_samples[:, -1, :].astype(np.float64).T.copy()
        it = args.warmup + args.steps + 1
        n_iter, n_evals, t_cpu = 0, 0, 0.0
        while t_cpu < args.cpu_seconds and n_iter < 5000:
            ts = time.perf_counter()
            ev, pos = B.nuts_sample_from(om, "f64", pos, res.step_size, res.mass_matrix, it + n_iter, 1,
                                         seed=args.seed, threads=cores)
            t_cpu += time.perf_counter() - ts
            n_iter += 1
            n_evals += ev
        s_per_iter = t_cpu / n_iter
        ess_per_iter = min_ess / args.steps
        cpu = {"value": ess_per_iter / s_per_iter, "unit": "ESS/s (projected)", "cores": cores, "kind": "port",
               "sample": f"{n_iter} sampling iterations x {args.chains} chains of the binary64 oracle NUTS from the "
                         f"GPU run's adapted state ({t_cpu:.1f} s); projected = GPU bulk ESS per iteration / "
                         f"CPU s per iteration",
               "cpu_s_per_iteration": s_per_iter, "cpu_gradient_evals_per_s": n_evals / t_cpu}
        log(f"bench: cpu {s_per_iter:.3f} s/iter on {cores} threads -> projected {cpu['value']:.1f} ESS/s")

    line = {
        "metric": METRIC, "value": value, "unit": "ESS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * res.sampling_seconds / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": "hierarchical normal non-centred J=64x32, batched_nuts persistent, 1024 chains, "
                               "max depth 10, fp32 state + fp64 energies",
                   "chains": args.chains, "num_params": model.num_params, "max_tree_depth": 10},
        "rhat_max": rhat_max, "rhat_ok": rhat_max < 1.01, "min_bulk_ess": min_ess,
        "sampling_seconds": res.sampling_seconds, "warmup_seconds": res.warmup_seconds,
        "gradient_evals_per_s": evals / res.sampling_seconds,
        "mean_tree_depth": float(res.tree_depths.mean()), "divergence_rate": res.divergence_rate(),
        "roofline": roofline, "cpu_baseline": cpu,
        "vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_funnel_flow(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D

    steps = 2000 if args.steps is None else args.steps
    warmup = 1000 if args.warmup is None else args.warmup
    chains = 4096 if args.chains is None else args.chains
    dim, L, H, layers = 100, 10, 256, 6
    model = ut.NealsFunnel(dim, 3.0)

    def summarise(res, label):
        rh, eb = D.rhat_and_ess_bulk(res.constrained_samples)
        out = {"min_bulk_ess": float(np.nanmin(eb)), "rhat_max": float(np.nanmax(rh)),
               "ess_bulk_v": float(eb[0]), "sampling_seconds": res.sampling_seconds,
               "warmup_seconds": res.warmup_seconds, "value": float(np.nanmin(eb)) / res.sampling_seconds,
               "gradient_evals_per_s": res.sampling_gradient_evals / res.sampling_seconds,
               "divergence_rate": res.divergence_rate(), "step_size": float(np.median(res.step_size)),
               "v_mean": float(res.unconstrained_samples[:, :, 0].mean()),
               "v_sd": float(res.unconstrained_samples[:, :, 0].std())}
        out["rhat_ok"] = out["rhat_max"] < 1.01
        log(f"bench: {label}: min ESS {out['min_bulk_ess']:.1f} rhat {out['rhat_max']:.4f} -> {out['value']:.1f} "
            f"ESS/s; v mean {out['v_mean']:.3f} sd {out['v_sd']:.3f}; div {out['divergence_rate']:.4f}")
        return out

    log(f"bench: funnel D={dim} plain batched_hmc chains={chains} warmup={warmup} steps={steps}")
    plain = ut.batched_hmc(model, num_chains=chains, num_samples=steps, num_warmup=warmup, num_leapfrog_steps=L,
                           seed=args.seed)
    plain_s = summarise(plain, "plain")
    flow = ut.AffineCouplingFlow(model.num_params, layers, H, seed=args.seed)
    elbo = args.flow_fit == "elbo"
    if args.flow_target_accept is None:
        args.flow_target_accept = 0.97 if elbo else 0.9
    # 4096 chains need ~50 effective draws per chain for split R-hat < 1.01; the ELBO
    # run reaches 36-80 per 2000 draws, so its reported run samples twice as many
    final_steps = args.final_steps if args.final_steps else (2 * steps if elbo else steps)
    if elbo:
        return _funnel_elbo_flow(args, ut, D, model, flow, plain, plain_s, summarise, steps, warmup, chains, dim, L,
                                 H, layers, final_steps)
    draws = plain.unconstrained_samples.reshape(-1, model.num_params)
    t0 = time.perf_counter()
    train_steps = 2000
    trace = flow.fit(draws, steps=train_steps, batch=4096, seed=args.seed)
    t_train = time.perf_counter() - t0
    log(f"bench: flow trained on {draws.shape[0]} draws, {train_steps} steps in {t_train:.2f}s, "
        f"loss {trace[:20].mean():.2f} -> {trace[-100:].mean():.2f}")
    del draws
    nrounds = max(1, args.flow_rounds)
    # the reported (last) flow run samples final_steps per chain (default: steps).  With
    # 4096 chains split R-hat < 1.01 needs ~50 effective draws per chain; 4 x steps gives
    # R-hat 1.019 here, ESS grows 1.7x for 4x the draws
    fres = ut.batched_hmc(model, num_chains=chains, num_samples=final_steps if nrounds == 1 else steps,
                          num_warmup=warmup // 2, num_leapfrog_steps=L, per_chain_adaptation=False, flow=flow,
                          seed=args.seed + 1, target_accept=args.flow_target_accept)
    flow_s = summarise(fres, "flow")
    rounds = [dict(flow_s, training_seconds=t_train)]
    for r in range(2, nrounds + 1):
        # re-train (warm start) on the flow run's draws, which reach the neck the plain run missed
        draws = fres.unconstrained_samples.reshape(-1, model.num_params)
        t0 = time.perf_counter()
        # warm start, twice the first round's steps
        trace = flow.fit(draws, steps=2 * train_steps, batch=4096, seed=args.seed + r, init_base=False)
        t_r = time.perf_counter() - t0
        t_train += t_r
        del draws
        log(f"bench: flow round {r} re-trained in {t_r:.2f}s, loss {trace[:20].mean():.2f} -> {trace[-100:].mean():.2f}")
        fres = ut.batched_hmc(model, num_chains=chains, num_samples=final_steps if r == nrounds else steps,
                              num_warmup=warmup // 2, num_leapfrog_steps=L, per_chain_adaptation=False, flow=flow,
                              seed=args.seed + r,
                              target_accept=args.flow_target_accept)
        flow_s = summarise(fres, f"flow round {r}")
        rounds.append(dict(flow_s, training_seconds=t_r))
    desc = (f"Neal's funnel D={dim}, batched_hmc L={L}, {chains} chains, plain vs affine-coupling flow ({layers} "
            f"layers, width {H}) trained on the plain run's {steps} draws/chain, {max(1, args.flow_rounds)} training "
            f"round(s)")
    return _funnel_report(args, D, model, flow, plain, plain_s, fres, flow_s, rounds, t_train, steps, warmup, chains,
                          L, H, final_steps, desc)


def _funnel_elbo_flow(args, ut, D, model, flow, plain, plain_s, summarise, steps, warmup, chains, dim, L, H, layers,
                      final_steps) -> int:
    """Flow fitted by the reparameterised ELBO against the funnel itself (the
    reference's flow ADVI objective, run on the device in stages of decreasing
    learning rate), then flow-preconditioned shared-adaptation HMC."""
    stages = [(int(a.split(":")[0]), float(a.split(":")[1])) for a in args.elbo_schedule.split(",")]
    t_train, rounds = 0.0, []
    for k, (n, lr) in enumerate(stages):
        t0 = time.perf_counter()
        out = flow.fit_elbo(model, num_particles=4096, num_iterations=n, learning_rate=lr, seed=args.seed + 100 + k)
        dt = time.perf_counter() - t0
        t_train += dt
        h = out["elbo_history"]
        log(f"bench: ELBO stage {k}: {n} iterations at lr {lr} in {dt:.2f}s, ELBO {h[:50].mean():.3f} -> "
            f"{h[-200:].mean():.3f} (log Z = 0)")
        rounds.append({"iterations": n, "learning_rate": lr, "seconds": dt, "elbo_first": float(h[:50].mean()),
                       "elbo_last": float(h[-200:].mean()), "best_elbo": out["best_elbo"]})
    fres = ut.batched_hmc(model, num_chains=chains, num_samples=final_steps, num_warmup=warmup // 2,
                          num_leapfrog_steps=L, per_chain_adaptation=False, flow=flow, seed=args.seed + 1,
                          target_accept=args.flow_target_accept)
    flow_s = summarise(fres, "flow (ELBO fit)")
    total = sum(n for n, _ in stages)
    desc = (f"Neal's funnel D={dim}, batched_hmc L={L}, {chains} chains, plain vs affine-coupling flow ({layers} "
            f"layers, width {H}) fitted by the reparameterised ELBO ({total} Adam iterations x 4096 particles)")
    return _funnel_report(args, D, model, flow, plain, plain_s, fres, flow_s, rounds, t_train, steps, warmup, chains,
                          L, H, final_steps, desc)


def _funnel_report(args, D, model, flow, plain, plain_s, fres, flow_s, rounds, t_train, steps, warmup, chains, L, H,
                   final_steps, desc) -> int:
    # conditioner GEMM flops per chain-gradient: forward 2 GEMMs + backward 2 GEMMs per layer
    flops_grad = 0.0
    for npass, nact, _ in flow.layout():
        flops_grad += 2 * 2.0 * (npass * H + H * 2 * nact)
    achieved = fres.sampling_gradient_evals * flops_grad / fres.sampling_kernel_seconds / 1e12
    traffic = None
    pmc = ROOT / "profiles" / "k_flow_cfg4_pmc.json"
    if pmc.exists():  # HBM bytes of one forward + one backward launch, from the committed FETCH_SIZE/WRITE_SIZE runs
        rec = json.loads(pmc.read_text())
        traffic = sum(rec[k]["traffic_bytes_per_launch"] for k in ("k_flow_fwd_all", "k_flow_bwd_all"))
    roofline = {"bound": "mfma", "compute": "bf16 MFMA conditioner GEMMs (latency-bound at 4096 rows)",
                "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic,
                "kernel": "k_flow_fwd_all + k_flow_bwd_all (one launch of each per batched gradient)",
                "timing": "HIP events around all sampling launches of the flow run",
                "flops_per_chain_gradient": flops_grad}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        cores = cpu_cores()
        n_it = max(1, min(200, int(args.cpu_seconds * 20)))
        t = time.perf_counter()
        om.run("hmc", "f64", chains, num_samples=n_it, num_warmup=0, num_leapfrog_steps=L,
               step_size=float(np.median(plain.step_size)), adapt_step_size=False, seed=args.seed, threads=cores)
        s_iter = (time.perf_counter() - t) / n_it
        cpu = {"value": (plain_s["min_bulk_ess"] / steps) / s_iter, "unit": "ESS/s (projected, plain HMC)",
               "cores": cores, "kind": "port",
               "sample": f"{n_it} binary64 oracle HMC iterations x {chains} chains (L={L}); projected = plain GPU "
                         f"bulk ESS per iteration / CPU s per iteration", "cpu_s_per_iteration": s_iter}
        log(f"bench: cpu {s_iter:.3f} s/iter -> projected {cpu['value']:.2f} ESS/s")
    line = {
        "metric": METRIC, "value": flow_s["value"], "unit": "ESS/s", "n_gpus": 1, "steps": final_steps,
        "warmup": warmup // 2, "ms_per_step": 1e3 * fres.sampling_seconds / final_steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic (no data; seed %d)" % args.seed,
        "config": {"workload": desc, "chains": chains, "flow_fit": args.flow_fit},
        "rhat_max": flow_s["rhat_max"], "rhat_ok": flow_s["rhat_ok"], "min_bulk_ess": flow_s["min_bulk_ess"],
        "flow": flow_s, "plain": plain_s, "flow_training_seconds": t_train, "flow_rounds": rounds,
        "flow_vs_plain": flow_s["value"] / plain_s["value"] if plain_s["value"] > 0 else None,
        "flow_moments_vs_plain": moments_vs(D, fres.constrained_samples[:, ::max(1, final_steps // steps)],
                                            plain.constrained_samples),
        # the funnel's exact law: v ~ N(0, 3) and u_i = x_i e^(-v/2) ~ N(0, 1) independent of v
        # (the raw x_i have E x_i^2 = e^(9/2) carried by rare draws: an unreliable MCSE)
        "flow_moments_vs_exact": moments_vs_exact(
            D, _funnel_standardised(fres.constrained_samples[:, ::max(1, final_steps // steps)]),
            np.zeros(model.num_params), np.concatenate([[9.0], np.ones(model.num_params - 1)])),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_advi(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import synthetic as S

    iters = 2000 if args.steps is None else args.steps
    warm = 100 if args.warmup is None else args.warmup
    n = 1024 if args.chains is None else args.chains
    y, off, sig, _ = S.hierarchical_data(64, 32, seed=args.seed)
    model = ut.HierarchicalNormalNC(y, off, sig, mu_loc=0.0, mu_scale=5.0, tau_scale=5.0)
    P = model.num_params
    kw = dict(num_particles=n, learning_rate=0.01, guide=args.guide, num_flow_layers=4,
              initial_params=np.zeros(P), initial_log_scale=-1.0, seed=args.seed)
    log(f"bench: batched_advi guide={args.guide} hierarchical J=64x32 particles={n} iterations={iters}")
    ut.batched_advi(model, num_iterations=max(warm, 1), **kw)
    t0 = time.perf_counter()
    r = ut.batched_advi(model, num_iterations=iters, **kw)
    wall = time.perf_counter() - t0
    h = r.elbo_history
    value = iters / wall
    log(f"bench: {wall:.3f}s -> {value:.1f} iterations/s; elbo {h[:50].mean():.2f} -> {h[-50:].mean():.2f}")
    # end-to-end figure: model-gradient flops of all particles over the wall time
    achieved = iters * n * FLOPS_PER_GRADIENT / wall / 1e12
    roofline = {"bound": "mfma", "compute": "fp32 vector ALU (model gradient); end to end, not one kernel",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                "kernel": "whole ADVI iteration (per-kernel times in profiles/advi_hier_kernel_trace.md)",
                "timing": "wall time of batched_advi"}
    cpu = None
    if not args.no_cpu_baseline and args.guide == "fullrank":
        from oracle import binding as B
        om = B.OracleModel(model)
        rng = np.random.default_rng(args.seed)
        mu, ls, F = np.zeros(P), np.full(P, -1.0), np.zeros((P, P))
        theta = [mu, ls, F]
        m1 = [np.zeros_like(a) for a in theta]
        m2 = [np.zeros_like(a) for a in theta]
        k, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and k < iters:
            ts = time.perf_counter()
            z = rng.standard_normal((n, P))
            Lc = np.tril(F, -1) + np.diag(np.exp(ls))
            x = mu + z @ Lc.T
            lj, g = om.logjoint_grad_f64(x.T)
            g = g.T
            ok = np.isfinite(lj) & np.all(np.isfinite(g), axis=1)
            gv, zv = g[ok], z[ok]
            grads = [gv.mean(axis=0), 1.0 + np.exp(ls) * (gv * zv).mean(axis=0), np.tril(gv.T @ zv, -1) / ok.sum()]
            for a, gr, ma, va in zip(theta, grads, m1, m2):
                ma *= 0.9
                ma += 0.1 * gr
                va *= 0.999
                va += 0.001 * gr * gr
                a += 0.01 * (ma / (1 - 0.9 ** (k + 1))) / (np.sqrt(va / (1 - 0.999 ** (k + 1))) + 1e-8)
            t_cpu += time.perf_counter() - ts
            k += 1
        cpu = {"value": k / t_cpu, "unit": "iterations/s", "cores": cpu_cores(), "kind": "port",
               "sample": f"{k} full-rank ADVI iterations x {n} particles in numpy around the oracle's binary64 "
                         f"gradient ({t_cpu:.1f} s)"}
        log(f"bench: cpu {cpu['value']:.2f} iterations/s")
    line = {
        "metric": "batched_advi iterations/s", "value": value, "unit": "iterations/s", "n_gpus": 1,
        "steps": iters, "warmup": warm, "ms_per_step": 1e3 * wall / iters, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": f"batched_advi guide={args.guide}, hierarchical normal non-centred J=64x32 "
                               f"(P=66), {n} particles, standard ELBO, lr 0.01",
                   "particles": n, "num_params": P, "guide": args.guide},
        "elbo_first50": float(h[:50].mean()), "elbo_last50": float(h[-50:].mean()), "best_elbo": r.best_elbo,
        "particle_gradients_per_s": iters * n / wall,
        "roofline": roofline, "cpu_baseline": cpu, "vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_gp_meads(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D
    from uncertaintea_amd import synthetic as S

    steps = 1000 if args.steps is None else args.steps
    warmup = 2000 if args.warmup is None else args.warmup
    chains = 1024 if args.chains is None else args.chains
    N = 512
    X, y, _ = S.gp_poisson_data(N, 2, seed=args.seed)
    model = ut.GaussianProcessPoisson(X, y, noise=1e-2)
    # start at the prior medians (l = 1, a = 0.674, z = 0) with N(0, 0.1^2) jitter: from prior
    # draws a few chains with large exp(f) dominate the folds' gradient Gram and the steps collapse
    rng = np.random.default_rng(args.seed)
    init = 0.1 * rng.normal(size=(model.num_params, chains))
    init[2] += np.log(0.674)

    def summarise(res, label):
        rh, eb = D.rhat_and_ess_bulk(res.constrained_samples)
        out = {"min_bulk_ess": float(np.nanmin(eb)), "rhat_max": float(np.nanmax(rh)),
               "ess_bulk_hyper": [float(v) for v in eb[:3]], "sampling_seconds": res.sampling_seconds,
               "warmup_seconds": res.warmup_seconds, "value": float(np.nanmin(eb)) / res.sampling_seconds,
               "gradient_evals_per_s": res.sampling_gradient_evals / res.sampling_seconds,
               "divergence_rate": res.divergence_rate(), "step_size": float(res.step_size[0]),
               "hyper_mean": [float(v) for v in res.constrained_samples[:, :, :3].mean(axis=(0, 1))]}
        out["rhat_ok"] = out["rhat_max"] < 1.01
        log(f"bench: {label}: min ESS {out['min_bulk_ess']:.1f} rhat {out['rhat_max']:.4f} -> {out['value']:.2f} "
            f"ESS/s; hyper means {out['hyper_mean']}")
        return out

    log(f"bench: GP Poisson N={N} batched_meads chains={chains} warmup={warmup} steps={steps}")
    plain = ut.batched_meads(model, num_chains=chains, num_samples=steps, num_warmup=warmup, seed=args.seed,
                             initial_params=init)
    plain_s = summarise(plain, "plain")
    flow = ut.AffineCouplingFlow(model.num_params, 6, 256, seed=args.seed)
    t0 = time.perf_counter()
    flow.fit(plain.unconstrained_samples.reshape(-1, model.num_params), steps=1000, batch=1024, seed=args.seed)
    t_train = time.perf_counter() - t0
    fres = ut.batched_meads(model, num_chains=chains, num_samples=steps, num_warmup=warmup, seed=args.seed + 1,
                            initial_params=init, flow=flow)
    flow_s = summarise(fres, "flow")
    rounds = [dict(flow_s, training_seconds=t_train)]
    nrounds = max(1, args.flow_rounds)
    # the reported (last) flow run samples final_steps per chain (default: steps); R-hat
    # there is set by the effective draws per chain, which MEADS's one leapfrog step per
    # iteration accumulates slowly for the hyperparameters
    final_steps = args.final_steps if args.final_steps else steps
    for r in range(2, nrounds + 1):
        t0 = time.perf_counter()
        flow.fit(fres.unconstrained_samples.reshape(-1, model.num_params), steps=1000, batch=1024, seed=args.seed + r,
                 init_base=False)
        t_r = time.perf_counter() - t0
        t_train += t_r
        fres = ut.batched_meads(model, num_chains=chains, num_samples=final_steps if r == nrounds else steps,
                                num_warmup=warmup, seed=args.seed + r, initial_params=init, flow=flow)
        log(f"bench: flow round {r} sampled {fres.num_samples} draws/chain in {fres.sampling_seconds:.1f}s "
            f"(warmup {fres.warmup_seconds:.1f}s); diagnostics")
        flow_s = summarise(fres, f"flow round {r}")
        rounds.append(dict(flow_s, training_seconds=t_r))
    flops = 7.0 * N ** 3 / 3.0 + 8.0 * N * N  # factorisation with three tangents + the four triangular products
    achieved = plain.sampling_gradient_evals * flops / plain.sampling_kernel_seconds / 1e12
    roofline = {"bound": "mfma", "compute": "binary64 batched Cholesky with tangents (vector FP64)",
                "achieved": achieved, "peak": 78.6, "unit": "TFLOP/s", "frac": achieved / 78.6, "traffic": None,
                "kernel": "k_gp_column / k_gp_panel", "flops_per_chain_gradient": flops,
                "timing": "HIP events around all sampling launches of the plain run (MEADS kernels included)"}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        sub = 16
        params = plain.unconstrained_samples[:sub, -1, :].astype(np.float64).T.copy()
        n, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n < 20:
            ts = time.perf_counter()
            om.gp_eval(params)
            t_cpu += time.perf_counter() - ts
            n += 1
        s_iter = t_cpu / n * chains / sub  # MEADS: one gradient per chain per iteration
        cores = min(int(os.environ.get("OMP_NUM_THREADS") or cpu_cores()), os.cpu_count() or 1)
        cpu = {"value": (plain_s["min_bulk_ess"] / steps) / s_iter, "unit": "ESS/s (projected)", "cores": cores,
               "kind": "port", "sample": f"{n} binary64 oracle GP gradients of {sub} chains ({t_cpu:.1f} s), scaled "
                                          f"to {chains} chains (one gradient per MEADS iteration)",
               "cpu_s_per_iteration": s_iter}
        log(f"bench: cpu {s_iter:.1f} s/iter -> projected {cpu['value']:.4f} ESS/s")
    line = {
        "metric": METRIC, "value": flow_s["value"], "unit": "ESS/s", "n_gpus": 1, "steps": fres.num_samples,
        "warmup": warmup, "ms_per_step": 1e3 * fres.sampling_seconds / fres.num_samples, "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (factorisation) + f32", "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": f"latent GP Poisson N={N} (2-D inputs, ARD SE kernel), batched_meads {chains} chains, "
                               "without / with an affine-coupling flow", "chains": chains,
                   "init": "prior medians + N(0, 0.1^2) jitter"},
        "rhat_max": flow_s["rhat_max"], "rhat_ok": flow_s["rhat_ok"], "min_bulk_ess": flow_s["min_bulk_ess"],
        "flow": flow_s, "plain": plain_s, "flow_training_seconds": t_train, "flow_rounds": rounds,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_glm_chees(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D
    from uncertaintea_amd import synthetic as S

    steps = 500 if args.steps is None else args.steps
    warmup = 500 if args.warmup is None else args.warmup
    chains = 4096 if args.chains is None else args.chains
    N, Dm = 131072, 256
    x, y, _ = S.logistic_data(N, Dm, seed=args.seed, cfg=3, x_scale=1.0 / 16.0)
    model = ut.LogisticRegressionBF16(x, y)
    log(f"bench: batched_chees logistic N={N} D={Dm} bf16 chains={chains} warmup={warmup} steps={steps}")
    res = ut.batched_chees(model, num_chains=chains, num_samples=steps, num_warmup=warmup, seed=args.seed,
                           forward_velocity=args.forward_velocity)
    log(f"bench: sampler done warmup {res.warmup_seconds:.3f}s sampling {res.sampling_seconds:.3f}s "
        f"T {res.trajectory_trace[-1] if warmup else 0:.4f} eps {res.step_size[0]:.5f}")
    rh, eb = D.rhat_and_ess_bulk(res.constrained_samples)
    min_ess, rhat_max = float(np.nanmin(eb)), float(np.nanmax(rh))
    value = min_ess / res.sampling_seconds
    log(f"bench: rhat max {rhat_max:.4f} min bulk ESS {min_ess:.1f} -> {value:.1f} ESS/s")
    evals = res.sampling_gradient_evals
    flops = evals * 4.0 * N * Dm
    ksec = res.sampling_kernel_seconds
    achieved = flops / ksec / 1e12 if ksec > 0 else 0.0
    mean_steps = float(res.integration_steps.mean())
    roofline = {"bound": "mfma", "compute": "bf16 MFMA, fp32 accumulation (hi+lo operand pairs: 2x issued)",
                "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None, "kernel": "k_glm_grad",
                "timing": "HIP events around all sampling launches (gradient + trajectory kernels)",
                "flops_per_batch_gradient": 4.0 * N * Dm * chains}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        # the oracle gradient runs an OpenMP loop over chains with the default team
        cores = min(int(os.environ.get("OMP_NUM_THREADS") or cpu_cores()), os.cpu_count() or 1)
        sub = 64
        params = res.unconstrained_samples[:sub, -1, :].astype(np.float64).T.copy()
        n, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n < 50:
            ts = time.perf_counter()
            om.logjoint_grad_f64(params)
            t_cpu += time.perf_counter() - ts
            n += 1
        s_batch = t_cpu / n * chains / sub
        s_iter = s_batch * mean_steps
        cpu = {"value": (min_ess / steps) / s_iter, "unit": "ESS/s (projected)", "cores": cores, "kind": "port",
               "sample": f"{n} binary64 oracle gradients of {sub} chains ({t_cpu:.1f} s), scaled to {chains} "
                         f"chains x {mean_steps:.2f} leapfrog steps per iteration; projected = GPU bulk ESS per "
                         f"iteration / CPU s per iteration",
               "cpu_s_per_batch_gradient": s_batch}
        log(f"bench: cpu {s_batch:.2f} s per batch gradient -> projected {cpu['value']:.3f} ESS/s")
    line = {
        "metric": METRIC, "value": value, "unit": "ESS/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * res.sampling_seconds / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": "logistic regression N=131072 D=256 bf16 X, batched_chees lock-step, "
                               f"{chains} chains, fused MFMA gradient",
                   "chains": chains, "forward_velocity": bool(args.forward_velocity)},
        "rhat_max": rhat_max, "rhat_ok": rhat_max < 1.01, "min_bulk_ess": min_ess,
        "sampling_seconds": res.sampling_seconds, "warmup_seconds": res.warmup_seconds,
        "gradient_evals_per_s": evals / res.sampling_seconds, "mean_leapfrog_steps": mean_steps,
        "trajectory_time": float(res.trajectory_trace[-1]) if warmup else None,
        "step_size": float(res.step_size[0]), "divergence_rate": res.divergence_rate(),
        "roofline": roofline, "cpu_baseline": cpu, "vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    if args.flow:
        flow = ut.AffineCouplingFlow(model.num_params, 6, 256, seed=args.seed)
        t0 = time.perf_counter()
        trace = flow.fit(res.unconstrained_samples.reshape(-1, model.num_params), steps=2000, batch=4096,
                         seed=args.seed)
        t_train = time.perf_counter() - t0
        log(f"bench: flow trained in {t_train:.2f}s, loss {trace[:20].mean():.2f} -> {trace[-100:].mean():.2f}")
        fres = ut.batched_chees(model, num_chains=chains, num_samples=steps, num_warmup=warmup, seed=args.seed + 1,
                                forward_velocity=args.forward_velocity, flow=flow)
        frh, feb = D.rhat_and_ess_bulk(fres.constrained_samples)
        fmin = float(np.nanmin(feb))
        fval = fmin / fres.sampling_seconds
        mom = moments_vs(D, fres.constrained_samples, res.constrained_samples)
        log(f"bench: flow: rhat {float(np.nanmax(frh)):.4f} min ESS {fmin:.1f} -> {fval:.1f} ESS/s, "
            f"{float(fres.integration_steps.mean()):.2f} steps/iter; vs plain {mom}")
        line["flow"] = {"value": fval, "min_bulk_ess": fmin, "rhat_max": float(np.nanmax(frh)),
                        "sampling_seconds": fres.sampling_seconds, "warmup_seconds": fres.warmup_seconds,
                        "mean_leapfrog_steps": float(fres.integration_steps.mean()),
                        "gradient_evals_per_s": fres.sampling_gradient_evals / fres.sampling_seconds,
                        "step_size": float(fres.step_size[0]), "training_seconds": t_train,
                        "moments_vs_plain": mom}
        line["flow_vs_plain"] = fval / value if value > 0 else None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
