"""Benchmark: min bulk-ESS/s at R-hat < 1.01, plain vs flow-preconditioned HMC, one MI355X.

Default workload (`--workload funnel_flow`, the line the driver records): Neal's funnel
D = 100 (P = 101, v ~ normal(0, 3), x_i ~ normal(0, exp(v/2))), batched_hmc with 4096
chains and L = 10 leapfrog steps, plain and flow-preconditioned.

  plain  per-chain adaptation (the persistent kernel): 2 x W x 100 warmup iterations,
         then 2000 draws per chain (8.19 M draws in all).
  flow   a 6-layer, width-256 affine-coupling flow (fp32 weights, bf16 GEMM operands on
         the matrix cores), trained on the GPU by the project's own kernels, then
         shared-adaptation lock-step HMC in the flow's base space: W x 100 warmup
         iterations and K x 200 sampling iterations.  `--flow-fit elbo` (default):
         the reparameterised ELBO against the model (the reference's flow-ADVI
         objective), 50k Adam iterations of 4096 particles from the identity flow.
         `--flow-fit ml`: maximum likelihood on the plain run's 2000 draws per chain,
         re-trained on each flow run's draws (`--flow-rounds`); `ml+elbo`: that fit
         refined by the ELBO.  The plain draws miss the funnel's neck (v sd 2.81
         against 3), so a flow fitted to them inherits the gap: the ML fit stalls
         near R-hat 1.03 and, as a start for the ELBO, leaves it at 1.13 after 30k
         iterations, while the ELBO from the identity reaches 1.0075.

One step is one block of 200 sampling iterations of all 4096 chains and one warmup
unit 100 warmup iterations, so the driver's `--steps 20 --warmup 5` samples 4000 draws
per chain after 500 warmup iterations.  `value` is the flow run's min over parameters
of rank-normalised bulk ESS (constrained space) divided by its sampling wall time
(chain state resident in HBM; the time includes copying the draws back), and is null
unless the run passes the reference harness's gate (bench/crossppl/python/analyze.py:
51-97): every parameter's rank R-hat < 1.01 and its mean and 5 % / 95 % quantiles within
4 MCSE of the reference -- here the funnel's exact law, so z is against the run's own
MCSE -- with the harness's ArviZ 0.23.4 estimators (bulk / tail ESS, quantile MCSE);
`ungated_value` keeps the raw quotient.
The plain run is reported beside it under the same gate.

Also reported: gradient evaluations/s; the roofline of the dominant kernel pair (the
flow's all-layer forward and backward launches, bf16 MFMA) from HIP-event device time of
back-to-back launches at the sampling launch shape (4096 rows) after the flow run, beside
the sampling run's device time per batched gradient (flow pair + model gradient +
leapfrog); the HBM bytes per launch when a committed counter record of these kernels
exists; and the CPU baseline -- the oracle's binary64 plain
HMC on every host core in the affinity mask (and on one thread), timed on a bounded
sample of iterations, giving CPU s/iteration and a projected plain ESS/s.

`--workload hier_nuts`: the non-centred hierarchical normal (J = 64 groups x 32
observations), batched_nuts with 1024 chains, max tree depth 10, 1000 + 3000
iterations (here a step is one NUTS iteration); the CPU baseline runs the same
sampler in full (binary64 oracle NUTS, 1024 chains, 1000 + 3000, all cores) and is
measured, not projected; a 1-thread figure is projected from a bounded sample.

`--workload glm_chees`: the large logistic regression (N = 131072, D = 256, X in bf16,
4096 chains) under batched_chees with the fused matrix-core gradient (steps =
iterations); CPU baseline projected from binary64 oracle gradients (a full CPU run
is infeasible: 5.5e11 flops per batch gradient).

`--workload gp_meads`: the latent GP with Poisson counts (N = 512, P = 515),
batched_meads with 1024 chains, without and with a flow; CPU baseline projected.

`--workload advi`: batched_advi on the hierarchical model (P = 66, 1024 particles);
`value` is Adam iterations/s.

`--workload flow_kernels`: the funnel preconditioner's sampling-path kernel pair alone (4096
rows; `--steps` = back-to-back launches per direction), for profiling those kernels by
themselves.

`--workload logistic_nuts`: few chains over many rows -- logistic regression N = 65536,
D = 32 (fp32), batched_nuts with 64 chains, 500 + 500 iterations, each chain's rows split over
a workgroup of 8 wavefronts (the automatic layout); also times the one-wavefront layout on a
shorter run; CPU baseline projected from binary64 oracle NUTS iterations from the GPU run's
adapted state.
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
# the other workloads' own metric strings (the default line above is the headline)
METRIC_NUTS = "min bulk-ESS/sec at R̂<1.01, batched NUTS, 1 MI355X"
METRIC_CHEES = "min bulk-ESS/sec at R̂<1.01, batched ChEES-HMC (plain; flow-preconditioned with --flow), 1 MI355X"
METRIC_MEADS = "min bulk-ESS/sec at R̂<1.01, plain vs flow-preconditioned MEADS, 1 MI355X"
RHAT_MAX = 1.01
Z_MAX = 4.0  # analyze.py's moment gate: combined-MCSE z of mean, q05 and q95
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
# funnel workload: one step = STEP_ITERS sampling iterations, one warmup unit = WARMUP_ITERS
STEP_ITERS = 200
WARMUP_ITERS = 100
TRAIN_DRAWS = 2000
# Algorithmic work per NUTS leaf of the J=64 x 32 model: one chain-gradient
# (6 N + 10 J + 20 = 12,948 flops) plus kick/drift/U-turn/energy (12 P = 792 flops).
FLOPS_PER_GRADIENT = 6 * 2048 + 10 * 64 + 20
FLOPS_PER_LEAF_EXTRA = 12 * 66


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gated(value, rhat_max):
    """The metric counts a throughput only when every parameter has R-hat < 1.01."""
    return value if (value is not None and np.isfinite(rhat_max) and rhat_max < RHAT_MAX) else None


def cpu_info() -> dict:
    """Host cores in this process's affinity mask and the CPU model string."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"affinity_cores": n, "cpu_model": model, "nproc": os.cpu_count(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_cores() -> int:
    return max(1, cpu_info()["affinity_cores"])


def harness_rhat_ess(D, draws):
    """Rank R-hat and bulk ESS per parameter with the harness's ArviZ 0.23.4 estimators."""
    st = D.harness_stats(draws)
    return st["rhat"], st["ess_bulk"]


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


def funnel_exact_stats(dim: int, v_scale: float) -> dict:
    """Mean, 5 % and 95 % quantiles of (v, x_1..x_dim) under Neal's funnel: v ~ N(0, s^2) and
    x_i | v ~ N(0, e^v); the x_i quantile q_p solves E_v[Phi(q e^(-v/2))] = p (quadrature)."""
    from scipy import integrate, optimize, stats

    def cdf(q):
        f = lambda v: stats.norm.cdf(q * np.exp(-0.5 * v)) * stats.norm.pdf(v, 0.0, v_scale)  # noqa: E731
        return integrate.quad(f, -12 * v_scale, 12 * v_scale, limit=400, epsabs=1e-12)[0]
    qx = {p: optimize.brentq(lambda q: cdf(q) - p, -1e4, 1e4, xtol=1e-12) for p in (0.05, 0.95)}
    zq = stats.norm.ppf(0.95)
    return {"mean": np.zeros(dim + 1),
            "q05": np.concatenate([[-zq * v_scale], np.full(dim, qx[0.05])]),
            "q95": np.concatenate([[zq * v_scale], np.full(dim, qx[0.95])])}


def _funnel_standardised(draws):
    """(v, x_1 e^(-v/2), ...) per draw: exactly N(0, 9) x N(0, 1)^dim under the funnel."""
    a = np.asarray(draws, dtype=np.float64)
    out = a.copy()
    out[..., 1:] *= np.exp(-0.5 * a[..., :1])
    return out


def _cpu_hmc_seconds_per_iteration(om, chains, L, step, seconds, threads, seed, chunk=20):
    """Wall seconds per binary64 oracle HMC iteration of `chains` chains (fixed step,
    no adaptation), from chunks of iterations until `seconds` have been spent."""
    it, t = 0, 0.0
    while t < seconds or it == 0:
        ts = time.perf_counter()
        om.run("hmc", "f64", chains, num_samples=chunk, num_warmup=0, num_leapfrog_steps=L, step_size=step,
               adapt_step_size=False, seed=seed + it, threads=threads)
        t += time.perf_counter() - ts
        it += chunk
    return t / it, it, t


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("funnel_flow", "hier_nuts", "glm_chees", "gp_meads", "advi", "flow_kernels",
                                           "logistic_nuts"),
                    default="funnel_flow")
    ap.add_argument("--guide", choices=("meanfield", "fullrank", "flow"), default="fullrank",
                    help="advi: the variational guide")
    ap.add_argument("--steps", type=int, default=None,
                    help="funnel_flow: blocks of 200 sampling iterations (default 20); other workloads: "
                         "sampling iterations")
    ap.add_argument("--warmup", type=int, default=None,
                    help="funnel_flow: blocks of 100 warmup iterations (default 5); other workloads: warmup "
                         "iterations")
    ap.add_argument("--chains", type=int, default=None)
    ap.add_argument("--flow", action="store_true",
                    help="glm_chees: also train a flow on the plain run's draws and run ChEES with it")
    ap.add_argument("--forward-velocity", action="store_true",
                    help="glm_chees: ChEES criterion with the forward end velocity (not the reference's sign)")
    ap.add_argument("--flow-rounds", type=int, default=3,
                    help="funnel_flow --flow-fit ml / gp_meads: train the flow on the plain run's draws, then "
                         "re-train it on each flow run's draws this many times in all (iterative NeuTra)")
    ap.add_argument("--flow-target-accept", type=float, default=None,
                    help="funnel_flow: target acceptance of the flow run's step-size adaptation (default 0.97 "
                         "when the fit ends with the ELBO: with L fixed at 10 the smaller step keeps the "
                         "trajectory near a quarter period of the near-Gaussian base; 0.9 after maximum "
                         "likelihood alone)")
    ap.add_argument("--flow-fit", choices=("elbo", "ml", "ml+elbo"), default="elbo",
                    help="funnel_flow: the reparameterised ELBO against the model from the identity flow "
                         "(default); maximum likelihood on the plain run's draws with --flow-rounds re-training "
                         "rounds on the flow runs' draws; or that maximum-likelihood fit refined by the ELBO")
    ap.add_argument("--ml-steps", type=int, default=2000,
                    help="funnel_flow: Adam steps of the maximum-likelihood fit (4096-draw minibatches)")
    ap.add_argument("--elbo-schedule", default=None,
                    help="funnel_flow: ELBO stages iterations:learning_rate, 4096 particles each (default "
                         "30000:1e-3,10000:3e-4,10000:1e-4 from the identity flow, "
                         "20000:3e-4,10000:1e-4 after maximum likelihood)")
    ap.add_argument("--final-steps", type=int, default=None,
                    help="gp_meads: sampling iterations of the reported flow run (default --steps)")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="wall seconds of the bounded all-core CPU sample (the 1-thread sample gets a third)")
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
    return {"funnel_flow": bench_funnel_flow, "hier_nuts": bench_hier_nuts, "glm_chees": bench_glm_chees,
            "gp_meads": bench_gp_meads, "advi": bench_advi, "flow_kernels": bench_flow_kernels,
            "logistic_nuts": bench_logistic_nuts}[args.workload](args)


def bench_logistic_nuts(args) -> int:
    """Few chains over many rows: persistent NUTS with each chain's rows split over the
    wavefronts of its workgroup."""
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D
    from uncertaintea_amd import synthetic as S
    from uncertaintea_amd.device import DeviceBatchedWorkspace

    N, Dd = 65536, 32
    args.steps = 500 if args.steps is None else args.steps
    args.warmup = 500 if args.warmup is None else args.warmup
    args.chains = 64 if args.chains is None else args.chains
    x, y, _ = S.logistic_data(N, Dd, seed=args.seed)
    model = ut.LogisticRegression(x, y)
    waves = DeviceBatchedWorkspace(model, args.chains).waves_per_chain
    log(f"bench: batched_nuts logistic N={N} D={Dd} chains={args.chains} waves/chain={waves} "
        f"warmup={args.warmup} steps={args.steps}")
    t0 = time.perf_counter()
    res = ut.batched_nuts(model, num_chains=args.chains, num_samples=args.steps, num_warmup=args.warmup,
                          max_tree_depth=10, seed=args.seed)
    wall = time.perf_counter() - t0
    log(f"bench: sampler done warmup {res.warmup_seconds:.3f}s sampling {res.sampling_seconds:.3f}s wall {wall:.3f}s")
    st = D.harness_stats(res.constrained_samples)
    min_ess = float(np.nanmin(st["ess_bulk"]))
    rhat_max = float(np.nanmax(st["rhat"]))
    raw = min_ess / res.sampling_seconds
    log(f"bench: rhat max {rhat_max:.4f} min bulk ESS {min_ess:.1f} -> {raw:.1f} ESS/s")
    evals = res.sampling_gradient_evals
    # per chain-gradient: 4 N D flops (X beta and X^T r) + ~6 N + 3 D (the logistic-regression
    # unit of work); algorithmic bytes 4 N D + 4 N (X and y, fp32, from L2 after the first chain)
    flops_per_grad = 4 * N * Dd + 6 * N + 3 * Dd
    ksec = res.sampling_kernel_seconds
    launches = max(res.sampling_launches, 1)
    achieved = evals * flops_per_grad / ksec / 1e12 if ksec > 0 else 0.0
    roofline = {"bound": "fp32-valu", "compute": "fp32 vector ALU and L2 reads of X (latency-bound: "
                f"{args.chains} chains x {waves} wavefronts on 1024 SIMDs)",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                "traffic": None, "kernel": f"k_nuts<TiledLogisticModel<1, 32, {waves}>>" if waves > 1 else "k_nuts<LogisticModel<1, 32>>", "launches": res.sampling_launches,
                "avg_launch_ms": 1e3 * ksec / launches, "flops_per_chain_gradient": flops_per_grad,
                "l2_bytes_per_chain_gradient": 4 * N * Dd + 4 * N,
                "timing": "HIP events around every sampling launch on the workspace stream"}
    # the one-wavefront layout on a shorter run, for the speedup of the workgroup layout
    short = min(100, args.steps)
    r1 = ut.batched_nuts(model, num_chains=args.chains, num_samples=short, num_warmup=min(200, args.warmup),
                         max_tree_depth=10, seed=args.seed, waves_per_chain=1)
    one = {"waves_per_chain": 1, "sampling_seconds": r1.sampling_seconds, "num_samples": short,
           "gradient_evals_per_s": r1.sampling_gradient_evals / r1.sampling_seconds,
           "ms_per_gradient_eval_all_chains": 1e3 * r1.sampling_seconds * args.chains / max(1, r1.sampling_gradient_evals)}
    mine = 1e3 * res.sampling_seconds * args.chains / max(1, evals)
    log(f"bench: per leapfrog step of all chains {mine:.3f} ms with {waves} wavefronts per chain, "
        f"{one['ms_per_gradient_eval_all_chains']:.3f} ms with one")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        info = cpu_info()
        cores = max(1, info["affinity_cores"])
        pos = res.unconstrained_samples[:, -1, :].astype(np.float64).T.copy()
        it = args.warmup + args.steps + 1
        n, t, ev = 0, 0.0, 0
        while t < args.cpu_seconds and n < 500:
            ts = time.perf_counter()
            e, pos = B.nuts_sample_from(om, "f64", pos, res.step_size, res.mass_matrix, it + n, 1, seed=args.seed,
                                        threads=cores)
            t += time.perf_counter() - ts
            ev += e
            n += 1
        s_it = t / n
        cpu = {"value": (min_ess / args.steps) / s_it, "unit": "ESS/s (projected, binary64 oracle NUTS)",
               "cores": cores, "kind": "port", "cpu_model": info["cpu_model"], "host": info,
               "sample": f"{n} sampling iterations x {args.chains} chains of binary64 oracle NUTS on {cores} threads "
                         f"({t:.1f} s) from the GPU run's adapted state; projected = GPU run's bulk ESS per "
                         f"iteration / CPU s per iteration",
               "cpu_s_per_iteration": s_it, "cpu_gradient_evals_per_s": ev / t}
        log(f"bench: cpu {cores} threads {s_it:.3f} s/iter -> projected {cpu['value']:.1f} ESS/s")

    ok = rhat_max < RHAT_MAX
    line = {
        "metric": METRIC_NUTS, "value": raw if ok else None, "unit": "ESS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * res.sampling_seconds / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": f"logistic regression N={N} D={Dd} fp32, batched_nuts persistent, {args.chains} chains, "
                               f"max depth 10, {waves} wavefronts per chain (plain NUTS, no flow)",
                   "chains": args.chains, "num_params": model.num_params, "waves_per_chain": waves},
        "ungated_value": raw, "rhat_max": rhat_max, "rhat_ok": ok, "min_bulk_ess": min_ess,
        "sampling_seconds": res.sampling_seconds, "warmup_seconds": res.warmup_seconds,
        "gradient_evals_per_s": evals / res.sampling_seconds, "mean_tree_depth": float(res.tree_depths.mean()),
        "ms_per_gradient_eval_all_chains": mine, "one_wavefront_per_chain": one,
        "roofline": roofline, "cpu_baseline": cpu,
        "vs_cpu": (raw / cpu["value"]) if cpu and cpu["value"] else None,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_hier_nuts(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D
    from uncertaintea_amd import synthetic as S

    args.steps = 3000 if args.steps is None else args.steps
    args.warmup = 1000 if args.warmup is None else args.warmup
    args.chains = 1024 if args.chains is None else args.chains
    y, off, sig, _ = S.hierarchical_data(64, 32, seed=args.seed)
    model = ut.HierarchicalNormalNC(y, off, sig, mu_loc=0.0, mu_scale=5.0, tau_scale=5.0)
    log(f"bench: batched_nuts hierarchical J=64x32 chains={args.chains} warmup={args.warmup} steps={args.steps}")
    t0 = time.perf_counter()
    res = ut.batched_nuts(model, num_chains=args.chains, num_samples=args.steps, num_warmup=args.warmup,
                          max_tree_depth=10, seed=args.seed)
    wall = time.perf_counter() - t0
    log(f"bench: sampler done warmup {res.warmup_seconds:.3f}s sampling {res.sampling_seconds:.3f}s wall {wall:.3f}s")
    st = D.harness_stats(res.constrained_samples)
    min_ess = float(np.nanmin(st["ess_bulk"]))
    rhat_max = float(np.nanmax(st["rhat"]))
    raw = min_ess / res.sampling_seconds
    log(f"bench: rhat max {rhat_max:.4f} min bulk ESS {min_ess:.1f} -> {raw:.1f} ESS/s")

    evals = res.sampling_gradient_evals
    flops = evals * (FLOPS_PER_GRADIENT + FLOPS_PER_LEAF_EXTRA)
    ksec = res.sampling_kernel_seconds
    launches = max(res.sampling_launches, 1)
    achieved = flops / ksec / 1e12 if ksec > 0 else 0.0
    # HBM bytes of a launch: the tree runs on chip, so a launch's traffic is its draws
    # and per-draw statistics; the committed FETCH_SIZE/WRITE_SIZE runs give bytes per
    # sampling iteration, scaled here to the iterations of the launches timed
    traffic = None
    pmc = ROOT / "profiles" / "k_nuts_hier_cfg2_pmc.json"
    if pmc.exists():
        rec = json.loads(pmc.read_text())
        per_it = rec.get("traffic_bytes_per_launch", 0.0) / max(1, rec.get("iterations_per_launch", 500))
        traffic = per_it * args.steps / launches
    roofline = {"bound": "fp32-valu", "compute": "fp32 vector ALU (the tree and the model gradient stay on chip)",
                "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                "kernel": "k_nuts<HierModel>", "launches": res.sampling_launches,
                "avg_launch_ms": 1e3 * ksec / launches, "flops_per_launch": flops / launches,
                "timing": "HIP events around every sampling launch on the workspace stream"}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        info = cpu_info()
        cores = max(1, info["affinity_cores"])
        log(f"bench: cpu full run: binary64 oracle NUTS, {args.chains} chains, {args.warmup} + {args.steps}, "
            f"{cores} threads")
        t0 = time.perf_counter()
        orc = om.run("nuts", "f64", args.chains, num_samples=args.steps, num_warmup=args.warmup, max_tree_depth=10,
                     seed=args.seed, threads=cores)
        t_full = time.perf_counter() - t0
        cst = D.harness_stats(model.constrain(orc["samples"]))
        c_ess, c_rhat = float(np.nanmin(cst["ess_bulk"])), float(np.nanmax(cst["rhat"]))
        c_raw = c_ess / orc["sampling_seconds"]
        # one thread: a bounded sample of sampling iterations from the GPU run's adapted state
        pos = res.unconstrained_samples[:, -1, :].astype(np.float64).T.copy()
        it = args.warmup + args.steps + 1
        n1, t1 = 0, 0.0
        while t1 < args.cpu_seconds / 3 and n1 < 2000:
            ts = time.perf_counter()
            B.nuts_sample_from(om, "f64", pos, res.step_size, res.mass_matrix, it + n1, 1, seed=args.seed, threads=1)
            t1 += time.perf_counter() - ts
            n1 += 1
        s1 = t1 / n1
        cpu = {"value": gated(c_raw, c_rhat), "ungated_value": c_raw, "unit": "ESS/s (measured, full run)",
               "cores": cores, "kind": "port", "cpu_model": info["cpu_model"], "host": info,
               "sample": f"the whole sampler run on the host: binary64 oracle NUTS, {args.chains} chains, "
                         f"{args.warmup} warmup + {args.steps} sampling iterations ({t_full:.1f} s wall)",
               "warmup_seconds": orc["warmup_seconds"], "sampling_seconds": orc["sampling_seconds"],
               "rhat_max": c_rhat, "min_bulk_ess": c_ess,
               "one_thread": {"s_per_iteration": s1, "projected_ess_per_s": (c_ess / args.steps) / s1,
                              "sample": f"{n1} sampling iterations x {args.chains} chains on 1 thread "
                                        f"({t1:.1f} s); projected = CPU run's bulk ESS per iteration / s per "
                                        f"iteration"}}
        log(f"bench: cpu {cores} threads: warmup {orc['warmup_seconds']:.1f}s sampling {orc['sampling_seconds']:.1f}s "
            f"rhat {c_rhat:.4f} -> {c_raw:.1f} ESS/s; 1 thread {s1:.3f} s/iter")

    # the harness gate: R-hat, and mean / q05 / q95 within 4 combined MCSE of the reference
    # path's run (the CPU oracle's full run) when it was made
    ok, fails, zmax = D.harness_gate(st, cst if cpu is not None else None, RHAT_MAX, Z_MAX)
    value = raw if ok else None
    log(f"bench: gate {'passed' if ok else 'failed: ' + '; '.join(fails[:3])} (max z vs the CPU run {zmax:.2f})")
    gate = {"passed": ok, "max_z_vs_cpu_run": zmax if cpu is not None else None, "failures": fails[:8],
            "rule": f"rank R-hat < {RHAT_MAX}" + (f"; mean, q05, q95 within {Z_MAX} combined MCSE of the CPU "
                                                 f"reference run" if cpu is not None else "")
                    + " (analyze.py:76-97; ArviZ 0.23.4 estimators)"}
    line = {
        "metric": METRIC_NUTS, "value": value, "unit": "ESS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * res.sampling_seconds / args.steps, "gate": gate,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": "hierarchical normal non-centred J=64x32, batched_nuts persistent, 1024 chains, "
                               "max depth 10, fp32 state + fp64 energies (plain NUTS, no flow)",
                   "chains": args.chains, "num_params": model.num_params, "max_tree_depth": 10},
        "ungated_value": raw, "rhat_max": rhat_max, "rhat_ok": rhat_max < RHAT_MAX, "min_bulk_ess": min_ess,
        "sampling_seconds": res.sampling_seconds, "warmup_seconds": res.warmup_seconds,
        "gradient_evals_per_s": evals / res.sampling_seconds,
        "mean_tree_depth": float(res.tree_depths.mean()), "divergence_rate": res.divergence_rate(),
        "roofline": roofline, "cpu_baseline": cpu,
        "vs_cpu": (value / cpu["value"]) if (cpu and value is not None and cpu["value"]) else None,
    }
    print(json.dumps(line), flush=True)
    return 0


FLOW16_PMC = "k_flow16_cfg4_pmc.json"
FLOW16_GRAD_PMC = "k_flow16_grad_cfg4_pmc.json"


def flow_roofline(flow, rows, reps=200) -> dict:
    """Roofline of the flow's sampling-path kernel pair (one forward and one backward launch per
    batched gradient of `rows` rows): algorithmic conditioner-GEMM flops over the kernels' own
    device time (HIP events around `reps` back-to-back launches of each direction on the flow's
    stream), and their HBM bytes when a committed FETCH_SIZE / WRITE_SIZE record exists."""
    H = flow.hidden
    # conditioner GEMM flops per chain-gradient: forward 2 GEMMs + backward 2 GEMMs per layer
    flops_grad = 0.0
    for npass, nact, _ in flow.layout():
        flops_grad += 2 * 2.0 * (npass * H + H * 2 * nact)
    fwd_us, bwd_us = flow.kernel_times(rows, reps=reps)
    achieved = rows * flops_grad / ((fwd_us + bwd_us) * 1e-6) / 1e12
    traffic, traffic_note = None, None
    pmc = ROOT / "profiles" / FLOW16_PMC
    if pmc.exists():  # HBM bytes of one forward + one backward launch at 4096 rows
        rec = json.loads(pmc.read_text())
        traffic = rec["traffic_bytes_per_launch_pair"]
        traffic_note = rec.get("note")
    else:
        traffic_note = "no FETCH_SIZE / WRITE_SIZE record for these kernels"
    return {"bound": "mfma", "compute": f"bf16 MFMA conditioner GEMMs, fp32 accumulation (latency-bound at {rows} rows)",
            "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic, "traffic_note": traffic_note,
            "kernel": "k_flow16_fwd + k_flow16_bwd (one launch of each per batched gradient of all chains)",
            "timing": f"HIP events around {reps} back-to-back launches of each direction on the flow's stream, "
                      f"{rows} device-resident rows",
            "forward_us": fwd_us, "backward_us": bwd_us,
            "flops_per_launch_pair": rows * flops_grad, "flops_per_chain_gradient": flops_grad}


def bench_flow_kernels(args) -> int:
    """Micro-workload: the funnel preconditioner's sampling-path kernel pair alone (P = 101, 6
    coupling layers of width 256, 4096 rows of N(0, I) draws, weights at a small random
    perturbation of the identity flow), for profiling the dominant kernels by themselves."""
    import uncertaintea_amd as ut
    rows = 4096 if args.chains is None else args.chains
    flow = ut.AffineCouplingFlow(101, 6, 256, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    w = flow.weights.copy()
    w += (rng.normal(size=w.size) * 0.02).astype(np.float32)
    flow.weights = w
    reps = 200 if args.steps is None else args.steps
    roof = flow_roofline(flow, rows, reps=reps)
    # the fused gradient-and-step kernel on the funnel with this flow (step size 0)
    from uncertaintea_amd import _lib as UL
    from uncertaintea_amd.device import DeviceBatchedWorkspace
    model = ut.NealsFunnel(100, 3.0)
    ws = DeviceBatchedWorkspace(model, rows)
    UL.check(UL.lib().ut_workspace_set_flow(ws.handle, flow.handle))
    roof["fused_step_us"] = ws.gradient_step_kernel_time(np.zeros((model.num_params, rows), np.float32), reps=reps)
    roof["fused_step_kernel"] = "k_flow16_grad (forward, funnel gradient, backward, leapfrog step)"
    del ws
    line = {"metric": "flow kernel pair throughput (rows/s through forward + backward)",
            "value": rows / ((roof["forward_us"] + roof["backward_us"]) * 1e-6), "unit": "rows/s", "n_gpus": 1,
            "steps": reps, "warmup": 1, "ms_per_step": 1e-3 * (roof["forward_us"] + roof["backward_us"]),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (MFMA operands)",
            "data": "synthetic (N(0, I) rows)",
            "config": {"workload": "flow kernel pair, P=101, 6 coupling layers, H=256", "rows": rows},
            "roofline": roof, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def bench_funnel_flow(args) -> int:
    import uncertaintea_amd as ut
    from uncertaintea_amd import diagnostics as D

    steps_blocks = 20 if args.steps is None else args.steps
    warm_blocks = 5 if args.warmup is None else args.warmup
    S_flow = STEP_ITERS * steps_blocks
    W_flow = WARMUP_ITERS * warm_blocks
    W_plain = 2 * W_flow
    chains = 4096 if args.chains is None else args.chains
    dim, L, H, layers = 100, 10, 256, 6
    model = ut.NealsFunnel(dim, 3.0)

    exact = funnel_exact_stats(dim, 3.0)

    def summarise(res, label):
        draws = res.constrained_samples
        st = D.harness_stats(draws)
        ok, fails, zmax = D.harness_gate(st, exact, RHAT_MAX, Z_MAX)
        eb = st["ess_bulk"]
        raw = float(np.nanmin(eb)) / res.sampling_seconds
        rmax = float(np.nanmax(st["rhat"]))
        _, eb_jl = D.rhat_and_ess_bulk(draws)
        out = {"min_bulk_ess": float(np.nanmin(eb)), "min_tail_ess": float(np.nanmin(st["ess_tail"])),
               "rhat_max": rmax, "rhat_ok": rmax < RHAT_MAX,
               "gate": {"passed": ok, "max_z_vs_exact": zmax, "failures": fails[:8], "num_failures": len(fails),
                        "rule": f"rank R-hat < {RHAT_MAX} and mean, q05, q95 of every parameter within {Z_MAX} "
                                f"MCSE of the funnel's exact law (analyze.py:76-97; ArviZ 0.23.4 estimators)"},
               "min_bulk_ess_diagnostics_jl": float(np.nanmin(eb_jl)),
               "ess_bulk_v": float(eb[0]), "sampling_seconds": res.sampling_seconds,
               "warmup_seconds": res.warmup_seconds, "value": raw if ok else None, "ungated_value": raw,
               "num_samples": res.num_samples,
               "gradient_evals_per_s": res.sampling_gradient_evals / res.sampling_seconds,
               "divergence_rate": res.divergence_rate(), "step_size": float(np.median(res.step_size)),
               "v_mean": float(res.unconstrained_samples[:, :, 0].mean()),
               "v_sd": float(res.unconstrained_samples[:, :, 0].std())}
        log(f"bench: {label}: min bulk ESS {out['min_bulk_ess']:.1f} rhat {rmax:.4f} max z {zmax:.2f} -> "
            f"{raw:.1f} ESS/s ({'gated in' if ok else 'fails the gate: ' + '; '.join(fails[:3])}); v mean "
            f"{out['v_mean']:.3f} sd {out['v_sd']:.3f}; div {out['divergence_rate']:.4f}")
        return out

    log(f"bench: funnel D={dim} plain batched_hmc chains={chains} warmup={W_plain} draws={TRAIN_DRAWS}")
    plain = ut.batched_hmc(model, num_chains=chains, num_samples=TRAIN_DRAWS, num_warmup=W_plain,
                           num_leapfrog_steps=L, seed=args.seed)
    plain_s = summarise(plain, "plain")
    flow = ut.AffineCouplingFlow(model.num_params, layers, H, seed=args.seed)
    fit = args.flow_fit
    if args.flow_target_accept is None:
        args.flow_target_accept = 0.9 if fit == "ml" else 0.97
    if args.elbo_schedule is None:
        args.elbo_schedule = "20000:3e-4,10000:1e-4" if fit == "ml+elbo" else "30000:1e-3,10000:3e-4,10000:1e-4"
    stages = [(int(a.split(":")[0]), float(a.split(":")[1])) for a in args.elbo_schedule.split(",") if a]
    fit_log, t_train = [], 0.0
    if fit in ("ml", "ml+elbo"):
        draws = plain.unconstrained_samples.reshape(-1, model.num_params)
        t0 = time.perf_counter()
        trace = flow.fit(draws, steps=args.ml_steps, batch=4096, seed=args.seed)
        dt = time.perf_counter() - t0
        t_train += dt
        del draws
        log(f"bench: flow fitted by maximum likelihood on the plain run's {chains} x {TRAIN_DRAWS} draws, "
            f"{args.ml_steps} steps in {dt:.2f}s, loss {trace[:20].mean():.3f} -> {trace[-100:].mean():.3f}")
        fit_log.append({"objective": "maximum likelihood on the plain draws", "steps": args.ml_steps,
                        "seconds": dt, "loss_first": float(trace[:20].mean()), "loss_last": float(trace[-100:].mean())})
    if fit in ("elbo", "ml+elbo"):
        for k, (n, lr) in enumerate(stages):
            t0 = time.perf_counter()
            out = flow.fit_elbo(model, num_particles=4096, num_iterations=n, learning_rate=lr, seed=args.seed + 100 + k)
            dt = time.perf_counter() - t0
            t_train += dt
            h = out["elbo_history"]
            log(f"bench: ELBO stage {k}: {n} iterations at lr {lr} in {dt:.2f}s, ELBO {h[:50].mean():.4f} -> "
                f"{h[-200:].mean():.4f} (log Z = 0)")
            fit_log.append({"objective": "reparameterised ELBO", "iterations": n, "learning_rate": lr, "seconds": dt,
                            "elbo_first": float(h[:50].mean()), "elbo_last": float(h[-200:].mean()),
                            "best_elbo": out["best_elbo"]})
    rounds = max(1, args.flow_rounds) if fit == "ml" else 1
    fres = None
    for r in range(1, rounds + 1):
        if r > 1:
            # iterative NeuTra: re-train (warm start) on the previous flow run's draws
            t0 = time.perf_counter()
            trace = flow.fit(fres.unconstrained_samples.reshape(-1, model.num_params), steps=2 * args.ml_steps,
                             batch=4096, seed=args.seed + r, init_base=False)
            dt = time.perf_counter() - t0
            t_train += dt
            fit_log.append({"objective": f"maximum likelihood on flow round {r - 1}'s draws",
                            "steps": 2 * args.ml_steps, "seconds": dt, "loss_last": float(trace[-100:].mean())})
        fres = ut.batched_hmc(model, num_chains=chains, num_samples=S_flow, num_warmup=W_flow, num_leapfrog_steps=L,
                              per_chain_adaptation=False, flow=flow, seed=args.seed + r,
                              target_accept=args.flow_target_accept)
    flow_s = summarise(fres, "flow")

    roofline = flow_roofline(flow, chains)
    ksec = fres.sampling_kernel_seconds
    evals = fres.sampling_gradient_evals
    # the sampling path's dominant kernel: the flow-wrapped funnel gradient and leapfrog step
    # in one launch per 16-row tile (k_flow16_grad), timed on its own with HIP events
    from uncertaintea_amd import _lib as UL
    from uncertaintea_amd.device import DeviceBatchedWorkspace
    ws = DeviceBatchedWorkspace(model, chains)
    UL.check(UL.lib().ut_workspace_set_flow(ws.handle, flow.handle))
    step_us = ws.gradient_step_kernel_time(np.zeros((model.num_params, chains), np.float32), reps=200)
    del ws
    achieved = chains * roofline["flops_per_chain_gradient"] / (step_us * 1e-6) / 1e12
    grad_pmc = ROOT / "profiles" / FLOW16_GRAD_PMC
    grad_rec = json.loads(grad_pmc.read_text()) if grad_pmc.exists() else None
    separate_traffic = roofline.get("traffic")
    roofline.update({
        "achieved": achieved, "frac": achieved / PEAK_BF16_TFLOPS,
        "kernel": "k_flow16_grad (forward, the funnel's gradient, backward and the leapfrog step of 16 rows per "
                  "workgroup, one launch per leapfrog step after a trajectory's first)",
        "timing": f"HIP events around 200 back-to-back leapfrog steps of {chains} device-resident rows on the "
                  f"workspace stream (ut_gradient_step_kernel_time, step size 0)",
        "us_per_launch": step_us,
        "flops_per_launch": chains * roofline["flops_per_chain_gradient"],
        "separate_kernels": {"forward_us": roofline.pop("forward_us"), "backward_us": roofline.pop("backward_us"),
                             "traffic": separate_traffic,
                             "note": "k_flow16_fwd + k_flow16_bwd alone (the path of a trajectory's last step)"},
        "traffic": grad_rec["traffic_bytes_per_launch"] if grad_rec else None,
        "traffic_note": grad_rec.get("note") if grad_rec else "no FETCH_SIZE / WRITE_SIZE record for the fused kernel",
        "us_per_batched_gradient_in_sampling": 1e6 * ksec / max(1, evals // chains),
        "per_gradient_note": "the sampling run's device time per leapfrog step of all chains (HIP events around the "
                             "timed region on the workspace stream): the fused kernel, and per trajectory the "
                             "separate forward, model and backward kernels of the last step and the lock-step "
                             "begin / end / record kernels"})
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        info = cpu_info()
        cores = max(1, info["affinity_cores"])
        step = float(np.median(plain.step_size))
        s_iter, n_it, t_all = _cpu_hmc_seconds_per_iteration(om, chains, L, step, args.cpu_seconds, cores, args.seed)
        s_iter1, n_it1, t_1 = _cpu_hmc_seconds_per_iteration(om, chains // 16, L, step, args.cpu_seconds / 3, 1,
                                                             args.seed, chunk=10)
        s_iter1 *= 16  # scaled from chains/16 chains to all chains
        ess_it = plain_s["min_bulk_ess"] / TRAIN_DRAWS
        cpu = {"value": ess_it / s_iter, "unit": "ESS/s (projected, plain HMC, not R-hat gated)", "cores": cores,
               "kind": "port", "cpu_model": info["cpu_model"], "host": info,
               "sample": f"{n_it} binary64 oracle plain-HMC iterations x {chains} chains (L={L}, the plain run's "
                         f"median step) on {cores} threads ({t_all:.1f} s); projected = the plain GPU run's bulk ESS "
                         f"per iteration / CPU s per iteration (the reference CPU path is plain HMC; like the plain "
                         f"GPU run it misses the funnel's neck, so this figure carries no R-hat gate)",
               "cpu_s_per_iteration": s_iter, "cpu_gradient_evals_per_s": chains * L / s_iter,
               "one_thread": {"s_per_iteration": s_iter1, "projected_ess_per_s": ess_it / s_iter1,
                              "sample": f"{n_it1} iterations x {chains // 16} chains on 1 thread ({t_1:.1f} s), "
                                        f"scaled to {chains} chains"}}
        log(f"bench: cpu {s_iter:.4f} s/iter on {cores} threads ({info['cpu_model']}) -> projected "
            f"{cpu['value']:.1f} plain ESS/s; 1 thread {s_iter1:.3f} s/iter")
    how = {"elbo": "fitted by the reparameterised ELBO", "ml": f"fitted by maximum likelihood on the plain run's "
           f"{TRAIN_DRAWS} draws/chain", "ml+elbo": f"fitted by maximum likelihood on the plain run's {TRAIN_DRAWS} "
           f"draws/chain, then the ELBO"}[fit]
    desc = (f"Neal's funnel D={dim}, batched_hmc L={L}, {chains} chains: plain (per-chain adaptation, {W_plain} "
            f"warmup + {TRAIN_DRAWS} draws) vs an affine-coupling flow ({layers} layers, width {H}) {how}, "
            f"shared-adaptation HMC in its base space ({W_flow} warmup + {S_flow} draws)")
    thin = max(1, S_flow // TRAIN_DRAWS)
    line = {
        "metric": METRIC, "value": flow_s["value"], "unit": "ESS/s", "n_gpus": 1, "steps": steps_blocks,
        "warmup": warm_blocks, "ms_per_step": 1e3 * fres.sampling_seconds / steps_blocks, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (bf16 flow GEMMs)",
        "data": "synthetic (no data; seed %d)" % args.seed,
        "config": {"workload": desc, "chains": chains, "flow_fit": fit, "step": f"{STEP_ITERS} sampling iterations",
                   "warmup_unit": f"{WARMUP_ITERS} warmup iterations", "num_samples": S_flow, "num_warmup": W_flow},
        "ungated_value": flow_s["ungated_value"], "rhat_max": flow_s["rhat_max"], "rhat_ok": flow_s["rhat_ok"],
        "min_bulk_ess": flow_s["min_bulk_ess"], "flow": flow_s, "plain": plain_s,
        "flow_training_seconds": t_train, "flow_fit": fit_log,
        "flow_vs_plain": (flow_s["value"] / plain_s["value"]) if (flow_s["value"] and plain_s["value"]) else None,
        "flow_vs_plain_ungated": flow_s["ungated_value"] / plain_s["ungated_value"],
        "flow_moments_vs_plain": moments_vs(D, fres.constrained_samples[:, ::thin], plain.constrained_samples),
        # the funnel's exact law: v ~ N(0, 3) and u_i = x_i e^(-v/2) ~ N(0, 1) independent of v
        # (the raw x_i have E x_i^2 = e^(9/2) carried by rare draws: an unreliable MCSE)
        "flow_moments_vs_exact": moments_vs_exact(
            D, _funnel_standardised(fres.constrained_samples[:, ::thin]), np.zeros(model.num_params),
            np.concatenate([[9.0], np.ones(model.num_params - 1)])),
        "roofline": roofline, "cpu_baseline": cpu,
        "vs_cpu": (flow_s["value"] / cpu["value"]) if (cpu and flow_s["value"]) else None,
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
    roofline = {"bound": "fp32-valu", "compute": "fp32 vector ALU (model gradient); end to end, not one kernel",
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
               "cpu_model": cpu_info()["cpu_model"],
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
        rh, eb = harness_rhat_ess(D, res.constrained_samples)
        out = {"min_bulk_ess": float(np.nanmin(eb)), "rhat_max": float(np.nanmax(rh)),
               "ess_bulk_hyper": [float(v) for v in eb[:3]], "sampling_seconds": res.sampling_seconds,
               "warmup_seconds": res.warmup_seconds, "value": float(np.nanmin(eb)) / res.sampling_seconds,
               "gradient_evals_per_s": res.sampling_gradient_evals / res.sampling_seconds,
               "divergence_rate": res.divergence_rate(), "step_size": float(res.step_size[0]),
               "hyper_mean": [float(v) for v in res.constrained_samples[:, :, :3].mean(axis=(0, 1))]}
        out["rhat_ok"] = out["rhat_max"] < RHAT_MAX
        out["ungated_value"] = out["value"]
        out["value"] = gated(out["value"], out["rhat_max"])
        log(f"bench: {label}: min ESS {out['min_bulk_ess']:.1f} rhat {out['rhat_max']:.4f} -> "
            f"{out['ungated_value']:.2f} ESS/s; hyper means {out['hyper_mean']}")
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
    roofline = {"bound": "fp64-valu", "compute": "binary64 batched Cholesky with tangents (vector FP64)",
                "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "traffic": None,
                "kernel": "k_gp_column / k_gp_panel", "flops_per_chain_gradient": flops,
                "timing": "HIP events around all sampling launches of the plain run (MEADS kernels included)"}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        info = cpu_info()
        cores = max(1, info["affinity_cores"])
        B.set_threads(cores)
        sub = min(chains, max(16, cores))
        params = plain.unconstrained_samples[:sub, -1, :].astype(np.float64).T.copy()
        n, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n < 20:
            ts = time.perf_counter()
            om.gp_eval(params)
            t_cpu += time.perf_counter() - ts
            n += 1
        s_iter = t_cpu / n * chains / sub  # MEADS: one gradient per chain per iteration
        cpu = {"value": (plain_s["min_bulk_ess"] / steps) / s_iter, "unit": "ESS/s (projected, not R-hat gated)",
               "cores": cores, "kind": "port", "cpu_model": info["cpu_model"], "host": info, "sample": f"{n} binary64 oracle GP gradients of {sub} chains ({t_cpu:.1f} s), scaled "
                                          f"to {chains} chains (one gradient per MEADS iteration)",
               "cpu_s_per_iteration": s_iter}
        log(f"bench: cpu {s_iter:.1f} s/iter -> projected {cpu['value']:.4f} ESS/s")
    line = {
        "metric": METRIC_MEADS, "value": flow_s["value"], "ungated_value": flow_s["ungated_value"], "unit": "ESS/s",
        "n_gpus": 1, "steps": fres.num_samples,
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
    rh, eb = harness_rhat_ess(D, res.constrained_samples)
    min_ess, rhat_max = float(np.nanmin(eb)), float(np.nanmax(rh))
    raw = min_ess / res.sampling_seconds
    value = gated(raw, rhat_max)
    log(f"bench: rhat max {rhat_max:.4f} min bulk ESS {min_ess:.1f} -> {raw:.1f} ESS/s")
    evals = res.sampling_gradient_evals
    flops = evals * 4.0 * N * Dm
    ksec = res.sampling_kernel_seconds
    sampling_tflops = flops / ksec / 1e12 if ksec > 0 else 0.0
    mean_steps = float(res.integration_steps.mean())
    # the batch gradient's own device time: HIP events around back-to-back launches of
    # the samplers' gradient (k_split_bf16 + k_glm_grad + k_glm_reduce) on the workspace
    # stream, at the run's last draws
    ws = ut.DeviceBatchedWorkspace(model, chains)
    last = res.unconstrained_samples[:, -1, :].T
    grad_ms = {nl: 1e-3 * ws.gradient_kernel_time(last, reps=20, need_logjoint=nl) for nl in (False, True)}
    del ws
    fl_batch = 4.0 * N * Dm * chains
    achieved = fl_batch / (grad_ms[False] * 1e-3) / 1e12
    traffic = None
    pmc = ROOT / "profiles" / "k_glm_grad_cfg3_pmc.json"
    if pmc.exists():
        traffic = json.loads(pmc.read_text()).get("traffic_bytes_per_launch")
    roofline = {"bound": "mfma", "compute": "bf16 MFMA, fp32 accumulation (hi+lo operand pairs: 2x issued)",
                "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic, "kernel": "k_glm_grad",
                "timing": "HIP events around 20 back-to-back batch gradients of all chains (k_split_bf16 + "
                          "k_glm_grad<false> + k_glm_reduce) on the workspace stream, at the run's last draws "
                          "(ut_gradient_kernel_time)",
                "ms_per_batch_gradient": grad_ms[False], "ms_per_batch_gradient_with_logjoint": grad_ms[True],
                "flops_per_batch_gradient": fl_batch,
                "sampling_tflops": sampling_tflops,
                "sampling_note": "algorithmic flops of all sampling gradients over the device time of every "
                                 "sampling launch (gradient + trajectory kernels)"}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import binding as B
        om = B.OracleModel(model)
        # the oracle gradient runs an OpenMP loop over chains with the default team
        info = cpu_info()
        cores = max(1, info["affinity_cores"])
        B.set_threads(cores)
        sub = min(chains, max(64, cores))
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
               "cpu_model": info["cpu_model"], "host": info,
               "sample": f"{n} binary64 oracle gradients of {sub} chains ({t_cpu:.1f} s), scaled to {chains} "
                         f"chains x {mean_steps:.2f} leapfrog steps per iteration; projected = GPU bulk ESS per "
                         f"iteration / CPU s per iteration",
               "cpu_s_per_batch_gradient": s_batch}
        log(f"bench: cpu {s_batch:.2f} s per batch gradient -> projected {cpu['value']:.3f} ESS/s")
    line = {
        "metric": METRIC_CHEES, "value": value, "ungated_value": raw, "unit": "ESS/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * res.sampling_seconds / steps, "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic (Philox4x32-10, seed %d)" % args.seed,
        "config": {"workload": "logistic regression N=131072 D=256 bf16 X, batched_chees lock-step, "
                               f"{chains} chains, fused MFMA gradient",
                   "chains": chains, "forward_velocity": bool(args.forward_velocity)},
        "rhat_max": rhat_max, "rhat_ok": rhat_max < 1.01, "min_bulk_ess": min_ess,
        "sampling_seconds": res.sampling_seconds, "warmup_seconds": res.warmup_seconds,
        "gradient_evals_per_s": evals / res.sampling_seconds, "mean_leapfrog_steps": mean_steps,
        "trajectory_time": float(res.trajectory_trace[-1]) if warmup else None,
        "step_size": float(res.step_size[0]), "divergence_rate": res.divergence_rate(),
        "roofline": roofline, "cpu_baseline": cpu,
        "vs_cpu": (value / cpu["value"]) if (cpu and value is not None) else None,
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
        frh, feb = harness_rhat_ess(D, fres.constrained_samples)
        fmin = float(np.nanmin(feb))
        fval_raw = fmin / fres.sampling_seconds
        fval = gated(fval_raw, float(np.nanmax(frh)))
        mom = moments_vs(D, fres.constrained_samples, res.constrained_samples)
        log(f"bench: flow: rhat {float(np.nanmax(frh)):.4f} min ESS {fmin:.1f} -> {fval:.1f} ESS/s, "
            f"{float(fres.integration_steps.mean()):.2f} steps/iter; vs plain {mom}")
        line["flow"] = {"value": fval, "ungated_value": fval_raw, "min_bulk_ess": fmin, "rhat_max": float(np.nanmax(frh)),
                        "sampling_seconds": fres.sampling_seconds, "warmup_seconds": fres.warmup_seconds,
                        "mean_leapfrog_steps": float(fres.integration_steps.mean()),
                        "gradient_evals_per_s": fres.sampling_gradient_evals / fres.sampling_seconds,
                        "step_size": float(fres.step_size[0]), "training_seconds": t_train,
                        "moments_vs_plain": mom}
        line["flow_vs_plain"] = fval / value if (fval is not None and value) else None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
