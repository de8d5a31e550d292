"""Benchmark: min bulk-ESS/s at R-hat < 1.01 of batched NUTS on one MI355X.

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
# Algorithmic work per NUTS leaf of the J=64 x 32 model: one chain-gradient
# (6 N + 10 J + 20 = 12,948 flops) plus kick/drift/U-turn/energy (12 P = 792 flops).
FLOPS_PER_GRADIENT = 6 * 2048 + 10 * 64 + 20
FLOPS_PER_LEAF_EXTRA = 12 * 66


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000, help="sampling iterations")
    ap.add_argument("--warmup", type=int, default=1000, help="warmup iterations")
    ap.add_argument("--chains", type=int, default=1024)
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


if __name__ == "__main__":
    sys.exit(main())
