"""Benchmark of the PNJL equilibrium-scan hot path on one MI355X.

A "step" is one full scan of the workload's (T, mu, xi) grid through the
persistent-kernel solver, with the grid already resident in HBM.  The value
is points solved per second at |res| = ||grad_x P||_inf < 1e-10.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads (grids as linspace including endpoints, MeV -> fm^-1 with 197.327):
  c2_aniso_plain  64 x 64 x 9 (T 20-300, mu 0-450 MeV, xi -0.4..0.4 step 0.1),
                  36,864 points, p = 64, t = 4, DefaultSeed per point
                  (auto hint -> Newton -> trust region -> MultiSeed), iterations 40
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HB = 197.327
PEAK_FP64_TFLOPS = 78.6
# FP64 FLOPs of one node x flavour evaluation of the fused integrand (value,
# gradient, Hessian and thermodynamic sums), counted on the gfx950 ISA of the
# node loop: 94 add + 65 mul + 98 FMA (x2) + 5 rcp/rsq = 360.
C_NODE = 360.0


def grid_c2():
    xis = np.round(np.arange(-0.4, 0.41, 0.1), 10)
    T, mu, xi = np.meshgrid(np.linspace(20, 300, 64), np.linspace(0, 450, 64), xis, indexing="ij")
    return T.ravel() / HB, mu.ravel() / HB, xi.ravel()


WORKLOADS = {
    "c2_aniso_plain": dict(grid=grid_c2, p_num=64, t_num=4, policy=0, max_iter=40,
                           desc="64x64x9 (T 20-300 MeV, mu 0-450 MeV, xi -0.4..0.4), p=64 t=4, "
                                "DefaultSeed per point, FP64, |res|<1e-10"),
}


def cpu_baseline(w, T, mu, xi, n_sample=12000, threads=None):
    """The oracle (plain-C restatement of the reference algorithm) on a seeded
    sample of the same workload, OpenMP over points."""
    from oracle import oracle as O
    threads = threads or min(os.cpu_count() or 1, 16)
    rng = np.random.default_rng(20261202)
    idx = rng.choice(T.size, min(n_sample, T.size), replace=False)
    opts = O.default_opts(ftol=1e-10, max_iter=w["max_iter"])
    t0 = time.perf_counter()
    r = O.solve_batch(T[idx], mu[idx], xi[idx], policy=w["policy"], p_num=w["p_num"], t_num=w["t_num"],
                      opts=opts, nthreads=threads)
    dt = time.perf_counter() - t0
    solved = int(np.sum((r["converged"] == 1) & (r["residual_norm"] < 1e-10)))
    return dict(value=solved / dt, unit="points/s", cores=threads, kind="port",
                sample=f"{idx.size} seeded-random points of the same grid (rng 20261202), {threads} OpenMP threads, "
                       f"{dt:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2_aniso_plain", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the scan would shard by (xi, T) slabs with no collectives, which is not built", file=sys.stderr)
        sys.exit(2)

    import torch

    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd.pnjl import load_phase_boundary

    w = WORKLOADS[args.workload]
    T, mu, xi = w["grid"]()
    n = T.size
    dev = torch.device("cuda", 0)
    ctx = _lib.Context(0)
    rows, cep, _ = load_phase_boundary(0.0)
    ctx.set_phase_boundary(rows, cep)
    opts = _lib.default_opts(ftol=1e-10, max_iter=w["max_iter"])
    Td, mud, xid = (torch.tensor(a, device=dev) for a in (T, mu, xi))
    out = None
    stream = torch.cuda.current_stream(dev)

    def step():
        nonlocal out
        out = ctx.solve_batch(Td, mud, xid, policy=w["policy"], p_num=w["p_num"], t_num=w["t_num"], opts=opts,
                              out=out, stream=stream, sync=True)
        return ctx.last_kernel_ms()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kms.append(step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    conv = out["converged"].cpu().numpy()
    res = out["residual_norm"].cpu().numpy()
    passes = out["passes"].cpu().numpy().astype(np.float64)
    solved = int(np.sum((conv == 1) & (res < 1e-10)))
    ms_per_step = elapsed / args.steps * 1e3
    t_eff = np.where(xi == 0.0, 1, w["t_num"])
    node_evals = float(np.sum(passes * w["p_num"] * t_eff))
    flops = node_evals * C_NODE
    kms_avg = float(np.mean(kms))
    achieved = flops / (kms_avg * 1e-3) / 1e12
    line = {
        "metric": "PNJL gap points solved/s at |res|<1e-10",
        "value": solved / (ms_per_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "points": n, "solved": solved, "p_num": w["p_num"],
                   "t_num": w["t_num"], "seed_policy": "DefaultSeed", "detail": w["desc"],
                   "launch": ctx.launch_cfg(), "kernel_ms": kms_avg,
                   "mean_iterations": float(np.mean(out["iterations"].cpu().numpy())),
                   "fallback_fraction": float(np.mean(out["flags"].cpu().numpy() != 0))},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "node_evals_per_launch": node_evals, "flops_per_node": C_NODE},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, T, mu, xi)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
