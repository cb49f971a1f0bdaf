"""Benchmark of the PNJL equilibrium-scan hot path on one MI355X.

A "step" is one full scan of the workload's (T, mu, xi) grid, with the grid
already resident in HBM; the value is points solved per second at
|res| = ||grad_x P||_inf < 1e-10 (converged, physical).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

Workloads (grids are linspace with endpoints; MeV -> fm^-1 with 197.327):
  c4_surrogate    (default) 256 x 256 x 17: T 20-300 MeV, mu 0-450 MeV,
                  xi -0.4..0.4 step 0.05 = 1,114,112 points, p = 64, t = 4,
                  iterations 40.  One step = MLP inference of the seeds
                  (BF16 MFMA) + the learned-seed scan (8-step cutoff,
                  DefaultSeed fallback, reference judge in the learned
                  ambiguity mask).  Set-up (surrogate.learned_seed),
                  untimed and reported: 200,000 plain roots at
                  U([20,300] x [0,450] MeV x [-0.4,0.4]) (default_rng(4)),
                  Adam batch 8192, 8,000 steps (lr 3e-3, 1.5e-3, 6e-4,
                  2e-4 for 2,000 steps each) on the GPU, judge mask
                  from those and 10^6 more probes (default_rng(5)).  The plain
                  DefaultSeed scan of the same grid is timed with the same K/W
                  and reported beside it.  The default run also times every
                  other workload below (child processes, 3 steps after 1
                  warm-up, no CPU leg) into config.other_workloads
                  (--no-others skips them).
  c5_multiroot    512 x 512 x 5: T 10-250 MeV, mu 200-450 MeV, xi -0.2..0.2
                  step 0.1 = 1,310,720 points, p = 64, t = 4, iterations 40.
                  One step = MLP inference + MULTIROOT3 (learned seed, HADRON
                  and QUARK seeds, minimum-Omega root, first-order flag);
                  learned seed trained as for c4_surrogate.  The plain
                  minimum-Omega policy, MultiSeed (6 seeds), is timed beside it.
  c1_continuity   32 x 32 at xi = 0 (T 50-300 MeV, mu 0-400 MeV) = 1,024 points,
                  p = 64, t = 8, one ContinuitySeed chain in scan order (T
                  outer, mu inner, no reset; first seed DefaultSeed): a
                  latency-bound serial chain, one lane group.
  c2_aniso_plain  64 x 64 x 9 (xi step 0.1) = 36,864 points, DefaultSeed.
  c3_polarization 256 equilibrium points drawn without replacement
                  (default_rng(20261202)) from the converged 32 x 32 (xi = 0,
                  T 50-300 MeV, mu 0-400 MeV, p = 64, t = 8) and 64 x 64 x 9
                  grids; Pi for {P, S} x {u ubar, u sbar, s sbar} on k0 =
                  linspace(0, 4, 128) x k = linspace(0, 3, 64) fm^-1 =
                  12,582,912 complex values; A_u, A_s with 64 momentum nodes
                  on [0, 10] (4 cos nodes when xi != 0).  One step = one
                  replay of the captured grid graph (A kernel + Pi kernel).
"""
import argparse
import ctypes as C
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HB = 197.327


def host_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS when the environment sets
    it (the host's CPU share for one GPU), else every core this process may
    run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_cpu() -> dict:
    """The host the CPU baseline ran on: model name, logical CPUs, the ones
    this process may use, and OMP_NUM_THREADS."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(cpu_model=model, nproc=len(os.sched_getaffinity(0)), logical_cpus=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"))
METRIC = "PNJL gap points solved/s at |res|<1e-10: surrogate-seeded vs plain Newton"
PEAK_FP64_TFLOPS = 78.6
# FP64 FLOPs of one node x flavour evaluation of the fused integrand (value,
# gradient, Hessian and thermodynamic sums) as the device routine executes it:
# FP64 FMA = 2, add/mul/transcendental = 1, measured with rocprofv3 FP64 VALU
# counters on the residual kernel at p = 64 and p = 32
# (python3 bench.py --flop-probe node): 176.0 + 2.0.
C_NODE = 178.0


def grid(nT, nmu, xis, T_range=(20, 300), mu_range=(0, 450)):
    T, mu, xi = np.meshgrid(np.linspace(*T_range, nT), np.linspace(*mu_range, nmu), np.asarray(xis, float),
                            indexing="ij")
    return T.ravel() / HB, mu.ravel() / HB, xi.ravel()


WORKLOADS = {
    "c4_surrogate": dict(grid=lambda: grid(256, 256, np.round(np.arange(-0.4, 0.4001, 0.05), 10)), p_num=64,
                         t_num=4, max_iter=40, learned=True, policy="SURROGATE", plain_policy="DEFAULT",
                         desc="256x256x17 (T 20-300 MeV, mu 0-450 MeV, xi -0.4..0.4 step 0.05), p=64 t=4, "
                              "learned seed (MLP 3-128-128-128-4, BF16 MFMA) -> Newton, 8-step cutoff, "
                              "DefaultSeed fallback, reference judge; FP64; |res|<1e-10"),
    "c2_aniso_plain": dict(grid=lambda: grid(64, 64, np.round(np.arange(-0.4, 0.41, 0.1), 10)), p_num=64, t_num=4,
                           max_iter=40, learned=False, plain_policy="DEFAULT",
                           desc="64x64x9 (T 20-300 MeV, mu 0-450 MeV, xi -0.4..0.4), p=64 t=4, "
                                "DefaultSeed per point, FP64, |res|<1e-10"),
    "c1_continuity": dict(grid=lambda: grid(32, 32, [0.0], (50, 300), (0, 400)), p_num=64, t_num=8, max_iter=1000,
                          learned=False, plain_policy="CONTINUITY", chain=True,
                          desc="32x32 at xi=0 (T 50-300 MeV, mu 0-400 MeV), p=64 t=8, one ContinuitySeed chain "
                               "(T outer, mu inner, no reset), FP64, |res|<1e-10"),
    "c3_polarization": dict(polar=True, desc="256 points x {P,S} x {uu,us,ss} x 128 k0 x 64 k = 12,582,912 complex "
                                              "Pi (polarization_aniso), hipGraph replay, FP64"),
    "propagators": dict(props=True, desc="1,048,576 calculate_propagator_by_channel queries: the 34 (process, "
                                         "channel) entries of the process table uniformly, at the 256 equilibrium "
                                         "points of the Pi workload, (k0, k) on its axes; every Pi evaluated "
                                         "(no memo), couplings from A; FP64"),
    "meson_masses": dict(masses=True, desc="solve_meson_mass for the 8 mesons (pi, K, sigma_pi, sigma_K, eta, eta', "
                                           "sigma, sigma') at every converged point of the xi = 0 slice of the 256 x 256 "
                                           "(T 20-300 MeV, mu 0-450 MeV) scan: k = 0, default guess, NLsolve defaults "
                                           "(trust region, central-difference Jacobian, ftol 1e-8, 1000 iterations); "
                                           "FP64"),
    "dual_branch": dict(dual=True, desc="run_dual_branch_scan for 655 (T, xi) lines (T 20-150 MeV step 1 x xi -0.2..0.2 "
                                        "step 0.1) over mu 0-450 MeV step 2 (226 points): hadron branch forward, "
                                        "quark branch backward, continuation seeds, p24 t8, FP64"),
    "maxwell_line": dict(maxwell=True, desc="S-curves (hadron and quark branches followed in mu, the unstable "
                         "branch from 7 starts between them) and Maxwell construction for 45 (T, xi) lines below the "
                         "CEP (xi 0, 0.2, 0.4; T 50 MeV up in 5 MeV steps) x mu 250-420 MeV step 0.25, p24 t8, FP64"),
    "tmu_scan": dict(tmu=True, desc="run_tmu_scan over T 20-250 MeV step 1 (231 lines) x mu 0-450 MeV step 2 "
                                      "(226 points), xi = 0: phase-aware / continuation / hadron / quark candidates, "
                                      "refine and promote, p24 t8, FP64"),
    "c5_multiroot": dict(grid=lambda: grid(512, 512, [-0.2, -0.1, 0.0, 0.1, 0.2], (10, 250), (200, 450)), p_num=64,
                         t_num=4, max_iter=40, learned=True, policy="MULTIROOT3", plain_policy="MULTI",
                         desc="512x512x5 (T 10-250 MeV, mu 200-450 MeV, xi -0.2..0.2 step 0.1), p=64 t=4, "
                              "MULTIROOT3: learned seed + HADRON + QUARK Newton solves, minimum Omega, "
                              "first-order flag; FP64; |res|<1e-10"),
}


C3_METRIC = "anisotropic one-loop polarization values/s (12,582,912-value Pi grid)"
# FP64 FLOPs per complex Pi value of this grid, data-dependent and therefore
# counted once: rocprofv3 FP64 VALU counters (FMA = 2) on 131,072 seeded cells
# of the same grid evaluated one per wavefront, so that inactive lanes of
# divergent waves are not counted (python3 bench.py --flop-probe pi):
# 86,571 per cell (one symmetrised B0 + correction and one Pi), two Pi values
# per cell.  The A kernel (512 values) is negligible.
C3_FLOPS_PER_VALUE = 43286.0


def c3_points(ctx, opts):
    """The 256 equilibrium points of the Pi workload (T, mu, xi, M_u, M_s, Phi, Phib)."""
    from julia_relaxtime_amd.workloads import pi_grid_points
    return pi_grid_points(ctx, opts)


def c3_cpu_baseline(pts, k0, k, n_sample=2_000_000, threads=None):
    """The oracle's polarization_aniso (C restatement, OpenMP over values) on
    a seeded sample of the grid's Pi values, A from the oracle as well."""
    from oracle import oracle as O
    threads = threads or host_threads()
    xp, wp = O.gauleg(0.0, 10.0, 64)
    xc, wc = np.polynomial.legendre.leggauss(4)
    t0 = time.perf_counter()
    A = np.array([[O.A_aniso(m, p[1], p[0], p[5], p[6], p[2], xp, wp, xc, wc) for m in (p[3], p[4])] for p in pts])
    rng = np.random.default_rng(20261202)
    n = n_sample
    pt, pair, ch = rng.integers(0, len(pts), n), rng.integers(0, 3, n), rng.integers(1, 3, n)
    i0, j = rng.integers(0, k0.size, n), rng.integers(0, k.size, n)
    P = pts[pt]
    m1 = np.where(pair == 2, P[:, 4], P[:, 3])
    m2 = np.where(pair == 0, P[:, 3], P[:, 4])
    A1 = np.where(pair == 2, A[pt, 1], A[pt, 0])
    A2 = np.where(pair == 0, A[pt, 0], A[pt, 1])
    qa = np.zeros(n, dtype=np.dtype([(f, np.int32 if t is C.c_int else np.float64) for f, t in O.PolQuery._fields_]))
    for f, v in dict(channel=ch, num_s=pair, k0=k0[i0], k=k[j], m1=m1, m2=m2, mu1=P[:, 1], mu2=P[:, 1], T=P[:, 0],
                     Phi=P[:, 5], Phib=P[:, 6], xi=P[:, 2], A1=A1, A2=A2).items():
        qa[f] = v
    q = (O.PolQuery * n).from_buffer(qa)
    t1 = time.perf_counter()
    O.polarization_batch(q, nthreads=threads)
    dt = time.perf_counter() - t1
    return dict(value=n / dt, unit="Pi values/s", cores=threads, kind="port",
                sample=f"{n} seeded-random values of the same grid (rng 20261202), oracle polarization_aniso, "
                       f"{threads} OpenMP threads, {dt:.2f} s wall (+{t1 - t0:.2f} s for 512 A values)")


def run_polar(args, w):
    import torch
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import polarization as PZ
    dev = torch.device("cuda", 0)
    ctx = _lib.Context(0)
    pts = c3_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    from julia_relaxtime_amd.workloads import PI_K, PI_K0
    k0, k = PI_K0, PI_K
    ptd = torch.tensor(pts, device=dev)
    k0d, kd = torch.tensor(k0, device=dev), torch.tensor(k, device=dev)
    stream = torch.cuda.current_stream(dev)
    state = {}
    pctx = PZ._ctx(0)

    def step():
        state["out"] = PZ.polarization_grid(ptd, k0d, kd, out=state.get("out"), stream=stream)
        return pctx.last_kernel_ms()

    ms_step, graph_ms = timed_steps(step, args.steps, args.warmup)
    re, im = (t.cpu().numpy() for t in state["out"])
    nval = re.size
    graph_avg = float(np.mean(graph_ms))
    flops = None if C3_FLOPS_PER_VALUE is None else C3_FLOPS_PER_VALUE * nval
    achieved = None if flops is None else flops / (graph_avg * 1e-3) / 1e12
    line = {
        "metric": C3_METRIC, "value": nval / (ms_step * 1e-3), "unit": "Pi values/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, points=len(pts), values=nval, detail=w["desc"],
                       nan_imag_values=int(np.isnan(im).sum()), nan_real_values=int(np.isnan(re).sum()),
                       inf_values=int(np.isinf(re).sum() + np.isinf(im).sum()), graph_ms_per_step=graph_avg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": None if achieved is None else achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "polar::grid_pi_kernel (+ grid_a_kernel)", "kernel_ms_per_step": graph_avg,
                     "flops_per_value": C3_FLOPS_PER_VALUE},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(c3_cpu_baseline(pts, k0, k), **host_cpu())
    print(json.dumps(line), flush=True)


def cpu_baseline(w, T, mu, xi, n_sample=20000, threads=None):
    """The oracle (plain-C restatement of the reference's solve with the
    workload's plain policy, DefaultSeed or MultiSeed) on a seeded sample of
    the same grid, OpenMP over points."""
    from oracle import oracle as O
    threads = threads or host_threads()
    rng = np.random.default_rng(20261202)
    idx = rng.choice(T.size, min(n_sample, T.size), replace=False)
    opts = O.default_opts(ftol=1e-10, max_iter=w["max_iter"])
    policy = getattr(O, "SEED_" + w["plain_policy"])
    if w.get("chain"):
        # a chain is sequential: the whole chain, one thread
        threads, idx = 1, np.arange(T.size)
        opts = O.default_opts(ftol=1e-10, max_iter=w["max_iter"], chain_len=T.size)
    t0 = time.perf_counter()
    r = O.solve_batch(T[idx], mu[idx], xi[idx], policy=policy, p_num=w["p_num"], t_num=w["t_num"],
                      opts=opts, nthreads=threads)
    dt = time.perf_counter() - t0
    solved = int(np.sum((r["converged"] == 1) & (r["residual_norm"] < 1e-10)))
    name = {"DEFAULT": "DefaultSeed", "MULTI": "MultiSeed", "CONTINUITY": "ContinuitySeed"}[w["plain_policy"]]
    what = "the whole chain" if w.get("chain") else f"{idx.size} seeded-random points of the same grid (rng 20261202)"
    return dict(value=solved / dt, unit="points/s", cores=threads, kind="port",
                sample=f"{what}, plain {name} solve, {threads} OpenMP threads, {dt:.2f} s wall")


# HBM bytes per scan_kernel launch of the learned-seed scan (c4_surrogate):
# FETCH_SIZE + WRITE_SIZE of its two launches per step, from two rocprofv3
# --pmc runs of this bench over three timed steps (FETCH 0.1297 + 0.0071 GiB,
# WRITE 0.2935 + 0.0386 GiB per step; profiles/scan_kernel_c4_surrogate_pmc.md),
# divided by the two launches.  A 1 GiB device copy reads 0.5 GiB in
# FETCH_SIZE and exactly 1 GiB in WRITE_SIZE, so FETCH is scaled by 2.  The
# algorithmic bytes are ~0.25 KB per point (inputs, seeds, outputs); the rest
# is the per-point attempt records.  Re-measure with the profile whenever the
# scan kernel changes.
C4_TRAFFIC_BYTES_PER_LAUNCH = (2.0 * (0.1297 + 0.0071) + 0.2935 + 0.0386) / 2 * 2 ** 30

def train_learned_seed(ctx, w, opts):
    """The learned seed's set-up (surrogate.learned_seed): training roots,
    Adam, judge mask; `info` goes into the line's config."""
    from julia_relaxtime_amd.surrogate import learned_seed
    s, info = learned_seed(ctx, p_num=w["p_num"], t_num=w["t_num"], opts=opts)
    info.pop("loss")
    return s, info


PROP_METRIC = "meson propagator (D_S, D_P) queries/s (calculate_propagator_by_channel, 1,048,576-query batch)"
# FP64 FLOPs per query of this batch, counted once like C3_FLOPS_PER_VALUE:
# FP64 VALU counters on 65,536 of its queries run one (pair, query) per
# wavefront (python3 bench.py --flop-probe prop): 119,772 per
# query (1.68 quark pairs of B0 + correction per query on average).  The
# bucketing and composition kernels (~10^2 operations per query) are not
# counted; their time is.
PROP_FLOPS_PER_QUERY = 119772.0


def prop_cpu_baseline(q, G, K, threads=None):
    """The oracle's calculate_propagator_by_channel on the whole batch: Pi in
    C (OpenMP over values), queries built and propagators composed as numpy
    array operations."""
    from oracle import mesons as OM
    threads = threads or host_threads()
    t0 = time.perf_counter()
    OM.propagator_by_channel_vec(q, G, K, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=q.size / dt, unit="queries/s", cores=threads, kind="port",
                sample=f"the whole {q.size}-query batch, oracle Pi (C, {threads} OpenMP threads) + numpy "
                       f"composition, {dt:.2f} s wall")


def run_propagators(args, w):
    import torch
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import propagators as PR
    from julia_relaxtime_amd.workloads import pi_grid_points, propagator_queries
    dev = torch.device("cuda", 0)
    ctx = _lib.Context(0)
    pts = pi_grid_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    q = propagator_queries(ctx, pts)
    n = q.size
    qd = torch.from_numpy(q.view(np.uint8)).to(dev)
    out = torch.empty((n, 4), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    pctx = PR._ctx(0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def step():
        ev[0].record(stream)
        PR.propagator_by_channel_device(qd, out, stream=stream)
        ev[1].record(stream)
        return ev

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kms = []
    with no_gc():
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            ev[1].synchronize()
            kms.append(ev[0].elapsed_time(ev[1]))
        torch.cuda.synchronize()
        ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    o = out.cpu().numpy()
    kavg = float(np.mean(kms))
    flops = None if PROP_FLOPS_PER_QUERY is None else PROP_FLOPS_PER_QUERY * n
    achieved = None if flops is None else flops / (kavg * 1e-3) / 1e12
    line = {
        "metric": PROP_METRIC, "value": n / (ms_step * 1e-3), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, queries=n, points=len(pts), detail=w["desc"],
                       nonfinite=int((~np.isfinite(o)).sum()), kernel_ms_per_step=kavg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": None if achieved is None else achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "polar::prop_pi_kernel (+ bucket count / scan / scatter, compose)",
                     "kernel_ms_per_step": kavg,
                     "flops_per_query": PROP_FLOPS_PER_QUERY},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(prop_cpu_baseline(q, pctx.params.G, pctx.params.K), **host_cpu())
    print(json.dumps(line), flush=True)


MASS_METRIC = "meson pole masses solved/s (solve_meson_mass, 8 mesons x 65,536 (T, mu) points)"
# FP64 FLOPs per pole-mass query of this batch, counted once on 8,192 seeded
# queries run one per wavefront (python3 bench.py --flop-probe mass), summed
# over the phase launches: 822,180 per query (38.7 trust-region iterations on
# average; 36 % of the queries stop at the start on a non-finite residual).
# The kernel's time is set by its longest serial chains (queries running the
# full 1000 iterations), not by FP64 throughput.
MASS_FLOPS_PER_QUERY = 822180.0


def _mass_worker(args):
    """One process of the pole-mass CPU baseline: the oracle's
    solve_meson_mass over its share of the sampled queries."""
    rows, G, K = args
    from oracle import mesons as OM
    for r in rows:
        qp = dict(m=dict(u=r["m_u"], s=r["m_s"]), mu=dict(u=r["mu_u"], s=r["mu_s"]), A=dict(u=r["A_u"], s=r["A_s"]))
        tp = dict(T=r["T"], Phi=r["Phi"], Phib=r["Phib"], xi=r["xi"])
        OM.solve_meson_mass(OM.MASS_MESONS[int(r["meson"])], qp, tp, G=G, K=K, k=r["k"])
    return len(rows)


def mass_cpu_baseline(q, G, K, n_sample=None, seed=11, threads=None):
    """The oracle's solve_meson_mass (Pi in C, trust region in Python) on a
    seeded sample of the batch, spread over host_threads() processes
    (spawned: the workers never touch the GPU)."""
    import multiprocessing as mp
    threads = threads or host_threads()
    n_sample = n_sample or 50 * threads
    sel = np.random.default_rng(seed).choice(q.size, n_sample, replace=False)
    rows = [{k: q[i][k] for k in q.dtype.names} for i in sel]
    parts = [(rows[j::threads], G, K) for j in range(threads)]
    with mp.get_context("spawn").Pool(threads) as pool:
        pool.map(_mass_worker, [(p[0][:1], G, K) for p in parts])   # start-up (imports) outside the timing
        t0 = time.perf_counter()
        done = sum(pool.map(_mass_worker, parts))
        dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="queries/s", cores=threads, kind="port",
                sample=f"{done} seeded-random queries of the batch (rng {seed}), oracle solve_meson_mass, "
                       f"{threads} processes, {dt:.2f} s wall")


def run_masses(args, w):
    import torch
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import propagators as PR
    from julia_relaxtime_amd.workloads import meson_mass_points, meson_mass_queries
    ctx = _lib.Context(0)
    pts = meson_mass_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    q = meson_mass_queries(pts)
    n = q.size
    pctx = PR._ctx(0)
    res = {}

    def step():
        res["r"] = PR.meson_mass_batch(q)
        return pctx.last_kernel_ms()

    ms_step, kms = timed_steps(step, args.steps, args.warmup)
    r = res["r"]
    kavg = float(np.mean(kms))
    flops = None if MASS_FLOPS_PER_QUERY is None else MASS_FLOPS_PER_QUERY * n
    achieved = None if flops is None else flops / (kavg * 1e-3) / 1e12
    line = {
        "metric": MASS_METRIC, "value": n / (ms_step * 1e-3), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, queries=n, points=int((pts[:, 8] == 1).sum()), detail=w["desc"],
                       converged_fraction=float(r["converged"].mean()),
                       exception_fraction=float(r["exception"].mean()),
                       mean_iterations=float(r["iterations"].mean()), max_iterations=int(r["iterations"].max()),
                       kernel_ms_per_step=kavg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": None if achieved is None else achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "polar::mass_kernel", "kernel_ms_per_step": kavg,
                     "flops_per_query": MASS_FLOPS_PER_QUERY},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(mass_cpu_baseline(q, pctx.params.G, pctx.params.K), **host_cpu())
    print(json.dumps(line), flush=True)


DUAL_METRIC = "dual-branch scan points/s (run_dual_branch_scan, 655 lines x 226 mu x 2 branches)"


def dual_lines():
    T = np.repeat(np.arange(20.0, 150.0 + 1e-9, 1.0), 5)
    xi = np.tile(np.round(np.arange(-0.2, 0.2001, 0.1), 10), T.size // 5)
    return T, xi, np.arange(0.0, 450.0 + 1e-9, 2.0)


def dual_cpu_baseline(n_lines=40, threads=None):
    """The same dual-branch logic driven by the oracle's solves (C, OpenMP over
    the lines of each mu step) on the first n_lines lines."""
    from julia_relaxtime_amd import dual_branch as DB
    from oracle import oracle as O
    threads = threads or host_threads()
    T, xi, mu = dual_lines()
    pts = [0]

    def solve(Tq, muq, xiq, policy, seed):
        pts[0] += Tq.size
        return O.solve_batch(Tq, muq, xiq, policy=policy, seed=seed, p_num=24, t_num=8, nthreads=threads)
    t0 = time.perf_counter()
    DB.run_dual_branch_scan_batch(T[:n_lines], mu, xi[:n_lines], solver=solve)
    dt = time.perf_counter() - t0
    return dict(value=pts[0] / dt, unit="points/s", cores=threads, kind="port",
                sample=f"the first {n_lines} lines ({pts[0]} branch points), oracle solves with {threads} OpenMP "
                       f"threads per mu step, {dt:.2f} s wall")


def run_dual(args, w):
    import torch
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import dual_branch as DB
    from julia_relaxtime_amd.pnjl import context
    ctx = context(0)
    T, xi, mu = dual_lines()
    gsolve = DB.gpu_chain_solver(0, 24, 8)
    stats = dict(ms=0.0, passes=0.0, points=0, launches=0)
    t_eff = lambda x: np.where(x == 0.0, 1, 8)

    def solve(Tq, muq, xiq, policy, chain_len):
        r = gsolve(Tq, muq, xiq, policy, chain_len)
        stats["ms"] += ctx.last_kernel_ms()
        stats["passes"] += float(np.sum(r["passes"] * 24.0 * t_eff(xiq)))
        stats["points"] += int(np.sum((r["flags"] & _lib.PT_SKIPPED) == 0))
        stats["launches"] += 1
        return r
    res = {}

    def step():
        for k in stats:
            stats[k] = 0
        res["r"] = DB.run_dual_branch_scan_batch(T, mu, xi, chain_solver=solve)
        return stats["ms"]

    ms_step, kms = timed_steps(step, args.steps, args.warmup)
    found = sum(DB.find_phase_transition(r)["found"] for r in res["r"])
    kavg = float(np.mean(kms))
    achieved = stats["passes"] * C_NODE / (kavg * 1e-3) / 1e12
    line = {
        "metric": DUAL_METRIC, "value": stats["points"] / (ms_step * 1e-3), "unit": "points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, lines=int(T.size), mu_points=int(mu.size), branch_points=stats["points"],
                       launches=stats["launches"], first_order_lines=int(found), detail=w["desc"],
                       kernel_ms_per_step=kavg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "pnjl::scan_kernel",
                     "kernel_ms_per_step": kavg, "flops_per_node": C_NODE},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(dual_cpu_baseline(), **host_cpu())
    print(json.dumps(line), flush=True)


TMU_METRIC = "T-mu scan points/s (run_tmu_scan, 231 T x 226 mu)"


def tmu_grid():
    return [20.0 + i for i in range(231)], [2.0 * i for i in range(226)]


def tmu_cpu_baseline(n_lines=24, threads=None):
    """The same TmuScan logic over the oracle's fixed-seed solves (C, OpenMP
    over the lines of each mu step) on the first n_lines temperatures."""
    from julia_relaxtime_amd import tmu_scan as TS
    from oracle import oracle as O
    threads = threads or host_threads()
    Ts, mus = tmu_grid()

    def solve(T, mu, xi, seeds):
        return O.solve_batch(T, mu, xi, policy=O.SEED_EXPLICIT, seed=seeds, p_num=24, t_num=8, nthreads=threads)
    t0 = time.perf_counter()
    st = TS.run_tmu_scan(T_values=Ts[:n_lines], mu_values=mus, solver=solve)
    dt = time.perf_counter() - t0
    return dict(value=st["total"] / dt, unit="points/s", cores=threads, kind="port",
                sample=f"the first {n_lines} temperatures ({st['total']} points), oracle solves with {threads} "
                       f"OpenMP threads per mu step, {dt:.2f} s wall")


def run_tmu(args, w):
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import tmu_scan as TS
    from julia_relaxtime_amd.pnjl import context
    ctx = context(0)
    Ts, mus = tmu_grid()
    opts = _lib.default_opts()
    stats = dict(ms=0.0, passes=0.0, launches=0)

    def solve(T, mu, xi, seeds):
        saved = ctx.launch_cfg()
        ctx.set_launch_cfg(lanes_per_point=32, chunk=1, waves_per_cu=8, tail_lanes_per_point=32)
        try:
            r = ctx.solve_batch(T, mu, xi, policy=_lib.SEED_EXPLICIT, seed=seeds, p_num=24, t_num=8, opts=opts)
        finally:
            ctx.set_launch_cfg(**saved)
        stats["ms"] += ctx.last_kernel_ms()
        stats["passes"] += float(np.sum(r["passes"])) * 24.0
        stats["launches"] += 1
        return r
    res = {}

    def step():
        for k in stats:
            stats[k] = 0
        res["st"] = TS.run_tmu_scan(T_values=Ts, mu_values=mus, solver=solve)
        return stats["ms"]

    ms_step, kms = timed_steps(step, args.steps, args.warmup)
    st = res["st"]
    kavg = float(np.mean(kms))
    achieved = stats["passes"] * C_NODE / (kavg * 1e-3) / 1e12
    line = {
        "metric": TMU_METRIC, "value": st["total"] / (ms_step * 1e-3), "unit": "points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, points=st["total"], success=st["success"], failure=st["failure"],
                       launches=stats["launches"], detail=w["desc"], kernel_ms_per_step=kavg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "pnjl::scan_kernel",
                     "kernel_ms_per_step": kavg, "flops_per_node": C_NODE},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(tmu_cpu_baseline(), **host_cpu())
    print(json.dumps(line), flush=True)


MAXWELL_METRIC = "Maxwell-line S-curve points/s (maxwell_line: S-curves + Maxwell construction, 45 lines x 681 mu)"


def maxwell_lines():
    """(T, xi) lines below each xi's CEP: T 50 MeV up in 5 MeV steps (130,
    120, 110 MeV at xi 0, 0.2, 0.4), mu 250-420 MeV in 0.25 MeV steps."""
    T, xi = [], []
    for x, tmax in ((0.0, 130.0), (0.2, 120.0), (0.4, 110.0)):
        for t in np.arange(50.0, tmax + 1e-9, 5.0):
            T.append(t)
            xi.append(x)
    return np.array(T), np.array(xi), np.arange(250.0, 420.0 + 1e-9, 0.25)


def maxwell_cpu_baseline(n_lines=45, threads=None):
    """The same S-curve + Maxwell logic over the oracle's solves (C, OpenMP
    over the lines of each mu step) for the first n_lines lines."""
    from julia_relaxtime_amd import phase_transition as PT
    from oracle import oracle as O
    threads = threads or host_threads()
    T, xi, mu = maxwell_lines()
    count = dict(points=0)

    def solve(Tq, muq, xiq, policy, seed):
        count["points"] += Tq.size
        return O.solve_batch(Tq, muq, xiq, policy=policy, seed=seed, p_num=24, t_num=8, nthreads=threads)
    t0 = time.perf_counter()
    PT.maxwell_line(T[:n_lines], mu, xi[:n_lines], solver=solve, min_samples=8, detect_min_points=5)
    dt = time.perf_counter() - t0
    return dict(value=count["points"] / dt, unit="points/s", cores=threads, kind="port",
                sample=f"the first {n_lines} lines ({count['points']} FixedMu solves), oracle solves with {threads} "
                       f"OpenMP threads per batch, {dt:.2f} s wall")


def run_maxwell(args, w):
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import dual_branch as DB
    from julia_relaxtime_amd import phase_transition as PT
    from julia_relaxtime_amd.pnjl import context
    ctx = context(0)
    T, xi, mu = maxwell_lines()
    gsolve, gchain = DB.gpu_solver(0, 24, 8), DB.gpu_chain_solver(0, 24, 8)
    stats = dict(ms=0.0, passes=0.0, points=0, launches=0)
    t_eff = lambda x: np.where(x == 0.0, 1, 8)

    def count(r, xiq):
        stats["ms"] += ctx.last_kernel_ms()
        stats["passes"] += float(np.sum(r["passes"] * 24.0 * t_eff(xiq)))
        stats["points"] += int(np.sum((r["flags"] & _lib.PT_SKIPPED) == 0))
        stats["launches"] += 1
        return r

    def solve(Tq, muq, xiq, policy, seed):
        return count(gsolve(Tq, muq, xiq, policy, seed), xiq)

    def chains(Tq, muq, xiq, policy, chain_len):
        return count(gchain(Tq, muq, xiq, policy, chain_len), xiq)
    res = {}

    def step():
        for k in stats:
            stats[k] = 0
        ms = PT.maxwell_line(T, mu, xi, chain_solver=chains, point_solver=solve, min_samples=8, detect_min_points=5)
        res["m"] = list(ms.values())
        return stats["ms"]

    ms_step, kms = timed_steps(step, args.steps, args.warmup)
    kavg = float(np.mean(kms))
    achieved = stats["passes"] * C_NODE / (kavg * 1e-3) / 1e12
    line = {
        "metric": MAXWELL_METRIC, "value": stats["points"] / (ms_step * 1e-3), "unit": "points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(workload=args.workload, lines=int(T.size), mu_points=int(mu.size), solves=stats["points"],
                       launches=stats["launches"], maxwell_converged=int(sum(m.converged for m in res["m"])),
                       detail=w["desc"], kernel_ms_per_step=kavg),
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "pnjl::scan_kernel",
                     "kernel_ms_per_step": kavg, "flops_per_node": C_NODE},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = dict(maxwell_cpu_baseline(), **host_cpu())
    print(json.dumps(line), flush=True)


def root_agreement(ctx, w, opts, T, mu, xi, Td, mud, xid, Q, P, cfg, plain_policy):
    """Points where the learned-seed scan Q reports another root than the plain
    scan P (DualBranchScan's sameness test, DualBranchScan.jl:409-419, and
    Omega to 1e-10), split into points whose plain root is itself rounding-
    sensitive -- it changes when the plain scan runs at the learned scan's
    launch geometry (another summation order) or when T or mu moves by one
    ulp -- and mask misses; and the points on the same root whose x differ by
    more than 1e-9 (next to the critical end point the Hessian is nearly
    singular)."""
    from julia_relaxtime_amd.surrogate import same_root
    saved = ctx.launch_cfg()
    ctx.set_launch_cfg(**cfg)
    P2 = {k: v.cpu().numpy() for k, v in ctx.solve_batch(Td, mud, xid, policy=plain_policy, p_num=w["p_num"],
                                                          t_num=w["t_num"], opts=opts).items()}
    ctx.set_launch_cfg(**saved)
    same = same_root(Q, P)
    off = np.flatnonzero(~same)
    unstable = ~same_root({k: P2[k][off] for k in ("x", "masses", "omega")},
                          {k: P[k][off] for k in ("x", "masses", "omega")})
    if off.size:
        for dT, dm in ((np.inf, None), (-np.inf, None), (None, np.inf), (None, -np.inf)):
            Tp = T[off] if dT is None else np.nextafter(T[off], dT)
            mp = mu[off] if dm is None else np.nextafter(mu[off], dm)
            r = ctx.solve_batch(Tp, mp, xi[off], policy=plain_policy, p_num=w["p_num"], t_num=w["t_num"], opts=opts)
            unstable |= ~same_root(r, {k: P[k][off] for k in ("x", "masses", "omega")})
    apart = same & (np.max(np.abs(Q["x"] - P["x"]) / np.maximum(1, np.abs(P["x"])), axis=1) >= 1e-9)
    return dict(points_on_another_root=int(off.size), of_which_plain_root_rounding_sensitive=int(unstable.sum()),
                mask_misses=int((~unstable).sum()),
                mask_misses_higher_omega=int(np.sum(Q["omega"][off[~unstable]] > P["omega"][off[~unstable]])),
                same_root_x_apart_1e9=int(apart.sum()), same_root_as_plain=float(same.mean()),
                same_root_as_plain_equal_geometry=float(same_root(Q, P2).mean()))


def judge_everywhere(ctx, w, opts, model, Td, mud, xid, seeds, cfg, stream, args):
    """The learned-seed scan with the Omega judge on every point (the whole box
    in the mask): each converged learned root is checked against the roots of
    the HADRON and hint-QUARK seeds (all six MultiSeed candidates where one of
    them fails) and the minimum-Omega root is kept.  Timed like the step, for
    the saving of the learned mask."""
    from julia_relaxtime_amd import _lib
    full = np.ones_like(model.judge_mask)
    o = _lib.default_opts(ftol=1e-10, max_iter=w["max_iter"], surrogate_judge=_lib.JUDGE_OMEGA)
    ctx.set_judge_mask(full, model.in_lo, model.in_hi)
    ctx.set_launch_cfg(**cfg)
    holder = {}
    try:
        def step():
            holder["out"] = ctx.solve_batch(Td, mud, xid, policy=_lib.SEED_SURROGATE, seed=seeds, p_num=w["p_num"],
                                            t_num=w["t_num"], opts=o, out=holder.get("out"), stream=stream, sync=True)
            return float(np.sum(ctx.last_scan_ms()))
        ms, _ = timed_steps(step, max(1, min(args.steps, 3)), 1)
    finally:
        ctx.set_judge_mask(model.judge_mask, model.in_lo, model.in_hi)
    R = holder["out"]
    solved = int(((R["converged"] == 1) & (R["residual_norm"] < 1e-10)).sum().item())
    return dict(judge_everywhere_ms_per_step=ms, judge_everywhere_points_per_s=solved / (ms * 1e-3))


class no_gc:
    """Python's cyclic garbage collector off inside a timed region (as
    timeit does): a collection that lands between two launches of a step
    leaves the GPU idle for tens of milliseconds."""

    def __enter__(self):
        gc.collect()
        self.was = gc.isenabled()
        gc.disable()

    def __exit__(self, *exc):
        if self.was:
            gc.enable()


def timed_steps(step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    scan_ms = []
    with no_gc():
        t0 = time.perf_counter()
        for _ in range(steps):
            scan_ms.append(step())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    return dt / steps * 1e3, scan_ms


# ---------------------------------------------------------------------------
# FLOP probes (run under rocprofv3 FP64 VALU counters; each prints what it ran)
# ---------------------------------------------------------------------------
def flop_probe_node(n=65536, xi=0.3, t_num=4, seed=3):
    """FP64 FLOPs of one node x flavour evaluation of the gap integrand on the
    device routine itself: pnjl_gap_residual_batch (one node pass of every
    flavour plus the dense assembly per state) twice on the same 65,536
    symmetric states at xi = 0.3, with p = 64 and with p = 32 (t = 4).  The
    difference of the two dispatches' FP64 counts (64 lanes each, FMA = 2,
    transcendental = 1) over the difference in node evaluations (states x 2
    flavours x 32 x 4) is C_NODE; the remainder is the per-state assembly."""
    from julia_relaxtime_amd import _lib
    rng = np.random.default_rng(seed)
    T = rng.uniform(50, 300, n) / HB
    mu = rng.uniform(0, 450, n) / HB
    x = np.stack([rng.uniform(-1.9, -0.1, n), np.zeros(n), rng.uniform(-2.3, -0.5, n), rng.uniform(0.1, 0.9, n),
                  rng.uniform(0.1, 0.9, n)], 1)
    x[:, 1] = x[:, 0]
    ctx = _lib.Context(0)
    out = {}
    for p_num in (64, 32):
        ctx.residual_batch(x, T, mu, np.full(n, xi), p_num, t_num)
        out[p_num] = dict(states=n, flavours=2, node_evals=n * 2 * p_num * t_num, kernel_ms=ctx.last_kernel_ms())
    return out


def flop_probe_pi(n_sample=131072, seed=7):
    """The polarization grid's cells as single channel-P queries with
    PNJL_FLAG_ONE_PER_WAVE (one query per wavefront, so 1 lane per
    wave-instruction is that query's own FP64 work; the grid kernel's own
    counts include the inactive lanes of divergent waves).  Two channels share
    one B0: per grid value the algorithmic cost is half a query's."""
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import polarization as PZ
    from julia_relaxtime_amd.pnjl import context
    from julia_relaxtime_amd.workloads import PI_K, PI_K0, pi_grid_points
    ctx = context(0)
    pts = pi_grid_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    rng = np.random.default_rng(seed)
    pt = rng.integers(0, len(pts), n_sample)
    pair = rng.integers(0, 3, n_sample)
    i0 = rng.integers(0, PI_K0.size, n_sample)
    j = rng.integers(0, PI_K.size, n_sample)
    P = pts[pt]
    q = np.zeros(n_sample, dtype=PZ.QUERY_DTYPE)
    q["channel"] = PZ.CHANNEL_P
    q["num_s"] = pair
    q["k0"], q["k"] = PI_K0[i0], PI_K[j]
    q["m1"] = np.where(pair == 2, P[:, 4], P[:, 3])
    q["m2"] = np.where(pair == 0, P[:, 3], P[:, 4])
    q["mu1"] = q["mu2"] = P[:, 1]
    q["T"], q["xi"], q["Phi"], q["Phib"] = P[:, 0], P[:, 2], P[:, 5], P[:, 6]
    PZ.polarization_batch(q, one_per_wave=True)
    return {"queries": int(n_sample), "seed": seed, "cells": int(len(pts) * 3 * PI_K0.size * PI_K.size),
            "kernel_ms": ctx.last_kernel_ms()}


def flop_probe_prop(n_sample=65536):
    """The first n_sample queries of the propagator workload with
    PNJL_FLAG_ONE_PER_WAVE: one (pair, query) per wavefront."""
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import propagators as PR
    from julia_relaxtime_amd.pnjl import context
    from julia_relaxtime_amd.workloads import pi_grid_points, propagator_queries
    ctx = context(0)
    pts = pi_grid_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    q = propagator_queries(ctx, pts, n=1 << 20)[:n_sample]
    PR.propagator_by_channel_batch(q, one_per_wave=True)
    return {"queries": int(n_sample), "workload_queries": 1 << 20, "seed": 20261203,
            "kernel_ms": ctx.last_kernel_ms()}


def flop_probe_mass(n_sample=8192, seed=11):
    """A seeded sample of the pole-mass workload's queries with
    PNJL_FLAG_ONE_PER_WAVE (one query per wavefront)."""
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import propagators as PR
    from julia_relaxtime_amd.pnjl import context
    from julia_relaxtime_amd.workloads import meson_mass_points, meson_mass_queries
    ctx = context(0)
    q = meson_mass_queries(meson_mass_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40)))
    sel = np.sort(np.random.default_rng(seed).choice(q.size, n_sample, replace=False))
    r = PR.meson_mass_batch(q[sel], one_per_wave=True)
    return {"queries": int(n_sample), "workload_queries": int(q.size), "seed": seed,
            "iterations": int(r["iterations"].sum()), "kernel_ms": ctx.last_kernel_ms()}


FLOP_PROBES = {"node": flop_probe_node, "pi": flop_probe_pi, "prop": flop_probe_prop, "mass": flop_probe_mass}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4_surrogate", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-others", action="store_true",
                    help="default workload only: skip timing the other workloads (child processes)")
    ap.add_argument("--launch", default=None,
                    help="lanes,chunk,waves_per_cu[,tail_lanes] for both scans instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--profile", action="store_true",
                    help="profiling runs: skip the root-agreement and judge-everywhere comparisons")
    ap.add_argument("--flop-probe", default=None, choices=sorted(FLOP_PROBES),
                    help="run one FLOP probe (for rocprofv3 FP64 counters) instead of a workload")
    args = ap.parse_args()
    if args.flop_probe:
        import torch  # noqa: F401  (loads the HIP runtime before the library)
        print(json.dumps(FLOP_PROBES[args.flop_probe]()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the scan would shard by (xi, T) slabs with no collectives, which is not built", file=sys.stderr)
        sys.exit(2)

    import torch

    from julia_relaxtime_amd import _lib

    w = WORKLOADS[args.workload]
    if w.get("polar"):
        return run_polar(args, w)
    if w.get("props"):
        return run_propagators(args, w)
    if w.get("masses"):
        return run_masses(args, w)
    if w.get("dual"):
        return run_dual(args, w)
    if w.get("maxwell"):
        return run_maxwell(args, w)
    if w.get("tmu"):
        return run_tmu(args, w)
    T, mu, xi = w["grid"]()
    n = T.size
    dev = torch.device("cuda", 0)
    ctx = _lib.Context(0)
    opts = _lib.default_opts(ftol=1e-10, max_iter=w["max_iter"], chain_len=n if w.get("chain") else 0)
    Td, mud, xid = (torch.tensor(a, device=dev) for a in (T, mu, xi))
    stream = torch.cuda.current_stream(dev)
    t_eff = np.where(xi == 0.0, 1, w["t_num"])

    def scan_flops(out):
        passes = out["passes"].cpu().numpy().astype(np.float64)
        return float(np.sum(passes * w["p_num"] * t_eff)) * C_NODE

    fixed = None
    if args.launch:
        v = [int(t) for t in args.launch.split(",")] + [0]
        fixed = dict(lanes_per_point=v[0], chunk=v[1], waves_per_cu=v[2], tail_lanes_per_point=v[3])
    # the plain scan (DefaultSeed, or MultiSeed for minimum-Omega workloads),
    # its launch geometry autotuned on the full grid
    plain_policy = getattr(_lib, "SEED_" + w["plain_policy"])
    plain_cfg = fixed or ctx.autotune(T, mu, xi, policy=plain_policy, p_num=w["p_num"], t_num=w["t_num"],
                                      opts=opts)[0]
    plain = {}

    def plain_step():
        ctx.set_launch_cfg(**plain_cfg)
        plain["out"] = ctx.solve_batch(Td, mud, xid, policy=plain_policy, p_num=w["p_num"], t_num=w["t_num"],
                                       opts=opts, out=plain.get("out"), stream=stream, sync=True)
        return float(np.sum(ctx.last_scan_ms()))

    ms_plain, scan_plain = timed_steps(plain_step, args.steps, args.warmup)
    P = {k: v.cpu().numpy() for k, v in plain["out"].items()}
    solved_plain = int(np.sum((P["converged"] == 1) & (P["residual_norm"] < 1e-10)))
    flops_plain = scan_flops(plain["out"])

    extra = {}
    if w["learned"]:
        model, extra = train_learned_seed(ctx, w, opts)
        ctx.set_judge_mask(model.judge_mask, model.in_lo, model.in_hi)
        state = {"seeds": torch.empty((n, 5), dtype=torch.float64, device=dev)}
        policy = getattr(_lib, "SEED_" + w["policy"])
        cfg = fixed or ctx.autotune(T, mu, xi, policy=policy, seed=model.infer(T, mu, xi),
                                    p_num=w["p_num"], t_num=w["t_num"], opts=opts)[0]

        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        gpu_ms = []

        def step():
            ctx.set_launch_cfg(**cfg)
            ev[0].record(stream)
            model.infer(Td, mud, xid, out=state["seeds"], stream=stream, sync=False)
            state["out"] = ctx.solve_batch(Td, mud, xid, policy=policy, seed=state["seeds"],
                                           p_num=w["p_num"], t_num=w["t_num"], opts=opts, out=state.get("out"),
                                           stream=stream, sync=True)
            ev[1].record(stream)
            ev[1].synchronize()
            gpu_ms.append(ev[0].elapsed_time(ev[1]))
            state["stage_ms"] = ctx.last_scan_ms()
            return float(np.sum(state["stage_ms"]))

        ms_step, scan_ms = timed_steps(step, args.steps, args.warmup)
        gpu_ms = gpu_ms[args.warmup:]
        Q = {k: v.cpu().numpy() for k, v in state["out"].items()}
        flops = scan_flops(state["out"])
        solved = int(np.sum((Q["converged"] == 1) & (Q["residual_norm"] < 1e-10)))
        seed_err = np.abs(state["seeds"].cpu().numpy() - Q["x"]).max(axis=1)
        if w["policy"] == "SURROGATE" and not args.profile:
            extra.update(root_agreement(ctx, w, opts, T, mu, xi, Td, mud, xid, Q, P, cfg, plain_policy))
            extra.update(judge_everywhere(ctx, w, opts, model, Td, mud, xid, state["seeds"], cfg, stream, args))
        extra.update(
            plain_points_per_s=solved_plain / (ms_plain * 1e-3), plain_ms_per_step=ms_plain,
            speedup_vs_plain=(solved / (ms_step * 1e-3)) / (solved_plain / (ms_plain * 1e-3)),
            break_even_grids=(extra["setup_s"] * 1e3 / (ms_plain - ms_step)) if ms_plain > ms_step else None,
            fallback_fraction=float(np.mean((Q["flags"] & _lib.PT_SURR_FALLBACK) != 0)),
            judged_fraction=float(np.mean((Q["flags"] & _lib.PT_SURR_MASKED) != 0)),
            seed_to_root_err_median=float(np.median(seed_err)), seed_to_root_err_p90=float(np.quantile(seed_err, 0.9)),
            node_passes=int(Q["passes"].sum()), plain_node_passes=int(P["passes"].sum()),
            plain_scan_kernel_ms=float(np.mean(scan_plain)), plain_launch=plain_cfg,
            scan_stage_ms=[round(float(v), 3) for v in state["stage_ms"]],
            gpu_ms_per_step=float(np.mean(gpu_ms)), gpu_ms_steps=[round(float(v), 2) for v in gpu_ms],
            scan_ms_steps=[round(float(v), 2) for v in scan_ms],
            plain_fp64_tflops=flops_plain / (np.mean(scan_plain) * 1e-3) / 1e12)
        if w["policy"] == "MULTIROOT3":
            eo = (Q["omega"] - P["omega"]) / np.abs(P["omega"])
            extra.update(
                plain_policy="MultiSeed", first_order_fraction=float(np.mean((Q["flags"] & _lib.PT_FIRST_ORDER) != 0)),
                omega_at_most_multiseed=float(np.mean(eo <= 1e-12)), multiseed_lower_omega=int(np.sum(eo > 1e-12)))
            for k in ("fallback_fraction", "judged_fraction"):
                extra.pop(k)
            seed_policy = "MULTIROOT3 (learned seed + HADRON + QUARK, min Omega)"
        else:
            seed_policy = "SurrogateSeed(judge=reference)"
        # implicit thermodynamic derivatives at every root of the step (one
        # pass per point after the solve; kernel time, outside the timed step)
        okr = Q["converged"] == 1
        ctx.thermo_derivatives_batch(Q["x"][okr], T[okr], mu[okr], xi[okr], w["p_num"], w["t_num"])
        extra.update(thermo_derivatives_kernel_ms=ctx.last_kernel_ms(),
                     thermo_derivatives_points_per_s=float(okr.sum()) / (ctx.last_kernel_ms() * 1e-3))
    else:
        ms_step, scan_ms, flops, solved = ms_plain, scan_plain, flops_plain, solved_plain
        Q = P
        seed_policy = {"DEFAULT": "DefaultSeed", "CONTINUITY": "ContinuitySeed (one chain)"}[w["plain_policy"]]

    scan_avg = float(np.mean(scan_ms))
    achieved = flops / (scan_avg * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": solved / (ms_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(workload=args.workload, points=n, solved=solved, p_num=w["p_num"], t_num=w["t_num"],
                       seed_policy=seed_policy, detail=w["desc"], launch=ctx.launch_cfg(),
                       mean_iterations=float(np.mean(Q["iterations"])), **extra),
        # dominant kernel: the persistent scan kernel (all its launches in one step)
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": C4_TRAFFIC_BYTES_PER_LAUNCH if args.workload == "c4_surrogate" else None,
                     "kernel": "pnjl::scan_kernel",
                     "kernel_ms_per_step": scan_avg, "flops_per_step": flops, "flops_per_node": C_NODE},
    }
    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        line["config"]["copy_probe_bytes"] = int(src.numel() * 4)
    if not args.no_cpu:
        line["cpu_baseline"] = dict(cpu_baseline(w, T, mu, xi), **host_cpu())
    if args.workload == "c4_surrogate" and not args.profile and not args.no_others:
        line["config"]["other_workloads"] = other_workloads()
    print(json.dumps(line), flush=True)


# the rest of the workloads, timed in the default run too (each a child
# process: 3 timed steps after 1 warm-up, no CPU leg); their figures go into
# the default line's config, so that one run records all of them.  The scan
# workloads take the launch geometry the autotuner picks for them (its search
# over the 1.3 M-point C5 grid alone takes about a minute).
OTHER_WORKLOADS = {"c5_multiroot": "8,1,16,0", "c3_polarization": None, "propagators": None, "meson_masses": None,
                   "dual_branch": None, "maxwell_line": None, "tmu_scan": None, "c1_continuity": "32,4,16,0"}


def other_workloads(timeout_s=150):
    import subprocess
    out = {}
    for name, launch in OTHER_WORKLOADS.items():
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", name, "--steps", "3", "--warmup", "1",
               "--no-cpu"] + (["--launch", launch] if launch else [])
        t0 = time.perf_counter()
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
            rows = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
            if p.returncode != 0 or not rows:
                out[name] = dict(error=f"exit {p.returncode}", stderr_tail=p.stderr[-300:])
                continue
            d = json.loads(rows[-1])
            r = d.get("roofline") or {}
            out[name] = dict(metric=d["metric"], value=d["value"], unit=d["unit"], ms_per_step=d["ms_per_step"],
                             steps=d["steps"], roofline_frac=r.get("frac"), kernel=r.get("kernel"),
                             kernel_ms_per_step=r.get("kernel_ms_per_step"), wall_s=time.perf_counter() - t0)
        except subprocess.TimeoutExpired:
            out[name] = dict(error=f"timeout after {timeout_s} s")
        print(f"bench: {name} {out[name].get('value', out[name].get('error'))}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
