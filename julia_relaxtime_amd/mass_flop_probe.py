"""FLOP probe of the pole-mass batch (run under rocprofv3 FP64 counters).

mass_kernel runs one query's trust-region iteration per thread; queries
differ in iteration count and quadrature branches, so the kernel's own
hardware counts include the inactive lanes of divergent waves.  This probe
runs a seeded sample of the pole-mass workload's queries with
PNJL_FLAG_ONE_PER_WAVE (one query per wavefront): (2 FMA + ADD + MUL) x 1
lane per wave-instruction, divided by the sample size, is the algorithmic
cost per query.

    python3 -m julia_relaxtime_amd.mass_flop_probe [n_sample]
"""
from __future__ import annotations

import json
import sys

import numpy as np

from . import _lib
from . import propagators as PR
from .pnjl import context
from .workloads import meson_mass_points, meson_mass_queries


def main(n_sample=8192, seed=11):
    ctx = context(0)
    q = meson_mass_queries(meson_mass_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40)))
    sel = np.sort(np.random.default_rng(seed).choice(q.size, n_sample, replace=False))
    r = PR.meson_mass_batch(q[sel], one_per_wave=True)
    print(json.dumps({"queries": int(n_sample), "workload_queries": int(q.size), "seed": seed,
                      "iterations": int(r["iterations"].sum()), "kernel_ms": ctx.last_kernel_ms()}), flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
