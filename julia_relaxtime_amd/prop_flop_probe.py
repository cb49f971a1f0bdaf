"""FLOP probe of the propagator batch (run under rocprofv3 FP64 counters).

prop_pi_kernel evaluates, for each query, the symmetrised B0 (+ correction)
of every quark pair its exchanged mesons need and Pi in the needed channels.
Its waves are divergent (queries of different processes need different
pairs; lanes take different singularity branches), so its own hardware
counts include inactive lanes.  This probe runs a seeded sample of the
propagator workload's queries with PNJL_FLAG_ONE_PER_WAVE: one (pair, query)
per wavefront, so (2 FMA + ADD + MUL) x 1 lane per wave-instruction is the
sample's own FP64 work; divided by the sample size it is the algorithmic
cost per query.

    python3 -m julia_relaxtime_amd.prop_flop_probe [n_sample]
"""
from __future__ import annotations

import json
import sys

from . import _lib
from . import propagators as PR
from .pnjl import context
from .workloads import pi_grid_points, propagator_queries


def main(n_sample=65536):
    ctx = context(0)
    pts = pi_grid_points(ctx, _lib.default_opts(ftol=1e-10, max_iter=40))
    q = propagator_queries(ctx, pts, n=1 << 20)
    q = q[:n_sample]
    PR.propagator_by_channel_batch(q, one_per_wave=True)
    print(json.dumps({"queries": int(n_sample), "workload_queries": 1 << 20, "seed": 20261203,
                      "kernel_ms": ctx.last_kernel_ms()}), flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
