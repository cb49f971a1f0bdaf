"""FLOP probe of the polarization grid (run under rocprofv3 FP64 counters).

The grid kernel computes one symmetrised B0 (+ B0_correction) per (point,
pair, k0, k) and from it Pi in both channels.  Its waves are divergent (the
lanes of a wave take different singularity branches), so its own hardware
counts include inactive lanes.  This probe evaluates a seeded sample of the
same (point, pair, k0, k) cells as single channel-P queries with
PNJL_FLAG_ONE_PER_WAVE: each wave then carries exactly one query, and
(2 FMA + ADD + MUL) x 1 lane per wave-instruction is that query's own FP64
work.  Per grid value (two channels share one B0) the algorithmic cost is
half a query's.

    python3 -m julia_relaxtime_amd.pi_flop_probe [n_sample]
"""
from __future__ import annotations

import json
import sys

import numpy as np

from . import _lib
from . import polarization as PZ
from .pnjl import context
from .workloads import PI_K, PI_K0, pi_grid_points


def main(n_sample=131072, seed=7):
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
    print(json.dumps({"queries": int(n_sample), "seed": seed, "cells": int(len(pts) * 3 * PI_K0.size * PI_K.size),
                      "kernel_ms": ctx.last_kernel_ms()}), flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 131072)
