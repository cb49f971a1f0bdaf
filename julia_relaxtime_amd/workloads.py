"""Synthetic scan workloads (grids are linspace with endpoints, MeV -> fm^-1
with hbar c = 197.327 MeV fm) shared by the benchmark and the FLOP probes."""
from __future__ import annotations

import numpy as np

HB = 197.327


def grid(nT, nmu, xis, T_range=(20, 300), mu_range=(0, 450)):
    """(T, mu, xi) of a T x mu x xi grid in fm^-1, T outermost."""
    T, mu, xi = np.meshgrid(np.linspace(*T_range, nT), np.linspace(*mu_range, nmu), np.asarray(xis, float),
                            indexing="ij")
    return T.ravel() / HB, mu.ravel() / HB, xi.ravel()


def scan_grid_32x32():
    """32 x 32 at xi = 0: T 50-300 MeV, mu 0-400 MeV (p = 64, t = 8)."""
    return grid(32, 32, [0.0], (50, 300), (0, 400))


def scan_grid_64x64x9():
    """64 x 64 x 9: T 20-300 MeV, mu 0-450 MeV, xi -0.4..0.4 step 0.1 (p = 64, t = 4)."""
    return grid(64, 64, np.round(np.arange(-0.4, 0.41, 0.1), 10))


PI_K0 = np.linspace(0, 4, 128)
PI_K = np.linspace(0, 3, 64)


def pi_grid_points(ctx, opts, n=256, seed=20261202):
    """The equilibrium points of the Pi grid workload: n points drawn without
    replacement (default_rng(seed)) from the converged roots of the 32 x 32
    and 64 x 64 x 9 scans.  Returns an (n, 7) array of (T, mu, xi, M_u, M_s,
    Phi, Phib) in fm^-1."""
    from . import _lib
    rows = []
    for (T, mu, xi), t_num in ((scan_grid_32x32(), 8), (scan_grid_64x64x9(), 4)):
        r = ctx.solve_batch(T, mu, xi, policy=_lib.SEED_DEFAULT, p_num=64, t_num=t_num, opts=opts)
        ok = r["converged"] == 1
        rows.append(np.stack([T, mu, xi, r["masses"][:, 0], r["masses"][:, 2], r["x"][:, 3], r["x"][:, 4]], 1)[ok])
    allp = np.concatenate(rows)
    idx = np.random.default_rng(seed).choice(len(allp), n, replace=False)
    return allp[idx]
