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


def propagator_queries(ctx, pts, n=1 << 20, seed=20261203):
    """Propagator workload: n (process, channel) queries drawn uniformly from
    the 34 channels of the process table, each at one of the Pi workload's
    equilibrium points (pts from pi_grid_points) with A_aniso(m_u), A_aniso(m_s)
    of that point (GL 64 on [0, 10] fm^-1 x GL 4 in cos) and (k0, k) drawn
    from the Pi grid's axes.  Returns a PROP_QUERY_DTYPE array."""
    from . import _lib
    from . import polarization as PZ
    from . import propagators as PR
    xp, wp = _lib.gauleg(0.0, 10.0, 64)
    xc, wc = np.polynomial.legendre.leggauss(4)
    args = np.concatenate([np.stack([pts[:, 3], pts[:, 1], pts[:, 0], pts[:, 5], pts[:, 6], pts[:, 2]], 1),
                           np.stack([pts[:, 4], pts[:, 1], pts[:, 0], pts[:, 5], pts[:, 6], pts[:, 2]], 1)])
    A = PZ.A_batch(args, xp, wp, xc, wc).reshape(2, len(pts))
    pc = [(PR.PROCESS_ID[p], PR.CHANNEL_ID[c]) for p in PR.PROCESSES for c in PR.process_channels(p)]
    rng = np.random.default_rng(seed)
    sel, pt = rng.integers(0, len(pc), n), rng.integers(0, len(pts), n)
    q = np.zeros(n, dtype=PR.PROP_QUERY_DTYPE)
    q["process"] = np.array([c[0] for c in pc])[sel]
    q["channel"] = np.array([c[1] for c in pc])[sel]
    q["k0"], q["k"] = PI_K0[rng.integers(0, PI_K0.size, n)], PI_K[rng.integers(0, PI_K.size, n)]
    P = pts[pt]
    q["T"], q["mu_u"], q["mu_s"], q["xi"] = P[:, 0], P[:, 1], P[:, 1], P[:, 2]
    q["m_u"], q["m_s"], q["Phi"], q["Phib"] = P[:, 3], P[:, 4], P[:, 5], P[:, 6]
    q["A_u"], q["A_s"] = A[0, pt], A[1, pt]
    return q


def meson_mass_points(ctx, opts, nT=256, nmu=256):
    """Equilibrium points of the pole-mass workload: the xi = 0 slice of the
    256 x 256 (T 20-300 MeV, mu 0-450 MeV) scan, DefaultSeed p64 t4, with
    A(m_u), A(m_s) on GL(0, 20, 16) (ensure_quark_params_has_A at xi = 0).
    Returns (n, 9) rows (T, mu, m_u, m_s, Phi, Phib, A_u, A_s, converged)."""
    from . import _lib
    from . import polarization as PZ
    T, mu, xi = grid(nT, nmu, [0.0])
    r = ctx.solve_batch(T, mu, xi, policy=_lib.SEED_DEFAULT, p_num=64, t_num=4, opts=opts)
    m_u, m_s, Phi, Phib = r["masses"][:, 0], r["masses"][:, 2], r["x"][:, 3], r["x"][:, 4]
    xp, wp = _lib.gauleg(0.0, 20.0, 16)
    args = np.concatenate([np.stack([m_u, mu, T, Phi, Phib], 1), np.stack([m_s, mu, T, Phi, Phib], 1)])
    A = PZ.A_batch(args, xp, wp).reshape(2, T.size)
    return np.stack([T, mu, m_u, m_s, Phi, Phib, A[0], A[1], r["converged"].astype(float)], 1)


def meson_mass_queries(pts):
    """All eight mesons at every converged point (k = 0, default guess,
    NLsolve defaults): a MASS_QUERY_DTYPE array, point-major."""
    from . import propagators as PR
    P = pts[pts[:, 8] == 1.0]
    n = len(P) * len(PR.MASS_MESONS)
    rep = lambda c: np.repeat(P[:, c], len(PR.MASS_MESONS))
    return PR.make_mass_queries(np.tile(np.arange(len(PR.MASS_MESONS)), len(P)), rep(0), rep(4), rep(5), 0.0,
                                rep(2), rep(3), rep(1), rep(1), rep(6), rep(7))
