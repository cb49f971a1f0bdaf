"""The gap part of the reference's scan driver
(scripts/relaxtime/run_gap_transport_scan.jl `run_scan` :362-550) as one
batched GPU call per chain length, writing the driver's scan_csv_v1 file.

The reference loops xi -> mu_B -> T; along each (xi, mu_B) line it solves
the first point with MultiSeed and the next ones with a
PhaseAwareContinuitySeed fed by the previous converged solution, then writes
one row per point (p_num 12, t_num 6, iterations 40 by default).  Here every
line of the grid is one chain of PNJL_SEED_PHASE_AWARE in a single
`pnjl_solve_fixedmu_batch` (chains run in parallel on the GPU), and the rows
are written in the reference's loop order, columns and number formats
(Julia's `string`).  The relaxation-time and transport columns (tau_*,
tauinv_*, eta, sigma, zeta, eta_over_s, zeta_over_s) belong to the
transport workflow, which this path does not compute: they are written as
NaN so that the file keeps the driver's schema.  Resume follows the
driver: rows whose (T_MeV, muB_MeV, xi) key is already in the file are
skipped, and a line's chain continues from the last point it solves.
"""
from __future__ import annotations

import os
from fractions import Fraction

import numpy as np

from . import _lib
from .scan_csv import julia_bool, julia_string, read_existing_keys, write_header, write_metadata

HBARC = 197.327
RHO0 = 0.16                      # rho0 fm^-3 (src/Constants_PNJL.jl)
DRIVER_COLUMNS = [
    "T_MeV", "muq_MeV", "muB_MeV", "xi", "T_fm", "muq_fm", "converged", "iterations", "residual_norm", "Phi",
    "Phibar", "m_u", "m_d", "m_s", "rho_baryon", "rho_norm", "omega_fm4inv", "P_fm4inv", "epsilon_fm4inv",
    "s_fm3inv", "omega_MeV_fm3", "P_MeV_fm3", "epsilon_MeV_fm3", "eps_minus_3P_over_T4", "n_u", "n_d", "n_s",
    "n_ubar", "n_dbar", "n_sbar", "tau_u", "tau_d", "tau_s", "tau_ubar", "tau_dbar", "tau_sbar", "tauinv_u",
    "tauinv_d", "tauinv_s", "tauinv_ubar", "tauinv_dbar", "tauinv_sbar", "eta", "sigma", "zeta", "eta_over_s",
    "zeta_over_s"]                                  # write_header_if_needed  run_gap_transport_scan.jl:266-283
TRANSPORT_COLUMNS = DRIVER_COLUMNS[30:]
KEY_COLUMNS = ["T_MeV", "muB_MeV", "xi"]           # read_existing_keys  run_gap_transport_scan.jl:244-246


def julia_range(start, stop, step):
    """collect(range(start; stop, step)) for decimal inputs: start + i*step
    rounded once from the exact decimal values, as Julia's ranges do."""
    a, s = Fraction(str(float(start))), Fraction(str(float(step)))
    n = int((Fraction(str(float(stop))) - a) / s)
    return [float(a + i * s) for i in range(n + 1)]


def _row(T_mev, muB_mev, xi, r, j):
    """One driver row (run_gap_transport_scan.jl:494-521) from solve result j."""
    T_fm = T_mev / HBARC
    muq_mev = muB_mev / 3.0
    muq_fm = muq_mev / HBARC
    nq, nqb = r["n_q"][j], r["n_qbar"][j]
    rho_baryon = ((nq[0] - nqb[0]) + (nq[1] - nqb[1]) + (nq[2] - nqb[2])) / 3.0
    om, P, eps, s = r["omega"][j], r["pressure"][j], r["energy"][j], r["entropy"][j]
    fin = np.isfinite(eps) and np.isfinite(P) and np.isfinite(T_fm) and T_fm != 0.0
    e3p = (eps - 3.0 * P) / T_fm ** 4 if fin else float("nan")
    m, x = r["masses"][j], r["x"][j]
    vals = [T_mev, muq_mev, muB_mev, xi, T_fm, muq_fm]
    out = [julia_string(v) for v in vals]
    out += [julia_bool(r["converged"][j] == 1), str(int(r["iterations"][j])), julia_string(r["residual_norm"][j])]
    out += [julia_string(v) for v in (x[3], x[4], m[0], m[1], m[2], rho_baryon, rho_baryon / RHO0, om, P, eps, s,
                                      om * HBARC, P * HBARC, eps * HBARC, e3p, nq[0], nq[1], nq[2], nqb[0], nqb[1],
                                      nqb[2])]
    out += ["NaN"] * len(TRANSPORT_COLUMNS)
    return ",".join(out)


def run_gap_scan(output, xi_values=(0.0,), tmin=50.0, tmax=200.0, tstep=10.0, mubmin=0.0, mubmax=1200.0,
                 mubstep=60.0, overwrite=False, resume=True, p_num=12, t_num=6, max_iter=40, device=0,
                 metadata=None, solver=None):
    """The driver's scan (defaults of run_gap_transport_scan.jl:99-125) with
    the gap solves batched on the GPU.  Returns the number of rows written.
    `solver(T, mu, xi, chain_len)` replaces the GPU batch solve (tests drive
    the same logic with other solvers)."""
    if solver is None:
        from .pnjl import context
        ctx = context(device)

        def solver(T, mu, xi, L):
            return ctx.solve_batch(T, mu, xi, policy=_lib.SEED_PHASE_AWARE, p_num=p_num, t_num=t_num,
                                   opts=_lib.default_opts(max_iter=max_iter, chain_len=L))
    existing = read_existing_keys(output, KEY_COLUMNS) if (resume and os.path.isfile(output) and not overwrite) \
        else set()
    if overwrite and os.path.isfile(output):
        os.remove(output)
    d = os.path.dirname(output)
    if d:
        os.makedirs(d, exist_ok=True)
    new_file = not os.path.isfile(output) or os.path.getsize(output) == 0
    T_values = julia_range(tmin, tmax, tstep)
    muB_values = sorted(set(julia_range(mubmin, mubmax, mubstep)))
    # the chains: one per (xi, mu_B), over the T values not already written
    chains = []
    for xi in xi_values:
        for muB in muB_values:
            Ts = [T for T in T_values if not (resume and (T, muB, float(xi)) in existing)]
            if Ts:
                chains.append((float(xi), muB, Ts))
    results = {}
    for L in sorted(set(len(c[2]) for c in chains)):
        sel = [c for c in chains if len(c[2]) == L]
        T = np.array([t / HBARC for c in sel for t in c[2]])
        mu = np.array([c[1] / 3.0 / HBARC for c in sel for _ in c[2]])
        xi = np.array([c[0] for c in sel for _ in c[2]])
        r = solver(T, mu, xi, L)
        for k, c in enumerate(sel):
            results[(c[0], c[1])] = (r, k * L)
    written = 0
    with open(output, "a") as f:
        if new_file:
            write_metadata(f, metadata or {"schema": "scan_csv_v1", "title": "gap_transport_scan",
                                           "script": "julia_relaxtime_amd.gap_scan", "p_num": str(p_num),
                                           "t_num": str(t_num), "max_iter": str(max_iter),
                                           "transport": "not computed (NaN columns)"})
            write_header(f, DRIVER_COLUMNS)
        for xi, muB, Ts in chains:
            r, base = results[(xi, muB)]
            for i, T in enumerate(Ts):
                f.write(_row(T, muB, xi, r, base + i) + "\n")
                written += 1
    return written


__all__ = ["run_gap_scan", "julia_range", "DRIVER_COLUMNS", "TRANSPORT_COLUMNS", "KEY_COLUMNS"]
