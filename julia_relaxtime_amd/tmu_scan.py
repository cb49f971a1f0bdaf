"""T-mu scans on the GPU -- the reference's TmuScan.run_tmu_scan
(src/pnjl/scans/TmuScan.jl:100-510) with every temperature line of a xi
followed at once.

Per point the reference tries seed candidates in order -- the
PhaseAwareContinuitySeed tracker's seed (reset at each T), the previous
solution of the same (T, xi), then HADRON / QUARK (QUARK first above 150 MeV
or 300 MeV) -- each a fixed-seed PNJL.solve; the first candidate whose
result converged or has a residual <= 1e-4 wins, a near-converged result is
re-solved from its own solution ("refine") and then marked converged
("promote"); the tracker and the continuation cache take the winner.  Here
each mu step solves candidate k of all lines still open in one batched
`pnjl_solve_fixedmu_batch` (SEED_EXPLICIT: the fixed-seed solve with its
trust-region and MultiSeed fallbacks), then the refinements in one more.
Seeds, candidate order, success test, refine and promote are host logic on
the GPU results.  Rows are written in the reference's loop order (xi, T,
mu) and format (%.6f, `true`/`false`, quoted message); the message of a
candidate that raised names the failure without Julia's exception text.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from .pnjl import HADRON_SEED_5, QUARK_SEED_5, VERY_HIGH_TEMP_SEED_5, interpolate_mu_c, load_phase_boundary

HBARC = 197.327
DEFAULT_T_VALUES = [50.0 + 10.0 * i for i in range(16)]          # 50:10:200
DEFAULT_MU_VALUES = [10.0 * i for i in range(41)]                 # 0:10:400
ACCEPTABLE_RESIDUAL = 1e-4                                       # TmuScan.jl:60
HEADER = ["T_MeV", "mu_MeV", "xi", "pressure_fm4", "rho", "entropy_fm3", "energy_fm4", "phi_u", "phi_d", "phi_s",
          "Phi1", "Phi2", "M_u_MeV", "M_d_MeV", "M_s_MeV", "iterations", "residual_norm", "converged", "message"]


def _key(T, mu, xi):
    return (round(float(T), 6), round(float(mu), 6), round(float(xi), 6))


def _load_completed(path):
    """TmuScan.jl:236-258: keys of the rows already written."""
    done = set()
    with open(path) as f:
        next(f, None)
        for line in f:
            if not line.strip():
                continue
            c = line.split(",")
            if len(c) < 3:
                continue
            try:
                done.add(_key(float(c[0]), float(c[1]), float(c[2])))
            except ValueError:
                pass
    return done


def _fmt(x):
    return "%.6f" % x


def _gpu_solver(device, p_num, t_num, opts):
    from .pnjl import context
    ctx = context(device)

    def solve(T, mu, xi, seeds):
        return ctx.solve_batch(T, mu, xi, policy=_lib.SEED_EXPLICIT, seed=seeds, p_num=p_num, t_num=t_num,
                               opts=opts)
    return solve


_UNKNOWN, _HADRON, _QUARK, _CROSSOVER = 0, 1, 2, 3
_FIELDS = {"x": (5,), "masses": (3,), "pressure": (), "rho_norm": (), "entropy": (), "energy": (),
           "iterations": (), "residual_norm": (), "converged": ()}


def _line_mu_c(bd, T_MeV):
    """mu_c(T) of each temperature line (interpolate_mu_c, SeedStrategies.jl:446-475),
    NaN above the CEP or without boundary data."""
    if bd is None:
        return np.full(T_MeV.size, np.nan), np.ones(T_MeV.size, dtype=bool)
    cross = ~np.isnan(bd[2]) & (T_MeV > bd[2])
    return np.array([interpolate_mu_c(bd, float(t)) for t in T_MeV]), cross


def _phases(muc, cross, mu_MeV):
    """_get_current_phase (SeedStrategies.jl:762-782) of points of lines with
    the given mu_c and above-CEP flags."""
    ph = np.where(np.isnan(muc), _UNKNOWN, np.where(mu_MeV < muc, _HADRON, _QUARK))
    return np.where(cross, _CROSSOVER, ph)


def _default_seed(T_MeV, mu_MeV):
    """get_seed(DefaultSeed(:auto)) (SeedStrategies.jl:131-225)."""
    q = (T_MeV > 150) | (mu_MeV > 300)
    seeds = np.where(q[:, None], np.asarray(QUARK_SEED_5), np.asarray(HADRON_SEED_5))
    seeds[q & (T_MeV >= 300.0)] = VERY_HIGH_TEMP_SEED_5
    return seeds


def run_tmu_scan(T_values=DEFAULT_T_VALUES, mu_values=DEFAULT_MU_VALUES, xi_values=(0.0,), output_path=None,
                 overwrite=False, resume=True, use_phase_aware=True, p_num=24, t_num=8, opts=None, device=0,
                 solver=None):
    """run_tmu_scan (TmuScan.jl:100-225).  Returns dict(total, success,
    failure, skipped, output, tables) with tables[xi] a dict of arrays over
    (T line, mu index): the solution fields (x, masses, pressure, rho_norm,
    entropy, energy, iterations, residual_norm, converged), `solved` (False
    for failed and skipped points) and `message`.  `solver(T, mu, xi, seeds)`
    replaces the GPU batch solve.  The per-line trackers, continuation cache
    and candidate lists are kept as arrays over the temperature lines."""
    solve = solver or _gpu_solver(device, p_num, t_num, opts or _lib.default_opts())
    completed = set()
    if output_path is not None and resume and not overwrite and os.path.isfile(output_path):
        completed = _load_completed(output_path)
    stats = dict(total=0, success=0, failure=0, skipped=0)
    tables = {}
    H, Q = np.asarray(HADRON_SEED_5), np.asarray(QUARK_SEED_5)
    Ts = np.array([float(T) for T in T_values])
    mus = np.array([float(m) for m in mu_values])
    L, n = Ts.size, mus.size
    for xi in xi_values:
        xi = float(xi)
        bd = load_phase_boundary(xi)[2] if use_phase_aware else None
        muc, cross = _line_mu_c(bd, Ts)
        tab = {k: np.full((L, n) + sh, np.nan) for k, sh in _FIELDS.items()}
        tab["solved"] = np.zeros((L, n), dtype=bool)
        tab["open"] = np.zeros((L, n), dtype=bool)
        tab["message"] = np.full((L, n), "", dtype=object)
        prev = np.full((L, 5), np.nan)          # tracker.previous_solution (= the continuation cache)
        prev_phase = np.full(L, _UNKNOWN)
        have = np.zeros(L, dtype=bool)
        for k, mu in enumerate(mus):
            stats["total"] += L
            if completed:
                open_ = np.array([_key(T, mu, xi) not in completed for T in Ts], dtype=bool)
            else:
                open_ = np.ones(L, dtype=bool)
            stats["skipped"] += int((~open_).sum())
            lines = np.flatnonzero(open_)
            if lines.size == 0:
                continue
            tab["open"][lines, k] = True
            T_MeV = Ts[lines]
            T_fm, mu_fm = T_MeV / HBARC, np.full(lines.size, mu / HBARC)
            # candidate lists  _build_seed_candidates_v2  TmuScan.jl:261-292
            cand = []
            if use_phase_aware:
                cur = _phases(muc[lines], cross[lines], mu)
                h, pp = have[lines], prev_phase[lines]
                trans = ((pp == _HADRON) & (cur == _QUARK)) | ((pp == _QUARK) & (cur == _HADRON))
                s1 = np.where(h[:, None], prev[lines], _default_seed(T_MeV, np.full(lines.size, mu)))
                s1[(~h | trans) & (cur == _HADRON)] = H
                s1[(~h | trans) & (cur == _QUARK)] = Q
                cand.append((np.ones(lines.size, bool), s1,
                             np.where(h, "phase_aware_continuity", "phase_aware_default")))
            cand.append((have[lines], prev[lines], np.full(lines.size, "continuation")))
            qfirst = (T_MeV > 150) | (mu > 300)
            cand.append((np.ones(lines.size, bool), np.where(qfirst[:, None], Q, H), np.where(qfirst, "quark", "hadron")))
            cand.append((np.ones(lines.size, bool), np.where(qfirst[:, None], H, Q), np.where(qfirst, "hadron", "quark")))
            msgs = {}
            won = np.zeros(lines.size, dtype=bool)

            def take(r, idx, sel):
                rows = lines[idx[sel]]
                for f in _FIELDS:
                    tab[f][rows, k] = r[f][sel]
            for valid, seeds, lab in cand:
                idx = np.flatnonzero(~won & valid)
                if idx.size == 0:
                    continue
                r = solve(T_fm[idx], mu_fm[idx], np.full(idx.size, xi), np.ascontiguousarray(seeds[idx]))
                raised = (r["flags"] & (_lib.PT_EXCEPTION | _lib.PT_NO_ROOT)) != 0
                res = r["residual_norm"]
                ok = ~raised & ((r["converged"] == 1) | (np.isfinite(res) & (res <= ACCEPTABLE_RESIDUAL)))
                take(r, idx, ok)
                won[idx[ok]] = True
                for j in np.flatnonzero(~ok):
                    i = idx[j]
                    msgs.setdefault(i, []).append(
                        f"seed[{lab[i]}] failed: solver raised (non-finite residual or no root)" if raised[j] else
                        f"seed[{lab[i]}] failed (iterations={int(r['iterations'][j])}, residual={_fmt(res[j])}, "
                        "converged=false)")
            # refine near-converged winners from their own solution  TmuScan.jl:386-400
            near = np.flatnonzero(won & (tab["converged"][lines, k] != 1))
            if near.size:
                r = solve(T_fm[near], mu_fm[near], np.full(near.size, xi), tab["x"][lines[near], k])
                ok = ((r["flags"] & (_lib.PT_EXCEPTION | _lib.PT_NO_ROOT)) == 0) & (r["converged"] == 1)
                take(r, near, ok)
                for i in near[ok]:
                    msgs.setdefault(i, []).append("refined from near-converged seed")
                # _promote_success  TmuScan.jl:420-458
                for i in near[~ok]:
                    msgs.setdefault(i, []).append(
                        "force-marked converged (residual " + _fmt(tab["residual_norm"][lines[i], k]) + ")")
                tab["converged"][lines[near], k] = 1
            wl = lines[won]
            tab["solved"][wl, k] = True
            stats["success"] += int(won.sum())
            stats["failure"] += int((~won).sum())
            prev[wl] = tab["x"][wl, k]
            have[wl] = True
            if use_phase_aware:                                   # update!(tracker, sol, T, mu)
                prev_phase[wl] = _phases(muc[wl], cross[wl], mu)
            for i, m in msgs.items():
                tab["message"][lines[i], k] = " | ".join(m)
        tables[xi] = tab
    if output_path is not None:
        d = os.path.dirname(output_path)
        if d:
            os.makedirs(d, exist_ok=True)
        new = overwrite or not os.path.isfile(output_path)
        with open(output_path, "w" if new else "a") as f:
            if new:
                f.write(",".join(HEADER) + "\n")
            for xi, tab in tables.items():
                for l in range(L):
                    for k in range(n):
                        if tab["open"][l, k]:
                            f.write(format_row(Ts[l], mus[k], xi, tab, l, k) + "\n")
    return dict(**stats, output=output_path, tables=tables)


def format_row(T, mu, xi, tab, l, k):
    """_write_row  TmuScan.jl:461-504."""
    message = tab["message"][l, k]
    q = f'"{message}"' if message else ""
    if not tab["solved"][l, k]:
        return ",".join([_fmt(T), _fmt(mu), _fmt(xi)] + ["NaN"] * 12 + ["-1", "NaN", "false", q])
    x, m = tab["x"][l, k], tab["masses"][l, k] * HBARC
    vals = [T, mu, xi, tab["pressure"][l, k], tab["rho_norm"][l, k], tab["entropy"][l, k], tab["energy"][l, k],
            x[0], x[1], x[2], x[3], x[4], m[0], m[1], m[2]]
    return ",".join([_fmt(v) for v in vals] + [str(int(tab["iterations"][l, k])), _fmt(tab["residual_norm"][l, k]),
                                                "true" if tab["converged"][l, k] == 1 else "false", q])


__all__ = ["run_tmu_scan", "format_row", "HEADER", "DEFAULT_T_VALUES", "DEFAULT_MU_VALUES"]
