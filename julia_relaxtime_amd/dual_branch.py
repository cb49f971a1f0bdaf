"""Dual-branch scans on the GPU -- the API of the reference's DualBranchScan
(src/pnjl/scans/DualBranchScan.jl): run_dual_branch_scan,
find_phase_transition, merge_branches.

For each (T, xi) line over a mu grid, the hadron branch is followed from
low to high mu (ContinuitySeed with a DefaultSeed(:hadron) fallback) and the
quark branch from high to low mu (DefaultSeed(:quark) fallback), each
stopping at the first point that does not converge, raises (non-finite
residual, MultiSeed without a root) or jumps (|d phi_u| > 0.5 or
|d M_u| > 50 MeV); the physical branch is the one of lower Omega
(DualBranchScan.jl:104-205, :330-445).  `run_dual_branch_scan_batch` follows
every line of a batch at once.  On the GPU each branch of all lines is ONE
launch of branch chains (PNJL_SEED_BRANCH_HADRON / _QUARK): every lane group
follows one line's mu grid point after point -- the first point from
DefaultSeed(:hadron / :quark), the next ones from the previous solution --
and stops at the branch end (failure, exception or jump, tested in the
kernel), so a dual-branch scan of any number of lines is two launches.  With
a `solver` argument (the oracle in the tests) the same branches are followed
one batched solve per mu step instead, the hadron step k and the quark step
n-1-k in one call.  Branch selection, merge and the Omega crossing are host
logic on the results.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib

HBARC = 197.327
_FIELDS = ("x", "omega", "pressure", "rho_norm", "entropy", "energy", "masses", "iterations", "residual_norm")


def gpu_solver(device=0, p_num=24, t_num=8, opts=None, lanes_per_point=32):
    """The batched GPU solve of one mu step.  The batches here are one point
    per line (hundreds, not millions), so each point gets a wide group of
    lanes: a Newton pass then takes a few hundred instructions per lane and
    the chains that run to the iteration limit at a branch end finish
    sooner (32 lanes: 2.6x faster than 4 over 655 lines x 226 mu).  The
    context's own launch geometry is restored after each call."""
    from .pnjl import context
    ctx = context(device)
    o = opts or _lib.default_opts()

    def solve(T, mu, xi, policy, seed):
        saved = ctx.launch_cfg()
        ctx.set_launch_cfg(lanes_per_point=lanes_per_point, chunk=1, waves_per_cu=8,
                           tail_lanes_per_point=lanes_per_point)
        try:
            return ctx.solve_batch(T, mu, xi, policy=policy, seed=seed, p_num=p_num, t_num=t_num, opts=o)
        finally:
            ctx.set_launch_cfg(**saved)
    return solve


def gpu_chain_solver(device=0, p_num=24, t_num=8, opts=None, lanes_per_point=32):
    """Whole branches on the GPU: `solve(T, mu, xi, policy, chain_len)` runs
    every chain of a PNJL_SEED_BRANCH_HADRON / _QUARK batch in one
    persistent-kernel launch (each lane group follows its chain point after
    point and stops at the branch end), with the context's launch geometry
    restored after the call."""
    from .pnjl import context
    ctx = context(device)
    base = opts or _lib.default_opts()

    def solve(T, mu, xi, policy, chain_len):
        o = _lib.default_opts()
        for f, _ in _lib.SolveOpts._fields_:
            setattr(o, f, getattr(base, f))
        o.chain_len = int(chain_len)
        saved = ctx.launch_cfg()
        ctx.set_launch_cfg(lanes_per_point=lanes_per_point, chunk=1, waves_per_cu=8,
                           tail_lanes_per_point=lanes_per_point)
        try:
            return ctx.solve_batch(T, mu, xi, policy=policy, p_num=p_num, t_num=t_num, opts=o)
        finally:
            ctx.set_launch_cfg(**saved)
    return solve


def _follow_chains(solve_chains, T, xi, mu_fm):
    """Both branches of every line as branch chains: one launch per branch
    (the hadron chains over mu ascending, the quark chains over mu
    descending).  Same result as _follow_both: a point is on the branch
    when it converged, did not raise and did not jump, up to the branch end."""
    L, n = T.size, mu_fm.size
    out = []
    for policy, order in ((_lib.SEED_BRANCH_HADRON, np.arange(n)), (_lib.SEED_BRANCH_QUARK, np.arange(n)[::-1])):
        r = solve_chains(np.repeat(T, n), np.tile(mu_fm[order], L), np.repeat(xi, n), policy, n)
        br = {}
        for f in _FIELDS:
            v = r[f].reshape((L, n) + r[f].shape[1:])
            br[f] = np.empty_like(v, dtype=np.float64)
            br[f][:, order] = v
        fl = r["flags"].reshape(L, n)
        ok = ((r["converged"].reshape(L, n) == 1) &
              ((fl & (_lib.PT_EXCEPTION | _lib.PT_NO_ROOT | _lib.PT_BRANCH_END | _lib.PT_SKIPPED)) == 0))
        valid = np.zeros((L, n), dtype=bool)
        valid[:, order] = ok
        for f in _FIELDS:
            mask = valid.reshape(valid.shape + (1,) * (br[f].ndim - 2))
            br[f] = np.where(mask, br[f], np.nan)
        br["valid"] = valid
        out.append(br)
    return out[0], out[1]


def _follow_both(solve, T, xi, mu_fm):
    """Both branches of every line: the hadron branch over mu ascending and the
    quark branch over mu descending.  Step k of the two is independent (mu
    index k and n-1-k), so after the first points (HADRON / QUARK policies)
    each step is ONE batched solve of the active lines of both branches.
    Returns two dicts of arrays (L, n, ...) with their `valid` masks."""
    L, n = T.size, mu_fm.size
    shape = lambda k: (L, n, 5) if k == "x" else (L, n, 3) if k == "masses" else (L, n)
    brs = [{k: np.full(shape(k), np.nan) for k in _FIELDS} for _ in range(2)]
    valid = [np.zeros((L, n), dtype=bool) for _ in range(2)]
    active = [np.ones(L, dtype=bool) for _ in range(2)]
    prev_x = [np.zeros((L, 5)) for _ in range(2)]
    prev_m = [np.zeros(L) for _ in range(2)]
    for k in range(n):
        cols = (k, n - 1 - k)
        idx = [np.flatnonzero(a) for a in active]
        if idx[0].size + idx[1].size == 0:
            break
        if k == 0:
            rs = [solve(T[idx[b]], np.full(idx[b].size, mu_fm[cols[b]]), xi[idx[b]], pol, None) if idx[b].size
                  else None for b, pol in ((0, _lib.SEED_HADRON), (1, _lib.SEED_QUARK))]
        else:
            r = solve(np.concatenate([T[idx[0]], T[idx[1]]]),
                      np.concatenate([np.full(idx[0].size, mu_fm[cols[0]]), np.full(idx[1].size, mu_fm[cols[1]])]),
                      np.concatenate([xi[idx[0]], xi[idx[1]]]), _lib.SEED_EXPLICIT,
                      np.concatenate([prev_x[0][idx[0]], prev_x[1][idx[1]]]))
            cut = idx[0].size
            rs = [{f: v[:cut] for f, v in r.items()}, {f: v[cut:] for f, v in r.items()}]
        for b in range(2):
            if idx[b].size == 0:
                continue
            r, i = rs[b], cols[b]
            ok = (r["converged"] == 1) & ((r["flags"] & (_lib.PT_EXCEPTION | _lib.PT_NO_ROOT)) == 0)
            if k > 0:   # _is_solution_jump  DualBranchScan.jl:410-415
                jump = (np.abs(prev_x[b][idx[b], 0] - r["x"][:, 0]) > 0.5) | \
                    (np.abs(prev_m[b][idx[b]] - r["masses"][:, 0]) * HBARC > 50.0)
                ok &= ~jump
            keep = idx[b][ok]
            for f in _FIELDS:
                brs[b][f][keep, i] = r[f][ok]
            valid[b][keep, i] = True
            prev_x[b][keep] = r["x"][ok]
            prev_m[b][keep] = r["masses"][ok, 0]
            active[b][idx[b][~ok]] = False
    for b in range(2):
        brs[b]["valid"] = valid[b]
    return brs[0], brs[1]


def _same_solution(h, q, tol=0.01):
    """_is_same_solution  DualBranchScan.jl:400-405 (its mass test compares fm^-1
    values against tol * 197.327, kept as written)."""
    return ((np.abs(h["x"][..., 0] - q["x"][..., 0]) < tol) & (np.abs(h["x"][..., 3] - q["x"][..., 3]) < tol)
            & (np.abs(h["masses"][..., 0] - q["masses"][..., 0]) < tol * HBARC))


def run_dual_branch_scan_batch(T_mev, mu_values, xi=0.0, p_num=24, t_num=8, opts=None, device=0, solver=None,
                               chain_solver=None):
    """run_dual_branch_scan (DualBranchScan.jl:104-169) for every (T_mev[l],
    xi[l]) line over the same mu grid (MeV).  Returns a list of result dicts
    (T_mev, xi, mu_values, hadron, quark, physical), each branch a dict of
    per-mu arrays with a `valid` mask.  `solver(T, mu, xi, policy, seed)`
    replaces the GPU batch solve (tests drive the same logic with other
    solvers) and follows the branches one mu step per call; without it each
    branch of all lines is one launch of branch chains (`chain_solver(T, mu,
    xi, policy, chain_len)`, by default `gpu_chain_solver`)."""
    T_mev = np.atleast_1d(np.asarray(T_mev, dtype=np.float64))
    xi = np.broadcast_to(np.asarray(xi, dtype=np.float64), T_mev.shape).copy()
    mu_values = np.asarray(list(mu_values), dtype=np.float64)
    T = T_mev / HBARC
    mu_fm = mu_values / HBARC
    if solver is not None:
        had, qrk = _follow_both(solver, T, xi, mu_fm)
    else:
        had, qrk = _follow_chains(chain_solver or gpu_chain_solver(device, p_num, t_num, opts), T, xi, mu_fm)
    # _select_physical_branch  DualBranchScan.jl:378-398
    both = had["valid"] & qrk["valid"]
    same = both & _same_solution(had, qrk)
    take_q = (~had["valid"] & qrk["valid"]) | (both & ~same & (had["omega"] > qrk["omega"]))
    phys = {}
    for k in _FIELDS:
        sel = take_q.reshape(take_q.shape + (1,) * (had[k].ndim - 2))
        phys[k] = np.where(sel, qrk[k], had[k])
    phys["valid"] = had["valid"] | qrk["valid"]
    out = []
    for l in range(T_mev.size):
        pick = lambda b: {k: v[l] for k, v in b.items()}
        out.append(dict(T_mev=float(T_mev[l]), xi=float(xi[l]), mu_values=mu_values.copy(), hadron=pick(had),
                        quark=pick(qrk), physical=pick(phys), same=same[l].copy()))
    return out


def run_dual_branch_scan(T_mev, mu_range, xi=0.0, p_num=24, t_num=8, opts=None, device=0):
    """DualBranchScan.jl:104-169 for one line."""
    return run_dual_branch_scan_batch([T_mev], mu_range, xi, p_num, t_num, opts, device)[0]


def _omega_crossing(mu, oh, oq):
    """_find_omega_crossing  DualBranchScan.jl:447-485 (linear interpolation)."""
    delta = oh - oq
    sig = [i for i in range(mu.size) if abs(delta[i]) > max(1e-6 * (0.5 * (abs(oh[i]) + abs(oq[i]))), 1e-10)]
    if len(sig) < 2:
        return math.nan, math.nan
    for a, b in zip(sig[:-1], sig[1:]):
        if delta[a] * delta[b] < 0:
            t = delta[a] / (delta[a] - delta[b])
            return mu[a] + t * (mu[b] - mu[a]), oh[a] + t * (oh[b] - oh[a])
    return math.nan, math.nan


def find_phase_transition(result):
    """DualBranchScan.jl:178-228: dict(found, mu_c, omega_at_transition,
    mu_hadron_spinodal, mu_quark_spinodal, coexistence_region)."""
    mu = result["mu_values"]
    h, q = result["hadron"], result["quark"]
    hv, qv = np.flatnonzero(h["valid"]), np.flatnonzero(q["valid"])
    mu_h = float(mu[hv[-1]]) if hv.size else math.nan       # last point of the forward branch
    mu_q = float(mu[qv[0]]) if qv.size else math.nan        # first point of the backward branch
    co = np.flatnonzero(h["valid"] & q["valid"] & ~result["same"])
    if co.size == 0:
        return dict(found=False, mu_c=math.nan, omega_at_transition=math.nan, mu_hadron_spinodal=mu_h,
                    mu_quark_spinodal=mu_q, coexistence_region=(math.nan, math.nan))
    mu_c, om_c = _omega_crossing(mu[co], h["omega"][co], q["omega"][co])
    return dict(found=not math.isnan(mu_c), mu_c=mu_c, omega_at_transition=om_c, mu_hadron_spinodal=mu_h,
                mu_quark_spinodal=mu_q, coexistence_region=(float(mu[co[0]]), float(mu[co[-1]])))


MERGED_COLUMNS = ["T_MeV", "mu_MeV", "xi", "branch", "omega", "pressure", "rho", "entropy", "energy", "phi_u", "phi_d",
                  "phi_s", "Phi1", "Phi2", "M_u_MeV", "M_d_MeV", "M_s_MeV", "delta_omega"]


def merge_branches(result, output_path=None):
    """DualBranchScan.jl:237-305: one row per mu with a physical point
    (branch hadron / quark / same), written with the reference's formats
    (%.6f for T, mu, xi and masses, %.10f otherwise) when output_path is given."""
    mu = result["mu_values"]
    h, q, p = result["hadron"], result["quark"], result["physical"]
    rows = []
    for i in range(mu.size):
        if not p["valid"][i]:
            continue
        if h["valid"][i] and q["valid"][i]:
            if result["same"][i]:
                branch = "same"
            elif abs(p["omega"][i] - h["omega"][i]) < 1e-10:
                branch = "hadron"
            else:
                branch = "quark"
        else:
            branch = "hadron" if h["valid"][i] else "quark"
        dom = h["omega"][i] - q["omega"][i] if (h["valid"][i] and q["valid"][i] and not result["same"][i]) else math.nan
        x, m = p["x"][i], p["masses"][i] * HBARC
        rows.append(dict(T_MeV=result["T_mev"], mu_MeV=float(mu[i]), xi=result["xi"], branch=branch,
                         omega=p["omega"][i], pressure=p["pressure"][i], rho=p["rho_norm"][i],
                         entropy=p["entropy"][i], energy=p["energy"][i], phi_u=x[0], phi_d=x[1], phi_s=x[2],
                         Phi1=x[3], Phi2=x[4], M_u_MeV=m[0], M_d_MeV=m[1], M_s_MeV=m[2], delta_omega=dom))
    if output_path is not None:
        six = {"T_MeV", "mu_MeV", "xi", "M_u_MeV", "M_d_MeV", "M_s_MeV"}
        with open(output_path, "w") as f:
            f.write(",".join(MERGED_COLUMNS) + "\n")
            for r in rows:
                f.write(",".join(r[c] if c == "branch" else "NaN" if math.isnan(r[c]) else
                                 ("%.6f" % r[c] if c in six else "%.10f" % r[c]) for c in MERGED_COLUMNS) + "\n")
    return rows


__all__ = ["run_dual_branch_scan", "run_dual_branch_scan_batch", "find_phase_transition", "merge_branches",
           "gpu_solver", "gpu_chain_solver", "MERGED_COLUMNS"]
