"""First-order transition analysis of mu(rho) curves -- the reference's
PhaseTransition module (src/pnjl/analysis/PhaseTransition.jl:18-437):
S-shape detection (spinodals), the Maxwell equal-area construction and the
grouping of scan rows into curves, plus `s_curve` and `maxwell_line`, which
build the curve of a temperature from batched FixedMu solves on the GPU.

The analysis itself is a few hundred floating-point operations per curve and
runs on the host; the curves are what the GPU produces.  On an S-shaped
mu(rho) curve (first-order region, T < T_CEP) the hadron branch rises to a
local maximum of mu (hadron spinodal), the unstable branch falls to a local
minimum (quark spinodal) and the quark branch rises again; the Maxwell mu_c
is the level whose signed area between the curve and mu = mu_c, taken
between the outermost crossings, vanishes -- the equal-pressure condition
P_hadron(mu_c) = P_quark(mu_c), since dP = rho dmu at fixed T.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

HBARC = 197.327
EPS_SLOPE = 0.0                 # PhaseTransition.jl:29-35
DEFAULT_AREA_TOL = 1e-4
DEFAULT_MAX_ITER = 60
DEFAULT_CANDIDATE_STEPS = 64
MAX_CANDIDATE_STEPS = 1024
BRACKET_SHRINK_REL = 1e-3
BRACKET_SHRINK_ABS = 1e-3


@dataclass
class SShapeResult:
    """SShapeResult  PhaseTransition.jl:52-60 (None where the reference has nothing)."""
    has_s_shape: bool = False
    mu_spinodal_hadron: float | None = None
    mu_spinodal_quark: float | None = None
    rho_spinodal_hadron: float | None = None
    rho_spinodal_quark: float | None = None
    derivative_sign_changes: int = 0


@dataclass
class MaxwellResult:
    """MaxwellResult  PhaseTransition.jl:185-195."""
    converged: bool = False
    mu_transition: float | None = None
    rho_hadron: float | None = None
    rho_quark: float | None = None
    area_residual: float | None = None
    iterations: int = 0
    details: dict = field(default_factory=dict)


def _by_rho(mu_vals, rho_vals):
    """The first min(len) samples ordered by rho (stable, as Julia's sortperm)."""
    n = min(len(mu_vals), len(rho_vals))
    mu = np.asarray(mu_vals[:n], dtype=np.float64)
    rho = np.asarray(rho_vals[:n], dtype=np.float64)
    o = np.argsort(rho, kind="stable")
    return mu[o], rho[o]


def _midpoint(mu, rho, i):
    return (rho[i] + rho[i + 1]) / 2, (mu[i] + mu[i + 1]) / 2


def _refine_extremum(mu, rho, idx, is_maximum):
    """_refine_extremum  PhaseTransition.jl:73-118: the vertex of the parabola
    through the samples around idx (0-based: the slope from idx to idx + 1
    changes sign), or the midpoint of that segment when the parabola opens the
    wrong way or its vertex leaves the three samples."""
    n = rho.size
    if idx < 1 or idx >= n - 1:
        return _midpoint(mu, rho, idx)
    i1, i2, i3 = idx - 1, idx, idx + 1
    r1, r2, r3 = rho[i1], rho[i2], rho[i3]
    m1, m2, m3 = mu[i1], mu[i2], mu[i3]
    den = (r1 - r2) * (r1 - r3) * (r2 - r3)
    if abs(den) < 1e-15:
        return _midpoint(mu, rho, idx)
    a = (r3 * (m2 - m1) + r2 * (m1 - m3) + r1 * (m3 - m2)) / den
    b = (r3 ** 2 * (m1 - m2) + r2 ** 2 * (m3 - m1) + r1 ** 2 * (m2 - m3)) / den
    c = (r2 * r3 * (r2 - r3) * m1 + r3 * r1 * (r3 - r1) * m2 + r1 * r2 * (r1 - r2) * m3) / den
    if (is_maximum and a >= 0) or (not is_maximum and a <= 0):
        return _midpoint(mu, rho, idx)
    r_ext = -b / (2 * a)
    if r_ext < min(r1, r2, r3) or r_ext > max(r1, r2, r3):
        return _midpoint(mu, rho, idx)
    return r_ext, a * r_ext ** 2 + b * r_ext + c


def detect_s_shape(mu_vals, rho_vals, eps=EPS_SLOPE, min_points=5):
    """detect_s_shape  PhaseTransition.jl:121-181: the sign pattern of the
    secant slopes d mu / d rho of the rho-ordered curve; the first + -> -
    change is the hadron spinodal (local maximum of mu), the next - -> + the
    quark spinodal (local minimum), each refined by a parabola."""
    n = min(len(mu_vals), len(rho_vals))
    if n < min_points:
        return SShapeResult()
    mu, rho = _by_rho(mu_vals, rho_vals)
    runs = []                              # (sign, sample index of the run's first slope)
    last = 0
    for i in range(rho.size - 1):
        dr = rho[i + 1] - rho[i]
        if abs(dr) <= eps:
            continue
        s = (mu[i + 1] - mu[i]) / dr
        sg = 0 if abs(s) <= EPS_SLOPE else (1 if s > 0 else -1)
        if sg != 0 and sg != last:
            runs.append((sg, i))
            last = sg
    if not runs:
        return SShapeResult()
    if len(runs) < 3:
        return SShapeResult(derivative_sign_changes=max(0, len(runs) - 1))
    i_max = i_min = None
    for (s0, _), (s1, i1) in zip(runs[:-1], runs[1:]):
        if s0 == 1 and s1 == -1 and i_max is None:
            i_max = i1
        elif s0 == -1 and s1 == 1 and i_max is not None and i_min is None:
            i_min = i1
    if i_max is None or i_min is None:
        return SShapeResult(derivative_sign_changes=len(runs) - 1)
    r_h, m_h = _refine_extremum(mu, rho, i_max, True)
    r_q, m_q = _refine_extremum(mu, rho, i_min, False)
    return SShapeResult(True, float(m_h), float(m_q), float(r_h), float(r_q), len(runs) - 1)


def _finite_curve(mu_vals, rho_vals):
    """_prepare_curve  PhaseTransition.jl:197-207: finite samples by rho."""
    n = min(len(mu_vals), len(rho_vals))
    mu = np.asarray(mu_vals[:n], dtype=np.float64)
    rho = np.asarray(rho_vals[:n], dtype=np.float64)
    ok = np.isfinite(mu) & np.isfinite(rho)
    mu, rho = mu[ok], rho[ok]
    o = np.argsort(rho, kind="stable")
    return rho[o], mu[o]


def _outer_crossings(mu0, rho, mu, atol=1e-9):
    """_find_outer_intersections  PhaseTransition.jl:227-255: the smallest and
    largest rho where the piecewise-linear curve meets mu = mu0.  Per segment
    (in rho order) the events are: its left sample within atol of mu0, then a
    sign change (the interpolated crossing) or, failing that, its right
    sample within atol; the first and last events are the outer crossings."""
    if rho.size < 2:
        return None, None
    r1, r2 = rho[:-1], rho[1:]
    f1, f2 = mu[:-1] - mu0, mu[1:] - mu0
    ev_a = np.abs(f1) < atol
    differ = f1 != f2
    cross = differ & (f1 * f2 < 0)
    ev_b = cross | (differ & ~cross & (np.abs(f2) < atol))
    with np.errstate(divide="ignore", invalid="ignore"):
        xb = np.where(cross, r1 + f1 / (f1 - f2) * (r2 - r1), r2)
    ev = np.stack([ev_a, ev_b], axis=1).ravel()
    if not ev.any():
        return None, None
    val = np.stack([r1, xb], axis=1).ravel()
    hit = np.flatnonzero(ev)
    return float(val[hit[0]]), float(val[hit[-1]])


def _lerp(r1, r2, m1, m2, r):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(r2 == r1, m1, m1 + (r - r1) / (r2 - r1) * (m2 - m1))


def _signed_area(mu0, rho, mu):
    """_area_difference  PhaseTransition.jl:257-276: integral of (mu - mu0)
    d rho over the outer crossings (trapezoids of the linear interpolant,
    summed in rho order as the reference does), None without two distinct
    crossings."""
    lo, hi = _outer_crossings(mu0, rho, mu)
    if lo is None or hi is None or hi - lo <= 1e-9:
        return None
    r1, r2 = rho[:-1], rho[1:]
    a, b = np.maximum(r1, lo), np.minimum(r2, hi)
    use = ~((r2 <= lo) | (r1 >= hi)) & (b > a)
    if not use.any():
        return 0.0
    r1, r2, a, b = r1[use], r2[use], a[use], b[use]
    m1, m2 = mu[:-1][use], mu[1:][use]
    ma, mb = _lerp(r1, r2, m1, m2, a), _lerp(r1, r2, m1, m2, b)
    terms = 0.5 * ((ma - mu0) + (mb - mu0)) * (b - a)
    return float(np.add.accumulate(terms)[-1])     # sequential, left to right


def _linspace(a, b, n):
    """range(a, b; length=n) of Julia: a + (b - a) * k / (n - 1), endpoints exact."""
    if n == 1:
        return [a]
    out = [a + (b - a) * k / (n - 1) for k in range(n)]
    out[-1] = b
    return out


def _scan_for_sign_change(rho, mu, lo, hi, steps, tol):
    """_find_mu_bracket  PhaseTransition.jl:278-295: equally spaced mu0 over
    [lo, hi] (steps, doubled up to 1024) until the area changes sign or
    vanishes to tol."""
    k = max(steps, 3)
    while k <= MAX_CANDIDATE_STEPS:
        pm = pa = None
        for m0 in _linspace(lo, hi, k):
            a = _signed_area(m0, rho, mu)
            if a is None:
                continue
            if abs(a) <= tol:
                return m0, m0, a, a
            if pa is not None and a * pa < 0:
                return pm, m0, pa, a
            pm, pa = m0, a
        k *= 2
    return None


def _bisect(rho, mu, a, b, fa, fb, tol, max_iter):
    """_bisection_solve  PhaseTransition.jl:297-319."""
    if a == b:
        return a, fa, 0
    if fa * fb > 0:
        return None, None, 0
    for it in range(1, max_iter + 1):
        m = 0.5 * (a + b)
        fm = _signed_area(m, rho, mu)
        if fm is None:
            return None, None, it
        if abs(fm) <= tol:
            return m, fm, it
        if fa * fm < 0:
            b, fb = m, fm
        else:
            a, fa = m, fm
    m = 0.5 * (a + b)
    fm = _signed_area(m, rho, mu)
    if fm is None:
        return None, None, max_iter
    return m, fm, max_iter


def _failure(reason, **kw):
    return MaxwellResult(details=dict(reason=reason, **kw))


def maxwell_construction(mu_vals, rho_vals, min_samples=12, detect_min_points=6, detect_eps=1e-6,
                         candidate_steps=DEFAULT_CANDIDATE_STEPS, max_iter=DEFAULT_MAX_ITER,
                         tol_area=DEFAULT_AREA_TOL, spinodal_hint=None):
    """maxwell_construction  PhaseTransition.jl:329-369: the spinodals bracket
    mu_c (shrunk by max(1e-3 * width, 1e-3)); a scan of the bracket finds a
    sign change of the signed area, bisection (to |area| <= tol_area) the
    root; the outer crossings at mu_c are rho_hadron and rho_quark.  Failures
    carry details["reason"] as in the reference."""
    rho, mu = _finite_curve(mu_vals, rho_vals)
    if rho.size < min_samples:
        return _failure("insufficient_points", count=int(rho.size))
    hint = spinodal_hint if spinodal_hint is not None else \
        detect_s_shape(mu_vals, rho_vals, eps=detect_eps, min_points=detect_min_points)
    if not hint.has_s_shape:
        return _failure("no_s_shape")
    if hint.mu_spinodal_hadron is None or hint.mu_spinodal_quark is None:
        return _failure("invalid_mu_bracket")
    lo, hi = sorted((hint.mu_spinodal_quark, hint.mu_spinodal_hadron))
    w = hi - lo
    if not w > 0:
        return _failure("degenerate_bracket", bracket=(lo, hi))
    d = max(w * BRACKET_SHRINK_REL, BRACKET_SHRINK_ABS)
    if lo + d >= hi - d:
        d = w / 4
        if not lo + d < hi - d:
            return _failure("degenerate_bracket", bracket=(lo, hi))
    lo, hi = lo + d, hi - d
    br = _scan_for_sign_change(rho, mu, lo, hi, candidate_steps, tol_area)
    if br is None:
        return _failure("no_sign_change", bracket=(lo, hi))
    a, b, fa, fb = br
    root, area, it = _bisect(rho, mu, a, b, fa, fb, tol_area, max_iter)
    if root is None:
        return _failure("bisection_failed", bracket=(a, b))
    left, right = _outer_crossings(root, rho, mu)
    if left is None or right is None or not left < right:
        return _failure("no_crossings", mu_transition=root, bracket=(a, b))
    return MaxwellResult(True, float(root), float(left), float(right), float(abs(area)), it,
                         dict(mu_bracket=(float(a), float(b)), rho_interval=(float(left), float(right)),
                              spinodal_hint=(hint.mu_spinodal_hadron, hint.mu_spinodal_quark)))


def group_curves_by_temperature(rows, xi=0.0, tol=1e-6):
    """group_curves_by_temperature  PhaseTransition.jl:371-387: {T_MeV: [(mu,
    rho), ...]} from rows (dicts of strings) of xi; mu is mu_avg_MeV, else
    mu_MeV; rows that do not parse are skipped."""
    def num(row, k):
        try:
            return float(row[k])
        except (KeyError, TypeError, ValueError):
            return math.nan
    out = {}
    for row in rows:
        T, x, rho = num(row, "T_MeV"), num(row, "xi"), num(row, "rho")
        if math.isnan(T) or math.isnan(x) or abs(x - xi) > tol or math.isnan(rho):
            continue
        mu = num(row, "mu_avg_MeV")
        if math.isnan(mu):
            mu = num(row, "mu_MeV")
        if math.isnan(mu) or not math.isfinite(rho):
            continue
        out.setdefault(T, []).append((mu, rho))
    return out


def load_curves_from_csv(path, xi):
    """load_curves_from_csv  scripts/pnjl/calculate_phase_structure.jl:226-273:
    {T_MeV: (mu_avg values, rho values)} of the converged rows of a TrhoScan
    file at xi, temperatures with fewer than 3 samples dropped."""
    import csv
    grouped = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            try:
                T, rho, x, mu = (float(r[k]) for k in ("T_MeV", "rho", "xi", "mu_avg_MeV"))
            except (KeyError, TypeError, ValueError):
                continue
            if str(r.get("converged", "true")).strip().lower() not in ("true", "1") or abs(x - xi) > 1e-6:
                continue
            grouped.setdefault(T, []).append((mu, rho))
    return {T: ([s[0] for s in v], [s[1] for s in v]) for T, v in grouped.items() if len(v) >= 3}


SADDLE_SEED_FRACTIONS = (0.5, 0.25, 0.75, 0.1, 0.9, 0.05, 0.95)


def s_curves(T_mev, mu_values, xi=0.0, p_num=24, t_num=8, opts=None, device=0, solver=None,
             fractions=SADDLE_SEED_FRACTIONS, chain_solver=None, point_solver=None, return_lines=False):
    """The mu(rho) curves of several (T, xi) lines from FixedMu solves on the
    GPU: the hadron branch followed up in mu and the quark branch down (the
    dual-branch scan of all lines at once: one launch of branch chains per
    branch, or one batched solve per mu step when `solver` is given),
    then the unstable branch of every line in one more batch: at every mu
    where both stable branches exist and differ, Newton is started from the
    points x_h + f (x_q - x_h) of the segment between the two stable states
    (f in `fractions`); the first start (in that order) that converges to a
    density strictly between the two stable ones is the unstable state (the
    saddle of Omega between the minima).  Returns one (mu_MeV, rho) pair per
    line with all three branches together -- the S-curve that TrhoScan
    traces at fixed rho.  `solver(T, mu, xi, policy, seed)` replaces every
    GPU solve (the branches then go one mu step per call); `chain_solver` and
    `point_solver` replace the branch-chain and the per-point GPU solves."""
    from .dual_branch import gpu_solver, run_dual_branch_scan_batch
    solve = solver or point_solver or gpu_solver(device, p_num, t_num, opts)
    T_mev = np.atleast_1d(np.asarray(T_mev, dtype=np.float64))
    xi = np.broadcast_to(np.asarray(xi, dtype=np.float64), T_mev.shape)
    lines = run_dual_branch_scan_batch(T_mev, mu_values, xi, p_num, t_num, opts, device, solver=solver,
                                       chain_solver=chain_solver)
    f = np.asarray(fractions, dtype=np.float64)
    nf = f.size
    seeds, Ts, mus, xis, owner = [], [], [], [], []
    for l, res in enumerate(lines):
        h, q, mu = res["hadron"], res["quark"], res["mu_values"]
        both = np.flatnonzero(h["valid"] & q["valid"] & ~res["same"])
        xh, xq = h["x"][both], q["x"][both]
        seeds.append((xh[:, None, :] + f[None, :, None] * (xq - xh)[:, None, :]).reshape(-1, 5))
        mus.append(np.repeat(mu[both], nf))
        Ts.append(np.full(both.size * nf, T_mev[l]))
        xis.append(np.full(both.size * nf, xi[l]))
        owner.append(both)
    out = []
    n_mid = sum(o.size for o in owner)
    r = None
    if n_mid:
        r = solve(np.concatenate(Ts) / HBARC, np.concatenate(mus) / HBARC, np.concatenate(xis), _lib.SEED_EXPLICIT,
                  np.concatenate(seeds))
        good = (r["converged"] == 1) & ((r["flags"] & (_lib.PT_EXCEPTION | _lib.PT_NO_ROOT)) == 0)
    base = 0
    for res, both in zip(lines, owner):
        h, q, mu = res["hadron"], res["quark"], res["mu_values"]
        m_list = [mu[h["valid"]], mu[q["valid"]]]
        r_list = [h["rho_norm"][h["valid"]], q["rho_norm"][q["valid"]]]
        if both.size:
            sl = slice(base * nf, (base + both.size) * nf)
            rr = r["rho_norm"][sl].reshape(-1, nf)
            ok = good[sl].reshape(-1, nf) & (rr > h["rho_norm"][both][:, None] + 1e-6) & \
                (rr < q["rho_norm"][both][:, None] - 1e-6)
            has = ok.any(axis=1)
            first = np.argmax(ok, axis=1)
            m_list.append(mu[both][has])
            r_list.append(rr[np.flatnonzero(has), first[has]])
            base += both.size
        out.append((np.concatenate(m_list), np.concatenate(r_list)))
    return (out, lines) if return_lines else out


def s_curve(T_mev, mu_values, xi=0.0, p_num=24, t_num=8, opts=None, device=0, solver=None,
            fractions=SADDLE_SEED_FRACTIONS, chain_solver=None):
    """`s_curves` of one temperature: (mu_MeV, rho)."""
    return s_curves([T_mev], mu_values, xi, p_num, t_num, opts, device, solver, fractions, chain_solver)[0]


def maxwell_line(T_values, mu_values, xi=0.0, p_num=24, t_num=8, opts=None, device=0, solver=None,
                 chain_solver=None, point_solver=None, refine=16, **kw):
    """{T_MeV: MaxwellResult} of the GPU S-curves of the temperatures (all
    lines solved together; with one xi per line the keys are (T_MeV, xi));
    keyword arguments go to maxwell_construction.
    A line whose construction fails although its two stable branches coexist
    (close to the CEP the S-region can be narrower than the mu grid) is
    solved again on a grid `refine` times finer over its coexistence window
    (between the first quark-branch point and the last hadron-branch point,
    also when these miss each other by up to two coarse steps, with two
    coarse steps of margin on each side, clipped to the mu range), merged
    with its coarse curve, and constructed again; the details record
    `refined`.  A line whose two branches are the same root wherever both
    exist (a crossover line above the CEP) has no coexistence window and is
    not refined."""
    T_values = [float(T) for T in np.atleast_1d(T_values)]
    mu_values = np.asarray(list(mu_values), dtype=np.float64)
    per_line = np.ndim(xi) > 0
    xi_l = np.broadcast_to(np.asarray(xi, dtype=np.float64), (len(T_values),))
    curves, lines = s_curves(T_values, mu_values, xi_l, p_num, t_num, opts, device, solver, chain_solver=chain_solver,
                             point_solver=point_solver, return_lines=True)
    out = {}
    step = float(np.min(np.diff(mu_values))) if mu_values.size > 1 else 0.0
    for T, x, (mu, rho), res in zip(T_values, xi_l, curves, lines):
        m = maxwell_construction(mu, rho, **kw)
        h, q = res["hadron"]["valid"], res["quark"]["valid"]
        overlap = h & q
        crossover = overlap.any() and bool(np.all(res["same"][overlap]))
        if not m.converged and refine and step > 0 and h.any() and q.any() and not crossover:
            lo, hi = mu_values[np.flatnonzero(q)[0]], mu_values[np.flatnonzero(h)[-1]]
            if lo <= hi + 2 * step:      # coexisting, or meeting between two grid points
                lo, hi = min(lo, hi), max(lo, hi)
                lo, hi = max(lo - 2 * step, mu_values[0]), min(hi + 2 * step, mu_values[-1])
                fine = np.arange(lo, hi + 1e-9, step / refine)
                mf, rf = s_curves([T], fine, x, p_num, t_num, opts, device, solver, chain_solver=chain_solver,
                                  point_solver=point_solver)[0]
                m = maxwell_construction(np.concatenate([mu, mf]), np.concatenate([rho, rf]), **kw)
                m.details["refined"] = (float(fine[0]), float(fine[-1]), step / refine)
        out[(T, float(x)) if per_line else T] = m
    return out


__all__ = ["SShapeResult", "MaxwellResult", "detect_s_shape", "maxwell_construction", "group_curves_by_temperature",
           "load_curves_from_csv", "s_curve", "s_curves", "maxwell_line",
           "SADDLE_SEED_FRACTIONS"]
