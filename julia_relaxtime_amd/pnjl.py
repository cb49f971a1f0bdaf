"""Host mirror of the reference's PNJL solver API, backed by libpnjl_hip.so.

Names, argument meaning and error behaviour follow the Julia module
(src/pnjl/PNJL.jl:64-118):

    solve(FixedMu(), T_fm, mu_fm; xi, seed_strategy, p_num, t_num, ...)  ImplicitSolver.jl:211-328
    SolverResult                                                          ImplicitSolver.jl:178-193
    DefaultSeed, MultiSeed, ContinuitySeed, PhaseAwareContinuitySeed,
    get_seed, update!, reset!                                            SeedStrategies.jl:183-888

Julia's `update!`/`reset!` are `update`/`reset` here.  `solve_batch` is the
batched form the scan driver switches to with one keyword: a whole grid (or
its continuity chains) in one persistent-kernel launch.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (SEED_CONTINUITY, SEED_DEFAULT, SEED_EXPLICIT, SEED_HADRON, SEED_MULTI, SEED_MULTIROOT3,
                   SEED_PHASE_AWARE, SEED_QUARK, SEED_SURROGATE, PT_EXCEPTION, PT_NO_ROOT, PnjlError)

HBARC = 197.327
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# SeedStrategies.jl:56-107
HADRON_SEED_5 = [-1.84329, -1.84329, -2.22701, 1.0e-5, 4.0e-5]
MEDIUM_SEED_5 = [-1.3647, -1.3647, -2.14502, 0.10594, 0.15569]
HIGH_DENSITY_SEED_5 = [-0.21695, -0.21695, -2.01372, 0.18601, 0.22333]
HIGH_TEMP_SEED_5 = [-0.73192, -0.73192, -1.79539, 0.60532, 0.60532]
VERY_HIGH_TEMP_SEED_5 = [-0.30, -0.30, -0.90, 0.90, 0.90]
HT_GUESS_0p8_SEED_5 = [-0.50, -0.50, -1.20, 0.80, 0.80]
HT_GUESS_0p9_SEED_5 = [-0.30, -0.30, -0.90, 0.90, 0.90]
HT_GUESS_0p95_SEED_5 = [-0.20, -0.20, -0.70, 0.95, 0.95]
WEAK_CHIRAL_CONF_SEED_5 = [-0.50, -0.50, -1.20, 1e-3, 1e-3]
QUARK_SEED_5 = HIGH_TEMP_SEED_5


class FixedMu:
    """Fixed chemical potential mode: theta = (T, mu), x = (phi_u, phi_d, phi_s, Phi, Phibar)."""

    def __repr__(self):
        return "FixedMu()"


# ---------------------------------------------------------------------------
# seed strategies
# ---------------------------------------------------------------------------
def auto_phase_hint(T_fm, mu_fm):
    """quark if T > 150 MeV or mu > 300 MeV (SeedStrategies.jl:131-136)."""
    return "quark" if (T_fm * HBARC > 150 or mu_fm * HBARC > 300) else "hadron"


class SeedStrategy:
    pass


@dataclass
class DefaultSeed(SeedStrategy):
    hadron_seed: list = field(default_factory=lambda: list(HADRON_SEED_5))
    quark_seed: list = field(default_factory=lambda: list(QUARK_SEED_5))
    phase_hint: str = "auto"

    def is_builtin(self):
        return self.hadron_seed == HADRON_SEED_5 and self.quark_seed == QUARK_SEED_5


@dataclass
class MultiSeed(SeedStrategy):
    """Six candidates; the converged root with minimum Omega wins (SeedStrategies.jl:251-284)."""

    def candidates(self):
        return [DefaultSeed(phase_hint="hadron"), DefaultSeed(phase_hint="quark"),
                DefaultSeed(list(WEAK_CHIRAL_CONF_SEED_5), list(WEAK_CHIRAL_CONF_SEED_5), "hadron"),
                DefaultSeed(list(HT_GUESS_0p8_SEED_5), list(HT_GUESS_0p8_SEED_5), "hadron"),
                DefaultSeed(list(HT_GUESS_0p9_SEED_5), list(HT_GUESS_0p9_SEED_5), "hadron"),
                DefaultSeed(list(HT_GUESS_0p95_SEED_5), list(HT_GUESS_0p95_SEED_5), "hadron")]


@dataclass
class ContinuitySeed(SeedStrategy):
    previous_solution: list | None = None
    fallback: SeedStrategy = field(default_factory=DefaultSeed)


def load_phase_boundary(xi, boundary_path=None, cep_path=None):
    """Rows of boundary.csv / cep.csv within 1e-6 of xi (SeedStrategies.jl:388-436).
    Returns (rows[k,3] of (xi, T_MeV, mu_MeV), cep[m,3]) for all xi, and the
    (T_values, mu_values, T_CEP, mu_CEP) of the requested xi."""
    bpath = boundary_path or os.path.join(_DATA, "boundary.csv")
    cpath = cep_path or os.path.join(_DATA, "cep.csv")
    rows = np.loadtxt(bpath, delimiter=",", skiprows=1, usecols=(0, 1, 2), ndmin=2)
    cep = np.loadtxt(cpath, delimiter=",", skiprows=1, ndmin=2)
    sel = rows[np.abs(rows[:, 0] - xi) <= 1e-6]
    sel = sel[np.argsort(sel[:, 1], kind="stable")]
    c = cep[np.abs(cep[:, 0] - xi) <= 1e-6]
    T_CEP, mu_CEP = (c[0, 1], c[0, 2]) if len(c) else (np.nan, np.nan)
    return rows, cep, (sel[:, 1].copy(), sel[:, 2].copy(), T_CEP, mu_CEP)


def interpolate_mu_c(data, T_MeV):
    """SeedStrategies.jl:446-475"""
    Ts, mus, T_CEP, _ = data
    if not np.isnan(T_CEP) and T_MeV > T_CEP:
        return np.nan
    if len(Ts) == 0:
        return np.nan
    if T_MeV <= Ts[0]:
        return mus[0]
    if T_MeV >= Ts[-1]:
        return mus[-1]
    for i in range(len(Ts) - 1):
        if Ts[i] <= T_MeV <= Ts[i + 1]:
            t = (T_MeV - Ts[i]) / (Ts[i + 1] - Ts[i])
            return mus[i] + t * (mus[i + 1] - mus[i])
    return np.nan


@dataclass
class PhaseAwareContinuitySeed(SeedStrategy):
    """Continuity tracking that re-seeds with the phase default when the
    hadron/quark boundary is crossed (SeedStrategies.jl:679-888)."""
    xi: float = 0.0
    previous_solution: list | None = None
    previous_phase: str = "unknown"
    bootstrap_multiseed: bool = False
    boundary: tuple | None = None

    def __post_init__(self):
        if self.boundary is None:
            self.boundary = load_phase_boundary(self.xi)[2]

    def current_phase(self, T_MeV, mu_MeV):
        Ts, _, T_CEP, _ = self.boundary
        if not np.isnan(T_CEP) and T_MeV > T_CEP:
            return "crossover"
        muc = interpolate_mu_c(self.boundary, T_MeV)
        if np.isnan(muc):
            return "unknown"
        return "hadron" if mu_MeV < muc else "quark"


@dataclass
class SurrogateSeed(SeedStrategy):
    """Learned seed (julia_relaxtime_amd.surrogate.Surrogate).  Newton from the
    learned seed gets `max_steps` steps; a point it does not solve in time is
    handed to the plain DefaultSeed policy.  Inside the model's ambiguity mask
    (`model.judge_mask`, from Surrogate.build_judge_mask) the judge decides:
      "reference"  DefaultSeed solves the point (the reference's own branch);
      "omega"      the minimum-Omega root among the learned root and the
                   MultiSeed candidates is kept (the stable branch);
      "multiroot3" every point, mask or not, is solved from the learned,
                   HADRON and QUARK seeds and the minimum-Omega root kept."""
    model: object = None
    max_steps: int = 8
    judge: str = "reference"

    def policy(self):
        if self.judge not in ("reference", "omega", "multiroot3"):
            raise ValueError(f"unknown judge {self.judge!r}")
        return SEED_MULTIROOT3 if self.judge == "multiroot3" else SEED_SURROGATE

    def judge_mode(self):
        return _lib.JUDGE_OMEGA if self.judge == "omega" else _lib.JUDGE_REFERENCE

    def install(self, ctx):
        m = getattr(self.model, "judge_mask", None)
        if m is None or self.judge == "multiroot3":
            ctx.set_judge_mask(None, None, None)
        else:
            ctx.set_judge_mask(m, self.model.in_lo, self.model.in_hi)


def get_seed(s: SeedStrategy, theta, mode=None):
    """get_seed(strategy, [T, mu], FixedMu()) -> 5-vector (SeedStrategies.jl:204-839)."""
    T_fm, mu_fm = float(theta[0]), float(theta[1]) if len(theta) > 1 else 0.0
    if isinstance(s, DefaultSeed):
        hint = s.phase_hint
        if hint == "auto":
            hint = auto_phase_hint(T_fm, mu_fm)
        if hint == "quark":
            return list(VERY_HIGH_TEMP_SEED_5) if T_fm * HBARC >= 300.0 else list(s.quark_seed)
        return list(s.hadron_seed)
    if isinstance(s, MultiSeed):
        return get_seed(s.candidates()[0], theta)
    if isinstance(s, ContinuitySeed):
        if s.previous_solution is not None:
            return list(s.previous_solution)[:5]
        return get_seed(s.fallback, theta)
    if isinstance(s, SurrogateSeed):
        return list(s.model.infer(np.array([T_fm]), np.array([mu_fm]), np.array([0.0]))[0])
    if isinstance(s, PhaseAwareContinuitySeed):
        cur = s.current_phase(T_fm * HBARC, mu_fm * HBARC)
        if s.previous_solution is None:
            if cur == "hadron":
                return list(HADRON_SEED_5)
            if cur == "quark":
                return list(QUARK_SEED_5)
            return get_seed(DefaultSeed(), theta)
        if (s.previous_phase, cur) in (("hadron", "quark"), ("quark", "hadron")):
            return list(HADRON_SEED_5) if cur == "hadron" else list(QUARK_SEED_5)
        return list(s.previous_solution)[:5]
    raise TypeError(f"unsupported seed strategy {type(s).__name__}")


def update(s: SeedStrategy, solution, T_MeV=None, mu_MeV=None):
    """update!(strategy, solution[, T_MeV, mu_MeV])"""
    s.previous_solution = [float(v) for v in solution]
    if isinstance(s, PhaseAwareContinuitySeed) and T_MeV is not None:
        s.previous_phase = s.current_phase(T_MeV, mu_MeV)
    return s


def reset(s: SeedStrategy):
    """reset!(strategy)"""
    s.previous_solution = None
    if isinstance(s, PhaseAwareContinuitySeed):
        s.previous_phase = "unknown"
    return s


# ---------------------------------------------------------------------------
# results
# ---------------------------------------------------------------------------
@dataclass
class SolverResult:
    mode: FixedMu
    converged: bool
    solution: np.ndarray
    x_state: np.ndarray
    mu_vec: np.ndarray
    omega: float
    pressure: float
    rho_norm: float
    entropy: float
    energy: float
    masses: np.ndarray
    iterations: int
    residual_norm: float
    xi: float
    n_q: np.ndarray | None = None
    n_qbar: np.ndarray | None = None
    flags: int = 0


_CTX = {}


def context(device: int = 0) -> _lib.Context:
    """The process-wide context of a device, with the phase-boundary tables loaded."""
    c = _CTX.get(device)
    if c is None:
        c = _lib.Context(device)
        rows, cep, _ = load_phase_boundary(0.0)
        c.set_phase_boundary(rows, cep)
        _CTX[device] = c
    return c


def _opts(iterations, trust_region_fallback, auto_multiseed_fallback, residual_norm_max, xtol, ftol, chain_len=0,
          surrogate_max_steps=8, surrogate_judge=0):
    return _lib.default_opts(max_iter=int(iterations), trust_region_fallback=int(trust_region_fallback),
                             auto_multiseed_fallback=int(auto_multiseed_fallback),
                             residual_norm_max=float(residual_norm_max), xtol=float(xtol), ftol=float(ftol),
                             chain_len=int(chain_len), surrogate_max_steps=int(surrogate_max_steps),
                             surrogate_judge=int(surrogate_judge))


def _policy_for(seed_strategy, T_fm, mu_fm):
    """-> (policy, explicit seed or None)"""
    s = seed_strategy
    if isinstance(s, DefaultSeed):
        if s.is_builtin():
            return {"auto": SEED_DEFAULT, "hadron": SEED_HADRON, "quark": SEED_QUARK}[s.phase_hint], None
        return SEED_EXPLICIT, get_seed(s, [T_fm, mu_fm])
    if isinstance(s, MultiSeed):
        return SEED_MULTI, None
    if isinstance(s, ContinuitySeed):
        if s.previous_solution is None:
            return _policy_for(s.fallback, T_fm, mu_fm)
        return SEED_EXPLICIT, get_seed(s, [T_fm, mu_fm])
    if isinstance(s, PhaseAwareContinuitySeed):
        if s.bootstrap_multiseed and s.previous_solution is None:
            return SEED_MULTI, None
        g = get_seed(s, [T_fm, mu_fm])
        return SEED_EXPLICIT, g
    if isinstance(s, SurrogateSeed):
        return s.policy(), None
    raise TypeError(f"unsupported seed strategy {type(s).__name__}")


def solve(mode, T_fm, mu_fm, *, xi=0.0, seed_strategy=None, p_num=64, t_num=8, iterations=1000,
          trust_region_fallback=True, auto_multiseed_fallback=True, residual_norm_max=1e-6, xtol=1e-9, ftol=1e-9,
          device=0) -> SolverResult:
    """PNJL.solve(FixedMu(), T_fm, mu_fm; kwargs...) on the GPU.

    Like the reference, non-convergence is reported through `converged`;
    a non-finite residual at the seed raises (NLsolve's IsFiniteException) and
    MultiSeed raises when no candidate converges (ImplicitSolver.jl:553-556)."""
    if not isinstance(mode, FixedMu):
        raise TypeError("only FixedMu() is supported by this backend")
    s = seed_strategy if seed_strategy is not None else DefaultSeed()
    policy, seed = _policy_for(s, T_fm, mu_fm)
    steps, judge = 8, _lib.JUDGE_REFERENCE
    ctx = context(device)
    if isinstance(s, SurrogateSeed):
        seed = s.model.infer(np.array([float(T_fm)]), np.array([float(mu_fm)]), np.array([float(xi)]))[0]
        steps, judge = s.max_steps, s.judge_mode()
        s.install(ctx)
    r = ctx.solve_batch(np.array([float(T_fm)]), np.array([float(mu_fm)]), np.array([float(xi)]),
                        policy=policy, seed=None if seed is None else np.array([seed]), p_num=p_num,
                        t_num=t_num, opts=_opts(iterations, trust_region_fallback, auto_multiseed_fallback,
                                                residual_norm_max, xtol, ftol, surrogate_max_steps=steps,
                                                surrogate_judge=judge))
    fl = int(r["flags"][0])
    if fl & PT_EXCEPTION:
        raise FloatingPointError("non-finite residual at the seed (NLsolve IsFiniteException)")
    if fl & PT_NO_ROOT:
        raise RuntimeError("All seeds failed to converge to a physical solution")
    return _result(r, 0, float(mu_fm), float(xi))


def _result(r, i, mu, xi):
    x = r["x"][i].copy()
    return SolverResult(mode=FixedMu(), converged=bool(r["converged"][i]), solution=x.copy(), x_state=x,
                        mu_vec=np.array([mu, mu, mu]), omega=float(r["omega"][i]), pressure=float(r["pressure"][i]),
                        rho_norm=float(r["rho_norm"][i]), entropy=float(r["entropy"][i]),
                        energy=float(r["energy"][i]), masses=r["masses"][i].copy(),
                        iterations=int(r["iterations"][i]), residual_norm=float(r["residual_norm"][i]), xi=xi,
                        n_q=r["n_q"][i].copy(), n_qbar=r["n_qbar"][i].copy(), flags=int(r["flags"][i]))


def solve_batch(mode, T_fm, mu_fm, *, xi=0.0, seed_strategy=None, seeds=None, chain_len=0, p_num=64, t_num=8,
                iterations=1000, trust_region_fallback=True, auto_multiseed_fallback=True, residual_norm_max=1e-6,
                xtol=1e-9, ftol=1e-9, device=0, as_results=False):
    """Batched PNJL.solve over arrays of points in one launch.

    seed_strategy: DefaultSeed / MultiSeed applied per point; ContinuitySeed or
    PhaseAwareContinuitySeed run consecutive blocks of `chain_len` points as
    continuity chains (the driver's loop order); `seeds` (n,5) gives explicit
    per-point seeds.  Returns a dict of numpy arrays (or SolverResults)."""
    if not isinstance(mode, FixedMu):
        raise TypeError("only FixedMu() is supported by this backend")
    T = np.ascontiguousarray(T_fm, dtype=np.float64).ravel()
    n = T.size
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu_fm, dtype=np.float64), (n,)))
    xiv = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, dtype=np.float64), (n,)))
    s = seed_strategy if seed_strategy is not None else DefaultSeed()
    steps, judge = 8, _lib.JUDGE_REFERENCE
    ctx = context(device)
    if seeds is not None:
        policy = SEED_EXPLICIT
    elif isinstance(s, SurrogateSeed):
        policy = s.policy()
        seeds = s.model.infer(T, mu, xiv)
        steps, judge = s.max_steps, s.judge_mode()
        s.install(ctx)
    elif isinstance(s, ContinuitySeed):
        policy = SEED_CONTINUITY
    elif isinstance(s, PhaseAwareContinuitySeed):
        policy = SEED_PHASE_AWARE
    elif isinstance(s, MultiSeed):
        policy = SEED_MULTI
    elif isinstance(s, DefaultSeed) and s.is_builtin():
        policy = {"auto": SEED_DEFAULT, "hadron": SEED_HADRON, "quark": SEED_QUARK}[s.phase_hint]
    else:
        policy = SEED_EXPLICIT
        seeds = np.array([get_seed(s, [T[i], mu[i]]) for i in range(n)])
    r = ctx.solve_batch(T, mu, xiv, policy=policy, seed=seeds, p_num=p_num, t_num=t_num,
                        opts=_opts(iterations, trust_region_fallback, auto_multiseed_fallback, residual_norm_max,
                                   xtol, ftol, chain_len, steps, judge))
    if as_results:
        return [_result(r, i, float(mu[i]), float(xiv[i])) for i in range(n)]
    return r


__all__ = ["FixedMu", "DefaultSeed", "MultiSeed", "ContinuitySeed", "PhaseAwareContinuitySeed", "SurrogateSeed",
           "SolverResult",
           "solve", "solve_batch", "update", "reset", "get_seed", "context", "HBARC", "PnjlError",
           "load_phase_boundary", "interpolate_mu_c", "auto_phase_hint"]
