"""Implicit thermodynamic derivatives on the GPU -- the API of the reference's
ThermoDerivatives module (src/pnjl/derivatives/ThermoDerivatives.jl):
mass_derivatives, thermo_derivatives, bulk_derivative_coeffs,
bulk_viscosity_coefficients, dP_dT, dP_dmu.

Each call solves the gap equations at (T, mu) as the reference's forward solve
does (PNJL.solve with its default seed strategy, ThermoDerivatives.jl:80-86)
and then evaluates dx/dtheta = -(d2 Omega/dx2)^-1 d2 Omega/dx dtheta and the
chain rules in one pass over the nodes (pnjl_thermo_derivatives_batch).  The
*_batch forms take arrays of points (and optionally the roots) in one launch.

Derivative orders: mass_derivatives and solve_with_derivatives orders 1 and
2 (the second order from x_ab = -H^-1 [(D_wb H) x_a + D_wb F_a] with the
directional derivatives of the closed forms in dual-number arithmetic,
pnjl_implicit_derivatives2_batch), dP_dT and dP_dmu orders 1 and 2; other
orders raise ValueError, as the reference's `error("order must be 1 or 2")`."""
from __future__ import annotations

import numpy as np

from . import _lib
from .pnjl import DefaultSeed, FixedMu, context, solve_batch


def thermo_derivatives_batch(T_fm, mu_fm, *, xi=0.0, roots=None, p_num=64, t_num=8, seed_strategy=None, device=0):
    """Derivatives at every (T, mu, xi) point; roots (n, 5) skip the solve.
    Returns the dict of pnjl_thermo_derivatives_batch columns plus the solve's
    converged / iterations / residual_norm."""
    T = np.atleast_1d(np.asarray(T_fm, dtype=np.float64))
    mu = np.broadcast_to(np.asarray(mu_fm, dtype=np.float64), T.shape).copy()
    xiv = np.broadcast_to(np.asarray(xi, dtype=np.float64), T.shape).copy()
    extra = {}
    if roots is None:
        r = solve_batch(FixedMu(), T, mu, xi=xiv, seed_strategy=seed_strategy or DefaultSeed(), p_num=p_num,
                        t_num=t_num, device=device)
        roots = r["x"]
        extra = dict(converged=r["converged"].astype(bool), iterations=r["iterations"],
                     residual_norm=r["residual_norm"])
    d = context(device).thermo_derivatives_batch(roots, T, mu, xiv, p_num, t_num)
    d.update(extra)
    return d


def _one(T_fm, mu_fm, xi, p_num, t_num):
    d = thermo_derivatives_batch([float(T_fm)], [float(mu_fm)], xi=xi, p_num=p_num, t_num=t_num)
    return {k: (v[0] if np.ndim(v) == 1 else v[0].copy()) for k, v in d.items()}


def implicit_derivatives2_batch(T_fm, mu_fm, *, xi=0.0, roots=None, p_num=64, t_num=8, seed_strategy=None,
                                device=0):
    """x, its first and second (T, mu) derivatives and those of the masses at
    every point (pnjl_implicit_derivatives2_batch); roots (n, 5) skip the solve."""
    T = np.atleast_1d(np.asarray(T_fm, dtype=np.float64))
    mu = np.broadcast_to(np.asarray(mu_fm, dtype=np.float64), T.shape).copy()
    xiv = np.broadcast_to(np.asarray(xi, dtype=np.float64), T.shape).copy()
    if roots is None:
        roots = solve_batch(FixedMu(), T, mu, xi=xiv, seed_strategy=seed_strategy or DefaultSeed(), p_num=p_num,
                            t_num=t_num, device=device)["x"]
    return context(device).implicit_derivatives2_batch(roots, T, mu, xiv, p_num, t_num)


def _check_order(order):
    if order not in (1, 2):
        raise ValueError(f"order must be 1 or 2, got {order}")


def solve_with_derivatives(T_fm, mu_fm, *, order=1, xi=0.0, p_num=64, t_num=8):
    """ImplicitSolver.jl:660-701: x with dx_dT, dx_dmu (order 1) and
    d2x_dT2, d2x_dmu2, d2x_dTdmu (order 2)."""
    _check_order(order)
    d = implicit_derivatives2_batch([float(T_fm)], [float(mu_fm)], xi=xi, p_num=p_num, t_num=t_num)
    r = dict(x=d["x"][0], dx_dT=d["dx_dT"][0], dx_dmu=d["dx_dmu"][0])
    if order == 2:
        r.update(d2x_dT2=d["d2x_dT2"][0], d2x_dmu2=d["d2x_dmu2"][0], d2x_dTdmu=d["d2x_dTdmu"][0])
    return r


def mass_derivatives(T_fm, mu_fm, *, order=1, xi=0.0, p_num=64, t_num=8):
    """ThermoDerivatives.jl:132-179: masses, dM_dT, dM_dmu (order 1) and
    d2M_dT2, d2M_dTdmu, d2M_dmu2 (order 2)."""
    _check_order(order)
    if order == 1:
        d = _one(T_fm, mu_fm, xi, p_num, t_num)
        return dict(masses=d["masses"], dM_dT=d["dM_dT"], dM_dmu=d["dM_dmu"])
    d = implicit_derivatives2_batch([float(T_fm)], [float(mu_fm)], xi=xi, p_num=p_num, t_num=t_num)
    return {k: d[k][0] for k in ("masses", "dM_dT", "dM_dmu", "d2M_dT2", "d2M_dTdmu", "d2M_dmu2")}


def thermo_derivatives(T_fm, mu_fm, *, xi=0.0, p_num=64, t_num=8):
    """ThermoDerivatives.jl:186-245: pressure, energy, rho, rho_norm, entropy,
    their T / mu derivatives, dP_depsilon_n, dP_dn_epsilon, masses, dM_dT,
    dM_dmu, converged, iterations, residual_norm."""
    d = _one(T_fm, mu_fm, xi, p_num, t_num)
    keys = ["pressure", "energy", "rho", "rho_norm", "entropy", "dP_dT", "dP_dmu", "dEpsilon_dT", "dEpsilon_dmu",
            "dn_dT", "dn_dmu", "dP_depsilon_n", "dP_dn_epsilon", "masses", "dM_dT", "dM_dmu", "converged",
            "iterations", "residual_norm"]
    return {k: d[k] for k in keys}


def bulk_derivative_coeffs(T_fm, mu_fm, **kw):
    """ThermoDerivatives.jl:252-260."""
    d = thermo_derivatives(T_fm, mu_fm, **kw)
    return dict(dP_depsilon_n=d["dP_depsilon_n"], dP_dn_epsilon=d["dP_dn_epsilon"], dM_dT=d["dM_dT"],
                dM_dmu=d["dM_dmu"])


def bulk_viscosity_coefficients(T_fm, mu_fm, *, xi=0.0, p_num=64, t_num=8):
    """ThermoDerivatives.jl:342-466: v_n_sq, dmuB_dT_sigma, masses, dM_dT,
    dM_dmuB, s, n_B (mu_B = 3 mu_q)."""
    d = _one(T_fm, mu_fm, xi, p_num, t_num)
    return dict(v_n_sq=d["v_n_sq"], dmuB_dT_sigma=d["dmuB_dT_sigma"], masses=d["masses"], dM_dT=d["dM_dT"],
                dM_dmuB=d["dM_dmu"] / 3.0, s=d["entropy"], n_B=d["rho"])


def dP_dT(T_fm, mu_fm, *, order=1, xi=0.0, p_num=64, t_num=8):
    """ThermoDerivatives.jl:278-293: total dP/dT (order 1: s; order 2: ds/dT)."""
    if order not in (1, 2):
        raise NotImplementedError("dP_dT: orders 1 and 2 are evaluated on this backend")
    d = _one(T_fm, mu_fm, xi, p_num, t_num)
    return d["entropy"] if order == 1 else d["ds_dT"]


def dP_dmu(T_fm, mu_fm, *, order=1, xi=0.0, p_num=64, t_num=8):
    """ThermoDerivatives.jl:300-319: total dP/dmu (order 1: sum_i rho_i;
    order 2: its mu derivative)."""
    if order not in (1, 2):
        raise NotImplementedError("dP_dmu: orders 1 and 2 are evaluated on this backend")
    d = _one(T_fm, mu_fm, xi, p_num, t_num)
    return d["dP_dmu"] if order == 1 else 3.0 * d["dn_dmu"]


__all__ = ["thermo_derivatives_batch", "implicit_derivatives2_batch", "solve_with_derivatives", "mass_derivatives", "thermo_derivatives", "bulk_derivative_coeffs",
           "bulk_viscosity_coefficients", "dP_dT", "dP_dmu"]
