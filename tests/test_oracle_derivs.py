"""Pins the oracle's implicit thermodynamic derivatives (ThermoDerivatives.jl:
mass_derivatives order 1, thermo_derivatives, bulk_viscosity_coefficients)
before it serves as the checker of the GPU derivative pass:
  * finite differences of re-solved roots and of the (pinned) thermodynamics,
  * exact identities (Maxwell relation d s/d mu = d n/d T, dP/dT = s, dP/dmu = n),
  * the reference's own output bulk_viscosity_xi0.0.csv (p24/t8, 7 digits)."""
import numpy as np
import pytest

from conftest import golden_rows

HB = 197.327

STATES = [(150, 0, 0.0), (150, 50, 0.0), (160, 50, 0.0), (120, 250, 0.3), (60, 360, 0.0), (200, 300, -0.2),
          (90, 150, 0.4)]


def _solve(oracle, T, mu, xi, seed=None, pn=24, tn=8):
    T, mu = np.atleast_1d(float(T)), np.atleast_1d(float(mu))
    opts = oracle.default_opts(ftol=1e-13)
    if seed is None:
        return oracle.solve_batch(T, mu, xi, policy=oracle.SEED_DEFAULT, p_num=pn, t_num=tn, opts=opts)
    return oracle.solve_batch(T, mu, xi, policy=oracle.SEED_EXPLICIT, seed=np.atleast_2d(seed), p_num=pn,
                              t_num=tn, opts=opts)


@pytest.mark.parametrize("state", STATES)
def test_derivatives_match_finite_differences(oracle, state):
    Tm, mm, xi = state
    T, mu = Tm / HB, mm / HB
    r = _solve(oracle, T, mu, xi)
    assert r["converged"][0]
    x = r["x"][0]
    d = oracle.thermo_derivs(x, T, mu, xi, 24, 8)
    h = 1e-4
    s = {k: _solve(oracle, T + dT, mu + dm, xi, x) for k, (dT, dm) in
         {"T+": (h, 0), "T-": (-h, 0), "m+": (0, h), "m-": (0, -h)}.items()}
    fd = lambda key, a, b: (s[a][key][0] - s[b][key][0]) / (2 * h)
    close = lambda a, b, tol: np.all(np.abs(np.asarray(a) - b) <= tol * np.maximum(1e-3, np.max(np.abs(b))))
    assert close(d["dx_dT"], fd("x", "T+", "T-"), 2e-5)
    assert close(d["dx_dmu"], fd("x", "m+", "m-"), 2e-5)
    assert close(d["dM_dT"], fd("masses", "T+", "T-"), 2e-5)
    assert close(d["dM_dmu"], fd("masses", "m+", "m-"), 2e-5)
    assert close(d["entropy"], fd("pressure", "T+", "T-"), 1e-6)
    n_of = lambda k: s[k]["rho_norm"][0] * 0.16         # n = sum rho_i / 3 = rho_norm rho0
    assert close(d["rho"] * 3, fd("pressure", "m+", "m-"), 1e-6)
    assert close(d["ds_dT"], fd("entropy", "T+", "T-"), 2e-5)
    assert close(d["ds_dmu"], fd("entropy", "m+", "m-"), 2e-5)
    assert close(d["dn_dmu"], (n_of("m+") - n_of("m-")) / (2 * h), 2e-5)
    assert close(d["dn_dT"], (n_of("T+") - n_of("T-")) / (2 * h), 2e-5)
    assert close(d["dEpsilon_dT"], fd("energy", "T+", "T-"), 2e-5)
    assert close(d["dEpsilon_dmu"], fd("energy", "m+", "m-"), 2e-5)
    # exact identities of the closed forms
    assert abs(d["ds_dmu"] - 3 * d["dn_dT"]) <= 1e-11 * max(1.0, abs(d["ds_dmu"]))     # Maxwell relation
    assert d["dP_dT"] == d["entropy"] and d["dP_dmu"] == 3 * d["rho"]
    assert abs(d["pressure"] - r["pressure"][0]) <= 1e-14 * abs(d["pressure"])
    assert abs(d["rho_norm"] - r["rho_norm"][0]) <= 1e-12 * max(1.0, abs(d["rho_norm"]))


def test_bulk_viscosity_coefficients_match_reference_output(oracle):
    """bulk_viscosity_xi0.0.csv rows at T = 150 MeV (the T = 160 MeV rows hold
    no root of the gap equations: test_reference_T160_rows_are_not_roots).
    The stored masses use the bare masses 5.5 and 140.0 MeV of
    ThermoDerivatives.jl:111-121."""
    rows = [r for r in golden_rows("bulk_viscosity_xi0.0.csv") if float(r["T_MeV"]) == 150.0]
    assert len(rows) == 2
    for row in rows:
        T, mu = float(row["T_MeV"]) / HB, float(row["mu_MeV"]) / HB
        r = _solve(oracle, T, mu, 0.0)
        d = oracle.thermo_derivs(r["x"][0], T, mu, 0.0, 24, 8)
        ref = lambda k: float(row[k])
        rel = lambda a, b: abs(a - b) <= 2e-6 * abs(b)
        assert rel(d["v_n_sq"], ref("v_n_sq"))
        assert rel(d["entropy"], ref("entropy_fm3"))
        for i, f in enumerate("uds"):
            assert abs(d["masses"][i] * HB - ref(f"M_{f}_MeV")) < 2e-6
            assert rel(d["dM_dT"][i], ref(f"dM_{f}_dT"))
        if float(row["mu_MeV"]) > 0:                      # mu = 0: n_B and d/dmu_B vanish (stored noise)
            assert rel(d["dmuB_dT_sigma"], ref("dmuB_dT_sigma"))
            assert rel(d["rho"], ref("n_B_fm3"))
            for i, f in enumerate("uds"):
                assert rel(d["dM_dmu"][i] / 3, ref(f"dM_{f}_dmuB"))
        else:
            assert abs(d["rho"]) < 1e-12 and np.all(np.abs(d["dM_dmu"]) < 1e-12)


DERIV_COLS = [("dM_u_dT", "dM_dT", 0), ("dM_s_dT", "dM_dT", 2), ("dM_u_dmu", "dM_dmu", 0), ("dM_s_dmu", "dM_dmu", 2),
              ("dP_dT", "dP_dT", None), ("dP_dmu", "dP_dmu", None), ("dEpsilon_dT", "dEpsilon_dT", None),
              ("dEpsilon_dmu", "dEpsilon_dmu", None), ("dn_dT", "dn_dT", None), ("dn_dmu", "dn_dmu", None),
              ("pressure_fm4", "pressure", None), ("energy_fm4", "energy", None), ("entropy_fm3", "entropy", None)]


def check_derivative_row(d, row):
    """One row of the reference's derivatives_xi0.0.csv (thermo_derivatives +
    mass_derivatives, 7 significant digits; masses and rho_norm with 6
    decimals).  Columns that vanish at mu = 0 hold rounding noise there and
    are checked for |value| < 1e-12 on both sides."""
    HBm = 197.327
    for i, f in ((0, "u"), (2, "s")):
        assert abs(d["masses"][i] * HBm - float(row[f"M_{f}_MeV"])) < 2e-6
    assert abs(d["rho_norm"] - float(row["rho_norm"])) < 6e-7
    for col, key, idx in DERIV_COLS:
        ref = float(row[col])
        got = d[key] if idx is None else d[key][idx]
        if abs(ref) < 1e-12:
            assert abs(got) < 1e-12, col
        else:
            assert abs(got - ref) <= 6e-7 * abs(ref), (col, got, ref)


def test_derivatives_match_reference_output(oracle):
    """derivatives_xi0.0.csv rows at T = 150 MeV (p24/t8); its T = 160 MeV rows
    hold no root of the gap equations (test_reference_T160_rows_are_not_roots)."""
    rows = [r for r in golden_rows("derivatives_xi0.0.csv") if float(r["T_MeV"]) == 150.0]
    assert len(rows) == 2
    for row in rows:
        T, mu = float(row["T_MeV"]) / HB, float(row["mu_MeV"]) / HB
        r = _solve(oracle, T, mu, 0.0)
        check_derivative_row(oracle.thermo_derivs(r["x"][0], T, mu, 0.0, 24, 8), row)


def _condensates_from_masses(Mu, Ms):
    """(phi_u = phi_d, phi_s) with M_u = m_u0 - 4 G phi_u + 2 K phi_u phi_s and
    M_s = m_s0 - 4 G phi_s + 2 K phi_u^2 (fm^-1), the bare masses 5.5 and
    140.0 MeV with which ThermoDerivatives.jl:111-121 writes the masses: every
    solution found from a grid of starts."""
    from scipy.optimize import fsolve
    L = 602.3 / HB
    G, K, m0, ms0 = 1.835 / L ** 2, 12.36 / L ** 5, 5.5 / HB, 140.0 / HB
    f = lambda v: [m0 - 4 * G * v[0] + 2 * K * v[0] * v[1] - Mu, ms0 - 4 * G * v[1] + 2 * K * v[0] ** 2 - Ms]
    sols = {}
    for a in np.linspace(-12, 12, 13):
        for b in np.linspace(-12, 12, 13):
            v, _, ier, _ = fsolve(f, [a, b], full_output=True)
            if ier == 1 and max(abs(np.array(f(v)))) < 1e-12:
                sols[(round(v[0], 6), round(v[1], 6))] = v
    return list(sols.values())


def _best_gradient(oracle, pu, ps, T, mu, n=25):
    """min over (Phi, Phib) in [0, 1.2]^2 (Phi = Phib at mu = 0) of
    max |dOmega/dx| at (pu, pu, ps, Phi, Phib): a grid, then a bounded local
    minimisation of |dOmega/dx|^2 from the three best grid points"""
    from scipy.optimize import minimize
    g2 = lambda P: oracle.omega_derivs(np.array([pu, pu, ps, P[0], P[-1]]), np.full(3, mu), T, 0.0, 24, 8)[1]
    pts = [(a, b) for a in np.linspace(0.0, 1.2, n) for b in (np.linspace(0.0, 1.2, n) if mu > 0 else [a])]
    vals = sorted((float(np.max(np.abs(g2(P)))), P) for P in pts)
    best = vals[0][0]
    dim = 2 if mu > 0 else 1
    for _, P in vals[:3]:
        r = minimize(lambda q: float(np.sum(g2(q) ** 2)), np.array(P[:dim]), method="L-BFGS-B",
                     bounds=[(0.0, 1.2)] * dim, options={"ftol": 1e-30, "gtol": 1e-14, "maxiter": 500})
        best = min(best, float(np.max(np.abs(g2(r.x)))))
    return best


@pytest.mark.parametrize("name", ["derivatives_xi0.0.csv", "bulk_viscosity_xi0.0.csv"])
def test_reference_T160_rows_are_not_roots(oracle, name):
    """The T = 160 MeV rows of the reference's derivative outputs (M_u = -636.6
    MeV, M_s = 367.5 MeV, reported converged) hold no root of the gap
    equations: the condensates that reproduce their stored masses leave
    max |dOmega/dx| >= 1 for every Polyakov loop (Phi, Phib) in [0, 1.2]^2,
    and a seed sweep over phi_u in [-9, 9], phi_s in [-6, 6] finds only the
    M_u = 358.9 and 650.3 MeV roots at (160 MeV, 0).  The same inversion on
    the T = 150 MeV rows lands on a root (max |dOmega/dx| below the 6-decimal
    rounding of the stored masses) with the stored pressure -- so the
    reference's run wrote a state its solver flagged converged without
    solving the equations there, and those rows cannot be reproduced by any
    root (the rows of T = 150 MeV are checked in full above)."""
    for row in golden_rows(name):
        Tm, mm = float(row["T_MeV"]), float(row["mu_MeV"])
        T, mu = Tm / HB, mm / HB
        sols = _condensates_from_masses(float(row["M_u_MeV"]) / HB, float(row["M_s_MeV"]) / HB)
        assert sols
        g = min(_best_gradient(oracle, v[0], v[1], T, mu) for v in sols)
        print(f"[{name} T={Tm} mu={mm}] {len(sols)} condensate pairs from the stored masses, "
              f"min max|dOmega/dx| over Phi, Phib: {g:.3e}", flush=True)
        if Tm == 160.0:
            assert g > 1.0
        else:
            assert Tm == 150.0
            assert g < 1e-5          # the stored masses' 6 decimals
    # T = 150, mu = 0: the physical pair refined in Phi reaches the stored root and pressure
    row = [r for r in golden_rows("derivatives_xi0.0.csv") if float(r["T_MeV"]) == 150 and float(r["mu_MeV"]) == 0][0]
    from scipy.optimize import minimize_scalar
    T = 150.0 / HB
    pu, ps = [v for v in _condensates_from_masses(float(row["M_u_MeV"]) / HB, float(row["M_s_MeV"]) / HB)
              if v[0] < 0 and v[1] < 0][0]
    gn = lambda P: float(np.max(np.abs(oracle.omega_derivs(np.array([pu, pu, ps, P, P]), np.zeros(3), T, 0.0,
                                                             24, 8)[1])))
    r = minimize_scalar(gn, bounds=(0.0, 0.2), method="bounded", options={"xatol": 1e-12})
    assert r.fun < 1e-6
    th = oracle.thermo(np.array([pu, pu, ps, r.x, r.x]), T, 0.0, 0.0, 24, 8)
    assert abs(th["pressure"] - float(row["pressure_fm4"])) < 1e-5
