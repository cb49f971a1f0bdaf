"""Pins the CPU oracle of the gap path before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign, transposed index, wrong node layout) fails it:
  * golden data the reference itself produced (dual-branch scan, driver scans);
  * library routines (numpy Gauss-Legendre, scipy quad);
  * finite differences of Omega for the analytic gradient / Hessian;
  * identities of the reference's own tests.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from conftest import golden_rows

HB = 197.327


# ---------------------------------------------------------------------------
# Gauss-Legendre (GaussLegendre.jl; test_gausslegendre.jl:45-178)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 4, 6, 8, 12, 16, 24, 32, 64, 128])
def test_gl_matches_numpy(oracle, n):
    x, w = oracle.gl_std(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 2e-15
    assert np.max(np.abs(w - wr)) < 2e-14          # numpy's own endpoint weights carry ~1e-14
    assert np.all(np.diff(x) > 0)
    assert abs(w.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("n", [16, 64, 128])
def test_gl_matches_high_precision(oracle, n):
    import mpmath as mp
    mp.mp.dps = 40
    x, w = oracle.gl_std(n)
    for i in (0, 1, n // 3, n // 2):
        z = mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(x[i]))
        dp = mp.diff(lambda t: mp.legendre(n, t), z)
        assert abs(x[i] - float(z)) < 2e-16
        assert abs(w[i] - float(2 / ((1 - z ** 2) * dp ** 2))) < 1e-15


def test_gauleg_polynomial_exactness(oracle):
    x, w = oracle.gauleg(0.0, 10.0, 12)
    for k in range(0, 24):
        exact = 10.0 ** (k + 1) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) <= 1e-12 * exact
    # numpy reference of the default momentum table: first node, weight sum
    x64, w64 = oracle.gauleg(0.0, 10.0, 64)
    assert abs(x64[0] - 0.0034747913211390369) < 1e-16
    assert abs(w64.sum() - 10.0) < 1e-13


def test_node_mesh_layout(oracle):
    p, c, C = oracle.nodes(24, 8)
    pn, wn = oracle.gauleg(0.0, 10.0, 24)
    tn, tw = oracle.gauleg(0.0, 1.0, 8)
    P = p.reshape(8, 24)      # column-major p x t: index j + l*p_num
    Cc = c.reshape(8, 24)
    W = C.reshape(8, 24)
    assert np.all(P == pn[None, :])
    assert np.all(Cc == tn[:, None])
    assert np.allclose(W, (wn[None, :] * 2 * tw[:, None]) * pn[None, :] ** 2 / (2 * math.pi) ** 2, rtol=1e-15)
    # sum C g = int d^3p/(2pi)^3 g for g even in cos(theta): g = 1 on p < 10
    assert abs(C.sum() - 1000.0 / 3.0 * 2.0 / (4 * math.pi ** 2)) < 1e-12


# ---------------------------------------------------------------------------
# vacuum term: closed form vs quadrature (test_aniso_gap_solver.jl:22-34)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("M", [0.05, 0.5, 1.86, 2.78, 4.0])
def test_vacuum_integral_closed_form(oracle, M):
    L = 602.3 / HB
    I, dI, d2I = oracle.vacuum_integral(M)
    q = integrate.quad(lambda p: p * p * math.sqrt(p * p + M * M), 0.0, L, epsabs=0, epsrel=1e-13)[0]
    Mt = M + 1e-12
    qt = integrate.quad(lambda p: p * p * math.sqrt(p * p + Mt * Mt), 0.0, L, epsabs=0, epsrel=1e-13)[0]
    assert abs(I - qt / (2 * math.pi ** 2)) < 1e-12 * abs(I)
    assert abs(I - q / (2 * math.pi ** 2)) < 1e-9 * abs(I)
    # derivatives: dI/dM = M/(2 pi^2) int p^2/E dp (exactly differentiating the closed form)
    dq = integrate.quad(lambda p: p * p / math.sqrt(p * p + Mt * Mt), 0.0, L, epsabs=0, epsrel=1e-13)[0]
    assert abs(dI - Mt * dq / (2 * math.pi ** 2)) < 1e-11 * abs(dI)
    h = 1e-4
    fd = (oracle.vacuum_integral(M + h)[1] - oracle.vacuum_integral(M - h)[1]) / (2 * h)
    assert abs(d2I - fd) < 1e-7 * abs(d2I)


# ---------------------------------------------------------------------------
# distributions and the log term (test_pnjl_quark_distribution_sum.jl:10-21)
# ---------------------------------------------------------------------------
def test_distribution_sum_rule(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(7)
    for _ in range(200):
        E, mu, T = rng.uniform(0.1, 3), rng.uniform(-1, 2), rng.uniform(0.1, 1.5)
        Phi, Phib = rng.uniform(0, 1, 2)
        s = L.oracle_quark_distribution(E, mu, T, Phi, Phib) + L.oracle_antiquark_distribution(-E, mu, T, Phi, Phib)
        assert abs(s - 1.0) < 1e-12


def test_log_term_both_branches_agree(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for _ in range(300):
        mu, T = rng.uniform(0.5, 2), rng.uniform(0.1, 1.0)
        Phi, Phib = rng.uniform(0, 1, 2)
        E = mu + rng.uniform(-2, 2) * T
        a, b = -(E - mu) / T, -(E + mu) / T
        direct = math.log(1 + 3 * Phi * math.exp(a) + 3 * Phib * math.exp(2 * a) + math.exp(3 * a)) + \
            math.log(1 + 3 * Phib * math.exp(b) + 3 * Phi * math.exp(2 * b) + math.exp(3 * b))
        assert abs(L.oracle_log_term(E, mu, T, Phi, Phib) - direct) < 1e-13 * max(1, abs(direct))
    # continuity across a = 0 (the LSE switch)
    mu, T, Phi, Phib = 1.2, 0.4, 0.3, 0.5
    lo = L.oracle_log_term(mu * (1 + 1e-15), mu, T, Phi, Phib)
    hi = L.oracle_log_term(mu * (1 - 1e-15), mu, T, Phi, Phib)
    assert abs(lo - hi) < 1e-13


# ---------------------------------------------------------------------------
# gradient / Hessian of Omega vs finite differences
# ---------------------------------------------------------------------------
STATES = [
    # (T, mu, xi, x)
    (100 / HB, 200 / HB, 0.0, [-1.84, -1.84, -2.22, 0.02, 0.03]),
    (180 / HB, 100 / HB, 0.3, [-0.9, -0.9, -1.85, 0.45, 0.5]),
    (60 / HB, 380 / HB, -0.2, [-0.3, -0.3, -2.1, 0.05, 0.09]),   # a > 0 nodes (LSE branch)
    (250 / HB, 350 / HB, 0.4, [-0.2, -0.25, -1.2, 0.8, 0.85]),   # phi_u != phi_d
]


def _fd_grad(oracle, x, mu, T, xi, pn, tn, h):
    g = np.zeros(5)
    for i in range(5):
        e = np.zeros(5)
        e[i] = h
        f = lambda s: oracle.omega(np.array(x) + s * e, mu, T, xi, pn, tn)
        g[i] = (8 * (f(1) - f(-1)) - (f(2) - f(-2))) / (12 * h)
    return g


@pytest.mark.parametrize("state", STATES)
def test_gradient_and_hessian_vs_fd(oracle, state):
    T, mu, xi, x = state
    x = np.array(x, dtype=float)
    om, g, H = oracle.omega_derivs(x, mu, T, xi, 24, 4)
    assert abs(om - oracle.omega(x, mu, T, xi, 24, 4)) == 0.0
    gfd = _fd_grad(oracle, x, mu, T, xi, 24, 4, 1e-4)
    assert np.allclose(g, gfd, rtol=1e-7, atol=1e-9)
    # Hessian: central differences of the analytic gradient
    Hfd = np.zeros((5, 5))
    h = 1e-5
    for j in range(5):
        e = np.zeros(5)
        e[j] = h
        Hfd[:, j] = (oracle.omega_derivs(x + e, mu, T, xi, 24, 4)[1] - oracle.omega_derivs(x - e, mu, T, xi, 24, 4)[1]) / (2 * h)
    assert np.allclose(H, Hfd, rtol=1e-6, atol=1e-8)
    assert np.allclose(H, H.T, rtol=0, atol=1e-12 * np.abs(H).max())


def test_mu_vector_gradient_is_density(oracle):
    """rho_i = dP/dmu_i = n_q,i - n_qbar,i (Thermodynamics.jl:215-220 vs :255-281)."""
    x = np.array([-0.9, -0.9, -1.85, 0.45, 0.5])
    T, mu, xi = 160 / HB, 250 / HB, 0.2
    th = oracle.thermo(x, T, mu, xi, 24, 8)
    rho = th["n_q"] - th["n_qbar"]
    assert abs(rho.sum() / (3 * 0.16) - th["rho_norm"]) < 1e-13
    h = 1e-5
    for i in range(3):
        mv = np.array([mu, mu, mu])
        mp, mm = mv.copy(), mv.copy()
        mp[i] += h
        mm[i] -= h
        dP = -(oracle.omega_derivs(x, mp, T, xi, 24, 8)[0] - oracle.omega_derivs(x, mm, T, xi, 24, 8)[0]) / (2 * h)
        assert abs(dP - rho[i]) < 1e-8
    # entropy = dP/dT at fixed x; P = -Omega; energy = -P + sum mu rho + T s
    dPdT = -(oracle.omega(x, mu, T + h, xi, 24, 8) - oracle.omega(x, mu, T - h, xi, 24, 8)) / (2 * h)
    assert abs(dPdT - th["entropy"]) < 1e-8 * max(1, abs(dPdT))
    assert th["pressure"] == -th["omega"]
    assert abs(th["energy"] - (-th["pressure"] + mu * rho.sum() + T * th["entropy"])) < 1e-13 * abs(th["energy"])


def test_polyakov_potential_symmetry(oracle):
    """U(Phi, Phibar) = U(Phibar, Phi) (test_core_thermodynamics.jl:80-84): Omega at mu=0 is symmetric."""
    x1 = np.array([-1.0, -1.0, -2.0, 0.2, 0.6])
    x2 = np.array([-1.0, -1.0, -2.0, 0.6, 0.2])
    T = 150 / HB
    assert abs(oracle.omega(x1, 0.0, T, 0.0, 24, 4) - oracle.omega(x2, 0.0, T, 0.0, 24, 4)) < 1e-13


def test_xi_zero_independent_of_cos_nodes(oracle):
    x = np.array([-1.2, -1.2, -2.0, 0.3, 0.35])
    T, mu = 140 / HB, 200 / HB
    a = oracle.omega(x, mu, T, 0.0, 32, 8)
    b = oracle.omega(x, mu, T, 0.0, 32, 3)
    assert abs(a - b) < 1e-13 * abs(a)
    # cos(theta) = 0 direction: xi only enters through (p cos)^2 -> small-xi continuity
    c = oracle.omega(x, mu, T, 1e-12, 32, 8)
    assert abs(a - c) < 1e-12 * abs(a)


# ---------------------------------------------------------------------------
# roots against data the reference produced
# ---------------------------------------------------------------------------
def test_dual_branch_T50_omega_and_roots(oracle):
    """data/outputs/results/pnjl/dual_branch_T50.csv: T=50 MeV, xi=0, p24/t8, 10 decimals."""
    rows = golden_rows("dual_branch_T50.csv")
    T = np.array([float(r["T_MeV"]) for r in rows]) / HB
    mu = np.array([float(r["mu_MeV"]) for r in rows]) / HB
    x = np.array([[float(r[k]) for k in ("phi_u", "phi_d", "phi_s", "Phi1", "Phi2")] for r in rows])
    om = np.array([float(r["omega"]) for r in rows])
    mine = np.array([oracle.omega(x[i], mu[i], T[i], 0.0, 24, 8) for i in range(len(rows))])
    assert np.max(np.abs(mine - om)) < 1e-10          # stored to 10 decimals
    r = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_EXPLICIT, seed=x, p_num=24, t_num=8,
                           opts=oracle.default_opts(ftol=1e-12))
    assert r["converged"].all()
    assert np.max(np.abs(r["x"] - x)) < 3e-9         # 10-decimal rounding + the reference's 1e-9 gate
    assert np.max(np.abs(r["omega"] - om)) < 1e-10
    # the branch label: hadron rows have M_u > 250 MeV, quark rows below
    Mu = r["masses"][:, 0] * HB
    for i, row in enumerate(rows):
        assert (Mu[i] > 250) == (row["branch"] == "hadron")


def _driver_rows(name):
    rows = golden_rows(name)
    keys = sorted(set((float(r["xi"]), float(r["muB_MeV"])) for r in rows))
    order = []
    L = None
    for k in keys:
        sel = sorted([i for i, r in enumerate(rows) if (float(r["xi"]), float(r["muB_MeV"])) == k],
                     key=lambda i: float(rows[i]["T_MeV"]))
        L = len(sel)
        order += sel
    return [rows[i] for i in order], L


def load_boundary_into(oracle):
    import os
    import julia_relaxtime_amd.pnjl as P
    rows, cep, _ = P.load_phase_boundary(0.0)
    oracle.set_boundary(rows, cep)


@pytest.mark.parametrize("name", ["driver_gap_scan.csv", "driver_gap_scan_xi0p2.csv",
                                  "driver_gap_scan_xi-0p6to0p6.csv"])
def test_driver_chains_reproduce_reference_scan(oracle, name):
    """run_gap_transport_scan.jl chains (PhaseAwareContinuitySeed, first point
    MultiSeed, p12/t6, iterations=40) reproduce every stored row."""
    load_boundary_into(oracle)
    rows, L = _driver_rows(name)
    f = lambda k: np.array([float(r[k]) for r in rows])
    T, mu, xi = f("T_fm"), f("muq_fm"), f("xi")
    r = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_PHASE_AWARE, p_num=12, t_num=6,
                           opts=oracle.default_opts(max_iter=40, chain_len=L))
    assert r["converged"].all()
    om = f("omega_fm4inv")
    assert np.max(np.abs(r["omega"] - om) / np.abs(om)) < 1e-12
    M = np.stack([f("m_u"), f("m_d"), f("m_s")], 1)
    assert np.max(np.abs(r["masses"] - M)) < 1e-8      # stored residuals reach 1e-9
    assert np.max(np.abs(r["x"][:, 3] - f("Phi"))) < 1e-8
    assert np.max(np.abs(r["x"][:, 4] - f("Phibar"))) < 1e-8
    assert np.max(np.abs(r["entropy"] - f("s_fm3inv"))) < 1e-8
    assert np.max(np.abs(r["energy"] - f("epsilon_fm4inv"))) < 1e-8
    assert np.max(np.abs(r["rho_norm"] - f("rho_norm"))) < 1e-8
    nq = np.stack([f("n_u"), f("n_d"), f("n_s")], 1)
    nqb = np.stack([f("n_ubar"), f("n_dbar"), f("n_sbar")], 1)
    assert np.max(np.abs(r["n_q"] - nq)) < 1e-8
    assert np.max(np.abs(r["n_qbar"] - nqb)) < 1e-8
    # NLsolve iteration counts: equal except where the stored residual sits at the ftol edge
    assert np.mean(r["iterations"] == f("iterations").astype(int)) > 0.95


def test_tmu_scan_coarse_values(oracle):
    """tmu_scan.csv (6 decimals): the roots of the stored points are roots here."""
    rows = golden_rows("tmu_scan.csv")[::7]
    T = np.array([float(r["T_MeV"]) for r in rows]) / HB
    mu = np.array([float(r["mu_MeV"]) for r in rows]) / HB
    x = np.array([[float(r[k]) for k in ("phi_u", "phi_d", "phi_s", "Phi1", "Phi2")] for r in rows])
    r = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_EXPLICIT, seed=x, p_num=24, t_num=8)
    ok = r["converged"].astype(bool)
    assert ok.mean() > 0.95
    assert np.max(np.abs(r["x"][ok] - x[ok])) < 5e-5
    P = np.array([float(r2["pressure_fm4"]) for r2 in rows])
    assert np.max(np.abs(r["pressure"][ok] - P[ok])) < 5e-6


def test_trho_scan_states_are_fixedmu_roots(oracle):
    """trho_scan_xi{0.0,0.2,0.4}.csv (TrhoScan, p24/t8, 6 decimals) at T = 50,
    80, 110 MeV: each FixedRho state is a FixedMu root at mu = mu_avg, on the
    hadron, unstable and quark branches of the S-curve alike; P, rho, s and
    epsilon of the root match the stored columns."""
    rows = golden_rows("trho_scan_T50_80_110.csv")
    f = lambda k: np.array([float(r[k]) for r in rows])
    T, mu, xi = f("T_MeV") / HB, f("mu_avg_MeV") / HB, f("xi")
    x = np.stack([f(k) for k in ("phi_u", "phi_d", "phi_s", "Phi1", "Phi2")], 1)
    r = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_EXPLICIT, seed=x, p_num=24, t_num=8, nthreads=8,
                           opts=oracle.default_opts(ftol=1e-12))
    assert r["converged"].all()
    dx = np.abs(r["x"] - x).max(1)
    assert np.median(dx) < 1e-6 and dx.max() < 5e-5     # 6-decimal mu, ill-conditioned near spinodals
    assert np.max(np.abs(r["pressure"] - f("pressure_fm4"))) < 1e-6
    assert np.max(np.abs(r["rho_norm"] - f("rho"))) < 1e-4
    assert np.max(np.abs(r["entropy"] - f("entropy_fm3"))) < 1e-5
    assert np.max(np.abs(r["energy"] - f("energy_fm4"))) < 1e-4
    assert np.max(np.abs(r["masses"][:, 0] * HB - f("M_u_MeV"))) < 1e-2
    # the unstable branch is covered: d mu / d rho < 0 between the spinodals
    for key in {(a, b) for a, b in zip(f("xi"), f("T_MeV"))}:
        sel = (f("xi") == key[0]) & (f("T_MeV") == key[1])
        o = np.argsort(f("rho")[sel])
        if key[1] <= 80:
            assert np.any(np.diff(f("mu_avg_MeV")[sel][o]) < 0)


def _T_cep(xi):
    """T_CEP (MeV) of the package's cep.csv, the reference's data/reference/pnjl/cep.csv."""
    import julia_relaxtime_amd.pnjl as P
    return P.load_phase_boundary(xi)[2][2]


def test_maxwell_transition_matches_reference_boundary(oracle):
    """boundary.csv (Maxwell construction on the T-rho S-curves): at mu_transition
    the hadron and quark roots have equal Omega to within 0.03 MeV of mu, and
    their densities match the stored rho_hadron / rho_quark."""
    rows = [r for r in golden_rows("phase_boundary.csv")]
    checked = 0
    for r in rows:
        xi, Tm, mt = float(r["xi"]), float(r["T_MeV"]), float(r["mu_transition_MeV"])
        if Tm > _T_cep(xi) - 10:
            continue
        mus = np.array([mt - 0.2, mt, mt + 0.2]) / HB
        Ts = np.full(3, Tm / HB)
        h = oracle.solve_batch(Ts, mus, xi, policy=oracle.SEED_HADRON, p_num=24, t_num=8)
        q = oracle.solve_batch(Ts, mus, xi, policy=oracle.SEED_QUARK, p_num=24, t_num=8)
        assert h["converged"].all() and q["converged"].all()
        assert np.all(h["masses"][:, 0] - q["masses"][:, 0] > 50 / HB)      # two distinct branches
        d = h["omega"] - q["omega"]
        assert d[0] < 0 < d[2]                                             # hadron stable below, quark above
        mu_star = mus[1] - d[1] * (mus[2] - mus[0]) / (d[2] - d[0])
        assert abs(mu_star * HB - mt) < 0.03, (xi, Tm, mu_star * HB - mt)
        assert abs(h["rho_norm"][1] - float(r["rho_hadron"])) < 0.01
        assert abs(q["rho_norm"][1] - float(r["rho_quark"])) < 0.01
        checked += 1
    assert checked >= 30


def _branch_end(oracle, T, xi, mus, first_policy):
    """Follows one branch by continuation along mus; returns the first mu at
    which the root leaves it (|delta M_u| > 20 MeV or no convergence)."""
    x, M_prev = None, None
    for m in mus:
        if x is None:
            s = oracle.solve_batch(np.array([T]), np.array([m]), xi, policy=first_policy, p_num=24, t_num=8)
        else:
            s = oracle.solve_batch(np.array([T]), np.array([m]), xi, policy=oracle.SEED_EXPLICIT, seed=x[None],
                                   p_num=24, t_num=8, opts=oracle.default_opts(auto_multiseed_fallback=0))
        M = s["masses"][0, 0] * HB
        if not s["converged"][0] or (M_prev is not None and abs(M - M_prev) > 20):
            return m
        x, M_prev = s["x"][0], M
    return None


@pytest.mark.parametrize("xi", [0.0, 0.2, 0.4])
def test_spinodals_match_reference(oracle, xi):
    """spinodals.csv: following the hadron branch up in mu (the quark branch
    down) by continuation, the branch ends within 0.1 MeV of the stored
    spinodal, for T at least 15 MeV below the CEP."""
    rows = [r for r in golden_rows("spinodals.csv") if float(r["xi"]) == xi]
    Tc = _T_cep(xi)
    rows = [r for r in rows if float(r["T_MeV"]) <= Tc - 15][::3]
    assert len(rows) >= 4
    step = 0.05
    for r in rows:
        T = float(r["T_MeV"]) / HB
        mh, mq = float(r["mu_spinodal_hadron_MeV"]), float(r["mu_spinodal_quark_MeV"])
        up = np.arange(mh - 1.5, mh + 1.5, step) / HB
        end_h = _branch_end(oracle, T, xi, up, oracle.SEED_HADRON)
        down = np.arange(mq + 1.5, mq - 1.5, -step) / HB
        end_q = _branch_end(oracle, T, xi, down, oracle.SEED_QUARK)
        assert end_h is not None and abs(end_h * HB - mh) < 2 * step + 0.02, (xi, r["T_MeV"], end_h * HB, mh)
        assert end_q is not None and abs(end_q * HB - mq) < 2 * step + 0.02, (xi, r["T_MeV"], end_q * HB, mq)


def test_seed_policy_semantics(oracle):
    """DefaultSeed auto hint / VERY_HIGH_TEMP switch, MultiSeed argmin Omega."""
    T = np.array([100, 160, 100, 320, 50]) / HB
    mu = np.array([100, 100, 320, 0, 350]) / HB
    r = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_DEFAULT, p_num=24, t_num=8)
    assert r["converged"].all()
    m = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_MULTI, p_num=24, t_num=8)
    assert m["converged"].all()
    # MultiSeed returns the lowest-Omega converged root, never above the default's
    assert np.all(m["omega"] <= r["omega"] + 1e-12)
    # in the coexistence region (T=50, mu=350 MeV) hadron and quark roots both exist
    h = oracle.solve_batch(T[4:], mu[4:], 0.0, policy=oracle.SEED_HADRON, p_num=24, t_num=8)
    q = oracle.solve_batch(T[4:], mu[4:], 0.0, policy=oracle.SEED_QUARK, p_num=24, t_num=8)
    assert h["converged"][0] and q["converged"][0]
    assert abs(h["masses"][0, 0] - q["masses"][0, 0]) > 0.5
    assert m["omega"][4] == min(h["omega"][0], q["omega"][0])


def test_trust_region_fallback_and_exception(oracle):
    """Unphysical seed forces the fallback chain; a NaN seed is the reference's
    IsFiniteException (ImplicitSolver.jl:112, :130-150)."""
    T = np.array([150 / HB])
    mu = np.array([100 / HB])
    bad = np.array([[-0.5, -0.5, -1.0, 1.5, 1.5]])
    r = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_EXPLICIT, seed=bad, p_num=24, t_num=8)
    assert r["converged"][0]
    nan = np.array([[np.nan, np.nan, -1.0, 0.1, 0.1]])
    r = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_EXPLICIT, seed=nan, p_num=24, t_num=8)
    assert r["flags"][0] & 0x04 and not r["converged"][0]
    # iterations=1 cannot converge from a far seed -> fallbacks engage
    r1 = oracle.solve_batch(T, mu, 0.0, policy=oracle.SEED_HADRON, p_num=24, t_num=8,
                            opts=oracle.default_opts(max_iter=1, auto_multiseed_fallback=0,
                                                     trust_region_fallback=0))
    assert not r1["converged"][0] and r1["iterations"][0] == 1
