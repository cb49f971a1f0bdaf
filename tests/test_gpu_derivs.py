"""GPU implicit thermodynamic derivatives (pnjl_thermo_derivatives_batch via
the C ABI) against the oracle's restatement on the same roots, and against the
reference's bulk_viscosity_xi0.0.csv through the host mirror of
ThermoDerivatives.jl."""
import numpy as np
import pytest

from conftest import golden_rows

pytestmark = pytest.mark.gpu
HB = 197.327


def _roots(oracle, n, seed, pn, tn):
    rng = np.random.default_rng(seed)
    T = rng.uniform(30, 300, n) / HB
    mu = rng.uniform(0, 420, n) / HB
    xi = rng.choice([0.0, 0.3, -0.2], n)
    r = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_DEFAULT, p_num=pn, t_num=tn, nthreads=8,
                           opts=oracle.default_opts(ftol=1e-12))
    ok = r["converged"] == 1
    return T[ok], mu[ok], xi[ok], r["x"][ok]


# the ratio columns: (name, numerator terms, denominator terms) as functions of the primitive columns
def _ratio_terms(d, T):
    s, n, ntot = d["entropy"], d["rho"], d["dP_dmu"]
    ET, Emu, nT, nmu = d["dEpsilon_dT"], d["dEpsilon_dmu"], d["dn_dT"], d["dn_dmu"]
    sT, smuB, nmuB = d["ds_dT"], d["ds_dmu"] / 3, nmu / 3
    return {
        "dP_depsilon_n": ((s * nmu, ntot * nT), (ET * nmu, Emu * nT)),
        "dP_dn_epsilon": ((s * Emu, ntot * ET), (nT * Emu, nmu * ET)),
        "v_n_sq": ((s * nmuB, n * nT), (T * sT * nmuB, T * smuB * nT)),
        "dmuB_dT_sigma": ((n * sT, s * nT), (n * smuB, s * nmuB)),
    }


@pytest.mark.parametrize("pn,tn", [(24, 8), (64, 4)])
def test_derivatives_match_oracle(gpu_ctx, oracle, pn, tn):
    """Every primitive column at 1e-9 relative (scale max(|oracle|, column
    typical size)); the implicit derivatives inherit the Hessian's condition,
    so roots next to a spinodal (|dx/dtheta| > 1e3) are held to 1e-7.  The
    ratio columns (dP/deps|n, dP/dn|eps, v_n^2, dmu_B/dT|sigma) are differences
    of products over differences of products: the GPU's own primitives must
    give them to 1e-12, and the oracle's to 1e-9 divided by the cancellation
    |a - b| / (|a| + |b|) of their numerator and denominator."""
    T, mu, xi, x = _roots(oracle, 300, 3, pn, tn)
    # three-flavour states too: phi_d != phi_u exercises the separate d sums
    x2 = x[:40].copy()
    x2[:, 1] *= 1.01
    T, mu, xi, x = np.r_[T, T[:40]], np.r_[mu, mu[:40]], np.r_[xi, xi[:40]], np.r_[x, x2]
    g = gpu_ctx.thermo_derivatives_batch(x, T, mu, xi, pn, tn)
    o = [oracle.thermo_derivs(x[i], T[i], mu[i], xi[i], pn, tn) for i in range(len(T))]
    stiff = np.array([max(np.abs(oi["dx_dT"]).max(), np.abs(oi["dx_dmu"]).max()) > 1e3 for oi in o])
    ratio = _ratio_terms(g, T)
    # where the entropy itself is a rounding-level remainder of cancelling
    # O(1e-9) terms (hadronic points at T ~ 30 MeV: s ~ 1e-12) the ratios built
    # from it carry that noise on both sides and are not compared
    s_ok = np.abs(g["entropy"]) > 1e-6 * np.median(np.abs(g["entropy"]))
    for k, ((a1, a2), (b1, b2)) in ratio.items():
        mine = (a1 - a2) / (b1 - b2)
        if k == "dmuB_dT_sigma":
            mine = -mine
        fin = np.isfinite(mine)
        assert np.all(np.abs(g[k][fin] - mine[fin]) <= 1e-12 * np.abs(mine[fin])), k
        ov = np.array([oi[k] for oi in o])
        canc = np.minimum(np.abs(a1 - a2) / (np.abs(a1) + np.abs(a2)), np.abs(b1 - b2) / (np.abs(b1) + np.abs(b2)))
        tol = np.minimum(1e-9 / np.maximum(canc, 1e-300), 1.0)
        use = np.isfinite(ov) & (canc > 1e-6) & s_ok   # 0/0 at mu = 0 (n = 0) is noise on both sides
        assert np.all(np.abs(g[k][use] - ov[use]) <= tol[use] * np.abs(ov[use])), k
    for k in g:
        if k in ratio:
            continue
        ov = np.array([oi[k] for oi in o])
        gv = g[k]
        scale = np.maximum(np.abs(ov), np.nanmedian(np.abs(ov)) if np.isfinite(ov).any() else 1.0)
        tol = np.where(stiff.reshape((-1,) + (1,) * (ov.ndim - 1)), 1e-7, 1e-9)
        both_nan = np.isnan(gv) & np.isnan(ov)
        ok = both_nan | (np.abs(gv - ov) <= tol * scale)
        assert ok.all(), (k, np.flatnonzero(~ok.reshape(len(T), -1).all(1))[:5])


def test_bulk_viscosity_coefficients_match_reference_output(gpu_ctx):
    from julia_relaxtime_amd import thermo_derivatives as TD
    rows = [r for r in golden_rows("bulk_viscosity_xi0.0.csv") if float(r["T_MeV"]) == 150.0]
    for row in rows:
        T, mu = float(row["T_MeV"]) / HB, float(row["mu_MeV"]) / HB
        b = TD.bulk_viscosity_coefficients(T, mu, xi=0.0, p_num=24, t_num=8)
        ref = lambda k: float(row[k])
        rel = lambda a, c: abs(a - c) <= 2e-6 * abs(c)
        assert rel(b["v_n_sq"], ref("v_n_sq"))
        assert rel(b["s"], ref("entropy_fm3"))
        for i, f in enumerate("uds"):
            assert abs(b["masses"][i] * HB - ref(f"M_{f}_MeV")) < 2e-6
            assert rel(b["dM_dT"][i], ref(f"dM_{f}_dT"))
            if float(row["mu_MeV"]) > 0:
                assert rel(b["dM_dmuB"][i], ref(f"dM_{f}_dmuB"))
        if float(row["mu_MeV"]) > 0:
            assert rel(b["dmuB_dT_sigma"], ref("dmuB_dT_sigma"))
            assert rel(b["n_B"], ref("n_B_fm3"))


def test_host_api_is_consistent(gpu_ctx):
    """thermo_derivatives / dP_dT / dP_dmu / mass_derivatives agree with the
    batch call and with each other; unsupported orders raise."""
    from julia_relaxtime_amd import thermo_derivatives as TD
    T, mu = 130 / HB, 320 / HB
    d = TD.thermo_derivatives(T, mu, xi=0.0, p_num=48, t_num=12)
    assert d["converged"]
    assert TD.dP_dT(T, mu, p_num=48, t_num=12) == d["dP_dT"] == d["entropy"]
    assert TD.dP_dmu(T, mu, p_num=48, t_num=12) == d["dP_dmu"]
    b = TD.thermo_derivatives_batch([T], [mu], p_num=48, t_num=12)
    assert TD.dP_dT(T, mu, order=2, p_num=48, t_num=12) == b["ds_dT"][0]
    m = TD.mass_derivatives(T, mu, p_num=48, t_num=12)
    assert np.array_equal(m["dM_dT"], d["dM_dT"])
    m2 = TD.mass_derivatives(T, mu, order=2, p_num=48, t_num=12)
    assert np.allclose(m2["dM_dT"], d["dM_dT"], rtol=1e-10) and np.allclose(m2["masses"], d["masses"], rtol=1e-14)
    assert set(m2) == {"masses", "dM_dT", "dM_dmu", "d2M_dT2", "d2M_dTdmu", "d2M_dmu2"}
    s2 = TD.solve_with_derivatives(T, mu, order=2, p_num=48, t_num=12)
    assert set(s2) == {"x", "dx_dT", "dx_dmu", "d2x_dT2", "d2x_dmu2", "d2x_dTdmu"}
    assert set(TD.solve_with_derivatives(T, mu, p_num=48, t_num=12)) == {"x", "dx_dT", "dx_dmu"}
    with pytest.raises(ValueError):
        TD.mass_derivatives(T, mu, order=3)
    with pytest.raises(ValueError):
        TD.solve_with_derivatives(T, mu, order=0)
    with pytest.raises(NotImplementedError):
        TD.dP_dT(T, mu, order=3)


def test_derivatives_match_reference_output(gpu_ctx):
    """derivatives_xi0.0.csv rows at T = 150 MeV through the host mirror of
    ThermoDerivatives.jl (solve with DefaultSeed, then the GPU derivative pass)."""
    from julia_relaxtime_amd import thermo_derivatives as TD
    from test_oracle_derivs import check_derivative_row
    rows = [r for r in golden_rows("derivatives_xi0.0.csv") if float(r["T_MeV"]) == 150.0]
    for row in rows:
        d = TD.thermo_derivatives(float(row["T_MeV"]) / HB, float(row["mu_MeV"]) / HB, xi=0.0, p_num=24, t_num=8)
        assert d["converged"]
        check_derivative_row(d, row)


@pytest.mark.parametrize("pn,tn", [(24, 8), (64, 4)])
def test_second_order_derivatives_match_oracle(gpu_ctx, oracle, pn, tn):
    """pnjl_implicit_derivatives2_batch (dual numbers over the closed-form
    Hessian and mixed terms) against the oracle's complex-step restatement on
    the same roots: x and the masses with their first and second (T, mu)
    derivatives (solve_with_derivatives order 2, ImplicitSolver.jl:660-701;
    mass_derivatives order 2, ThermoDerivatives.jl:153-172).  1e-9 relative
    to max(|oracle|, 1e-3 of the block's largest value); roots next to a
    spinodal (|dx/dtheta| > 1e3: the second derivatives carry the Hessian's
    condition twice) at 1e-6.  Three-flavour states (phi_d != phi_u) included."""
    T, mu, xi, x = _roots(oracle, 200, 5, pn, tn)
    x2 = x[:30].copy()
    x2[:, 1] *= 1.01
    T, mu, xi, x = np.r_[T, T[:30]], np.r_[mu, mu[:30]], np.r_[xi, xi[:30]], np.r_[x, x2]
    g = gpu_ctx.implicit_derivatives2_batch(x, T, mu, xi, pn, tn)
    o = [oracle.implicit_derivs2(x[i], T[i], mu[i], xi[i], pn, tn) for i in range(len(T))]
    stiff = np.array([max(np.abs(oi["dx_dT"]).max(), np.abs(oi["dx_dmu"]).max()) > 1e3 for oi in o])
    worst = {}
    for k in g:
        ov = np.array([oi[k] for oi in o])
        sc = np.maximum(np.abs(ov), 1e-3 * np.abs(ov).max(axis=1, keepdims=True))
        err = np.max(np.abs(g[k] - ov) / sc, axis=1)
        tol = np.where(stiff, 1e-6, 1e-9)
        worst[k] = float(err[~stiff].max()) if (~stiff).any() else 0.0
        assert np.all(err <= tol), (k, np.flatnonzero(err > tol)[:5], err.max())
    print(f"[deriv2 p{pn} t{tn}] {len(T)} roots ({int(stiff.sum())} next to a spinodal); worst relative "
          f"difference per block: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()), flush=True)
