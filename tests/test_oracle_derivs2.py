"""The oracle's second-order implicit derivatives (oracle_implicit_derivs2:
x_ab = -H^-1 [(D_wb H) x_a + D_wb F_a] with the directional derivatives by
complex step on the closed forms) pinned by what the definition fixes
independently of any third derivative of Omega: central differences, with
Richardson extrapolation, of the oracle's first derivatives (dx/dT, dx/dmu,
dM/dT, dM/dmu at a root: already pinned against the reference's derivative
CSVs) along roots re-solved at T +- h and mu +- h.  That is the reference's
own definition of the second order (ForwardDiff of the implicit solve's
derivative, ImplicitSolver.jl:679-697, ThermoDerivatives.jl:153-172).  The
mixed derivative is checked both ways (d/dmu of x_T and d/dT of x_mu).
Tolerance 1e-6 relative (floored at 1e-3 of the block's largest value): the
differences agree to 1e-9 - 1e-12 at relative step 1e-4, while a wrong term
or sign shows at O(1)."""
import numpy as np
import pytest

HB = 197.327
POINTS = [(150, 100, 0.0), (120, 300, 0.2), (200, 50, -0.3), (60, 250, 0.0), (100, 330, 0.0), (250, 400, 0.4),
          (170, 200, -0.1), (40, 320, 0.3)]


def _resolve(oracle, x0, T, mu, xi, pn, tn):
    o = oracle.default_opts(ftol=1e-14, xtol=1e-16, max_iter=60, trust_region_fallback=0,
                            auto_multiseed_fallback=0)
    r = oracle.solve_batch(np.array([T]), np.array([mu]), np.array([xi]), policy=oracle.SEED_EXPLICIT,
                           seed=np.array([x0]), p_num=pn, t_num=tn, opts=o)
    return r["x"][0]


def _first(oracle, x, T, mu, xi, pn, tn):
    d = oracle.thermo_derivs(x, T, mu, xi, pn, tn)
    return np.r_[d["dx_dT"], d["dx_dmu"], d["dM_dT"], d["dM_dmu"]]


def _richardson(oracle, x, T, mu, xi, var, pn, tn, rel=1e-4):
    base = T if var == "T" else mu
    h0 = rel * max(abs(base), 0.05)
    D = []
    for h in (h0, h0 / 2):
        f = []
        for s in (1, -1):
            Tp, mp = (T + s * h, mu) if var == "T" else (T, mu + s * h)
            f.append(_first(oracle, _resolve(oracle, x, Tp, mp, xi, pn, tn), Tp, mp, xi, pn, tn))
        D.append((f[0] - f[1]) / (2 * h))
    return (4 * D[1] - D[0]) / 3


@pytest.mark.parametrize("pn,tn", [(64, 8), (24, 4)])
def test_second_order_against_differences_of_resolved_roots(oracle, pn, tn):
    worst = 0.0
    for Tm, mm, xi in POINTS:
        T, mu = Tm / HB, mm / HB
        r = oracle.solve_batch(np.array([T]), np.array([mu]), np.array([xi]), policy=oracle.SEED_DEFAULT,
                               p_num=pn, t_num=tn, opts=oracle.default_opts(ftol=1e-10, max_iter=100))
        assert r["converged"][0] == 1
        x = _resolve(oracle, r["x"][0], T, mu, xi, pn, tn)
        c = oracle.implicit_derivs2(x, T, mu, xi, pn, tn)
        fT = _richardson(oracle, x, T, mu, xi, "T", pn, tn)
        fm = _richardson(oracle, x, T, mu, xi, "mu", pn, tn)
        checks = {"d2x_dT2": (fT[0:5], c["d2x_dT2"]), "d2x_dTdmu": (fm[0:5], c["d2x_dTdmu"]),
                  "d2x_dmudT": (fT[5:10], c["d2x_dTdmu"]), "d2x_dmu2": (fm[5:10], c["d2x_dmu2"]),
                  "d2M_dT2": (fT[10:13], c["d2M_dT2"]), "d2M_dTdmu": (fm[10:13], c["d2M_dTdmu"]),
                  "d2M_dmudT": (fT[13:16], c["d2M_dTdmu"]), "d2M_dmu2": (fm[13:16], c["d2M_dmu2"])}
        for k, (fd, cs) in checks.items():
            sc = np.maximum(np.abs(cs), 1e-3 * np.abs(cs).max() + 1e-12)
            err = float(np.max(np.abs(fd - cs) / sc))
            worst = max(worst, err)
            assert err < 1e-6, (Tm, mm, xi, k, fd, cs)
        # the first-order blocks are the first-order routine's (the mass chain
        # rule in another operation order)
        d1 = oracle.thermo_derivs(x, T, mu, xi, pn, tn)
        assert np.array_equal(c["dx_dT"], d1["dx_dT"]) and np.array_equal(c["dx_dmu"], d1["dx_dmu"])
        for k in ("masses", "dM_dT", "dM_dmu"):
            assert np.allclose(c[k], d1[k], rtol=1e-13, atol=1e-15 * np.abs(d1[k]).max()), k
    print(f"[oracle order 2 p{pn} t{tn}] worst relative difference to the extrapolated differences {worst:.1e}")
