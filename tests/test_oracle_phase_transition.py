"""The first-order transition analysis (julia_relaxtime_amd.phase_transition):
S-shape detection and the Maxwell construction against the reference's unit
tests (tests/unit/pnjl/test_phase_transition.jl), against the reference's own
phase-diagram outputs -- boundary.csv and spinodals.csv, which
calculate_phase_structure.jl computed with these functions from the TrhoScan
curves kept in tests/golden/ -- and against what the mathematics fixes (an
antisymmetric cubic S-curve, equal pressures at mu_c).  The GPU S-curve
logic (`s_curve`) is driven here by the oracle's solves."""
import math

import numpy as np
import pytest

from conftest import golden_rows

HB = 197.327
MAXWELL_KW = dict(min_samples=8, detect_min_points=5, detect_eps=1e-6, candidate_steps=64, max_iter=60,
                  tol_area=1e-4)          # step3_maxwell_construction  calculate_phase_structure.jl:463-470


def PT():
    from julia_relaxtime_amd import phase_transition
    return phase_transition


def s_shape_curve():
    rho = [0.0, 0.2, 0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6, 1.8, 2.0]
    mu = [250.0, 260.0, 270.0, 275.0, 272.0, 268.0, 265.0, 270.0, 280.0, 290.0, 300.0]
    return mu, rho


def monotonic_curve():
    return [250.0, 260.0, 270.0, 280.0, 290.0], [0.0, 0.5, 1.0, 1.5, 2.0]


def synthetic_maxwell_curve():
    rho = [0.1 * i for i in range(1, 16)]
    mu = [250.0, 255.0, 260.0, 265.0, 268.0, 266.0, 262.0, 258.0, 260.0, 265.0, 272.0, 280.0, 288.0, 296.0, 305.0]
    return mu, rho


def test_reference_unit_cases():
    """test_phase_transition.jl:38-150, assertion for assertion."""
    P = PT()
    r = P.detect_s_shape(*s_shape_curve())
    assert r.has_s_shape and None not in (r.mu_spinodal_hadron, r.mu_spinodal_quark, r.rho_spinodal_hadron,
                                          r.rho_spinodal_quark)
    assert r.mu_spinodal_hadron > r.mu_spinodal_quark and r.rho_spinodal_hadron < r.rho_spinodal_quark
    r = P.detect_s_shape(*monotonic_curve())
    assert not r.has_s_shape and r.mu_spinodal_hadron is None and r.mu_spinodal_quark is None
    assert not P.detect_s_shape([250.0, 260.0], [0.0, 1.0], min_points=5).has_s_shape
    m = P.maxwell_construction(*synthetic_maxwell_curve(), min_samples=8)
    assert m.converged and m.rho_hadron < m.rho_quark and m.area_residual < 0.1
    hint = P.detect_s_shape(*synthetic_maxwell_curve())
    assert P.maxwell_construction(*synthetic_maxwell_curve(), spinodal_hint=hint, min_samples=8).converged
    m = P.maxwell_construction(*monotonic_curve(), min_samples=3)
    assert not m.converged and m.details["reason"] == "no_s_shape"
    m = P.maxwell_construction([250.0, 260.0, 270.0], [0.0, 0.5, 1.0], min_samples=12)
    assert not m.converged and m.details["reason"] == "insufficient_points"
    rows = [dict(T_MeV="70", xi="0.0", rho="0.1", mu_avg_MeV="250.0"),
            dict(T_MeV="70", xi="0.0", rho="0.2", mu_avg_MeV="260.0"),
            dict(T_MeV="80", xi="0.0", rho="0.1", mu_avg_MeV="255.0"),
            dict(T_MeV="70", xi="0.5", rho="0.1", mu_avg_MeV="245.0")]
    g = P.group_curves_by_temperature(rows, xi=0.0)
    assert sorted(g) == [70.0, 80.0] and len(g[70.0]) == 2 and len(g[80.0]) == 1


def test_hand_worked_extremum_and_area():
    """The s_shape_curve by hand: the + -> - change is at rho 0.6 (mu 275), so
    the hadron spinodal is the vertex of the parabola through (0.4, 270),
    (0.6, 275), (0.8, 272), solved here as a 3x3 linear system; the - -> +
    change at rho 1.2 (265) gives the vertex through (1.0, 268), (1.2, 265),
    (1.4, 270)."""
    P = PT()
    r = P.detect_s_shape(*s_shape_curve())

    def vertex(r1, r2, r3, m1, m2, m3):
        A = np.array([[r1 * r1, r1, 1], [r2 * r2, r2, 1], [r3 * r3, r3, 1]])
        a, b, c = np.linalg.solve(A, [m1, m2, m3])
        x = -b / (2 * a)
        return x, a * x * x + b * x + c
    x, m = vertex(0.4, 0.6, 0.8, 270.0, 275.0, 272.0)
    assert abs(r.rho_spinodal_hadron - x) < 1e-12 and abs(r.mu_spinodal_hadron - m) < 1e-9
    x, m = vertex(1.0, 1.2, 1.4, 268.0, 265.0, 270.0)
    assert abs(r.rho_spinodal_quark - x) < 1e-12 and abs(r.mu_spinodal_quark - m) < 1e-9
    assert r.derivative_sign_changes == 2


def test_antisymmetric_cubic_gives_the_centre():
    """mu = mu0 - a x + b x^3, x = rho - rho_c: by symmetry the equal-area
    level is mu0 and the outer crossings are x = -+sqrt(a / b); the hadron
    spinodal (maximum) is at x = -sqrt(a / 3b), mu0 + (2a/3) sqrt(a / 3b), the
    quark spinodal (minimum) at the mirror point."""
    P = PT()
    mu0, a, b, rc = 340.0, 12.0, 30.0, 1.3
    rho = np.linspace(0.0, 2.6, 2601)
    x = rho - rc
    mu = mu0 - a * x + b * x ** 3
    r = P.maxwell_construction(mu, rho, tol_area=1e-12, max_iter=200)
    assert r.converged
    assert abs(r.mu_transition - mu0) < 1e-9
    w = math.sqrt(a / b)
    assert abs(r.rho_hadron - (rc - w)) < 1e-6 and abs(r.rho_quark - (rc + w)) < 1e-6
    s = P.detect_s_shape(mu, rho)
    xs = math.sqrt(a / (3 * b))
    dm = 2 * a / 3 * xs
    assert s.has_s_shape and s.derivative_sign_changes == 2
    assert abs(s.rho_spinodal_hadron - (rc - xs)) < 1e-6 and abs(s.rho_spinodal_quark - (rc + xs)) < 1e-6
    # the parabola through three samples of the cubic: O(h^2) from the true extremum
    assert abs(s.mu_spinodal_hadron - (mu0 + dm)) < 1e-6 and abs(s.mu_spinodal_quark - (mu0 - dm)) < 1e-6
    # a shifted level breaks the symmetry: the area is no longer zero there
    r2 = P.maxwell_construction(mu + 0.5 * (rho > rc) * (rho - rc), rho, tol_area=1e-12, max_iter=200)
    assert r2.converged and r2.mu_transition > mu0 + 1e-3


def _stored_curves(xi):
    rows = golden_rows("trho_scan_T50_80_110.csv")
    return rows, PT().group_curves_by_temperature(rows, xi=xi)


@pytest.mark.parametrize("xi", [0.0, 0.2, 0.4])
def test_stored_curves_reproduce_reference_boundary(xi):
    """maxwell_construction / detect_s_shape on the reference's TrhoScan curves
    at T = 50, 80, 110 MeV give the reference's boundary.csv (mu_c, rho_hadron,
    rho_quark) and spinodals.csv to the last bit: same curve, same steps."""
    P = PT()
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    sp = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("spinodals.csv")}
    _, curves = _stored_curves(xi)
    assert sorted(curves) == [50.0, 80.0, 110.0]
    for T, samples in curves.items():
        mu, rho = [s[0] for s in samples], [s[1] for s in samples]
        m = P.maxwell_construction(mu, rho, **MAXWELL_KW)
        b = bd[(xi, T)]
        assert m.converged
        assert m.mu_transition == float(b["mu_transition_MeV"])
        assert m.rho_hadron == float(b["rho_hadron"]) and m.rho_quark == float(b["rho_quark"])
        s = P.detect_s_shape(mu, rho)
        p = sp[(xi, T)]
        assert s.has_s_shape
        assert (s.mu_spinodal_hadron, s.mu_spinodal_quark, s.rho_spinodal_hadron, s.rho_spinodal_quark) == tuple(
            float(p[k]) for k in ("mu_spinodal_hadron_MeV", "mu_spinodal_quark_MeV", "rho_spinodal_hadron",
                                  "rho_spinodal_quark"))


@pytest.mark.parametrize("xi", [0.0, 0.4])
def test_maxwell_level_has_equal_pressures(xi):
    """dP = rho dmu at fixed T: at the Maxwell mu_c the stored pressures of the
    hadron branch (rho below the hadron spinodal) and of the quark branch
    (rho above the quark spinodal), interpolated in mu, agree."""
    P = PT()
    rows, curves = _stored_curves(xi)
    for T in (50.0, 80.0):
        sel = [r for r in rows if float(r["T_MeV"]) == T and abs(float(r["xi"]) - xi) < 1e-9]
        mu = np.array([float(r["mu_avg_MeV"]) for r in sel])
        rho = np.array([float(r["rho"]) for r in sel])
        p = np.array([float(r["pressure_fm4"]) for r in sel])
        m = P.maxwell_construction(mu, rho, **MAXWELL_KW)
        s = P.detect_s_shape(mu, rho)
        h = (rho < s.rho_spinodal_hadron) & (rho > 0)
        q = rho > s.rho_spinodal_quark
        ph = np.interp(m.mu_transition, *zip(*sorted(zip(mu[h], p[h]))))
        pq = np.interp(m.mu_transition, *zip(*sorted(zip(mu[q], p[q]))))
        assert abs(ph - pq) < 2e-4 * max(abs(ph), 1.0)
        # and away from mu_c they differ: the check has teeth
        ph2 = np.interp(m.mu_transition + 3.0, *zip(*sorted(zip(mu[h], p[h]))))
        pq2 = np.interp(m.mu_transition + 3.0, *zip(*sorted(zip(mu[q], p[q]))))
        assert abs(ph2 - pq2) > 20 * abs(ph - pq)


def _oracle_solver(oracle):
    def solve(T, mu, xi, policy, seed):
        return oracle.solve_batch(T, mu, xi, policy=policy, seed=seed, p_num=24, t_num=8, nthreads=8)
    return solve


def test_s_curve_maxwell_matches_reference_boundary(oracle):
    """The three FixedMu branches (hadron, unstable, quark) over mu 250-420 MeV
    in 0.25 MeV steps form the S-curve; its Maxwell mu_c lies within 0.06 MeV
    of the reference's boundary.csv (which used TrhoScan's rho-sampled curve),
    rho_hadron / rho_quark within 0.01."""
    P = PT()
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    mus = np.arange(250.0, 420.0 + 1e-9, 0.25)
    for xi, T in ((0.0, 50.0), (0.0, 80.0), (0.4, 80.0)):
        mu, rho = P.s_curve(T, mus, xi, solver=_oracle_solver(oracle))
        # the unstable branch is there: rho between the spinodal densities
        s = P.detect_s_shape(mu, rho)
        assert s.has_s_shape
        m = P.maxwell_construction(mu, rho, **MAXWELL_KW)
        b = bd[(xi, T)]
        assert m.converged
        assert abs(m.mu_transition - float(b["mu_transition_MeV"])) < 0.06
        assert abs(m.rho_hadron - float(b["rho_hadron"])) < 0.01
        assert abs(m.rho_quark - float(b["rho_quark"])) < 0.01


def test_load_curves_from_csv_feeds_the_boundary(tmp_path):
    """load_curves_from_csv (calculate_phase_structure.jl:226-273) reads a
    TrhoScan file -- converged rows of one xi, temperatures with >= 3
    samples -- and its curves give the boundary.csv rows; an unconverged row
    and a short temperature are dropped."""
    import csv
    import os
    P = PT()
    src = os.path.join(os.path.dirname(__file__), "golden", "trho_scan_T50_80_110.csv")
    rows = list(csv.DictReader(open(src)))
    path = tmp_path / "trho.csv"
    with open(path, "w") as f:
        f.write(",".join(list(rows[0].keys()) + ["converged"]) + "\n")
        for r in rows:
            f.write(",".join(list(r.values()) + ["true"]) + "\n")
        bad = dict(rows[0], mu_avg_MeV="999.0")
        f.write(",".join(list(bad.values()) + ["false"]) + "\n")      # not converged: ignored
        for k in range(2):                                               # 2 samples only: dropped
            f.write(",".join(["60.000000", f"{k}.0", "0.000000"] + ["1.0"] * (len(rows[0]) - 3) + ["true"]) + "\n")
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    for xi in (0.0, 0.4):
        curves = P.load_curves_from_csv(str(path), xi)
        assert sorted(curves) == [50.0, 80.0, 110.0]
        for T, (mu, rho) in curves.items():
            m = P.maxwell_construction(mu, rho, **MAXWELL_KW)
            assert m.converged and m.mu_transition == float(bd[(xi, T)]["mu_transition_MeV"])


def test_maxwell_line_refines_next_to_the_cep(oracle):
    """Just below each xi's CEP (T_CEP 130.9, 123.0, 113.0 MeV at xi 0, 0.2,
    0.4) the S-region is narrower than a 0.25 MeV mu grid; maxwell_line
    re-solves such a line 16x finer over its coexistence window and then
    lands within 0.01 MeV of the reference's boundary.csv."""
    P = PT()
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    mus = np.arange(250.0, 420.0 + 1e-9, 0.25)
    for xi, T in ((0.0, 130.0), (0.2, 120.0), (0.4, 110.0)):
        m = P.maxwell_line([T], mus, xi, solver=_oracle_solver(oracle), min_samples=8, detect_min_points=5)[T]
        assert m.converged and "refined" in m.details
        assert abs(m.mu_transition - float(bd[(xi, T)]["mu_transition_MeV"])) < 0.01


def test_maxwell_line_does_not_refine_crossover_lines(oracle):
    """Above each xi's CEP the hadron and quark branches are the same root
    wherever both exist: the construction fails (no S-shape) and the line is
    not re-solved on a finer grid; no solve leaves the requested mu range."""
    P = PT()
    mus = np.arange(250.0, 420.0 + 1e-9, 0.25)
    solve = _oracle_solver(oracle)
    for xi, T in ((0.0, 140.0), (0.2, 130.0), (0.4, 120.0)):
        seen = {"points": 0, "mu_min": np.inf, "mu_max": -np.inf}

        def counting(Tq, muq, xiq, policy, seed):
            seen["points"] += np.size(Tq)
            seen["mu_min"] = min(seen["mu_min"], float(np.min(muq)) * 197.327)
            seen["mu_max"] = max(seen["mu_max"], float(np.max(muq)) * 197.327)
            return solve(Tq, muq, xiq, policy, seed)
        m = P.maxwell_line([T], mus, xi, solver=counting, min_samples=8, detect_min_points=5)[T]
        in_line = seen["points"]
        seen["points"] = 0
        P.s_curves([T], mus, xi, solver=counting)
        assert not m.converged and "refined" not in m.details
        assert in_line == seen["points"], (in_line, seen["points"])    # no second, finer pass
        assert seen["mu_min"] >= 250.0 - 1e-9 and seen["mu_max"] <= 420.0 + 1e-9
