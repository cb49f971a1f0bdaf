"""The dual-branch scan logic (julia_relaxtime_amd.dual_branch, DualBranchScan.jl)
driven by the oracle's solves on the CPU, against the reference's own
output data/outputs/results/pnjl/dual_branch_T50.csv (T = 50 MeV, xi = 0,
mu 0-400 MeV step 5, p24/t8, 10 decimals) and its Maxwell line
(boundary.csv, mu_transition at T = 50 MeV)."""
import numpy as np
import pytest

from conftest import golden_rows

HB = 197.327


def _oracle_solver(oracle, p_num=24, t_num=8):
    def solve(T, mu, xi, policy, seed):
        return oracle.solve_batch(T, mu, xi, policy=policy, seed=seed, p_num=p_num, t_num=t_num, nthreads=8)
    return solve


@pytest.fixture(scope="module")
def T50(oracle):
    from julia_relaxtime_amd import dual_branch as DB
    return DB.run_dual_branch_scan_batch([50.0], np.arange(0.0, 400.0 + 1e-9, 5.0), 0.0,
                                         solver=_oracle_solver(oracle))[0]


def test_branches_reproduce_reference_output(T50, tmp_path):
    from julia_relaxtime_amd import dual_branch as DB
    rows = golden_rows("dual_branch_T50.csv")
    merged = DB.merge_branches(T50, output_path=str(tmp_path / "merged.csv"))
    assert [r["mu_MeV"] for r in merged] == [float(r["mu_MeV"]) for r in rows]
    assert [r["branch"] for r in merged] == [r["branch"] for r in rows]
    for m, r in zip(merged, rows):
        for k in ("omega", "pressure", "rho", "entropy", "energy", "phi_u", "phi_d", "phi_s", "Phi1", "Phi2"):
            assert abs(m[k] - float(r[k])) < 3e-9, (r["mu_MeV"], k)
        for k in ("M_u_MeV", "M_d_MeV", "M_s_MeV"):
            assert abs(m[k] - float(r[k])) < 2e-6
    # the written file has the reference's columns and formats
    lines = open(tmp_path / "merged.csv").read().splitlines()
    assert lines[0] == ",".join(DB.MERGED_COLUMNS)
    first = lines[1].split(",")
    assert first[:4] == ["50.000000", "0.000000", "0.000000", "hadron"] and first[-1] == "NaN"


def test_phase_transition_matches_maxwell_line(T50):
    from julia_relaxtime_amd import dual_branch as DB
    t = DB.find_phase_transition(T50)
    ref = [r for r in golden_rows("phase_boundary.csv") if float(r["xi"]) == 0.0 and float(r["T_MeV"]) == 50.0][0]
    assert t["found"] and 355.0 < t["mu_c"] < 360.0
    # linear interpolation of the Omega difference on a 5 MeV grid
    assert abs(t["mu_c"] - float(ref["mu_transition_MeV"])) < 0.1
    lo, hi = t["coexistence_region"]
    assert t["mu_quark_spinodal"] <= lo < t["mu_c"] < hi <= t["mu_hadron_spinodal"]


def test_batch_lines_are_independent(oracle):
    """Lines of one batch give what each gives alone (a line that breaks early
    does not disturb the others)."""
    from julia_relaxtime_amd import dual_branch as DB
    mu = np.arange(200.0, 400.0 + 1e-9, 10.0)
    solve = _oracle_solver(oracle)
    both = DB.run_dual_branch_scan_batch([40.0, 120.0], mu, [0.0, 0.2], solver=solve)
    for l, (T, xi) in enumerate(((40.0, 0.0), (120.0, 0.2))):
        one = DB.run_dual_branch_scan_batch([T], mu, xi, solver=solve)[0]
        for b in ("hadron", "quark", "physical"):
            assert np.array_equal(both[l][b]["valid"], one[b]["valid"])
            v = one[b]["valid"]
            assert np.array_equal(both[l][b]["omega"][v], one[b]["omega"][v])
