"""GPU dual-branch scans (julia_relaxtime_amd.dual_branch over the C ABI's
batched solves) against the reference's dual_branch_T50.csv and Maxwell line,
and against the same logic driven by the oracle on a batch of lines."""
import numpy as np
import pytest

from conftest import golden_rows

pytestmark = pytest.mark.gpu


def test_T50_reproduces_reference_output(gpu_ctx):
    from julia_relaxtime_amd import dual_branch as DB
    res = DB.run_dual_branch_scan(50.0, np.arange(0.0, 400.0 + 1e-9, 5.0), 0.0)
    rows = golden_rows("dual_branch_T50.csv")
    merged = DB.merge_branches(res)
    assert [r["branch"] for r in merged] == [r["branch"] for r in rows]
    for m, r in zip(merged, rows):
        for k in ("omega", "pressure", "rho", "entropy", "energy", "phi_u", "phi_d", "phi_s", "Phi1", "Phi2"):
            assert abs(m[k] - float(r[k])) < 3e-9, (r["mu_MeV"], k)
    t = DB.find_phase_transition(res)
    ref = [r for r in golden_rows("phase_boundary.csv") if float(r["xi"]) == 0.0 and float(r["T_MeV"]) == 50.0][0]
    assert t["found"] and abs(t["mu_c"] - float(ref["mu_transition_MeV"])) < 0.1


def test_batch_matches_oracle(gpu_ctx, oracle):
    """36 lines (T 30-140 MeV, xi 0, +-0.2) over mu 200-420 MeV: the same
    branch extents, Omega to 1e-10 relative where both followed a point, the
    same transitions."""
    from julia_relaxtime_amd import dual_branch as DB
    T = np.repeat(np.arange(30.0, 141.0, 10.0), 3)
    xi = np.tile([0.0, 0.2, -0.2], T.size // 3)
    mu = np.arange(200.0, 420.0 + 1e-9, 5.0)
    g = DB.run_dual_branch_scan_batch(T, mu, xi)

    def solve(Tq, muq, xiq, policy, seed):
        return oracle.solve_batch(Tq, muq, xiq, policy=policy, seed=seed, p_num=24, t_num=8, nthreads=8)
    o = DB.run_dual_branch_scan_batch(T, mu, xi, solver=solve)
    agree = 0
    for gl, ol in zip(g, o):
        for b in ("hadron", "quark"):
            agree += np.array_equal(gl[b]["valid"], ol[b]["valid"])
            v = gl[b]["valid"] & ol[b]["valid"]
            assert np.all(np.abs(gl[b]["omega"][v] - ol[b]["omega"][v]) <= 1e-10 * np.abs(ol[b]["omega"][v]))
        tg, to = DB.find_phase_transition(gl), DB.find_phase_transition(ol)
        assert tg["found"] == to["found"]
        if tg["found"]:
            assert abs(tg["mu_c"] - to["mu_c"]) < 1e-6
    assert agree >= 2 * len(g) - 1     # a branch end decided by a last-bit difference is allowed once


def test_branch_chains_match_per_step_solves(gpu_ctx):
    """One launch of branch chains per branch (PNJL_SEED_BRANCH_HADRON /
    _QUARK: the branch end and the jump test inside the kernel) against the
    same branches followed one batched launch per mu step: same extents,
    same states; the points after a branch end are reported skipped."""
    from julia_relaxtime_amd import _lib
    from julia_relaxtime_amd import dual_branch as DB
    T = np.repeat(np.arange(20.0, 151.0, 10.0), 2)
    xi = np.tile([0.0, 0.3], T.size // 2)
    mu = np.arange(0.0, 450.0 + 1e-9, 3.0)
    seen = {}
    base = DB.gpu_chain_solver()

    def chains(Tq, muq, xiq, policy, n):
        r = base(Tq, muq, xiq, policy, n)
        seen[policy] = r
        return r
    c = DB.run_dual_branch_scan_batch(T, mu, xi, chain_solver=chains)
    s = DB.run_dual_branch_scan_batch(T, mu, xi, solver=DB.gpu_solver())
    for cl, sl in zip(c, s):
        for b in ("hadron", "quark"):
            assert np.array_equal(cl[b]["valid"], sl[b]["valid"])
            v = cl[b]["valid"]
            for k in ("x", "omega", "pressure", "rho_norm", "masses"):
                assert np.allclose(cl[b][k][v], sl[b][k][v], rtol=1e-12, atol=1e-13), k
    for pol, r in seen.items():
        fl = r["flags"].reshape(T.size, mu.size)
        ends = (fl & _lib.PT_BRANCH_END) != 0
        skipped = (fl & _lib.PT_SKIPPED) != 0
        assert np.all(ends.sum(axis=1) <= 1)
        for l in range(T.size):
            e = np.flatnonzero(ends[l])
            if e.size:
                assert skipped[l, e[0] + 1:].all() and not skipped[l, :e[0] + 1].any()
            else:
                assert not skipped[l].any()
        assert np.all(r["converged"].reshape(T.size, mu.size)[skipped] == 0)
