"""The oracle's learned-seed policies (SURROGATE, MULTIROOT3) pinned by
composition: each must reduce exactly to reference policies the gap tests
already pin (DefaultSeed, explicit seeds, HADRON/QUARK seeds)."""
import numpy as np
import pytest

HB = 197.327
HADRON = np.array([-1.84329, -1.84329, -2.22701, 1.0e-5, 4.0e-5])
HIGH_TEMP = np.array([-0.73192, -0.73192, -1.79539, 0.60532, 0.60532])


@pytest.fixture(scope="module")
def case(oracle):
    rng = np.random.default_rng(2)
    n = 160
    T = rng.uniform(20, 290, n) / HB
    mu = rng.uniform(0, 440, n) / HB
    xi = rng.choice([0.0, 0.2, -0.3], n)
    opts = oracle.default_opts(ftol=1e-10, max_iter=40)
    plain = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_DEFAULT, p_num=32, t_num=4, opts=opts, nthreads=8)
    seeds = plain["x"] * (1 + rng.normal(0, 0.03, (n, 5)))
    far = rng.random(n) < 0.35
    seeds[far] = np.stack([rng.uniform(-2, 0, n), np.zeros(n), rng.uniform(-2.3, -0.5, n), rng.uniform(0, 1, n),
                           rng.uniform(0, 1, n)], 1)[far]
    seeds[:, 1] = seeds[:, 0]
    return T, mu, xi, seeds, opts, plain


def _solve(oracle, case, policy, seed=None, **kw):
    T, mu, xi, seeds, opts, _ = case
    o = oracle.default_opts(ftol=1e-10, max_iter=40, **kw)
    return oracle.solve_batch(T, mu, xi, policy=policy, seed=seed, p_num=32, t_num=4, opts=o, nthreads=8)


def _same(a, b, keys=("x", "omega", "pressure", "entropy", "residual_norm", "converged")):
    return np.all([np.array_equal(a[k], b[k], equal_nan=True) for k in keys])


def test_reference_judge_in_mask_is_default_seed(oracle, case):
    T, mu, xi, seeds, opts, plain = case
    oracle.set_judge_mask(np.ones((2, 2, 2), np.uint8), [0, 0, -1], [2, 3, 1])
    try:
        r = _solve(oracle, case, oracle.SEED_SURROGATE, seeds)
    finally:
        oracle.set_judge_mask(None)
    assert _same(r, plain)
    assert np.array_equal(r["flags"], plain["flags"] | 0x100)


def test_learned_seed_cutoff_and_fallback(oracle, case):
    """No mask: a learned run that converges within 8 steps is the explicit-seed
    solve itself (same iterates); one that does not is DefaultSeed's solve."""
    T, mu, xi, seeds, opts, plain = case
    r = _solve(oracle, case, oracle.SEED_SURROGATE, seeds)
    ex = _solve(oracle, case, oracle.SEED_EXPLICIT, seeds, trust_region_fallback=0, auto_multiseed_fallback=0)
    fb = (r["flags"] & 0x10) != 0
    assert 0.1 < fb.mean() < 0.7
    for k in ("x", "omega", "residual_norm", "iterations"):
        assert np.array_equal(r[k][~fb], ex[k][~fb]), k
        assert np.array_equal(r[k][fb], plain[k][fb], equal_nan=True), k
    assert np.all(r["iterations"][~fb] <= 8)
    # the fallback points are those whose explicit run needs more than 8 steps or
    # fails, or lands farther from its seed than DualBranchScan's jump test allows
    # (|d phi_u| > 0.5 or |d M_u| > 50 MeV, DualBranchScan.jl:410-415)
    P = oracle.default_params()
    m_seed = P.m_ud0 - 4 * P.G * seeds[:, 0] + 2 * P.K * seeds[:, 1] * seeds[:, 2]
    jump = (np.abs(ex["x"][:, 0] - seeds[:, 0]) > 0.5) | (np.abs(ex["masses"][:, 0] - m_seed) * 197.327 > 50.0)
    slow = (ex["iterations"] > 8) | (ex["converged"] != 1) | ((ex["converged"] == 1) & jump)
    assert np.array_equal(fb, slow)


def test_multiroot3_is_min_omega_of_three_solves(oracle, case):
    T, mu, xi, seeds, opts, _ = case
    n = T.size
    r = _solve(oracle, case, oracle.SEED_MULTIROOT3, seeds)
    runs = [_solve(oracle, case, oracle.SEED_EXPLICIT, s, auto_multiseed_fallback=0)
            for s in (seeds, np.tile(HADRON, (n, 1)), np.tile(HIGH_TEMP, (n, 1)))]
    om = np.stack([np.where(d["converged"] == 1, d["omega"], np.inf) for d in runs], 1)
    b = np.argmin(om, 1)
    for i in range(n):
        if np.isinf(om[i, b[i]]):
            assert r["converged"][i] == 0 and (r["flags"][i] & 0x08)
            continue
        assert np.array_equal(r["x"][i], runs[b[i]]["x"][i]), i
    # first-order flag <=> two converged runs on distinct roots
    # (DualBranchScan.jl:409-419: |d phi_u|, |d Phi| >= 0.01 or |d M_u| >= 0.01 hbar c)
    dist = np.zeros(n, bool)
    for a in range(3):
        for c in range(a + 1, 3):
            A, C = runs[a], runs[c]
            both = (A["converged"] == 1) & (C["converged"] == 1)
            differ = (np.abs(A["x"][:, 0] - C["x"][:, 0]) >= 0.01) | (np.abs(A["x"][:, 3] - C["x"][:, 3]) >= 0.01) | (
                np.abs(A["masses"][:, 0] - C["masses"][:, 0]) >= 0.01)
            dist |= both & differ
    assert np.array_equal((r["flags"] & 0x40) != 0, dist)
    assert dist.any()


def test_omega_judge_never_keeps_a_higher_omega_root(oracle, case):
    T, mu, xi, seeds, opts, _ = case
    n = T.size
    oracle.set_judge_mask(np.ones((2, 2, 2), np.uint8), [0, 0, -1], [2, 3, 1])
    try:
        r = _solve(oracle, case, oracle.SEED_SURROGATE, seeds, surrogate_judge=1)
        base = _solve(oracle, case, oracle.SEED_SURROGATE, seeds, surrogate_judge=1)
    finally:
        oracle.set_judge_mask(None)
    nomask = _solve(oracle, case, oracle.SEED_SURROGATE, seeds)
    runs = [_solve(oracle, case, oracle.SEED_EXPLICIT, s, auto_multiseed_fallback=0)
            for s in (np.tile(HADRON, (n, 1)), np.tile(HIGH_TEMP, (n, 1)))]
    assert _same(r, base)                                   # deterministic
    ok = nomask["converged"] == 1
    assert np.all(r["omega"][ok] <= nomask["omega"][ok])
    for d in runs:
        c = ok & (d["converged"] == 1)
        assert np.all(r["omega"][c] <= d["omega"][c])
    moved = ok & ~np.all(r["x"] == nomask["x"], axis=1)
    assert moved.any()


def c5_lines(oracle, every=1):
    """mu-lines of the multi-root scan at the temperatures of the reference's
    Maxwell table (boundary.csv), T at least 15 MeV below the CEP: mu =
    linspace(200, 450, 512) MeV (the spacing of the 512-point scan), p24/t8 as
    the table.  Returns (lines, T, mu, xi) with lines = [(xi, T_MeV, mu_tr,
    mu_spinodal_quark, mu_spinodal_hadron)]."""
    import julia_relaxtime_amd.pnjl as P
    from conftest import golden_rows
    sp = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("spinodals.csv")}
    lines = []
    for r in golden_rows("phase_boundary.csv"):
        xi, Tm = float(r["xi"]), float(r["T_MeV"])
        if Tm > P.load_phase_boundary(xi)[2][2] - 15 or (xi, Tm) not in sp:
            continue
        s = sp[(xi, Tm)]
        lines.append((xi, Tm, float(r["mu_transition_MeV"]), float(s["mu_spinodal_quark_MeV"]),
                      float(s["mu_spinodal_hadron_MeV"])))
    lines = lines[::every]
    mus = np.linspace(200, 450, 512)
    T = np.concatenate([np.full(512, l[1]) for l in lines]) / HB
    mu = np.tile(mus, len(lines)) / HB
    xi = np.concatenate([np.full(512, l[0]) for l in lines])
    return lines, T, mu, xi


def check_c5_lines(lines, r):
    """On every line: all points converge; the min-Omega root jumps (|dM_u| >
    50 MeV) exactly once, within one grid step of the reference's Maxwell
    mu_transition; every point inside the reference's spinodal interval
    (1 MeV margin) carries the first-order flag."""
    mus = np.linspace(200, 450, 512)
    for k, (xi, Tm, mt, mq, mh) in enumerate(lines):
        s = slice(512 * k, 512 * (k + 1))
        assert np.all(r["converged"][s] == 1), (xi, Tm)
        dM = np.abs(np.diff(r["masses"][s, 0])) * HB
        assert np.sum(dM > 50) == 1, (xi, Tm)
        j = int(np.argmax(dM))
        assert mus[j] - 0.05 <= mt <= mus[j + 1] + 0.05, (xi, Tm, mus[j], mt)
        inside = (mus > mq + 1) & (mus < mh - 1)
        assert np.all((r["flags"][s][inside] & 0x40) != 0), (xi, Tm)


def test_multiroot3_lines_switch_at_reference_transition(oracle):
    """MULTIROOT3 (learned seed + HADRON + QUARK, min Omega) along mu at fixed
    T reproduces the reference's first-order line (Maxwell construction on
    the T-rho S-curves) and flags its coexistence region.  The learned-seed
    candidate is the DefaultSeed root here, as any seed is admissible."""
    lines, T, mu, xi = c5_lines(oracle, every=4)
    assert len(lines) >= 8 and {l[0] for l in lines} == {0.0, 0.2, 0.4}
    opts = oracle.default_opts(ftol=1e-10, max_iter=40)
    plain = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_DEFAULT, p_num=24, t_num=8, opts=opts, nthreads=8)
    r = oracle.solve_batch(T, mu, xi, policy=oracle.SEED_MULTIROOT3, seed=plain["x"], p_num=24, t_num=8,
                           opts=opts, nthreads=8)
    check_c5_lines(lines, r)
