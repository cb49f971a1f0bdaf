"""GPU parity of the gap path: libpnjl_hip.so (through its C ABI) against the
CPU oracle on the same seeded inputs, and against the reference's own stored
scans.  Tolerances: roots / Omega to 1e-10 relative;
the node pass itself to ~1e-13."""
import numpy as np
import pytest

from conftest import golden_rows

pytestmark = pytest.mark.gpu
HB = 197.327


def _grid(nT, nmu, xis, Tlo=20, Thi=300, mlo=0, mhi=450):
    T, mu, xi = np.meshgrid(np.linspace(Tlo, Thi, nT), np.linspace(mlo, mhi, nmu), np.asarray(xis, float),
                            indexing="ij")
    return T.ravel() / HB, mu.ravel() / HB, xi.ravel()


def _x_close(a, b, rtol=1e-10):
    """state vectors: relative to max(|x|, 1) per component"""
    return np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b))


def _xdiff(a, b):
    return np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)), axis=1)


def ulp_unstable(rerun, T, mu, idx, o, rtol=1e-10):
    """Points idx at which the oracle's own root moves when T or mu moves by
    one ulp (nextafter up or down): the oracle is re-run there by
    rerun(T', mu', idx) and compared with its root o at the exact input."""
    idx = np.asarray(idx, dtype=np.int64)
    out = np.zeros(idx.size, dtype=bool)
    if idx.size == 0:
        return out
    up, dn = np.inf, -np.inf
    for tT, tm in ((up, None), (dn, None), (None, up), (None, dn)):
        Tp = T[idx] if tT is None else np.nextafter(T[idx], tT)
        mp = mu[idx] if tm is None else np.nextafter(mu[idx], tm)
        q = rerun(Tp, mp, idx)
        c0 = o["converged"][idx] == 1
        out |= (q["converged"] != o["converged"][idx]) | (c0 & (_xdiff(q["x"], o["x"][idx]) > rtol))
    return out


def assert_parity(r, o, T=None, mu=None, rerun=None, rtol=1e-10, selection=False, label=""):
    """Per-point parity of GPU result r with oracle result o.

    Every converged point agrees to rtol in x and Omega and to 1e-9 in the
    thermodynamics, and the converged flags agree, except at points where the
    oracle itself is not stable: if T or mu moves by one ulp, the oracle's
    own root (or its converged flag) changes there (a Newton run that
    wanders from a far seed amplifies last-bit differences, and the GPU's
    node sums are added in another order).  Those points are found by
    re-running the oracle through rerun(T', mu', idx) at the mismatches only;
    the counts are printed.

    selection=True (policies that keep the minimum-Omega root of several
    runs): two runs that stop within ftol of the same root tie in Omega to
    rounding, and either may be kept; such a point (Omega equal to 1e-12,
    x to 1e-8, a different winning run) counts as the same root, and its
    thermodynamics (first derivatives of P, steep in x at low T) are held to
    1e-6 instead of 1e-9."""
    n = r["x"].shape[0]
    err = _xdiff(r["x"], o["x"])
    eo = np.abs(r["omega"] - o["omega"]) / np.abs(o["omega"])
    bad = (r["converged"] != o["converged"]) | ((o["converged"] == 1) & ((err > rtol) | (eo > rtol)))
    tie = np.zeros_like(bad)
    if selection:
        tie = (eo < 1e-12) & (err < 1e-8) & (r["iterations"] != o["iterations"]) & bad
        bad &= ~tie
    unstable = np.zeros(n, dtype=bool)
    mism = np.flatnonzero(bad)
    if mism.size and rerun is not None:
        unstable[mism] = ulp_unstable(rerun, T, mu, mism, o, rtol)
    stable_bad = bad & ~unstable
    print(f"[parity {label}] {n} points: {int(bad.sum())} root mismatches, {int(unstable.sum())} of them at "
          f"ulp-unstable points, {int(stable_bad.sum())} at stable points; {int(tie.sum())} Omega ties",
          flush=True)
    worst = np.flatnonzero(stable_bad)[:5]
    info = [(int(i), float(err[i]), int(r["iterations"][i]), int(o["iterations"][i]), int(r["flags"][i]),
             int(o["flags"][i])) for i in worst]
    assert stable_bad.sum() == 0, f"mismatches at ulp-stable points: {info}"
    ok = (o["converged"] == 1) & ~bad
    tol = np.where(tie, 1e-6, 1e-9)
    for k in ("pressure", "entropy", "energy", "rho_norm"):
        d = np.abs(r[k] - o[k]) / np.maximum(1, np.abs(o[k]))
        w = np.flatnonzero(ok & (d > tol))
        assert w.size == 0, (k, [(int(i), float(d[i]), float(err[i]), bool(tie[i])) for i in w[:5]])
    dn = np.max(np.abs(r["n_q"] - o["n_q"]), axis=1)
    assert np.all(dn[ok] <= tol[ok])
    return ok


def oracle_rerun(oracle, xi, policy, p_num, t_num, opts, seed=None):
    """rerun(T', mu', idx) for assert_parity: the oracle at points idx with
    the same xi, policy, seeds and options."""
    def run(Tp, mp, idx):
        return oracle.solve_batch(Tp, mp, xi[idx], policy=policy, seed=None if seed is None else seed[idx],
                                  p_num=p_num, t_num=t_num, opts=opts, nthreads=8)
    return run


def test_residual_and_jacobian_match_oracle(gpu_ctx, oracle):
    rng = np.random.default_rng(11)
    n = 300
    T = rng.uniform(15, 320, n) / HB
    mu = rng.uniform(0, 460, n) / HB
    xi = rng.choice([0.0, -0.4, 0.2, 0.4], n)
    x = np.stack([rng.uniform(-2, -0.1, n), rng.uniform(-2, -0.1, n), rng.uniform(-2.3, -0.5, n),
                  rng.uniform(0, 1, n), rng.uniform(0, 1, n)], 1)
    x[: n // 2, 1] = x[: n // 2, 0]          # half on the phi_u == phi_d manifold
    for pn, tn in ((64, 4), (24, 8)):
        F, J, om = gpu_ctx.residual_batch(x, T, mu, xi, pn, tn)
        for i in range(0, n, 3):
            o, g, H = oracle.omega_derivs(x[i], mu[i], T[i], xi[i], pn, tn)
            assert abs(om[i] - o) <= 1e-13 * max(1.0, abs(o)), i
            scale = max(1.0, np.abs(g).max())
            assert np.max(np.abs(F[i] + g)) <= 1e-12 * scale, i
            hs = max(1.0, np.abs(H).max())
            assert np.max(np.abs(J[i] + H)) <= 1e-12 * hs, i


def test_residual_at_floor_and_extreme_exponents(gpu_ctx, oracle):
    """The integrand's edge branches: negative Polyakov loops that push a
    log argument under the reference's 1e-16 floor, exponents (mu - E)/T up to
    150 (T = 3 MeV, mu = 450 MeV) beyond the unscaled fast path, mu < 0
    (exp(-2 mu/T) > 1) and xi != 0 -- value, gradient and Hessian as the
    oracle's shifted-log restatement.  A negative loop can leave a log
    argument D just above the floor, where the ratios t^k/D and their squares
    in the Hessian amplify rounding (max |H| reaches ~1e4 here): the Hessian
    is held to 1e-9 of its largest entry, a wrong branch would miss by O(1)."""
    import itertools
    rows = list(itertools.product([3, 5, 20, 150], [-300, 0, 450], [-1.2, -0.4, 0.3], [0.05, 0.9], [0.0, 0.3]))
    T = np.array([r[0] for r in rows]) / HB
    mu = np.array([r[1] for r in rows]) / HB
    xi = np.array([r[4] for r in rows], dtype=float)
    x = np.array([[-0.5, -0.5, -1.9, r[2], r[3]] for r in rows])
    F, J, om = gpu_ctx.residual_batch(x, T, mu, xi, 64, 4)
    for i in range(len(rows)):
        o, g, H = oracle.omega_derivs(x[i], mu[i], T[i], xi[i], 64, 4)
        assert abs(om[i] - o) <= 1e-13 * max(1.0, abs(o)), rows[i]
        assert np.max(np.abs(F[i] + g)) <= 1e-11 * max(1.0, np.abs(g).max()), rows[i]
        assert np.max(np.abs(J[i] + H)) <= 1e-9 * max(1.0, np.abs(H).max()), rows[i]


@pytest.mark.parametrize("policy,name", [(0, "default"), (3, "multi"), (1, "hadron"), (2, "quark")])
def test_policy_parity_small_grid(gpu_ctx, oracle, policy, name):
    from julia_relaxtime_amd import _lib
    T, mu, xi = _grid(10, 10, [0.0, -0.4, 0.3])
    # |res| < 1e-10 gate: both solvers stop within 1e-10 of the same root
    r = gpu_ctx.solve_batch(T, mu, xi, policy=policy, p_num=64, t_num=4,
                            opts=_lib.default_opts(ftol=1e-10, max_iter=40))
    oo = oracle.default_opts(ftol=1e-10, max_iter=40)
    o = oracle.solve_batch(T, mu, xi, policy=policy, p_num=64, t_num=4, nthreads=8, opts=oo)
    assert o["converged"].mean() > 0.9
    ok = assert_parity(r, o, T, mu, oracle_rerun(oracle, xi, policy, 64, 4, oo), selection=(policy == 3),
                       label=name)
    # integer by-products: identical except at points whose residual sits on the ftol edge.
    # MultiSeed: several candidates reach the same root and argmin Omega then
    # ties to rounding, so the winning candidate (and its count) may differ.
    # Points that went through the MultiSeed fallback are excluded for the same reason.
    plain = (r["flags"] == 0) & (o["flags"] == 0)
    same_it = np.mean(r["iterations"][ok & plain] == o["iterations"][ok & plain])
    same_fl = np.mean(r["flags"][ok] == o["flags"][ok])
    print(f"[parity {name}] iterations equal on {same_it:.4f} of plain points, flags on {same_fl:.4f}", flush=True)
    if policy != 3:
        assert same_it > 0.97
        assert same_fl > 0.9


def test_explicit_seed_and_asymmetric_states(gpu_ctx, oracle):
    """Explicit per-point seeds, half of them with phi_u != phi_d (the
    three-flavour path): seeds are the roots perturbed by up to 5 %."""
    from julia_relaxtime_amd import _lib
    rng = np.random.default_rng(5)
    n = 120
    T = rng.uniform(30, 300, n) / HB
    mu = rng.uniform(0, 450, n) / HB
    xi = rng.uniform(-0.4, 0.4, n)
    base = oracle.solve_batch(T, mu, xi, policy=0, p_num=32, t_num=4, opts=oracle.default_opts(ftol=1e-10),
                              nthreads=8)["x"]
    seed = base * (1 + rng.uniform(-0.05, 0.05, base.shape))
    seed[: n // 2, 1] = seed[: n // 2, 0]
    r = gpu_ctx.solve_batch(T, mu, xi, policy=4, seed=seed, p_num=32, t_num=4, opts=_lib.default_opts(ftol=1e-10))
    oo = oracle.default_opts(ftol=1e-10)
    o = oracle.solve_batch(T, mu, xi, policy=4, seed=seed, p_num=32, t_num=4, nthreads=8, opts=oo)
    ok = assert_parity(r, o, T, mu, oracle_rerun(oracle, xi, 4, 32, 4, oo, seed=seed), label="explicit seeds")
    # asymmetric seeds converge to the symmetric root: phi_u = phi_d to 1e-9
    assert np.all(np.abs(r["x"][ok, 0] - r["x"][ok, 1]) < 1e-9)


def _driver(name):
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


@pytest.mark.parametrize("name", ["driver_gap_scan.csv", "driver_gap_scan_xi-0p6to0p6.csv"])
def test_driver_chains_match_reference_scan(gpu_ctx, name):
    """PhaseAwareContinuitySeed chains on the GPU reproduce the reference's
    own run_gap_transport_scan.jl output (p12/t6, iterations=40)."""
    from julia_relaxtime_amd import _lib
    rows, L = _driver(name)
    f = lambda k: np.array([float(r[k]) for r in rows])
    r = gpu_ctx.solve_batch(f("T_fm"), f("muq_fm"), f("xi"), policy=6, p_num=12, t_num=6,
                            opts=_lib.default_opts(max_iter=40, chain_len=L))
    assert r["converged"].all()
    om = f("omega_fm4inv")
    assert np.max(np.abs(r["omega"] - om) / np.abs(om)) < 1e-12
    M = np.stack([f("m_u"), f("m_d"), f("m_s")], 1)
    assert np.max(np.abs(r["masses"] - M)) < 1e-8
    assert np.max(np.abs(r["entropy"] - f("s_fm3inv"))) < 1e-8
    assert np.mean(r["iterations"] == f("iterations").astype(int)) > 0.95


def test_continuity_chain_c1(gpu_ctx, oracle):
    """C1: 32x32 (T 50-300, mu 0-400 MeV), xi = 0, p = 64, one ContinuitySeed
    chain in scan order (T outer, mu inner, no reset), gate |res| < 1e-10.

    Near the first-order line the chain can step onto a metastable branch
    (phi_s > 0) and follow it; whether it does is decided by the last bit of
    the arithmetic -- the oracle itself flips when T moves by 1 ulp.  Parity
    is asserted on every point where the oracle is stable under +-1 ulp of T."""
    from julia_relaxtime_amd import _lib
    T, mu, xi = _grid(32, 32, [0.0], 50, 300, 0, 400)
    n = T.size
    r = gpu_ctx.solve_batch(T, mu, xi, policy=5, p_num=64, t_num=8,
                            opts=_lib.default_opts(ftol=1e-10, chain_len=n))
    run = lambda TT: oracle.solve_batch(TT, mu, xi, policy=5, p_num=64, t_num=8,
                                        opts=oracle.default_opts(ftol=1e-10, chain_len=n))
    o = run(T)
    assert o["converged"].all() and r["converged"].all()
    assert np.all(r["residual_norm"] < 1e-10)
    bad = ~_x_close(r["x"], o["x"]).all(axis=1) | (np.abs(r["omega"] - o["omega"]) > 1e-10 * np.abs(o["omega"]))
    # the whole chain is re-run with every T (or mu) moved by one ulp
    unstable = np.zeros(n, dtype=bool)
    if bad.any():
        for Tp, mp in ((np.nextafter(T, np.inf), mu), (np.nextafter(T, -np.inf), mu),
                       (T, np.nextafter(mu, np.inf)), (T, np.nextafter(mu, -np.inf))):
            q = oracle.solve_batch(Tp, mp, xi, policy=5, p_num=64, t_num=8,
                                   opts=oracle.default_opts(ftol=1e-10, chain_len=n))
            unstable |= _xdiff(q["x"], o["x"]) > 1e-10
    print(f"[parity c1 chain] {n} points: {int(bad.sum())} root mismatches, {int((bad & unstable).sum())} of them "
          f"at ulp-unstable points", flush=True)
    assert not np.any(bad & ~unstable), np.flatnonzero(bad & ~unstable)[:10]


def test_c2_full_grid_properties(gpu_ctx, oracle):
    """C2 at full size (64x64x9 = 36,864 points, p=64, t=4, DefaultSeed per
    point, |res| < 1e-10): size-independent properties on every point plus
    oracle parity on a seeded sample."""
    from julia_relaxtime_amd import _lib
    xis = np.round(np.arange(-0.4, 0.41, 0.1), 10)
    T, mu, xi = _grid(64, 64, xis)
    opts = _lib.default_opts(ftol=1e-10, max_iter=40)
    r = gpu_ctx.solve_batch(T, mu, xi, policy=0, p_num=64, t_num=4, opts=opts)
    conv = r["converged"].astype(bool)
    assert conv.mean() > 0.99
    assert np.all(r["residual_norm"][conv] < 1e-10)
    assert np.all(r["x"][conv, 0] == r["x"][conv, 1])
    assert np.all((r["x"][conv, 3] >= -1e-8) & (r["x"][conv, 3] <= 1 + 1e-8))
    assert np.all(r["masses"][conv] > 0)
    assert np.allclose(r["pressure"], -r["omega"], rtol=0, atol=0, equal_nan=True)
    rng = np.random.default_rng(20261202)
    idx = rng.choice(T.size, 400, replace=False)
    oo = oracle.default_opts(ftol=1e-10, max_iter=40)
    o = oracle.solve_batch(T[idx], mu[idx], xi[idx], policy=0, p_num=64, t_num=4, opts=oo, nthreads=8)
    assert_parity({k: v[idx] for k, v in r.items()}, o, T[idx], mu[idx],
                  oracle_rerun(oracle, xi[idx], 0, 64, 4, oo), label="c2 sample")


def test_launch_geometry_does_not_change_results(gpu_ctx):
    """Queue chunk size and persistent grid size only move work between
    groups: bit-identical results.  Lanes per point change the summation
    order: equal to rounding."""
    T, mu, xi = _grid(12, 12, [0.0, 0.3])
    base_cfg = gpu_ctx.launch_cfg()
    try:
        from julia_relaxtime_amd import _lib
        opts = _lib.default_opts(ftol=1e-10)
        outs = []
        for lpp, chunk, wpc in ((4, 1, 4), (4, 16, 8), (4, 3, 2)):
            gpu_ctx.set_launch_cfg(lpp, chunk, wpc)
            outs.append(gpu_ctx.solve_batch(T, mu, xi, policy=0, p_num=64, t_num=4, opts=opts))
        for o in outs[1:]:
            for k in ("x", "omega", "iterations", "converged", "entropy"):
                assert np.array_equal(o[k], outs[0][k], equal_nan=True), k
        for lpp, tail in ((1, 0), (2, 0), (8, 0), (16, 0), (64, 0), (4, 64), (8, 32)):
            gpu_ctx.set_launch_cfg(lpp, 4, 8, tail)
            o = gpu_ctx.solve_batch(T, mu, xi, policy=0, p_num=64, t_num=4, opts=opts)
            ok = outs[0]["converged"].astype(bool)
            assert np.array_equal(o["converged"], outs[0]["converged"])
            assert np.all(_x_close(o["x"][ok], outs[0]["x"][ok], 1e-10))
    finally:
        gpu_ctx.set_launch_cfg(**base_cfg)


def test_device_pointer_path(gpu_ctx):
    """Inputs already in HBM (torch tensors) give the same results as the host path."""
    import torch
    T, mu, xi = _grid(6, 6, [0.0, 0.2])
    h = gpu_ctx.solve_batch(T, mu, xi, policy=0, p_num=64, t_num=4)
    dev = torch.device("cuda", 0)
    d = gpu_ctx.solve_batch(torch.tensor(T, device=dev), torch.tensor(mu, device=dev), torch.tensor(xi, device=dev),
                            policy=0, p_num=64, t_num=4)
    torch.cuda.synchronize()
    for k in ("x", "omega", "iterations", "converged"):
        assert np.array_equal(d[k].cpu().numpy(), h[k], equal_nan=True), k


def test_error_paths(gpu_ctx):
    from julia_relaxtime_amd import _lib
    T = np.array([0.5])
    with pytest.raises(_lib.PnjlError):
        gpu_ctx.solve_batch(T, T, T, policy=4, seed=None)          # EXPLICIT needs seeds
    with pytest.raises(_lib.PnjlError):
        gpu_ctx.solve_batch(T, T, T, policy=42)
    with pytest.raises(_lib.PnjlError):
        gpu_ctx.solve_batch(T, T, T, policy=0, p_num=4096, t_num=2)
    with pytest.raises(_lib.PnjlError):
        gpu_ctx.set_launch_cfg(3, 1, 1)
