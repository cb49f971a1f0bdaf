"""GPU consumers of Pi (pnjl_effective_couplings_batch,
pnjl_meson_propagators_batch, pnjl_propagator_by_channel_batch via the C ABI)
against the oracle's restatement (oracle/mesons.py).

Tolerances: the couplings and the mixing-matrix elements are the same
operations in the same order on both sides (no FMA contraction) and must
match bit for bit; propagators differ only in how a complex quotient is
rounded (1e-14 relative).  The per-process D_S / D_P compare (a) with the
oracle's composition applied to the GPU's own Pi values
(pnjl_polarization_batch, the same device arithmetic) at 1e-13, which pins
the table, flavour factors, slots and composition, and (b) end to end with
the oracle's Pi at 1e-9 relative, scaled by the propagator's sensitivity to
Pi (|4 K Pi / (1 - 4 K Pi)| for the simple and the det M cancellation for the
mixed mesons)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HB = 197.327


@pytest.fixture(scope="module")
def P():
    from julia_relaxtime_amd import propagators
    return propagators


@pytest.fixture(scope="module")
def M():
    from oracle import mesons
    return mesons


def _GK(gpu_ctx):
    return gpu_ctx.params.G, gpu_ctx.params.K


def test_couplings_bit_exact(gpu_ctx, P, M):
    G, K = _GK(gpu_ctx)
    rng = np.random.default_rng(1)
    n = 4000
    A_u, A_s = rng.uniform(-16, 2, n), rng.uniform(-16, 2, n)
    m_u, m_s = rng.uniform(0.0, 2.0, n), rng.uniform(0.3, 2.6, n)
    Gf, kc = P.couplings_from_A(A_u, A_s, m_u, m_s)
    for i in range(n):
        gu, gs = M.G_from_A(A_u[i], m_u[i]), M.G_from_A(A_s[i], m_s[i])
        assert Gf[i, 0] == gu and Gf[i, 1] == gs
        o = M.effective_couplings(G, K, gu, gs)
        assert np.array_equal(kc[i], [o[k] for k in M.KC_FIELDS])
    for Kv in (0.0, K, -2 * K):
        c = P.calculate_effective_couplings(G, Kv, -0.3, -0.2)
        assert c == M.effective_couplings(G, Kv, -0.3, -0.2)
    assert P.calculate_G_from_A(1.5, 0.8) == M.G_from_A(1.5, 0.8)


def _close(g, o, rtol):
    return np.abs(g - o) <= rtol * np.maximum(np.abs(o), 1e-300)


def test_meson_propagators_match_oracle(gpu_ctx, P, M):
    G, K = _GK(gpu_ctx)
    rng = np.random.default_rng(2)
    n = 3000
    kc = np.array([[M.effective_couplings(G, K, rng.uniform(-3, 3), rng.uniform(-3, 3))[k] for k in M.KC_FIELDS]
                   for _ in range(n)])
    pi = rng.normal(0, 1, (n, 6)) + 1j * rng.normal(0, 0.5, (n, 6))
    pi[::7] = pi[::7].real                             # real Pi: real propagators
    # exact poles of the simple propagators and non-finite Pi: the guards
    pi[1, 0] = 1 / (4 * kc[1, 2])
    pi[2, 2] = np.inf
    pi[3, 4] = np.nan
    g = P.meson_propagators_batch(kc, pi)
    for i in range(n):
        c = dict(zip(M.KC_FIELDS, kc[i]))
        o = [M.meson_propagator_simple(m, c, pi[i, s]) for m, s in (("pi", 0), ("K", 2), ("sigma_pi", 1),
                                                                      ("sigma_K", 3))]
        assert np.all(_close(g[i, :4], np.array(o), 1e-14)), i
        for ch, b, s in (("P", 4, 0), ("S", 7, 1)):
            M00, M08, M88, dK = M.mixing_matrix_elements(pi[i, s], pi[i, 4 + s], c, ch)
            mo = np.array([M00, M08, M88])
            same = (g[i, b:b + 3] == mo) | (np.isnan(g[i, b:b + 3]) & np.isnan(mo))
            assert same.all(), (i, ch)
            quarks = {(0, 0): "uuuu", (0, 1): "uuss", (1, 0): "ssuu", (1, 1): "ssss"}
            for j, (a, bq) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
                q = quarks[a, bq]
                om = M.meson_propagator_mixed(dK, M00, M08, M88, q[0], q[2], q[1], q[3], "t")
                gm = g[i, (10 if ch == "P" else 14) + j]
                assert _close(gm, om, 1e-14) or gm == om, (i, ch, j)
    assert np.all(g[::7, :4].imag == 0)


def _queries(P, M, n, seed, all_masks=False):
    rng = np.random.default_rng(seed)
    pc = [(p, c) for p in M.PROCESSES for c in M.PROCESS_TABLE[p]]
    sel = rng.integers(0, len(pc), n)
    mu = rng.uniform(0, 300, n) / HB
    mu_s = np.where(rng.random(n) < 0.2, mu + rng.uniform(-20, 20, n) / HB, mu)
    masks = rng.choice([0, 0, 0, 1, 2, 4, 8, 16, 32, 5, 10, 48, 63], n) if all_masks else np.zeros(n, int)
    q = P.make_queries([pc[j][0] for j in sel], [pc[j][1] for j in sel], rng.uniform(0, 4, n), rng.uniform(0, 3, n),
                       rng.uniform(100, 250, n) / HB, rng.uniform(0.05, 0.9, n), rng.uniform(0.05, 0.9, n),
                       rng.choice([0.0, 0.3, -0.2], n), rng.uniform(5, 350, n) / HB, rng.uniform(100, 500, n) / HB,
                       mu, mu_s, rng.uniform(-15, -5, n), rng.uniform(-14, -4, n), masks)
    return q


def _gpu_pi(P, q):
    """The GPU's Pi for every slot name a query can need (pnjl_polarization_batch)."""
    from julia_relaxtime_amd import polarization as POL
    from oracle import mesons as M
    out = []
    for name in ("uu_P", "uu_S", "K", "sigma_K", "ss_P", "ss_S"):
        cols = {k: [] for k in ("channel", "num_s", "k0", "k", "m1", "m2", "mu1", "mu2", "T", "Phi", "Phib", "xi",
                                "A1", "A2")}
        for r in q:
            qp = dict(m=dict(u=r["m_u"], s=r["m_s"]), mu=dict(u=r["mu_u"], s=r["mu_s"]), A=dict(u=r["A_u"], s=r["A_s"]))
            p = M.polarization_params(name, qp)
            for k in ("m1", "m2", "mu1", "mu2", "A1", "A2", "num_s"):
                cols[k].append(p[k])
            cols["channel"].append(POL.CHANNEL_P if p["channel"] == "P" else POL.CHANNEL_S)
            for k in ("k0", "k", "T", "Phi", "Phib", "xi"):
                cols[k].append(r[k])
        pq = POL.make_queries(**{k: np.array(v) for k, v in cols.items()})
        re, im = POL.polarization_batch(pq)
        out.append(re + 1j * im)
    return dict(zip(("uu_P", "uu_S", "K", "sigma_K", "ss_P", "ss_S"), out))


def test_by_channel_composition_on_gpu_pi(gpu_ctx, P, M):
    G, K = _GK(gpu_ctx)
    q = _queries(P, M, 1500, 3, all_masks=True)
    DS, DP = P.propagator_by_channel_batch(q)
    pis = _gpu_pi(P, q)
    for i, r in enumerate(q):
        kc = M.effective_couplings(G, K, M.G_from_A(r["A_u"], r["m_u"]), M.G_from_A(r["A_s"], r["m_s"]))
        qp = dict(m=dict(u=r["m_u"], s=r["m_s"]), mu=dict(u=r["mu_u"], s=r["mu_s"]), A=dict(u=r["A_u"], s=r["A_s"]))
        tp = dict(T=r["T"], Phi=r["Phi"], Phib=r["Phib"], xi=r["xi"])
        pi = {k: v[i] for k, v in pis.items()}
        oS, oP = M.propagator_by_channel(M.PROCESSES[r["process"]], "tsu"[r["channel"]], r["k0"], r["k"], qp, tp, kc,
                                         int(r["mesons"]), pi)
        for g, o in ((DS[i], oS), (DP[i], oP)):
            assert abs(g - o) <= 1e-13 * max(abs(o), 1e-300) or g == o, (i, g, o)


def _sensitivity(q, M, G, K, pis):
    """max over a query's mesons of |dD/dPi| |Pi| / |D| (simple: |4KPi/(1-4KPi)|;
    mixed: the cancellation in det M)."""
    s = np.ones(len(q))
    for i, r in enumerate(q):
        c = M.effective_couplings(G, K, M.G_from_A(r["A_u"], r["m_u"]), M.G_from_A(r["A_s"], r["m_s"]))
        for key, name in (("K123_plus", "uu_P"), ("K123_minus", "uu_S"), ("K4567_plus", "K"),
                          ("K4567_minus", "sigma_K")):
            x = 4 * c[key] * pis[name][i]
            s[i] = max(s[i], abs(x) / max(abs(1 - x), 1e-300))
        for ch in "PS":
            M00, M08, M88, _ = M.mixing_matrix_elements(pis[f"uu_{ch}"][i], pis[f"ss_{ch}"][i], c, ch)
            dM = M00 * M88 - M08 * M08
            s[i] = max(s[i], (abs(M00 * M88) + abs(M08 * M08)) / max(abs(dM), 1e-300) * 10)
    return s


def test_by_channel_matches_oracle_end_to_end(gpu_ctx, P, M):
    G, K = _GK(gpu_ctx)
    q = _queries(P, M, 800, 4)
    DS, DP = P.propagator_by_channel_batch(q)
    oS, oP = M.propagator_by_channel_batch(q, G, K)
    sens = _sensitivity(q, M, G, K, _gpu_pi(P, q))
    for g, o in ((DS, oS), (DP, oP)):
        scale = np.maximum(np.abs(o), 1e-3 * np.median(np.abs(o)))
        assert np.all(np.abs(g - o) <= 1e-9 * sens * scale)


def test_explicit_couplings_and_host_api(gpu_ctx, P, M):
    G, K = _GK(gpu_ctx)
    q = _queries(P, M, 200, 5)
    kc = np.array([[M.effective_couplings(G, 0.0, -0.2, -0.1)[k] for k in M.KC_FIELDS]] * len(q))
    DS, DP = P.propagator_by_channel_batch(q, K_coeffs=kc)
    oS, oP = M.propagator_by_channel_batch(q, G, K, kc=kc)
    assert np.all(np.abs(DS - oS) <= 1e-8 * np.maximum(np.abs(oS), 1e-6))
    assert np.all(np.abs(DP - oP) <= 1e-8 * np.maximum(np.abs(oP), 1e-6))
    # scalar forms against the batch
    qp = dict(m=dict(u=320 / HB, s=510 / HB), mu=dict(u=40 / HB, s=40 / HB), A=dict(u=-13.2, s=-10.8))
    tp = dict(T=160 / HB, Phi=0.45, Phibar=0.5, xi=0.3)
    D_S, D_P = P.calculate_propagator_by_channel("uu_to_uu", "t", 1.1, 0.6, qp, tp)
    mP = P.total_propagator_mixed("uu_to_uu", "t", "P", 1.1, 0.6, qp, tp)
    mS = P.total_propagator_mixed("uu_to_uu", "t", "S", 1.1, 0.6, qp, tp)
    sim = P.total_propagator_simple("uu_to_uu", "t", ["pi", "sigma_pi"], 1.1, 0.6, qp, tp)
    assert abs(D_S + D_P - (mP + mS + sim)) <= 1e-13 * abs(D_S + D_P)
    kc1 = P.calculate_effective_couplings(G, K, M.G_from_A(-13.2, 320 / HB), M.G_from_A(-10.8, 510 / HB))
    Pi = 0.02 + 0.003j
    assert P.meson_propagator_simple("pi", kc1, Pi) == pytest.approx(M.meson_propagator_simple("pi", kc1, Pi),
                                                                     rel=1e-14)
    e = P.mixing_matrix_elements(Pi, 0.01, kc1, "S")
    o = M.mixing_matrix_elements(Pi, 0.01, kc1, "S")
    assert (e["M00"], e["M08"], e["M88"], e["det_K"]) == o
    mix = P.meson_propagator_mixed(kc1, Pi, 0.01, "S", "u", "ubar", "s", "sbar", "s")
    assert mix == pytest.approx(M.meson_propagator_mixed(o[3], *o[:3], "u", "ubar", "s", "sbar", "s"), rel=1e-14)
    assert P.meson_propagator_mixed(kc1, Pi, 0.01, "P", "u", "d", "u", "d", "u") == 0
    # errors: the reference's KeyError / ErrorException cases
    with pytest.raises(KeyError):
        P.calculate_propagator_by_channel("uu_to_dd", "t", 1.0, 0.5, qp, tp)
    bad = q[:3].copy()
    bad["channel"][1] = 1                                  # s channel of a quark-quark process
    bad["process"][1] = 0
    from julia_relaxtime_amd._lib import PnjlError
    with pytest.raises(PnjlError):
        P.propagator_by_channel_batch(bad)
    with pytest.raises(ValueError):
        P.meson_propagator_simple("eta", kc1, Pi)


def test_device_pointer_and_one_per_wave_paths(gpu_ctx, P, M):
    """The HBM-resident call (PNJL_FLAG_DEVICE_PTRS | ASYNC on a torch stream)
    and the one-query-per-wavefront FLOP-counting launch give the host call's
    results bit for bit."""
    import torch
    q = _queries(P, M, 700, 6, all_masks=True)
    DS, DP = P.propagator_by_channel_batch(q)
    DS1, DP1 = P.propagator_by_channel_batch(q, one_per_wave=True)
    assert np.array_equal(DS, DS1) and np.array_equal(DP, DP1)
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q.view(np.uint8).copy()).to(dev)
    out = torch.full((q.size, 4), np.nan, dtype=torch.float64, device=dev)
    P.propagator_by_channel_device(qd, out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.array_equal(o[:, 0] + 1j * o[:, 1], DS) and np.array_equal(o[:, 2] + 1j * o[:, 3], DP)


def _mass_states(oracle):
    """(quark_params, thermo_params) at solved equilibrium points; A on GL(0, 20, 16)."""
    out = []
    x, w = oracle.gauleg(0.0, 20.0, 16)
    for TM, mM, xi in ((120, 0, 0.0), (150, 0, 0.0), (150, 100, 0.3), (200, 100, 0.0), (250, 0, -0.2),
                       (100, 250, 0.0), (180, 200, 0.2)):
        T, mu = TM / HB, mM / HB
        r = oracle.solve_batch(np.array([T]), np.array([mu]), xi, policy=oracle.SEED_DEFAULT, p_num=64, t_num=8,
                               opts=oracle.default_opts(ftol=1e-12))
        ms, X = r["masses"][0], r["x"][0]
        qp = dict(m=dict(u=ms[0], s=ms[2]), mu=dict(u=mu, s=mu))
        qp["A"] = dict(u=oracle.A(ms[0], mu, T, X[3], X[4], x, w), s=oracle.A(ms[2], mu, T, X[3], X[4], x, w))
        out.append((qp, dict(T=T, Phi=X[3], Phib=X[4], xi=xi)))
    return out


def test_meson_masses_match_oracle(gpu_ctx, oracle, P, M):
    """Every meson at seven equilibrium points, from the default guess and from
    two other starts: the same convergence / exception outcome as the oracle's
    restatement of NLsolve's trust region; converged poles to 1e-9 relative
    (1e-7 absolute where a rounding difference changed the trust-region path)."""
    G, K = _GK(gpu_ctx)
    rows, ref = [], []
    for qp, tp in _mass_states(oracle):
        for mes in M.MASS_MESONS:
            for m0, g0 in ((None, 0.0), (1.2, 0.0), (0.9, 0.1)):
                rows.append((mes, qp, tp, m0, g0))
                ref.append(M.solve_meson_mass(mes, qp, tp, initial_mass=m0, initial_gamma=g0, G=G, K=K))
    q = np.concatenate([P.make_mass_queries(mes, tp["T"], tp["Phi"], tp["Phib"], tp["xi"], qp["m"]["u"], qp["m"]["s"],
                                            qp["mu"]["u"], qp["mu"]["s"], qp["A"]["u"], qp["A"]["s"],
                                            mass0=np.nan if m0 is None else m0, gamma0=g0)
                        for mes, qp, tp, m0, g0 in rows])
    g = P.meson_mass_batch(q)
    exc = np.array([r["exception"] for r in ref])
    conv = np.array([r["converged"] for r in ref])
    assert np.array_equal(g["exception"], exc)
    assert np.array_equal(g["converged"], conv)
    om, og = np.array([r["mass"] for r in ref]), np.array([r["gamma"] for r in ref])
    it = np.array([r["iterations"] for r in ref])
    c = conv & ~exc
    assert c.sum() > len(rows) // 3
    same_path = c & (g["iterations"] == it)
    tight = np.abs(g["mass"] - om) <= 1e-9 * np.maximum(np.abs(om), 0.1)
    tight &= np.abs(g["gamma"] - og) <= 1e-9 * np.maximum(np.abs(om), 0.1)
    assert np.all(tight[same_path]) and same_path.sum() >= 0.9 * c.sum()
    loose = (np.abs(g["mass"] - om) <= 1e-7) & (np.abs(g["gamma"] - og) <= 1e-7)
    assert np.all(loose[c])
    assert np.all(g["residual_norm"][c] < 1e-8)
    # the host forms
    qp, tp = _mass_states(oracle)[3]
    r = P.solve_meson_mass("sigma", qp, dict(tp, Phibar=tp["Phib"]))
    o = M.solve_meson_mass("sigma", qp, tp, G=G, K=K)
    assert r["converged"] == o["converged"] and abs(r["mass"] - o["mass"]) <= 1e-7
    with pytest.raises(ValueError):
        P.solve_meson_mass("rho", qp, tp)
