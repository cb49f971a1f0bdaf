"""Pins the oracle's consumers of Pi (oracle/mesons.py: effective couplings,
meson propagators, per-process D_S / D_P) before they serve as the checker of
the GPU path:
  * the reference's own unit-test identities (tests/unit/relaxtime/
    test_effective_couplings.jl, test_meson_propagator.jl),
  * closed forms the algebra fixes: the Dyson form of the simple propagator,
    the mixed propagator as 2 J^T (K^-1 - 4 Pi_08)^-1 J' with the 0-8 basis
    polarization matrix built from traces of the lambda matrices, and its
    't Hooft-free limit (K = 0: ideal mixing, u and s decouple),
  * the process table's structure (charge conjugates, the C ABI's order)."""
import math
import os
import re

import numpy as np
import pytest

HB = 197.327
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def M():
    from oracle import mesons
    return mesons


@pytest.fixture(scope="module")
def GK(oracle):
    p = oracle.default_params()
    return p.G, p.K


def _A(oracle, m, mu, T, Phi=0.5, Phib=0.5):
    x, w = oracle.gauleg(0.0, 20.0, 64)
    return oracle.A(m, mu, T, Phi, Phib, x, w)


def test_G_from_A_reference_identity(M):
    # test_effective_couplings.jl:45-57
    A, m = 1.5, 0.8
    G = M.G_from_A(A, m)
    assert G < 0 and abs(G - (-3 / (4.0 * math.pi ** 2) * (m * A))) < 1e-10
    assert M.G_from_A(2 * A, m) == 2 * G and M.G_from_A(A, 2 * m) == 2 * G


def test_couplings_limits_and_determinant(M, GK):
    G, K = GK
    # K = 0 and G^f = 0: every channel degenerates to G  (test_effective_couplings.jl:98-141)
    for c in (M.effective_couplings(G, 0.0, -0.3, -0.2), M.effective_couplings(G, K, 0.0, 0.0)):
        for k in M.KC_FIELDS[:8]:
            assert abs(c[k] - G) < 1e-15
        assert abs(c["K08_plus"]) < 1e-15 and abs(c["K08_minus"]) < 1e-15
    # SU(3)-symmetric condensates: pi, K and eta_8 degenerate, no 0-8 mixing  (:143-160)
    c = M.effective_couplings(G, K, -0.25, -0.25)
    for s in ("plus", "minus"):
        assert abs(c[f"K123_{s}"] - c[f"K4567_{s}"]) < 1e-15 and abs(c[f"K123_{s}"] - c[f"K8_{s}"]) < 1e-15
        assert abs(c[f"K08_{s}"]) < 1e-15
        assert c[f"det_K_{s}"] == M.coupling_matrix_determinant(c[f"K0_{s}"], c[f"K8_{s}"], c[f"K08_{s}"])
    assert abs(M.coupling_matrix_determinant(1e-5, 1.1e-5, 1e-7) - (1e-5 * 1.1e-5 - 1e-14)) < 1e-20
    # magnitudes  (:189-214)
    c = M.effective_couplings(G, K, -0.3, -0.2)
    for k, v in c.items():
        if k.startswith("det"):
            assert np.isfinite(v) and v > 0
        elif k.startswith("K08"):
            assert abs(v) < 0.1 * G
        else:
            assert abs(v - G) / G < 0.5
    # the P / S couplings differ only in the sign of the 't Hooft terms
    for a in ("K0", "K123", "K4567", "K8"):
        assert abs((c[f"{a}_plus"] + c[f"{a}_minus"]) / 2 - G) < 1e-16
    assert c["K08_plus"] == -c["K08_minus"]


@pytest.mark.parametrize("T_MeV,mu_MeV", [(50, 0), (100, 0), (150, 0), (200, 0), (250, 0), (150, 100), (150, 200),
                                          (150, 300)])
def test_detK_positive_on_reference_scan(oracle, M, GK, T_MeV, mu_MeV):
    """test_effective_couplings.jl:165-187, :216-290 (A with gauleg(0, 20, 64),
    m_u = 300, m_s = 500 MeV, Phi = Phibar = 0.5)."""
    G, K = GK
    T, mu, mu_, ms = T_MeV / HB, mu_MeV / HB, 300 / HB, 500 / HB
    Gu, Gs = M.G_from_A(_A(oracle, mu_, mu, T), mu_), M.G_from_A(_A(oracle, ms, mu, T), ms)
    c = M.effective_couplings(G, K, Gu, Gs)
    assert np.isfinite(Gu) and np.isfinite(Gs) and np.isfinite(c["K123_minus"])
    assert c["det_K_plus"] > 0 and c["det_K_minus"] > 0


def _kc(oracle, M, GK):
    G, K = GK
    T, mu_, ms = 150 / HB, 300 / HB, 500 / HB
    return M.effective_couplings(G, K, M.G_from_A(_A(oracle, mu_, 0.0, T), mu_), M.G_from_A(_A(oracle, ms, 0.0, T), ms))


def test_simple_propagator(oracle, M, GK):
    kc = _kc(oracle, M, GK)
    rng = np.random.default_rng(11)
    key = {"pi": "K123_plus", "K": "K4567_plus", "sigma_pi": "K123_minus", "sigma_K": "K4567_minus"}
    for _ in range(200):
        Pi = complex(rng.normal(0, 2), rng.normal(0, 1))
        for m, k in key.items():
            D = M.meson_propagator_simple(m, kc, Pi)
            Kv = kc[k]
            # Dyson form: D (1 - 4 K Pi) = 2 K
            assert abs(D * (1 - 4 * Kv * Pi) - 2 * Kv) <= 1e-14 * abs(2 * Kv) * max(1.0, abs(4 * Kv * Pi))
    # real Pi -> real D; nearer the pole -> larger |D|  (test_meson_propagator.jl:230-244)
    Ds, Dl = M.meson_propagator_simple("pi", kc, 1e-6), M.meson_propagator_simple("pi", kc, 5e-5)
    assert abs(Dl) > abs(Ds) and abs(Ds.imag) < 1e-10
    # exactly on the pole: the denominator is moved by i * 1e-12 (MesonPropagator.jl:285-287)
    Kp = kc["K123_plus"]
    Pi0 = 1 / (4 * Kp)
    if 1.0 - 4.0 * Kp * Pi0 == 0.0:
        D = M.meson_propagator_simple("pi", kc, Pi0)
        assert D == (2 * Kp) / complex(0.0, 1e-12)
    assert M.meson_propagator_simple("pi", kc, complex(np.inf, 0)) == 0
    assert M.meson_propagator_simple("K", kc, complex(np.nan, 0)) == 0
    with pytest.raises(ValueError):
        M.meson_propagator_simple("invalid", kc, 1e-5)


def test_current_vectors(M):
    """test_meson_propagator.jl:114-131 and the lambda entries."""
    assert M.current_vector("u", "d") == (0.0, 0.0)
    assert M.current_vector("u", "dbar") == (0.0, 0.0)
    assert M.current_vector("u", "ubar") == (math.sqrt(2 / 3), 1 / math.sqrt(3))
    assert M.current_vector("d", "d") == (math.sqrt(2 / 3), 1 / math.sqrt(3))
    assert M.current_vector("sbar", "s") == (math.sqrt(2 / 3), -2 / math.sqrt(3))


def _pi08(Puu, Pss, M):
    """0-8 basis polarization: Pi_ab = tr(lambda_a Pi lambda_b) / 2 with Pi = diag(Puu, Puu, Pss)."""
    P = np.diag([Puu, Puu, Pss])
    L = (M.LAMBDA0, M.LAMBDA8)
    return np.array([[np.trace(L[a] @ P @ L[b]) / 2 for b in range(2)] for a in range(2)])


def test_mixed_propagator_is_rpa_matrix_inverse(M, GK):
    """D = 2 det K / det M J^T M J' equals 2 J^T (K^-1 - 4 Pi_08)^-1 J' for
    every current pair, channel and random complex Pi: fixes the three M
    elements' coefficients (4/3, 4 sqrt2 / 3), their signs and the 0/8 index
    order."""
    G, K = GK
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(300):
        c = M.effective_couplings(G, K, rng.uniform(-3, 3), rng.uniform(-3, 3))
        Puu, Pss = complex(*rng.normal(0, 1, 2)), complex(*rng.normal(0, 1, 2))
        for ch, s in (("P", "plus"), ("S", "minus")):
            Km = np.array([[c[f"K0_{s}"], c[f"K08_{s}"]], [c[f"K08_{s}"], c[f"K8_{s}"]]])
            X = np.linalg.inv(np.linalg.inv(Km) - 4 * _pi08(Puu, Pss, M))
            M00, M08, M88, dK = M.mixing_matrix_elements(Puu, Pss, c, ch)
            for q in (("u", "u", "u", "u"), ("u", "s", "u", "s"), ("s", "s", "s", "s"), ("d", "ubar", "d", "ubar")):
                for chan in "tsu":
                    D = M.meson_propagator_mixed(dK, M00, M08, M88, *q, chan)
                    legs = {"t": ((0, 2), (1, 3)), "s": ((0, 1), (2, 3)), "u": ((0, 3), (1, 2))}[chan]
                    J = np.array(M.current_vector(q[legs[0][0]], q[legs[0][1]]))
                    Jp = np.array(M.current_vector(q[legs[1][0]], q[legs[1][1]]))
                    ref = 2 * (J @ X @ Jp)
                    scale = np.abs(X).max() * 2
                    worst = max(worst, abs(D - ref) / scale)
    assert worst < 1e-10, worst


def test_mixed_propagator_without_t_hooft_term(M, GK):
    """K = 0: the 0-8 system is ideally mixed -- u ubar and s sbar propagate
    separately (2G / (1 - 4 G Pi_uu), 2 x 2G / (1 - 4 G Pi_ss), nothing
    between them)."""
    G, _ = GK
    c = M.effective_couplings(G, 0.0, -0.4, -0.1)
    rng = np.random.default_rng(9)
    for _ in range(50):
        Puu, Pss = complex(*rng.normal(0, 1, 2)), complex(*rng.normal(0, 1, 2))
        for ch in "PS":
            M00, M08, M88, dK = M.mixing_matrix_elements(Puu, Pss, c, ch)
            duu = M.meson_propagator_mixed(dK, M00, M08, M88, "u", "u", "u", "u", "t")
            dss = M.meson_propagator_mixed(dK, M00, M08, M88, "s", "s", "s", "s", "t")
            dus = M.meson_propagator_mixed(dK, M00, M08, M88, "u", "u", "s", "s", "t")
            assert abs(duu - 2 * G / (1 - 4 * G * Puu)) <= 1e-12 * abs(duu)
            assert abs(dss - 4 * G / (1 - 4 * G * Pss)) <= 1e-12 * abs(dss)
            assert abs(dus) <= 1e-12 * (abs(duu) + abs(dss))
    with pytest.raises(ValueError):
        M.meson_propagator_mixed(dK, M00, M08, M88, "u", "d", "u", "d", "x")


def test_mixing_matrix_reference_values(oracle, M, GK):
    """test_meson_propagator.jl:134-168: symmetric M, the documented
    coefficient 4 sqrt2 / 3 ~ 1.8856 of M08."""
    kc = _kc(oracle, M, GK)
    Puu, Pss = 1.0e-5 + 0j, 8.0e-6 + 0j
    M00, M08, M88, dK = M.mixing_matrix_elements(Puu, Pss, kc, "P")
    assert abs(M08 - (kc["K08_plus"] + 1.8856180831641267 * dK * (Puu - Pss))) < 1e-15
    assert abs(M00 - kc["K0_plus"]) < 1e-4 and abs(M88 - kc["K8_plus"]) < 1e-4
    D = {q: M.meson_propagator_mixed(dK, M00, M08, M88, *q, "t") for q in (("u",) * 4, ("s",) * 4, ("u", "s") * 2)}
    assert D[("u",) * 4] != D[("s",) * 4] and D[("u",) * 4] != D[("u", "s") * 2]
    assert all(np.isfinite(v) and abs(v) > 1e-10 for v in (D[("u",) * 4], D[("s",) * 4]))


def test_process_table(M):
    assert len(M.PROCESSES) == 17
    for p in M.PROCESSES:
        q = M.parse_process(p)
        assert "".join(q[:2]) + "_to_" + "".join(q[2:]) == p
        if not any(x.endswith("bar") for x in q):
            assert set(M.PROCESS_TABLE[p]) == {"t", "u"}      # quark-quark: t and u channels
        elif all(x.endswith("bar") for x in q):
            quark = "_to_".join(2 * ["".join(x[0] for x in q[:2])])
            assert M.PROCESS_TABLE[p] == M.PROCESS_TABLE[quark]   # charge conjugate of a qq process
        else:
            assert set(M.PROCESS_TABLE[p]) == {"t", "s"}      # quark-antiquark: t and s channels
    # same names and order as the C ABI enum and the host mirror
    hdr = open(os.path.join(ROOT, "include", "pnjl_hip.h")).read()
    enum = re.search(r"enum \{\s*(PNJL_PROC_UD_TO_UD.*?)PNJL_PROC_COUNT", hdr, re.S).group(1)
    names = [n.strip()[len("PNJL_PROC_"):].lower() for n in enum.replace("= 0", "").split(",") if n.strip()]
    assert names == M.PROCESSES
    from julia_relaxtime_amd import propagators as P
    assert P.PROCESSES == M.PROCESSES and P.KC_FIELDS == M.KC_FIELDS


def test_flavour_factors(M):
    """Table 5.3 (TotalPropagator.jl:239-275)."""
    assert (M.flavour_factor("u", "ubar"), M.flavour_factor("d", "d"), M.flavour_factor("s", "sbar")) == (1, -1, 2)
    assert M.flavour_factor("u", "s") == math.sqrt(2) == M.flavour_factor("d", "u")
    assert M.flavour_factors_for_channel("ud_to_ud", "t") == (1.0, -1.0)
    assert M.flavour_factors_for_channel("ud_to_ud", "u") == (math.sqrt(2), math.sqrt(2))
    assert M.flavour_factors_for_channel("uubar_to_ssbar", "s") == (1.0, 2.0)
    assert M.flavour_factors_for_channel("uubar_to_ssbar", "t") == (math.sqrt(2), math.sqrt(2))


def test_by_channel_composition(oracle, M, GK):
    """D_S / D_P = flavour-weighted simple mesons + mixed propagators, with the
    table's meson sets; charge-conjugate and isospin-mirror processes agree."""
    kc = _kc(oracle, M, GK)
    T = 150 / HB
    qp = dict(m=dict(u=300 / HB, s=500 / HB), mu=dict(u=30 / HB, s=30 / HB), A=dict(u=-13.0, s=-11.0))
    tp = dict(T=T, Phi=0.5, Phib=0.6, xi=0.2)
    k0, k = 1.3, 0.7
    res = {}
    for p in M.PROCESSES:
        for ch in M.PROCESS_TABLE[p]:
            res[p, ch] = M.propagator_by_channel(p, ch, k0, k, qp, tp, kc)
            simple, mP, mS = M.PROCESS_TABLE[p][ch]
            parts = [M.propagator_by_channel(p, ch, k0, k, qp, tp, kc, mesons=M.MESON_BITS[m]) for m in simple]
            mix = [M.propagator_by_channel(p, ch, k0, k, qp, tp, kc, mesons=b) for on, b in ((mP, 16), (mS, 32)) if on]
            DS = sum(x[0] for x in parts + mix)
            DP = sum(x[1] for x in parts + mix)
            assert abs(res[p, ch][0] - DS) <= 1e-14 * max(abs(DS), 1e-30)
            assert abs(res[p, ch][1] - DP) <= 1e-14 * max(abs(DP), 1e-30)
    for qq, bar in (("ud_to_ud", "ubardbar_to_ubardbar"), ("us_to_us", "ubarsbar_to_ubarsbar"),
                    ("uu_to_uu", "ubarubar_to_ubarubar"), ("ss_to_ss", "sbarsbar_to_sbarsbar")):
        for ch in "tu":
            assert res[qq, ch] == res[bar, ch]
    # mirrored legs: the same terms, the bilinear form J^T M J' taken in the other order
    for a, b in (("udbar_to_udbar", "dubar_to_dubar"), ("usbar_to_usbar", "subar_to_subar")):
        for ch in "ts":
            for x, y in zip(res[a, ch], res[b, ch]):
                assert abs(x - y) <= 1e-14 * abs(x)
    # pi exchange in ud -> ud (t): flavour factors 1 x -1
    DSp, DPp = M.propagator_by_channel("ud_to_ud", "t", k0, k, qp, tp, kc, mesons=1)
    assert DSp == 0 and DPp == -M.meson_propagator_simple("pi", kc, complex(*oracle.polarization(
        oracle.CHANNEL_P, k0, k, qp["m"]["u"], qp["m"]["u"], qp["mu"]["u"], qp["mu"]["u"], T, 0.5, 0.6, 0.2,
        qp["A"]["u"], qp["A"]["u"], 0)))


def test_vectorised_form_matches_scalar_restatement(oracle, M, GK):
    """The numpy form (the CPU baseline's timed path) reproduces the scalar,
    reference-shaped restatement query by query."""
    from julia_relaxtime_amd import propagators as P
    G, K = GK
    rng = np.random.default_rng(21)
    n = 300
    pc = [(p, c) for p in M.PROCESSES for c in M.PROCESS_TABLE[p]]
    sel = rng.integers(0, len(pc), n)
    mu = rng.uniform(0, 300, n) / HB
    q = P.make_queries([pc[j][0] for j in sel], [pc[j][1] for j in sel], rng.uniform(0, 4, n), rng.uniform(0, 3, n),
                       rng.uniform(100, 250, n) / HB, rng.uniform(0.05, 0.9, n), rng.uniform(0.05, 0.9, n),
                       rng.choice([0.0, 0.3], n), rng.uniform(5, 350, n) / HB, rng.uniform(100, 500, n) / HB, mu,
                       np.where(rng.random(n) < 0.3, mu + 0.05, mu), rng.uniform(-15, -5, n), rng.uniform(-14, -4, n),
                       rng.choice([0, 0, 0, 1, 2, 16, 32, 63], n))
    s_S, s_P = M.propagator_by_channel_batch(q, G, K, nthreads=4)
    v_S, v_P = M.propagator_by_channel_vec(q, G, K, nthreads=4)
    for a, b in ((s_S, v_S), (s_P, v_P)):
        assert np.all(np.abs(a - b) <= 1e-13 * np.maximum(np.abs(a), 1e-300)), np.max(np.abs(a - b) / np.abs(a))


# ---- pole masses (MesonMass.jl / MottTransition.jl) ---------------------------------
def _qp_with_A(oracle, m_u, m_d, m_s, T, Phi, Phib, mu=0.0):
    """ensure_quark_params_has_A (MesonMass.jl:88-120) at xi = 0: A on GL(0, 20, 16)."""
    x, w = oracle.gauleg(0.0, 20.0, 16)
    qp = dict(m=dict(u=m_u, d=m_d, s=m_s), mu=dict(u=mu, d=mu, s=mu))
    qp["A"] = {f: oracle.A(qp["m"][f], mu, T, Phi, Phib, x, w) for f in "uds"}
    return qp


def test_mass_guess_and_mott_thresholds(oracle, M):
    """test_meson_mass_mott_transition.jl:19-49."""
    from julia_relaxtime_amd import propagators as P
    qp = dict(m=dict(u=0.3, d=0.31, s=0.5), mu=dict(u=0.0, d=0.0, s=0.0))
    assert P.default_meson_mass_guess("pi", qp) == 0.3 + 0.31
    assert P.default_meson_mass_guess("K", qp) == 0.3 + 0.5
    assert P.default_meson_mass_guess("eta", qp) == 0.6 and P.default_meson_mass_guess("eta_prime", qp) == 1.0
    thr = P.mott_threshold_mass("K", qp)
    assert thr == 0.8 and P.is_mott_point("K", thr, qp, atol=1e-12)
    assert P.mott_threshold_mass("eta", qp) == min(0.6, 1.0) and P.mott_gap("sigma", 0.7, qp) == 0.7 - 0.6
    qp_ud = dict(m=dict(u=0.3, s=0.5))
    for mes in M.MASS_MESONS:
        assert M.default_meson_mass_guess(mes, qp_ud) == P.default_meson_mass_guess(mes, qp_ud)


def test_mixed_residual_is_twice_the_eigenvalues(oracle, M, GK):
    """The eta / eta' (sigma / sigma') residuals are the two eigenvalues of 2M
    (numpy eigvals), their product 4 det M, and at an SU(3)-symmetric point
    M08 = 0 (test_meson_mass_mott_transition.jl:51-69)."""
    G, K = GK
    qp = _qp_with_A(oracle, 0.35, 0.35, 0.35, 0.12, 0.6, 0.6)
    tp = dict(T=0.12, Phi=0.6, Phib=0.6, xi=0.0)
    kc = M.effective_couplings(G, K, M.G_from_A(qp["A"]["u"], 0.35), M.G_from_A(qp["A"]["s"], 0.35))
    Puu = M._pi_width("P", "uu", 0.7, 0.0, 0.0, qp, tp)
    Pss = M._pi_width("P", "ss", 0.7, 0.0, 0.0, qp, tp)
    M00, M08, M88, _ = M.mixing_matrix_elements(Puu, Pss, kc, "P")
    assert abs(M08) <= 1e-10
    fe = M.meson_mass_equation("eta", 0.7, 0.0, 0.0, qp, tp, kc)
    fp = M.meson_mass_equation("eta_prime", 0.7, 0.0, 0.0, qp, tp, kc)
    assert abs(fe * fp - 4.0 * (M00 * M88 - M08 * M08)) <= 1e-10 * abs(fe * fp)
    qp = _qp_with_A(oracle, 0.3, 0.3, 0.5, 0.75, 0.5, 0.55, mu=0.2)
    tp = dict(T=0.75, Phi=0.5, Phib=0.55, xi=0.0)
    kc = M.effective_couplings(G, K, M.G_from_A(qp["A"]["u"], 0.3), M.G_from_A(qp["A"]["s"], 0.5))
    for k0, gam, ch, pair in ((0.9, 0.0, "P", ("eta", "eta_prime")), (1.7, 0.3, "S", ("sigma", "sigma_prime")),
                              (2.6, 0.8, "P", ("eta", "eta_prime"))):
        Puu = M._pi_width(ch, "uu", k0, gam, 0.0, qp, tp)
        Pss = M._pi_width(ch, "ss", k0, gam, 0.0, qp, tp)
        M00, M08, M88, _ = M.mixing_matrix_elements(Puu, Pss, kc, ch)
        ev = 2 * np.linalg.eigvals(np.array([[M00, M08], [M08, M88]]))
        f = [M.meson_mass_equation(m, k0, gam, 0.0, qp, tp, kc) for m in pair]
        assert sorted(abs(f[0] - e) for e in ev)[0] <= 1e-12 * abs(f[0]) + 1e-15
        assert sorted(abs(f[1] - e) for e in ev)[0] <= 1e-12 * abs(f[1]) + 1e-15
        assert abs(f[0] * f[1] - 4.0 * (M00 * M88 - M08 * M08)) <= 1e-10 * abs(f[0] * f[1])


def test_pole_solver_finds_the_dyson_pole(oracle, M, GK):
    """Converged (mass, gamma) are zeros of the residual to ftol; the central-
    difference Jacobian matches the analytic gamma derivative of
    polarization_with_width (Pi is a quadratic in gamma at fixed k0); the
    solve is independent of the starting point within the basin."""
    G, K = GK
    HBm = 197.327
    T = 200 / HBm
    qp = _qp_with_A(oracle, 80 / HBm, 80 / HBm, 400 / HBm, T, 0.5, 0.5, mu=100 / HBm)
    tp = dict(T=T, Phi=0.5, Phib=0.5, xi=0.0)
    kc = M.effective_couplings(G, K, M.G_from_A(qp["A"]["u"], qp["m"]["u"]), M.G_from_A(qp["A"]["s"], qp["m"]["s"]))
    for mes in ("pi", "sigma_pi", "eta", "sigma_prime"):
        r = M.solve_meson_mass(mes, qp, tp, kc=kc)
        assert r["converged"] and r["residual_norm"] < 1e-8
        f = M.meson_mass_equation(mes, r["mass"], r["gamma"], 0.0, qp, tp, kc)
        assert abs(f.real) < 1e-8 and abs(f.imag) < 1e-8
        r2 = M.solve_meson_mass(mes, qp, tp, kc=kc, initial_mass=r["mass"] * 1.05, initial_gamma=r["gamma"])
        assert r2["converged"] and abs(r2["mass"] - r["mass"]) < 1e-7 * abs(r["mass"])
    # d/dgamma of 1 - 4 K Pi: Pi = c [A12 + (k^2 - lam^2 + (m1 -/+ m2)^2 + g^2/4) B + i g lam B]
    k0, g = 1.1, 0.2
    Pa = M._pi_width("P", "uu", k0, g, 0.0, qp, tp)
    Pb = M._pi_width("P", "uu", k0, 0.0, 0.0, qp, tp)
    h = 1e-3
    Pp, Pm = M._pi_width("P", "uu", k0, g + h, 0.0, qp, tp), M._pi_width("P", "uu", k0, g - h, 0.0, qp, tp)
    quad = (Pp - 2 * Pa + Pm) / (h * h)             # = 2c for Pi = a + b g + c g^2
    lin = (Pp - Pm) / (2 * h)                       # = b + 2 c g
    assert abs((Pa - Pb) - (lin - quad * g) * g - quad * g * g / 2) <= 1e-9 * abs(Pa - Pb)
    # the solver's own central difference at its step agrees with the wide-step derivative of the quadratic
    step = max(M.FD_STEP * abs(g), M.FD_STEP)
    fd = (M._pi_width("P", "uu", k0, g + step, 0.0, qp, tp) - M._pi_width("P", "uu", k0, g - step, 0.0, qp, tp)) / (
        2 * step)
    assert abs(fd - lin) <= 1e-6 * abs(lin)
