"""CPU oracle of the consumers of the polarization Pi: effective couplings,
meson propagators and the per-process D_S / D_P.

TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path
(pnjl_effective_couplings_batch, pnjl_meson_propagators_batch,
pnjl_propagator_by_channel_batch) is compared against; only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it.

A Python restatement, with complex scalars, of
  src/relaxtime/EffectiveCouplings.jl   calculate_G_from_A :56-60,
      calculate_effective_couplings :240-278, coupling_matrix_determinant
      :111-115, mixing_matrix_elements :137-160
  src/relaxtime/MesonPropagator.jl      calculate_current_vector :133-141,
      meson_propagator_simple :268-291, meson_propagator_mixed :365-419
  src/relaxtime/TotalPropagator.jl      get_flavor_factor :239-258,
      get_flavor_factors_for_channel :260-275, get_polarization_params
      :308-336, total_propagator_simple :423-470, total_propagator_mixed
      :514-560, calculate_propagator_by_channel :894-963
  src/Constants_PNJL.jl                 lambda0 / lambda8 :106-117, the
      process table SCATTERING_MESON_MAP :168-400
The Pi values come from the oracle's C restatement of polarization_aniso
(oracle.polarization_batch).  Flavour structure is built the reference's way
(lambda matrices between flavour vectors), not from a precomputed table.
"""
from __future__ import annotations

import cmath
import math

import numpy as np

from . import oracle as O

DENOM_EPS = 1e-12                      # PROPAGATOR_DENOM_EPS  MesonPropagator.jl:22

# Constants_PNJL.jl:106-117
LAMBDA0 = np.eye(3) * math.sqrt(2 / 3)
LAMBDA8 = np.diag([1.0, 1.0, -2.0]) / math.sqrt(3)
PSI = {"u": np.array([1.0, 0.0, 0.0]), "d": np.array([0.0, 1.0, 0.0]), "s": np.array([0.0, 0.0, 1.0])}

KC_FIELDS = ["K0_plus", "K0_minus", "K123_plus", "K123_minus", "K4567_plus", "K4567_minus", "K8_plus", "K8_minus",
             "K08_plus", "K08_minus", "det_K_plus", "det_K_minus"]

# SCATTERING_MESON_MAP (Constants_PNJL.jl:168-400), in its order:
# process -> {channel: (simple mesons, mixed_P, mixed_S)}
_PI = ("pi", "sigma_pi")
_K = ("K", "sigma_K")
PROCESS_TABLE = {
    "ud_to_ud": {"t": (_PI, True, True), "u": (_PI, False, False)},
    "uu_to_uu": {"t": (_PI, True, True), "u": (_PI, True, True)},
    "us_to_us": {"t": ((), True, True), "u": (_K, False, False)},
    "ss_to_ss": {"t": ((), True, True), "u": ((), True, True)},
    "ubardbar_to_ubardbar": {"t": (_PI, True, True), "u": (_PI, False, False)},
    "ubarubar_to_ubarubar": {"t": (_PI, True, True), "u": (_PI, True, True)},
    "ubarsbar_to_ubarsbar": {"t": ((), True, True), "u": (_K, False, False)},
    "sbarsbar_to_sbarsbar": {"t": ((), True, True), "u": ((), True, True)},
    "udbar_to_udbar": {"t": (_PI, True, True), "s": (_PI, False, False)},
    "dubar_to_dubar": {"t": (_PI, True, True), "s": (_PI, False, False)},
    "uubar_to_uubar": {"t": (_PI, True, True), "s": (_PI, True, True)},
    "uubar_to_ddbar": {"t": (_PI, False, False), "s": (_PI, True, True)},
    "usbar_to_usbar": {"t": ((), True, True), "s": (_K, False, False)},
    "subar_to_subar": {"t": ((), True, True), "s": (_K, False, False)},
    "uubar_to_ssbar": {"t": (_K, False, False), "s": ((), True, True)},
    "ssbar_to_uubar": {"t": (_K, False, False), "s": ((), True, True)},
    "ssbar_to_ssbar": {"t": ((), True, True), "s": ((), True, True)},
}
PROCESSES = list(PROCESS_TABLE)
CHANNELS = {"t": 0, "s": 1, "u": 2}
MESON_BITS = {"pi": 1, "K": 2, "sigma_pi": 4, "sigma_K": 8, "mixed_P": 16, "mixed_S": 32}


def parse_process(name):
    """'udbar_to_udbar' -> ('u', 'dbar', 'u', 'dbar')  (ParticleSymbols.parse_scattering_process)."""
    out = []
    for side in name.split("_to_"):
        i = 0
        while i < len(side):
            f = side[i]
            if side.startswith("bar", i + 1):
                out.append(f + "bar")
                i += 4
            else:
                out.append(f)
                i += 1
    assert len(out) == 4, name
    return tuple(out)


def flavour(q):
    return q[0]


# ---- EffectiveCouplings.jl --------------------------------------------------
def G_from_A(A, m, Nc=3):
    """calculate_G_from_A  EffectiveCouplings.jl:56-60."""
    return -Nc / (4.0 * (math.pi * math.pi)) * (m * A)     # Julia's literal x^2 is x * x


def coupling_matrix_determinant(K0, K8, K08):
    """EffectiveCouplings.jl:111-115."""
    return K0 * K8 - K08 * K08


def effective_couplings(G, K, Gu, Gs):
    """calculate_effective_couplings  EffectiveCouplings.jl:240-278."""
    t0 = (1.0 / 3.0) * K * (2.0 * Gu + Gs)
    t123 = 0.5 * K * Gs
    t4567 = 0.5 * K * Gu
    t8 = (1.0 / 6.0) * K * (4.0 * Gu - Gs)
    t08 = (1.0 / 6.0) * math.sqrt(2.0) * K * (Gu - Gs)
    c = dict(K0_plus=G - t0, K0_minus=G + t0, K123_plus=G + t123, K123_minus=G - t123, K4567_plus=G + t4567,
             K4567_minus=G - t4567, K8_plus=G + t8, K8_minus=G - t8, K08_plus=t08, K08_minus=-t08)
    c["det_K_plus"] = coupling_matrix_determinant(c["K0_plus"], c["K8_plus"], c["K08_plus"])
    c["det_K_minus"] = coupling_matrix_determinant(c["K0_minus"], c["K8_minus"], c["K08_minus"])
    return c


def mixing_matrix_elements(Puu, Pss, kc, channel):
    """EffectiveCouplings.jl:137-160 -> (M00, M08, M88, det_K)."""
    sfx = {"P": "plus", "S": "minus"}[channel]
    K0, K8, K08, dK = (kc[f"{k}_{sfx}"] for k in ("K0", "K8", "K08", "det_K"))
    M00 = K0 - (4.0 / 3.0) * dK * (Puu + 2.0 * Pss)
    M08 = K08 + (4.0 / 3.0) * math.sqrt(2.0) * dK * (Puu - Pss)
    M88 = K8 - (4.0 / 3.0) * dK * (2.0 * Puu + Pss)
    return M00, M08, M88, dK


# ---- MesonPropagator.jl -----------------------------------------------------
def current_vector(q1, q2):
    """(psibar_{q2} lambda0 psi_{q1}, psibar_{q2} lambda8 psi_{q1})  MesonPropagator.jl:133-141."""
    a, b = PSI[flavour(q1)], PSI[flavour(q2)]
    return float(b @ LAMBDA0 @ a), float(b @ LAMBDA8 @ a)


def _finite(z):
    return math.isfinite(z.real) and math.isfinite(z.imag)


def meson_propagator_simple(meson, kc, Pi):
    """MesonPropagator.jl:268-291: D = 2K / (1 - 4 K Pi)."""
    key = {"pi": "K123_plus", "K": "K4567_plus", "sigma_pi": "K123_minus", "sigma_K": "K4567_minus"}
    if meson not in key:
        raise ValueError(f"Unknown meson type: {meson}")
    K = kc[key[meson]]
    den = 1.0 - 4.0 * K * complex(Pi)
    if not _finite(den):
        return 0j
    if abs(den) < DENOM_EPS:
        den += complex(0.0, DENOM_EPS)
    return (2.0 * K) / den


def meson_propagator_mixed(det_K, M00, M08, M88, q1, q2, q3, q4, channel):
    """MesonPropagator.jl:365-419: D = 2 det K / det M * J^T M J'."""
    legs = {"t": ((q1, q3), (q2, q4)), "s": ((q1, q2), (q3, q4)), "u": ((q1, q4), (q2, q3))}
    if channel not in legs:
        raise ValueError(f"Unknown channel: {channel}")
    (a, b), (c, d) = legs[channel]
    J0, J8 = current_vector(a, b)
    Jp0, Jp8 = current_vector(c, d)
    det_M = M00 * M88 - M08 * M08
    if not _finite(det_M):
        return 0j
    if abs(det_M) < DENOM_EPS:
        det_M += complex(0.0, DENOM_EPS)
    r = J0 * (M00 * Jp0 + M08 * Jp8) + J8 * (M08 * Jp0 + M88 * Jp8)
    if not _finite(r):
        return 0j
    return 2.0 * det_K / det_M * r


# ---- TotalPropagator.jl -----------------------------------------------------
def flavour_factor(q1, q2):
    """Table 5.3  TotalPropagator.jl:239-258."""
    f1, f2 = flavour(q1), flavour(q2)
    if f1 == f2:
        return {"u": 1.0, "d": -1.0, "s": 2.0}[f1]
    return math.sqrt(2.0)


def flavour_factors_for_channel(process, channel):
    """TotalPropagator.jl:260-275."""
    q1, q2, q3, q4 = parse_process(process)
    pairs = {"t": ((q1, q3), (q2, q4)), "u": ((q1, q4), (q2, q3)), "s": ((q1, q2), (q3, q4))}
    (a, b), (c, d) = pairs[channel]
    return flavour_factor(a, b), flavour_factor(c, d)


def polarization_params(meson, qp):
    """get_polarization_params  TotalPropagator.jl:308-336 (+ the s sbar pair
    of total_propagator_mixed :530-537 as 'ss_P' / 'ss_S')."""
    u, s = "u", "s"
    table = {"pi": ("P", u, u, 0), "K": ("P", u, s, 1), "sigma_pi": ("S", u, u, 0), "sigma_K": ("S", u, s, 1),
             "uu_P": ("P", u, u, 0), "uu_S": ("S", u, u, 0), "ss_P": ("P", s, s, 2), "ss_S": ("S", s, s, 2)}
    ch, f1, f2, ns = table[meson]
    return dict(channel=ch, m1=qp["m"][f1], m2=qp["m"][f2], mu1=qp["mu"][f1], mu2=qp["mu"][f2], A1=qp["A"][f1],
                A2=qp["A"][f2], num_s=ns)


def _mesons_of(process, channel, mask):
    if mask == 0:
        simple, mP, mS = PROCESS_TABLE[process][channel]
        return list(simple), mP, mS
    simple = [m for m in ("pi", "K", "sigma_pi", "sigma_K") if mask & MESON_BITS[m]]
    return simple, bool(mask & 16), bool(mask & 32)


def propagator_by_channel(process, channel, k0, k, qp, tp, kc, mesons=0, pi=None):
    """calculate_propagator_by_channel  TotalPropagator.jl:894-963 -> (D_S, D_P).
    qp = dict(m=..., mu=..., A=...) keyed by flavour, tp = dict(T, Phi, Phib, xi);
    pi: optional {name: complex} of precomputed Pi (names of polarization_params)."""
    simple, mP, mS = _mesons_of(process, channel, mesons)
    T1, T2 = flavour_factors_for_channel(process, channel)
    pi = pi if pi is not None else {}

    def Pi(name):
        if name not in pi:
            p = polarization_params(name, qp)
            re, im = O.polarization(O.CHANNEL_P if p["channel"] == "P" else O.CHANNEL_S, k0, k, p["m1"], p["m2"],
                                    p["mu1"], p["mu2"], tp["T"], tp["Phi"], tp["Phib"], tp["xi"], p["A1"], p["A2"],
                                    p["num_s"])
            pi[name] = complex(re, im)
        return pi[name]

    DS, DP = 0j, 0j
    for m in simple:
        D = meson_propagator_simple(m, kc, Pi({"pi": "uu_P", "sigma_pi": "uu_S", "K": "K", "sigma_K": "sigma_K"}[m]))
        if m in ("pi", "K"):
            DP += T1 * D * T2
        else:
            DS += T1 * D * T2
    q = parse_process(process)
    for on, ch in ((mP, "P"), (mS, "S")):
        if on:
            M00, M08, M88, dK = mixing_matrix_elements(Pi(f"uu_{ch}"), Pi(f"ss_{ch}"), kc, ch)
            D = meson_propagator_mixed(dK, M00, M08, M88, *q, channel)
            if ch == "P":
                DP += D
            else:
                DS += D
    return DS, DP


def needed_pi(process, channel, mesons=0):
    """Names of the Pi values a query evaluates."""
    simple, mP, mS = _mesons_of(process, channel, mesons)
    names = {"pi": "uu_P", "sigma_pi": "uu_S", "K": "K", "sigma_K": "sigma_K"}
    out = {names[m] for m in simple}
    if mP:
        out |= {"uu_P", "ss_P"}
    if mS:
        out |= {"uu_S", "ss_S"}
    return sorted(out)


def propagator_by_channel_batch(q, G, K, kc=None, nthreads=8, Nc=3):
    """(D_S, D_P) for a structured array of queries (fields of
    pnjl_prop_query); all Pi values in one multi-threaded oracle batch.
    kc = None derives the couplings from A as the reference's callers do."""
    n = len(q)
    names = [needed_pi(PROCESSES[int(r["process"])], "tsu"[int(r["channel"])], int(r["mesons"])) for r in q]
    rows = []
    for i, r in enumerate(q):
        qp = dict(m=dict(u=r["m_u"], s=r["m_s"]), mu=dict(u=r["mu_u"], s=r["mu_s"]), A=dict(u=r["A_u"], s=r["A_s"]))
        for nm in names[i]:
            p = polarization_params(nm, qp)
            rows.append((O.CHANNEL_P if p["channel"] == "P" else O.CHANNEL_S, p["num_s"], r["k0"], r["k"], p["m1"],
                         p["m2"], p["mu1"], p["mu2"], r["T"], r["Phi"], r["Phib"], r["xi"], p["A1"], p["A2"]))
    pq = (O.PolQuery * len(rows))()
    for j, t in enumerate(rows):
        x = pq[j]
        x.channel, x.num_s, x.width = int(t[0]), int(t[1]), 0
        x.k0, x.gamma, x.k, x.m1, x.m2, x.mu1, x.mu2 = t[2], 0.0, t[3], t[4], t[5], t[6], t[7]
        x.T, x.Phi, x.Phib, x.xi, x.A1, x.A2 = t[8], t[9], t[10], t[11], t[12], t[13]
    re, im = O.polarization_batch(pq, nthreads=nthreads) if rows else (np.zeros(0), np.zeros(0))
    DS, DP = np.zeros(n, complex), np.zeros(n, complex)
    j = 0
    for i, r in enumerate(q):
        pi = {}
        for nm in names[i]:
            pi[nm] = complex(re[j], im[j])
            j += 1
        if kc is None:
            c = effective_couplings(G, K, G_from_A(r["A_u"], r["m_u"], Nc), G_from_A(r["A_s"], r["m_s"], Nc))
        else:
            c = dict(zip(KC_FIELDS, kc[i]))
        qp = dict(m=dict(u=r["m_u"], s=r["m_s"]), mu=dict(u=r["mu_u"], s=r["mu_s"]), A=dict(u=r["A_u"], s=r["A_s"]))
        tp = dict(T=r["T"], Phi=r["Phi"], Phib=r["Phib"], xi=r["xi"])
        DS[i], DP[i] = propagator_by_channel(PROCESSES[int(r["process"])], "tsu"[int(r["channel"])], r["k0"], r["k"],
                                             qp, tp, c, int(r["mesons"]), pi)
    return DS, DP


# ---- vectorised form (the CPU baseline's timing path) ----------------------------
_SLOT_NAMES = ["uu_P", "uu_S", "K", "sigma_K", "ss_P", "ss_S"]


def _simple_vec(K, Pi):
    den = 1.0 - (4.0 * K) * Pi
    fin = np.isfinite(den.real) & np.isfinite(den.imag)
    den = np.where(fin & (np.abs(den) < DENOM_EPS), den + 1j * DENOM_EPS, den)
    with np.errstate(all="ignore"):
        D = (2.0 * K) / np.where(fin, den, 1.0)
    return np.where(fin, D, 0j)


def _mixed_vec(dK, Puu, Pss, K0, K8, K08, J0, J8, Jp0, Jp8):
    a, b = (4.0 / 3.0) * dK, (4.0 / 3.0) * math.sqrt(2.0) * dK
    M00 = K0 - a * (Puu + 2.0 * Pss)
    M08 = K08 + b * (Puu - Pss)
    M88 = K8 - a * (2.0 * Puu + Pss)
    det = M00 * M88 - M08 * M08
    fin = np.isfinite(det.real) & np.isfinite(det.imag)
    det = np.where(fin & (np.abs(det) < DENOM_EPS), det + 1j * DENOM_EPS, det)
    r = J0 * (M00 * Jp0 + M08 * Jp8) + J8 * (M08 * Jp0 + M88 * Jp8)
    ok = fin & np.isfinite(r.real) & np.isfinite(r.imag)
    with np.errstate(all="ignore"):
        D = (2.0 * dK) / np.where(ok, det, 1.0) * r
    return np.where(ok, D, 0j)


def propagator_by_channel_vec(q, G, K, nthreads=8, Nc=3):
    """The same (D_S, D_P) as propagator_by_channel_batch (couplings from A),
    with the Pi queries built and the propagators composed as numpy array
    operations: the CPU baseline's timed path.  Agreement with the scalar
    restatement is checked in the tests."""
    n = len(q)
    keys = q["process"].astype(np.int64) * 256 + q["channel"].astype(np.int64) * 64 + q["mesons"].astype(np.int64)
    combos = {}
    for key in np.unique(keys):
        p, ch, mk = PROCESSES[int(key // 256)], "tsu"[int(key // 64 % 4)], int(key % 64)
        simple, mP, mS = _mesons_of(p, ch, mk)
        T1, T2 = flavour_factors_for_channel(p, ch)
        q4 = parse_process(p)
        legs = {"t": ((0, 2), (1, 3)), "s": ((0, 1), (2, 3)), "u": ((0, 3), (1, 2))}[ch]
        J = current_vector(q4[legs[0][0]], q4[legs[0][1]])
        Jp = current_vector(q4[legs[1][0]], q4[legs[1][1]])
        combos[int(key)] = (simple, mP, mS, T1, T2, J, Jp, needed_pi(p, ch, mk))
    need = np.zeros((n, 6), bool)
    for key, c in combos.items():
        sel = keys == key
        for nm in c[7]:
            need[sel, _SLOT_NAMES.index(nm)] = True
    idx, slot = np.nonzero(need)
    m = len(idx)
    qa = np.zeros(m, dtype=np.dtype([(f, np.int32 if t is O.C.c_int else np.float64) for f, t in O.PolQuery._fields_]))
    u_first = np.isin(slot, (0, 1, 2, 3))            # pairs starting with u: uu and us
    second_u = np.isin(slot, (0, 1))
    r = q[idx]
    qa["channel"] = np.where(np.isin(slot, (0, 2, 4)), O.CHANNEL_P, O.CHANNEL_S)
    qa["num_s"] = np.where(slot < 2, 0, np.where(slot < 4, 1, 2))
    qa["k0"], qa["k"] = r["k0"], r["k"]
    qa["m1"] = np.where(u_first, r["m_u"], r["m_s"])
    qa["m2"] = np.where(second_u, r["m_u"], r["m_s"])
    qa["mu1"] = np.where(u_first, r["mu_u"], r["mu_s"])
    qa["mu2"] = np.where(second_u, r["mu_u"], r["mu_s"])
    qa["A1"] = np.where(u_first, r["A_u"], r["A_s"])
    qa["A2"] = np.where(second_u, r["A_u"], r["A_s"])
    qa["T"], qa["Phi"], qa["Phib"], qa["xi"] = r["T"], r["Phi"], r["Phib"], r["xi"]
    pq = (O.PolQuery * m).from_buffer(qa)
    re, im = O.polarization_batch(pq, nthreads=nthreads) if m else (np.zeros(0), np.zeros(0))
    pi = np.zeros((n, 6), complex)
    pi[idx, slot] = re + 1j * im
    Gu = -Nc / (4.0 * (math.pi * math.pi)) * (q["m_u"] * q["A_u"])
    Gs = -Nc / (4.0 * (math.pi * math.pi)) * (q["m_s"] * q["A_s"])
    c = effective_couplings(G, K, Gu, Gs)
    DS, DP = np.zeros(n, complex), np.zeros(n, complex)
    simple_k = {"pi": ("K123_plus", 0, True), "K": ("K4567_plus", 2, True), "sigma_pi": ("K123_minus", 1, False),
                "sigma_K": ("K4567_minus", 3, False)}
    for key, (simple, mP, mS, T1, T2, J, Jp, _) in combos.items():
        sel = np.nonzero(keys == key)[0]
        for mes in simple:
            kk, s, is_p = simple_k[mes]
            D = T1 * _simple_vec(c[kk][sel], pi[sel, s]) * T2
            if is_p:
                DP[sel] += D
            else:
                DS[sel] += D
        for on, sfx, s0, tgt in ((mP, "plus", 0, DP), (mS, "minus", 1, DS)):
            if on:
                tgt[sel] += _mixed_vec(c[f"det_K_{sfx}"][sel], pi[sel, s0], pi[sel, s0 + 4], c[f"K0_{sfx}"][sel],
                                       c[f"K8_{sfx}"][sel], c[f"K08_{sfx}"][sel], J[0], J[1], Jp[0], Jp[1])
    return DS, DP


# ---- meson pole masses (MesonMass.jl) ----------------------------------------------
MASS_MESONS = ["pi", "K", "sigma_pi", "sigma_K", "eta", "eta_prime", "sigma", "sigma_prime"]
FD_STEP = 6.055454452393343e-06          # cbrt(eps(Float64)): FiniteDiff's central relstep = absstep


def default_meson_mass_guess(meson, qp):
    """MesonMass.jl:183-204 (d = u)."""
    mu_, ms = qp["m"]["u"], qp["m"]["s"]
    return {"pi": mu_ + mu_, "K": mu_ + ms, "sigma_pi": 2.0 * mu_, "sigma_K": mu_ + ms, "eta": 2.0 * mu_,
            "eta_prime": 2.0 * ms, "sigma": 2.0 * mu_, "sigma_prime": 2.0 * ms}[meson]


def _pi_width(channel, pair, k0, gamma, k, qp, tp):
    """polarization_with_width for the quark pair uu / us / ss (num_s 0, 1, 0:
    MesonMass.jl passes 0 for s sbar, equivalent to 2)."""
    f1, f2, ns = {"uu": ("u", "u", 0), "us": ("u", "s", 1), "ss": ("s", "s", 0)}[pair]
    re, im = O.polarization(O.CHANNEL_P if channel == "P" else O.CHANNEL_S, k0, k, qp["m"][f1], qp["m"][f2],
                            qp["mu"][f1], qp["mu"][f2], tp["T"], tp["Phi"], tp["Phib"], tp["xi"], qp["A"][f1],
                            qp["A"][f2], ns, gamma=gamma)
    return complex(re, im)


def meson_mass_equation(meson, k0, gamma, k, qp, tp, kc):
    """MesonMass.jl:211-233: 1 - 4 K Pi(k0, gamma) for pi, K, sigma_pi, sigma_K;
    (M00 + M88) -/+ sqrt((M00 - M88)^2 + 4 M08^2) (the light / heavy
    eigenvalue of 2M) for eta / eta' (channel P) and sigma / sigma' (S),
    MesonMass.jl:34-87."""
    simple = {"pi": ("P", "uu", "K123_plus"), "K": ("P", "us", "K4567_plus"), "sigma_pi": ("S", "uu", "K123_minus"),
              "sigma_K": ("S", "us", "K4567_minus")}
    if meson in simple:
        ch, pair, key = simple[meson]
        return 1.0 - 4.0 * kc[key] * _pi_width(ch, pair, k0, gamma, k, qp, tp)
    ch = "P" if meson in ("eta", "eta_prime") else "S"
    Puu = _pi_width(ch, "uu", k0, gamma, k, qp, tp)
    Pss = _pi_width(ch, "ss", k0, gamma, k, qp, tp)
    M00, M08, M88, _ = mixing_matrix_elements(Puu, Pss, kc, ch)
    sq = cmath.sqrt((M00 - M88) * (M00 - M88) + 4.0 * (M08 * M08))
    return (M00 + M88) - sq if meson in ("eta", "sigma") else (M00 + M88) + sq


def _wnorm(d, v):
    return math.sqrt(sum((di * vi) * (di * vi) for di, vi in zip(d, v)))


def _div(a, b):
    """IEEE a / b (inf / nan instead of ZeroDivisionError), as Julia divides."""
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def _dogleg(r, d, J, delta):
    """NLsolve dogleg! (2 x 2; a singular Jacobian takes the scaled
    steepest-descent branch, as the gap path's restatement does)."""
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0]
    if det != 0.0 and math.isfinite(det):
        pi_ = [-_div(J[1][1] * r[0] - J[0][1] * r[1], det), -_div(-J[1][0] * r[0] + J[0][0] * r[1], det)]
    else:
        pi_ = [math.inf, math.inf]
    if _wnorm(d, pi_) <= delta:
        return pi_
    g = [(J[0][j] * r[0] + J[1][j] * r[1]) / (d[j] * d[j]) for j in range(2)]
    Jg = [J[i][0] * g[0] + J[i][1] * g[1] for i in range(2)]
    wg = _wnorm(d, g)
    coef = _div(-(wg * wg), Jg[0] * Jg[0] + Jg[1] * Jg[1])
    pc = [coef * g[0], coef * g[1]]
    if _wnorm(d, pc) >= delta:
        sc = _div(-delta, wg)
        return [g[0] * sc, g[1] * sc]
    pdiff = [pi_[0] - pc[0], pi_[1] - pc[1]]
    b = 2.0 * sum((d[i] * pc[i]) * (d[i] * pdiff[i]) for i in range(2))
    wa = _wnorm(d, pdiff)
    a = wa * wa
    wpc = _wnorm(d, pc)
    tau = _div(-b + math.sqrt(b * b - 4.0 * a * (wpc * wpc - delta * delta)), 2.0 * a)
    return [pc[0] + tau * pdiff[0], pc[1] + tau * pdiff[1]]


def solve_meson_mass(meson, qp, tp, kc=None, k=0.0, initial_mass=None, initial_gamma=0.0, ftol=1e-8, xtol=0.0,
                     iterations=1000, G=None, K=None):
    """MesonMass.solve_meson_mass (MesonMass.jl:241-276) with NLsolve's
    defaults: trust region with dogleg and autoscaling (factor 1), central
    finite-difference Jacobian (FiniteDiff, step max(cbrt(eps) |x|, cbrt(eps))),
    ftol 1e-8, xtol 0, 1000 iterations.  kc None: couplings from qp's A.
    Returns dict(mass, gamma, converged, residual_norm, iterations, exception)."""
    if kc is None:
        kc = effective_couplings(G, K, G_from_A(qp["A"]["u"], qp["m"]["u"]), G_from_A(qp["A"]["s"], qp["m"]["s"]))
    x = [default_meson_mass_guess(meson, qp) if initial_mass is None else float(initial_mass), float(initial_gamma)]

    def F(v):
        f = meson_mass_equation(meson, v[0], v[1], k, qp, tp, kc)
        return [f.real, f.imag]

    def jac(v):
        J = [[0.0, 0.0], [0.0, 0.0]]
        for j in range(2):
            eps = max(FD_STEP * abs(v[j]), FD_STEP)
            vp, vm = list(v), list(v)
            vp[j] = v[j] + eps
            vm[j] = v[j] - eps
            fp, fm = F(vp), F(vm)
            for i in range(2):
                J[i][j] = (fp[i] - fm[i]) / (2 * eps)
        return J

    r = F(x)
    if not all(math.isfinite(t) for t in r):
        return dict(mass=x[0], gamma=x[1], converged=False, residual_norm=math.nan, iterations=0, exception=True)
    J = jac(x)
    it = 0
    conv = max(abs(r[0]), abs(r[1])) < ftol
    stopped = any(math.isnan(t) for t in x + r)
    if not conv:
        d = [math.sqrt(J[0][j] * J[0][j] + J[1][j] * J[1][j]) for j in range(2)]
        d = [dj if dj != 0.0 else 1.0 for dj in d]
        delta = 1.0 * _wnorm(d, x)
        if delta == 0.0:
            delta = 1.0
        while not stopped and not conv and it < iterations:
            it += 1
            p = _dogleg(r, d, J, delta)
            xold = list(x)
            x = [x[0] + p[0], x[1] + p[1]]
            fn = F(x)
            rpred = [J[i][0] * p[0] + J[i][1] * p[1] + r[i] for i in range(2)]
            ssr = r[0] * r[0] + r[1] * r[1]
            rho = _div(ssr - (fn[0] * fn[0] + fn[1] * fn[1]), ssr - (rpred[0] * rpred[0] + rpred[1] * rpred[1]))
            if rho > 1e-4:
                r = fn
                J = jac(x)
                d = [max(0.1 * d[j], math.sqrt(J[0][j] * J[0][j] + J[1][j] * J[1][j])) for j in range(2)]
                xc = max(abs(x[0] - xold[0]), abs(x[1] - xold[1])) < xtol
                conv = xc or max(abs(r[0]), abs(r[1])) < ftol
            else:
                x = [x[0] - p[0], x[1] - p[1]]
                conv = False
            if rho < 0.1:
                delta = delta / 2.0
            elif rho >= 0.9:
                delta = 2.0 * _wnorm(d, p)
            elif rho >= 0.5:
                delta = max(delta, 2.0 * _wnorm(d, p))
            stopped = any(math.isnan(t) for t in x + fn)
    return dict(mass=x[0], gamma=x[1], converged=bool(conv), residual_norm=max(abs(r[0]), abs(r[1])),
                iterations=it, exception=False)
