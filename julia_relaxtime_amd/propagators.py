"""Host mirror of the reference's consumers of the polarization Pi on the HIP
library (libpnjl_hip.so):

    calculate_G_from_A(A_f, m_f; Nc)            EffectiveCouplings.jl:56-60
    calculate_effective_couplings(G, K, G_u, G_s)   :240-278 (with det_K+-)
    mixing_matrix_elements(Pi_uu, Pi_ss, K_coeffs, channel)   :137-160
    meson_propagator_simple(meson, K_coeffs, Pi)    MesonPropagator.jl:268-291
    meson_propagator_mixed(...)                     MesonPropagator.jl:365-419
    calculate_propagator_by_channel(process, channel, k0, k_norm,
                                    quark_params, thermo_params, K_coeffs)
                                                    TotalPropagator.jl:894-963
    total_propagator_simple(...)  :423-470,  total_propagator_mixed(...)  :514-560

plus `propagator_by_channel_batch`, which takes a structured array of queries
(PROP_QUERY_DTYPE) and evaluates every Pi it needs and the propagators in two
kernel launches.  Everything is computed on the GPU; there is no CPU path and
no host-side memo of Pi (the reference's polarization_aniso_cached Dict is
replaced by evaluating each Pi in the batch).

Arguments follow the reference: K_coeffs is a dict with its NamedTuple's
keys; quark_params is dict(m=dict(u=, s=), mu=dict(u=, s=), A=dict(u=, s=))
(d = u); thermo_params is dict(T=, Phi=, Phibar=, xi=).  Two deviations, both
from the form of the batched call: meson_propagator_mixed takes
(K_coeffs, Pi_uu, Pi_ss, meson_channel) in place of a prebuilt (det_K, M)
matrix -- the only caller, total_propagator_mixed, builds M from exactly
those -- and total_propagator_simple returns its sum as D_S + D_P (the
same terms; the reference adds them before scaling by the flavour factors).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

# SCATTERING_MESON_MAP order (src/Constants_PNJL.jl:168-400) = PNJL_PROC_* ids
PROCESSES = ["ud_to_ud", "uu_to_uu", "us_to_us", "ss_to_ss", "ubardbar_to_ubardbar", "ubarubar_to_ubarubar",
             "ubarsbar_to_ubarsbar", "sbarsbar_to_sbarsbar", "udbar_to_udbar", "dubar_to_dubar", "uubar_to_uubar",
             "uubar_to_ddbar", "usbar_to_usbar", "subar_to_subar", "uubar_to_ssbar", "ssbar_to_uubar",
             "ssbar_to_ssbar"]
PROCESS_ID = {p: i for i, p in enumerate(PROCESSES)}
CHANNEL_ID = {"t": 0, "s": 1, "u": 2}
MESON_BITS = {"pi": 0x1, "K": 0x2, "sigma_pi": 0x4, "sigma_K": 0x8, "mixed_P": 0x10, "mixed_S": 0x20}
KC_FIELDS = ["K0_plus", "K0_minus", "K123_plus", "K123_minus", "K4567_plus", "K4567_minus", "K8_plus", "K8_minus",
             "K08_plus", "K08_minus", "det_K_plus", "det_K_minus"]
# Pi slots of pnjl_meson_propagators_batch / pnjl_polarization_grid
PI_SLOTS = ["uu_P", "uu_S", "us_P", "us_S", "ss_P", "ss_S"]
MESON_PROP_FIELDS = 36

PROP_QUERY_DTYPE = np.dtype([("process", np.int32), ("channel", np.int32), ("mesons", np.int32),
                             ("reserved", np.int32), ("k0", np.float64), ("k", np.float64), ("T", np.float64),
                             ("Phi", np.float64), ("Phib", np.float64), ("xi", np.float64), ("m_u", np.float64),
                             ("m_s", np.float64), ("mu_u", np.float64), ("mu_s", np.float64), ("A_u", np.float64),
                             ("A_s", np.float64)])
assert PROP_QUERY_DTYPE.itemsize == 4 * 4 + 12 * 8


@_lib.register_extra
def _register(L):
    vp = C.c_void_p
    L.pnjl_effective_couplings_batch.argtypes = [vp, C.c_int64, vp, C.c_int, C.c_double, C.c_double, vp, vp]
    L.pnjl_meson_propagators_batch.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.pnjl_propagator_by_channel_batch.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp]
    for n in ("pnjl_effective_couplings_batch", "pnjl_meson_propagators_batch", "pnjl_propagator_by_channel_batch"):
        getattr(L, n).restype = C.c_int


def _ctx(device=0):
    from .pnjl import context
    return context(device)


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _params(device):
    return _ctx(device).params


# ---- couplings --------------------------------------------------------------
def couplings_from_A(A_u, A_s, m_u, m_s, G=None, K=None, device=0):
    """G_u, G_s (n, 2) and the couplings (n, 12) from A and the masses in one
    launch; G, K default to the model's G_fm2, K_fm5."""
    ctx = _ctx(device)
    cols = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (A_u, A_s, m_u, m_s)))
    n = cols[0].size
    rows = _d(np.stack([c.ravel() for c in cols], axis=1))
    Gf, kc = np.zeros((n, 2)), np.zeros((n, 12))
    p = ctx.params
    rc = _lib.load().pnjl_effective_couplings_batch(ctx.h, n, _ptr(rows), 0, p.G if G is None else float(G),
                                                     p.K if K is None else float(K), _ptr(Gf), _ptr(kc))
    ctx._check(rc, "pnjl_effective_couplings_batch")
    return Gf, kc


def calculate_G_from_A(A_f, m_f, Nc=None, device=0):
    """calculate_G_from_A(A_f, m_f; Nc)  EffectiveCouplings.jl:56-60 (scalar or array)."""
    if Nc is not None and int(Nc) != _ctx(device).params.N_color:
        raise NotImplementedError("calculate_G_from_A: Nc is the context's N_color")
    A, m = np.broadcast_arrays(np.asarray(A_f, dtype=np.float64), np.asarray(m_f, dtype=np.float64))
    Gf, _ = couplings_from_A(A.ravel(), A.ravel(), m.ravel(), m.ravel(), device=device)
    out = Gf[:, 0].reshape(A.shape)
    return float(out) if out.ndim == 0 else out


def _kc_dict(kc):
    kc = np.asarray(kc)
    if kc.ndim == 1:
        return {k: float(kc[j]) for j, k in enumerate(KC_FIELDS)}
    return {k: kc[:, j].copy() for j, k in enumerate(KC_FIELDS)}


def _kc_rows(K_coeffs, n):
    if isinstance(K_coeffs, dict):
        return _d(np.stack([np.broadcast_to(np.asarray(K_coeffs[k], dtype=np.float64), (n,)) for k in KC_FIELDS],
                           axis=1))
    return _d(np.broadcast_to(np.asarray(K_coeffs, dtype=np.float64), (n, 12)))


def calculate_effective_couplings(G, K, G_u, G_s, device=0):
    """calculate_effective_couplings  EffectiveCouplings.jl:240-278: dict of
    the 12 couplings (K0_plus ... det_K_minus), scalars or arrays."""
    ctx = _ctx(device)
    gu, gs = np.broadcast_arrays(np.asarray(G_u, dtype=np.float64), np.asarray(G_s, dtype=np.float64))
    n = gu.size
    rows = _d(np.stack([gu.ravel(), gs.ravel()], axis=1))
    kc = np.zeros((n, 12))
    rc = _lib.load().pnjl_effective_couplings_batch(ctx.h, n, _ptr(rows), 1, float(G), float(K), None, _ptr(kc))
    ctx._check(rc, "pnjl_effective_couplings_batch")
    return _kc_dict(kc[0] if gu.ndim == 0 else kc)


# ---- meson propagators from given Pi ------------------------------------------
def meson_propagators_batch(K_coeffs, pi, device=0):
    """pnjl_meson_propagators_batch: K_coeffs (dict or (n, 12)), pi (n, 6)
    complex in PI_SLOTS order -> (n, 18) complex (the 36-field layout)."""
    ctx = _ctx(device)
    pi = np.asarray(pi, dtype=complex)
    pi = pi.reshape(-1, 6)
    n = pi.shape[0]
    kc = _kc_rows(K_coeffs, n)
    pr = _d(np.stack([pi.real, pi.imag], axis=2).reshape(n, 12))
    out = np.zeros((n, MESON_PROP_FIELDS))
    rc = _lib.load().pnjl_meson_propagators_batch(ctx.h, n, _ptr(kc), _ptr(pr), _ptr(out))
    ctx._check(rc, "pnjl_meson_propagators_batch")
    return out[:, 0::2] + 1j * out[:, 1::2]


def _one(K_coeffs, slots, device):
    pi = np.zeros(6, complex)
    for k, v in slots.items():
        pi[PI_SLOTS.index(k)] = complex(v)
    return meson_propagators_batch(K_coeffs, pi[None, :], device)[0]


def meson_propagator_simple(meson, K_coeffs, Pi, device=0):
    """MesonPropagator.jl:268-291 (meson :pi, :K, :sigma_pi, :sigma_K)."""
    slot = {"pi": ("uu_P", 0), "K": ("us_P", 1), "sigma_pi": ("uu_S", 2), "sigma_K": ("us_S", 3)}
    m = str(meson).lstrip(":")
    if m not in slot:
        raise ValueError(f"Unknown meson type: {meson}. Use :pi, :K, :sigma_pi, or :sigma_K")
    name, j = slot[m]
    return complex(_one(K_coeffs, {name: Pi}, device)[j])


def _channel(ch):
    c = str(ch).lstrip(":").upper()
    if c not in ("P", "S"):
        raise ValueError(f"Unknown channel: {ch}. Use :P (pseudoscalar) or :S (scalar)")
    return c


def mixing_matrix_elements(Pi_uu, Pi_ss, K_coeffs, channel, device=0):
    """EffectiveCouplings.jl:137-160 -> dict(M00, M08, M88, det_K)."""
    c = _channel(channel)
    o = _one(K_coeffs, {f"uu_{c}": Pi_uu, f"ss_{c}": Pi_ss}, device)
    b = 4 if c == "P" else 7
    kc = K_coeffs if isinstance(K_coeffs, dict) else _kc_dict(K_coeffs)
    return dict(M00=complex(o[b]), M08=complex(o[b + 1]), M88=complex(o[b + 2]),
                det_K=kc["det_K_plus" if c == "P" else "det_K_minus"])


def calculate_coupling_matrix(Pi_uu, Pi_ss, K_coeffs, channel, device=0):
    """MesonPropagator.jl:207-213: the symmetric 2x2 complex M."""
    e = mixing_matrix_elements(Pi_uu, Pi_ss, K_coeffs, channel, device)
    return np.array([[e["M00"], e["M08"]], [e["M08"], e["M88"]]])


def _flavour(q):
    q = str(q).lstrip(":")
    if q not in ("u", "d", "s", "ubar", "dbar", "sbar"):
        raise ValueError(f"Unknown flavor: {q}")
    return q[0]


def meson_propagator_mixed(K_coeffs, Pi_uu, Pi_ss, meson_channel, q1, q2, q3, q4, channel, device=0):
    """MesonPropagator.jl:365-419 with M built from (Pi_uu, Pi_ss) on the GPU."""
    c = _channel(meson_channel)
    legs = {"t": ((q1, q3), (q2, q4)), "s": ((q1, q2), (q3, q4)), "u": ((q1, q4), (q2, q3))}
    ch = str(channel).lstrip(":")
    if ch not in legs:
        raise ValueError(f"Unknown channel: {channel}. Use :t, :s, or :u")
    (a, b), (x, y) = legs[ch]
    fa, fb, fx, fy = (_flavour(v) for v in (a, b, x, y))
    o = _one(K_coeffs, {f"uu_{c}": Pi_uu, f"ss_{c}": Pi_ss}, device)
    if fa != fb or fx != fy:
        return 0j                           # no current between different flavours
    j = 2 * (fa == "s") + (fx == "s")       # (J, J') = (a|b, a|b)
    return complex(o[(10 if c == "P" else 14) + j])


# ---- D_S / D_P ----------------------------------------------------------------
def make_queries(process, channel, k0, k, T, Phi, Phib, xi, m_u, m_s, mu_u, mu_s, A_u, A_s, mesons=0):
    """Structured PROP_QUERY_DTYPE array from broadcastable columns; process
    and channel by name ('ud_to_ud', 't') or id."""
    cols = dict(k0=k0, k=k, T=T, Phi=Phi, Phib=Phib, xi=xi, m_u=m_u, m_s=m_s, mu_u=mu_u, mu_s=mu_s, A_u=A_u,
                A_s=A_s)
    n = int(max([np.size(v) for v in cols.values()] + [np.size(process), np.size(channel), np.size(mesons)]))
    q = np.zeros(n, dtype=PROP_QUERY_DTYPE)
    pr = np.broadcast_to(np.asarray(process, dtype=object), (n,))
    q["process"] = [PROCESS_ID[str(p).lstrip(":")] if not isinstance(p, (int, np.integer)) else int(p) for p in pr]
    chs = np.broadcast_to(np.asarray(channel, dtype=object), (n,))
    q["channel"] = [CHANNEL_ID[str(c).lstrip(":")] if not isinstance(c, (int, np.integer)) else int(c) for c in chs]
    q["mesons"] = np.broadcast_to(np.asarray(mesons), (n,))
    for f, v in cols.items():
        q[f] = np.broadcast_to(np.asarray(v, dtype=np.float64), (n,))
    return q


def process_channels(process):
    """The Mandelstam channels SCATTERING_MESON_MAP lists for a process: t and u
    for quark-quark (and antiquark-antiquark), t and s for quark-antiquark."""
    left = process.split("_to_")[0]
    return ("t", "s") if left.count("bar") == 1 else ("t", "u")


def propagator_by_channel_batch(q, K_coeffs=None, device=0, one_per_wave=False):
    """(D_S, D_P) complex arrays for a PROP_QUERY_DTYPE array.  K_coeffs None:
    couplings from each query's A (the reference callers' K_coeffs).
    one_per_wave: one (Pi pair, query) per wavefront, for FLOP counting."""
    ctx = _ctx(device)
    q = np.ascontiguousarray(q, dtype=PROP_QUERY_DTYPE)
    n = q.size
    kc = None if K_coeffs is None else _kc_rows(K_coeffs, n)
    out = np.zeros((n, 4))
    flags = _lib.FLAG_ONE_PER_WAVE if one_per_wave else 0
    rc = _lib.load().pnjl_propagator_by_channel_batch(ctx.h, n, _ptr(q), None if kc is None else _ptr(kc),
                                                       _ptr(out), flags, None)
    ctx._check(rc, "pnjl_propagator_by_channel_batch")
    return out[:, 0] + 1j * out[:, 1], out[:, 2] + 1j * out[:, 3]


def propagator_by_channel_device(q, out, K_coeffs=None, stream=None, device=0):
    """Device-resident form: q a torch uint8 tensor holding PROP_QUERY_DTYPE
    rows in HBM, out a float64 tensor (n, 4) = (D_S re, im, D_P re, im),
    K_coeffs None or a float64 (n, 12) tensor; launches on `stream` (a
    torch.cuda.Stream, default the current one) and returns without a host
    copy (PNJL_FLAG_DEVICE_PTRS | PNJL_FLAG_ASYNC)."""
    import torch
    ctx = _ctx(device)
    n = q.numel() // PROP_QUERY_DTYPE.itemsize
    if out.dtype != torch.float64 or out.numel() != 4 * n or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 tensor of n x 4")
    st = stream if stream is not None else torch.cuda.current_stream(q.device)
    flags = _lib.FLAG_DEVICE_PTRS | _lib.FLAG_ASYNC
    rc = _lib.load().pnjl_propagator_by_channel_batch(ctx.h, n, C.c_void_p(q.data_ptr()),
                                                       None if K_coeffs is None else C.c_void_p(K_coeffs.data_ptr()),
                                                       C.c_void_p(out.data_ptr()), flags, C.c_void_p(st.cuda_stream))
    ctx._check(rc, "pnjl_propagator_by_channel_batch")
    return out


def _query(process, channel, k0, k_norm, qp, tp, mesons):
    return make_queries(process, channel, k0, k_norm, tp["T"], tp["Phi"], tp.get("Phibar", tp.get("Phib")), tp["xi"],
                        qp["m"]["u"], qp["m"]["s"], qp["mu"]["u"], qp["mu"]["s"], qp["A"]["u"], qp["A"]["s"], mesons)


def calculate_propagator_by_channel(process, channel, k0, k_norm, quark_params, thermo_params, K_coeffs=None,
                                    device=0):
    """TotalPropagator.jl:894-963 -> (D_S, D_P)."""
    p = str(process).lstrip(":")
    if p not in PROCESS_ID:
        raise KeyError(process)
    DS, DP = propagator_by_channel_batch(_query(p, channel, k0, k_norm, quark_params, thermo_params, 0), K_coeffs,
                                         device)
    return complex(DS[0]), complex(DP[0])


def total_propagator_simple(process, channel, meson_list, k0, k_norm, quark_params, thermo_params, K_coeffs=None,
                            device=0):
    """TotalPropagator.jl:423-470: T1 * sum(D_meson) * T2 over meson_list."""
    mask = 0
    for m in meson_list:
        m = str(m).lstrip(":")
        if m not in ("pi", "K", "sigma_pi", "sigma_K"):
            raise ValueError(f"Unknown meson type: {m}")
        mask |= MESON_BITS[m]
    if not mask:
        return 0j
    DS, DP = propagator_by_channel_batch(_query(process, channel, k0, k_norm, quark_params, thermo_params, mask),
                                         K_coeffs, device)
    return complex(DS[0] + DP[0])


def total_propagator_mixed(process, channel, meson_channel, k0, k_norm, quark_params, thermo_params, K_coeffs=None,
                           device=0):
    """TotalPropagator.jl:514-560: the eta/eta' (P) or sigma/sigma' (S) propagator."""
    c = _channel(meson_channel)
    DS, DP = propagator_by_channel_batch(
        _query(process, channel, k0, k_norm, quark_params, thermo_params, MESON_BITS[f"mixed_{c}"]), K_coeffs,
        device)
    return complex(DP[0] if c == "P" else DS[0])


__all__ = ["PROCESSES", "PROP_QUERY_DTYPE", "KC_FIELDS", "PI_SLOTS", "couplings_from_A", "calculate_G_from_A",
           "calculate_effective_couplings", "meson_propagators_batch", "meson_propagator_simple",
           "mixing_matrix_elements", "calculate_coupling_matrix", "meson_propagator_mixed", "make_queries",
           "propagator_by_channel_batch", "propagator_by_channel_device", "process_channels",
           "calculate_propagator_by_channel", "total_propagator_simple",
           "total_propagator_mixed"]


# ---- meson pole masses (MesonMass.jl / MottTransition.jl) ---------------------------------
MASS_MESONS = ["pi", "K", "sigma_pi", "sigma_K", "eta", "eta_prime", "sigma", "sigma_prime"]   # PNJL_MESON_ID_*
MASS_QUERY_DTYPE = np.dtype([("meson", np.int32), ("max_iter", np.int32), ("reserved", np.int32, 2),
                             ("k", np.float64), ("mass0", np.float64), ("gamma0", np.float64), ("ftol", np.float64),
                             ("xtol", np.float64), ("T", np.float64), ("Phi", np.float64), ("Phib", np.float64),
                             ("xi", np.float64), ("m_u", np.float64), ("m_s", np.float64), ("mu_u", np.float64),
                             ("mu_s", np.float64), ("A_u", np.float64), ("A_s", np.float64)])
assert MASS_QUERY_DTYPE.itemsize == 16 + 15 * 8


@_lib.register_extra
def _register_mass(L):
    vp = C.c_void_p
    L.pnjl_meson_mass_batch.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp]
    L.pnjl_meson_mass_batch.restype = C.c_int


def make_mass_queries(meson, T, Phi, Phib, xi, m_u, m_s, mu_u, mu_s, A_u, A_s, k=0.0, mass0=np.nan, gamma0=0.0,
                      ftol=1e-8, xtol=0.0, max_iter=1000):
    """MASS_QUERY_DTYPE array from broadcastable columns (meson by name or id)."""
    cols = dict(T=T, Phi=Phi, Phib=Phib, xi=xi, m_u=m_u, m_s=m_s, mu_u=mu_u, mu_s=mu_s, A_u=A_u, A_s=A_s, k=k,
                mass0=mass0, gamma0=gamma0, ftol=ftol, xtol=xtol)
    n = int(max([np.size(v) for v in cols.values()] + [np.size(meson), np.size(max_iter)]))
    q = np.zeros(n, dtype=MASS_QUERY_DTYPE)
    ms = np.broadcast_to(np.asarray(meson, dtype=object), (n,))
    q["meson"] = [MASS_MESONS.index(str(m).lstrip(":")) if not isinstance(m, (int, np.integer)) else int(m)
                  for m in ms]
    q["max_iter"] = np.broadcast_to(np.asarray(max_iter), (n,))
    for f, v in cols.items():
        q[f] = np.broadcast_to(np.asarray(v, dtype=np.float64), (n,))
    return q


def meson_mass_batch(q, K_coeffs=None, device=0, one_per_wave=False):
    """Pole masses for a MASS_QUERY_DTYPE array -> dict of arrays (mass, gamma,
    residual_norm, converged, iterations, exception).  one_per_wave: one
    query per wavefront, for FLOP counting."""
    ctx = _ctx(device)
    q = np.ascontiguousarray(q, dtype=MASS_QUERY_DTYPE)
    n = q.size
    kc = None if K_coeffs is None else _kc_rows(K_coeffs, n)
    out = np.zeros((n, 6))
    flags = _lib.FLAG_ONE_PER_WAVE if one_per_wave else 0
    rc = _lib.load().pnjl_meson_mass_batch(ctx.h, n, _ptr(q), None if kc is None else _ptr(kc), _ptr(out), flags,
                                           None)
    ctx._check(rc, "pnjl_meson_mass_batch")
    return dict(mass=out[:, 0].copy(), gamma=out[:, 1].copy(), residual_norm=out[:, 2].copy(),
                converged=out[:, 3] == 1.0, iterations=out[:, 4].astype(int), exception=out[:, 5] == 1.0)


def default_meson_mass_guess(meson, quark_params):
    """MesonMass.jl:183-204."""
    m = quark_params["m"]
    md = m.get("d", m["u"])
    g = {"pi": m["u"] + md, "K": m["u"] + m["s"], "sigma_pi": 2.0 * m["u"], "sigma_K": m["u"] + m["s"],
         "eta": 2.0 * m["u"], "eta_prime": 2.0 * m["s"], "sigma": 2.0 * m["u"], "sigma_prime": 2.0 * m["s"]}
    name = str(meson).lstrip(":")
    if name not in g:
        raise ValueError(f"Unknown meson type: {meson}")
    return g[name]


def solve_meson_mass(meson, quark_params, thermo_params, K_coeffs=None, k_norm=0.0, initial_mass=None,
                     initial_gamma=0.0, ftol=1e-8, xtol=0.0, iterations=1000, device=0):
    """MesonMass.jl:241-276 -> dict(mass, gamma, converged, residual_norm,
    iterations).  quark_params must carry A (as TotalPropagator's do); a
    non-finite residual at the initial point raises, as NLsolve does."""
    name = str(meson).lstrip(":")
    if name not in MASS_MESONS:
        raise ValueError(f"Unknown meson type: {meson}")
    qp, tp = quark_params, thermo_params
    q = make_mass_queries(name, tp["T"], tp["Phi"], tp.get("Phibar", tp.get("Phib")), tp.get("xi", 0.0),
                          qp["m"]["u"], qp["m"]["s"], qp["mu"]["u"], qp["mu"]["s"], qp["A"]["u"], qp["A"]["s"],
                          k=k_norm, mass0=np.nan if initial_mass is None else initial_mass, gamma0=initial_gamma,
                          ftol=ftol, xtol=xtol, max_iter=iterations)
    r = meson_mass_batch(q, K_coeffs, device)
    if r["exception"][0]:
        raise FloatingPointError("solve_meson_mass: non-finite residual at the initial point")
    return dict(mass=float(r["mass"][0]), gamma=float(r["gamma"][0]), converged=bool(r["converged"][0]),
                residual_norm=float(r["residual_norm"][0]), iterations=int(r["iterations"][0]))


def mott_threshold_mass(meson, quark_params):
    """MottTransition.jl: the lowest quark-pair threshold of a meson."""
    m = quark_params["m"]
    md = m.get("d", m["u"])
    name = str(meson).lstrip(":")
    thr = {"pi": m["u"] + md, "K": m["u"] + m["s"], "sigma_pi": m["u"] + md, "sigma_K": m["u"] + m["s"]}
    if name in thr:
        return thr[name]
    if name in ("eta", "eta_prime", "sigma", "sigma_prime"):
        return min(2.0 * m["u"], 2.0 * m["s"])
    raise ValueError(f"Unknown meson type: {meson}")


def mott_gap(meson, meson_mass, quark_params):
    """MottTransition.jl: M_meson - threshold."""
    return meson_mass - mott_threshold_mass(meson, quark_params)


def is_mott_point(meson, meson_mass, quark_params, atol=1e-6):
    return abs(mott_gap(meson, meson_mass, quark_params)) <= atol


__all__ += ["MASS_MESONS", "MASS_QUERY_DTYPE", "make_mass_queries", "meson_mass_batch", "default_meson_mass_guess",
            "solve_meson_mass", "mott_threshold_mass", "mott_gap", "is_mott_point"]
