"""ctypes front end of the CPU oracle (libpnjl_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP path is compared
against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module.  It is a plain-C FP64 restatement of the reference's
PNJL FixedMu gap path (function-by-function citations in pnjl_gap_oracle.c) and
of its one-loop polarization path (pnjl_polar_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpnjl_oracle.so")
_lib = None

# policy codes (same values as include/pnjl_hip.h)
SEED_DEFAULT, SEED_HADRON, SEED_QUARK, SEED_MULTI, SEED_EXPLICIT = 0, 1, 2, 3, 4
SEED_CONTINUITY, SEED_PHASE_AWARE, SEED_SURROGATE, SEED_MULTIROOT3 = 5, 6, 7, 8

HBARC = 197.327


class ModelParams(C.Structure):
    _fields_ = [("hbarc", C.c_double), ("Lambda", C.c_double), ("G", C.c_double), ("K", C.c_double),
                ("m_ud0", C.c_double), ("m_s0", C.c_double), ("T0", C.c_double), ("a0", C.c_double),
                ("a1", C.c_double), ("a2", C.c_double), ("b3", C.c_double), ("rho0", C.c_double),
                ("N_color", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("xtol", C.c_double), ("ftol", C.c_double), ("residual_norm_max", C.c_double),
                ("phi_tol", C.c_double), ("max_iter", C.c_int), ("trust_region_fallback", C.c_int),
                ("auto_multiseed_fallback", C.c_int), ("surrogate_max_steps", C.c_int),
                ("chain_len", C.c_int), ("surrogate_judge", C.c_int)]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class SolveOut(C.Structure):
    _fields_ = [("x", _DP), ("omega", _DP), ("pressure", _DP), ("rho_norm", _DP), ("entropy", _DP),
                ("energy", _DP), ("masses", _DP), ("n_q", _DP), ("n_qbar", _DP), ("residual_norm", _DP),
                ("iterations", _IP), ("converged", _IP), ("flags", _IP), ("newton_steps", _IP),
                ("passes", _IP)]


def default_opts(**kw) -> SolveOpts:
    """Reference defaults: NLsolve xtol=ftol=1e-9 (ImplicitSolver.jl:112),
    residual_norm_max=1e-6, iterations=1000 (NLsolve default), fallbacks on."""
    o = SolveOpts(xtol=1e-9, ftol=1e-9, residual_norm_max=1e-6, phi_tol=1e-8, max_iter=1000,
                  trust_region_fallback=1, auto_multiseed_fallback=1, surrogate_max_steps=8,
                  chain_len=0, surrogate_judge=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def lib():
    global _lib
    if _lib is None:
        srcs = [os.path.join(_HERE, f) for f in os.listdir(_HERE) if f.endswith(".c")]
        stale = (not os.path.exists(_LIB_PATH)) or any(
            os.path.getmtime(s) > os.path.getmtime(_LIB_PATH) for s in srcs)
        if stale:
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = C.CDLL(_LIB_PATH)
        L.oracle_default_params.argtypes = [C.POINTER(ModelParams)]
        L.oracle_set_params.argtypes = [C.POINTER(ModelParams)]
        L.oracle_gl_std.argtypes = [C.c_int, _DP, _DP]
        L.oracle_gauleg.argtypes = [C.c_double, C.c_double, C.c_int, _DP, _DP]
        L.oracle_nodes.argtypes = [C.c_int, C.c_int, _DP, _DP, _DP]
        L.oracle_omega.argtypes = [_DP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.oracle_omega.restype = C.c_double
        L.oracle_omega_derivs.argtypes = [_DP, _DP, C.c_double, C.c_double, C.c_int, C.c_int, _DP, _DP, _DP]
        L.oracle_vacuum_integral.argtypes = [C.c_double, _DP]
        L.oracle_quark_distribution.argtypes = [C.c_double] * 5
        L.oracle_quark_distribution.restype = C.c_double
        L.oracle_antiquark_distribution.argtypes = [C.c_double] * 5
        L.oracle_antiquark_distribution.restype = C.c_double
        L.oracle_log_term.argtypes = [C.c_double] * 5
        L.oracle_log_term.restype = C.c_double
        L.oracle_solve_batch.argtypes = [C.c_int64, _DP, _DP, _DP, _DP, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(SolveOpts), C.POINTER(SolveOut), C.c_int]
        L.oracle_solve_batch.restype = C.c_int
        L.oracle_newton_capped.argtypes = [C.c_double, C.c_double, C.c_double, _DP, C.c_int, C.c_int,
                                           C.POINTER(SolveOpts), C.c_int, _DP, _DP, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        L.oracle_newton_capped.restype = C.c_int
        L.oracle_thermo.argtypes = [_DP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _DP]
        L.oracle_thermo_derivs.argtypes = [_DP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _DP]
        L.oracle_thermo_derivs.restype = C.c_int
        L.oracle_implicit_derivs2.argtypes = [_DP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _DP]
        L.oracle_implicit_derivs2.restype = C.c_int
        L.oracle_set_boundary.argtypes = [C.c_int, _DP, _DP, _DP, C.c_int, _DP, _DP, _DP]
        L.oracle_interp_mu_c.argtypes = [C.c_double, C.c_double]
        L.oracle_interp_mu_c.restype = C.c_double
        L.oracle_current_phase.argtypes = [C.c_double, C.c_double, C.c_double]
        L.oracle_current_phase.restype = C.c_int
        _register_polar(L)
        _lib = L
    return _lib


def _register_polar(L):
    if not hasattr(L, "oracle_polarization_aniso"):
        return
    d = C.c_double
    L.oracle_polarization_aniso.argtypes = [C.c_int, d, d, d, d, d, d, d, d, d, d, d, d, C.c_int, _DP]
    L.oracle_polarization_aniso.restype = C.c_int
    L.oracle_B0.argtypes = [d, d, d, d, d, d, d, d, d, _DP]
    L.oracle_B0_correction.argtypes = [d, d, d, d, d, d, d, d, d, d, _DP]
    L.oracle_A.argtypes = [d, d, d, d, d, _DP, _DP, C.c_int]
    L.oracle_A.restype = d
    L.oracle_A_aniso.argtypes = [d, d, d, d, d, d, _DP, _DP, C.c_int, _DP, _DP, C.c_int]
    L.oracle_A_aniso.restype = d
    L.oracle_tilde_B0.argtypes = [C.c_int, d, d, d, d, d, d, d, d, _DP]
    L.oracle_tilde_B0_correction.argtypes = [C.c_int, d, d, d, d, d, d, d, d, d, _DP]
    L.oracle_hybrid_integrate_test.argtypes = [C.c_int, C.c_int, d, d, C.c_int, _DP]
    L.oracle_hybrid_integrate_test.restype = d
    L.oracle_polarization_batch.argtypes = [C.c_int64, _DP, _DP, _DP, C.c_int]
    L.oracle_polarization_batch.restype = C.c_int


def _p(a):
    return a.ctypes.data_as(_DP)


def set_params(params: ModelParams | None = None):
    lib().oracle_set_params(C.byref(params) if params is not None else None)


def default_params() -> ModelParams:
    p = ModelParams()
    lib().oracle_default_params(C.byref(p))
    return p


def gl_std(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_gl_std(n, _p(x), _p(w))
    return x, w


def gauleg(a, b, n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_gauleg(a, b, n, _p(x), _p(w))
    return x, w


def nodes(p_num, t_num):
    n = p_num * t_num
    p = np.zeros(n)
    c = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_nodes(p_num, t_num, _p(p), _p(c), _p(w))
    return p, c, w


def omega(x, mu, T, xi=0.0, p_num=64, t_num=8):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_omega(_p(x), mu, T, xi, p_num, t_num)


def omega_derivs(x, mu_vec, T, xi=0.0, p_num=64, t_num=8):
    x = np.ascontiguousarray(x, dtype=np.float64)
    muv = np.ascontiguousarray(np.broadcast_to(np.asarray(mu_vec, dtype=np.float64), (3,)))
    om = C.c_double()
    g = np.zeros(5)
    h = np.zeros(25)
    lib().oracle_omega_derivs(_p(x), _p(muv), T, xi, p_num, t_num, C.byref(om), _p(g), _p(h))
    return om.value, g, h.reshape(5, 5)


def vacuum_integral(M):
    out = np.zeros(3)
    lib().oracle_vacuum_integral(M, _p(out))
    return out


def thermo(x, T, mu, xi=0.0, p_num=64, t_num=8):
    """-> dict(omega, pressure, rho_norm, entropy, energy, n_q[3], n_qbar[3], phys)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(12)
    lib().oracle_thermo(_p(x), T, mu, xi, p_num, t_num, _p(out))
    return dict(omega=out[0], pressure=out[1], rho_norm=out[2], entropy=out[3], energy=out[4],
                n_q=out[5:8].copy(), n_qbar=out[8:11].copy(), phys=bool(out[11]))


THERMO_DERIV_FIELDS = ["pressure", "energy", "rho", "rho_norm", "entropy", "dP_dT", "dP_dmu", "dEpsilon_dT",
                       "dEpsilon_dmu", "dn_dT", "dn_dmu", "dP_depsilon_n", "dP_dn_epsilon", "masses", "dM_dT",
                       "dM_dmu", "v_n_sq", "dmuB_dT_sigma", "dx_dT", "dx_dmu", "ds_dT", "ds_dmu"]
_VEC = {"masses": 3, "dM_dT": 3, "dM_dmu": 3, "dx_dT": 5, "dx_dmu": 5}


def unpack_thermo_derivs(out):
    """36 doubles (one point) or (n, 36) -> dict of the named fields."""
    out = np.asarray(out)
    d, k = {}, 0
    for f in THERMO_DERIV_FIELDS:
        w = _VEC.get(f, 1)
        d[f] = out[..., k] if w == 1 else out[..., k:k + w]
        k += w
    return d


def thermo_derivs(x, T, mu, xi=0.0, p_num=64, t_num=8):
    """Implicit thermodynamic derivatives at the root x (oracle_thermo_derivs)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(36)
    lib().oracle_thermo_derivs(_p(x), T, mu, xi, p_num, t_num, _p(out))
    return unpack_thermo_derivs(out)


# field layout of oracle_implicit_derivs2 / pnjl_implicit_derivatives2_batch
IMPLICIT2_LAYOUT = [("x", 5), ("dx_dT", 5), ("dx_dmu", 5), ("d2x_dT2", 5), ("d2x_dTdmu", 5), ("d2x_dmu2", 5),
                    ("masses", 3), ("dM_dT", 3), ("dM_dmu", 3), ("d2M_dT2", 3), ("d2M_dTdmu", 3), ("d2M_dmu2", 3)]


def unpack_implicit2(out):
    """dict of the named blocks of a 48-value row (or an (n, 48) array)"""
    out = np.asarray(out)
    d, o = {}, 0
    for name, w in IMPLICIT2_LAYOUT:
        d[name] = out[..., o:o + w]
        o += w
    return d


def implicit_derivs2(x, T, mu, xi=0.0, p_num=64, t_num=8):
    """Second-order implicit derivatives at the root x (oracle_implicit_derivs2):
    x_ab = -H^-1 [(D_wb H) x_a + D_wb F_a] by complex step on the closed forms."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(48)
    lib().oracle_implicit_derivs2(_p(x), T, mu, xi, p_num, t_num, _p(out))
    return unpack_implicit2(out)


def alloc_out(n):
    arrs = dict(x=np.zeros((n, 5)), omega=np.zeros(n), pressure=np.zeros(n), rho_norm=np.zeros(n),
                entropy=np.zeros(n), energy=np.zeros(n), masses=np.zeros((n, 3)), n_q=np.zeros((n, 3)),
                n_qbar=np.zeros((n, 3)), residual_norm=np.zeros(n), iterations=np.zeros(n, np.int32),
                converged=np.zeros(n, np.int32), flags=np.zeros(n, np.int32),
                newton_steps=np.zeros(n, np.int32), passes=np.zeros(n, np.int32))
    s = SolveOut()
    for k, a in arrs.items():
        setattr(s, k, a.ctypes.data_as(_IP if a.dtype == np.int32 else _DP))
    return arrs, s


def solve_batch(T, mu, xi, policy=SEED_DEFAULT, seed=None, p_num=64, t_num=8, opts=None, nthreads=1):
    """Batch FixedMu solve on the CPU; same argument meaning as
    pnjl_solve_fixedmu_batch.  Returns a dict of numpy arrays."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    n = T.size
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
    xi = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, dtype=np.float64), (n,)))
    sp = None
    if seed is not None:
        seed = np.ascontiguousarray(seed, dtype=np.float64).reshape(n, 5)
        sp = _p(seed)
    arrs, s = alloc_out(n)
    o = opts if opts is not None else default_opts()
    rc = lib().oracle_solve_batch(n, _p(T), _p(mu), _p(xi), sp, policy, p_num, t_num, C.byref(o),
                                  C.byref(s), nthreads)
    if rc != 0:
        raise ValueError(f"oracle_solve_batch failed with status {rc}")
    return arrs


def newton_capped(T, mu, xi, x0, cap, p_num=64, t_num=8, opts=None):
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    xo = np.zeros(5)
    rn = C.c_double()
    it = C.c_int()
    fc = C.c_int()
    o = opts if opts is not None else default_opts()
    exc = lib().oracle_newton_capped(T, mu, xi, _p(x0), p_num, t_num, C.byref(o), cap, _p(xo), C.byref(rn),
                                     C.byref(it), C.byref(fc))
    return dict(x=xo, residual_norm=rn.value, iterations=it.value, f_converged=bool(fc.value),
                exception=bool(exc))


def set_boundary(rows, cep_rows):
    """rows: (k,3) array of (xi, T_MeV, mu_MeV); cep_rows: (m,3) of (xi, T_CEP, mu_CEP)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 3)
    cep = np.ascontiguousarray(cep_rows, dtype=np.float64).reshape(-1, 3)
    a, b, c = (np.ascontiguousarray(rows[:, i]) for i in range(3))
    d, e, f = (np.ascontiguousarray(cep[:, i]) for i in range(3))
    lib().oracle_set_boundary(len(a), _p(a), _p(b), _p(c), len(d), _p(d), _p(e), _p(f))


def set_judge_mask(mask, lo=None, hi=None):
    """Ambiguity mask (uint8 [nT, nmu, nxi] over the box [lo, hi]) of the
    learned-seed policy's 3-seed Omega check; None clears it."""
    L = lib()
    L.oracle_set_judge_mask.argtypes = [C.c_int, C.c_int, C.c_int, _DP, _DP, C.POINTER(C.c_uint8)]
    if mask is None:
        L.oracle_set_judge_mask(0, 0, 0, None, None, None)
        return
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    rc = L.oracle_set_judge_mask(m.shape[0], m.shape[1], m.shape[2], _p(lo), _p(hi),
                                 m.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != 0:
        raise ValueError("oracle_set_judge_mask: bad geometry")


# ---------------------------------------------------------------------------
# anisotropic one-loop polarization (pnjl_polar_oracle.c)
# ---------------------------------------------------------------------------
CHANNEL_P, CHANNEL_S = 1, 2


class PolQuery(C.Structure):
    _fields_ = [("channel", C.c_int), ("num_s", C.c_int), ("width", C.c_int), ("reserved", C.c_int),
                ("k0", C.c_double), ("gamma", C.c_double), ("k", C.c_double), ("m1", C.c_double),
                ("m2", C.c_double), ("mu1", C.c_double), ("mu2", C.c_double), ("T", C.c_double),
                ("Phi", C.c_double), ("Phib", C.c_double), ("xi", C.c_double), ("A1", C.c_double),
                ("A2", C.c_double)]


POL_FIELDS = [f for f, _ in PolQuery._fields_ if f != "reserved"]


def pol_queries(**cols):
    """Array of PolQuery from equal-length columns (names as in PolQuery)."""
    n = max(np.size(v) for v in cols.values())
    q = (PolQuery * n)()
    for f in POL_FIELDS:
        v = np.broadcast_to(np.asarray(cols.get(f, 0)), (n,))
        for i in range(n):
            setattr(q[i], f, int(v[i]) if f in ("channel", "num_s", "width") else float(v[i]))
    return q


def _polar_lib():
    L = lib()
    if not getattr(L, "_polar_ready", False):
        d2 = [C.c_double] * 2
        L.oracle_B0.argtypes = [C.c_double] * 9 + [_DP, _DP]
        L.oracle_tilde_B0.argtypes = [C.c_int] + [C.c_double] * 8 + [_DP, _DP]
        L.oracle_B0_correction.argtypes = [C.c_double] * 10 + [_DP, _DP]
        L.oracle_tilde_B0_correction.argtypes = [C.c_int] + [C.c_double] * 9 + [_DP, _DP]
        L.oracle_real_tool.argtypes = [C.c_double, C.c_double]
        L.oracle_real_tool.restype = C.c_double
        L.oracle_imag_tool.argtypes = [C.c_double, C.c_double]
        L.oracle_imag_tool.restype = C.c_double
        L.oracle_A_const.argtypes = [C.c_double]
        L.oracle_A_const.restype = C.c_double
        L.oracle_A.argtypes = [C.c_double] * 5 + [C.c_int, _DP, _DP]
        L.oracle_A.restype = C.c_double
        L.oracle_A_correction.argtypes = [C.c_double] * 6 + [C.c_int, _DP, _DP]
        L.oracle_A_correction.restype = C.c_double
        L.oracle_A_aniso.argtypes = [C.c_double] * 6 + [C.c_int, _DP, _DP, C.c_int, _DP, _DP]
        L.oracle_A_aniso.restype = C.c_double
        L.oracle_polarization.argtypes = [C.c_int, C.c_int] + [C.c_double] * 13 + [C.c_int, _DP, _DP]
        L.oracle_polarization.restype = C.c_int
        L.oracle_polarization_batch.argtypes = [C.c_int64, C.POINTER(PolQuery), _DP, _DP, C.c_int]
        L.oracle_polarization_batch.restype = C.c_int
        L.oracle_polar_tables.argtypes = [_DP]
        del d2
        L._polar_ready = True
    return L


def _pair(fn, *args):
    re, im = C.c_double(), C.c_double()
    fn(*args, C.byref(re), C.byref(im))
    return re.value, im.value


def B0(lam, k, m1, mu1, m2, mu2, T, Phi=0.0, Phib=0.0):
    return _pair(_polar_lib().oracle_B0, lam, k, m1, mu1, m2, mu2, T, Phi, Phib)


def tilde_B0(plus, lam, k, m, mp, mu, T, Phi=0.0, Phib=0.0):
    return _pair(_polar_lib().oracle_tilde_B0, 0 if plus else 1, lam, k, m, mp, mu, T, Phi, Phib)


def B0_correction(lam, k, m1, m2, mu1, mu2, T, Phi, Phib, xi):
    return _pair(_polar_lib().oracle_B0_correction, lam, k, m1, m2, mu1, mu2, T, Phi, Phib, xi)


def tilde_B0_correction(quark, lam, k, m, mp, mu, T, Phi, Phib, xi):
    """One delta-B~ term (OneLoopIntegralsAniso.jl:380-390)."""
    return _pair(_polar_lib().oracle_tilde_B0_correction, 0 if quark else 1, lam, k, m, mp, mu, T, Phi, Phib, xi)


def real_tool(cx, dc):
    """real_integral_tool (OneLoopIntegralsAniso.jl:26-33)."""
    return _polar_lib().oracle_real_tool(cx, dc)


def imag_tool(cx, dc):
    """imag_integral_tool (OneLoopIntegralsAniso.jl:36-38)."""
    return _polar_lib().oracle_imag_tool(cx, dc)


def A_const(m):
    return _polar_lib().oracle_A_const(m)


def A(m, mu, T, Phi, Phib, nodes, weights):
    x, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (nodes, weights))
    return _polar_lib().oracle_A(m, mu, T, Phi, Phib, x.size, _p(x), _p(w))


def A_correction(m, mu, T, Phi, Phib, xi, nodes, weights):
    x, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (nodes, weights))
    return _polar_lib().oracle_A_correction(m, mu, T, Phi, Phib, xi, x.size, _p(x), _p(w))


def A_aniso(m, mu, T, Phi, Phib, xi, nodes_p, weights_p, nodes_c, weights_c):
    xp, wp, xc, wc = (np.ascontiguousarray(a, dtype=np.float64) for a in (nodes_p, weights_p, nodes_c, weights_c))
    return _polar_lib().oracle_A_aniso(m, mu, T, Phi, Phib, xi, xp.size, _p(xp), _p(wp), xc.size, _p(xc), _p(wc))


def polarization(channel, k0, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, A1, A2, num_s, gamma=None):
    """polarization_aniso (gamma None) or polarization_with_width."""
    re, im = C.c_double(), C.c_double()
    rc = _polar_lib().oracle_polarization(channel, 0 if gamma is None else 1, k0, 0.0 if gamma is None else gamma, k,
                                          m1, m2, mu1, mu2, T, Phi, Phib, xi, A1, A2, num_s, C.byref(re),
                                          C.byref(im))
    if rc != 0:
        raise ValueError("Unsupported channel")
    return re.value, im.value


def polarization_batch(q, nthreads=8):
    n = len(q)
    re, im = np.zeros(n), np.zeros(n)
    rc = _polar_lib().oracle_polarization_batch(n, q, _p(re), _p(im), nthreads)
    if rc != 0:
        raise ValueError("Unsupported channel")
    return re, im


def polar_tables():
    t = np.zeros(16 * 2 + 32 * 2 + (16 * 2 + 32 * 2) * 2 + 66)
    _polar_lib().oracle_polar_tables(_p(t))
    return t
