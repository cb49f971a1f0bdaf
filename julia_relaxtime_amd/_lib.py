"""ctypes binding of libpnjl_hip.so (the C ABI in include/pnjl_hip.h).

The library is built in-tree by julia_relaxtime_amd.build (hipcc, gfx950).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpnjl_hip.so")

# status codes
PNJL_OK, PNJL_ERR_ARG, PNJL_ERR_HIP, PNJL_ERR_NO_DEVICE, PNJL_ERR_CHANNEL, PNJL_ERR_STATE = range(6)
FLAG_DEVICE_PTRS, FLAG_ASYNC, FLAG_ONE_PER_WAVE = 0x1, 0x2, 0x4
# points per call of pnjl_solve_fixedmu_batch_noalloc
NOALLOC_MAX_POINTS = 8192

# seed policies
SEED_DEFAULT, SEED_HADRON, SEED_QUARK, SEED_MULTI, SEED_EXPLICIT = 0, 1, 2, 3, 4
SEED_CONTINUITY, SEED_PHASE_AWARE, SEED_SURROGATE, SEED_MULTIROOT3 = 5, 6, 7, 8
SEED_BRANCH_HADRON, SEED_BRANCH_QUARK = 9, 10

# per-point flags
PT_TR_USED, PT_MULTISEED, PT_EXCEPTION, PT_NO_ROOT = 0x01, 0x02, 0x04, 0x08
PT_SURR_FALLBACK, PT_SURR_OMEGA, PT_FIRST_ORDER, PT_CHAIN_RESET = 0x10, 0x20, 0x40, 0x80
PT_SURR_MASKED = 0x100
PT_BRANCH_END, PT_SKIPPED = 0x200, 0x400

# pnjl_implicit_derivatives2_batch: per-point output blocks (name, width)
IMPLICIT2_LAYOUT = [("x", 5), ("dx_dT", 5), ("dx_dmu", 5), ("d2x_dT2", 5), ("d2x_dTdmu", 5), ("d2x_dmu2", 5),
                    ("masses", 3), ("dM_dT", 3), ("dM_dmu", 3), ("d2M_dT2", 3), ("d2M_dTdmu", 3), ("d2M_dmu2", 3)]
IMPLICIT2_FIELDS = sum(w for _, w in IMPLICIT2_LAYOUT)
# pnjl_thermo_derivatives_batch: per-point output columns (name, width)
THERMO_DERIV_FIELDS = 36
THERMO_DERIV_LAYOUT = [("pressure", 1), ("energy", 1), ("rho", 1), ("rho_norm", 1), ("entropy", 1), ("dP_dT", 1),
                       ("dP_dmu", 1), ("dEpsilon_dT", 1), ("dEpsilon_dmu", 1), ("dn_dT", 1), ("dn_dmu", 1),
                       ("dP_depsilon_n", 1), ("dP_dn_epsilon", 1), ("masses", 3), ("dM_dT", 3), ("dM_dmu", 3),
                       ("v_n_sq", 1), ("dmuB_dT_sigma", 1), ("dx_dT", 5), ("dx_dmu", 5), ("ds_dT", 1), ("ds_dmu", 1)]


def unpack_thermo_derivs(out):
    """(n, 36) -> dict of named columns ((n,) or (n, w))."""
    d, k = {}, 0
    for name, w in THERMO_DERIV_LAYOUT:
        d[name] = out[:, k] if w == 1 else out[:, k:k + w]
        k += w
    return d
JUDGE_REFERENCE, JUDGE_OMEGA = 0, 1


class PnjlError(RuntimeError):
    pass


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


class LaunchCfg(C.Structure):
    _fields_ = [("lanes_per_point", C.c_int), ("chunk", C.c_int), ("waves_per_cu", C.c_int),
                ("tail_lanes_per_point", C.c_int)]

    def as_dict(self):
        return dict(lanes_per_point=self.lanes_per_point, chunk=self.chunk, waves_per_cu=self.waves_per_cu,
                    tail_lanes_per_point=self.tail_lanes_per_point)


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class SolveOut(C.Structure):
    _fields_ = [("x", _DP), ("omega", _DP), ("pressure", _DP), ("rho_norm", _DP), ("entropy", _DP),
                ("energy", _DP), ("masses", _DP), ("n_q", _DP), ("n_qbar", _DP), ("residual_norm", _DP),
                ("iterations", _IP), ("converged", _IP), ("flags", _IP), ("newton_steps", _IP),
                ("passes", _IP)]


OUT_FIELDS = [("x", 5, np.float64), ("omega", 1, np.float64), ("pressure", 1, np.float64),
              ("rho_norm", 1, np.float64), ("entropy", 1, np.float64), ("energy", 1, np.float64),
              ("masses", 3, np.float64), ("n_q", 3, np.float64), ("n_qbar", 3, np.float64),
              ("residual_norm", 1, np.float64), ("iterations", 1, np.int32), ("converged", 1, np.int32),
              ("flags", 1, np.int32), ("newton_steps", 1, np.int32),
              ("passes", 1, np.int32)]

_lib = None


def load():
    """Loads libpnjl_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PnjlError(f"{LIB_PATH} is missing: build it with julia_relaxtime_amd.build.build() "
                        "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    # torch wheels ship their own HIP/HSA runtime.  Loading torch first lets the
    # library bind to that same runtime (same sonames), so a process that also
    # uses torch for device memory holds one HIP runtime, not two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.pnjl_default_params.argtypes = [C.POINTER(ModelParams)]
    L.pnjl_default_params.restype = None
    L.pnjl_default_solve_opts.argtypes = [C.POINTER(SolveOpts)]
    L.pnjl_default_solve_opts.restype = None
    L.pnjl_ctx_create.argtypes = [C.POINTER(ModelParams), C.c_int, C.POINTER(vp)]
    L.pnjl_ctx_destroy.argtypes = [vp]
    L.pnjl_ctx_destroy.restype = None
    L.pnjl_last_error.argtypes = [vp]
    L.pnjl_last_error.restype = C.c_char_p
    L.pnjl_last_kernel_ms.argtypes = [vp]
    L.pnjl_last_kernel_ms.restype = C.c_double
    L.pnjl_last_scan_ms.argtypes = [vp, _DP, C.c_int]
    L.pnjl_last_scan_ms.restype = C.c_int
    L.pnjl_gauleg.argtypes = [C.c_double, C.c_double, C.c_int, _DP, _DP]
    L.pnjl_nodes.argtypes = [vp, C.c_int, C.c_int, _DP, _DP, _DP]
    L.pnjl_set_phase_boundary.argtypes = [vp, C.c_int, _DP, _DP, _DP, C.c_int, _DP, _DP, _DP]
    L.pnjl_set_judge_mask.argtypes = [vp, C.c_int, C.c_int, C.c_int, _DP, _DP, C.POINTER(C.c_uint8)]
    L.pnjl_set_launch_cfg.argtypes = [vp, C.POINTER(LaunchCfg)]
    L.pnjl_get_launch_cfg.argtypes = [vp, C.POINTER(LaunchCfg)]
    L.pnjl_solve_fixedmu_batch.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(SolveOpts), C.POINTER(SolveOut), C.c_int, vp]
    L.pnjl_solve_fixedmu_batch_noalloc.argtypes = L.pnjl_solve_fixedmu_batch.argtypes
    L.pnjl_gap_residual_batch.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp,
                                          C.c_int, vp]
    L.pnjl_thermo_derivatives_batch.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_int, C.c_int, vp, C.c_int, vp]
    L.pnjl_implicit_derivatives2_batch.argtypes = L.pnjl_thermo_derivatives_batch.argtypes
    L.pnjl_autotune.argtypes = [vp, C.c_int64, _DP, _DP, _DP, _DP, C.c_int, C.c_int, C.c_int,
                                C.POINTER(SolveOpts), C.POINTER(LaunchCfg), C.POINTER(C.c_double)]
    for name in ("pnjl_ctx_create", "pnjl_gauleg", "pnjl_nodes", "pnjl_set_phase_boundary", "pnjl_set_judge_mask",
                 "pnjl_set_launch_cfg", "pnjl_get_launch_cfg", "pnjl_solve_fixedmu_batch",
                 "pnjl_solve_fixedmu_batch_noalloc",
                 "pnjl_gap_residual_batch", "pnjl_thermo_derivatives_batch", "pnjl_implicit_derivatives2_batch",
                 "pnjl_autotune"):
        getattr(L, name).restype = C.c_int
    _register_extra(L)
    _lib = L
    return L


_extra_registrars = []


def register_extra(fn):
    """Hook for the other modules of the package to declare their symbols."""
    _extra_registrars.append(fn)
    if _lib is not None:
        fn(_lib)
    return fn


def _register_extra(L):
    for fn in _extra_registrars:
        fn(L)


def default_params() -> ModelParams:
    p = ModelParams()
    load().pnjl_default_params(C.byref(p))
    return p


def default_opts(**kw) -> SolveOpts:
    """Reference defaults (ImplicitSolver.jl:112, :211-222; NLsolve iterations=1000)."""
    o = SolveOpts()
    load().pnjl_default_solve_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown solver option {k!r}")
        setattr(o, k, v)
    return o


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class Context:
    """One pnjl_ctx (device buffers, node tables, launch geometry) on one GPU."""

    def __init__(self, device: int = 0, params: ModelParams | None = None):
        self._L = load()
        h = C.c_void_p()
        rc = self._L.pnjl_ctx_create(C.byref(params) if params is not None else None, device, C.byref(h))
        if rc != PNJL_OK:
            raise PnjlError(f"pnjl_ctx_create failed with status {rc} (no gfx950 device visible?)")
        self.h = h
        self.device = device
        self.params = params if params is not None else default_params()

    def close(self):
        if getattr(self, "h", None):
            self._L.pnjl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != PNJL_OK:
            msg = self._L.pnjl_last_error(self.h)
            raise PnjlError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}")

    # ------------------------------------------------------------------
    def nodes(self, p_num, t_num):
        n = p_num * t_num
        p, c, w = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self._L.pnjl_nodes(self.h, p_num, t_num, _dptr(p), _dptr(c), _dptr(w)), "pnjl_nodes")
        return p, c, w

    def set_phase_boundary(self, rows, cep_rows):
        rows = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1, 3)
        cep = np.ascontiguousarray(cep_rows, dtype=np.float64).reshape(-1, 3)
        cols = [np.ascontiguousarray(rows[:, i]) for i in range(3)]
        ccols = [np.ascontiguousarray(cep[:, i]) for i in range(3)]
        self._check(self._L.pnjl_set_phase_boundary(self.h, len(cols[0]), *[_dptr(c) for c in cols], len(ccols[0]),
                                                    *[_dptr(c) for c in ccols]), "pnjl_set_phase_boundary")

    def set_judge_mask(self, mask, lo, hi):
        """Ambiguity mask (uint8 [nT, nmu, nxi] over the box [lo, hi]) of the
        learned seed's 3-seed Omega check; None clears it."""
        if mask is None:
            self._check(self._L.pnjl_set_judge_mask(self.h, 0, 0, 0, None, None, None), "pnjl_set_judge_mask")
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        self._check(self._L.pnjl_set_judge_mask(self.h, m.shape[0], m.shape[1], m.shape[2], _dptr(lo), _dptr(hi),
                                                m.ctypes.data_as(C.POINTER(C.c_uint8))), "pnjl_set_judge_mask")

    def set_launch_cfg(self, lanes_per_point=4, chunk=4, waves_per_cu=8, tail_lanes_per_point=0):
        cfg = LaunchCfg(lanes_per_point, chunk, waves_per_cu, tail_lanes_per_point)
        self._check(self._L.pnjl_set_launch_cfg(self.h, C.byref(cfg)), "pnjl_set_launch_cfg")

    def launch_cfg(self):
        cfg = LaunchCfg()
        self._check(self._L.pnjl_get_launch_cfg(self.h, C.byref(cfg)), "pnjl_get_launch_cfg")
        return cfg.as_dict()

    def last_kernel_ms(self):
        return self._L.pnjl_last_kernel_ms(self.h)

    def last_scan_ms(self):
        """Durations (ms) of the persistent-kernel launches of the last solve."""
        ms = np.zeros(8)
        k = self._L.pnjl_last_scan_ms(self.h, _dptr(ms), ms.size)
        if k < 0:
            self._check(k, "pnjl_last_scan_ms")
        return ms[:min(k, ms.size)].copy()

    # ------------------------------------------------------------------
    def solve_batch(self, T, mu, xi, policy=SEED_DEFAULT, seed=None, p_num=64, t_num=8, opts=None, out=None,
                    stream=None, sync=True):
        """Batch FixedMu solve.  numpy inputs -> host staging, returns numpy
        arrays; torch.cuda tensors -> device pointers (inputs already in HBM),
        returns a dict of torch tensors (allocated here unless `out` is given)."""
        o = opts if opts is not None else default_opts()
        if _is_torch(T):
            return self._solve_torch(T, mu, xi, policy, seed, p_num, t_num, o, out, stream, sync)
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        n = T.size
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
        xi = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, dtype=np.float64), (n,)))
        sp = None
        if seed is not None:
            seed = np.ascontiguousarray(seed, dtype=np.float64).reshape(n, 5)
            sp = seed.ctypes.data
        arrs = {k: np.zeros((n, w) if w > 1 else n, dtype=dt) for k, w, dt in OUT_FIELDS}
        s = SolveOut(**{k: a.ctypes.data_as(_IP if a.dtype == np.int32 else _DP) for k, a in arrs.items()})
        rc = self._L.pnjl_solve_fixedmu_batch(self.h, n, T.ctypes.data, mu.ctypes.data, xi.ctypes.data, sp, policy,
                                              p_num, t_num, C.byref(o), C.byref(s), 0, None)
        self._check(rc, "pnjl_solve_fixedmu_batch")
        return arrs

    def solve_batch_noalloc(self, T, mu, xi, policy=SEED_DEFAULT, seed=None, p_num=64, t_num=8, opts=None, out=None,
                            stream=None, sync=True):
        """solve_batch through pnjl_solve_fixedmu_batch_noalloc: torch.cuda
        tensors only, at most NOALLOC_MAX_POINTS points; the library allocates
        nothing (pass `out` to allocate nothing here either)."""
        if not _is_torch(T):
            raise PnjlError("solve_batch_noalloc takes device tensors only")
        o = opts if opts is not None else default_opts()
        return self._solve_torch(T, mu, xi, policy, seed, p_num, t_num, o, out, stream, sync,
                                 entry="pnjl_solve_fixedmu_batch_noalloc")

    def _solve_torch(self, T, mu, xi, policy, seed, p_num, t_num, o, out, stream, sync,
                     entry="pnjl_solve_fixedmu_batch"):
        import torch
        n = T.numel()
        dev = T.device
        if out is None:
            out = {k: torch.empty((n, w) if w > 1 else (n,), dtype=torch.float64 if dt == np.float64 else torch.int32,
                                  device=dev) for k, w, dt in OUT_FIELDS}
        for t in [T, mu, xi] + ([seed] if seed is not None else []) + list(out.values()):
            if not t.is_contiguous() or t.device != dev:
                raise PnjlError("device inputs must be contiguous tensors on the context's device")
        s = SolveOut(**{k: C.cast(C.c_void_p(v.data_ptr()), _IP if v.dtype == torch.int32 else _DP)
                        for k, v in out.items()})
        st = stream if stream is not None else torch.cuda.current_stream(dev)
        flags = FLAG_DEVICE_PTRS | (0 if sync else FLAG_ASYNC)
        rc = getattr(self._L, entry)(self.h, n, T.data_ptr(), mu.data_ptr(), xi.data_ptr(),
                                     seed.data_ptr() if seed is not None else None, policy, p_num, t_num,
                                     C.byref(o), C.byref(s), flags, C.c_void_p(st.cuda_stream))
        self._check(rc, entry)
        return out

    def residual_batch(self, x, T, mu, xi, p_num=64, t_num=8):
        """F = grad_x P (n,5), J = dF/dx (n,5,5), Omega (n) at given states."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 5)
        n = x.shape[0]
        T = np.ascontiguousarray(np.broadcast_to(np.asarray(T, dtype=np.float64), (n,)))
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
        xi = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, dtype=np.float64), (n,)))
        F, J, om = np.zeros((n, 5)), np.zeros((n, 5, 5)), np.zeros(n)
        rc = self._L.pnjl_gap_residual_batch(self.h, n, x.ctypes.data, T.ctypes.data, mu.ctypes.data,
                                             xi.ctypes.data, p_num, t_num, F.ctypes.data, J.ctypes.data,
                                             om.ctypes.data, 0, None)
        self._check(rc, "pnjl_gap_residual_batch")
        return F, J, om

    def thermo_derivatives_batch(self, x, T, mu, xi, p_num=64, t_num=8):
        """Implicit thermodynamic derivatives at the roots x (n, 5) -> dict of
        the PNJL_THERMO_DERIV_FIELDS columns (THERMO_DERIV_LAYOUT)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 5)
        n = x.shape[0]
        T, mu, xi = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (T, mu, xi))
        out = np.zeros((n, THERMO_DERIV_FIELDS))
        rc = self._L.pnjl_thermo_derivatives_batch(self.h, n, T.ctypes.data, mu.ctypes.data, xi.ctypes.data,
                                                   x.ctypes.data, p_num, t_num, out.ctypes.data, 0, None)
        self._check(rc, "pnjl_thermo_derivatives_batch")
        return unpack_thermo_derivs(out)

    def implicit_derivatives2_batch(self, x, T, mu, xi, p_num=64, t_num=8):
        """First and second implicit derivatives of the roots x (n, 5) in (T, mu)
        and of the masses -> dict of the IMPLICIT2_LAYOUT blocks."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 5)
        n = x.shape[0]
        T, mu, xi = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (T, mu, xi))
        out = np.zeros((n, IMPLICIT2_FIELDS))
        rc = self._L.pnjl_implicit_derivatives2_batch(self.h, n, T.ctypes.data, mu.ctypes.data, xi.ctypes.data,
                                                      x.ctypes.data, p_num, t_num, out.ctypes.data, 0, None)
        self._check(rc, "pnjl_implicit_derivatives2_batch")
        d, o = {}, 0
        for name, w in IMPLICIT2_LAYOUT:
            d[name] = out[:, o:o + w].copy()
            o += w
        return d

    def autotune(self, T, mu, xi, policy=SEED_DEFAULT, seed=None, p_num=64, t_num=8, opts=None, sample=None):
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        n = T.size if sample is None else min(sample, T.size)
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (T.size,)))
        xi = np.ascontiguousarray(np.broadcast_to(np.asarray(xi, dtype=np.float64), (T.size,)))
        sp = None
        if seed is not None:
            seed = np.ascontiguousarray(seed, dtype=np.float64).reshape(-1, 5)
            sp = _dptr(seed)
        o = opts if opts is not None else default_opts()
        best = LaunchCfg()
        ns = C.c_double()
        self._check(self._L.pnjl_autotune(self.h, n, _dptr(T), _dptr(mu), _dptr(xi), sp, policy, p_num, t_num,
                                          C.byref(o), C.byref(best), C.byref(ns)), "pnjl_autotune")
        return best.as_dict(), ns.value


def gauleg(a, b, n):
    """GL nodes/weights on [a,b] from the library's host routine."""
    x, w = np.zeros(n), np.zeros(n)
    rc = load().pnjl_gauleg(a, b, n, _dptr(x), _dptr(w))
    if rc != PNJL_OK:
        raise PnjlError(f"pnjl_gauleg failed with status {rc}")
    return x, w
