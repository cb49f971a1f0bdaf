"""Host mirror of the reference's anisotropic one-loop polarization API on the
HIP library (libpnjl_hip.so):

    polarization_aniso(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar,
                       xi, A1_value, A2_value, num_s_quark)
        src/relaxtime/PolarizationAniso.jl:22-24
    polarization_with_width(...)   :88-91
    polarization_complex(...)      :127-134
    polarization_aniso_cached(...), reset_cache(), get_cache_stats()
        src/relaxtime/PolarizationCache.jl:203-300 (same key quantisation)
    B0(lam, k, m1, mu1, m2, mu2, T, Phi=0, Phibar=0)
        src/relaxtime/OneLoopIntegrals.jl:489-503
    B0_correction(lam, k, m1, m2, mu1, mu2, T, Phi, Phibar, xi)
        src/relaxtime/OneLoopIntegralsAniso.jl:396-413
    A(m, mu, T, Phi, Phibar, nodes_p, weights_p)      OneLoopIntegrals.jl:531
    A_aniso(m, mu, T, Phi, Phibar, xi, nodes_p, weights_p, nodes_cos, weights_cos)
        OneLoopIntegralsAniso.jl:446

plus the batched forms (`polarization_batch`, `B0_batch`, ...) and the
hipGraph-captured grid `polarization_grid`.  Everything is computed on the
GPU; there is no CPU path.  Channels are "P" / "S" (the reference's :P / :S);
anything else raises ValueError, as the reference raises ArgumentError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PnjlError

CHANNEL_P, CHANNEL_S = 1, 2
_CHANNELS = {"P": CHANNEL_P, "S": CHANNEL_S, CHANNEL_P: CHANNEL_P, CHANNEL_S: CHANNEL_S}


class PolQuery(C.Structure):
    _fields_ = [("channel", C.c_int), ("num_s", C.c_int), ("width", C.c_int), ("reserved", C.c_int),
                ("k0", C.c_double), ("gamma", C.c_double), ("k", C.c_double), ("m1", C.c_double),
                ("m2", C.c_double), ("mu1", C.c_double), ("mu2", C.c_double), ("T", C.c_double),
                ("Phi", C.c_double), ("Phib", C.c_double), ("xi", C.c_double), ("A1", C.c_double),
                ("A2", C.c_double)]


class PolPoint(C.Structure):
    _fields_ = [("T", C.c_double), ("mu", C.c_double), ("xi", C.c_double), ("m_u", C.c_double),
                ("m_s", C.c_double), ("Phi", C.c_double), ("Phib", C.c_double)]


# numpy mirror of PolQuery for building large batches without Python loops
QUERY_DTYPE = np.dtype([("channel", np.int32), ("num_s", np.int32), ("width", np.int32), ("reserved", np.int32),
                        ("k0", np.float64), ("gamma", np.float64), ("k", np.float64), ("m1", np.float64),
                        ("m2", np.float64), ("mu1", np.float64), ("mu2", np.float64), ("T", np.float64),
                        ("Phi", np.float64), ("Phib", np.float64), ("xi", np.float64), ("A1", np.float64),
                        ("A2", np.float64)])
POINT_DTYPE = np.dtype([(n, np.float64) for n, _ in PolPoint._fields_])
assert QUERY_DTYPE.itemsize == C.sizeof(PolQuery) and POINT_DTYPE.itemsize == C.sizeof(PolPoint)

_DP = C.POINTER(C.c_double)


@_lib.register_extra
def _register(L):
    vp = C.c_void_p
    L.pnjl_polarization_batch.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp]
    L.pnjl_B0_batch.argtypes = [vp, C.c_int64, _DP, _DP, _DP]
    L.pnjl_B0_correction_batch.argtypes = [vp, C.c_int64, _DP, _DP, _DP]
    L.pnjl_A_batch.argtypes = [vp, C.c_int64, _DP, C.c_int, _DP, _DP, C.c_int, _DP, _DP, _DP]
    L.pnjl_polarization_grid.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, C.c_int,
                                         vp]
    for n in ("pnjl_polarization_batch", "pnjl_B0_batch", "pnjl_B0_correction_batch", "pnjl_A_batch",
              "pnjl_polarization_grid"):
        getattr(L, n).restype = C.c_int


def _ctx(device=0):
    from .pnjl import context
    return context(device)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _channel(c):
    if isinstance(c, str):
        c = c.lstrip(":").upper()
    if c not in _CHANNELS:
        raise ValueError(f"Unsupported channel: {c}. Use :P or :S.")
    return _CHANNELS[c]


def make_queries(channel, k0, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, A1, A2, num_s, gamma=None):
    """Structured array of queries from broadcastable columns."""
    cols = dict(k0=k0, k=k, m1=m1, m2=m2, mu1=mu1, mu2=mu2, T=T, Phi=Phi, Phib=Phib, xi=xi, A1=A1, A2=A2,
                num_s=num_s)
    n = int(max(np.size(v) for v in cols.values()))
    q = np.zeros(n, dtype=QUERY_DTYPE)
    ch = np.broadcast_to(np.asarray(channel, dtype=object), (n,))
    q["channel"] = [_channel(c) for c in ch] if ch.dtype == object else ch
    for f, v in cols.items():
        q[f] = np.broadcast_to(np.asarray(v), (n,))
    if gamma is not None:
        q["width"] = 1
        q["gamma"] = np.broadcast_to(np.asarray(gamma, dtype=np.float64), (n,))
    return q


def polarization_batch(q, device=0, one_per_wave=False):
    """Pi for a structured array of queries (QUERY_DTYPE) -> (re, im) numpy.
    one_per_wave: one query per wavefront (FLOP counting with hardware
    counters; 64x slower)."""
    ctx = _ctx(device)
    q = np.ascontiguousarray(q, dtype=QUERY_DTYPE)
    bad = ~np.isin(q["channel"], (CHANNEL_P, CHANNEL_S))
    if bad.any():
        raise ValueError(f"Unsupported channel: {q['channel'][bad][0]}. Use :P or :S.")
    n = q.size
    re, im = np.zeros(n), np.zeros(n)
    flags = _lib.FLAG_ONE_PER_WAVE if one_per_wave else 0
    rc = ctx._L.pnjl_polarization_batch(ctx.h, n, q.ctypes.data, re.ctypes.data, im.ctypes.data, flags, None)
    ctx._check(rc, "pnjl_polarization_batch")
    return re, im


def polarization_aniso(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value, A2_value, num_s_quark,
                       device=0):
    re, im = polarization_batch(make_queries(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value,
                                             A2_value, num_s_quark), device)
    return float(re[0]), float(im[0])


def polarization_with_width(channel, k0, gamma, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value, A2_value,
                            num_s_quark, device=0):
    re, im = polarization_batch(make_queries(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value,
                                             A2_value, num_s_quark, gamma=gamma), device)
    return float(re[0]), float(im[0])


def polarization_complex(channel, p0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value, A2_value, num_s_quark,
                         device=0):
    p0 = complex(p0)
    return polarization_with_width(channel, p0.real, 2.0 * p0.imag, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi,
                                   A1_value, A2_value, num_s_quark, device)


def B0_batch(args, device=0):
    """args (n, 9): lam, k, m1, mu1, m2, mu2, T, Phi, Phibar -> (re, im)."""
    ctx = _ctx(device)
    a = _d(args).reshape(-1, 9)
    re, im = np.zeros(a.shape[0]), np.zeros(a.shape[0])
    ctx._check(ctx._L.pnjl_B0_batch(ctx.h, a.shape[0], a.ctypes.data_as(_DP), re.ctypes.data_as(_DP),
                                    im.ctypes.data_as(_DP)), "pnjl_B0_batch")
    return re, im


def B0(lam, k, m1, mu1, m2, mu2, T, Phi=0.0, Phibar=0.0, device=0):
    re, im = B0_batch([[lam, k, m1, mu1, m2, mu2, T, Phi, Phibar]], device)
    return float(re[0]), float(im[0])


def B0_correction_batch(args, device=0):
    """args (n, 10): lam, k, m1, m2, mu1, mu2, T, Phi, Phibar, xi -> (re, im)."""
    ctx = _ctx(device)
    a = _d(args).reshape(-1, 10)
    re, im = np.zeros(a.shape[0]), np.zeros(a.shape[0])
    ctx._check(ctx._L.pnjl_B0_correction_batch(ctx.h, a.shape[0], a.ctypes.data_as(_DP), re.ctypes.data_as(_DP),
                                               im.ctypes.data_as(_DP)), "pnjl_B0_correction_batch")
    return re, im


def B0_correction(lam, k, m1, m2, mu1, mu2, T, Phi, Phibar, xi, device=0):
    re, im = B0_correction_batch([[lam, k, m1, m2, mu1, mu2, T, Phi, Phibar, xi]], device)
    return float(re[0]), float(im[0])


def A_batch(args, nodes_p, weights_p, nodes_cos=None, weights_cos=None, device=0):
    """args (n, 5) = (m, mu, T, Phi, Phibar) -> A; with cos nodes, args (n, 6)
    with xi appended -> A_aniso."""
    ctx = _ctx(device)
    aniso = nodes_cos is not None
    a = _d(args).reshape(-1, 6 if aniso else 5)
    xp, wp = _d(nodes_p), _d(weights_p)
    xc, wc = (_d(nodes_cos), _d(weights_cos)) if aniso else (np.zeros(1), np.zeros(1))
    out = np.zeros(a.shape[0])
    ctx._check(ctx._L.pnjl_A_batch(ctx.h, a.shape[0], a.ctypes.data_as(_DP), xp.size, xp.ctypes.data_as(_DP),
                                   wp.ctypes.data_as(_DP), xc.size if aniso else 0, xc.ctypes.data_as(_DP),
                                   wc.ctypes.data_as(_DP), out.ctypes.data_as(_DP)), "pnjl_A_batch")
    return out


def A(m, mu, T, Phi, Phibar, nodes_p, weights_p, device=0):
    return float(A_batch([[m, mu, T, Phi, Phibar]], nodes_p, weights_p, device=device)[0])


def A_aniso(m, mu, T, Phi, Phibar, xi, nodes_p, weights_p, nodes_cos, weights_cos, device=0):
    return float(A_batch([[m, mu, T, Phi, Phibar, xi]], nodes_p, weights_p, nodes_cos, weights_cos, device)[0])


def polarization_grid(points, k0, k, np_nodes=64, nc_nodes=4, device=0, out=None, stream=None):
    """Pi on the (k0, k) grid for each point (structured POINT_DTYPE array or
    dict of columns T, mu, xi, m_u, m_s, Phi, Phib).  Returns (re, im) shaped
    (npts, 3 pairs [u ubar, u sbar, s sbar], 2 channels [P, S], nk0, nk).
    With torch.cuda tensors (points as a float64 (npts, 7) tensor, k0, k) the
    outputs stay in HBM."""
    ctx = _ctx(device)
    if type(k0).__module__.startswith("torch"):
        import torch
        npts, nk0, nk = points.shape[0], k0.numel(), k.numel()
        shape = (npts, 3, 2, nk0, nk)
        if out is None:
            out = (torch.empty(shape, dtype=torch.float64, device=k0.device),
                   torch.empty(shape, dtype=torch.float64, device=k0.device))
        st = stream if stream is not None else torch.cuda.current_stream(k0.device)
        rc = ctx._L.pnjl_polarization_grid(ctx.h, npts, points.data_ptr(), nk0, k0.data_ptr(), nk, k.data_ptr(),
                                           np_nodes, nc_nodes, out[0].data_ptr(), out[1].data_ptr(),
                                           _lib.FLAG_DEVICE_PTRS, C.c_void_p(st.cuda_stream))
        ctx._check(rc, "pnjl_polarization_grid")
        return out
    if isinstance(points, dict):
        n = np.size(points["T"])
        p = np.zeros(n, dtype=POINT_DTYPE)
        for f in POINT_DTYPE.names:
            p[f] = points[f]
        points = p
    pts = np.ascontiguousarray(points, dtype=POINT_DTYPE)
    k0, k = _d(k0), _d(k)
    shape = (pts.size, 3, 2, k0.size, k.size)
    re, im = np.zeros(shape), np.zeros(shape)
    rc = ctx._L.pnjl_polarization_grid(ctx.h, pts.size, pts.ctypes.data, k0.size, k0.ctypes.data, k.size,
                                       k.ctypes.data, np_nodes, nc_nodes, re.ctypes.data, im.ctypes.data, 0, None)
    ctx._check(rc, "pnjl_polarization_grid")
    return re, im


# ---------------------------------------------------------------------------
# PolarizationCache.jl semantics: memoisation keyed on the 14 arguments, each
# truncated to 40 mantissa bits (relative bucket ~9.1e-13); channel and
# num_s_quark exact.  The first call in a bucket computes and stores.
# ---------------------------------------------------------------------------
CACHE_MANTISSA_BITS = 40
_KEEP = np.uint64(0xFFF0000000000000) | (np.uint64(0x000FFFFFFFFFFFFF) &
                                         ~np.uint64((1 << (52 - CACHE_MANTISSA_BITS)) - 1))
_CACHE: dict = {}
_STATS = {"calls": 0, "hits": 0}


def _round_cache(x: float) -> float:
    """PolarizationCache.jl:49-56 (zero, NaN and Inf unchanged)."""
    x = float(x)
    if x == 0.0 or not np.isfinite(x):
        return x
    b = np.array([x]).view(np.uint64)[0] & _KEEP
    return float(np.array([b], dtype=np.uint64).view(np.float64)[0])


def polarization_aniso_cached(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value, A2_value,
                              num_s_quark, device=0):
    """PolarizationCache.jl:203-229"""
    _STATS["calls"] += 1
    code = _channel(channel)
    key = (code,) + tuple(_round_cache(v) for v in (k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value,
                                                    A2_value)) + (int(num_s_quark),)
    hit = _CACHE.get(key)
    if hit is not None:
        _STATS["hits"] += 1
        return hit
    r = polarization_aniso(code, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar, xi, A1_value, A2_value, num_s_quark,
                           device)
    _CACHE[key] = r
    return r


def reset_cache():
    """PolarizationCache.jl:256-261"""
    _CACHE.clear()
    _STATS["calls"] = _STATS["hits"] = 0


def get_cache_stats():
    """PolarizationCache.jl:286-300"""
    c, h = _STATS["calls"], _STATS["hits"]
    return {"total_calls": c, "cache_hits": h, "cache_misses": c - h, "hit_rate": (h / c) if c else 0.0,
            "cache_size": len(_CACHE)}


__all__ = ["polarization_aniso", "polarization_with_width", "polarization_complex", "polarization_batch",
           "polarization_grid", "polarization_aniso_cached", "reset_cache", "get_cache_stats", "B0", "B0_batch",
           "B0_correction", "B0_correction_batch", "A", "A_aniso", "A_batch", "make_queries", "PnjlError"]
