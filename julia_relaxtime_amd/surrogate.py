"""The learned seed: an MLP (T, mu, xi) -> (phi_u, phi_s, Phi, Phibar) trained
on the GPU (libpnjl_hip.so: BF16-MFMA forward/backward, FP32 Adam) from roots
the plain solver produced, and `SurrogateSeed`, the seed strategy that uses
it.  The reference seeds stay the fallback and the judge: a learned seed whose
Newton run does not converge in `surrogate_max_steps` steps hands the point
to the plain DefaultSeed policy (PNJL_SEED_SURROGATE); inside an ambiguity
mask learned on the training set (several roots, first-order region) the
root is checked against the six MultiSeed candidates and the minimum-Omega
root is kept; MULTIROOT3 keeps the minimum-Omega root among the learned,
hadron and quark seeds everywhere.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PnjlError

HBARC = 197.327
NPARAM = 34052
# the build's training domain: T 20-300 MeV, mu 0-450 MeV, xi -0.4..0.4
DOMAIN_LO = np.array([20.0 / HBARC, 0.0, -0.4])
DOMAIN_HI = np.array([300.0 / HBARC, 450.0 / HBARC, 0.4])

_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)


@_lib.register_extra
def _register(L):
    vp = C.c_void_p
    L.pnjl_surrogate_param_count.restype = C.c_int
    L.pnjl_surrogate_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
    L.pnjl_surrogate_create.restype = C.c_int
    L.pnjl_surrogate_destroy.argtypes = [vp]
    L.pnjl_surrogate_destroy.restype = None
    L.pnjl_surrogate_last_error.argtypes = [vp]
    L.pnjl_surrogate_last_error.restype = C.c_char_p
    L.pnjl_surrogate_last_ms.argtypes = [vp]
    L.pnjl_surrogate_last_ms.restype = C.c_double
    L.pnjl_surrogate_set_scaling.argtypes = [vp, _DP, _DP, _DP, _DP]
    L.pnjl_surrogate_get_params.argtypes = [vp, _FP]
    L.pnjl_surrogate_set_params.argtypes = [vp, _FP]
    L.pnjl_surrogate_train.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                       _FP, C.c_int, vp]
    L.pnjl_surrogate_loss_grad.argtypes = [vp, C.c_int64, _DP, _DP, _DP, _DP, _FP, _FP]
    L.pnjl_surrogate_infer.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_int, vp]
    for n in ("pnjl_surrogate_set_scaling", "pnjl_surrogate_get_params", "pnjl_surrogate_set_params",
              "pnjl_surrogate_train", "pnjl_surrogate_loss_grad", "pnjl_surrogate_infer"):
        getattr(L, n).restype = C.c_int


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Surrogate:
    """GPU-resident MLP seed model (one pnjl_surrogate)."""

    def __init__(self, device: int = 0, init_seed: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        rc = self._L.pnjl_surrogate_create(device, init_seed, C.byref(h))
        if rc != 0:
            raise PnjlError(f"pnjl_surrogate_create failed with status {rc}")
        self.h = h
        self.device = device
        self.in_lo, self.in_hi = DOMAIN_LO.copy(), DOMAIN_HI.copy()
        self.out_mean, self.out_std = np.zeros(4), np.ones(4)
        self._push_scaling()

    def close(self):
        if getattr(self, "h", None):
            self._L.pnjl_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            m = self._L.pnjl_surrogate_last_error(self.h)
            raise PnjlError(f"{what} failed with status {rc}: {m.decode() if m else ''}")

    def _push_scaling(self):
        a = [_d(v) for v in (self.in_lo, self.in_hi, self.out_mean, self.out_std)]
        self._check(self._L.pnjl_surrogate_set_scaling(self.h, *[v.ctypes.data_as(_DP) for v in a]), "set_scaling")

    def set_scaling(self, in_lo=None, in_hi=None, out_mean=None, out_std=None):
        if in_lo is not None:
            self.in_lo = np.asarray(in_lo, float)
        if in_hi is not None:
            self.in_hi = np.asarray(in_hi, float)
        if out_mean is not None:
            self.out_mean = np.asarray(out_mean, float)
        if out_std is not None:
            self.out_std = np.asarray(out_std, float)
        self._push_scaling()

    def fit_output_scaling(self, roots):
        y = np.stack([roots[:, 0], roots[:, 2], roots[:, 3], roots[:, 4]], 1)
        self.set_scaling(out_mean=y.mean(0), out_std=np.maximum(y.std(0), 1e-6))

    @property
    def params(self):
        p = np.zeros(NPARAM, dtype=np.float32)
        self._check(self._L.pnjl_surrogate_get_params(self.h, p.ctypes.data_as(_FP)), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters")
        self._check(self._L.pnjl_surrogate_set_params(self.h, p.ctypes.data_as(_FP)), "set_params")

    def train(self, T, mu, xi, roots, steps=20000, batch=4096, lr=1e-3, seed=4):
        """Adam on the GPU; returns the per-step loss (numpy) and stores the
        training kernel time in `last_ms`."""
        T, mu, xi, roots = _d(T), _d(mu), _d(xi), _d(roots).reshape(-1, 5)
        n = T.size
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        rc = self._L.pnjl_surrogate_train(self.h, n, T.ctypes.data, _d(mu).ctypes.data, xi.ctypes.data,
                                          roots.ctypes.data, steps, batch, lr, seed, loss.ctypes.data_as(_FP), 0,
                                          None)
        self._check(rc, "pnjl_surrogate_train")
        self.last_ms = self._L.pnjl_surrogate_last_ms(self.h)
        return loss[:steps]

    def loss_grad(self, T, mu, xi, roots):
        T, mu, xi, roots = _d(T), _d(mu), _d(xi), _d(roots).reshape(-1, 5)
        loss = C.c_float()
        g = np.zeros(NPARAM, dtype=np.float32)
        rc = self._L.pnjl_surrogate_loss_grad(self.h, T.size, T.ctypes.data_as(_DP), mu.ctypes.data_as(_DP),
                                              xi.ctypes.data_as(_DP), roots.ctypes.data_as(_DP), C.byref(loss),
                                              g.ctypes.data_as(_FP))
        self._check(rc, "pnjl_surrogate_loss_grad")
        return loss.value, g

    def infer(self, T, mu, xi, out=None, stream=None, sync=True):
        """Seeds (n,5) = (phi_u, phi_u, phi_s, Phi, Phibar).  numpy in -> numpy
        out; torch.cuda tensors in -> a torch tensor in HBM (sync=False: only
        enqueued on the stream, e.g. ahead of a solve on the same stream)."""
        if type(T).__module__.startswith("torch"):
            import torch
            n = T.numel()
            if out is None:
                out = torch.empty((n, 5), dtype=torch.float64, device=T.device)
            st = stream if stream is not None else torch.cuda.current_stream(T.device)
            rc = self._L.pnjl_surrogate_infer(self.h, n, T.data_ptr(), mu.data_ptr(), xi.data_ptr(), out.data_ptr(),
                                              _lib.FLAG_DEVICE_PTRS | (0 if sync else _lib.FLAG_ASYNC),
                                              C.c_void_p(st.cuda_stream))
            self._check(rc, "pnjl_surrogate_infer")
            return out
        T, mu, xi = _d(T).ravel(), _d(mu).ravel(), _d(xi).ravel()
        seeds = np.zeros((T.size, 5))
        rc = self._L.pnjl_surrogate_infer(self.h, T.size, T.ctypes.data, mu.ctypes.data, xi.ctypes.data,
                                          seeds.ctypes.data, 0, None)
        self._check(rc, "pnjl_surrogate_infer")
        return seeds

    def save(self, path):
        extra = {} if getattr(self, "judge_mask", None) is None else {"judge_mask": self.judge_mask}
        np.savez(path, params=self.params, in_lo=self.in_lo, in_hi=self.in_hi, out_mean=self.out_mean,
                 out_std=self.out_std, **extra)

    def load(self, path):
        z = np.load(path)
        self.set_scaling(z["in_lo"], z["in_hi"], z["out_mean"], z["out_std"])
        self.params = z["params"]
        self.judge_mask = z["judge_mask"] if "judge_mask" in z.files else None
        return self

    def build_judge_mask(self, ctx, T, mu, xi, plain_roots=None, p_num=64, t_num=4, opts=None, bins=(128, 128, 16),
                         dilate=1, multi_root=False, refine=8, refine_seed=6, neighbourhood="full",
                         multi_max_iter=12):
        """Ambiguity mask of the learned seed's judge, learned on probe points:
        a bin is set where the learned-seed root, the plain (DefaultSeed) root
        and the MultiSeed (argmin-Omega) root are not all the same root, or a
        solve failed; with `multi_root`, also where the learned, HADRON and
        QUARK seeds reach distinct roots.  The MultiSeed probe runs each
        candidate as at most `multi_max_iter` Newton iterations, without the
        trust-region fallback: it only has to find the competing roots, and
        its candidates that fail would otherwise run the whole budget twice
        over.  The probes run on the context's GPU; only the verdicts come
        back.

        Two passes: the given probes (the training set, optionally with more
        random points), then -- when `refine` > 0 -- `refine` further random
        probes in every bin within two bins of a set one (default_rng
        (refine_seed)), where ambiguity is likely.  The set bins are dilated by
        `dilate` bins over their 26 neighbours (`neighbourhood` "full") or 6
        faces ("face").  Stored in self.judge_mask and installed in `ctx`;
        returns (mask, ambiguous probes of the first pass)."""
        T, mu, xi = _d(T), _d(mu), _d(xi)
        nb = np.asarray(bins)
        amb = self._ambiguous(ctx, T, mu, xi, plain_roots, p_num, t_num, opts, multi_root, multi_max_iter)
        mask = np.zeros(tuple(nb), dtype=np.uint8)
        b = self._bins(T[amb], mu[amb], xi[amb], nb)
        mask[b[:, 0], b[:, 1], b[:, 2]] = 1
        if refine > 0 and mask.any():
            region = _dilate(mask, 2, full=True)
            cells = np.argwhere(region)
            rng = np.random.default_rng(refine_seed)
            pick = np.repeat(np.arange(cells.shape[0]), refine)
            u = (cells[pick] + rng.random((pick.size, 3))) / nb
            P = self.in_lo + u * (self.in_hi - self.in_lo)
            amb2 = self._ambiguous(ctx, P[:, 0], P[:, 1], P[:, 2], None, p_num, t_num, opts, multi_root,
                                   multi_max_iter)
            b2 = self._bins(P[amb2, 0], P[amb2, 1], P[amb2, 2], nb)
            mask[b2[:, 0], b2[:, 1], b2[:, 2]] = 1
        mask = _dilate(mask, dilate, full=(neighbourhood == "full"))
        self.judge_mask = mask
        ctx.set_judge_mask(mask, self.in_lo, self.in_hi)
        return mask, amb

    def _bins(self, T, mu, xi, nb):
        u = (np.stack([T, mu, xi], 1) - self.in_lo) / (self.in_hi - self.in_lo)
        return np.clip((u * nb).astype(int), 0, nb - 1)

    def _ambiguous(self, ctx, T, mu, xi, plain_roots, p_num, t_num, opts, multi_root, multi_max_iter=12):
        """Probes at which the learned-seed, DefaultSeed and MultiSeed solves do
        not all reach the same root (DualBranchScan's sameness test and Omega
        to 1e-10), or one of them failed: a numpy bool per probe.  The probes
        and their solves stay in HBM; the verdict is taken on the host from
        plain copies of the four fields it reads (no torch compute kernels,
        whose code a fresh process first has to load from disk: ~0.3 s)."""
        import torch
        from ._lib import PT_FIRST_ORDER, SEED_DEFAULT, SEED_MULTI, SEED_MULTIROOT3, SEED_SURROGATE
        o = opts if opts is not None else _lib.default_opts()
        mo = type(o).from_buffer_copy(o)
        if multi_max_iter:
            mo.max_iter = min(int(o.max_iter), int(multi_max_iter))
            mo.trust_region_fallback = 0
        dev = torch.device("cuda", ctx.device)
        Td, mud, xid = (torch.tensor(_d(a), device=dev) for a in (T, mu, xi))
        saved = getattr(self, "judge_mask", None)
        ctx.set_judge_mask(None, None, None)
        def host(r, keys=("x", "masses", "omega", "converged")):
            return {k: r[k].cpu().numpy() for k in keys}

        try:
            seeds = self.infer(Td, mud, xid)
            q = host(ctx.solve_batch(Td, mud, xid, policy=SEED_SURROGATE, seed=seeds, p_num=p_num, t_num=t_num,
                                     opts=o))
            m = host(ctx.solve_batch(Td, mud, xid, policy=SEED_MULTI, p_num=p_num, t_num=t_num, opts=mo))
            amb = (q["converged"] != 1) | (m["converged"] != 1) | ~same_root(q, m)
            if plain_roots is None:
                pr = host(ctx.solve_batch(Td, mud, xid, policy=SEED_DEFAULT, p_num=p_num, t_num=t_num, opts=o))
                amb |= (pr["converged"] != 1) | ~same_root(pr, m)
            elif isinstance(plain_roots, dict):
                amb |= (np.asarray(plain_roots["converged"]) != 1) | ~same_root(plain_roots, m)
            else:
                px = _d(plain_roots).reshape(-1, 5)
                amb |= (np.abs(px - m["x"]) / np.maximum(np.abs(m["x"]), 1.0)).max(axis=1) > 1e-8
            if multi_root:
                # also every probe where the HADRON and QUARK seeds reach distinct roots
                r3 = host(ctx.solve_batch(Td, mud, xid, policy=SEED_MULTIROOT3, seed=seeds, p_num=p_num,
                                          t_num=t_num, opts=o), ("flags",))
                amb |= (r3["flags"] & PT_FIRST_ORDER) != 0
            return amb
        finally:
            if saved is not None:
                ctx.set_judge_mask(saved, self.in_lo, self.in_hi)


def same_root(a, b, omega_rtol=1e-10):
    """Per point: do solve results a and b (dicts with x, masses, omega, as
    numpy arrays or both as torch tensors) hold the same root?
    DualBranchScan's sameness test (|d phi_u| < 0.01, |d Phi| < 0.01, |d M_u|
    < 0.01 fm^-1; DualBranchScan.jl:409-419) and Omega equal to omega_rtol --
    two Newton runs that stop within the solver's tolerances of one root can
    differ in x by more than 1e-9 where the Hessian is nearly singular (next
    to the critical end point), never in Omega, which is stationary there."""
    if type(a["x"]).__module__.startswith("torch"):
        return (((a["x"][:, 0] - b["x"][:, 0]).abs() < 0.01) & ((a["x"][:, 3] - b["x"][:, 3]).abs() < 0.01)
                & ((a["masses"][:, 0] - b["masses"][:, 0]).abs() < 0.01)
                & ((a["omega"] - b["omega"]).abs() <= omega_rtol * b["omega"].abs()))
    ax, bx = np.asarray(a["x"]), np.asarray(b["x"])
    am, bm = np.asarray(a["masses"]), np.asarray(b["masses"])
    ao, bo = np.asarray(a["omega"]), np.asarray(b["omega"])
    return ((np.abs(ax[:, 0] - bx[:, 0]) < 0.01) & (np.abs(ax[:, 3] - bx[:, 3]) < 0.01)
            & (np.abs(am[:, 0] - bm[:, 0]) < 0.01) & (np.abs(ao - bo) <= omega_rtol * np.abs(bo)))


def _dilate(mask, n, full=True):
    """n steps of binary dilation over the 26 (full) or 6 (face) neighbours."""
    m = mask.astype(bool)
    for _ in range(n):
        g = m.copy()
        p = np.pad(m, 1)
        nb = m.shape
        for d0 in (-1, 0, 1):
            for d1 in (-1, 0, 1):
                for d2 in (-1, 0, 1):
                    if not full and abs(d0) + abs(d1) + abs(d2) != 1:
                        continue
                    g |= p[1 + d0:1 + d0 + nb[0], 1 + d1:1 + d1 + nb[1], 1 + d2:1 + d2 + nb[2]]
        m = g
    return m.astype(np.uint8)


# launch geometry of the set-up solves (training targets, judge-mask probes):
# fixed, so that the trained network does not depend on the geometry a caller
# picked for its scans (lanes per point sets the order of the node sums, hence
# the last bits of the targets, hence the training run)
SETUP_LAUNCH = dict(lanes_per_point=8, chunk=1, waves_per_cu=16, tail_lanes_per_point=0)


# Adam phases (steps, lr) of learned_seed: a step decay.  With a constant
# rate the network that comes out is whatever state the last steps leave;
# a late loss spike then makes seeds that double the scan's work.
LR_SCHEDULE = ((2000, 3e-3), (2000, 1.5e-3), (2000, 6e-4), (2000, 2e-4))


def learned_seed(ctx, p_num=64, t_num=4, opts=None, n_train=200_000, schedule=LR_SCHEDULE, batch=8192,
                 n_probe=1_000_000, multi_max_iter=12):
    """The learned seed as the scans use it, built on the context's GPU: plain
    (DefaultSeed) roots at n_train points of the training box
    (training_points, default_rng(4)), Adam on them (the (steps, lr) phases
    of `schedule`, batch draws seeded 4, 5, ... per phase), and the judge
    mask from those points plus n_probe random probes (default_rng(5)).
    Returns (Surrogate, info) with the set-up's figures; the context's launch
    geometry is restored afterwards."""
    import threading
    import time
    t0 = time.perf_counter()
    saved = ctx.launch_cfg()
    ctx.set_launch_cfg(**SETUP_LAUNCH)
    try:
        T, mu, xi = training_points(n_train, seed=4)
        tr = ctx.solve_batch(T, mu, xi, policy=_lib.SEED_DEFAULT, p_num=p_num, t_num=t_num, opts=opts)
        ok = tr["converged"] == 1
        t1 = time.perf_counter()
        # the probe points are drawn on the host while the GPU trains (the
        # training call releases the interpreter lock)
        probes = {}

        def draw():
            Tp, mup, xip = training_points(n_probe, seed=5)
            probes.update(T=np.r_[T, Tp], mu=np.r_[mu, mup], xi=np.r_[xi, xip])

        drawer = threading.Thread(target=draw)
        drawer.start()
        s = Surrogate(ctx.device, init_seed=4)
        s.fit_output_scaling(tr["x"][ok])
        loss, train_ms = [], 0.0
        try:
            for i, (steps, lr) in enumerate(schedule):
                loss.append(s.train(T[ok], mu[ok], xi[ok], tr["x"][ok], steps=steps, batch=batch, lr=lr,
                                    seed=4 + i))
                train_ms += s.last_ms
        finally:
            drawer.join()
        loss = np.concatenate(loss)
        t2 = time.perf_counter()
        s.build_judge_mask(ctx, probes["T"], probes["mu"], probes["xi"], p_num=p_num, t_num=t_num, opts=opts,
                           multi_max_iter=multi_max_iter)
    finally:
        ctx.set_launch_cfg(**saved)
    info = dict(train_points=int(ok.sum()), adam_steps=int(sum(n for n, _ in schedule)), batch=batch,
                lr=[list(ph) for ph in schedule], train_kernel_ms=train_ms,
                final_loss=float(np.mean(loss[-200:])), judge_probes=n_train + n_probe,
                judge_mask_fraction=float(s.judge_mask.mean()), setup_s=time.perf_counter() - t0,
                setup_phases_s=dict(training_roots=t1 - t0, training=t2 - t1, judge_mask=time.perf_counter() - t2),
                loss=loss)
    return s, info


def training_points(n=200_000, seed=4):
    """(T, mu, xi) ~ U([20,300] MeV x [0,450] MeV x [-0.4,0.4]) from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(20.0, 300.0, n) / HBARC
    mu = rng.uniform(0.0, 450.0, n) / HBARC
    xi = rng.uniform(-0.4, 0.4, n)
    return T, mu, xi
