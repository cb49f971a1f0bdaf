"""pnjl_solve_fixedmu_batch_noalloc (the entry point that allocates nothing:
its node table, queue words, candidate records and per-point lists live in a
fixed device array) against pnjl_solve_fixedmu_batch on the same device
inputs: every output field bit-identical (both run the same kernels at the
same lanes per point; only the grid size and the buffers differ).  Calls in a
row on the shared array: the learned-seed policy with the Omega judge on every
point (the largest use of the array), then a smaller plain solve at other
node counts (the node table is rebuilt), then a still smaller continuity
chain scan.  Sizes beyond the array and host pointers are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HB = 197.327


def _grid(nT, nmu, xis):
    T, mu, xi = np.meshgrid(np.linspace(20, 300, nT), np.linspace(0, 450, nmu), np.asarray(xis, float),
                            indexing="ij")
    return T.ravel() / HB, mu.ravel() / HB, xi.ravel()


def _filled(torch, n, dev):
    """output tensors pre-filled with a sentinel, so an unwritten field shows"""
    from julia_relaxtime_amd._lib import OUT_FIELDS
    return {k: torch.full((n, w) if w > 1 else (n,), -7, dtype=torch.float64 if dt == np.float64 else torch.int32,
                          device=dev) for k, w, dt in OUT_FIELDS}


def _same_bits(a, b, label):
    for k in a:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        if x.dtype == np.float64:
            x, y = x.view(np.int64), y.view(np.int64)
        bad = np.flatnonzero((x != y).reshape(x.shape[0], -1).any(axis=1))
        print(f"[{label}] {k}: {bad.size} of {x.shape[0]} points differ", flush=True)
        assert bad.size == 0, (label, k, bad[:10])


def _both(ctx, torch, dev, T, mu, xi, policy, seed, p_num, t_num, opts, label):
    args = [torch.as_tensor(a, device=dev) for a in (T, mu, xi)]
    sd = torch.as_tensor(seed, device=dev) if seed is not None else None
    n = T.size
    ref = _filled(torch, n, dev)
    got = _filled(torch, n, dev)
    print(f"[{label}] n={n} policy={policy} p_num={p_num} t_num={t_num}: regular entry point", flush=True)
    ctx.solve_batch(*args, policy=policy, seed=sd, p_num=p_num, t_num=t_num, opts=opts, out=ref)
    print(f"[{label}] no-alloc entry point", flush=True)
    ctx.solve_batch_noalloc(*args, policy=policy, seed=sd, p_num=p_num, t_num=t_num, opts=opts, out=got)
    torch.cuda.synchronize()
    _same_bits(got, ref, label)
    conv = ref["converged"].cpu().numpy()
    fl = ref["flags"].cpu().numpy()
    from julia_relaxtime_amd import _lib
    kinds = {name: int(np.sum((fl & getattr(_lib, name)) != 0))
             for name in ("PT_SURR_FALLBACK", "PT_MULTISEED", "PT_SURR_OMEGA", "PT_FIRST_ORDER", "PT_TR_USED")}
    print(f"[{label}] converged {int(conv.sum())} of {n}; flags {kinds}", flush=True)
    return ref


def test_noalloc_matches_regular_entry_point(gpu_ctx):
    import torch
    from julia_relaxtime_amd import _lib
    ctx = gpu_ctx
    dev = torch.device("cuda:0")
    # 1) learned-seed policy, Omega judge on every point: seeds = plain roots
    #    plus a perturbation (some attempts fail and fall back)
    T, mu, xi = _grid(60, 50, (0.0, 0.2))
    assert T.size == 6000 <= _lib.NOALLOC_MAX_POINTS
    plain = ctx.solve_batch(T, mu, xi, policy=_lib.SEED_DEFAULT, p_num=64, t_num=4)
    rng = np.random.default_rng(20261203)
    scale = np.where(rng.random(T.size) < 0.3, 0.8, 0.05)[:, None]
    seeds = plain["x"] * (1.0 + scale * rng.standard_normal(plain["x"].shape))
    lo = np.array([T.min(), mu.min(), xi.min()]) - 1e-3
    hi = np.array([T.max(), mu.max(), xi.max()]) + 1e-3
    ctx.set_judge_mask(np.ones((4, 4, 2), dtype=np.uint8), lo, hi)
    try:
        o = _lib.default_opts(ftol=1e-10, max_iter=200, surrogate_judge=_lib.JUDGE_OMEGA)
        _both(ctx, torch, dev, T, mu, xi, _lib.SEED_SURROGATE, seeds, 64, 4, o, "learned+judge")
    finally:
        ctx.set_judge_mask(None, None, None)
    # 2) a smaller plain DefaultSeed solve at other node counts
    T2, mu2, xi2 = _grid(30, 25, (0.0, -0.3))
    assert T2.size < T.size
    _both(ctx, torch, dev, T2, mu2, xi2, _lib.SEED_DEFAULT, None, 32, 8, _lib.default_opts(max_iter=200), "plain")
    # 3) continuity chains (one chain per kernel group), smaller again
    T3, mu3, xi3 = _grid(20, 20, (0.1,))
    _both(ctx, torch, dev, T3, mu3, xi3, _lib.SEED_CONTINUITY, None, 32, 8,
          _lib.default_opts(max_iter=200, chain_len=20), "chains")


def test_noalloc_rejects_what_it_cannot_hold(gpu_ctx):
    import torch
    from julia_relaxtime_amd import _lib
    ctx = gpu_ctx
    dev = torch.device("cuda:0")
    n = _lib.NOALLOC_MAX_POINTS + 1
    T = torch.full((n,), 0.5, dtype=torch.float64, device=dev)
    with pytest.raises(_lib.PnjlError, match="exceeds the maximum of 8192 points"):
        ctx.solve_batch_noalloc(T, T.clone(), torch.zeros_like(T), p_num=32, t_num=8)
    with pytest.raises(_lib.PnjlError, match="device"):
        ctx.solve_batch_noalloc(np.full(4, 0.5), np.zeros(4), np.zeros(4))
