"""Pins the MLP oracle (the checker of the GPU training kernels)."""
import numpy as np

from oracle import mlp_oracle as M


def _data(n=40, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    t = rng.normal(0, 1, (n, 4))
    p = np.zeros(M.NPARAM)
    for k, (o, sz) in M.OFF.items():
        fan_in = M.SHAPES[k][1] if k.startswith("W") else 1
        p[o:o + sz] = rng.uniform(-1, 1, sz) * (np.sqrt(6.0 / fan_in) if k.startswith("W") else 0.1)
    return x, t, p


def test_exact_gradient_matches_finite_differences():
    x, t, p = _data()
    L, g = M.loss_grad(p, x, t, bf=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(o, o + n), min(6, n), replace=False) for k, (o, n) in M.OFF.items()])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss_grad(p + e, x, t)[0] - M.loss_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265, 1e-30])
    r = M.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1.0 + 2 ** -7            # above half -> up
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
    assert abs(r[5] - 1e-30) < 1e-32


def test_rounded_path_close_to_exact():
    x, t, p = _data(64, 3)
    L0, g0 = M.loss_grad(p, x, t, bf=False)
    L1, g1 = M.loss_grad(p, x, t, bf=True)
    assert abs(L1 - L0) < 2e-2 * abs(L0)
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)
