"""FP64 numpy restatement of the learned seed's MLP (forward, MSE loss,
backward).  TEST INFRASTRUCTURE ONLY: the checker of the GPU training kernels.

The network has no counterpart in the reference; it is specified by the
build: 3 -> 128 -> 128 -> 128 -> 4, SiLU, MSE over the four normalised
outputs, parameters packed W1[128][3] b1 W2[128][128] b2 W3[128][128] b3
W4[4][128] b4 (34052 values).

`bf16=True` rounds at the same points as the GPU path (weights, hidden
activations h_l, silu'(z_l), and dL/dz_l), so the GPU gradient can be
compared tightly; `bf16=False` is the exact network, pinned by finite
differences in the tests.
"""
import numpy as np

IN, H, OUT = 3, 128, 4
OFF = {}
_o = 0
for _name, _n in (("W1", H * IN), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H), ("W4", OUT * H),
                  ("b4", OUT)):
    OFF[_name] = (_o, _n)
    _o += _n
NPARAM = _o
SHAPES = {"W1": (H, IN), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (H, H), "b3": (H,), "W4": (OUT, H), "b4": (OUT,)}


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return {k: p[o:o + n].reshape(SHAPES[k]) for k, (o, n) in OFF.items()}


def pack(d):
    out = np.zeros(NPARAM)
    for k, (o, n) in OFF.items():
        out[o:o + n] = np.asarray(d[k], dtype=np.float64).ravel()
    return out


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(params, xn, bf=False):
    P = unpack(params)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    W = {k: (r(P[k]) if k.startswith("W") else P[k]) for k in P}
    cache = {"x": xn}
    z1 = xn @ W["W1"].T + W["b1"]
    h1, g1 = _silu(z1)
    h1, g1 = r(h1), r(g1)
    z2 = h1 @ W["W2"].T + W["b2"]
    h2, g2 = _silu(z2)
    h2, g2 = r(h2), r(g2)
    z3 = h2 @ W["W3"].T + W["b3"]
    h3, g3 = _silu(z3)
    h3, g3 = r(h3), r(g3)
    y = h3 @ W["W4"].T + W["b4"]
    cache.update(W=W, h1=h1, h2=h2, h3=h3, g1=g1, g2=g2, g3=g3)
    return y, cache


def loss_grad(params, xn, tn, bf=False):
    """MSE L = sum (y - t)^2 / (4 B) and its gradient (packed)."""
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    y, c = forward(params, xn, bf)
    B = xn.shape[0]
    d = y - tn
    L = float(np.sum(d * d) / (4.0 * B))
    dy = 2.0 * d / (4.0 * B)
    W = c["W"]
    g = {}
    g["W4"] = dy.T @ c["h3"]
    g["b4"] = dy.sum(0)
    dz3 = r((dy @ W["W4"]) * c["g3"])
    g["W3"] = dz3.T @ c["h2"]
    g["b3"] = dz3.sum(0)
    dz2 = r((dz3 @ W["W3"]) * c["g2"])
    g["W2"] = dz2.T @ c["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = r((dz2 @ W["W2"]) * c["g1"])
    g["W1"] = dz1.T @ c["x"]
    g["b1"] = dz1.sum(0)
    return L, pack(g)


def normalise(T, mu, xi, roots, in_lo, in_hi, out_mean, out_std):
    x = np.stack([T, mu, xi], 1)
    xn = (x - np.asarray(in_lo)) * (2.0 / (np.asarray(in_hi) - np.asarray(in_lo))) - 1.0
    y = np.stack([roots[:, 0], roots[:, 2], roots[:, 3], roots[:, 4]], 1)
    tn = (y - np.asarray(out_mean)) / np.asarray(out_std)
    return xn.astype(np.float32).astype(np.float64), tn.astype(np.float32).astype(np.float64)
