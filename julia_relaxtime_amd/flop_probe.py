"""FP64 FLOPs of one node x flavour evaluation of the gap integrand, measured
on the device routine itself.

Runs pnjl_gap_residual_batch (one node pass of every flavour plus the dense
assembly per state) twice on the same 65,536 symmetric states at xi = 0.3:
with p = 64 and with p = 32 momentum nodes (t = 4).  Under

    rocprofv3 --kernel-include-regex residual_kernel \\
        --pmc SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_TRANS_F64 \\
        -- python3 -m julia_relaxtime_amd.flop_probe

the difference of the two dispatches' FP64 counts (64 lanes each, FMA = 2
FLOPs, transcendental = 1) divided by the difference in node evaluations
(states x 2 flavours x 32 x 4) is the per-node cost C_node; the remainder is
the per-state assembly.  The script prints the node-evaluation counts it
used.
"""
import json

import numpy as np


def main(n=65536, xi=0.3, t_num=4, seed=3):
    import torch  # noqa: F401  (loads the HIP runtime before the library)

    from julia_relaxtime_amd import _lib
    rng = np.random.default_rng(seed)
    T = rng.uniform(50, 300, n) / 197.327
    mu = rng.uniform(0, 450, n) / 197.327
    x = np.stack([rng.uniform(-1.9, -0.1, n), np.zeros(n), rng.uniform(-2.3, -0.5, n), rng.uniform(0.1, 0.9, n),
                  rng.uniform(0.1, 0.9, n)], 1)
    x[:, 1] = x[:, 0]
    ctx = _lib.Context(0)
    out = {}
    for p_num in (64, 32):
        ctx.residual_batch(x, T, mu, np.full(n, xi), p_num, t_num)
        out[p_num] = dict(states=n, flavours=2, node_evals=n * 2 * p_num * t_num, kernel_ms=ctx.last_kernel_ms())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
