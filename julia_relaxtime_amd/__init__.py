"""MI355X-native (gfx950) backend of the PNJL equilibrium-scan hot path of
w5851/Julia_RelaxTime: the 2+1-flavour FixedMu gap equations solved by
PNJL.solve with its seed strategies, and the anisotropic one-loop integrals
behind polarization_aniso.

The compute runs in libpnjl_hip.so (hand-written HIP, FP64); this package is
the host layer that mirrors the reference's Julia API (`pnjl`,
`polarization`) over its C ABI (include/pnjl_hip.h).
"""
from . import _lib  # noqa: F401
from .pnjl import (FixedMu, DefaultSeed, MultiSeed, ContinuitySeed, PhaseAwareContinuitySeed,  # noqa: F401
                   SolverResult, solve, solve_batch, update, reset, get_seed, context, HBARC)

__all__ = ["FixedMu", "DefaultSeed", "MultiSeed", "ContinuitySeed", "PhaseAwareContinuitySeed", "SolverResult",
           "solve", "solve_batch", "update", "reset", "get_seed", "context", "HBARC"]
