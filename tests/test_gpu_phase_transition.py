"""GPU S-curves (julia_relaxtime_amd.phase_transition.s_curves over batched
FixedMu solves) and their Maxwell construction against the reference's
boundary.csv and against the same logic over the oracle's solves."""
import numpy as np
import pytest

from conftest import golden_rows
from test_oracle_phase_transition import MAXWELL_KW, _oracle_solver

pytestmark = pytest.mark.gpu

LINES = [(50.0, 0.0), (80.0, 0.0), (110.0, 0.0), (80.0, 0.4)]


def test_gpu_s_curves_maxwell(gpu_ctx, oracle):
    from julia_relaxtime_amd import phase_transition as PT
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    mus = np.arange(250.0, 420.0 + 1e-9, 0.25)
    T = [t for t, _ in LINES]
    xi = [x for _, x in LINES]
    print("s_curves on the GPU:", LINES, mus.size, "mu", flush=True)
    gpu = PT.s_curves(T, mus, xi)
    ref = PT.s_curves(T, mus, xi, solver=_oracle_solver(oracle))
    for (t, x), (mg, rg), (mo, ro) in zip(LINES, gpu, ref):
        # same states: the three branches agree point for point
        og, oo = np.lexsort((rg, mg)), np.lexsort((ro, mo))
        assert mg.size == mo.size and np.array_equal(mg[og], mo[oo])
        assert np.max(np.abs(rg[og] - ro[oo])) < 1e-8
        m = PT.maxwell_construction(mg, rg, **MAXWELL_KW)
        mr = PT.maxwell_construction(mo, ro, **MAXWELL_KW)
        b = bd[(x, t)]
        assert m.converged and mr.converged
        assert abs(m.mu_transition - mr.mu_transition) < 1e-6
        assert abs(m.mu_transition - float(b["mu_transition_MeV"])) < 0.06
        assert abs(m.rho_hadron - float(b["rho_hadron"])) < 0.01
        assert abs(m.rho_quark - float(b["rho_quark"])) < 0.01


def test_gpu_maxwell_line_next_to_the_cep(gpu_ctx):
    """The refined lines just below the CEP on the GPU (branch chains and
    batched unstable-branch solves): within 0.01 MeV of boundary.csv."""
    from julia_relaxtime_amd import phase_transition as PT
    bd = {(float(r["xi"]), float(r["T_MeV"])): r for r in golden_rows("phase_boundary.csv")}
    mus = np.arange(250.0, 420.0 + 1e-9, 0.25)
    for xi, T in ((0.0, 130.0), (0.2, 120.0), (0.4, 110.0)):
        print("maxwell_line on the GPU:", xi, T, flush=True)
        m = PT.maxwell_line([T], mus, xi, **MAXWELL_KW)[T]
        assert m.converged and "refined" in m.details
        assert abs(m.mu_transition - float(bd[(xi, T)]["mu_transition_MeV"])) < 0.01
