"""GPU T-mu scans (julia_relaxtime_amd.tmu_scan over batched fixed-seed
solves) against the reference's tmu_scan.csv and against the same logic
driven by the oracle."""
import csv

import numpy as np
import pytest

from test_oracle_tmu_scan import COLS, _oracle_solver, check_against_reference

pytestmark = pytest.mark.gpu


def test_gpu_tmu_scan(gpu_ctx, oracle, tmp_path):
    from julia_relaxtime_amd import tmu_scan as TS
    g_out, o_out = str(tmp_path / "g.csv"), str(tmp_path / "o.csv")
    TS.run_tmu_scan(output_path=g_out)
    TS.run_tmu_scan(output_path=o_out, solver=_oracle_solver(oracle))
    assert check_against_reference(oracle, g_out) > 0.9
    g, o = list(csv.DictReader(open(g_out))), list(csv.DictReader(open(o_out)))
    same = [all(abs(float(a[k]) - float(b[k])) <= 1.5e-6 for k in COLS) for a, b in zip(g, o)]
    assert np.mean(same) >= 0.99
    assert all(a["converged"] == b["converged"] for a, b, s in zip(g, o, same) if s)
