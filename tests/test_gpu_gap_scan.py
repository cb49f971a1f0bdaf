"""The gap-scan driver (julia_relaxtime_amd.gap_scan) on the GPU: all lines
of the reference's driver grid as PhaseAwareContinuitySeed chains in one
batched call, written as the driver's scan_csv_v1 rows, against the
reference's own gap_transport_scan outputs."""
import csv

import numpy as np
import pytest

from conftest import golden_rows

pytestmark = pytest.mark.gpu
HB = 197.327


@pytest.mark.parametrize("name,xis", [("driver_gap_scan.csv", (0.0,)),
                                      ("driver_gap_scan_xi-0p6to0p6.csv", (-0.6, -0.4, -0.2, 0.0, 0.2, 0.4, 0.6))])
def test_gpu_driver_rows_match_reference_output(gpu_ctx, tmp_path, name, xis):
    from julia_relaxtime_amd import gap_scan as GS
    out = str(tmp_path / "scan.csv")
    n = GS.run_gap_scan(out, xi_values=xis, tmin=120.0, tmax=400.0, tstep=10.0, mubmin=0.0, mubmax=800.0,
                        mubstep=800.0)
    ref = golden_rows(name)
    mine = list(csv.DictReader([l for l in open(out) if not l.startswith("#")]))
    assert n == len(ref) == len(mine)
    for a, b in zip(mine, ref):
        for k in ("T_MeV", "muq_MeV", "muB_MeV", "xi", "T_fm", "muq_fm", "converged"):
            assert a[k] == b[k], (k, a[k], b[k])
        for k in ("omega_fm4inv", "P_fm4inv"):
            assert abs(float(a[k]) - float(b[k])) <= 1e-12 * max(1.0, abs(float(b[k]))), k
        for k in ("m_u", "m_d", "m_s", "Phi", "Phibar"):
            assert abs(float(a[k]) - float(b[k])) <= 1e-8, k
        for k in ("epsilon_fm4inv", "s_fm3inv", "n_u", "n_d", "n_s", "n_ubar", "n_dbar", "n_sbar", "rho_norm"):
            assert abs(float(a[k]) - float(b[k])) <= 1e-8 * max(1e-3, abs(float(b[k]))), k
    assert np.mean([int(a["iterations"]) == int(b["iterations"]) for a, b in zip(mine, ref)]) > 0.95
