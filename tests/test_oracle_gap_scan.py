"""The scan_csv_v1 helpers and the gap-scan driver logic
(julia_relaxtime_amd.scan_csv / gap_scan, scripts/utils/scan_csv.jl and
run_gap_transport_scan.jl) driven by the oracle's PhaseAwareContinuitySeed
chains on the CPU, against the reference's own driver outputs."""
import csv
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_rows

HB = 197.327


def test_scan_csv_helpers(tmp_path):
    from julia_relaxtime_amd import scan_csv as S
    p = tmp_path / "a.csv"
    with open(p, "w") as f:
        S.write_metadata(f, {"schema": "scan_csv_v1", "title": "t"})
        S.write_header(f, ["T_MeV", "muB_MeV", "xi", "v"])
        f.write("120.0,0.0,0.0,1\n# comment\n\n130.0,800.0,0.2,x\nbad,1,2,3\n140.0,1\n")
    assert open(p).read().startswith("# schema: scan_csv_v1\n# title: t\nT_MeV,muB_MeV,xi,v\n")
    assert S.read_existing_keys(str(p), ["T_MeV", "muB_MeV", "xi"]) == {(120.0, 0.0, 0.0), (130.0, 800.0, 0.2)}
    assert S.read_existing_keys(str(p), ["T_MeV", "nope"]) == set()
    assert S.read_existing_keys(str(tmp_path / "missing.csv"), ["T_MeV"]) == set()
    # Julia's string(::Float64)
    for x, js in ((120.0, "120.0"), (1e-5, "1.0e-5"), (0.0001, "0.0001"), (6.299558097389024e-12,
                  "6.299558097389024e-12"), (1e6, "1.0e6"), (123456.0, "123456.0"), (float("nan"), "NaN"),
                  (-2.5e20, "-2.5e20"), (266.6666666666667, "266.6666666666667")):
        assert S.julia_string(x) == js


def test_julia_range():
    from julia_relaxtime_amd.gap_scan import julia_range
    assert julia_range(120.0, 400.0, 10.0)[:2] == [120.0, 130.0] and len(julia_range(120.0, 400.0, 10.0)) == 29
    assert julia_range(0.0, 0.3, 0.1) == [0.0, 0.1, 0.2, 0.3]          # not 0.30000000000000004
    assert julia_range(0.0, 800.0, 800.0) == [0.0, 800.0]


def _oracle_solver(oracle):
    def solve(T, mu, xi, L):
        return oracle.solve_batch(T, mu, xi, policy=oracle.SEED_PHASE_AWARE, p_num=12, t_num=6, nthreads=8,
                                  opts=oracle.default_opts(max_iter=40, chain_len=L))
    return solve


@pytest.mark.parametrize("name,xis", [("driver_gap_scan.csv", (0.0,)), ("driver_gap_scan_xi0p2.csv", (0.2,)),
                                      ("driver_gap_scan_xi-0p6to0p6.csv", (-0.6, -0.4, -0.2, 0.0, 0.2, 0.4, 0.6))])
def test_driver_rows_match_reference_output(oracle, tmp_path, name, xis):
    """The reference's gap_transport_scan*.csv (T 120-400 MeV step 10, mu_B 0
    and 800 MeV): same rows in the same order, the key columns byte for byte,
    the gap columns to 1e-12 relative (densities 1e-10)."""
    from julia_relaxtime_amd import gap_scan as GS
    out = str(tmp_path / "scan.csv")
    n = GS.run_gap_scan(out, xi_values=xis, tmin=120.0, tmax=400.0, tstep=10.0, mubmin=0.0, mubmax=800.0,
                        mubstep=800.0, solver=_oracle_solver(oracle))
    ref = golden_rows(name)
    lines = [l for l in open(out) if not l.startswith("#")]
    mine = list(csv.DictReader(lines))
    assert n == len(ref) == len(mine)
    assert lines[0].strip() == ",".join(GS.DRIVER_COLUMNS)
    for a, b in zip(mine, ref):
        for k in ("T_MeV", "muq_MeV", "muB_MeV", "xi", "T_fm", "muq_fm", "converged"):
            assert a[k] == b[k], (k, a[k], b[k])
        for k in ("omega_fm4inv", "P_fm4inv"):          # stationary in x: second order in the root's error
            assert abs(float(a[k]) - float(b[k])) <= 1e-12 * max(1.0, abs(float(b[k]))), k
        for k in ("m_u", "m_d", "m_s", "Phi", "Phibar"):  # roots agree to the Newton stop (ftol 1e-9)
            assert abs(float(a[k]) - float(b[k])) <= 1e-8, k
        for k in ("epsilon_fm4inv", "s_fm3inv", "n_u", "n_d", "n_s", "n_ubar", "n_dbar", "n_sbar", "rho_norm"):
            assert abs(float(a[k]) - float(b[k])) <= 1e-8 * max(1e-3, abs(float(b[k]))), k
        assert a["eta"] == "NaN" and float(a["omega_MeV_fm3"]) == float(a["omega_fm4inv"]) * HB
    iters = np.array([int(a["iterations"]) == int(b["iterations"]) for a, b in zip(mine, ref)])
    assert iters.mean() > 0.95


def test_resume_skips_written_keys(oracle, tmp_path):
    from julia_relaxtime_amd import gap_scan as GS
    out = str(tmp_path / "scan.csv")
    kw = dict(xi_values=(0.0, 0.2), tmin=120.0, tmax=200.0, tstep=20.0, mubmin=0.0, mubmax=800.0, mubstep=400.0,
              solver=_oracle_solver(oracle))
    full = str(tmp_path / "full.csv")
    n_full = GS.run_gap_scan(full, **kw)
    # keep the metadata, the header and the first 7 rows, then resume
    lines = open(full).read().splitlines(True)
    head = [l for l in lines if l.startswith("#")]
    data = [l for l in lines if not l.startswith("#")]
    open(out, "w").write("".join(head + data[:8]))
    n_more = GS.run_gap_scan(out, **kw)
    assert n_more == n_full - 7
    rows = [l for l in open(out) if not l.startswith("#")]
    assert rows.count(rows[0]) == 1                                    # one header
    keys = [tuple(r.split(",")[i] for i in (0, 2, 3)) for r in rows[1:]]
    assert len(keys) == len(set(keys)) == n_full
    assert GS.run_gap_scan(out, **kw) == 0                              # nothing left
    assert GS.run_gap_scan(out, overwrite=True, **kw) == n_full         # overwrite starts again
