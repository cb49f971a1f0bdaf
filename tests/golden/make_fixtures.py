"""Writes the golden fixtures of tests/golden/ from data files the reference
holds (its own scan outputs and phase-diagram tables).  Nothing of the
reference is run: the rows are copied as data, column-subset only.

    python tests/golden/make_fixtures.py [/path/to/reference]

Sources (paths relative to the reference root):
  data/outputs/results/pnjl/dual_branch_T50.csv        (DualBranchScan, p24/t8, 10 decimals)
  data/outputs/results/relaxtime/gap_transport_scan*.csv  (run_gap_transport_scan.jl,
      PhaseAwareContinuitySeed chains, p12/t6, full Float64 precision)
  data/outputs/results/pnjl/tmu_scan.csv               (TmuScan, 6 decimals)
"""
import csv
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def rows_of(path):
    with open(path) as f:
        lines = [l for l in f if not l.startswith("#")]
    return list(csv.DictReader(lines))


def write(name, header, rows):
    with open(os.path.join(HERE, name), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        for r in rows:
            w.writerow([r[h] for h in header])


def main(ref):
    write("dual_branch_T50.csv",
          ["T_MeV", "mu_MeV", "xi", "branch", "omega", "pressure", "rho", "entropy", "energy", "phi_u", "phi_d",
           "phi_s", "Phi1", "Phi2", "M_u_MeV", "M_d_MeV", "M_s_MeV"],
          rows_of(os.path.join(ref, "data/outputs/results/pnjl/dual_branch_T50.csv")))
    gap_cols = ["T_MeV", "muq_MeV", "muB_MeV", "xi", "T_fm", "muq_fm", "converged", "iterations", "residual_norm",
                "Phi", "Phibar", "m_u", "m_d", "m_s", "rho_norm", "omega_fm4inv", "P_fm4inv", "epsilon_fm4inv",
                "s_fm3inv", "n_u", "n_d", "n_s", "n_ubar", "n_dbar", "n_sbar"]
    for src in ["gap_transport_scan.csv", "gap_transport_scan_xi0p2.csv", "gap_transport_scan_xi-0p6to0p6.csv"]:
        write(src.replace("gap_transport_scan", "driver_gap_scan"), gap_cols,
              rows_of(os.path.join(ref, "data/outputs/results/relaxtime", src)))
    write("tmu_scan.csv",
          ["T_MeV", "mu_MeV", "xi", "pressure_fm4", "rho", "entropy_fm3", "energy_fm4", "phi_u", "phi_d", "phi_s",
           "Phi1", "Phi2", "iterations", "converged"],
          rows_of(os.path.join(ref, "data/outputs/results/pnjl/tmu_scan.csv")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
