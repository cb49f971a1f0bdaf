"""Writes the golden fixtures of tests/golden/ from data files the reference
holds (its own scan outputs and phase-diagram tables).  Nothing of the
reference is run: the rows are copied as data, column-subset only.

    python tests/golden/make_fixtures.py [/path/to/reference]

Sources (paths relative to the reference root):
  data/outputs/results/pnjl/dual_branch_T50.csv        (DualBranchScan, p24/t8, 10 decimals)
  data/outputs/results/relaxtime/gap_transport_scan*.csv  (run_gap_transport_scan.jl,
      PhaseAwareContinuitySeed chains, p12/t6, full Float64 precision)
  data/outputs/results/pnjl/tmu_scan.csv               (TmuScan, 6 decimals)
  data/reference/pnjl/trho_scan_xi{0.0,0.2,0.4}.csv    (TrhoScan, p24/t8, 6 decimals): the
      rows at T = 50, 80, 110 MeV, converged (the curves calculate_phase_structure.jl's
      load_curves_from_csv feeds to maxwell_construction) -- FixedRho states, which are
      FixedMu roots at mu = mu_avg on all three branches of the first-order S-curve
  data/reference/pnjl/spinodals.csv, boundary.csv      (calculate_phase_structure.jl: S-curve
      spinodals and the Maxwell-construction transition line, xi = 0, 0.2, 0.4)
  data/outputs/results/pnjl/bulk_viscosity_xi0.0.csv   (calculate_derivatives.jl, ThermoDerivatives
      bulk_viscosity_coefficients, p24/t8, 7 digits)
  data/outputs/results/pnjl/derivatives_xi0.0.csv      (calculate_derivatives.jl, ThermoDerivatives
      thermo_derivatives + mass_derivatives, p24/t8, 7 digits)
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
    trho = []
    for xi in ["0.0", "0.2", "0.4"]:
        for r in rows_of(os.path.join(ref, f"data/reference/pnjl/trho_scan_xi{xi}.csv")):
            if (r["converged"] == "true" and float(r["rho"]) >= 0
                    and round(float(r["T_MeV"])) in (50, 80, 110)):
                trho.append(r)
    write("trho_scan_T50_80_110.csv",
          ["T_MeV", "rho", "xi", "mu_avg_MeV", "pressure_fm4", "entropy_fm3", "energy_fm4", "phi_u", "phi_d",
           "phi_s", "Phi1", "Phi2", "M_u_MeV", "M_s_MeV"], trho)
    write("spinodals.csv",
          ["xi", "T_MeV", "mu_spinodal_hadron_MeV", "mu_spinodal_quark_MeV", "rho_spinodal_hadron",
           "rho_spinodal_quark"],
          rows_of(os.path.join(ref, "data/reference/pnjl/spinodals.csv")))
    write("phase_boundary.csv", ["xi", "T_MeV", "mu_transition_MeV", "rho_hadron", "rho_quark"],
          rows_of(os.path.join(ref, "data/reference/pnjl/boundary.csv")))
    bulk = os.path.join(ref, "data/outputs/results/pnjl/bulk_viscosity_xi0.0.csv")
    write("bulk_viscosity_xi0.0.csv", list(rows_of(bulk)[0].keys()), rows_of(bulk))
    der = os.path.join(ref, "data/outputs/results/pnjl/derivatives_xi0.0.csv")
    write("derivatives_xi0.0.csv", list(rows_of(der)[0].keys()), rows_of(der))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
