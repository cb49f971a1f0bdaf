/* sanitize_driver.c -- runs every exported oracle routine once over a small
 * set of inputs, edge cases included, for an AddressSanitizer +
 * UndefinedBehaviorSanitizer build of the oracle (make -C oracle sanitize).
 * Test infrastructure only.  Any out-of-bounds access, use of freed memory,
 * signed overflow, misaligned access, bad shift or non-finite-to-int
 * conversion aborts the run with the sanitizer's report; a clean run prints
 * the number of calls made and exits 0. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pnjl_hip.h"

/* oracle entry points (pnjl_gap_oracle.c, pnjl_polar_oracle.c) */
void oracle_default_params(pnjl_model_params* o);
void oracle_set_params(const pnjl_model_params* p);
void oracle_gauleg(double a, double b, int n, double* x, double* w);
void oracle_nodes(int p_num, int t_num, double* p_mesh, double* c_mesh, double* C_mesh);
double oracle_log_term(double E, double mu, double T, double Phi, double Phib);
double oracle_omega(const double* x, double mu, double T, double xi, int p_num, int t_num);
void oracle_vacuum_integral(double M, double* out3);
int oracle_set_judge_mask(int nT, int nmu, int nxi, const double* lo, const double* hi, const uint8_t* mask);
void oracle_set_boundary(int nrows, const double* xi_r, const double* T_r, const double* mu_r, int ncep,
                         const double* cxi, const double* cT, const double* cmu);
double oracle_interp_mu_c(double xi, double T_MeV);
int oracle_current_phase(double xi, double T_MeV, double mu_MeV);
int oracle_solve_batch(int64_t n, const double* T, const double* mu, const double* xi, const double* seed,
                       int policy, int p_num, int t_num, const pnjl_solve_opts* o, const pnjl_solve_out* out,
                       int nthreads);
void oracle_thermo(const double* x, double T, double mu, double xi, int p_num, int t_num, double* out12);
int oracle_thermo_derivs(const double* x, double T, double mu, double xi, int p_num, int t_num, double* out);
void oracle_B0(double lam, double k, double m1, double mu1, double m2, double mu2, double T, double Phi,
               double Phib, double* re, double* im);
void oracle_B0_correction(double lam, double k, double m1, double m2, double mu1, double mu2, double T, double Phi,
                          double Phib, double xi, double* re, double* im);
void oracle_tilde_B0_correction(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi,
                                double Phib, double xi, double* re, double* im);
double oracle_real_tool(double cx, double dc);
double oracle_imag_tool(double cx, double dc);
double oracle_A(double m, double mu, double T, double Phi, double Phib, int n, const double* xp, const double* wp);
double oracle_A_aniso(double m, double mu, double T, double Phi, double Phib, double xi, int np, const double* xp,
                      const double* wp, int nc, const double* xc, const double* wc);
int oracle_polarization_batch(int64_t n, const pnjl_pol_query* q, double* re, double* im, int nthreads);

static const double HB = 197.327;
static long g_calls = 0;
static double g_sink = 0.0;   /* keeps results live; non-finite values are expected at some inputs */

static void sink(double v)
{
    g_calls++;
    if (isfinite(v)) g_sink += v * 1e-30;
}

#define NPT 18

static void solve_all_policies(void)
{
    /* T, mu in MeV (converted below), xi: the plain grid's corners, the
     * transition region, very low T, large mu/T, both signs of xi */
    const double TM[NPT] = {20, 20, 300, 300, 150, 150, 120, 100, 80, 60, 5, 5, 250, 180, 140, 40, 90, 200};
    const double MM[NPT] = {0, 450, 0, 450, 200, 330, 310, 300, 320, 340, 100, 600, 10, 250, 280, 370, 0, 420};
    const double XI[NPT] = {0, 0.4, -0.4, 0, 0.2, -0.2, 0.05, 0, 0.3, -0.3, 0, 0.1, 0, -0.1, 0.4, 0, -0.4, 0.25};
    double T[NPT], mu[NPT], xi[NPT], seed[NPT * 5];
    for (int i = 0; i < NPT; i++) {
        T[i] = TM[i] / HB;
        mu[i] = MM[i] / HB;
        xi[i] = XI[i];
    }
    double x[NPT * 5], om[NPT], pr[NPT], rho[NPT], en[NPT], ee[NPT], ms[NPT * 3], nq[NPT * 3], nqb[NPT * 3],
        res[NPT];
    int32_t it[NPT], cv[NPT], fl[NPT], st[NPT], ps[NPT];
    pnjl_solve_out out = {x, om, pr, rho, en, ee, ms, nq, nqb, res, it, cv, fl, st, ps};
    pnjl_solve_opts o;
    memset(&o, 0, sizeof(o));
    o.xtol = 1e-9;
    o.ftol = 1e-10;
    o.residual_norm_max = 1e-6;
    o.phi_tol = 1e-8;
    o.max_iter = 60;
    o.trust_region_fallback = 1;
    o.auto_multiseed_fallback = 1;
    o.surrogate_max_steps = 8;
    o.chain_len = 6;
    o.surrogate_judge = PNJL_JUDGE_REFERENCE;

    /* phase tables and a judge mask for the policies that read them */
    const double bxi[6] = {0, 0, 0, 0.1, 0.1, 0.1}, bT[6] = {10, 80, 120, 10, 80, 120},
                 bmu[6] = {340, 320, 290, 345, 325, 295};
    const double cxi[2] = {0, 0.1}, cT[2] = {130, 128}, cmu[2] = {280, 285};
    oracle_set_boundary(6, bxi, bT, bmu, 2, cxi, cT, cmu);
    sink(oracle_interp_mu_c(0.0, 100.0));
    sink((double)oracle_current_phase(0.1, 60.0, 400.0));
    uint8_t mask[4 * 4 * 2];
    for (int i = 0; i < 32; i++) mask[i] = (uint8_t)(i % 3 == 0);
    const double lo[3] = {0.0, 0.0, -0.5}, hi[3] = {2.0, 3.5, 0.5};
    oracle_set_judge_mask(4, 4, 2, lo, hi, mask);

    const int sizes[2][2] = {{64, 8}, {24, 4}};
    const int policies[] = {PNJL_SEED_DEFAULT, PNJL_SEED_HADRON,    PNJL_SEED_QUARK,     PNJL_SEED_MULTI,
                            PNJL_SEED_EXPLICIT, PNJL_SEED_CONTINUITY, PNJL_SEED_PHASE_AWARE, PNJL_SEED_SURROGATE,
                            PNJL_SEED_MULTIROOT3};
    for (int s = 0; s < 2; s++) {
        const int pn = sizes[s][0], tn = sizes[s][1];
        /* seeds for the seeded policies: the DefaultSeed roots, perturbed */
        oracle_solve_batch(NPT, T, mu, xi, NULL, PNJL_SEED_DEFAULT, pn, tn, &o, &out, 1);
        for (int i = 0; i < NPT * 5; i++) seed[i] = x[i] * (1.0 + 0.3 * sin(1.7 * i));
        for (size_t k = 0; k < sizeof(policies) / sizeof(policies[0]); k++) {
            for (int j = 0; j < 2; j++) {
                o.surrogate_judge = j ? PNJL_JUDGE_OMEGA : PNJL_JUDGE_REFERENCE;
                int rc = oracle_solve_batch(NPT, T, mu, xi, seed, policies[k], pn, tn, &o, &out, 1);
                sink((double)rc);
                for (int i = 0; i < NPT; i++) {
                    sink(om[i]);
                    sink(res[i]);
                }
            }
        }
        /* thermodynamics and implicit derivatives at the roots (and at the seeds) */
        for (int i = 0; i < NPT; i++) {
            double th[12], d[PNJL_THERMO_DERIV_FIELDS];
            oracle_thermo(x + 5 * i, T[i], mu[i], xi[i], pn, tn, th);
            sink(th[0]);
            sink((double)oracle_thermo_derivs(x + 5 * i, T[i], mu[i], xi[i], pn, tn, d));
            sink(d[5]);
            sink((double)oracle_thermo_derivs(seed + 5 * i, T[i], mu[i], xi[i], pn, tn, d));
            sink(oracle_omega(x + 5 * i, mu[i], T[i], xi[i], pn, tn));
        }
    }
    /* node meshes and rules of several sizes */
    for (int n = 1; n <= 40; n += 13) {
        double px[64], pw[64];
        oracle_gauleg(0.0, 10.0, n, px, pw);
        sink(px[n - 1] + pw[0]);
    }
    double pm[64 * 8], cm[64 * 8], Cm[64 * 8];
    oracle_nodes(64, 8, pm, cm, Cm);
    sink(Cm[511]);
    /* the log term at the floor and at extreme exponents */
    const double Es[4] = {0.01, 1.0, 10.0, 60.0};
    for (int i = 0; i < 4; i++) {
        sink(oracle_log_term(Es[i], 2.0, 0.05, 0.0, 0.0));
        sink(oracle_log_term(Es[i], 0.0, 1.5, -0.2, 1.1));
    }
    double v3[3];
    oracle_vacuum_integral(1e-3, v3);
    oracle_vacuum_integral(-2.0, v3);
    sink(v3[0]);
}

static void polarization(void)
{
    double np[32], wp[32], nc[8], wc[8];
    oracle_gauleg(0.0, 10.0, 32, np, wp);
    oracle_gauleg(0.0, 1.0, 8, nc, wc);
    const double k0s[5] = {0.0, 0.3, 1.2, -0.8, 3.0}, ks[4] = {0.0, 0.2, 1.0, 2.5};
    const double xis[3] = {0.0, 0.2, -0.4};
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 4; b++)
            for (int c = 0; c < 3; c++) {
                double re, im;
                oracle_B0(k0s[a], ks[b], 1.6, 0.5, 2.4, 0.5, 0.8, 0.3, 0.35, &re, &im);
                sink(re);
                sink(im);
                oracle_B0_correction(k0s[a], ks[b], 1.6, 2.4, 0.5, 0.5, 0.8, 0.3, 0.35, xis[c], &re, &im);
                sink(re);
                sink(im);
                /* degenerate masses and a vanishing mass difference */
                oracle_B0_correction(k0s[a], ks[b], 0.24, 0.24, 0.0, 0.0, 0.17, 0.15, 0.15, xis[c], &re, &im);
                sink(re);
                for (int sg = 0; sg < 2; sg++) {
                    oracle_tilde_B0_correction(sg, k0s[a], ks[b], 0.38, 0.24, -0.05, 0.17, 0.15, 0.15, xis[c], &re,
                                               &im);
                    sink(re);
                }
            }
    const double cx[6] = {0.0, 0.5, -2.0, 3.0, 1e-9, -1e-9}, dc[6] = {1.0, -0.25, 0.5, 3.0, 2.0, -1.0};
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            sink(oracle_real_tool(cx[i], dc[j]));
            sink(oracle_imag_tool(cx[i], dc[j]));
        }
    for (int c = 0; c < 3; c++) {
        sink(oracle_A(1.6, 0.5, 0.8, 0.3, 0.35, 32, np, wp));
        sink(oracle_A_aniso(1.6, 0.5, 0.8, 0.3, 0.35, xis[c], 32, np, wp, 8, nc, wc));
    }
    /* the batch entry point over both channels, widths, the symmetrised B0
     * and the (k0, k) = (0, 0) cell, plus one bad channel */
    pnjl_pol_query q[41];
    memset(q, 0, sizeof(q));
    for (int i = 0; i < 40; i++) {
        q[i].channel = (i & 1) ? PNJL_CHANNEL_S : PNJL_CHANNEL_P;
        q[i].num_s = (i / 2) % 3;
        q[i].width = (i / 6) % 2;
        q[i].k0 = k0s[i % 5];
        q[i].gamma = 0.05;
        q[i].k = ks[(i / 5) % 4];
        q[i].m1 = 1.6;
        q[i].m2 = (i % 4 == 0) ? 1.6 : 2.4;
        q[i].mu1 = 0.5;
        q[i].mu2 = 0.5;
        q[i].T = 0.8;
        q[i].Phi = 0.3;
        q[i].Phib = 0.35;
        q[i].xi = xis[i % 3];
        q[i].A1 = -3.0;
        q[i].A2 = -4.0;
    }
    q[40] = q[0];
    q[40].channel = 99;
    double re[41], im[41];
    sink((double)oracle_polarization_batch(40, q, re, im, 1));
    sink((double)oracle_polarization_batch(41, q, re, im, 1));
    for (int i = 0; i < 40; i++) sink(re[i] + im[i]);
}

int main(void)
{
    pnjl_model_params p;
    oracle_default_params(&p);
    oracle_set_params(&p);
    solve_all_policies();
    polarization();
    printf("sanitizer run: %ld oracle calls, no report (checksum %.3e)\n", g_calls, g_sink);
    return 0;
}
