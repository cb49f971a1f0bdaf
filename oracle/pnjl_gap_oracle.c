/*
 * pnjl_gap_oracle.c -- CPU FP64 restatement of the reference's PNJL FixedMu gap
 * solver.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
 * leg).  No product code links or calls it.
 *
 * Follows, function by function:
 *   parameters        src/Constants_PNJL.jl:80-103, config/pnjl/default.toml
 *   gauleg            src/integration/GaussLegendre.jl:94-119 (FastGaussQuadrature
 *                     gausslegendre restated as Newton on the Legendre recurrence)
 *   nodes             src/pnjl/core/Integrals.jl:67-97 (theta weights x2, p on [0,10])
 *   vacuum_integral   src/pnjl/core/Integrals.jl:123-145
 *   calculate_log_term src/pnjl/core/Integrals.jl:193-238 (two LSE branches, 1e-16 floor)
 *   calculate_log_sum src/pnjl/core/Integrals.jl:248-259
 *   mass/chiral/U     src/pnjl/core/Thermodynamics.jl:81-131
 *   omega             src/pnjl/core/Thermodynamics.jl:181-195
 *   gap_conditions    src/pnjl/solver/Conditions.jl:72-81 (F = grad_x P, analytic)
 *   Jacobian          ForwardDiff jacobian of F == -Hessian(Omega), analytic
 *   thermo            src/pnjl/core/Thermodynamics.jl:215-244 (rho = dP/dmu, s = dP/dT)
 *   densities         src/pnjl/core/Thermodynamics.jl:255-281, src/QuarkDistribution.jl:14-51
 *   Newton            NLsolve 4.5.1 newton_ (method=:newton, Static line search,
 *                     xtol/ftol test, Levenberg step on a singular Jacobian), as
 *                     called at src/pnjl/solver/ImplicitSolver.jl:112
 *   trust region      NLsolve 4.5.1 trust_region_ (dogleg, autoscale, factor 1,
 *                     eta 1e-4), as called at ImplicitSolver.jl:144
 *   fallback/choice   src/pnjl/solver/ImplicitSolver.jl:50-151
 *   solve             src/pnjl/solver/ImplicitSolver.jl:211-328
 *   solve_multi       src/pnjl/solver/ImplicitSolver.jl:532-559
 *   seeds             src/pnjl/solver/SeedStrategies.jl:56-107, 131-136, 204-284,
 *                     303-345, 388-475, 762-888
 *   driver chains     scripts/relaxtime/run_gap_transport_scan.jl:410-442
 *
 * Everything is computed in 5 unknowns and 3 flavours exactly as the reference
 * does; the derivatives are the exact derivatives of the implemented Omega
 * (the reference gets them by forward-mode AD, which is exact to rounding).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pnjl_hip.h"

#define PI_D 3.14159265358979323846
#define POLYAKOV_EPS 1e-16

/* ------------------------------------------------------------------------- */
/* parameters                                                                */
/* ------------------------------------------------------------------------- */
static pnjl_model_params P;
static int P_set = 0;

void oracle_default_params(pnjl_model_params* o)
{
    /* src/Constants_PNJL.jl:80-103 */
    o->hbarc = 197.327;
    o->Lambda = 602.3 / o->hbarc;
    o->G = 1.835 / (o->Lambda * o->Lambda);
    double L2 = o->Lambda * o->Lambda;
    o->K = 12.36 / (L2 * L2 * o->Lambda);
    o->m_ud0 = 5.5 / o->hbarc;
    o->m_s0 = 140.7 / o->hbarc;
    o->T0 = 210.0 / o->hbarc;
    o->a0 = 3.51;
    o->a1 = -2.47;
    o->a2 = 15.2;
    o->b3 = -1.75;
    o->rho0 = 0.16;
    o->N_color = 3;
}

void oracle_set_params(const pnjl_model_params* p)
{
    if (p)
        P = *p;
    else
        oracle_default_params(&P);
    P_set = 1;
}

static void ensure_params(void)
{
    if (!P_set) oracle_set_params(NULL);
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre                                                            */
/* ------------------------------------------------------------------------- */
/* n-point rule on [-1,1], ascending; Newton on the three-term recurrence. */
void oracle_gl_std(int n, double* x, double* w)
{
    int half = (n + 1) / 2;
    for (int i = 0; i < half; i++) {
        double z = cos(PI_D * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            pp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) <= 1e-17 * (1.0 + fabs(z)) && it > 2) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            pp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        double wi = 2.0 / ((1.0 - z * z) * pp * pp);
        /* z is the i-th largest node */
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

/* gauleg(a, b, n)  src/integration/GaussLegendre.jl:94-119 */
void oracle_gauleg(double a, double b, int n, double* x, double* w)
{
    double* xs = (double*)malloc(sizeof(double) * n);
    double* ws = (double*)malloc(sizeof(double) * n);
    oracle_gl_std(n, xs, ws);
    double scale = (b - a) / 2.0, shift = (b + a) / 2.0;
    for (int i = 0; i < n; i++) {
        x[i] = scale * xs[i] + shift;
        w[i] = scale * ws[i];
    }
    free(xs);
    free(ws);
}

/* build_nodes(p_num, t_num)  src/pnjl/core/Integrals.jl:87-97
 * column-major p x t meshes, index j + l*p_num. */
void oracle_nodes(int p_num, int t_num, double* p_mesh, double* c_mesh, double* C_mesh)
{
    double* pn = (double*)malloc(sizeof(double) * p_num);
    double* pw = (double*)malloc(sizeof(double) * p_num);
    double* tn = (double*)malloc(sizeof(double) * t_num);
    double* tw = (double*)malloc(sizeof(double) * t_num);
    oracle_gauleg(0.0, 10.0, p_num, pn, pw);
    oracle_gauleg(0.0, 1.0, t_num, tn, tw);
    double twopi2 = (2.0 * PI_D) * (2.0 * PI_D);
    for (int l = 0; l < t_num; l++) {
        double wt = tw[l] * 2.0;
        for (int j = 0; j < p_num; j++) {
            int idx = j + l * p_num;
            p_mesh[idx] = pn[j];
            c_mesh[idx] = tn[l];
            C_mesh[idx] = (pw[j] * wt) * (pn[j] * pn[j]) / twopi2;
        }
    }
    free(pn);
    free(pw);
    free(tn);
    free(tw);
}

/* ------------------------------------------------------------------------- */
/* node tables cached per (p_num, t_num)                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    int p_num, t_num, n;
    double *p, *c, *C;
} nodes_t;

static void nodes_init(nodes_t* nd, int p_num, int t_num)
{
    nd->p_num = p_num;
    nd->t_num = t_num;
    nd->n = p_num * t_num;
    nd->p = (double*)malloc(sizeof(double) * nd->n);
    nd->c = (double*)malloc(sizeof(double) * nd->n);
    nd->C = (double*)malloc(sizeof(double) * nd->n);
    oracle_nodes(p_num, t_num, nd->p, nd->c, nd->C);
}

static void nodes_free(nodes_t* nd)
{
    free(nd->p);
    free(nd->c);
    free(nd->C);
}

/* ------------------------------------------------------------------------- */
/* the integrand: ln F(a) with its derivatives                               */
/* ------------------------------------------------------------------------- */
/* ln(1 + 3 c1 e^a + 3 c2 e^{2a} + e^{3a}) exactly as calculate_log_term
 * (src/pnjl/core/Integrals.jl:203-218): a > 0 uses the normalised form
 * 3a + ln(e^{-3a} + 3 c1 e^{-2a} + 3 c2 e^{-a} + 1); the argument of the log is
 * floored at 1e-16 (a floored term has zero derivative).
 * Quark: (c1, c2) = (Phi, Phibar); antiquark: (Phibar, Phi), argument b.     */
typedef struct {
    double v;              /* value                        */
    double a, c1, c2;      /* first derivatives            */
    double aa, ac1, ac2;   /* second derivatives           */
    double c1c1, c1c2, c2c2;
} lnF_t;

static void lnF_eval(double a, double c1, double c2, lnF_t* o)
{
    double D, r1, r2, r3;
    memset(o, 0, sizeof(*o));
    if (a > 0) {
        double m_a = 3.0 * a;
        double e_a_m = exp(-2.0 * a);
        double e_2a_m = exp(-a);
        double e_neg_m = exp(-m_a);
        D = e_neg_m + 3.0 * c1 * e_a_m + 3.0 * c2 * e_2a_m + 1.0;
        if (D < POLYAKOV_EPS) {
            o->v = m_a + log(POLYAKOV_EPS);
            o->a = 3.0;
            return;
        }
        o->v = m_a + log(D);
        r1 = e_a_m / D;
        r2 = e_2a_m / D;
        r3 = 1.0 / D;
    } else {
        double ea = exp(a);
        double e2 = ea * ea;
        double e3 = ea * e2;
        D = 1.0 + 3.0 * c1 * ea + 3.0 * c2 * e2 + e3;
        if (D < POLYAKOV_EPS) {
            o->v = log(POLYAKOV_EPS);
            return;
        }
        o->v = log(D);
        r1 = ea / D;
        r2 = e2 / D;
        r3 = e3 / D;
    }
    double f = c1 * r1 + 2.0 * c2 * r2 + r3;
    o->a = 3.0 * f;
    o->c1 = 3.0 * r1;
    o->c2 = 3.0 * r2;
    o->aa = 3.0 * (c1 * r1 + 4.0 * c2 * r2 + 3.0 * r3) - 9.0 * f * f;
    o->ac1 = 3.0 * r1 * (1.0 - 3.0 * f);
    o->ac2 = 3.0 * r2 * (2.0 - 3.0 * f);
    o->c1c1 = -9.0 * r1 * r1;
    o->c1c2 = -9.0 * r1 * r2;
    o->c2c2 = -9.0 * r2 * r2;
}

/* quark_distribution / antiquark_distribution  src/QuarkDistribution.jl:14-51 */
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

double oracle_quark_distribution(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1.0 / T;
    double e1 = clampd(exp(-(E - mu) * beta), 1e-200, 1e200);
    double e2 = e1 * e1, e3 = e2 * e1;
    return (Phi * e1 + 2.0 * Phib * e2 + e3) / (1.0 + 3.0 * Phi * e1 + 3.0 * Phib * e2 + e3);
}

double oracle_antiquark_distribution(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1.0 / T;
    double e1 = clampd(exp(-(E + mu) * beta), 1e-200, 1e200);
    double e2 = e1 * e1, e3 = e2 * e1;
    return (Phib * e1 + 2.0 * Phi * e2 + e3) / (1.0 + 3.0 * Phib * e1 + 3.0 * Phi * e2 + e3);
}

/* calculate_log_term  src/pnjl/core/Integrals.jl:193-238 (value only) */
double oracle_log_term(double E, double mu, double T, double Phi, double Phib)
{
    lnF_t qp, qm;
    double invT = 1.0 / T;
    lnF_eval(-(E - mu) * invT, Phi, Phib, &qp);
    lnF_eval(-(E + mu) * invT, Phib, Phi, &qm);
    return qp.v + qm.v;
}

/* Per-flavour node sums.  S = sum_nodes C * Q(E) where Q = lnF+(a) + lnF-(b),
 * with derivatives in (M, Phi, Phibar) and the sums the thermodynamics need. */
typedef struct {
    double S, SM, SP, SPb;                 /* value, d/dM, d/dPhi, d/dPhibar   */
    double SMM, SMP, SMPb, SPP, SPPb, SPbPb;/* second derivatives              */
    double Sga, Sgb;                       /* sum C dlnF+/da, sum C dlnF-/db   */
    double Sgab;                           /* sum C (a dlnF+/da + b dlnF-/db) */
    double Nq, Nqb;                        /* sum C f+, sum C f- (distributions) */
} flav_t;

static void flavour_sums(const nodes_t* nd, double M, double mu, double T, double Phi, double Phib,
                         double xi, flav_t* o)
{
    memset(o, 0, sizeof(*o));
    double invT = 1.0 / T;
    for (int idx = 0; idx < nd->n; idx++) {
        double p = nd->p[idx], t = nd->c[idx], C = nd->C[idx];
        double pt = p * t;
        double E = sqrt(p * p + M * M + xi * (pt * pt));
        double a = -(E - mu) * invT;
        double b = -(E + mu) * invT;
        lnF_t qp, qm;
        lnF_eval(a, Phi, Phib, &qp);
        lnF_eval(b, Phib, Phi, &qm);
        double Q = qp.v + qm.v;
        double QE = -(qp.a + qm.a) * invT;
        double QEE = (qp.aa + qm.aa) * invT * invT;
        double QP = qp.c1 + qm.c2;
        double QPb = qp.c2 + qm.c1;
        double QEP = -(qp.ac1 + qm.ac2) * invT;
        double QEPb = -(qp.ac2 + qm.ac1) * invT;
        double QPP = qp.c1c1 + qm.c2c2;
        double QPPb = qp.c1c2 + qm.c1c2;
        double QPbPb = qp.c2c2 + qm.c1c1;
        double EM = M / E;
        double EMM = (1.0 - EM * EM) / E;
        o->S += Q * C;
        o->SM += C * (QE * EM);
        o->SP += C * QP;
        o->SPb += C * QPb;
        o->SMM += C * (QEE * EM * EM + QE * EMM);
        o->SMP += C * (QEP * EM);
        o->SMPb += C * (QEPb * EM);
        o->SPP += C * QPP;
        o->SPPb += C * QPPb;
        o->SPbPb += C * QPbPb;
        o->Sga += C * qp.a;
        o->Sgb += C * qm.a;
        o->Sgab += C * (qp.a * a + qm.a * b);
        o->Nq += C * oracle_quark_distribution(E, mu, T, Phi, Phib);
        o->Nqb += C * oracle_antiquark_distribution(E, mu, T, Phi, Phib);
    }
}

/* ------------------------------------------------------------------------- */
/* Omega with its gradient and Hessian                                       */
/* ------------------------------------------------------------------------- */
/* vacuum_integral  src/pnjl/core/Integrals.jl:123-132 and its M derivatives */
static void vacuum_I(double M, double* I, double* dI, double* d2I)
{
    double L = P.Lambda;
    double sgn = (M < 0) ? -1.0 : 1.0;
    double Mt = fabs(M) + 1e-12;
    double s = sqrt(L * L + Mt * Mt);
    double lg = log((L + s) / Mt);
    double poly = L * s * (2.0 * L * L + Mt * Mt);
    double logt = (Mt * Mt * Mt * Mt) * lg;
    *I = (poly - logt) / (16.0 * PI_D * PI_D);
    *dI = sgn * Mt / (4.0 * PI_D * PI_D) * (L * s - Mt * Mt * lg);
    *d2I = (2.0 * L * Mt * Mt / s + L * s - 3.0 * Mt * Mt * lg) / (4.0 * PI_D * PI_D);
}

/* masses M_i(phi) and their derivatives  src/pnjl/core/Thermodynamics.jl:81-88 */
static void mass_vec(const double* x, double M[3], double dM[3][3], double d2M[3][3][3])
{
    double G = P.G, K = P.K;
    double pu = x[0], pd = x[1], ps = x[2];
    M[0] = P.m_ud0 - 4.0 * G * pu + 2.0 * K * pd * ps;
    M[1] = P.m_ud0 - 4.0 * G * pd + 2.0 * K * pu * ps;
    M[2] = P.m_s0 - 4.0 * G * ps + 2.0 * K * pu * pd;
    if (dM) {
        memset(dM, 0, sizeof(double) * 9);
        dM[0][0] = -4.0 * G; dM[0][1] = 2.0 * K * ps; dM[0][2] = 2.0 * K * pd;
        dM[1][0] = 2.0 * K * ps; dM[1][1] = -4.0 * G; dM[1][2] = 2.0 * K * pu;
        dM[2][0] = 2.0 * K * pd; dM[2][1] = 2.0 * K * pu; dM[2][2] = -4.0 * G;
    }
    if (d2M) {
        memset(d2M, 0, sizeof(double) * 27);
        d2M[0][1][2] = d2M[0][2][1] = 2.0 * K;
        d2M[1][0][2] = d2M[1][2][0] = 2.0 * K;
        d2M[2][0][1] = d2M[2][1][0] = 2.0 * K;
    }
}

/* U(T, Phi, Phibar) with derivatives  src/pnjl/core/Thermodynamics.jl:124-160 */
static void polyakov_U(double T, double Phi, double Phib, double* U, double g[2], double h[3], double* UT)
{
    double Tr = P.T0 / T;
    double Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr;
    double Tb = P.b3 * Tr * Tr * Tr;
    double val = 1.0 - 6.0 * Phib * Phi + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * (Phib * Phi) * (Phib * Phi);
    int floored = (val <= 0.0) || (val < POLYAKOV_EPS);
    double lg = floored ? log(POLYAKOV_EPS) : log(val);
    double T2 = T * T, T4 = T2 * T2;
    *U = T4 * (-0.5 * Ta * Phib * Phi + Tb * lg);
    double gP = -6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib;
    double gPb = -6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib;
    double gPP = 24.0 * Phi - 6.0 * Phib * Phib;
    double gPbPb = 24.0 * Phib - 6.0 * Phi * Phi;
    double gPPb = -6.0 - 12.0 * Phi * Phib;
    double lP = 0, lPb = 0, lPP = 0, lPPb = 0, lPbPb = 0;
    if (!floored) {
        lP = gP / val;
        lPb = gPb / val;
        lPP = gPP / val - lP * lP;
        lPPb = gPPb / val - lP * lPb;
        lPbPb = gPbPb / val - lPb * lPb;
    }
    g[0] = T4 * (-0.5 * Ta * Phib + Tb * lP);
    g[1] = T4 * (-0.5 * Ta * Phi + Tb * lPb);
    h[0] = T4 * (Tb * lPP);
    h[1] = T4 * (-0.5 * Ta + Tb * lPPb);
    h[2] = T4 * (Tb * lPbPb);
    if (UT) {
        double U_over_T4 = -0.5 * Ta * Phib * Phi + Tb * lg;
        double dTa = -P.a1 * P.T0 / (T * T) - 2.0 * P.a2 * P.T0 * P.T0 / (T * T * T);
        double dTb = -3.0 * P.b3 * P.T0 * P.T0 * P.T0 / (T4);
        *UT = 4.0 * T * T * T * U_over_T4 + T4 * (dTa * (-0.5 * Phib * Phi) + dTb * lg);
    }
}

typedef struct {
    double omega;
    double grad[5];
    double hess[5][5];
    double M[3];
    flav_t fl[3];
    double UT;
} eval_t;

/* Omega(x), grad Omega, Hessian Omega at fixed (mu_vec, T, xi). */
static void eval_point(const nodes_t* nd, const double* x, const double mu_vec[3], double T, double xi, eval_t* e)
{
    double M[3], dM[3][3], d2M[3][3][3];
    mass_vec(x, M, dM, d2M);
    double G = P.G, K = P.K;
    double pu = x[0], pd = x[1], ps = x[2], Phi = x[3], Phib = x[4];
    /* chiral condensate term  Thermodynamics.jl:112-114 */
    double chi = 2.0 * G * (pu * pu + pd * pd + ps * ps) - 4.0 * K * (pu * pd * ps);
    double dchi[3] = {4.0 * G * pu - 4.0 * K * pd * ps, 4.0 * G * pd - 4.0 * K * pu * ps,
                      4.0 * G * ps - 4.0 * K * pu * pd};
    double d2chi[3][3] = {{4.0 * G, -4.0 * K * ps, -4.0 * K * pd},
                          {-4.0 * K * ps, 4.0 * G, -4.0 * K * pu},
                          {-4.0 * K * pd, -4.0 * K * pu, 4.0 * G}};
    double U, gU[2], hU[3], UT;
    polyakov_U(T, Phi, Phib, &U, gU, hU, &UT);

    double Evac = 0.0;
    double VM[3], VMM[3];
    for (int i = 0; i < 3; i++) {
        double I, dI, d2I;
        vacuum_I(M[i], &I, &dI, &d2I);
        Evac += I;
        VM[i] = -2.0 * P.N_color * dI;
        VMM[i] = -2.0 * P.N_color * d2I;
    }
    Evac = -2.0 * P.N_color * Evac;

    double Ssum = 0.0;
    for (int i = 0; i < 3; i++) {
        flavour_sums(nd, M[i], mu_vec[i], T, Phi, Phib, xi, &e->fl[i]);
        Ssum += e->fl[i].S;
    }
    double Lth = -2.0 * T * Ssum;
    e->omega = chi + U + Evac + Lth;
    e->UT = UT;
    for (int i = 0; i < 3; i++) e->M[i] = M[i];

    double gM[3], hM[3];
    for (int i = 0; i < 3; i++) {
        gM[i] = VM[i] - 2.0 * T * e->fl[i].SM;
        hM[i] = VMM[i] - 2.0 * T * e->fl[i].SMM;
    }
    for (int a = 0; a < 3; a++) {
        double g = dchi[a];
        for (int i = 0; i < 3; i++) g += gM[i] * dM[i][a];
        e->grad[a] = g;
        for (int b = 0; b < 3; b++) {
            double h = d2chi[a][b];
            for (int i = 0; i < 3; i++) h += hM[i] * dM[i][a] * dM[i][b] + gM[i] * d2M[i][a][b];
            e->hess[a][b] = h;
        }
        double hP = 0.0, hPb = 0.0;
        for (int i = 0; i < 3; i++) {
            hP += -2.0 * T * e->fl[i].SMP * dM[i][a];
            hPb += -2.0 * T * e->fl[i].SMPb * dM[i][a];
        }
        e->hess[a][3] = e->hess[3][a] = hP;
        e->hess[a][4] = e->hess[4][a] = hPb;
    }
    double SP = 0, SPb = 0, SPP = 0, SPPb = 0, SPbPb = 0;
    for (int i = 0; i < 3; i++) {
        SP += e->fl[i].SP;
        SPb += e->fl[i].SPb;
        SPP += e->fl[i].SPP;
        SPPb += e->fl[i].SPPb;
        SPbPb += e->fl[i].SPbPb;
    }
    e->grad[3] = gU[0] - 2.0 * T * SP;
    e->grad[4] = gU[1] - 2.0 * T * SPb;
    e->hess[3][3] = hU[0] - 2.0 * T * SPP;
    e->hess[3][4] = e->hess[4][3] = hU[1] - 2.0 * T * SPPb;
    e->hess[4][4] = hU[2] - 2.0 * T * SPbPb;
}

/* ------------------------------------------------------------------------- */
/* exported evaluation entry points (used by the tests)                      */
/* ------------------------------------------------------------------------- */
static nodes_t* cache_nodes(int p_num, int t_num);

double oracle_omega(const double* x, double mu, double T, double xi, int p_num, int t_num)
{
    ensure_params();
    eval_t e;
    double muv[3] = {mu, mu, mu};
    eval_point(cache_nodes(p_num, t_num), x, muv, T, xi, &e);
    return e.omega;
}

/* omega with a general mu vector; returns gradient (5) and Hessian (25). */
void oracle_omega_derivs(const double* x, const double* mu_vec, double T, double xi, int p_num, int t_num,
                         double* omega, double* grad, double* hess)
{
    ensure_params();
    eval_t e;
    eval_point(cache_nodes(p_num, t_num), x, mu_vec, T, xi, &e);
    *omega = e.omega;
    for (int i = 0; i < 5; i++) {
        grad[i] = e.grad[i];
        for (int j = 0; j < 5; j++) hess[i * 5 + j] = e.hess[i][j];
    }
}

/* vacuum_integral and its M-derivatives (for the closed-form tests). */
void oracle_vacuum_integral(double M, double* out3)
{
    ensure_params();
    vacuum_I(M, &out3[0], &out3[1], &out3[2]);
}

/* ------------------------------------------------------------------------- */
/* thermodynamics of a candidate  (Thermodynamics.jl:215-281)                */
/* ------------------------------------------------------------------------- */
typedef struct {
    double x[5];
    double omega, pressure, rho_norm, entropy, energy;
    double M[3];
    double nq[3], nqb[3];
    int phys;
} cand_t;

static int isfin(double v) { return isfinite(v); }

static void postprocess(const eval_t* e, const double* x, const double mu_vec[3], double T, double phi_tol, cand_t* c)
{
    for (int i = 0; i < 5; i++) c->x[i] = x[i];
    c->omega = e->omega;
    c->pressure = -e->omega;
    double rho[3], rsum = 0.0, Sg = 0.0, S = 0.0;
    for (int i = 0; i < 3; i++) {
        rho[i] = 2.0 * (e->fl[i].Sga - e->fl[i].Sgb);
        rsum += rho[i];
        Sg += e->fl[i].Sgab;
        S += e->fl[i].S;
        c->M[i] = e->M[i];
        c->nq[i] = 2.0 * P.N_color * e->fl[i].Nq;
        c->nqb[i] = 2.0 * P.N_color * e->fl[i].Nqb;
    }
    c->rho_norm = rsum / (3.0 * P.rho0);
    /* s = dP/dT = -dU/dT - dLth/dT, dLth/dT = -2 S + 2 sum C (a g_a + b g_b) */
    c->entropy = -e->UT + 2.0 * S - 2.0 * Sg;
    c->energy = -c->pressure + (mu_vec[0] * rho[0] + mu_vec[1] * rho[1] + mu_vec[2] * rho[2]) + T * c->entropy;
    /* _default_is_physical_solution  ImplicitSolver.jl:50-60 + _all_finite_thermo :62-64 */
    int phys = 1;
    double Phi = x[3], Phib = x[4];
    if (!(isfin(Phi) && isfin(Phib) && (-phi_tol <= Phi && Phi <= 1 + phi_tol) && (-phi_tol <= Phib && Phib <= 1 + phi_tol)))
        phys = 0;
    for (int i = 0; i < 3; i++)
        if (!isfin(c->M[i]) || c->M[i] <= 0.0) phys = 0;
    if (!(isfin(c->omega) && isfin(c->pressure) && isfin(c->rho_norm) && isfin(c->entropy) && isfin(c->energy))) phys = 0;
    c->phys = phys;
}

/* ------------------------------------------------------------------------- */
/* NLsolve-like drivers                                                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    double x[5];
    double F[5];          /* residual at x (last accepted point for TR)       */
    double resnorm;
    int iterations;
    int f_converged, x_converged;
    int exception;        /* check_isfinite failed on the seed                */
    eval_t e;             /* evaluation at the returned x                     */
} nl_t;

static double norm_inf(const double* v, int n)
{
    double m = 0.0;
    for (int i = 0; i < n; i++) {
        double a = fabs(v[i]);
        if (a != a) return a;   /* NaN propagates like Julia's norm(x, Inf) */
        if (a > m) m = a;
    }
    return m;
}

static int any_nan(const double* v, int n)
{
    for (int i = 0; i < n; i++)
        if (v[i] != v[i]) return 1;
    return 0;
}

/* LU with partial pivoting (LAPACK getrf semantics: largest |pivot|, first
 * on ties).  Returns 0 on success, 1 on an exactly zero pivot (Julia's
 * SingularException). */
static int lu_solve(int n, const double* A_in, const double* b_in, double* x)
{
    double A[25], b[5];
    memcpy(A, A_in, sizeof(double) * n * n);
    memcpy(b, b_in, sizeof(double) * n);
    int piv[5];
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(A[i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (A[p * n + k] == 0.0) return 1;
        if (p != k) {
            for (int j = 0; j < n; j++) {
                double t = A[k * n + j];
                A[k * n + j] = A[p * n + j];
                A[p * n + j] = t;
            }
        }
        for (int i = k + 1; i < n; i++) {
            double l = A[i * n + k] / A[k * n + k];
            A[i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
        }
    }
    for (int k = 0; k < n; k++) {
        int p = piv[k];
        if (p != k) {
            double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= A[i * n + j] * b[j];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[i * n + j] * b[j];
        b[i] = s / A[i * n + i];
    }
    memcpy(x, b, sizeof(double) * n);
    return 0;
}

/* F = grad_x P = -grad Omega;  J = dF/dx = -Hessian(Omega) */
static void residual_jac(const eval_t* e, double F[5], double J[25])
{
    for (int i = 0; i < 5; i++) {
        F[i] = -e->grad[i];
        if (J)
            for (int j = 0; j < 5; j++) J[i * 5 + j] = -e->hess[i][j];
    }
}

typedef struct {
    const nodes_t* nd;
    double mu_vec[3];
    double T, xi;
    double xtol, ftol;
    int max_iter;
} prob_t;

static __thread int g_evals;  /* Omega evaluations of the current point */

static void evaluate(const prob_t* pb, const double* x, eval_t* e)
{
    g_evals++;
    eval_point(pb->nd, x, pb->mu_vec, pb->T, pb->xi, e);
}

/* NLsolve newton_ with Static line search. max_steps_cap (<0: none) stops
 * the loop early (used by the learned-seed cutoff). */
static void nl_newton(const prob_t* pb, const double* x0, nl_t* r)
{
    memset(r, 0, sizeof(*r));
    double x[5], xold[5], F[5], J[25];
    memcpy(x, x0, sizeof(x));
    for (int i = 0; i < 5; i++) xold[i] = NAN;
    evaluate(pb, x, &r->e);
    residual_jac(&r->e, F, J);
    for (int i = 0; i < 5; i++)
        if (!isfin(F[i])) {
            r->exception = 1;
            memcpy(r->x, x, sizeof(x));
            memcpy(r->F, F, sizeof(F));
            r->resnorm = NAN;
            return;
        }
    int it = 0;
    int f_conv = norm_inf(F, 5) < pb->ftol;
    int x_conv = 0;
    int stopped = any_nan(x, 5) || any_nan(F, 5);
    int converged = f_conv || x_conv;
    while (!stopped && !converged && it < pb->max_iter) {
        it++;
        /* J already at x (computed together with F) */
        double p[5];
        double negF[5];
        for (int i = 0; i < 5; i++) negF[i] = F[i];
        if (lu_solve(5, J, negF, p) == 0) {
            for (int i = 0; i < 5; i++) p[i] = -p[i];
        } else {
            /* Levenberg-modified direction on a singular Jacobian */
            double g[5], A[25];
            for (int i = 0; i < 5; i++) {
                double s = 0;
                for (int k = 0; k < 5; k++) s += J[k * 5 + i] * F[k];
                g[i] = s;
            }
            double n1 = 0.0;
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) {
                    double s = 0;
                    for (int k = 0; k < 5; k++) s += J[k * 5 + i] * J[k * 5 + j];
                    A[i * 5 + j] = s;
                    n1 += fabs(s);
                }
            double lambda = 1e6 * sqrt(5.0 * 2.220446049250313e-16) * n1;
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) A[i * 5 + j] = -(A[i * 5 + j] + (i == j ? lambda : 0.0));
            if (lu_solve(5, A, g, p) != 0)
                for (int i = 0; i < 5; i++) p[i] = NAN;
        }
        memcpy(xold, x, sizeof(x));
        for (int i = 0; i < 5; i++) x[i] = xold[i] + p[i];
        evaluate(pb, x, &r->e);
        residual_jac(&r->e, F, J);
        double cheb = 0.0;
        int xnan = 0;
        for (int i = 0; i < 5; i++) {
            if (xold[i] != xold[i]) xnan = 1;
            double d = fabs(x[i] - xold[i]);
            if (d > cheb || d != d) cheb = d;
        }
        x_conv = (!xnan && cheb < pb->xtol);
        f_conv = norm_inf(F, 5) < pb->ftol;
        stopped = any_nan(x, 5) || any_nan(F, 5);
        converged = x_conv || f_conv;
    }
    memcpy(r->x, x, sizeof(x));
    memcpy(r->F, F, sizeof(F));
    r->resnorm = norm_inf(F, 5);
    r->iterations = it;
    r->f_converged = f_conv;
    r->x_converged = x_conv;
}

static double wnorm(const double* d, const double* v)
{
    double s = 0.0;
    for (int i = 0; i < 5; i++) s += (d[i] * v[i]) * (d[i] * v[i]);
    return sqrt(s);
}

static double sumsq(const double* v)
{
    double s = 0.0;
    for (int i = 0; i < 5; i++) s += v[i] * v[i];
    return s;
}

/* NLsolve dogleg! */
static void dogleg(double* p, const double* r, const double* d, const double* J, double delta)
{
    double pi_[5], negr[5];
    for (int i = 0; i < 5; i++) negr[i] = r[i];
    if (lu_solve(5, J, negr, pi_) == 0) {
        for (int i = 0; i < 5; i++) pi_[i] = -pi_[i];
    } else {
        /* singular Jacobian: the reference takes a pseudo-inverse step; the
         * restatement falls back to the scaled steepest-descent branch. */
        for (int i = 0; i < 5; i++) pi_[i] = INFINITY;
    }
    if (wnorm(d, pi_) <= delta) {
        memcpy(p, pi_, sizeof(double) * 5);
        return;
    }
    double g[5];
    for (int j = 0; j < 5; j++) {
        double s = 0;
        for (int k = 0; k < 5; k++) s += J[k * 5 + j] * r[k];
        g[j] = s / (d[j] * d[j]);
    }
    double Jg[5];
    for (int i = 0; i < 5; i++) {
        double s = 0;
        for (int k = 0; k < 5; k++) s += J[i * 5 + k] * g[k];
        Jg[i] = s;
    }
    double wg = wnorm(d, g);
    double coef = -(wg * wg) / sumsq(Jg);
    double pc[5];
    for (int i = 0; i < 5; i++) pc[i] = coef * g[i];
    if (wnorm(d, pc) >= delta) {
        double sc = -delta / wg;
        for (int i = 0; i < 5; i++) p[i] = g[i] * sc;
        return;
    }
    double pdiff[5];
    for (int i = 0; i < 5; i++) pdiff[i] = pi_[i] - pc[i];
    double b = 0.0;
    for (int i = 0; i < 5; i++) b += (d[i] * pc[i]) * (d[i] * pdiff[i]);
    b *= 2.0;
    double a = wnorm(d, pdiff);
    a = a * a;
    double wpc = wnorm(d, pc);
    double tau = (-b + sqrt(b * b - 4.0 * a * (wpc * wpc - delta * delta))) / (2.0 * a);
    for (int i = 0; i < 5; i++) p[i] = pc[i] + tau * pdiff[i];
}

static double colnorm(const double* J, int j)
{
    double s = 0.0;
    for (int i = 0; i < 5; i++) s += J[i * 5 + j] * J[i * 5 + j];
    return sqrt(s);
}

/* NLsolve trust_region_ (autoscale=true, factor=1.0) */
static void nl_trust_region(const prob_t* pb, const double* x0, nl_t* r)
{
    memset(r, 0, sizeof(*r));
    double x[5], xold[5], rr[5], J[25], Fn[25];
    memcpy(x, x0, sizeof(x));
    for (int i = 0; i < 5; i++) xold[i] = NAN;
    eval_t e;
    evaluate(pb, x, &e);
    residual_jac(&e, rr, J);
    r->e = e;
    for (int i = 0; i < 5; i++)
        if (!isfin(rr[i])) {
            r->exception = 1;
            memcpy(r->x, x, sizeof(x));
            r->resnorm = NAN;
            return;
        }
    int it = 0;
    int f_conv = norm_inf(rr, 5) < pb->ftol, x_conv = 0;
    double Fcur[5];
    memcpy(Fcur, rr, sizeof(rr));
    int stopped = any_nan(x, 5) || any_nan(Fcur, 5);
    int converged = f_conv;
    if (!converged) {
        double d[5];
        for (int j = 0; j < 5; j++) {
            d[j] = colnorm(J, j);
            if (d[j] == 0.0) d[j] = 1.0;
        }
        double delta = 1.0 * wnorm(d, x);
        if (delta == 0.0) delta = 1.0;
        const double eta = 1e-4;
        while (!stopped && !converged && it < pb->max_iter) {
            it++;
            double p[5];
            dogleg(p, rr, d, J, delta);
            memcpy(xold, x, sizeof(x));
            for (int i = 0; i < 5; i++) x[i] += p[i];
            eval_t en;
            evaluate(pb, x, &en);
            residual_jac(&en, Fcur, Fn);
            double rpred[5];
            for (int i = 0; i < 5; i++) {
                double s = 0;
                for (int k = 0; k < 5; k++) s += J[i * 5 + k] * p[k];
                rpred[i] = s + rr[i];
            }
            double ssr = sumsq(rr);
            double rho = (ssr - sumsq(Fcur)) / (ssr - sumsq(rpred));
            if (rho > eta) {
                memcpy(rr, Fcur, sizeof(rr));
                memcpy(J, Fn, sizeof(J));
                r->e = en;
                for (int j = 0; j < 5; j++) {
                    double cn = colnorm(J, j);
                    double t = 0.1 * d[j];
                    d[j] = t > cn ? t : cn;
                }
                double cheb = 0.0;
                int xnan = 0;
                for (int i = 0; i < 5; i++) {
                    if (xold[i] != xold[i]) xnan = 1;
                    double dd = fabs(x[i] - xold[i]);
                    if (dd > cheb || dd != dd) cheb = dd;
                }
                x_conv = (!xnan && cheb < pb->xtol);
                f_conv = norm_inf(rr, 5) < pb->ftol;
                converged = x_conv || f_conv;
            } else {
                for (int i = 0; i < 5; i++) x[i] -= p[i];
                x_conv = 0;
                converged = 0;
            }
            if (rho < 0.1) {
                delta = delta / 2.0;
            } else if (rho >= 0.9) {
                delta = 2.0 * wnorm(d, p);
            } else if (rho >= 0.5) {
                double t = 2.0 * wnorm(d, p);
                delta = delta > t ? delta : t;
            }
            stopped = any_nan(x, 5) || any_nan(Fcur, 5);
        }
    }
    memcpy(r->x, x, sizeof(x));
    memcpy(r->F, rr, sizeof(rr));
    /* maximum(abs, r): NaN-propagating like norm_inf */
    r->resnorm = norm_inf(rr, 5);
    r->iterations = it;
    r->f_converged = f_conv;
    r->x_converged = x_conv;
    /* the candidate is post-processed at the returned x */
    evaluate(pb, x, &r->e);
}

/* ------------------------------------------------------------------------- */
/* one solve() call with a single seed  (ImplicitSolver.jl:260-304)          */
/* ------------------------------------------------------------------------- */
typedef struct {
    cand_t c;
    double resnorm;
    int iterations;
    int f_converged;
    int converged;
    int exception;       /* primary nlsolve threw */
    int tr_used;
    int steps;           /* Newton + TR steps spent */
} single_t;

static int cand_good(const nl_t* r, const cand_t* c, double resmax)
{
    return r->f_converged && isfin(r->resnorm) && r->resnorm <= resmax && c->phys;
}

static void solve_single(const prob_t* pb, const double* x0, const pnjl_solve_opts* o, single_t* out)
{
    memset(out, 0, sizeof(*out));
    nl_t prim;
    nl_newton(pb, x0, &prim);
    out->steps = prim.iterations;
    if (prim.exception) {
        out->exception = 1;
        for (int i = 0; i < 5; i++) out->c.x[i] = x0[i];
        out->c.omega = out->c.pressure = out->c.rho_norm = out->c.entropy = out->c.energy = NAN;
        for (int i = 0; i < 3; i++) out->c.M[i] = out->c.nq[i] = out->c.nqb[i] = NAN;
        out->resnorm = NAN;
        return;
    }
    /* the Newton loop leaves e evaluated at its final x */
    cand_t pc;
    postprocess(&prim.e, prim.x, pb->mu_vec, pb->T, o->phi_tol, &pc);
    int need_fb = o->trust_region_fallback &&
                  (!prim.f_converged || !isfin(prim.resnorm) || prim.resnorm > o->residual_norm_max || !pc.phys);
    const nl_t* rsel = &prim;
    const cand_t* csel = &pc;
    nl_t fb;
    cand_t fc;
    if (need_fb) {
        nl_trust_region(pb, x0, &fb);
        out->steps += fb.iterations;
        if (!fb.exception) {
            postprocess(&fb.e, fb.x, pb->mu_vec, pb->T, o->phi_tol, &fc);
            /* _choose_candidate  ImplicitSolver.jl:72-101 */
            int pg = cand_good(&prim, &pc, o->residual_norm_max);
            int fg = cand_good(&fb, &fc, o->residual_norm_max);
            int pick_fb;
            if (fg && !pg) pick_fb = 1;
            else if (pg && !fg) pick_fb = 0;
            else if (fg && pg) {
                if (fc.omega < pc.omega) pick_fb = 1;
                else if (fc.omega > pc.omega) pick_fb = 0;
                else pick_fb = fb.resnorm < prim.resnorm;
            } else {
                if (fb.f_converged && !prim.f_converged) pick_fb = 1;
                else if (prim.f_converged && !fb.f_converged) pick_fb = 0;
                else if (isfin(fb.resnorm) && isfin(prim.resnorm)) pick_fb = fb.resnorm < prim.resnorm;
                else pick_fb = 0;
            }
            if (pick_fb) {
                rsel = &fb;
                csel = &fc;
                out->tr_used = 1;
            }
        }
    }
    out->c = *csel;
    out->resnorm = rsel->resnorm;
    out->iterations = rsel->iterations;
    out->f_converged = rsel->f_converged;
    out->converged = rsel->f_converged && csel->phys && isfin(rsel->resnorm) && rsel->resnorm <= o->residual_norm_max;
}

/* ------------------------------------------------------------------------- */
/* seeds  (SeedStrategies.jl)                                                */
/* ------------------------------------------------------------------------- */
static const double HADRON_SEED_5[5] = {-1.84329, -1.84329, -2.22701, 1.0e-5, 4.0e-5};
static const double HIGH_TEMP_SEED_5[5] = {-0.73192, -0.73192, -1.79539, 0.60532, 0.60532};
static const double VERY_HIGH_TEMP_SEED_5[5] = {-0.30, -0.30, -0.90, 0.90, 0.90};
static const double HT_GUESS_0p8_SEED_5[5] = {-0.50, -0.50, -1.20, 0.80, 0.80};
static const double HT_GUESS_0p9_SEED_5[5] = {-0.30, -0.30, -0.90, 0.90, 0.90};
static const double HT_GUESS_0p95_SEED_5[5] = {-0.20, -0.20, -0.70, 0.95, 0.95};
static const double WEAK_CHIRAL_CONF_SEED_5[5] = {-0.50, -0.50, -1.20, 1e-3, 1e-3};

/* hint: 0 auto, 1 hadron, 2 quark.  get_seed(::DefaultSeed)  :204-225 */
static void default_seed(int hint, double T, double mu, double* x0)
{
    if (hint == 0) {
        double hb = 197.327;
        hint = (T * hb > 150 || mu * hb > 300) ? 2 : 1;  /* auto_phase_hint :131-136 */
    }
    const double* s;
    if (hint == 2)
        s = (T * 197.327 >= 300.0) ? VERY_HIGH_TEMP_SEED_5 : HIGH_TEMP_SEED_5;
    else
        s = HADRON_SEED_5;
    memcpy(x0, s, sizeof(double) * 5);
}

/* get_all_seeds(MultiSeed())  :257-284 */
static void multi_seeds(double T, double mu, double seeds[6][5])
{
    default_seed(1, T, mu, seeds[0]);
    default_seed(2, T, mu, seeds[1]);
    memcpy(seeds[2], WEAK_CHIRAL_CONF_SEED_5, sizeof(double) * 5);
    memcpy(seeds[3], HT_GUESS_0p8_SEED_5, sizeof(double) * 5);
    memcpy(seeds[4], HT_GUESS_0p9_SEED_5, sizeof(double) * 5);
    memcpy(seeds[5], HT_GUESS_0p95_SEED_5, sizeof(double) * 5);
}

/* point result (what the batch writes) */
typedef struct {
    cand_t c;
    double resnorm;
    int iterations, converged, flags, steps;
} point_t;

static void from_single(const single_t* s, point_t* p)
{
    p->c = s->c;
    p->resnorm = s->resnorm;
    p->iterations = s->iterations;
    p->converged = s->converged;
    p->flags = (s->tr_used ? PNJL_PT_TR_USED : 0) | (s->exception ? PNJL_PT_EXCEPTION : 0);
}

/* solve_multi  ImplicitSolver.jl:532-559 (selector: argmin Omega, first on ties).
 * Returns 1 if some seed converged, 0 otherwise (the reference throws). */
static int solve_multi(const prob_t* pb, const pnjl_solve_opts* o, point_t* out, int* steps)
{
    double seeds[6][5];
    multi_seeds(pb->T, pb->mu_vec[0], seeds);
    int have = 0;
    point_t best;
    memset(&best, 0, sizeof(best));
    for (int k = 0; k < 6; k++) {
        single_t s;
        solve_single(pb, seeds[k], o, &s);
        *steps += s.steps;
        if (s.exception) continue;
        if (!s.converged) continue;
        if (!have || s.c.omega < best.c.omega) {
            from_single(&s, &best);
            have = 1;
        }
    }
    if (have) {
        *out = best;
        out->flags |= PNJL_PT_MULTISEED;
    }
    return have;
}

/* solve(FixedMu(), T, mu; seed_strategy=<single seed>)  ImplicitSolver.jl:260-328 */
static void solve_with_seed(const prob_t* pb, const double* x0, const pnjl_solve_opts* o, point_t* out)
{
    single_t s;
    solve_single(pb, x0, o, &s);
    int steps = s.steps;
    from_single(&s, out);
    if (s.exception) {
        /* the primary nlsolve throws out of solve() */
        out->steps = steps;
        return;
    }
    if (!s.converged && o->auto_multiseed_fallback) {
        point_t m;
        if (solve_multi(pb, o, &m, &steps)) *out = m;
    }
    out->steps = steps;
}

/* ------------------------------------------------------------------------- */
/* learned-seed policies.  No counterpart in the reference: they compose its   */
/* pieces -- a Newton run (nl_newton, NLsolve :newton) from the given seed     */
/* capped at surrogate_max_steps, solve(DefaultSeed) (ImplicitSolver.jl:260-  */
/* 328) as the fallback, and solve_multi's argmin-Omega choice (ImplicitSolver */
/* .jl:532-559): over the learned root and the MultiSeed candidates inside    */
/* the ambiguity mask (SURROGATE), over the learned, HADRON and HIGH_TEMP     */
/* seeds (MULTIROOT3).                                                         */
/* ------------------------------------------------------------------------- */
static struct {
    int dim[3];
    double lo[3], hi[3];
    uint8_t* mask;
} g_judge;

int oracle_set_judge_mask(int nT, int nmu, int nxi, const double* lo, const double* hi, const uint8_t* mask)
{
    free(g_judge.mask);
    g_judge.mask = NULL;
    if (!mask) return 0;
    if (nT <= 0 || nmu <= 0 || nxi <= 0) return PNJL_ERR_ARG;
    size_t nb = (size_t)nT * nmu * nxi;
    g_judge.mask = (uint8_t*)malloc(nb);
    memcpy(g_judge.mask, mask, nb);
    g_judge.dim[0] = nT;
    g_judge.dim[1] = nmu;
    g_judge.dim[2] = nxi;
    for (int k = 0; k < 3; k++) {
        g_judge.lo[k] = lo[k];
        g_judge.hi[k] = hi[k];
    }
    return 0;
}

/* bins of the mask over [lo, hi]; a point outside the box is always judged */
static int judge_region(double T, double mu, double xi)
{
    const double v[3] = {T, mu, xi};
    int64_t idx = 0;
    for (int k = 0; k < 3; k++) {
        double u = (v[k] - g_judge.lo[k]) / (g_judge.hi[k] - g_judge.lo[k]);
        if (!(u >= 0.0 && u <= 1.0)) return 1;
        int b = (int)(u * g_judge.dim[k]);
        if (b >= g_judge.dim[k]) b = g_judge.dim[k] - 1;
        idx = idx * g_judge.dim[k] + b;
    }
    return g_judge.mask[idx] != 0;
}

/* DualBranchScan._is_same_solution  DualBranchScan.jl:409-419 */
static int same_root(const cand_t* a, const cand_t* b)
{
    return fabs(a->x[0] - b->x[0]) < 0.01 && fabs(a->x[3] - b->x[3]) < 0.01 &&
           fabs(a->M[0] - b->M[0]) * 197.327 < 0.01 * 197.327;
}

/* A learned-seed run that lands farther from its seed than DualBranchScan's
 * jump test allows a continuation seed (|d phi_u| > 0.5 or |d M_u| > 50 MeV,
 * DualBranchScan.jl:410-415) has left the seed's basin: its root is not
 * trusted and the plain policy takes the point. */
static int learned_jump(const double* seed, const cand_t* c)
{
    double M0[3];
    mass_vec(seed, M0, NULL, NULL);
    return fabs(c->x[0] - seed[0]) > 0.5 || fabs(c->M[0] - M0[0]) * 197.327 > 50.0;
}

static void solve_surrogate(const prob_t* pb, const double* seed, const pnjl_solve_opts* o, point_t* out)
{
    const int in_mask = g_judge.mask && judge_region(pb->T, pb->mu_vec[0], pb->xi);
    if (in_mask && o->surrogate_judge == PNJL_JUDGE_REFERENCE) {
        /* ambiguity region: solve(DefaultSeed) takes the point, learned seed unused */
        double x0[5];
        default_seed(0, pb->T, pb->mu_vec[0], x0);
        solve_with_seed(pb, x0, o, out);
        out->flags |= PNJL_PT_SURR_MASKED;
        return;
    }
    /* 1. Newton from the learned seed, at most surrogate_max_steps steps, no fallback */
    prob_t capped = *pb;
    if (o->surrogate_max_steps < pb->max_iter) capped.max_iter = o->surrogate_max_steps;
    pnjl_solve_opts oc = *o;
    oc.trust_region_fallback = 0;
    single_t s;
    solve_single(&capped, seed, &oc, &s);
    int steps = s.steps, judged;
    if (!s.exception && s.converged && !learned_jump(seed, &s.c)) {
        from_single(&s, out);
        judged = 1;
    } else {
        /* 2. the learned seed is abandoned: solve(DefaultSeed) takes the point */
        double x0[5];
        default_seed(0, pb->T, pb->mu_vec[0], x0);
        single_t p;
        solve_single(pb, x0, o, &p);
        steps += p.steps;
        from_single(&p, out);
        judged = !p.exception && p.converged;
        if (!p.exception && !p.converged && o->auto_multiseed_fallback) {
            point_t m;
            if (solve_multi(pb, o, &m, &steps)) *out = m;
        }
        out->flags |= PNJL_PT_SURR_FALLBACK;
    }
    /* 3. inside the ambiguity mask: minimum Omega among this root and the
     *    roots of the six MultiSeed candidates (get_all_seeds(MultiSeed())) */
    if (judged && in_mask) {
        double js[6][5];
        multi_seeds(pb->T, pb->mu_vec[0], js);
        point_t r0 = *out, best = *out;
        int distinct = 0, moved = 0;
        single_t jr[6];
        /* HADRON and hint-QUARK first; the other four candidates only where
         * one of them does not converge (its own MultiSeed fallback) */
        solve_single(pb, js[0], o, &jr[0]);
        solve_single(pb, js[1], o, &jr[1]);
        int nk = 2;
        for (int k = 0; k < 2; k++)
            if (jr[k].exception || !jr[k].converged) nk = 6;
        for (int k = 2; k < nk; k++) solve_single(pb, js[k], o, &jr[k]);
        for (int k = 0; k < nk; k++) {
            const single_t j = jr[k];
            steps += j.steps;
            if (j.exception || !j.converged) continue;
            if (!same_root(&j.c, &r0.c)) distinct = 1;
            if (j.c.omega < best.c.omega) {
                from_single(&j, &best);
                best.flags |= r0.flags & PNJL_PT_SURR_FALLBACK;
                moved = 1;
            }
        }
        *out = best;
        if (distinct) out->flags |= PNJL_PT_FIRST_ORDER;
        if (moved && !same_root(&best.c, &r0.c)) out->flags |= PNJL_PT_SURR_OMEGA;
    }
    out->steps = steps;
}

static void solve_multiroot3(const prob_t* pb, const double* seed, const pnjl_solve_opts* o, point_t* out)
{
    const double* sd[3] = {seed, HADRON_SEED_5, HIGH_TEMP_SEED_5};
    single_t r[3];
    int steps = 0, b = -1;
    for (int k = 0; k < 3; k++) {
        solve_single(pb, sd[k], o, &r[k]);
        steps += r[k].steps;
        if (!r[k].exception && r[k].converged && (b < 0 || r[k].c.omega < r[b].c.omega)) b = k;
    }
    if (b < 0) {
        from_single(&r[0], out);
        out->converged = 0;
        out->flags |= PNJL_PT_NO_ROOT;
    } else {
        from_single(&r[b], out);
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++)
                if (!r[i].exception && r[i].converged && !r[j].exception && r[j].converged &&
                    !same_root(&r[i].c, &r[j].c))
                    out->flags |= PNJL_PT_FIRST_ORDER;
        if (r[0].exception || !r[0].converged) out->flags |= PNJL_PT_SURR_FALLBACK;
        else if (b != 0 && !same_root(&r[0].c, &r[b].c)) out->flags |= PNJL_PT_SURR_OMEGA;
    }
    out->steps = steps;
}

static void solve_policy_point(const prob_t* pb, int policy, const double* seed, const pnjl_solve_opts* o,
                               point_t* out)
{
    double x0[5];
    memset(out, 0, sizeof(*out));
    switch (policy) {
    case PNJL_SEED_DEFAULT:
    case PNJL_SEED_HADRON:
    case PNJL_SEED_QUARK:
        default_seed(policy == PNJL_SEED_DEFAULT ? 0 : (policy == PNJL_SEED_HADRON ? 1 : 2), pb->T, pb->mu_vec[0], x0);
        solve_with_seed(pb, x0, o, out);
        break;
    case PNJL_SEED_EXPLICIT:
        memcpy(x0, seed, sizeof(x0));
        solve_with_seed(pb, x0, o, out);
        break;
    case PNJL_SEED_MULTI: {
        int steps = 0;
        if (!solve_multi(pb, o, out, &steps)) {
            /* reference: ErrorException; report the first seed's attempt */
            double seeds[6][5];
            multi_seeds(pb->T, pb->mu_vec[0], seeds);
            single_t s;
            solve_single(pb, seeds[0], o, &s);
            from_single(&s, out);
            out->flags |= PNJL_PT_NO_ROOT;
        }
        out->steps = steps;
        break;
    }
    case PNJL_SEED_SURROGATE:
        solve_surrogate(pb, seed, o, out);
        break;
    case PNJL_SEED_MULTIROOT3:
        solve_multiroot3(pb, seed, o, out);
        break;
    default:
        memset(out, 0, sizeof(*out));
        out->flags = PNJL_PT_EXCEPTION;
        break;
    }
}

/* ------------------------------------------------------------------------- */
/* phase boundary (PhaseAwareContinuitySeed)  SeedStrategies.jl:388-475, 762-888 */
/* ------------------------------------------------------------------------- */
#define MAX_BND_SETS 16
#define MAX_BND_ROWS 256
typedef struct {
    double xi;
    int n;
    double T[MAX_BND_ROWS], mu[MAX_BND_ROWS];
    double T_CEP, mu_CEP;
} boundary_t;

static boundary_t BND[MAX_BND_SETS];
static int BND_N = 0;

/* Installs the phase-boundary tables: rows (xi, T_MeV, mu_MeV) of boundary.csv
 * and rows (xi, T_CEP, mu_CEP) of cep.csv.  load_phase_boundary (:388-436)
 * keeps, for a requested xi, the rows with |xi_row - xi| <= 1e-6 sorted by T,
 * and the first matching CEP row. */
void oracle_set_boundary(int nrows, const double* xi_r, const double* T_r, const double* mu_r, int ncep,
                         const double* xi_c, const double* T_c, const double* mu_c)
{
    BND_N = 0;
    for (int i = 0; i < nrows; i++) {
        int s = -1;
        for (int k = 0; k < BND_N; k++)
            if (BND[k].xi == xi_r[i]) s = k;
        if (s < 0) {
            if (BND_N == MAX_BND_SETS) continue;
            s = BND_N++;
            BND[s].xi = xi_r[i];
            BND[s].n = 0;
            BND[s].T_CEP = NAN;
            BND[s].mu_CEP = NAN;
        }
        if (BND[s].n < MAX_BND_ROWS) {
            BND[s].T[BND[s].n] = T_r[i];
            BND[s].mu[BND[s].n] = mu_r[i];
            BND[s].n++;
        }
    }
    for (int k = 0; k < BND_N; k++) {
        /* insertion sort by T (stable, like sortperm) */
        for (int i = 1; i < BND[k].n; i++) {
            double t = BND[k].T[i], m = BND[k].mu[i];
            int j = i - 1;
            while (j >= 0 && BND[k].T[j] > t) {
                BND[k].T[j + 1] = BND[k].T[j];
                BND[k].mu[j + 1] = BND[k].mu[j];
                j--;
            }
            BND[k].T[j + 1] = t;
            BND[k].mu[j + 1] = m;
        }
        for (int c = 0; c < ncep; c++)
            if (fabs(xi_c[c] - BND[k].xi) <= 1e-6) {
                BND[k].T_CEP = T_c[c];
                BND[k].mu_CEP = mu_c[c];
                break;
            }
    }
}

static const boundary_t* find_boundary(double xi)
{
    for (int k = 0; k < BND_N; k++)
        if (fabs(BND[k].xi - xi) <= 1e-6) return &BND[k];
    return NULL;
}

/* interpolate_mu_c  :446-475 */
static double interp_mu_c(const boundary_t* B, double T)
{
    if (!isnan(B->T_CEP) && T > B->T_CEP) return NAN;
    if (B->n == 0) return NAN;
    if (T <= B->T[0]) return B->mu[0];
    if (T >= B->T[B->n - 1]) return B->mu[B->n - 1];
    for (int i = 0; i < B->n - 1; i++) {
        if (B->T[i] <= T && T <= B->T[i + 1]) {
            double t = (T - B->T[i]) / (B->T[i + 1] - B->T[i]);
            return B->mu[i] + t * (B->mu[i + 1] - B->mu[i]);
        }
    }
    return NAN;
}

/* phases: 0 unknown, 1 hadron, 2 quark, 3 crossover.  _get_current_phase :762-782
 * (an xi without boundary rows behaves as the reference's empty
 * PhaseBoundaryData: T_CEP = NaN, no rows -> unknown). */
static int current_phase(double xi, double T_MeV, double mu_MeV)
{
    const boundary_t* B = find_boundary(xi);
    if (!B) return 0;
    if (!isnan(B->T_CEP) && T_MeV > B->T_CEP) return 3;
    double muc = interp_mu_c(B, T_MeV);
    if (isnan(muc)) return 0;
    return mu_MeV < muc ? 1 : 2;
}

double oracle_interp_mu_c(double xi, double T_MeV)
{
    const boundary_t* B = find_boundary(xi);
    return B ? interp_mu_c(B, T_MeV) : NAN;
}

int oracle_current_phase(double xi, double T_MeV, double mu_MeV) { return current_phase(xi, T_MeV, mu_MeV); }

/* ------------------------------------------------------------------------- */
/* batch entry points                                                        */
/* ------------------------------------------------------------------------- */
#define MAX_NODE_CACHE 16
static nodes_t NODE_CACHE[MAX_NODE_CACHE];
static int NODE_CACHE_N = 0;

static nodes_t* cache_nodes(int p_num, int t_num)
{
    for (int i = 0; i < NODE_CACHE_N; i++)
        if (NODE_CACHE[i].p_num == p_num && NODE_CACHE[i].t_num == t_num) return &NODE_CACHE[i];
    if (NODE_CACHE_N == MAX_NODE_CACHE) {
        nodes_free(&NODE_CACHE[0]);
        memmove(&NODE_CACHE[0], &NODE_CACHE[1], sizeof(nodes_t) * (MAX_NODE_CACHE - 1));
        NODE_CACHE_N--;
    }
    nodes_init(&NODE_CACHE[NODE_CACHE_N], p_num, t_num);
    return &NODE_CACHE[NODE_CACHE_N++];
}

static void write_point(const point_t* r, int64_t i, const pnjl_solve_out* out)
{
    for (int k = 0; k < 5; k++) out->x[i * 5 + k] = r->c.x[k];
    out->omega[i] = r->c.omega;
    out->pressure[i] = r->c.pressure;
    out->rho_norm[i] = r->c.rho_norm;
    out->entropy[i] = r->c.entropy;
    out->energy[i] = r->c.energy;
    for (int k = 0; k < 3; k++) {
        out->masses[i * 3 + k] = r->c.M[k];
        if (out->n_q) out->n_q[i * 3 + k] = r->c.nq[k];
        if (out->n_qbar) out->n_qbar[i * 3 + k] = r->c.nqb[k];
    }
    out->residual_norm[i] = r->resnorm;
    out->iterations[i] = r->iterations;
    out->converged[i] = r->converged;
    if (out->flags) out->flags[i] = r->flags;
    if (out->newton_steps) out->newton_steps[i] = r->steps;
    if (out->passes) out->passes[i] = 3 * g_evals;  /* three flavours per evaluation */
    g_evals = 0;
}

static void make_prob(prob_t* pb, const nodes_t* nd, double T, double mu, double xi, const pnjl_solve_opts* o)
{
    pb->nd = nd;
    pb->mu_vec[0] = pb->mu_vec[1] = pb->mu_vec[2] = mu;
    pb->T = T;
    pb->xi = xi;
    pb->xtol = o->xtol;
    pb->ftol = o->ftol;
    pb->max_iter = o->max_iter;
}

/* One chain of consecutive points: ContinuitySeed (policy CONTINUITY, as in
 * tests/perf/pnjl/scan_perf.jl:58-75: update! only on convergence) or the
 * driver's PhaseAwareContinuitySeed chain (policy PHASE_AWARE,
 * run_gap_transport_scan.jl:410-442: first point MultiSeed, then the tracker). */
static void solve_chain(const nodes_t* nd, int64_t i0, int64_t len, const double* T, const double* mu,
                        const double* xi, int policy, const pnjl_solve_opts* o, const pnjl_solve_out* out)
{
    double prev[5];
    int have_prev = 0;
    int prev_phase = 0;
    const double hb = 197.327;
    for (int64_t i = i0; i < i0 + len; i++) {
        prob_t pb;
        make_prob(&pb, nd, T[i], mu[i], xi[i], o);
        point_t r;
        memset(&r, 0, sizeof(r));
        g_evals = 0;
        if (policy == PNJL_SEED_CONTINUITY) {
            double x0[5];
            if (have_prev) memcpy(x0, prev, sizeof(x0));
            else default_seed(0, T[i], mu[i], x0);
            solve_with_seed(&pb, x0, o, &r);
            if (r.converged && !(r.flags & PNJL_PT_EXCEPTION)) {
                memcpy(prev, r.c.x, sizeof(prev));
                have_prev = 1;
            }
        } else {
            double TM = T[i] * hb, muM = mu[i] * hb;
            if (!have_prev) {
                int steps = 0;
                if (!solve_multi(&pb, o, &r, &steps)) {
                    double seeds[6][5];
                    multi_seeds(pb.T, pb.mu_vec[0], seeds);
                    single_t s;
                    solve_single(&pb, seeds[0], o, &s);
                    from_single(&s, &r);
                    r.flags |= PNJL_PT_NO_ROOT;
                    r.converged = 0;
                }
                r.steps = steps;
            } else {
                int cur = current_phase(xi[i], TM, muM);
                double x0[5];
                int reset = ((prev_phase == 1 && cur == 2) || (prev_phase == 2 && cur == 1));
                if (reset)
                    memcpy(x0, cur == 1 ? HADRON_SEED_5 : HIGH_TEMP_SEED_5, sizeof(x0));
                else
                    memcpy(x0, prev, sizeof(x0));
                solve_with_seed(&pb, x0, o, &r);
                if (reset) r.flags |= PNJL_PT_CHAIN_RESET;
            }
            if (r.converged && !(r.flags & (PNJL_PT_EXCEPTION | PNJL_PT_NO_ROOT))) {
                memcpy(prev, r.c.x, sizeof(prev));
                have_prev = 1;
                prev_phase = current_phase(xi[i], TM, muM);
            }
        }
        write_point(&r, i, out);
    }
}

/* Batch solve with the same argument meaning as pnjl_solve_fixedmu_batch
 * (host pointers).  nthreads > 1 -> OpenMP over points (or chains). */
int oracle_solve_batch(int64_t n, const double* T, const double* mu, const double* xi, const double* seed,
                       int policy, int p_num, int t_num, const pnjl_solve_opts* o, const pnjl_solve_out* out,
                       int nthreads)
{
    ensure_params();
    const nodes_t* nd = cache_nodes(p_num, t_num);
    if (policy == PNJL_SEED_CONTINUITY || policy == PNJL_SEED_PHASE_AWARE) {
        int64_t L = o->chain_len > 0 ? o->chain_len : n;
        int64_t nchains = (n + L - 1) / L;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t c = 0; c < nchains; c++) {
            int64_t i0 = c * L;
            int64_t len = (i0 + L <= n) ? L : n - i0;
            solve_chain(nd, i0, len, T, mu, xi, policy, o, out);
        }
        return 0;
    }
    if (policy != PNJL_SEED_DEFAULT && policy != PNJL_SEED_HADRON && policy != PNJL_SEED_QUARK &&
        policy != PNJL_SEED_MULTI && policy != PNJL_SEED_EXPLICIT && policy != PNJL_SEED_SURROGATE &&
        policy != PNJL_SEED_MULTIROOT3)
        return PNJL_ERR_ARG;
    if ((policy == PNJL_SEED_EXPLICIT || policy == PNJL_SEED_SURROGATE || policy == PNJL_SEED_MULTIROOT3) && !seed)
        return PNJL_ERR_ARG;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < n; i++) {
        prob_t pb;
        make_prob(&pb, nd, T[i], mu[i], xi[i], o);
        point_t r;
        g_evals = 0;
        solve_policy_point(&pb, policy, seed ? seed + i * 5 : NULL, o, &r);
        write_point(&r, i, out);
    }
    return 0;
}

/* Newton from an explicit seed, capped at `cap` steps, no fallbacks: returns
 * the state after the cap (used to pin the learned-seed cutoff semantics). */
int oracle_newton_capped(double T, double mu, double xi, const double* x0, int p_num, int t_num,
                         const pnjl_solve_opts* o, int cap, double* x_out, double* resnorm, int* iters,
                         int* f_conv)
{
    ensure_params();
    prob_t pb;
    make_prob(&pb, cache_nodes(p_num, t_num), T, mu, xi, o);
    pb.max_iter = cap;
    nl_t r;
    nl_newton(&pb, x0, &r);
    memcpy(x_out, r.x, sizeof(double) * 5);
    *resnorm = r.resnorm;
    *iters = r.iterations;
    *f_conv = r.f_converged;
    return r.exception ? 1 : 0;
}

/* Thermodynamics at a given state (calculate_thermo + number densities). */
void oracle_thermo(const double* x, double T, double mu, double xi, int p_num, int t_num, double* out12)
{
    ensure_params();
    prob_t pb;
    pnjl_solve_opts o;
    memset(&o, 0, sizeof(o));
    o.phi_tol = 1e-8;
    make_prob(&pb, cache_nodes(p_num, t_num), T, mu, xi, &o);
    eval_t e;
    evaluate(&pb, x, &e);
    cand_t c;
    postprocess(&e, x, pb.mu_vec, T, 1e-8, &c);
    out12[0] = c.omega;
    out12[1] = c.pressure;
    out12[2] = c.rho_norm;
    out12[3] = c.entropy;
    out12[4] = c.energy;
    for (int i = 0; i < 3; i++) {
        out12[5 + i] = c.nq[i];
        out12[8 + i] = c.nqb[i];
    }
    out12[11] = c.phys;
}

/* ------------------------------------------------------------------------- */
/* implicit thermodynamic derivatives                                        */
/* ThermoDerivatives.jl:132-245 (mass_derivatives order 1, thermo_derivatives)*/
/* and :342-466 (bulk_viscosity_coefficients): at a root x of F(x; T, mu) =  */
/* grad_x P = 0, dx/dtheta = -(dF/dx)^-1 dF/dtheta (the implicit function    */
/* theorem that ImplicitDifferentiation.jl applies), then chain rules.  The  */
/* reference obtains the partial derivatives by forward-mode AD; here they   */
/* are closed forms of the same node terms.                                  */
/* ------------------------------------------------------------------------- */
/* Per-flavour derivatives of S = sum C Q(E; mu, T, Phi, Phib), Q = lnF+(a) +
 * lnF-(b), a = (mu - E)/T, b = -(E + mu)/T, E = E(M): the mixed ones with
 * theta = (mu, T) for one common mu. */
typedef struct {
    double Smu, ST, Smumu, SmuT, STT, SMmu, SMT, SPmu, SPbmu, SPT, SPbT;
} flav_theta_t;

static void flavour_theta_sums(const nodes_t* nd, double M, double mu, double T, double Phi, double Phib,
                               double xi, flav_theta_t* o)
{
    memset(o, 0, sizeof(*o));
    double invT = 1.0 / T, invT2 = invT * invT;
    for (int idx = 0; idx < nd->n; idx++) {
        double p = nd->p[idx], t = nd->c[idx], C = nd->C[idx];
        double pt = p * t;
        double E = sqrt(p * p + M * M + xi * (pt * pt));
        double a = -(E - mu) * invT;
        double b = -(E + mu) * invT;
        lnF_t qp, qm;
        lnF_eval(a, Phi, Phib, &qp);
        lnF_eval(b, Phib, Phi, &qm);
        double EM = M / E;
        /* da/dmu = 1/T, db/dmu = -1/T, da/dT = -a/T, db/dT = -b/T, da/dE = db/dE = -1/T */
        o->Smu += C * (qp.a - qm.a) * invT;
        o->ST += C * (-(qp.a * a + qm.a * b) * invT);
        o->Smumu += C * (qp.aa + qm.aa) * invT2;
        o->SmuT += C * (-(qp.a - qm.a) - (qp.aa * a - qm.aa * b)) * invT2;
        o->STT += C * (2.0 * (qp.a * a + qm.a * b) + qp.aa * a * a + qm.aa * b * b) * invT2;
        o->SMmu += C * (-(qp.aa - qm.aa) * invT2) * EM;
        o->SMT += C * ((qp.a + qm.a + qp.aa * a + qm.aa * b) * invT2) * EM;
        o->SPmu += C * (qp.ac1 - qm.ac2) * invT;
        o->SPbmu += C * (qp.ac2 - qm.ac1) * invT;
        o->SPT += C * (-(qp.ac1 * a + qm.ac2 * b) * invT);
        o->SPbT += C * (-(qp.ac2 * a + qm.ac1 * b) * invT);
    }
}

/* d/dT and d^2/dT^2 of U at fixed (Phi, Phib), and d^2 U / dPhi dT, dPhib dT
 * (Thermodynamics.jl:124-131: U = T^4 [-a(T)/2 Phi Phib + b(T) log(...)],
 * a = a0 + a1 T0/T + a2 (T0/T)^2, b = b3 (T0/T)^3, log argument floored). */
static void polyakov_U_T(double T, double Phi, double Phib, double* UT, double* UTT, double* UPT, double* UPbT)
{
    double val = 1.0 - 6.0 * Phib * Phi + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * (Phib * Phi) * (Phib * Phi);
    int floored = (val <= 0.0) || (val < POLYAKOV_EPS);
    double lg = floored ? log(POLYAKOV_EPS) : log(val);
    double lP = 0.0, lPb = 0.0;
    if (!floored) {
        lP = (-6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib) / val;
        lPb = (-6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib) / val;
    }
    /* T^4 a(T) = a0 T^4 + a1 T0 T^3 + a2 T0^2 T^2;  T^4 b(T) = b3 T0^3 T */
    double T0 = P.T0;
    double A1 = 4.0 * P.a0 * T * T * T + 3.0 * P.a1 * T0 * T * T + 2.0 * P.a2 * T0 * T0 * T;
    double A2 = 12.0 * P.a0 * T * T + 6.0 * P.a1 * T0 * T + 2.0 * P.a2 * T0 * T0;
    double B1 = P.b3 * T0 * T0 * T0;
    *UT = -0.5 * Phi * Phib * A1 + B1 * lg;
    *UTT = -0.5 * Phi * Phib * A2;
    *UPT = -0.5 * Phib * A1 + B1 * lP;
    *UPbT = -0.5 * Phi * A1 + B1 * lPb;
}

/* The reference's mass function of ThermoDerivatives.jl:111-121 (bare masses
 * 5.5 and 140.0 MeV written out there). */
static void deriv_masses(const double* x, double M[3])
{
    double mu0 = 0.0055 / 0.197327, ms0 = 0.140 / 0.197327;
    M[0] = mu0 - 4.0 * P.G * x[0] + 2.0 * P.K * x[1] * x[2];
    M[1] = mu0 - 4.0 * P.G * x[1] + 2.0 * P.K * x[0] * x[2];
    M[2] = ms0 - 4.0 * P.G * x[2] + 2.0 * P.K * x[0] * x[1];
}

/* out[36]: pressure, energy, rho (= sum rho_i / 3), rho_norm, entropy,
 * dP_dT, dP_dmu, dEpsilon_dT, dEpsilon_dmu, dn_dT, dn_dmu, dP_depsilon_n,
 * dP_dn_epsilon, masses[3], dM_dT[3], dM_dmu[3], v_n_sq, dmuB_dT_sigma,
 * dx_dT[5], dx_dmu[5], ds_dT, ds_dmu.  Returns 1 if dF/dx is singular. */
int oracle_thermo_derivs(const double* x, double T, double mu, double xi, int p_num, int t_num, double* out)
{
    ensure_params();
    const nodes_t* nd = cache_nodes(p_num, t_num);
    double muv[3] = {mu, mu, mu};
    eval_t e;
    eval_point(nd, x, muv, T, xi, &e);
    double M[3], dM[3][3], d2M[3][3][3];
    mass_vec(x, M, dM, d2M);
    flav_theta_t th[3];
    for (int i = 0; i < 3; i++) flavour_theta_sums(nd, M[i], mu, T, x[3], x[4], xi, &th[i]);
    double UT, UTT, UPT, UPbT;
    polyakov_U_T(T, x[3], x[4], &UT, &UTT, &UPT, &UPbT);

    /* derivatives of Omega = chi + U + Evac - 2 T sum_i S_i in (mu, T) */
    double Om_mu = 0, Om_T = UT, Om_mumu = 0, Om_muT = 0, Om_TT = UTT;
    double Om_xmu[5] = {0}, Om_xT[5] = {0};
    for (int i = 0; i < 3; i++) {
        const flav_t* f = &e.fl[i];
        const flav_theta_t* g = &th[i];
        Om_mu += -2.0 * T * g->Smu;
        Om_T += -2.0 * f->S - 2.0 * T * g->ST;
        Om_mumu += -2.0 * T * g->Smumu;
        Om_muT += -2.0 * g->Smu - 2.0 * T * g->SmuT;
        Om_TT += -4.0 * g->ST - 2.0 * T * g->STT;
        double LMmu = -2.0 * T * g->SMmu;
        double LMT = -2.0 * f->SM - 2.0 * T * g->SMT;
        for (int a = 0; a < 3; a++) {
            Om_xmu[a] += dM[i][a] * LMmu;
            Om_xT[a] += dM[i][a] * LMT;
        }
        Om_xmu[3] += -2.0 * T * g->SPmu;
        Om_xmu[4] += -2.0 * T * g->SPbmu;
        Om_xT[3] += -2.0 * f->SP - 2.0 * T * g->SPT;
        Om_xT[4] += -2.0 * f->SPb - 2.0 * T * g->SPbT;
    }
    Om_xT[3] += UPT;
    Om_xT[4] += UPbT;

    /* dx/dtheta = -(d2 Omega/dx2)^-1 d2 Omega/dx dtheta */
    double H[25], rhs[5], dxdT[5], dxdmu[5];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) H[i * 5 + j] = e.hess[i][j];
    for (int i = 0; i < 5; i++) rhs[i] = -Om_xT[i];
    int sing = lu_solve(5, H, rhs, dxdT);
    for (int i = 0; i < 5; i++) rhs[i] = -Om_xmu[i];
    sing |= lu_solve(5, H, rhs, dxdmu);
    if (sing) {
        for (int k = 0; k < 36; k++) out[k] = NAN;
        return 1;
    }
    double Pr = -e.omega;
    double s = -Om_T;              /* dP/dT at fixed x          */
    double ntot = -Om_mu;          /* dP/dmu = sum_i rho_i      */
    double ds_dT = -Om_TT, ds_dmu = -Om_muT, dnt_dT = -Om_muT, dnt_dmu = -Om_mumu;
    for (int a = 0; a < 5; a++) {
        ds_dT += -Om_xT[a] * dxdT[a];
        ds_dmu += -Om_xT[a] * dxdmu[a];
        dnt_dT += -Om_xmu[a] * dxdT[a];
        dnt_dmu += -Om_xmu[a] * dxdmu[a];
    }
    double n = ntot / 3.0, dn_dT = dnt_dT / 3.0, dn_dmu = dnt_dmu / 3.0;
    double eps = -Pr + T * s + mu * ntot;
    double E_T = T * ds_dT + mu * dnt_dT, E_mu = T * ds_dmu + mu * dnt_dmu;
    double P_T = s, P_mu = ntot;
    double den_e = E_T * dn_dmu - E_mu * dn_dT;
    double den_n = dn_dT * E_mu - dn_dmu * E_T;
    out[0] = Pr;
    out[1] = eps;
    out[2] = n;
    out[3] = ntot / (3.0 * P.rho0);
    out[4] = s;
    out[5] = P_T;
    out[6] = P_mu;
    out[7] = E_T;
    out[8] = E_mu;
    out[9] = dn_dT;
    out[10] = dn_dmu;
    out[11] = den_e == 0.0 ? NAN : (P_T * dn_dmu - P_mu * dn_dT) / den_e;
    out[12] = den_n == 0.0 ? NAN : (P_T * E_mu - P_mu * E_T) / den_n;
    double Mr[3];
    deriv_masses(x, Mr);
    for (int i = 0; i < 3; i++) {
        double gT = 0.0, gmu = 0.0;
        for (int a = 0; a < 3; a++) {
            gT += dM[i][a] * dxdT[a];
            gmu += dM[i][a] * dxdmu[a];
        }
        out[13 + i] = Mr[i];
        out[16 + i] = gT;
        out[19 + i] = gmu;
    }
    /* bulk_viscosity_coefficients: mu_B = 3 mu_q, n_B = n */
    double ds_dmuB = ds_dmu / 3.0, dn_dmuB = dn_dmu / 3.0;
    out[22] = (s * dn_dmuB - n * dn_dT) / (T * (ds_dT * dn_dmuB - ds_dmuB * dn_dT));
    out[23] = -(n * ds_dT - s * dn_dT) / (n * ds_dmuB - s * dn_dmuB);
    for (int a = 0; a < 5; a++) {
        out[24 + a] = dxdT[a];
        out[29 + a] = dxdmu[a];
    }
    out[34] = ds_dT;
    out[35] = ds_dmu;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* second-order implicit derivatives                                         */
/* solve_with_derivatives(order=2)  src/pnjl/solver/ImplicitSolver.jl:660-701 */
/* mass_derivatives(order=2)  src/pnjl/derivatives/ThermoDerivatives.jl:153-172*/
/* Along the root x(theta), theta = (T, mu), F(x(theta), theta) = 0 with      */
/* F = grad_x Omega.  Differentiating H x_a + F_a = 0 (H = dF/dx, F_a =       */
/* dF/dtheta_a) once more along w_b = (x_b, e_b):                             */
/*   x_ab = -H^-1 [ (D_wb H) x_a + D_wb F_a ].                                */
/* The directional derivatives D_w of the closed-form H and F_a come from the */
/* complex-step method (Squire and Trapp, SIAM Review 40(1), 1998): the same  */
/* closed forms evaluated in complex arithmetic at z + i h w give D_w f =     */
/* Im f / h with no subtraction (h = 1e-20), so the result is exact to        */
/* rounding like the reference's nested forward-mode AD.  Branches (the       */
/* log-argument floor, the sign of a, |M|) follow the real part.              */
/* ------------------------------------------------------------------------- */
#include <complex.h>

typedef double complex cplx;

typedef struct {
    cplx a, c1, c2, aa, ac1, ac2, c1c1, c1c2, c2c2;
} lnF_c;

static void lnF_eval_c(cplx a, cplx c1, cplx c2, lnF_c* o)
{
    cplx D, r1, r2, r3;
    memset(o, 0, sizeof(*o));
    if (creal(a) > 0) {
        cplx e_a_m = cexp(-2.0 * a), e_2a_m = cexp(-a), e_neg_m = cexp(-3.0 * a);
        D = e_neg_m + 3.0 * c1 * e_a_m + 3.0 * c2 * e_2a_m + 1.0;
        if (creal(D) < POLYAKOV_EPS) {
            o->a = 3.0;
            return;
        }
        r1 = e_a_m / D;
        r2 = e_2a_m / D;
        r3 = 1.0 / D;
    } else {
        cplx ea = cexp(a), e2 = ea * ea, e3 = ea * e2;
        D = 1.0 + 3.0 * c1 * ea + 3.0 * c2 * e2 + e3;
        if (creal(D) < POLYAKOV_EPS) return;
        r1 = ea / D;
        r2 = e2 / D;
        r3 = e3 / D;
    }
    cplx f = c1 * r1 + 2.0 * c2 * r2 + r3;
    o->a = 3.0 * f;
    o->c1 = 3.0 * r1;
    o->c2 = 3.0 * r2;
    o->aa = 3.0 * (c1 * r1 + 4.0 * c2 * r2 + 3.0 * r3) - 9.0 * f * f;
    o->ac1 = 3.0 * r1 * (1.0 - 3.0 * f);
    o->ac2 = 3.0 * r2 * (2.0 - 3.0 * f);
    o->c1c1 = -9.0 * r1 * r1;
    o->c1c2 = -9.0 * r1 * r2;
    o->c2c2 = -9.0 * r2 * r2;
}

/* node sums of one flavour that H and dF/dtheta need */
typedef struct {
    cplx SM, SP, SPb, SMM, SMP, SMPb, SPP, SPPb, SPbPb, SMmu, SMT, SPmu, SPbmu, SPT, SPbT;
} flav_c;

static void flavour_sums_c(const nodes_t* nd, cplx M, cplx mu, cplx T, cplx Phi, cplx Phib, double xi, flav_c* o)
{
    memset(o, 0, sizeof(*o));
    cplx invT = 1.0 / T, invT2 = invT * invT;
    for (int idx = 0; idx < nd->n; idx++) {
        double p = nd->p[idx], t = nd->c[idx], C = nd->C[idx];
        double pt = p * t;
        cplx E = csqrt(p * p + M * M + xi * (pt * pt));
        cplx a = -(E - mu) * invT, b = -(E + mu) * invT;
        lnF_c qp, qm;
        lnF_eval_c(a, Phi, Phib, &qp);
        lnF_eval_c(b, Phib, Phi, &qm);
        cplx QE = -(qp.a + qm.a) * invT, QEE = (qp.aa + qm.aa) * invT2;
        cplx QEP = -(qp.ac1 + qm.ac2) * invT, QEPb = -(qp.ac2 + qm.ac1) * invT;
        cplx EM = M / E, EMM = (1.0 - EM * EM) / E;
        o->SM += C * (QE * EM);
        o->SP += C * (qp.c1 + qm.c2);
        o->SPb += C * (qp.c2 + qm.c1);
        o->SMM += C * (QEE * EM * EM + QE * EMM);
        o->SMP += C * (QEP * EM);
        o->SMPb += C * (QEPb * EM);
        o->SPP += C * (qp.c1c1 + qm.c2c2);
        o->SPPb += C * (qp.c1c2 + qm.c1c2);
        o->SPbPb += C * (qp.c2c2 + qm.c1c1);
        o->SMmu += C * (-(qp.aa - qm.aa) * invT2) * EM;
        o->SMT += C * ((qp.a + qm.a + qp.aa * a + qm.aa * b) * invT2) * EM;
        o->SPmu += C * (qp.ac1 - qm.ac2) * invT;
        o->SPbmu += C * (qp.ac2 - qm.ac1) * invT;
        o->SPT += C * (-(qp.ac1 * a + qm.ac2 * b) * invT);
        o->SPbT += C * (-(qp.ac2 * a + qm.ac1 * b) * invT);
    }
}

/* dI/dM and d2I/dM2 of the vacuum term (vacuum_I) */
static void vacuum_I_c(cplx M, cplx* dI, cplx* d2I)
{
    double L = P.Lambda;
    double sgn = (creal(M) < 0) ? -1.0 : 1.0;
    cplx Mt = sgn * M + 1e-12;
    cplx s = csqrt(L * L + Mt * Mt);
    cplx lg = clog((L + s) / Mt);
    *dI = sgn * Mt / (4.0 * PI_D * PI_D) * (L * s - Mt * Mt * lg);
    *d2I = (2.0 * L * Mt * Mt / s + L * s - 3.0 * Mt * Mt * lg) / (4.0 * PI_D * PI_D);
}

/* H = d2 Omega/dx2 and F_T, F_mu = d2 Omega/dx dT, d2 Omega/dx dmu at complex
 * (x, T, mu): the formulas of eval_point, polyakov_U, polyakov_U_T and
 * oracle_thermo_derivs. */
static void hess_theta_c(const nodes_t* nd, const cplx* x, cplx T, cplx mu, double xi, cplx H[5][5], cplx FT[5],
                         cplx Fmu[5])
{
    double G = P.G, K = P.K;
    cplx pu = x[0], pd = x[1], ps = x[2], Phi = x[3], Phib = x[4];
    cplx M[3] = {P.m_ud0 - 4.0 * G * pu + 2.0 * K * pd * ps, P.m_ud0 - 4.0 * G * pd + 2.0 * K * pu * ps,
                 P.m_s0 - 4.0 * G * ps + 2.0 * K * pu * pd};
    cplx dM[3][3] = {{-4.0 * G, 2.0 * K * ps, 2.0 * K * pd},
                     {2.0 * K * ps, -4.0 * G, 2.0 * K * pu},
                     {2.0 * K * pd, 2.0 * K * pu, -4.0 * G}};
    cplx d2chi[3][3] = {{4.0 * G, -4.0 * K * ps, -4.0 * K * pd},
                        {-4.0 * K * ps, 4.0 * G, -4.0 * K * pu},
                        {-4.0 * K * pd, -4.0 * K * pu, 4.0 * G}};
    flav_c f[3];
    cplx gM[3], hM[3];
    for (int i = 0; i < 3; i++) {
        flavour_sums_c(nd, M[i], mu, T, Phi, Phib, xi, &f[i]);
        cplx dI, d2I;
        vacuum_I_c(M[i], &dI, &d2I);
        gM[i] = -2.0 * P.N_color * dI - 2.0 * T * f[i].SM;
        hM[i] = -2.0 * P.N_color * d2I - 2.0 * T * f[i].SMM;
    }
    for (int a = 0; a < 3; a++) {
        for (int b = 0; b < 3; b++) {
            cplx h = d2chi[a][b];
            for (int i = 0; i < 3; i++) {
                double d2 = (a != b && a != i && b != i) ? 2.0 * K : 0.0;   /* d2M_i / dphi_a dphi_b */
                h += hM[i] * dM[i][a] * dM[i][b] + gM[i] * d2;
            }
            H[a][b] = h;
        }
        cplx hP = 0.0, hPb = 0.0;
        for (int i = 0; i < 3; i++) {
            hP += -2.0 * T * f[i].SMP * dM[i][a];
            hPb += -2.0 * T * f[i].SMPb * dM[i][a];
        }
        H[a][3] = H[3][a] = hP;
        H[a][4] = H[4][a] = hPb;
    }
    /* Polyakov potential: second derivatives in the loops and mixed ones with T */
    cplx Tr = P.T0 / T;
    cplx Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr, Tb = P.b3 * Tr * Tr * Tr;
    cplx val = 1.0 - 6.0 * Phib * Phi + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * (Phib * Phi) * (Phib * Phi);
    int floored = (creal(val) <= 0.0) || (creal(val) < POLYAKOV_EPS);
    cplx lP = 0, lPb = 0, lPP = 0, lPPb = 0, lPbPb = 0;
    if (!floored) {
        lP = (-6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib) / val;
        lPb = (-6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib) / val;
        lPP = (24.0 * Phi - 6.0 * Phib * Phib) / val - lP * lP;
        lPPb = (-6.0 - 12.0 * Phi * Phib) / val - lP * lPb;
        lPbPb = (24.0 * Phib - 6.0 * Phi * Phi) / val - lPb * lPb;
    }
    cplx T4 = T * T * T * T;
    cplx SPP = 0, SPPb = 0, SPbPb = 0;
    for (int i = 0; i < 3; i++) {
        SPP += f[i].SPP;
        SPPb += f[i].SPPb;
        SPbPb += f[i].SPbPb;
    }
    H[3][3] = T4 * (Tb * lPP) - 2.0 * T * SPP;
    H[3][4] = H[4][3] = T4 * (-0.5 * Ta + Tb * lPPb) - 2.0 * T * SPPb;
    H[4][4] = T4 * (Tb * lPbPb) - 2.0 * T * SPbPb;
    double T0 = P.T0;
    cplx A1 = 4.0 * P.a0 * T * T * T + 3.0 * P.a1 * T0 * T * T + 2.0 * P.a2 * T0 * T0 * T;
    double B1 = P.b3 * T0 * T0 * T0;
    for (int a = 0; a < 5; a++) FT[a] = Fmu[a] = 0.0;
    for (int i = 0; i < 3; i++) {
        cplx LMmu = -2.0 * T * f[i].SMmu, LMT = -2.0 * f[i].SM - 2.0 * T * f[i].SMT;
        for (int a = 0; a < 3; a++) {
            Fmu[a] += dM[i][a] * LMmu;
            FT[a] += dM[i][a] * LMT;
        }
        Fmu[3] += -2.0 * T * f[i].SPmu;
        Fmu[4] += -2.0 * T * f[i].SPbmu;
        FT[3] += -2.0 * f[i].SP - 2.0 * T * f[i].SPT;
        FT[4] += -2.0 * f[i].SPb - 2.0 * T * f[i].SPbT;
    }
    FT[3] += -0.5 * Phib * A1 + B1 * lP;
    FT[4] += -0.5 * Phi * A1 + B1 * lPb;
}

/* out[PNJL_IMPLICIT2_FIELDS]: x[5], dx_dT[5], dx_dmu[5], d2x_dT2[5],
 * d2x_dTdmu[5], d2x_dmu2[5], masses[3], dM_dT[3], dM_dmu[3], d2M_dT2[3],
 * d2M_dTdmu[3], d2M_dmu2[3] (masses with the bare 5.5 and 140.0 MeV of
 * ThermoDerivatives.jl:111-121).  Returns 1 if dF/dx is singular (all NaN). */
int oracle_implicit_derivs2(const double* x, double T, double mu, double xi, int p_num, int t_num, double* out)
{
    ensure_params();
    const nodes_t* nd = cache_nodes(p_num, t_num);
    double d1[36];
    if (oracle_thermo_derivs(x, T, mu, xi, p_num, t_num, d1)) {
        for (int k = 0; k < PNJL_IMPLICIT2_FIELDS; k++) out[k] = NAN;
        return 1;
    }
    const double* xT = d1 + 24;
    const double* xmu = d1 + 29;
    const double h = 1e-20;
    /* D_w of H and F_theta along w_T = (x_T, 1, 0) and w_mu = (x_mu, 0, 1) */
    double Hr[25], dH[2][5][5], dFT[2][5], dFmu[2][5];
    for (int dir = 0; dir < 2; dir++) {
        const double* w = dir == 0 ? xT : xmu;
        cplx xc[5], H[5][5], FT[5], Fmu[5];
        for (int a = 0; a < 5; a++) xc[a] = x[a] + I * (h * w[a]);
        cplx Tc = T + (dir == 0 ? I * h : 0.0), muc = mu + (dir == 1 ? I * h : 0.0);
        hess_theta_c(nd, xc, Tc, muc, xi, H, FT, Fmu);
        for (int a = 0; a < 5; a++) {
            for (int b = 0; b < 5; b++) {
                dH[dir][a][b] = cimag(H[a][b]) / h;
                if (dir == 0) Hr[a * 5 + b] = creal(H[a][b]);
            }
            dFT[dir][a] = cimag(FT[a]) / h;
            dFmu[dir][a] = cimag(Fmu[a]) / h;
        }
    }
    /* x_TT (w_T, a = T), x_Tmu (w_mu, a = T), x_mumu (w_mu, a = mu) */
    double x2[3][5];
    for (int k = 0; k < 3; k++) {
        const int dir = (k == 0) ? 0 : 1;
        const double* xa = (k == 2) ? xmu : xT;
        const double* Fa = (k == 2) ? dFmu[dir] : dFT[dir];
        double rhs[5];
        for (int a = 0; a < 5; a++) {
            double r = Fa[a];
            for (int b = 0; b < 5; b++) r += dH[dir][a][b] * xa[b];
            rhs[a] = -r;
        }
        if (lu_solve(5, Hr, rhs, x2[k])) {
            for (int q = 0; q < PNJL_IMPLICIT2_FIELDS; q++) out[q] = NAN;
            return 1;
        }
    }
    for (int a = 0; a < 5; a++) {
        out[a] = x[a];
        out[5 + a] = xT[a];
        out[10 + a] = xmu[a];
        out[15 + a] = x2[0][a];
        out[20 + a] = x2[1][a];
        out[25 + a] = x2[2][a];
    }
    /* masses: M_i = m_i0 - 4 G phi_i + 2 K phi_j phi_k (j, k the other two) */
    double Mr[3];
    deriv_masses(x, Mr);
    const double G = P.G, K = P.K;
    for (int i = 0; i < 3; i++) {
        const int j = (i == 0) ? 1 : 0, k = (i == 2) ? 1 : 2;
        double m1[2], m2[3];
        for (int d = 0; d < 2; d++) {
            const double* v = d == 0 ? xT : xmu;
            m1[d] = -4.0 * G * v[i] + 2.0 * K * (v[j] * x[k] + x[j] * v[k]);
        }
        for (int q = 0; q < 3; q++) {
            const double* va = (q == 2) ? xmu : xT;   /* TT: (T, T), Tmu: (T, mu), mumu: (mu, mu) */
            const double* vb = (q == 0) ? xT : xmu;
            m2[q] = -4.0 * G * x2[q][i] + 2.0 * K * (x2[q][j] * x[k] + x[j] * x2[q][k] + va[j] * vb[k] + vb[j] * va[k]);
        }
        out[30 + i] = Mr[i];
        out[33 + i] = m1[0];
        out[36 + i] = m1[1];
        out[39 + i] = m2[0];
        out[42 + i] = m2[1];
        out[45 + i] = m2[2];
    }
    return 0;
}
