/* pnjl_polar_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's anisotropic one-loop
 * polarization path, the checker of the HIP kernels in
 * julia_relaxtime_amd/csrc/pnjl_polar.hip.  Only tests/, smoke() and the
 * bench's CPU-baseline leg may load it; the product never does.
 *
 * Follows, function by function (file:line cited at each):
 *   src/QuarkDistribution.jl            PNJL f+/f-, antiderivatives
 *   src/QuarkDistribution_Aniso.jl      df/dE, first-order xi coefficient
 *   src/integration/IntervalQuadratureStrategies.jl   power/DE rules
 *   src/relaxtime/OneLoopIntegrals.jl   B~0, B0, A
 *   src/relaxtime/OneLoopIntegralsAniso.jl   delta B~0, B0_correction, A_aniso
 *   src/relaxtime/PolarizationAniso.jl  polarization_aniso / _with_width
 * Same segmentation, same nodes, same skip-non-finite rule as the
 * reference, in IEEE arithmetic (the reference's @fastmath last bits are not
 * reproduced; see the parity notes of the tests).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pnjl_hip.h"

void oracle_gl_std(int n, double* x, double* w);
void oracle_default_params(pnjl_model_params* o);

#define PI_D 3.14159265358979323846
#define EPS_K 1.0e-9        /* OneLoopIntegrals.jl:23 */
#define EPS_SEGMENT 1.0e-12 /* OneLoopIntegrals.jl:24 */
#define PV_GAP_REL 1e-6     /* OneLoopIntegrals.jl:35 */
#define POWER_ALPHA 0.35    /* IntervalQuadratureStrategies.jl:35 */
#define DE_H 0.15           /* IntervalQuadratureStrategies.jl:36 */

static double g_Lambda = 0.0; /* fm^-1 */
static double g_Nc = 3.0;

enum { SING_NONE = 0, SING_LEFT = 1, SING_RIGHT = 2, SING_BOTH = 3 };
enum { PLUS = 0, MINUS = 1 };       /* B0 signs (:plus, :minus) */
enum { QUARK = 0, ANTIQUARK = 1 };  /* aniso signs (:quark, :antiquark) */

/* quadrature tables  IntervalQuadratureStrategies.jl:56-85 */
static struct {
    int ready;
    double x16[16], w16[16], x32[32], w32[32];
    double pl_t16[16], pl_tp16[16], pl_t32[32], pl_tp32[32];
    double pr_t16[16], pr_tp16[16], pr_t32[32], pr_tp32[32];
    double de_phi[33], de_phip[33];
} Q;

static void power_tables(const double* x, int n, double* lt, double* ltp, double* rt, double* rtp)
{
    const double ia = 1.0 / POWER_ALPHA;
    for (int i = 0; i < n; i++) {
        double u = (x[i] + 1) / 2;
        lt[i] = pow(u, ia);
        ltp[i] = ia * pow(u, ia - 1) / 2;
        rt[i] = 1 - pow(1 - u, ia);
        rtp[i] = ia * pow(1 - u, ia - 1) / 2;
    }
}

static void ensure_tables(void)
{
    if (Q.ready) return;
    pnjl_model_params P;
    oracle_default_params(&P);
    g_Lambda = P.Lambda;
    g_Nc = (double)P.N_color;
    oracle_gl_std(16, Q.x16, Q.w16);
    oracle_gl_std(32, Q.x32, Q.w32);
    power_tables(Q.x16, 16, Q.pl_t16, Q.pl_tp16, Q.pr_t16, Q.pr_tp16);
    power_tables(Q.x32, 32, Q.pl_t32, Q.pl_tp32, Q.pr_t32, Q.pr_tp32);
    for (int k = -16; k <= 16; k++) {
        double t = k * DE_H;
        double s = PI_D / 2 * sinh(t);
        double sech = 1.0 / cosh(s);
        Q.de_phi[k + 16] = tanh(s);
        Q.de_phip[k + 16] = DE_H * (PI_D / 2) * cosh(t) * (sech * sech);
    }
    Q.ready = 1;
}

void oracle_polar_tables(double* out)
{
    /* x16 w16 x32 w32 | pl_t16 pl_tp16 pl_t32 pl_tp32 | pr_* | de_phi de_phip */
    ensure_tables();
    double* o = out;
    memcpy(o, Q.x16, sizeof Q.x16); o += 16;
    memcpy(o, Q.w16, sizeof Q.w16); o += 16;
    memcpy(o, Q.x32, sizeof Q.x32); o += 32;
    memcpy(o, Q.w32, sizeof Q.w32); o += 32;
    memcpy(o, Q.pl_t16, sizeof Q.pl_t16); o += 16;
    memcpy(o, Q.pl_tp16, sizeof Q.pl_tp16); o += 16;
    memcpy(o, Q.pl_t32, sizeof Q.pl_t32); o += 32;
    memcpy(o, Q.pl_tp32, sizeof Q.pl_tp32); o += 32;
    memcpy(o, Q.pr_t16, sizeof Q.pr_t16); o += 16;
    memcpy(o, Q.pr_tp16, sizeof Q.pr_tp16); o += 16;
    memcpy(o, Q.pr_t32, sizeof Q.pr_t32); o += 32;
    memcpy(o, Q.pr_tp32, sizeof Q.pr_tp32); o += 32;
    memcpy(o, Q.de_phi, sizeof Q.de_phi); o += 33;
    memcpy(o, Q.de_phip, sizeof Q.de_phip);
}

/* Julia eps(x): spacing of doubles at |x| */
static double jeps(double x)
{
    x = fabs(x);
    if (!(x >= DBL_MIN)) return 4.9406564584124654e-324;
    return ldexp(1.0, ilogb(x) - 52);
}

static double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

static double clamp_e(double e) { return e < 1e-200 ? 1e-200 : (e > 1e200 ? 1e200 : e); }

/* ---------------------------------------------------------------------- */
/* distributions  QuarkDistribution.jl:14-99                               */
/* ---------------------------------------------------------------------- */
static double quark_f(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1 / T;
    double e = clamp_e(exp(-(E - mu) * beta));
    double e2 = e * e, e3 = e2 * e;
    return (Phi * e + 2 * Phib * e2 + e3) / (1 + 3 * Phi * e + 3 * Phib * e2 + e3);
}

static double antiquark_f(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1 / T;
    double e = clamp_e(exp(-(E + mu) * beta));
    double e2 = e * e, e3 = e2 * e;
    return (Phib * e + 2 * Phi * e2 + e3) / (1 + 3 * Phib * e + 3 * Phi * e2 + e3);
}

static double safe_log300(double x) { return log(x <= 1e-300 ? 1e-300 : x); }

static double quark_F(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1 / T;
    double e = clamp_e(exp(-(E - mu) * beta));
    double e2 = e * e, e3 = e2 * e;
    return -T * safe_log300(1 + 3 * Phi * e + 3 * Phib * e2 + e3) / 3;
}

static double antiquark_F(double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1 / T;
    double e = clamp_e(exp(-(E + mu) * beta));
    double e2 = e * e, e3 = e2 * e;
    return -T * safe_log300(1 + 3 * Phib * e + 3 * Phi * e2 + e3) / 3;
}

/* distribution_value_b0 / distribution_integral_b0  OneLoopIntegrals.jl:68-104 */
static double dist_b0(int sg, double E, double mu, double T, double Phi, double Phib)
{
    return sg == PLUS ? quark_f(E, mu, T, Phi, Phib) : 1.0 - antiquark_f(E, mu, T, Phi, Phib);
}

static double dist_int_b0(int sg, double a, double b, double mu, double T, double Phi, double Phib)
{
    if (sg == PLUS) return quark_F(b, mu, T, Phi, Phib) - quark_F(a, mu, T, Phi, Phib);
    return (b - a) - (antiquark_F(b, mu, T, Phi, Phib) - antiquark_F(a, mu, T, Phi, Phib));
}

/* energy_cutoff / internal_momentum  OneLoopIntegrals.jl:37-48 */
static double e_cut(double m)
{
    double mp = m > 0 ? m : 0.0;
    return sqrt(mp * mp + g_Lambda * g_Lambda);
}

static double p_int(double E, double m)
{
    double mp = m > 0 ? m : 0.0;
    return sqrt(E * E - mp * mp);
}

/* ---------------------------------------------------------------------- */
/* hybrid interval rules  IntervalQuadratureStrategies.jl:194-310          */
/* ---------------------------------------------------------------------- */
typedef double (*integrand_fn)(const void* ctx, double E);

static double rule_power(integrand_fn f, const void* c, double a, double b, const double* t, const double* tp,
                         const double* W, int n)
{
    double len = b - a, total = 0.0;
    for (int i = 0; i < n; i++) {
        double x = a + len * t[i];
        double w = W[i] * len * tp[i];
        double v = f(c, x);
        if (isfinite(v)) total += w * v;
    }
    return total;
}

static double hybrid(integrand_fn f, const void* c, double a, double b, int pos, int n)
{
    if (!(isfinite(a) && isfinite(b)) || b <= a) return 0.0;
    if (pos == SING_BOTH) {
        double half = (b - a) / 2, center = (a + b) / 2, total = 0.0;
        for (int i = 0; i < 33; i++) {
            double x = center + half * Q.de_phi[i];
            double w = half * Q.de_phip[i];
            double v = f(c, x);
            if (isfinite(v)) total += w * v;
        }
        return total;
    }
    if (pos == SING_RIGHT)
        return n == 16 ? rule_power(f, c, a, b, Q.pr_t16, Q.pr_tp16, Q.w16, 16)
                       : rule_power(f, c, a, b, Q.pr_t32, Q.pr_tp32, Q.w32, 32);
    return n == 16 ? rule_power(f, c, a, b, Q.pl_t16, Q.pl_tp16, Q.w16, 16)
                   : rule_power(f, c, a, b, Q.pl_t32, Q.pl_tp32, Q.w32, 32);
}

/* ---------------------------------------------------------------------- */
/* B~0 at k = 0  OneLoopIntegrals.jl:117-224                                */
/* ---------------------------------------------------------------------- */
typedef struct {
    int sg;
    double lam, k, m, mp, dterm, mu, T, Phi, Phib, E0, f0, Emin, Emax;
} b0ctx;

static double b0_real_k0(const void* cv, double E)
{
    const b0ctx* c = (const b0ctx*)cv;
    double p = p_int(E, c->m);
    double dist = dist_b0(c->sg, E, c->mu, c->T, c->Phi, c->Phib);
    double den = c->lam * E + c->dterm;
    if (!isfinite(p) || !isfinite(dist) || !isfinite(den)) return 0.0;
    if (fabs(den) < EPS_SEGMENT) return 0.0;
    double v = p * dist / den;
    return isfinite(v) ? v : 0.0;
}

static double b0_numer(const b0ctx* c, double E)
{
    double p = p_int(E, c->m);
    double dist = dist_b0(c->sg, E, c->mu, c->T, c->Phi, c->Phib);
    if (!isfinite(p) || !isfinite(dist)) return 0.0;
    double v = p * dist;
    return isfinite(v) ? v : 0.0;
}

static double b0_regular(const void* cv, double E)
{
    const b0ctx* c = (const b0ctx*)cv;
    double d = E - c->E0;
    if (fabs(d) < 64 * jeps(c->E0)) {
        double del = fmax(1e-7 * (c->Emax - c->Emin), 64 * jeps(c->E0));
        double El = fmax(c->Emin, c->E0 - del), Er = fmin(c->Emax, c->E0 + del);
        if (Er > El) return (b0_numer(c, Er) - b0_numer(c, El)) / (Er - El);
        return 0.0;
    }
    return (b0_numer(c, E) - c->f0) / d;
}

static void tilde_B0_k0(int sg, double lam, double m, double mp, double mu, double T, double Phi, double Phib,
                        double* re, double* im)
{
    b0ctx c;
    c.sg = sg;
    c.lam = lam;
    c.m = m > 0 ? m : 0.0;
    c.mp = mp > 0 ? mp : 0.0;
    c.mu = mu;
    c.T = T;
    c.Phi = Phi;
    c.Phib = Phib;
    c.Emin = c.m;
    c.Emax = e_cut(c.m);
    c.dterm = (lam * lam + c.m * c.m - c.mp * c.mp) / 2.0;
    /* singularity_k_zero  :144-157 */
    double E0 = -c.dterm / lam;
    int has = 0;
    if (isfinite(E0)) {
        double tol = 1e-8 * fmax(1.0, fabs(E0));
        has = (E0 > c.Emin - tol) && (E0 < c.Emax + tol);
    }
    if (!has) {
        double r = hybrid(b0_real_k0, &c, c.Emin, c.Emax, SING_NONE, 16);
        *re = r * 2.0;
        *im = 0.0 / lam;
        return;
    }
    c.E0 = E0;
    c.f0 = b0_numer(&c, E0);
    double pv;
    double etol = 1e-10 * fmax(1.0, c.Emax - c.Emin);
    if (fabs(E0 - c.Emin) <= etol || fabs(c.Emax - E0) <= etol) {
        double gap = fmax(PV_GAP_REL * (c.Emax - c.Emin), 64 * jeps(E0));
        double left_end = fmax(c.Emin, E0 - gap), right_start = fmin(c.Emax, E0 + gap);
        pv = 0.0;
        if (left_end > c.Emin) pv += hybrid(b0_real_k0, &c, c.Emin, left_end, SING_NONE, 16);
        if (right_start < c.Emax) pv += hybrid(b0_real_k0, &c, right_start, c.Emax, SING_NONE, 16);
    } else {
        double regular = hybrid(b0_regular, &c, c.Emin, c.Emax, SING_NONE, 32);
        double logterm = c.f0 * (log(fabs(c.Emax - E0)) - log(fabs(c.Emin - E0)));
        pv = (regular + logterm) / lam;
    }
    double p0 = p_int(E0, c.m);
    double imag = 2.0 * PI_D * p0 * dist_b0(sg, E0, mu, T, Phi, Phib);
    *re = pv * 2.0;
    *im = imag / lam;
}

/* ---------------------------------------------------------------------- */
/* B~0 at k > 0  OneLoopIntegrals.jl:228-476                                */
/* ---------------------------------------------------------------------- */
static double b0_real_kpos(const void* cv, double E)
{
    const b0ctx* c = (const b0ctx*)cv;
    double p = p_int(E, c->m);
    double dist = dist_b0(c->sg, E, c->mu, c->T, c->Phi, c->Phib);
    double lpE = c->lam + E;
    double common = lpE * lpE - c->mp * c->mp;
    double pmk = p - c->k, ppk = p + c->k;
    double num = common - pmk * pmk;
    double den = common - ppk * ppk;
    if (!isfinite(p) || !isfinite(dist) || !isfinite(num) || !isfinite(den)) return 0.0;
    if (fabs(num) < EPS_SEGMENT && fabs(den) < EPS_SEGMENT) return 0.0;
    if (fabs(den) < EPS_SEGMENT) den = (den == 0.0) ? EPS_SEGMENT : jsign(den) * EPS_SEGMENT;
    double ratio = fabs(num / den);
    if (!isfinite(ratio)) return 0.0;
    ratio = fmax(ratio, 1e-300);
    double lt = log(ratio);
    if (!isfinite(lt)) return 0.0;
    double v = dist * lt;
    return isfinite(v) ? v : 0.0;
}

/* singularity_k_positive  :239-330; returns the number of intervals (0..2) */
static int sing_kpos(double lam, double k, double m, double mp, double Emin, double Emax, double iv[2][2])
{
    int n = 0;
    double a = lam * lam - k * k;
    double d0 = k * k - lam * lam;
    double denom = 2.0 * d0;
    if (fabs(a) < EPS_SEGMENT) {
        if (lam > 0.0 && mp > m) {
            double m2d = mp * mp - m * m;
            double Ec = m2d / (4.0 * lam) + lam * m * m / m2d;
            if (Ec < Emax) {
                iv[n][0] = fmax(Ec, Emin);
                iv[n][1] = Emax;
                n++;
            }
        }
        return n;
    }
    double A = lam * lam - k * k + m * m - mp * mp;
    double disc = A * A + 4.0 * m * m * d0;
    if (disc < 0.0) return 0;
    double sq = sqrt(disc);
    double E1 = (lam * A - k * sq) / denom, E2 = (lam * A + k * sq) / denom;
    if (E1 > E2) {
        double t = E1;
        E1 = E2;
        E2 = t;
    }
    if (E2 - E1 < EPS_SEGMENT) return 0;
    if (a > 0.0) {
        if (E2 > Emin && E1 < Emax) {
            double c1 = fmax(E1, Emin), c2 = fmin(E2, Emax);
            if (c2 > c1) {
                iv[n][0] = c1;
                iv[n][1] = c2;
                n++;
            }
        }
        return n;
    }
    if (E1 > Emin) {
        double c = fmin(E1, Emax);
        if (c > Emin) {
            iv[n][0] = Emin;
            iv[n][1] = c;
            n++;
        }
    }
    if (E2 < Emax) {
        double c = fmax(E2, Emin);
        if (Emax > c) {
            iv[n][0] = c;
            iv[n][1] = Emax;
            n++;
        }
    }
    return n;
}

static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* integrate_piecewise_hybrid  :335-421 */
static double piecewise(integrand_fn f, const void* c, double Emin, double Emax, int niv, double iv[2][2])
{
    if (niv == 0) return hybrid(f, c, Emin, Emax, SING_NONE, 16);
    double E1 = NAN, E2 = NAN;
    if (niv == 1) {
        E1 = iv[0][0];
        E2 = iv[0][1];
    } else {
        double a1 = iv[0][0], b1 = iv[0][1], a2 = iv[1][0], b2 = iv[1][1];
        double tol = 1e-12 * fmax(1.0, Emax - Emin);
        if (fabs(a1 - Emin) <= tol) E1 = b1;
        else if (fabs(a2 - Emin) <= tol) E1 = b2;
        if (fabs(b1 - Emax) <= tol) E2 = a1;
        else if (fabs(b2 - Emax) <= tol) E2 = a2;
    }
    if (isfinite(E1) && isfinite(E2)) {
        double total = 0.0;
        if (E1 > Emin) total += hybrid(f, c, Emin, E1, SING_RIGHT, 32);
        if (E2 > E1) total += hybrid(f, c, E1, E2, SING_BOTH, 32);
        if (Emax > E2) total += hybrid(f, c, E2, Emax, SING_LEFT, 16);
        return total;
    }
    /* general path: sorted unique breakpoints */
    double pts[6];
    int np = 0;
    pts[np++] = Emin;
    for (int i = 0; i < niv; i++) {
        pts[np++] = iv[i][0];
        pts[np++] = iv[i][1];
    }
    pts[np++] = Emax;
    qsort(pts, np, sizeof(double), cmp_d);
    int u = 0;
    for (int i = 0; i < np; i++)
        if (u == 0 || pts[i] != pts[u - 1]) pts[u++] = pts[i];
    int nseg = u - 1;
    if (nseg <= 0) return 0.0;
    if (nseg == 1) return hybrid(f, c, pts[0], pts[1], SING_NONE, 16);
    double total = 0.0;
    for (int i = 1; i <= nseg; i++) {
        double a = pts[i - 1], b = pts[i];
        if (!(isfinite(a) && isfinite(b)) || b <= a) continue;
        int pos = (i == 1) ? SING_RIGHT : (i == nseg ? SING_LEFT : SING_BOTH);
        int nn = (i == nseg) ? 16 : 32;
        total += hybrid(f, c, a, b, pos, nn);
    }
    return total;
}

static void tilde_B0_kpos(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi,
                          double Phib, double* re, double* im)
{
    b0ctx c;
    memset(&c, 0, sizeof c);
    c.sg = sg;
    c.lam = lam;
    c.k = k;
    c.m = m > 0 ? m : 0.0;
    c.mp = mp > 0 ? mp : 0.0;
    c.mu = mu;
    c.T = T;
    c.Phi = Phi;
    c.Phib = Phib;
    double Emin = c.m, Emax = e_cut(c.m);
    double iv[2][2];
    int niv = sing_kpos(lam, k, c.m, c.mp, Emin, Emax, iv);
    double real = piecewise(b0_real_kpos, &c, Emin, Emax, niv, iv);
    double imag = 0.0;
    if (niv > 0) {
        for (int i = 0; i < niv; i++) imag += dist_int_b0(sg, iv[i][0], iv[i][1], mu, T, Phi, Phib);
        imag *= PI_D * jsign(lam);
    }
    *re = real / k;
    *im = imag / k;
}

static void tilde_B0(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi, double Phib,
                     double* re, double* im)
{
    if (fabs(k) < EPS_K) tilde_B0_k0(sg, lam, m, mp, mu, T, Phi, Phib, re, im);
    else tilde_B0_kpos(sg, lam, k, m, mp, mu, T, Phi, Phib, re, im);
}

/* B0  OneLoopIntegrals.jl:489-503 */
void oracle_B0(double lam, double k, double m1, double mu1, double m2, double mu2, double T, double Phi,
               double Phib, double* re, double* im)
{
    ensure_tables();
    double r[4], i[4];
    tilde_B0(PLUS, -lam, k, m1, m2, mu1, T, Phi, Phib, &r[0], &i[0]);
    tilde_B0(MINUS, lam, k, m1, m2, mu1, T, Phi, Phib, &r[1], &i[1]);
    tilde_B0(PLUS, lam, k, m2, m1, mu2, T, Phi, Phib, &r[2], &i[2]);
    tilde_B0(MINUS, -lam, k, m2, m1, mu2, T, Phi, Phib, &r[3], &i[3]);
    *re = r[0] - r[1] + r[2] - r[3];
    *im = i[0] - i[1] + i[2] - i[3];
}

void oracle_tilde_B0(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi, double Phib,
                     double* re, double* im)
{
    ensure_tables();
    tilde_B0(sg, lam, k, m, mp, mu, T, Phi, Phib, re, im);
}

/* const_integral_term_A / A  OneLoopIntegrals.jl:507-543 */
double oracle_A_const(double m)
{
    ensure_tables();
    double mp = m > 0 ? m : 0.0;
    double L = g_Lambda;
    if (mp < 1e-14) return (L * L) / 2.0;
    double s = sqrt(L * L + mp * mp);
    double t1 = L * s;
    double t2 = mp * mp * log((L + s) / mp);
    return (t1 - t2) / 2.0;
}

double oracle_A(double m, double mu, double T, double Phi, double Phib, int n, const double* xp, const double* wp)
{
    double integral = -oracle_A_const(m);
    for (int i = 0; i < n; i++) {
        double p = xp[i], E = sqrt(p * p + m * m);
        integral += wp[i] * (p * p) / E * (quark_f(E, mu, T, Phi, Phib) + antiquark_f(E, mu, T, Phi, Phib));
    }
    return 4.0 * integral;
}

/* ---------------------------------------------------------------------- */
/* first-order anisotropy  QuarkDistribution_Aniso.jl, OneLoopIntegralsAniso.jl */
/* ---------------------------------------------------------------------- */
/* quark_df_dE / antiquark_df_dE  QuarkDistribution_Aniso.jl:14-66 */
static double df_dE(int sg, double E, double mu, double T, double Phi, double Phib)
{
    double beta = 1 / T;
    double c1 = sg == QUARK ? Phi : Phib, c2 = sg == QUARK ? Phib : Phi;
    double arg = sg == QUARK ? -(E - mu) * beta : -(E + mu) * beta;
    double e = clamp_e(exp(arg));
    double e2 = e * e, e3 = e2 * e;
    double num = c1 * e + 2 * c2 * e2 + e3;
    double den = 1 + 3 * c1 * e + 3 * c2 * e2 + e3;
    double dnum = -(c1 * e + 4 * c2 * e2 + 3 * e3);
    double dden = -(3 * c1 * e + 6 * c2 * e2 + 3 * e3);
    double dfdx = dnum * den - num * dden;
    dfdx /= den * den;
    return beta * dfdx;
}

/* correction_cos_theta_coefficient  :71-86 */
static double corr_coef(int sg, double p, double m, double mu, double T, double Phi, double Phib, double xi)
{
    double E = sqrt(p * p + m * m);
    double coeff = 0.5 * xi * (p * p) / E;
    return coeff * df_dE(sg, E, mu, T, Phi, Phib);
}

typedef struct {
    int sg;
    double lam, k, m, mp, mu, T, Phi, Phib, xi, A0, cE, dc;
} bcctx;

/* compute_coefficients  OneLoopIntegralsAniso.jl:50-61 */
static void coeffs(double lam, double k, double m, double mp, double E, double* cx, double* dc)
{
    if (k > EPS_K) {
        *dc = lam * lam + 2 * lam * E + m * m - mp * mp - k * k;
        *cx = 2 * p_int(E, m) * k;
    } else {
        *dc = lam * lam + m * m - mp * mp;
        *cx = 2 * lam;
    }
}

static double real_tool(double cx, double dc)
{
    double t1 = -2 * dc / (cx * cx);
    double t2 = dc * dc / (cx * cx * cx);
    double la = (cx + dc) / (cx - dc);
    return t1 + t2 * log(fabs(la));
}

static double imag_tool(double cx, double dc) { return PI_D * dc * dc / (cx * cx * cx); }

static double bc_numer_k0(const bcctx* c, double E)
{
    double p = p_int(E, c->m);
    return (4.0 / 3.0) * p * corr_coef(c->sg, p, c->m, c->mu, c->T, c->Phi, c->Phib, c->xi);
}

static double bc_imag_k0(const bcctx* c)
{
    double cE, dc;
    coeffs(c->lam, 0.0, c->m, c->mp, 0.0, &cE, &dc);
    double Ep = -dc / cE;
    double x = -dc / cE;
    if (!(x >= c->m && x <= e_cut(c->m))) return 0.0;
    double pp = p_int(Ep, c->m);
    return PI_D * (4.0 / 3.0) * pp / cE * corr_coef(c->sg, pp, c->m, c->mu, c->T, c->Phi, c->Phib, c->xi);
}

/* tilde_B0_correction_k_zero  :133-190 */
static void tilde_Bc_k0(const bcctx* c, double* re, double* im)
{
    double Emin = c->m, Emax = e_cut(c->m);
    double cE = 2.0 * c->lam;
    double dc = c->lam * c->lam + c->m * c->m - c->mp * c->mp;
    int has = 0;
    double E0 = 0.0;
    if (cE != 0.0) {
        double v = -dc / cE;
        if (isfinite(v) && v >= Emin && v <= Emax) {
            has = 1;
            E0 = v;
        }
    }
    double half = (Emax - Emin) / 2.0, sum = 0.0;
    if (has) {
        double A0 = bc_numer_k0(c, E0);
        for (int i = 0; i < 32; i++) {
            double E = ((Emax - Emin) * Q.x32[i] + (Emax + Emin)) / 2.0;
            double w = Q.w32[i] * (Emax - Emin) / 2.0;
            double den = cE * E + dc;
            double v = bc_numer_k0(c, E) - A0;
            sum += w * ((isfinite(den) && isfinite(v)) ? v / den : 0.0);
        }
        sum += A0 / cE * (log(fabs(cE * Emax + dc)) - log(fabs(cE * Emin + dc)));
    } else {
        for (int i = 0; i < 32; i++) {
            double E = ((Emax - Emin) * Q.x32[i] + (Emax + Emin)) / 2.0;
            double w = Q.w32[i] * (Emax - Emin) / 2.0;
            double num = bc_numer_k0(c, E);
            double den = cE * E + dc;
            sum += w * ((isfinite(den) && isfinite(num)) ? num / den : 0.0);
        }
    }
    (void)half;
    *re = sum;
    *im = bc_imag_k0(c);
}

static double bc_real_kpos(const void* cv, double E)
{
    const bcctx* c = (const bcctx*)cv;
    double cx, dc;
    coeffs(c->lam, c->k, c->m, c->mp, E, &cx, &dc);
    double p = p_int(E, c->m);
    return 2.0 * p * real_tool(cx, dc) * corr_coef(c->sg, p, c->m, c->mu, c->T, c->Phi, c->Phib, c->xi);
}

static double bc_imag_kpos(const bcctx* c, double E)
{
    double cx, dc;
    coeffs(c->lam, c->k, c->m, c->mp, E, &cx, &dc);
    double p = p_int(E, c->m);
    return 2.0 * p * imag_tool(cx, dc) * corr_coef(c->sg, p, c->m, c->mu, c->T, c->Phi, c->Phib, c->xi);
}

/* find_roots_AB  :207-272; returns the number of roots (ascending, unique) */
static int roots_AB(double lam, double k, double m, double mp, double Emin, double Emax, double r[2])
{
    double C = lam * lam + m * m - mp * mp - k * k;
    double a = k * k - lam * lam, b = -lam * C, c = -(k * k * m * m + C * C / 4);
    if (fabs(a) < 1e-14) {
        if (fabs(b) < 1e-14) return 0;
        double E = -c / b;
        if (E >= Emin && E <= Emax && E >= m) {
            double p = sqrt(fmax(0.0, E * E - m * m));
            double A = 2 * k * p, B = lam * lam + 2 * lam * E + m * m - mp * mp - k * k;
            if (fabs(A + B) < 1e-10 || fabs(A - B) < 1e-10) {
                r[0] = E;
                return 1;
            }
        }
        return 0;
    }
    double D = b * b - 4 * a * c;
    if (D < 0) return 0;
    double sD = sqrt(D);
    double E12[2] = {(-b - sD) / (2 * a), (-b + sD) / (2 * a)};
    int n = 0;
    for (int i = 0; i < 2; i++) {
        double E = E12[i];
        if (E >= m && E >= Emin && E <= Emax) {
            double p = sqrt(E * E - m * m);
            double A = 2 * k * p, B = lam * lam + 2 * lam * E + m * m - mp * mp - k * k;
            if (fabs(A + B) < 1e-10 || fabs(A - B) < 1e-10) r[n++] = E;
        }
    }
    if (n == 2) {
        if (r[0] > r[1]) {
            double t = r[0];
            r[0] = r[1];
            r[1] = t;
        }
        if (r[0] == r[1]) n = 1;
    }
    return n;
}

/* tilde_B0_correction_k_positive  :307-375 */
static void tilde_Bc_kpos(const bcctx* c, double* re, double* im)
{
    double Emin = c->m, Emax = e_cut(c->m);
    double ivi[2][2];
    int nimag = sing_kpos(c->lam, c->k, c->m, c->mp, Emin, Emax, ivi);
    double roots[2];
    int nr = roots_AB(c->lam, c->k, c->m, c->mp, Emin, Emax, roots);
    /* build_intervals_from_roots  :275-292 */
    double del = fmax(1e-10, 1e-8 * (Emax - Emin));
    double ep[6];
    int ne = 0;
    ep[ne++] = Emin;
    for (int i = 0; i < nr; i++) {
        ep[ne++] = roots[i] - del;
        ep[ne++] = roots[i] + del;
    }
    ep[ne++] = Emax;
    double iv[3][2];
    int niv = 0;
    for (int i = 0; i + 1 < ne; i += 2)
        if (ep[i + 1] > ep[i] + 1e-16) {
            iv[niv][0] = ep[i];
            iv[niv][1] = ep[i + 1];
            niv++;
        }
    double real = 0.0;
    for (int i = 0; i < niv; i++) {
        int pos;
        if (niv == 1) pos = nr == 0 ? SING_NONE : SING_BOTH;
        else if (i == 0) pos = SING_RIGHT;
        else if (i == niv - 1) pos = SING_LEFT;
        else pos = SING_BOTH;
        int n = (nr == 0) ? 16 : (pos == SING_LEFT ? 16 : 32);
        real += hybrid(bc_real_kpos, c, iv[i][0], iv[i][1], pos, n);
    }
    double imag = 0.0;
    if (nimag > 0) {
        for (int j = 0; j < nimag; j++) {
            double E1 = ivi[j][0], E2 = ivi[j][1];
            double half = (E2 - E1) / 2, center = (E1 + E2) / 2;
            for (int i = 0; i < 16; i++) {
                double x = center + half * Q.x16[i];
                double w = Q.w16[i] * half;
                double v = bc_imag_kpos(c, x);
                if (isfinite(v)) imag += w * v;
            }
        }
        imag *= jsign(c->lam);
    }
    *re = real;
    *im = imag;
}

static void tilde_Bc(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi, double Phib,
                     double xi, double* re, double* im)
{
    bcctx c;
    memset(&c, 0, sizeof c);
    c.sg = sg;
    c.lam = lam;
    c.k = k;
    c.m = m;
    c.mp = mp;
    c.mu = mu;
    c.T = T;
    c.Phi = Phi;
    c.Phib = Phib;
    c.xi = xi;
    if (k > EPS_K) tilde_Bc_kpos(&c, re, im);
    else tilde_Bc_k0(&c, re, im);
}

/* one delta-B~ term (tilde_B0_correction, OneLoopIntegralsAniso.jl:380-390); sg 0 quark, 1 antiquark */
void oracle_tilde_B0_correction(int sg, double lam, double k, double m, double mp, double mu, double T, double Phi,
                                double Phib, double xi, double* re, double* im)
{
    ensure_tables();
    tilde_Bc(sg == 0 ? QUARK : ANTIQUARK, lam, k, m, mp, mu, T, Phi, Phib, xi, re, im);
}

/* real_integral_tool / imag_integral_tool  OneLoopIntegralsAniso.jl:26-38 */
double oracle_real_tool(double cx, double dc) { return real_tool(cx, dc); }
double oracle_imag_tool(double cx, double dc) { return imag_tool(cx, dc); }

/* B0_correction  OneLoopIntegralsAniso.jl:396-413 */
void oracle_B0_correction(double lam, double k, double m1, double m2, double mu1, double mu2, double T, double Phi,
                          double Phib, double xi, double* re, double* im)
{
    ensure_tables();
    double r[4], i[4];
    tilde_Bc(QUARK, -lam, k, m1, m2, mu1, T, Phi, Phib, xi, &r[0], &i[0]);
    tilde_Bc(ANTIQUARK, lam, k, m1, m2, mu1, T, Phi, Phib, xi, &r[1], &i[1]);
    tilde_Bc(QUARK, lam, k, m2, m1, mu2, T, Phi, Phib, xi, &r[2], &i[2]);
    tilde_Bc(ANTIQUARK, -lam, k, m2, m1, mu2, T, Phi, Phib, xi, &r[3], &i[3]);
    *re = r[0] + r[1] + r[2] + r[3];
    *im = i[0] + i[1] + i[2] + i[3];
}

/* A_correction / A_aniso  OneLoopIntegralsAniso.jl:421-467 */
double oracle_A_correction(double m, double mu, double T, double Phi, double Phib, double xi, int n, const double* xp,
                           const double* wp)
{
    double integral = 0.0;
    for (int i = 0; i < n; i++) {
        double p = xp[i], E = sqrt(p * p + m * m);
        integral += wp[i] * (p * p) / E *
                    (corr_coef(QUARK, p, m, mu, T, Phi, Phib, xi) + corr_coef(ANTIQUARK, p, m, mu, T, Phi, Phib, xi));
    }
    return 4.0 * integral / 3.0;
}

double oracle_A_aniso(double m, double mu, double T, double Phi, double Phib, double xi, int np, const double* xp,
                      const double* wp, int nc, const double* xc, const double* wc)
{
    double integral = 0.0;
    for (int i = 0; i < np; i++) {
        double p = xp[i], aq = 0.0, aa = 0.0;
        double E = sqrt(p * p + m * m);
        for (int j = 0; j < nc; j++) {
            double c = xc[j];
            double pc = p * c;
            double Ea = sqrt(p * p + m * m + xi * (pc * pc));
            aq += wc[j] * (p * p) / E * quark_f(Ea, mu, T, Phi, Phib);
            aa += wc[j] * (p * p) / E * antiquark_f(Ea, mu, T, Phi, Phib);
        }
        integral += wp[i] * (aq + aa);
    }
    return 2.0 * integral - 4.0 * oracle_A_const(m);
}

/* ---------------------------------------------------------------------- */
/* polarization_aniso / polarization_with_width  PolarizationAniso.jl:22-118 */
/* ---------------------------------------------------------------------- */
int oracle_polarization(int channel, int width, double k0, double gamma, double k, double m1, double m2, double mu1,
                        double mu2, double T, double Phi, double Phib, double xi, double A1, double A2, int num_s,
                        double* re, double* im)
{
    ensure_tables();
    if (channel != PNJL_CHANNEL_P && channel != PNJL_CHANNEL_S) return PNJL_ERR_ARG;
    double factor = -g_Nc / (8 * (PI_D * PI_D));
    double lam = k0 + mu1 - mu2;
    double br, bi;
    oracle_B0(lam, k, m1, mu1, m2, mu2, T, Phi, Phib, &br, &bi);
    double lx = -k0 + mu1 - mu2;
    if (num_s == 1) {
        double r2, i2;
        oracle_B0(lx, k, m1, mu1, m2, mu2, T, Phi, Phib, &r2, &i2);
        br = 0.5 * (br + r2);
        bi = 0.5 * (bi + i2);
    }
    if (fabs(xi) > EPS_SEGMENT) {
        double cr, ci;
        oracle_B0_correction(lam, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, &cr, &ci);
        if (num_s == 1) {
            double r2, i2;
            oracle_B0_correction(lx, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, &r2, &i2);
            cr = 0.5 * (cr + r2);
            ci = 0.5 * (ci + i2);
        }
        br += cr;
        bi += ci;
    }
    double real = A1 + A2, imag = 0.0;
    double pref = k * k - lam * lam;
    pref += channel == PNJL_CHANNEL_P ? (m1 - m2) * (m1 - m2) : (m1 + m2) * (m1 + m2);
    if (width) {
        pref += (gamma * gamma) / 4.0;
        real += pref * br - gamma * lam * bi;
        imag += pref * bi + gamma * lam * br;
    } else {
        real += pref * br;
        imag += pref * bi;
    }
    *re = factor * real;
    *im = factor * imag;
    return 0;
}

/* batch over query structs (host pointers), OpenMP over queries */
int oracle_polarization_batch(int64_t n, const pnjl_pol_query* q, double* re, double* im, int nthreads)
{
    ensure_tables();
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : bad)
    for (int64_t i = 0; i < n; i++) {
        const pnjl_pol_query* a = q + i;
        if (oracle_polarization(a->channel, a->width, a->k0, a->gamma, a->k, a->m1, a->m2, a->mu1, a->mu2, a->T,
                                a->Phi, a->Phib, a->xi, a->A1, a->A2, a->num_s, re + i, im + i))
            bad = 1;
    }
    return bad ? PNJL_ERR_ARG : 0;
}
