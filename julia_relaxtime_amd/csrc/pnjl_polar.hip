// pnjl_polar.hip -- the anisotropic one-loop polarization Pi(k0, k) on gfx950.
//
// One GPU thread evaluates one B0 / B0_correction / polarization value with
// the reference's fixed quadrature: the same breakpoints (poles of the k = 0
// integrand, roots of the k > 0 logarithm, imaginary-part intervals), the
// same 16/32-node power-clustered and 33-node double-exponential rules, the
// same principal-value subtraction and the same skip-non-finite rule, so the
// results reproduce the reference's numbers, not the exact integrals.
// Quadrature tables sit in LDS.  For the grid call, the threads of a
// wavefront share (point, pair) and take an 8 x 8 tile of (k0, k), so their
// segment structures differ only smoothly; the propagator batch sorts its
// work into buckets of similar (pair, mass, lambda, k) for the same reason.
//
// Follows src/relaxtime/OneLoopIntegrals.jl, OneLoopIntegralsAniso.jl,
// PolarizationAniso.jl, src/integration/IntervalQuadratureStrategies.jl and
// src/QuarkDistribution{,_Aniso}.jl (file:line at each function).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pnjl_ctx.h"

using namespace pnjl_host;

// No contraction of a*b + c into FMA here: the reference evaluates these
// expressions unfused, and some of them cancel exactly (lambda^2 + m^2 - m'^2
// with m = m'), where a fused product leaves a 1e-17 residue instead of 0.
#pragma clang fp contract(off)

#include "pnjl_meson.h"

namespace polar {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEpsK = 1.0e-9;        // OneLoopIntegrals.jl:23
constexpr double kEpsSeg = 1.0e-12;     // OneLoopIntegrals.jl:24
constexpr double kPvGapRel = 1e-6;      // OneLoopIntegrals.jl:35
constexpr double kAlpha = 0.35;         // IntervalQuadratureStrategies.jl:35
constexpr double kDeH = 0.15;           // IntervalQuadratureStrategies.jl:36
constexpr int kTabLen = 354;            // see tab_* offsets below

enum { NONE = 0, LEFT = 1, RIGHT = 2, BOTH = 3 };
enum { PLUS = 0, MINUS = 1 };           // B0 signs (f+(E), f+(-E) = 1 - f-(E))
enum { QUARK = 0, ANTIQUARK = 1 };      // first-order anisotropy signs

// table layout: x16 w16 x32 w32 | pl_t16 pl_tp16 pl_t32 pl_tp32 | pr_t16 pr_tp16 pr_t32 pr_tp32 | de_phi de_phip
struct Tabs {
    const double *x16, *w16, *x32, *w32;
    const double *pl_t16, *pl_tp16, *pl_t32, *pl_tp32;
    const double *pr_t16, *pr_tp16, *pr_t32, *pr_tp32;
    const double *de_phi, *de_phip;
};

__device__ __forceinline__ Tabs make_tabs(const double* t)
{
    Tabs q;
    q.x16 = t; q.w16 = t + 16; q.x32 = t + 32; q.w32 = t + 64;
    q.pl_t16 = t + 96; q.pl_tp16 = t + 112; q.pl_t32 = t + 128; q.pl_tp32 = t + 160;
    q.pr_t16 = t + 192; q.pr_tp16 = t + 208; q.pr_t32 = t + 224; q.pr_tp32 = t + 256;
    q.de_phi = t + 288; q.de_phip = t + 321;
    return q;
}

// The pole-search kernels' quadrature tables.  Their B0 routine is out of
// line; it rebuilds its table pointers from this array rather than taking
// them from the caller, so that it reads the tables with LDS instructions
// (a pointer passed in would be a generic address: flat loads).
__shared__ double mass_tab[kTabLen];

struct Phys {
    double Lambda, Nc;
};

// Julia eps(x): spacing of doubles at |x|
__device__ __forceinline__ double jeps(double x)
{
    x = fabs(x);
    if (!(x >= 2.2250738585072014e-308)) return 4.9406564584124654e-324;
    return ldexp(1.0, ilogb(x) - 52);
}

__device__ __forceinline__ double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
__device__ __forceinline__ double clamp_e(double e) { return e < 1e-200 ? 1e-200 : (e > 1e200 ? 1e200 : e); }

// PNJL distributions and their antiderivatives  QuarkDistribution.jl:14-99
__device__ __forceinline__ double dist_ratio(double e, double c1, double c2)
{
    double e2 = e * e, e3 = e2 * e;
    return (c1 * e + 2 * c2 * e2 + e3) / (1 + 3 * c1 * e + 3 * c2 * e2 + e3);
}

__device__ __forceinline__ double quark_f(double E, double mu, double T, double Phi, double Phib)
{
    return dist_ratio(clamp_e(exp(-(E - mu) * (1 / T))), Phi, Phib);
}

__device__ __forceinline__ double antiquark_f(double E, double mu, double T, double Phi, double Phib)
{
    return dist_ratio(clamp_e(exp(-(E + mu) * (1 / T))), Phib, Phi);
}

__device__ __forceinline__ double anti_deriv(double e, double c1, double c2, double T)
{
    double e2 = e * e, e3 = e2 * e;
    double d = 1 + 3 * c1 * e + 3 * c2 * e2 + e3;
    return -T * log(d <= 1e-300 ? 1e-300 : d) / 3;
}

// distribution_value_b0 / distribution_integral_b0  OneLoopIntegrals.jl:68-104
__device__ __forceinline__ double dist_b0(int sg, double E, double mu, double T, double Phi, double Phib)
{
    return sg == PLUS ? quark_f(E, mu, T, Phi, Phib) : 1.0 - antiquark_f(E, mu, T, Phi, Phib);
}

__device__ double dist_int_b0(int sg, double a, double b, double mu, double T, double Phi, double Phib)
{
    const double beta = 1 / T;
    if (sg == PLUS)
        return anti_deriv(clamp_e(exp(-(b - mu) * beta)), Phi, Phib, T) -
               anti_deriv(clamp_e(exp(-(a - mu) * beta)), Phi, Phib, T);
    return (b - a) - (anti_deriv(clamp_e(exp(-(b + mu) * beta)), Phib, Phi, T) -
                      anti_deriv(clamp_e(exp(-(a + mu) * beta)), Phib, Phi, T));
}

// energy_cutoff / internal_momentum  OneLoopIntegrals.jl:37-48
__device__ __forceinline__ double e_cut(const Phys& P, double m)
{
    double mp = m > 0 ? m : 0.0;
    return sqrt(mp * mp + P.Lambda * P.Lambda);
}

__device__ __forceinline__ double p_int(double E, double m)
{
    double mp = m > 0 ? m : 0.0;
    return sqrt(E * E - mp * mp);
}

// Lane-split quadrature (L > 1, the pole search's long tail): the L
// consecutive lanes of a group each evaluate the nodes i = gl, gl + L, ... of
// an N-node rule and leave the terms in the group's LDS row, then every lane
// adds all N terms in the order i = 0 .. N-1 -- the one-thread loop's sum, bit
// for bit, on every lane.  A skipped (non-finite) node contributes +0.0 there,
// which leaves any sum the loop can hold unchanged: the sums start at +0.0
// and so never hold -0.0.  Callers keep the group's control flow uniform
// (same arguments on every lane of the group); the group's lanes are in one
// wavefront, whose LDS accesses complete in program order.  L = 1 is the
// one-thread loop itself.
constexpr int kQRow = 33;              // terms of the longest rule
constexpr int kQRows = 256 / 4;        // rows per 256-thread block at 4 lanes per group
__shared__ double qterms[kQRows][kQRow];

template <int L>
__device__ __forceinline__ int qlane()
{
    return L == 1 ? 0 : (int)(threadIdx.x & (L - 1));
}

template <int L, int N, class Term>
__device__ __forceinline__ double ordered_sum(double acc, const Term& term)
{
    static_assert(L > 1 && (L & (L - 1)) == 0 && 256 / L <= kQRows && N <= kQRow, "lane-split rule shape");
    double* row = qterms[threadIdx.x / L];
    const int gl = qlane<L>();
#pragma unroll 1
    for (int i = gl; i < N; i += L) row[i] = term(i);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int i = 0; i < N; i++) acc += row[i];   // the loads issue together, the adds keep their order
    // the row is written again by the group's next rule only after every lane has read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return acc;
}

// integrate_hybrid_interval  IntervalQuadratureStrategies.jl:194-310
template <int L, int N, class F>
__device__ double rule_power(const F& f, double a, double b, const double* t, const double* tp, const double* W)
{
    double len = b - a, total = 0.0;
    if constexpr (L > 1) {
        return ordered_sum<L, N>(total, [&](int i) {
            double x = a + len * t[i];
            double w = W[i] * len * tp[i];
            double v = f(x);
            return isfinite(v) ? w * v : 0.0;
        });
    } else {
        for (int i = 0; i < N; i++) {
            double x = a + len * t[i];
            double w = W[i] * len * tp[i];
            double v = f(x);
            if (isfinite(v)) total += w * v;
        }
        return total;
    }
}

template <int L, class F>
__device__ double hybrid(const F& f, const Tabs& q, double a, double b, int pos, int n)
{
    if (!(isfinite(a) && isfinite(b)) || b <= a) return 0.0;
    if (pos == BOTH) {
        double half = (b - a) / 2, center = (a + b) / 2, total = 0.0;
        if constexpr (L > 1) {
            return ordered_sum<L, 33>(total, [&](int i) {
                double x = center + half * q.de_phi[i];
                double w = half * q.de_phip[i];
                double v = f(x);
                return isfinite(v) ? w * v : 0.0;
            });
        } else {
            for (int i = 0; i < 33; i++) {
                double x = center + half * q.de_phi[i];
                double w = half * q.de_phip[i];
                double v = f(x);
                if (isfinite(v)) total += w * v;
            }
            return total;
        }
    }
    if (pos == RIGHT)
        return n == 16 ? rule_power<L, 16>(f, a, b, q.pr_t16, q.pr_tp16, q.w16)
                       : rule_power<L, 32>(f, a, b, q.pr_t32, q.pr_tp32, q.w32);
    return n == 16 ? rule_power<L, 16>(f, a, b, q.pl_t16, q.pl_tp16, q.w16)
                   : rule_power<L, 32>(f, a, b, q.pl_t32, q.pl_tp32, q.w32);
}

// ---- B~0 at k = 0  OneLoopIntegrals.jl:117-224 ------------------------------
struct B0K0 {
    int sg;
    double lam, m, dterm, mu, T, Phi, Phib;
    __device__ double operator()(double E) const
    {
        double p = p_int(E, m);
        double dist = dist_b0(sg, E, mu, T, Phi, Phib);
        double den = lam * E + dterm;
        if (!isfinite(p) || !isfinite(dist) || !isfinite(den)) return 0.0;
        if (fabs(den) < kEpsSeg) return 0.0;
        double v = p * dist / den;
        return isfinite(v) ? v : 0.0;
    }
};

struct B0Numer {
    int sg;
    double m, mu, T, Phi, Phib;
    __device__ double operator()(double E) const
    {
        double p = p_int(E, m);
        double dist = dist_b0(sg, E, mu, T, Phi, Phib);
        if (!isfinite(p) || !isfinite(dist)) return 0.0;
        double v = p * dist;
        return isfinite(v) ? v : 0.0;
    }
};

struct B0Regular {
    B0Numer num;
    double E0, f0, Emin, Emax;
    __device__ double operator()(double E) const
    {
        double d = E - E0;
        if (fabs(d) < 64 * jeps(E0)) {
            double del = fmax(1e-7 * (Emax - Emin), 64 * jeps(E0));
            double El = fmax(Emin, E0 - del), Er = fmin(Emax, E0 + del);
            return Er > El ? (num(Er) - num(El)) / (Er - El) : 0.0;
        }
        return (num(E) - f0) / d;
    }
};

template <int L>
__device__ void tilde_B0_k0(const Phys& P, const Tabs& q, int sg, double lam, double m_in, double mp_in, double mu,
                            double T, double Phi, double Phib, double& re, double& im)
{
    const double m = m_in > 0 ? m_in : 0.0, mp = mp_in > 0 ? mp_in : 0.0;
    const double Emin = m, Emax = e_cut(P, m);
    const double dterm = (lam * lam + m * m - mp * mp) / 2.0;
    const B0K0 f{sg, lam, m, dterm, mu, T, Phi, Phib};
    // singularity_k_zero  :144-157
    const double E0 = -dterm / lam;
    bool has = false;
    if (isfinite(E0)) {
        double tol = 1e-8 * fmax(1.0, fabs(E0));
        has = (E0 > Emin - tol) && (E0 < Emax + tol);
    }
    if (!has) {
        re = hybrid<L>(f, q, Emin, Emax, NONE, 16) * 2.0;
        im = 0.0 / lam;
        return;
    }
    const B0Numer num{sg, m, mu, T, Phi, Phib};
    const double f0 = num(E0);
    double pv;
    const double etol = 1e-10 * fmax(1.0, Emax - Emin);
    if (fabs(E0 - Emin) <= etol || fabs(Emax - E0) <= etol) {
        double gap = fmax(kPvGapRel * (Emax - Emin), 64 * jeps(E0));
        double left_end = fmax(Emin, E0 - gap), right_start = fmin(Emax, E0 + gap);
        pv = 0.0;
        if (left_end > Emin) pv += hybrid<L>(f, q, Emin, left_end, NONE, 16);
        if (right_start < Emax) pv += hybrid<L>(f, q, right_start, Emax, NONE, 16);
    } else {
        const B0Regular g{num, E0, f0, Emin, Emax};
        double regular = hybrid<L>(g, q, Emin, Emax, NONE, 32);
        double logterm = f0 * (log(fabs(Emax - E0)) - log(fabs(Emin - E0)));
        pv = (regular + logterm) / lam;
    }
    const double p0 = p_int(E0, m);
    re = pv * 2.0;
    im = 2.0 * kPi * p0 * dist_b0(sg, E0, mu, T, Phi, Phib) / lam;
}

// ---- B~0 at k > 0  OneLoopIntegrals.jl:228-476 ------------------------------
struct B0Kpos {
    int sg;
    double lam, k, m, mp, mu, T, Phi, Phib;
    __device__ double operator()(double E) const
    {
        double p = p_int(E, m);
        double dist = dist_b0(sg, E, mu, T, Phi, Phib);
        double lpE = lam + E;
        double common = lpE * lpE - mp * mp;
        double pmk = p - k, ppk = p + k;
        double num = common - pmk * pmk;
        double den = common - ppk * ppk;
        if (!isfinite(p) || !isfinite(dist) || !isfinite(num) || !isfinite(den)) return 0.0;
        if (fabs(num) < kEpsSeg && fabs(den) < kEpsSeg) return 0.0;
        if (fabs(den) < kEpsSeg) den = (den == 0.0) ? kEpsSeg : jsign(den) * kEpsSeg;
        double ratio = fabs(num / den);
        if (!isfinite(ratio)) return 0.0;
        ratio = fmax(ratio, 1e-300);
        double lt = log(ratio);
        if (!isfinite(lt)) return 0.0;
        double v = dist * lt;
        return isfinite(v) ? v : 0.0;
    }
};

// singularity_k_positive  :239-330 -> number of intervals in iv (0..2)
__device__ int sing_kpos(double lam, double k, double m, double mp, double Emin, double Emax, double iv[2][2])
{
    int n = 0;
    const double a = lam * lam - k * k;
    const double d0 = k * k - lam * lam;
    const double denom = 2.0 * d0;
    if (fabs(a) < kEpsSeg) {
        if (lam > 0.0 && mp > m) {
            double m2d = mp * mp - m * m;
            double Ec = m2d / (4.0 * lam) + lam * m * m / m2d;
            if (Ec < Emax) {
                iv[0][0] = fmax(Ec, Emin);
                iv[0][1] = Emax;
                n = 1;
            }
        }
        return n;
    }
    const double A = lam * lam - k * k + m * m - mp * mp;
    const double disc = A * A + 4.0 * m * m * d0;
    if (disc < 0.0) return 0;
    const double sq = sqrt(disc);
    double E1 = (lam * A - k * sq) / denom, E2 = (lam * A + k * sq) / denom;
    if (E1 > E2) {
        double t = E1;
        E1 = E2;
        E2 = t;
    }
    if (E2 - E1 < kEpsSeg) return 0;
    if (a > 0.0) {
        if (E2 > Emin && E1 < Emax) {
            double c1 = fmax(E1, Emin), c2 = fmin(E2, Emax);
            if (c2 > c1) {
                iv[0][0] = c1;
                iv[0][1] = c2;
                n = 1;
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

// integrate_piecewise_hybrid  :335-421
template <int L, class F>
__device__ double piecewise(const F& f, const Tabs& q, double Emin, double Emax, int niv, const double iv[2][2])
{
    if (niv == 0) return hybrid<L>(f, q, Emin, Emax, NONE, 16);
    double E1 = __builtin_nan(""), E2 = __builtin_nan("");
    if (niv == 1) {
        E1 = iv[0][0];
        E2 = iv[0][1];
    } else {
        const double a1 = iv[0][0], b1 = iv[0][1], a2 = iv[1][0], b2 = iv[1][1];
        const double tol = 1e-12 * fmax(1.0, Emax - Emin);
        if (fabs(a1 - Emin) <= tol) E1 = b1;
        else if (fabs(a2 - Emin) <= tol) E1 = b2;
        if (fabs(b1 - Emax) <= tol) E2 = a1;
        else if (fabs(b2 - Emax) <= tol) E2 = a2;
    }
    if (isfinite(E1) && isfinite(E2)) {
        double total = 0.0;
        if (E1 > Emin) total += hybrid<L>(f, q, Emin, E1, RIGHT, 32);
        if (E2 > E1) total += hybrid<L>(f, q, E1, E2, BOTH, 32);
        if (Emax > E2) total += hybrid<L>(f, q, E2, Emax, LEFT, 16);
        return total;
    }
    // general path: sorted unique breakpoints (insertion sort of <= 6 values)
    double pts[6];
    int np = 0;
    pts[np++] = Emin;
    for (int i = 0; i < niv; i++) {
        pts[np++] = iv[i][0];
        pts[np++] = iv[i][1];
    }
    pts[np++] = Emax;
    for (int i = 1; i < np; i++)
        for (int j = i; j > 0 && pts[j] < pts[j - 1]; j--) {
            double t = pts[j];
            pts[j] = pts[j - 1];
            pts[j - 1] = t;
        }
    int u = 0;
    for (int i = 0; i < np; i++)
        if (u == 0 || pts[i] != pts[u - 1]) pts[u++] = pts[i];
    const int nseg = u - 1;
    if (nseg <= 0) return 0.0;
    if (nseg == 1) return hybrid<L>(f, q, pts[0], pts[1], NONE, 16);
    double total = 0.0;
    for (int i = 1; i <= nseg; i++) {
        double a = pts[i - 1], b = pts[i];
        if (!(isfinite(a) && isfinite(b)) || b <= a) continue;
        int pos = (i == 1) ? RIGHT : (i == nseg ? LEFT : BOTH);
        total += hybrid<L>(f, q, a, b, pos, i == nseg ? 16 : 32);
    }
    return total;
}

template <int L>
__device__ void tilde_B0_kpos(const Phys& P, const Tabs& q, int sg, double lam, double k, double m_in, double mp_in,
                              double mu, double T, double Phi, double Phib, double& re, double& im)
{
    const double m = m_in > 0 ? m_in : 0.0, mp = mp_in > 0 ? mp_in : 0.0;
    const double Emin = m, Emax = e_cut(P, m);
    double iv[2][2];
    const int niv = sing_kpos(lam, k, m, mp, Emin, Emax, iv);
    const B0Kpos f{sg, lam, k, m, mp, mu, T, Phi, Phib};
    const double real = piecewise<L>(f, q, Emin, Emax, niv, iv);
    double imag = 0.0;
    if (niv > 0) {
        for (int i = 0; i < niv; i++) imag += dist_int_b0(sg, iv[i][0], iv[i][1], mu, T, Phi, Phib);
        imag *= kPi * jsign(lam);
    }
    re = real / k;
    im = imag / k;
}

template <int L>
__device__ void tilde_B0(const Phys& P, const Tabs& q, int sg, double lam, double k, double m, double mp, double mu,
                         double T, double Phi, double Phib, double& re, double& im)
{
    if (fabs(k) < kEpsK) tilde_B0_k0<L>(P, q, sg, lam, m, mp, mu, T, Phi, Phib, re, im);
    else tilde_B0_kpos<L>(P, q, sg, lam, k, m, mp, mu, T, Phi, Phib, re, im);
}

// B0  OneLoopIntegrals.jl:489-503: B~+(-l; m1, m2, mu1) - B~-(l; m1, m2, mu1)
// + B~+(l; m2, m1, mu2) - B~-(-l; m2, m1, mu2), summed in this order.  One
// loop over the four terms keeps a single copy of tilde_B0 in the kernel.
template <int L = 1>
__device__ void B0(const Phys& P, const Tabs& q, double lam, double k, double m1, double mu1, double m2, double mu2,
                   double T, double Phi, double Phib, double& re, double& im)
{
    re = 0.0;
    im = 0.0;
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
        const bool swap = t >= 2;
        const double l = (t == 0 || t == 3) ? -lam : lam;
        double r, i;
        tilde_B0<L>(P, q, (t & 1) ? MINUS : PLUS, l, k, swap ? m2 : m1, swap ? m1 : m2, swap ? mu2 : mu1, T, Phi, Phib,
                 r, i);
        if (t & 1) {
            re -= r;
            im -= i;
        } else {
            re += r;
            im += i;
        }
    }
}

// ---- first-order anisotropy  OneLoopIntegralsAniso.jl ---------------------
// quark_df_dE / antiquark_df_dE / correction_cos_theta_coefficient
//   QuarkDistribution_Aniso.jl:14-86
__device__ double corr_coef(int sg, double p, double m, double mu, double T, double Phi, double Phib, double xi)
{
    const double E = sqrt(p * p + m * m);
    const double coeff = 0.5 * xi * (p * p) / E;
    const double beta = 1 / T;
    const double c1 = sg == QUARK ? Phi : Phib, c2 = sg == QUARK ? Phib : Phi;
    const double arg = sg == QUARK ? -(E - mu) * beta : -(E + mu) * beta;
    const double e = clamp_e(exp(arg));
    const double e2 = e * e, e3 = e2 * e;
    const double num = c1 * e + 2 * c2 * e2 + e3;
    const double den = 1 + 3 * c1 * e + 3 * c2 * e2 + e3;
    const double dnum = -(c1 * e + 4 * c2 * e2 + 3 * e3);
    const double dden = -(3 * c1 * e + 6 * c2 * e2 + 3 * e3);
    double dfdx = dnum * den - num * dden;
    dfdx /= den * den;
    return coeff * (beta * dfdx);
}

// compute_coefficients  OneLoopIntegralsAniso.jl:50-61 (k > 0 branch)
struct BcKpos {
    int sg;
    double lam, k, m, mp, mu, T, Phi, Phib, xi;
    bool imag;
    __device__ double operator()(double E) const
    {
        const double dc = lam * lam + 2 * lam * E + m * m - mp * mp - k * k;
        const double p = p_int(E, m);
        const double cx = 2 * p * k;
        double tool;
        if (imag) {
            tool = kPi * dc * dc / (cx * cx * cx);
        } else {
            double t1 = -2 * dc / (cx * cx);
            double t2 = dc * dc / (cx * cx * cx);
            tool = t1 + t2 * log(fabs((cx + dc) / (cx - dc)));
        }
        return 2.0 * p * tool * corr_coef(sg, p, m, mu, T, Phi, Phib, xi);
    }
};

// find_roots_AB  :207-272 -> number of roots (ascending, unique)
__device__ int roots_AB(double lam, double k, double m, double mp, double Emin, double Emax, double r[2])
{
    const double C = lam * lam + m * m - mp * mp - k * k;
    const double a = k * k - lam * lam, b = -lam * C, c = -(k * k * m * m + C * C / 4);
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
    const double D = b * b - 4 * a * c;
    if (D < 0) return 0;
    const double sD = sqrt(D);
    const double E12[2] = {(-b - sD) / (2 * a), (-b + sD) / (2 * a)};
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

// tilde_B0_correction_k_zero  :133-190
template <int L>
__device__ void tilde_Bc_k0(const Phys& P, const Tabs& q, int sg, double lam, double m, double mp, double mu, double T,
                            double Phi, double Phib, double xi, double& re, double& im)
{
    const double Emin = m, Emax = e_cut(P, m);
    const double cE = 2.0 * lam;
    const double dc = lam * lam + m * m - mp * mp;
    bool has = false;
    double E0 = 0.0;
    if (cE != 0.0) {
        double v = -dc / cE;
        if (isfinite(v) && v >= Emin && v <= Emax) {
            has = true;
            E0 = v;
        }
    }
    auto numer = [&](double E) {
        double p = p_int(E, m);
        return (4.0 / 3.0) * p * corr_coef(sg, p, m, mu, T, Phi, Phib, xi);
    };
    double sum = 0.0;
    const double A0 = has ? numer(E0) : 0.0;
    auto term = [&](int i) {
        double E = ((Emax - Emin) * q.x32[i] + (Emax + Emin)) / 2.0;
        double w = q.w32[i] * (Emax - Emin) / 2.0;
        double den = cE * E + dc;
        double v = numer(E) - A0;
        return w * ((isfinite(den) && isfinite(v)) ? v / den : 0.0);
    };
    if constexpr (L > 1)
        sum = ordered_sum<L, 32>(sum, term);
    else
        for (int i = 0; i < 32; i++) sum += term(i);
    if (has) sum += A0 / cE * (log(fabs(cE * Emax + dc)) - log(fabs(cE * Emin + dc)));
    re = sum;
    // imag_integrand_k_zero  :98-112
    const double Ep = -dc / cE;
    if (!(Ep >= m && Ep <= e_cut(P, m))) {
        im = 0.0;
    } else {
        const double pp = p_int(Ep, m);
        im = kPi * (4.0 / 3.0) * pp / cE * corr_coef(sg, pp, m, mu, T, Phi, Phib, xi);
    }
}

// tilde_B0_correction_k_positive  :307-375
template <int L>
__device__ void tilde_Bc_kpos(const Phys& P, const Tabs& q, int sg, double lam, double k, double m, double mp,
                              double mu, double T, double Phi, double Phib, double xi, double& re, double& im)
{
    const double Emin = m, Emax = e_cut(P, m);
    double ivi[2][2];
    const int nimag = sing_kpos(lam, k, m, mp, Emin, Emax, ivi);
    double roots[2];
    const int nr = roots_AB(lam, k, m, mp, Emin, Emax, roots);
    // build_intervals_from_roots  :275-292
    const double del = fmax(1e-10, 1e-8 * (Emax - Emin));
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
    const BcKpos fr{sg, lam, k, m, mp, mu, T, Phi, Phib, xi, false};
    double real = 0.0;
    for (int i = 0; i < niv; i++) {
        int pos;
        if (niv == 1) pos = nr == 0 ? NONE : BOTH;
        else if (i == 0) pos = RIGHT;
        else if (i == niv - 1) pos = LEFT;
        else pos = BOTH;
        const int n = (nr == 0) ? 16 : (pos == LEFT ? 16 : 32);
        real += hybrid<L>(fr, q, iv[i][0], iv[i][1], pos, n);
    }
    double imag = 0.0;
    if (nimag > 0) {
        const BcKpos fi{sg, lam, k, m, mp, mu, T, Phi, Phib, xi, true};
        for (int j = 0; j < nimag; j++) {
            const double E1 = ivi[j][0], E2 = ivi[j][1];
            const double half = (E2 - E1) / 2, center = (E1 + E2) / 2;
            if constexpr (L > 1) {
                imag = ordered_sum<L, 16>(imag, [&](int i) {
                    double x = center + half * q.x16[i];
                    double w = q.w16[i] * half;
                    double v = fi(x);
                    return isfinite(v) ? w * v : 0.0;
                });
            } else {
                for (int i = 0; i < 16; i++) {
                    double x = center + half * q.x16[i];
                    double w = q.w16[i] * half;
                    double v = fi(x);
                    if (isfinite(v)) imag += w * v;
                }
            }
        }
        imag *= jsign(lam);
    }
    re = real;
    im = imag;
}

template <int L>
__device__ void tilde_Bc(const Phys& P, const Tabs& q, int sg, double lam, double k, double m, double mp, double mu,
                         double T, double Phi, double Phib, double xi, double& re, double& im)
{
    if (k > kEpsK) tilde_Bc_kpos<L>(P, q, sg, lam, k, m, mp, mu, T, Phi, Phib, xi, re, im);
    else tilde_Bc_k0<L>(P, q, sg, lam, m, mp, mu, T, Phi, Phib, xi, re, im);
}

// B0_correction  OneLoopIntegralsAniso.jl:396-413 (all four terms added):
// dB~q(-l; m1, m2, mu1) + dB~qbar(l; m1, m2, mu1) + dB~q(l; m2, m1, mu2)
// + dB~qbar(-l; m2, m1, mu2).
template <int L = 1>
__device__ void B0_corr(const Phys& P, const Tabs& q, double lam, double k, double m1, double m2, double mu1,
                        double mu2, double T, double Phi, double Phib, double xi, double& re, double& im)
{
    re = 0.0;
    im = 0.0;
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
        const bool swap = t >= 2;
        const double l = (t == 0 || t == 3) ? -lam : lam;
        double r, i;
        tilde_Bc<L>(P, q, (t & 1) ? ANTIQUARK : QUARK, l, k, swap ? m2 : m1, swap ? m1 : m2, swap ? mu2 : mu1, T, Phi,
                 Phib, xi, r, i);
        re += r;
        im += i;
    }
}

// The symmetrised B0 (+ B0_correction) of polarization_aniso
//   PolarizationAniso.jl:29-54
template <int L = 1>
__device__ void b0_total(const Phys& P, const Tabs& q, double k0, double k, double m1, double m2, double mu1,
                         double mu2, double T, double Phi, double Phib, double xi, int num_s, double& br, double& bi)
{
    // lambda = k0 + mu1 - mu2; num_s = 1 averages with lambda' = -k0 + mu1 - mu2
    const int nv = num_s == 1 ? 2 : 1;
    double vr[2], vi[2];
#pragma unroll 1
    for (int v = 0; v < nv; v++) {
        const double l = (v == 0 ? k0 : -k0) + mu1 - mu2;
        B0<L>(P, q, l, k, m1, mu1, m2, mu2, T, Phi, Phib, vr[v], vi[v]);
    }
    br = nv == 2 ? 0.5 * (vr[0] + vr[1]) : vr[0];
    bi = nv == 2 ? 0.5 * (vi[0] + vi[1]) : vi[0];
    if (fabs(xi) > kEpsSeg) {
#pragma unroll 1
        for (int v = 0; v < nv; v++) {
            const double l = (v == 0 ? k0 : -k0) + mu1 - mu2;
            B0_corr<L>(P, q, l, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, vr[v], vi[v]);
        }
        br += nv == 2 ? 0.5 * (vr[0] + vr[1]) : vr[0];
        bi += nv == 2 ? 0.5 * (vi[0] + vi[1]) : vi[0];
    }
}

// polarization_aniso / polarization_with_width  PolarizationAniso.jl:22-118
__device__ void pi_value(const Phys& P, int channel, int width, double k0, double gamma, double k, double m1,
                         double m2, double lam, double A12, double br, double bi, double& re, double& im)
{
    const double factor = -P.Nc / (8 * (kPi * kPi));
    double real = A12, imag = 0.0;
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
    re = factor * real;
    im = factor * imag;
}

// A / A_aniso  OneLoopIntegrals.jl:507-543, OneLoopIntegralsAniso.jl:446-467
__device__ double A_const(const Phys& P, double m)
{
    const double mp = m > 0 ? m : 0.0;
    const double L = P.Lambda;
    if (mp < 1e-14) return (L * L) / 2.0;
    const double s = sqrt(L * L + mp * mp);
    return (L * s - mp * mp * log((L + s) / mp)) / 2.0;
}

__device__ double A_value(const Phys& P, double m, double mu, double T, double Phi, double Phib, double xi, int np,
                          const double* xp, const double* wp, int nc, const double* xc, const double* wc)
{
    double integral = 0.0;
    if (nc <= 0) {
        integral = -A_const(P, m);
        for (int i = 0; i < np; i++) {
            double p = xp[i], E = sqrt(p * p + m * m);
            integral += wp[i] * (p * p) / E * (quark_f(E, mu, T, Phi, Phib) + antiquark_f(E, mu, T, Phi, Phib));
        }
        return 4.0 * integral;
    }
    for (int i = 0; i < np; i++) {
        double p = xp[i], aq = 0.0, aa = 0.0;
        double E = sqrt(p * p + m * m);
        for (int j = 0; j < nc; j++) {
            double pc = p * xc[j];
            double Ea = sqrt(p * p + m * m + xi * (pc * pc));
            aq += wc[j] * (p * p) / E * quark_f(Ea, mu, T, Phi, Phib);
            aa += wc[j] * (p * p) / E * antiquark_f(Ea, mu, T, Phi, Phib);
        }
        integral += wp[i] * (aq + aa);
    }
    return 2.0 * integral - 4.0 * A_const(P, m);
}

// ---- kernels -----------------------------------------------------------------
__device__ __forceinline__ void load_tabs(double* lds, const double* tab)
{
    for (int i = threadIdx.x; i < kTabLen; i += blockDim.x) lds[i] = tab[i];
    __syncthreads();
}

// query i = thread / stride (stride 1, or 64: one query per wavefront)
__global__ void __launch_bounds__(256, 3) query_kernel(Phys P, const double* tab, int64_t n, const pnjl_pol_query* qs,
                                                    double* re, double* im, int stride)
{
    __shared__ double lds[kTabLen];
    load_tabs(lds, tab);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t % stride) return;
    const int64_t i = t / stride;
    if (i >= n) return;
    const Tabs q = make_tabs(lds);
    const pnjl_pol_query a = qs[i];
    if (a.channel != PNJL_CHANNEL_P && a.channel != PNJL_CHANNEL_S) {
        re[i] = im[i] = __builtin_nan("");   // unsupported channel
        return;
    }
    double br, bi;
    b0_total(P, q, a.k0, a.k, a.m1, a.m2, a.mu1, a.mu2, a.T, a.Phi, a.Phib, a.xi, a.num_s, br, bi);
    pi_value(P, a.channel, a.width, a.k0, a.gamma, a.k, a.m1, a.m2, a.k0 + a.mu1 - a.mu2, a.A1 + a.A2, br, bi, re[i],
             im[i]);
}

__global__ void __launch_bounds__(256, 3) b0_kernel(Phys P, const double* tab, int64_t n, const double* args, int corr,
                                                 double* re, double* im)
{
    __shared__ double lds[kTabLen];
    load_tabs(lds, tab);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Tabs q = make_tabs(lds);
    if (!corr) {
        const double* a = args + i * 9;
        B0(P, q, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], re[i], im[i]);
    } else {
        const double* a = args + i * 10;
        B0_corr(P, q, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], re[i], im[i]);
    }
}

__global__ void __launch_bounds__(256) a_kernel(Phys P, int64_t n, const double* args, int np, const double* xp,
                                                const double* wp, int nc, const double* xc, const double* wc,
                                                double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int w = nc > 0 ? 6 : 5;
    const double* a = args + i * w;
    out[i] = A_value(P, a[0], a[1], a[2], a[3], a[4], nc > 0 ? a[5] : 0.0, np, xp, wp, nc, xc, wc);
}

// grid: A_aniso(m_u), A_aniso(m_s) per point -> A[pt*2 + f]
__global__ void __launch_bounds__(256) grid_a_kernel(Phys P, int npts, const pnjl_pol_point* pts, int np,
                                                     const double* xp, const double* wp, int nc, const double* xc,
                                                     const double* wc, double* A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * npts) return;
    const pnjl_pol_point s = pts[i >> 1];
    const double m = (i & 1) ? s.m_s : s.m_u;
    A[i] = A_value(P, m, s.mu, s.T, s.Phi, s.Phib, s.xi, np, xp, wp, nc, xc, wc);
}

// k extent of a wavefront's (k0, k) tile: 8 x 8 tiles keep the lanes of a
// wavefront closer in both momenta than a 1 x 64 row (measured on the 256 x
// 128 x 64 grid: 31.1 ms per replay against 42.7 ms for 1 x 64 rows, 31.1
// for 4 x 16, 32.1 for 16 x 4, 35.1 for 2 x 32).
#ifndef PNJL_GRID_TILE_K
#define PNJL_GRID_TILE_K 8
#endif
// grid: one thread per (point, pair, k0, k); both channels written.
// The quadrature kernels are held to 3 waves per SIMD (168 VGPRs, a few
// spills): latency hiding pays more than the spill traffic costs.
__global__ void __launch_bounds__(256, 3) grid_pi_kernel(Phys P, const double* tab, int npts, const pnjl_pol_point* pts,
                                                      const double* A, int nk0, const double* k0s, int nk,
                                                      const double* ks, double* re, double* im)
{
    __shared__ double lds[kTabLen];
    load_tabs(lds, tab);
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)nk0 * nk;
    if (tid >= (int64_t)npts * 3 * plane) return;
    const Tabs q = make_tabs(lds);
    int64_t j, i0;
    int pair, pt;
    constexpr int TK = PNJL_GRID_TILE_K, T0 = 64 / TK;
    if (TK < 64 && nk % TK == 0 && nk0 % T0 == 0) {
        // a wavefront takes a T0 (k0) x TK (k) tile of one (point, pair) plane
        const int64_t lane = tid & 63, tile = tid >> 6;
        const int64_t ntk = nk / TK, tiles = ntk * (nk0 / T0);
        const int64_t pp = tile / tiles, tl = tile - pp * tiles;
        j = (tl % ntk) * TK + lane % TK;
        i0 = (tl / ntk) * T0 + lane / TK;
        pair = (int)(pp % 3);
        pt = (int)(pp / 3);
    } else {
        j = tid % nk, i0 = (tid / nk) % nk0;
        pair = (int)((tid / plane) % 3);
        pt = (int)(tid / (3 * plane));
    }
    const pnjl_pol_point s = pts[pt];
    const double m1 = pair == 2 ? s.m_s : s.m_u, m2 = pair == 0 ? s.m_u : s.m_s;
    const double A12 = A[pt * 2 + (pair == 2 ? 1 : 0)] + A[pt * 2 + (pair == 0 ? 0 : 1)];
    const double k0 = k0s[i0], k = ks[j];
    double br, bi;
    b0_total(P, q, k0, k, m1, m2, s.mu, s.mu, s.T, s.Phi, s.Phib, s.xi, pair, br, bi);
    const int64_t base = ((int64_t)pt * 3 + pair) * 2 * plane + i0 * nk + j;
    for (int c = 0; c < 2; c++) {
        double r, m;
        pi_value(P, c == 0 ? PNJL_CHANNEL_P : PNJL_CHANNEL_S, 0, k0, 0.0, k, m1, m2, k0, A12, br, bi, r, m);
        re[base + c * plane] = r;
        im[base + c * plane] = m;
    }
}

// ---- consumers of Pi ---------------------------------------------------------------
__device__ __forceinline__ void store_kc(double* o, const meson::KCoef& c)
{
    o[0] = c.K0p, o[1] = c.K0m, o[2] = c.K123p, o[3] = c.K123m, o[4] = c.K4567p, o[5] = c.K4567m;
    o[6] = c.K8p, o[7] = c.K8m, o[8] = c.K08p, o[9] = c.K08m, o[10] = c.detKp, o[11] = c.detKm;
}

__device__ __forceinline__ meson::KCoef load_kc(const double* o)
{
    return meson::KCoef{o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11]};
}

__device__ __forceinline__ void store_c(double* o, meson::C v) { o[0] = v.r, o[1] = v.i; }

// in[i*4] = (A_u, A_s, m_u, m_s), or in[i*2] = (G_u, G_s) -> G_f[i*2] (optional), kc[i*12]
__global__ void __launch_bounds__(256) couplings_kernel(int64_t n, const double* in, int from_gf, double G, double K,
                                                        int Nc, double* Gf, double* kc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double Gu, Gs;
    if (from_gf) {
        Gu = in[i * 2], Gs = in[i * 2 + 1];
    } else {
        const double* a = in + i * 4;
        Gu = meson::G_from_A(a[0], a[2], Nc), Gs = meson::G_from_A(a[1], a[3], Nc);
    }
    if (Gf) Gf[i * 2] = Gu, Gf[i * 2 + 1] = Gs;
    store_kc(kc + i * 12, meson::couplings(G, K, Gu, Gs));
}

// kc[i*12], pi[i*12] (six slots) -> out[i*36] (layout of pnjl_meson_propagators_batch)
__global__ void __launch_bounds__(256) meson_kernel(int64_t n, const double* kc, const double* pi, double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using meson::C;
    const meson::KCoef k = load_kc(kc + i * 12);
    C P[6];
    for (int s = 0; s < 6; s++) P[s] = C{pi[i * 12 + 2 * s], pi[i * 12 + 2 * s + 1]};
    double* o = out + i * 36;
    store_c(o + 0, meson::simple(k.K123p, P[0]));
    store_c(o + 2, meson::simple(k.K4567p, P[2]));
    store_c(o + 4, meson::simple(k.K123m, P[1]));
    store_c(o + 6, meson::simple(k.K4567m, P[3]));
    double J[2][2];   // currents a (u ubar) and b (s sbar)
    meson::current(0, 0, J[0][0], J[0][1]);
    meson::current(2, 2, J[1][0], J[1][1]);
    for (int ch = 0; ch < 2; ch++) {
        C M00, M08, M88;
        const double detK = ch == 0 ? k.detKp : k.detKm;
        if (ch == 0)
            meson::mixing(P[0], P[4], k.K0p, k.K8p, k.K08p, k.detKp, M00, M08, M88);
        else
            meson::mixing(P[1], P[5], k.K0m, k.K8m, k.K08m, k.detKm, M00, M08, M88);
        store_c(o + 8 + 6 * ch, M00);
        store_c(o + 10 + 6 * ch, M08);
        store_c(o + 12 + 6 * ch, M88);
        for (int c = 0; c < 4; c++) {
            const double* Ja = J[c >> 1];
            const double* Jb = J[c & 1];
            store_c(o + 20 + 8 * ch + 2 * c, meson::mixed(detK, M00, M08, M88, Ja[0], Ja[1], Jb[0], Jb[1]));
        }
    }
}

// the meson mask of a query, or MS_NONE for a process / channel / mask outside the table
__device__ __host__ __forceinline__ unsigned query_mesons(const pnjl_prop_query& q, const meson::ProcRow* tab)
{
    if (q.process < 0 || q.process >= PNJL_PROC_COUNT || q.channel < 0 || q.channel > 2) return meson::MS_NONE;
    if (q.mesons) return (q.mesons & ~0x3f) ? meson::MS_NONE : (unsigned)q.mesons;
    return tab[q.process].ch[q.channel];
}

// ---- work order of the propagator batch ----------------------------------------
// The (pair, query) items the queries' mesons need are bucketed by what sets
// the branch structure of B0 -- pair, m1, |lambda| = |k0 + mu1 - mu2| and k --
// and evaluated bucket by bucket, so that the lanes of a wavefront take the
// same singularity branches (a counting sort: count, scan, scatter).  Items
// of no need are not evaluated at all.  The order inside a bucket follows the
// atomics and varies from call to call; each item's arithmetic does not.
constexpr int kBM = 8, kBL = 16, kBK = 16;
constexpr int kPropBuckets = 3 * kBM * kBL * kBK;

__device__ __forceinline__ int bin_of(double x, double scale, int nb)
{
    const double v = x * scale;
    if (!(v > 0.0)) return 0;   // negative or NaN
    return v >= (double)(nb - 1) ? nb - 1 : (int)v;
}

__device__ __forceinline__ unsigned item_need(const pnjl_prop_query& a, int pair)
{
    const unsigned mes = query_mesons(a, meson::kProcs);
    return mes == meson::MS_NONE ? 0u : (meson::slots_for(mes) >> (2 * pair)) & 3u;
}

__device__ __forceinline__ int item_bucket(const pnjl_prop_query& a, int pair)
{
    const double m1 = pair == 2 ? a.m_s : a.m_u;
    const double mu1 = pair == 2 ? a.mu_s : a.mu_u, mu2 = pair == 0 ? a.mu_u : a.mu_s;
    const int bm = bin_of(m1, 4.0, kBM), bl = bin_of(fabs(a.k0 + mu1 - mu2), 4.0, kBL), bk = bin_of(a.k, 5.0, kBK);
    return ((pair * kBM + bm) * kBL + bl) * kBK + bk;
}

__global__ void __launch_bounds__(256) zero_u32(unsigned* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// item t = pair * n + i: bucket counts; the Pi slots of items of no need are zeroed
__global__ void __launch_bounds__(256) prop_count_kernel(int64_t n, const pnjl_prop_query* qs, unsigned* counts,
                                                         double* pib)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int pair = (int)(t / n);
    const int64_t i = t - pair * n;
    const pnjl_prop_query a = qs[i];
    if (item_need(a, pair)) {
        atomicAdd(&counts[item_bucket(a, pair)], 1u);
    } else {
        double* o = pib + i * 12 + 4 * pair;
        o[0] = o[1] = o[2] = o[3] = 0.0;
    }
}

// exclusive scan of the bucket counts (one block of 256); offs[kPropBuckets] = total
__global__ void __launch_bounds__(256) prop_scan_kernel(const unsigned* counts, unsigned* offs)
{
    constexpr int per = kPropBuckets / 256;
    static_assert(kPropBuckets % 256 == 0, "bucket count");
    __shared__ unsigned part[256];
    const int b0 = threadIdx.x * per;
    unsigned s = 0;
    for (int j = 0; j < per; j++) s += counts[b0 + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {   // Hillis-Steele inclusive scan
        const unsigned v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = part[threadIdx.x] - s;
    for (int j = 0; j < per; j++) {
        offs[b0 + j] = run;
        run += counts[b0 + j];
    }
    if (threadIdx.x == 255) offs[kPropBuckets] = part[255];
}

__global__ void __launch_bounds__(256) prop_scatter_kernel(int64_t n, const pnjl_prop_query* qs, const unsigned* offs,
                                                           unsigned* cursor, unsigned* items)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int pair = (int)(t / n);
    const int64_t i = t - pair * n;
    const pnjl_prop_query a = qs[i];
    if (!item_need(a, pair)) return;
    const int b = item_bucket(a, pair);
    items[offs[b] + atomicAdd(&cursor[b], 1u)] = (unsigned)(i * 3 + pair);
}

// Pi of one worklist item (query i, pair): one B0 serves both channels of the
// pair.  pib[i*12 + 2*slot] as in pnjl_polarization_grid's order.  Threads
// past the worklist's length (offs[kPropBuckets]) exit; stride 64 puts one
// item per wavefront, for counting FLOPs.
__global__ void __launch_bounds__(256, 3) prop_pi_kernel(Phys P, const double* tab, const pnjl_prop_query* qs,
                                                      const unsigned* items, const unsigned* offs, double* pib,
                                                      int stride)
{
    __shared__ double lds[kTabLen];
    load_tabs(lds, tab);
    const int64_t ts = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ts % stride) return;
    const int64_t t = ts / stride;
    if (t >= (int64_t)offs[kPropBuckets]) return;
    const unsigned code = items[t];
    const int pair = (int)(code % 3u);
    const int64_t i = code / 3u;
    const pnjl_prop_query a = qs[i];
    const unsigned need = item_need(a, pair);
    double* o = pib + i * 12 + 4 * pair;
    const Tabs q = make_tabs(lds);
    const double m1 = pair == 2 ? a.m_s : a.m_u, m2 = pair == 0 ? a.m_u : a.m_s;
    const double mu1 = pair == 2 ? a.mu_s : a.mu_u, mu2 = pair == 0 ? a.mu_u : a.mu_s;
    const double A1 = pair == 2 ? a.A_s : a.A_u, A2 = pair == 0 ? a.A_u : a.A_s;
    double br, bi;
    b0_total(P, q, a.k0, a.k, m1, m2, mu1, mu2, a.T, a.Phi, a.Phib, a.xi, pair, br, bi);
    for (int c = 0; c < 2; c++) {
        if (need & (1u << c))
            pi_value(P, c == 0 ? PNJL_CHANNEL_P : PNJL_CHANNEL_S, 0, a.k0, 0.0, a.k, m1, m2, a.k0 + mu1 - mu2, A1 + A2,
                     br, bi, o[2 * c], o[2 * c + 1]);
        else
            o[2 * c] = o[2 * c + 1] = 0.0;
    }
}

// (D_S, D_P) per query from its Pi slots; kc == nullptr: couplings from A
__global__ void __launch_bounds__(256) prop_compose_kernel(int64_t n, const pnjl_prop_query* qs, const double* kc,
                                                           double G, double K, int Nc, const double* pib,
                                                           double* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using meson::C;
    const pnjl_prop_query a = qs[i];
    const unsigned mes = query_mesons(a, meson::kProcs);
    double* o = out + i * 4;
    if (mes == meson::MS_NONE) {
        o[0] = o[1] = o[2] = o[3] = __builtin_nan("");
        return;
    }
    const meson::KCoef k = kc ? load_kc(kc + i * 12)
                              : meson::couplings(G, K, meson::G_from_A(a.A_u, a.m_u, Nc),
                                                 meson::G_from_A(a.A_s, a.m_s, Nc));
    C pi[6];
    for (int s = 0; s < 6; s++) pi[s] = C{pib[i * 12 + 2 * s], pib[i * 12 + 2 * s + 1]};
    C DS, DP;
    meson::by_channel(meson::kProcs[a.process], a.channel, mes, k, pi, DS, DP);
    store_c(o, DS);
    store_c(o + 2, DP);
}

// ---- meson pole masses (MesonMass.jl) -------------------------------------------------
constexpr double kFdStep = 6.055454452393343e-06;   // cbrt(eps): FiniteDiff's central relstep = absstep

// principal square root
__device__ __forceinline__ meson::C csqrt_p(meson::C z)
{
    if (z.r == 0.0 && z.i == 0.0) return meson::C{0.0, z.i};
    const double t = sqrt((fabs(z.r) + hypot(z.r, z.i)) * 0.5);
    if (z.r >= 0.0) return meson::C{t, z.i / (2.0 * t)};
    return meson::C{fabs(z.i) / (2.0 * t), copysign(t, z.i)};
}

// The residual at (k0, gamma) needs B0 of one quark pair (two for the mixed
// mesons) at k0 only: gamma enters polarization_with_width algebraically.
// B0 is therefore computed once per k0 and reused for the gamma column of
// the finite-difference Jacobian (the same values a recomputation gives).
struct MassB0 {
    double r[2], i[2];
};

__device__ __forceinline__ void mass_pair(const pnjl_mass_query& a, int pair, double& m1, double& m2, double& mu1,
                                          double& mu2, double& A12)
{
    m1 = pair == 2 ? a.m_s : a.m_u, m2 = pair == 0 ? a.m_u : a.m_s;
    mu1 = pair == 2 ? a.mu_s : a.mu_u, mu2 = pair == 0 ? a.mu_u : a.mu_s;
    A12 = (pair == 2 ? a.A_s : a.A_u) + (pair == 0 ? a.A_u : a.A_s);
}

// the pairs of a meson: pi, sigma_pi -> uu; K, sigma_K -> us; mixed -> uu, ss
__device__ __forceinline__ int mass_pairs(int m, int* pairs)
{
    if (m >= PNJL_MESON_ID_ETA) {
        pairs[0] = 0, pairs[1] = 2;
        return 2;
    }
    pairs[0] = (m == PNJL_MESON_ID_K || m == PNJL_MESON_ID_SIGMA_K) ? 1 : 0;
    return 1;
}

// B0 (symmetrised for u sbar, + correction) of the meson's pairs at k0
//   MesonMass.jl:34-60, :222-231 (MesonMass passes num_s 0 for s sbar, equivalent to 2)
template <int L = 1>
__device__ __noinline__ void mass_b0(const Phys& P, const Tabs&, const pnjl_mass_query& a, double k0, MassB0& B)
{
    const Tabs T = make_tabs(mass_tab);
    int pairs[2];
    const int np = mass_pairs(a.meson, pairs);
#pragma unroll 1
    for (int j = 0; j < np; j++) {
        double m1, m2, mu1, mu2, A12;
        mass_pair(a, pairs[j], m1, m2, mu1, mu2, A12);
        b0_total<L>(P, T, k0, a.k, m1, m2, mu1, mu2, a.T, a.Phi, a.Phib, a.xi, pairs[j] == 1 ? 1 : 0, B.r[j], B.i[j]);
    }
}

// meson_mass_equation  MesonMass.jl:211-233 from the B0 at k0: 1 - 4 K Pi, or
// for the mixed mesons the light / heavy eigenvalue of 2M,
// (M00 + M88) -/+ sqrt((M00 - M88)^2 + 4 M08^2)
__device__ void mass_residual(const Phys& P, const pnjl_mass_query& a, const meson::KCoef& k, const MassB0& B,
                              double k0, double gamma, double* f)
{
    using meson::C;
    const int m = a.meson;
    int pairs[2];
    const int np = mass_pairs(m, pairs);
    const bool pch = m == PNJL_MESON_ID_PI || m == PNJL_MESON_ID_K || m == PNJL_MESON_ID_ETA ||
                     m == PNJL_MESON_ID_ETA_PRIME;
    const int ch = pch ? PNJL_CHANNEL_P : PNJL_CHANNEL_S;
    C Pi[2];
    for (int j = 0; j < np; j++) {
        double m1, m2, mu1, mu2, A12;
        mass_pair(a, pairs[j], m1, m2, mu1, mu2, A12);
        pi_value(P, ch, 1, k0, gamma, a.k, m1, m2, k0 + mu1 - mu2, A12, B.r[j], B.i[j], Pi[j].r, Pi[j].i);
    }
    C res;
    if (m < PNJL_MESON_ID_ETA) {
        const double K = m == PNJL_MESON_ID_PI ? k.K123p
                       : m == PNJL_MESON_ID_K  ? k.K4567p
                       : m == PNJL_MESON_ID_SIGMA_PI ? k.K123m : k.K4567m;
        const double k4 = 4.0 * K;
        res = C{1.0 - k4 * Pi[0].r, -(k4 * Pi[0].i)};
    } else {
        C M00, M08, M88;
        if (pch)
            meson::mixing(Pi[0], Pi[1], k.K0p, k.K8p, k.K08p, k.detKp, M00, M08, M88);
        else
            meson::mixing(Pi[0], Pi[1], k.K0m, k.K8m, k.K08m, k.detKm, M00, M08, M88);
        const C dm = meson::csub(M00, M88);
        const C sq = csqrt_p(meson::cadd(meson::cmul(dm, dm), meson::rmul(4.0, meson::cmul(M08, M08))));
        const C tr = meson::cadd(M00, M88);
        res = (m == PNJL_MESON_ID_ETA || m == PNJL_MESON_ID_SIGMA) ? meson::csub(tr, sq) : meson::cadd(tr, sq);
    }
    f[0] = res.r;
    f[1] = res.i;
}

// central finite differences (FiniteDiff): J[:, j] = (f(x + e_j h) - f(x - e_j h)) / 2h;
// Bx holds B0 at x[0] (the gamma column needs no new B0)
__device__ void mass_jac(const Phys& P, const Tabs& T, const pnjl_mass_query& a, const meson::KCoef& k,
                         const double* x, const MassB0& Bx, double J[2][2])
{
    for (int j = 0; j < 2; j++) {
        const double h = fmax(kFdStep * fabs(x[j]), kFdStep);
        double xp[2] = {x[0], x[1]}, xm[2] = {x[0], x[1]}, fp[2], fm[2];
        xp[j] = x[j] + h;
        xm[j] = x[j] - h;
        if (j == 0) {
            MassB0 Bp, Bm;
            mass_b0(P, T, a, xp[0], Bp);
            mass_b0(P, T, a, xm[0], Bm);
            mass_residual(P, a, k, Bp, xp[0], xp[1], fp);
            mass_residual(P, a, k, Bm, xm[0], xm[1], fm);
        } else {
            mass_residual(P, a, k, Bx, xp[0], xp[1], fp);
            mass_residual(P, a, k, Bx, xm[0], xm[1], fm);
        }
        for (int i = 0; i < 2; i++) J[i][j] = (fp[i] - fm[i]) / (2.0 * h);
    }
}

__device__ __forceinline__ double wnorm2(const double* d, const double* v)
{
    return sqrt((d[0] * v[0]) * (d[0] * v[0]) + (d[1] * v[1]) * (d[1] * v[1]));
}

// NLsolve dogleg! for 2 x 2 (a singular Jacobian takes the scaled steepest-descent branch)
__device__ void dogleg2(double* p, const double* r, const double* d, const double J[2][2], double delta)
{
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double pi_[2];
    if (det != 0.0 && isfinite(det)) {
        pi_[0] = -((J[1][1] * r[0] - J[0][1] * r[1]) / det);
        pi_[1] = -((-J[1][0] * r[0] + J[0][0] * r[1]) / det);
    } else {
        pi_[0] = pi_[1] = __builtin_inf();
    }
    if (wnorm2(d, pi_) <= delta) {
        p[0] = pi_[0], p[1] = pi_[1];
        return;
    }
    double g[2], Jg[2];
    for (int j = 0; j < 2; j++) g[j] = (J[0][j] * r[0] + J[1][j] * r[1]) / (d[j] * d[j]);
    for (int i = 0; i < 2; i++) Jg[i] = J[i][0] * g[0] + J[i][1] * g[1];
    const double wg = wnorm2(d, g);
    const double coef = -(wg * wg) / (Jg[0] * Jg[0] + Jg[1] * Jg[1]);
    const double pc[2] = {coef * g[0], coef * g[1]};
    if (wnorm2(d, pc) >= delta) {
        const double sc = -delta / wg;
        p[0] = g[0] * sc, p[1] = g[1] * sc;
        return;
    }
    const double pd[2] = {pi_[0] - pc[0], pi_[1] - pc[1]};
    const double b = 2.0 * ((d[0] * pc[0]) * (d[0] * pd[0]) + (d[1] * pc[1]) * (d[1] * pd[1]));
    const double wa = wnorm2(d, pd);
    const double aa = wa * wa;
    const double wpc = wnorm2(d, pc);
    const double tau = (-b + sqrt(b * b - 4.0 * aa * (wpc * wpc - delta * delta))) / (2.0 * aa);
    p[0] = pc[0] + tau * pd[0], p[1] = pc[1] + tau * pd[1];
}

// default_meson_mass_guess  MesonMass.jl:183-204 (d = u)
__device__ __forceinline__ double mass_guess(const pnjl_mass_query& a)
{
    switch (a.meson) {
        case PNJL_MESON_ID_PI: return a.m_u + a.m_u;
        case PNJL_MESON_ID_K:
        case PNJL_MESON_ID_SIGMA_K: return a.m_u + a.m_s;
        case PNJL_MESON_ID_ETA_PRIME:
        case PNJL_MESON_ID_SIGMA_PRIME: return 2.0 * a.m_s;
        default: return 2.0 * a.m_u;
    }
}

// Trust-region state of a pole search between the launches of
// pnjl_meson_mass_batch (every field it carries from one iteration to the next).
struct MassState {
    double x[2], r[2], J[2][2], d[2], delta;
    int it, conv, stopped, pad;
};

// solve_meson_mass  MesonMass.jl:241-276 with NLsolve's trust region (dogleg,
// autoscale, factor 1) and a central-difference Jacobian; one thread per query.
// out[i*6] = (mass, gamma, residual_norm (inf), converged, iterations, exception)
//
// The queries run in phases of growing iteration budgets (16, 16, 32, ...,
// then unlimited): phase 0 starts every query; a query that is not done when
// its phase's budget is spent saves its state and joins the list of the next
// phase.  The searches differ by orders of magnitude in length (most stop in
// a few dozen iterations, some run NLsolve's full 1000), so one launch would
// hold every wavefront for its longest query; the phases keep the long ones
// packed together.  A query's arithmetic is the same as in one uninterrupted
// loop.  stride 64 puts one query per wavefront, for counting FLOPs.
// 3 waves per SIMD (168 VGPRs, some spilled): the extra waves hide the
// quadratures' latency (1.5x over 256 VGPRs at 1 wave).
__global__ void __launch_bounds__(256, 3) mass_kernel(Phys P, const double* tab, int64_t n,
                                                   const pnjl_mass_query* qs, const double* kc, double G, double K,
                                                   int Nc, double* out, int stride, MassState* states,
                                                   const int32_t* in_list, const unsigned* in_count,
                                                   int32_t* out_list, unsigned* out_count, int budget)
{
    double* const lds = mass_tab;
    load_tabs(lds, tab);
    const int64_t ts = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ts % stride) return;
    const int64_t slot = ts / stride;
    const bool first = (in_list == nullptr);
    if (first ? slot >= n : slot >= (int64_t)*in_count) return;
    const int64_t i = first ? slot : (int64_t)in_list[slot];
    const Tabs T = make_tabs(lds);
    const pnjl_mass_query a = qs[i];
    double* o = out + i * 6;
    const meson::KCoef k = kc ? load_kc(kc + i * 12)
                              : meson::couplings(G, K, meson::G_from_A(a.A_u, a.m_u, Nc),
                                                 meson::G_from_A(a.A_s, a.m_s, Nc));
    double x[2], r[2], J[2][2], d[2], delta;
    int it;
    bool conv, stopped;
    if (first) {
        if (a.meson < 0 || a.meson >= PNJL_MESON_ID_COUNT || a.max_iter < 0) {
            for (int j = 0; j < 6; j++) o[j] = __builtin_nan("");
            return;
        }
        x[0] = isnan(a.mass0) ? mass_guess(a) : a.mass0;
        x[1] = a.gamma0;
        MassB0 B;
        mass_b0(P, T, a, x[0], B);
        mass_residual(P, a, k, B, x[0], x[1], r);
        if (!isfinite(r[0]) || !isfinite(r[1])) {   // NLsolve's IsFiniteException at the start
            o[0] = x[0], o[1] = x[1], o[2] = __builtin_nan(""), o[3] = 0.0, o[4] = 0.0, o[5] = 1.0;
            return;
        }
        mass_jac(P, T, a, k, x, B, J);
        it = 0;
        conv = fmax(fabs(r[0]), fabs(r[1])) < a.ftol;
        stopped = isnan(x[0]) || isnan(x[1]) || isnan(r[0]) || isnan(r[1]);
        d[0] = d[1] = delta = 0.0;
        if (!conv) {
            for (int j = 0; j < 2; j++) {
                d[j] = sqrt(J[0][j] * J[0][j] + J[1][j] * J[1][j]);
                if (d[j] == 0.0) d[j] = 1.0;
            }
            delta = 1.0 * wnorm2(d, x);
            if (delta == 0.0) delta = 1.0;
        }
    } else {
        const MassState& S = states[i];
        x[0] = S.x[0], x[1] = S.x[1], r[0] = S.r[0], r[1] = S.r[1];
        J[0][0] = S.J[0][0], J[0][1] = S.J[0][1], J[1][0] = S.J[1][0], J[1][1] = S.J[1][1];
        d[0] = S.d[0], d[1] = S.d[1], delta = S.delta;
        it = S.it, conv = S.conv != 0, stopped = S.stopped != 0;
    }
    int spent = 0;
    while (!stopped && !conv && it < a.max_iter && (budget < 0 || spent < budget)) {
        it++;
        spent++;
        double p[2], fn[2];
        dogleg2(p, r, d, J, delta);
        const double xold[2] = {x[0], x[1]};
        x[0] += p[0], x[1] += p[1];
        MassB0 Bn;
        mass_b0(P, T, a, x[0], Bn);
        mass_residual(P, a, k, Bn, x[0], x[1], fn);
        double rp[2];
        for (int q = 0; q < 2; q++) rp[q] = J[q][0] * p[0] + J[q][1] * p[1] + r[q];
        const double ssr = r[0] * r[0] + r[1] * r[1];
        const double rho = (ssr - (fn[0] * fn[0] + fn[1] * fn[1])) / (ssr - (rp[0] * rp[0] + rp[1] * rp[1]));
        if (rho > 1e-4) {
            r[0] = fn[0], r[1] = fn[1];
            mass_jac(P, T, a, k, x, Bn, J);
            for (int j = 0; j < 2; j++) d[j] = fmax(0.1 * d[j], sqrt(J[0][j] * J[0][j] + J[1][j] * J[1][j]));
            const bool xc = fmax(fabs(x[0] - xold[0]), fabs(x[1] - xold[1])) < a.xtol;
            conv = xc || fmax(fabs(r[0]), fabs(r[1])) < a.ftol;
        } else {
            x[0] -= p[0], x[1] -= p[1];
            conv = false;
        }
        if (rho < 0.1)
            delta = delta / 2.0;
        else if (rho >= 0.9)
            delta = 2.0 * wnorm2(d, p);
        else if (rho >= 0.5)
            delta = fmax(delta, 2.0 * wnorm2(d, p));
        stopped = isnan(x[0]) || isnan(x[1]) || isnan(fn[0]) || isnan(fn[1]);
    }
    if (!stopped && !conv && it < a.max_iter) {
        // budget spent: save the state for the next phase
        MassState& S = states[i];
        S.x[0] = x[0], S.x[1] = x[1], S.r[0] = r[0], S.r[1] = r[1];
        S.J[0][0] = J[0][0], S.J[0][1] = J[0][1], S.J[1][0] = J[1][0], S.J[1][1] = J[1][1];
        S.d[0] = d[0], S.d[1] = d[1], S.delta = delta;
        S.it = it, S.conv = 0, S.stopped = 0;
        out_list[atomicAdd(out_count, 1u)] = (int32_t)i;
        return;
    }
    o[0] = x[0], o[1] = x[1], o[2] = fmax(fabs(r[0]), fabs(r[1])), o[3] = conv ? 1.0 : 0.0, o[4] = (double)it;
    o[5] = 0.0;
}

// The long tail of the pole searches (the phases after the first two: a few
// percent of the queries, most of them running NLsolve's full 1000
// iterations) with 16 lanes per query, four sub-groups of four: every
// trust-region iteration needs B0 at the trial k0 and, when the step is
// accepted, at k0 +- h for the Jacobian's k0 column; sub-groups 0, 1, 2
// evaluate the three at once (sub-group 3 repeats sub-group 0), each with its
// quadrature rules split over its four lanes and summed in the one-thread
// order (ordered_sum), and the group exchanges them.  So an iteration costs a
// quarter of one B0 evaluation of latency instead of three evaluations, the
// phase runs 16x more lanes to hide it, and every lane runs the same control
// logic on the same values, bit for bit, as the one-thread loop of
// mass_kernel.  (Splitting a mixed meson's two quark pairs over sub-groups
// as well was slower: the pairs' masses differ, so their quadrature branches
// diverge within the wavefront.)  2 waves per SIMD: the 256 VGPRs spill less
// than 168 at 3 waves, 8 % faster.
constexpr int kMassLanes = 16;
constexpr int kMassQLanes = 4;   // quadrature lanes per B0 evaluation

__device__ __forceinline__ void shfl_b0(MassB0& dst, const MassB0& src, int lane)
{
    for (int j = 0; j < 2; j++) {
        dst.r[j] = __shfl(src.r[j], lane, kMassLanes);
        dst.i[j] = __shfl(src.i[j], lane, kMassLanes);
    }
}

__global__ void __launch_bounds__(256, 2) mass_tail_kernel(Phys P, const double* tab, const pnjl_mass_query* qs,
                                                        const double* kc, double G, double K, int Nc, double* out,
                                                        MassState* states, const int32_t* in_list,
                                                        const unsigned* in_count, int32_t* out_list,
                                                        unsigned* out_count, int budget)
{
    // blocks past the list leave before the table load (uniform per block)
    if ((int64_t)blockIdx.x * (blockDim.x / kMassLanes) >= (int64_t)*in_count) return;
    double* const lds = mass_tab;
    load_tabs(lds, tab);
    const int64_t ts = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t slot = ts / kMassLanes;
    const int gl = (int)(ts % kMassLanes);
    const int sub = gl / kMassQLanes;
    if (slot >= (int64_t)*in_count) return;   // whole groups leave together
    const int64_t i = (int64_t)in_list[slot];
    const Tabs T = make_tabs(lds);
    const pnjl_mass_query a = qs[i];
    const meson::KCoef k = kc ? load_kc(kc + i * 12)
                              : meson::couplings(G, K, meson::G_from_A(a.A_u, a.m_u, Nc),
                                                 meson::G_from_A(a.A_s, a.m_s, Nc));
    const MassState& S = states[i];
    double x[2] = {S.x[0], S.x[1]}, r[2] = {S.r[0], S.r[1]};
    double J[2][2] = {{S.J[0][0], S.J[0][1]}, {S.J[1][0], S.J[1][1]}};
    double d[2] = {S.d[0], S.d[1]}, delta = S.delta;
    int it = S.it;
    bool conv = S.conv != 0, stopped = S.stopped != 0;
    int spent = 0;
    while (!stopped && !conv && it < a.max_iter && (budget < 0 || spent < budget)) {
        it++;
        spent++;
        double p[2], fn[2];
        dogleg2(p, r, d, J, delta);
        const double xold[2] = {x[0], x[1]};
        x[0] += p[0], x[1] += p[1];
        // B0 at the trial k0 (sub-groups 0, 3) and at k0 +- h (sub-groups 1, 2), h as mass_jac's k0 column
        const double h = fmax(kFdStep * fabs(x[0]), kFdStep);
        const double k0 = sub == 1 ? x[0] + h : (sub == 2 ? x[0] - h : x[0]);
        MassB0 Bl, Bn, Bp, Bm;
        mass_b0<kMassQLanes>(P, T, a, k0, Bl);
        shfl_b0(Bn, Bl, 0);   // lanes of the query's group (__shfl's lane is relative to its width)
        shfl_b0(Bp, Bl, kMassQLanes);
        shfl_b0(Bm, Bl, 2 * kMassQLanes);
        mass_residual(P, a, k, Bn, x[0], x[1], fn);
        double rp[2];
        for (int q = 0; q < 2; q++) rp[q] = J[q][0] * p[0] + J[q][1] * p[1] + r[q];
        const double ssr = r[0] * r[0] + r[1] * r[1];
        const double rho = (ssr - (fn[0] * fn[0] + fn[1] * fn[1])) / (ssr - (rp[0] * rp[0] + rp[1] * rp[1]));
        if (rho > 1e-4) {
            r[0] = fn[0], r[1] = fn[1];
            // mass_jac with its three B0 sets already at hand
            for (int j = 0; j < 2; j++) {
                const double hj = fmax(kFdStep * fabs(x[j]), kFdStep);
                double xp[2] = {x[0], x[1]}, xm[2] = {x[0], x[1]}, fp[2], fm[2];
                xp[j] = x[j] + hj;
                xm[j] = x[j] - hj;
                if (j == 0) {
                    mass_residual(P, a, k, Bp, xp[0], xp[1], fp);
                    mass_residual(P, a, k, Bm, xm[0], xm[1], fm);
                } else {
                    mass_residual(P, a, k, Bn, xp[0], xp[1], fp);
                    mass_residual(P, a, k, Bn, xm[0], xm[1], fm);
                }
                for (int q = 0; q < 2; q++) J[q][j] = (fp[q] - fm[q]) / (2.0 * hj);
            }
            for (int j = 0; j < 2; j++) d[j] = fmax(0.1 * d[j], sqrt(J[0][j] * J[0][j] + J[1][j] * J[1][j]));
            const bool xc = fmax(fabs(x[0] - xold[0]), fabs(x[1] - xold[1])) < a.xtol;
            conv = xc || fmax(fabs(r[0]), fabs(r[1])) < a.ftol;
        } else {
            x[0] -= p[0], x[1] -= p[1];
            conv = false;
        }
        if (rho < 0.1)
            delta = delta / 2.0;
        else if (rho >= 0.9)
            delta = 2.0 * wnorm2(d, p);
        else if (rho >= 0.5)
            delta = fmax(delta, 2.0 * wnorm2(d, p));
        stopped = isnan(x[0]) || isnan(x[1]) || isnan(fn[0]) || isnan(fn[1]);
    }
    if (gl != 0) return;
    if (!stopped && !conv && it < a.max_iter) {
        MassState& W = states[i];
        W.x[0] = x[0], W.x[1] = x[1], W.r[0] = r[0], W.r[1] = r[1];
        W.J[0][0] = J[0][0], W.J[0][1] = J[0][1], W.J[1][0] = J[1][0], W.J[1][1] = J[1][1];
        W.d[0] = d[0], W.d[1] = d[1], W.delta = delta;
        W.it = it, W.conv = 0, W.stopped = 0;
        out_list[atomicAdd(out_count, 1u)] = (int32_t)i;
        return;
    }
    double* o = out + i * 6;
    o[0] = x[0], o[1] = x[1], o[2] = fmax(fabs(r[0]), fabs(r[1])), o[3] = conv ? 1.0 : 0.0, o[4] = (double)it;
    o[5] = 0.0;
}

}  // namespace polar

// =============================================================================
// host side
// =============================================================================
struct PolarGraph {
    int npts = 0, nk0 = 0, nk = 0, np = 0, nc = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    pnjl_pol_point* pts = nullptr;
    double *k0 = nullptr, *k = nullptr, *A = nullptr, *re = nullptr, *im = nullptr, *nodes = nullptr;
};

namespace pnjl_host {

void polar_release(pnjl_ctx* c)
{
    if (c->pgraph) {
        PolarGraph* g = c->pgraph;
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        for (void* p : {(void*)g->pts, (void*)g->k0, (void*)g->k, (void*)g->A, (void*)g->re, (void*)g->im,
                        (void*)g->nodes})
            if (p) (void)hipFree(p);
        if (g->cap) (void)hipStreamDestroy(g->cap);
        delete g;
        c->pgraph = nullptr;
    }
    if (c->polar_tab) (void)hipFree(c->polar_tab);
    c->polar_tab = nullptr;
}

}  // namespace pnjl_host

namespace {

polar::Phys phys(const pnjl_ctx* c) { return polar::Phys{c->params.Lambda, (double)c->params.N_color}; }

// x16 w16 x32 w32 | power-left 16/32 (t, t') | power-right 16/32 | DE 33 (phi, phi')
//   IntervalQuadratureStrategies.jl:56-85
int polar_tab(pnjl_ctx* ctx, double** out)
{
    if (!ctx->polar_tab) {
        std::vector<double> t(polar::kTabLen), x, w;
        double* o = t.data();
        std::vector<double> x16, w16, x32, w32;
        gl_rule(16, x16, w16);
        gl_rule(32, x32, w32);
        auto put = [&](const std::vector<double>& v) {
            memcpy(o, v.data(), v.size() * sizeof(double));
            o += v.size();
        };
        put(x16);
        put(w16);
        put(x32);
        put(w32);
        const double ia = 1.0 / polar::kAlpha;
        for (int side = 0; side < 2; side++)
            for (const std::vector<double>* xs : {&x16, &x32}) {
                const size_t n = xs->size();
                std::vector<double> tt(n), tp(n);
                for (size_t i = 0; i < n; i++) {
                    double u = ((*xs)[i] + 1) / 2;
                    if (side == 0) {
                        tt[i] = pow(u, ia);
                        tp[i] = ia * pow(u, ia - 1) / 2;
                    } else {
                        tt[i] = 1 - pow(1 - u, ia);
                        tp[i] = ia * pow(1 - u, ia - 1) / 2;
                    }
                }
                put(tt);
                put(tp);
            }
        std::vector<double> phi(33), phip(33);
        for (int k = -16; k <= 16; k++) {
            double tk = k * polar::kDeH;
            double s = polar::kPi / 2 * sinh(tk);
            double sech = 1.0 / cosh(s);
            phi[k + 16] = tanh(s);
            phip[k + 16] = polar::kDeH * (polar::kPi / 2) * cosh(tk) * (sech * sech);
        }
        put(phi);
        put(phip);
        HIPCHK(ctx, hipMalloc(&ctx->polar_tab, t.size() * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(ctx->polar_tab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    *out = ctx->polar_tab;
    return PNJL_OK;
}

int blocks_for(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

extern "C" {

int pnjl_polarization_batch(pnjl_ctx* ctx, int64_t n, const pnjl_pol_query* q, double* re, double* im, int flags,
                            void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!q || !re || !im)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_polarization_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    if (!dev)
        for (int64_t i = 0; i < n; i++)
            if (q[i].channel != PNJL_CHANNEL_P && q[i].channel != PNJL_CHANNEL_S)
                return set_err(ctx, PNJL_ERR_ARG, "pnjl_polarization_batch: unsupported channel (use P or S)");
    double* tab = nullptr;
    if (int rc = polar_tab(ctx, &tab)) return rc;
    const pnjl_pol_query* dq = q;
    double *dre = re, *dim = im;
    if (!dev) {
        void *a, *b, *c;
        if (int rc = stage_buf(ctx, 60, (size_t)n * sizeof(pnjl_pol_query), &a)) return rc;
        if (int rc = stage_buf(ctx, 61, (size_t)n * sizeof(double), &b)) return rc;
        if (int rc = stage_buf(ctx, 62, (size_t)n * sizeof(double), &c)) return rc;
        HIPCHK(ctx, hipMemcpyAsync(a, q, (size_t)n * sizeof(pnjl_pol_query), hipMemcpyHostToDevice, st));
        dq = (const pnjl_pol_query*)a;
        dre = (double*)b;
        dim = (double*)c;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    const int stride = (flags & PNJL_FLAG_ONE_PER_WAVE) ? 64 : 1;
    hipLaunchKernelGGL(polar::query_kernel, dim3(blocks_for(n * stride)), dim3(256), 0, st, phys(ctx), tab, n, dq, dre,
                       dim, stride);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!dev) {
        HIPCHK(ctx, hipMemcpyAsync(re, dre, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(im, dim, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

static int b0_common(pnjl_ctx* ctx, int64_t n, const double* args, int corr, double* re, double* im)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!args || !re || !im))) return set_err(ctx, PNJL_ERR_ARG, "B0 batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    double* tab = nullptr;
    if (int rc = polar_tab(ctx, &tab)) return rc;
    const int w = corr ? 10 : 9;
    void *a, *b, *c;
    if (int rc = stage_buf(ctx, 60, (size_t)n * w * sizeof(double), &a)) return rc;
    if (int rc = stage_buf(ctx, 61, (size_t)n * sizeof(double), &b)) return rc;
    if (int rc = stage_buf(ctx, 62, (size_t)n * sizeof(double), &c)) return rc;
    HIPCHK(ctx, hipMemcpy(a, args, (size_t)n * w * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(polar::b0_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, phys(ctx), tab, n, (const double*)a,
                       corr, (double*)b, (double*)c);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpy(re, b, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(im, c, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_B0_batch(pnjl_ctx* ctx, int64_t n, const double* args, double* re, double* im)
{
    return b0_common(ctx, n, args, 0, re, im);
}

int pnjl_B0_correction_batch(pnjl_ctx* ctx, int64_t n, const double* args, double* re, double* im)
{
    return b0_common(ctx, n, args, 1, re, im);
}

int pnjl_A_batch(pnjl_ctx* ctx, int64_t n, const double* args, int np, const double* nodes_p, const double* weights_p,
                 int nc, const double* nodes_cos, const double* weights_cos, double* out)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || np <= 0 || np > 65536 || nc < 0 || nc > 4096 || (n > 0 && (!args || !out)) || !nodes_p ||
        !weights_p || (nc > 0 && (!nodes_cos || !weights_cos)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_A_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int w = nc > 0 ? 6 : 5;
    void *a, *nd, *o;
    if (int rc = stage_buf(ctx, 60, (size_t)n * w * sizeof(double), &a)) return rc;
    if (int rc = stage_buf(ctx, 61, (size_t)(2 * np + 2 * nc) * sizeof(double), &nd)) return rc;
    if (int rc = stage_buf(ctx, 62, (size_t)n * sizeof(double), &o)) return rc;
    double* d = (double*)nd;
    HIPCHK(ctx, hipMemcpy(a, args, (size_t)n * w * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(d, nodes_p, np * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(d + np, weights_p, np * sizeof(double), hipMemcpyHostToDevice));
    if (nc > 0) {
        HIPCHK(ctx, hipMemcpy(d + 2 * np, nodes_cos, nc * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(ctx, hipMemcpy(d + 2 * np + nc, weights_cos, nc * sizeof(double), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(polar::a_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, phys(ctx), n, (const double*)a, np, d,
                       d + np, nc, d + 2 * np, d + 2 * np + nc, (double*)o);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpy(out, o, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_polarization_grid(pnjl_ctx* ctx, int npts, const pnjl_pol_point* pts, int nk0, const double* k0, int nk,
                           const double* k, int np, int nc, double* re, double* im, int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (npts <= 0 || nk0 <= 0 || nk <= 0 || np <= 0 || np > 4096 || nc <= 0 || nc > 64 || !pts || !k0 || !k ||
        !re || !im)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_polarization_grid: bad argument");
    const int64_t nout = (int64_t)npts * 6 * nk0 * nk;
    if (nout > ((int64_t)1 << 31)) return set_err(ctx, PNJL_ERR_ARG, "pnjl_polarization_grid: grid too large");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    double* tab = nullptr;
    if (int rc = polar_tab(ctx, &tab)) return rc;
    PolarGraph* g = ctx->pgraph;
    if (g && (g->npts != npts || g->nk0 != nk0 || g->nk != nk || g->np != np || g->nc != nc)) {
        polar_release(ctx);   // a new shape: rebuild the graph (tables are rebuilt too)
        if (int rc = polar_tab(ctx, &tab)) return rc;
        g = nullptr;
    }
    if (!g) {
        g = new PolarGraph();
        ctx->pgraph = g;
        g->npts = npts;
        g->nk0 = nk0;
        g->nk = nk;
        g->np = np;
        g->nc = nc;
        HIPCHK(ctx, hipStreamCreate(&g->cap));
        HIPCHK(ctx, hipMalloc(&g->pts, (size_t)npts * sizeof(pnjl_pol_point)));
        HIPCHK(ctx, hipMalloc(&g->k0, (size_t)nk0 * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&g->k, (size_t)nk * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&g->A, (size_t)npts * 2 * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&g->re, (size_t)nout * sizeof(double)));
        HIPCHK(ctx, hipMalloc(&g->im, (size_t)nout * sizeof(double)));
        // momentum nodes GL on [0, 10] fm^-1, cos nodes GL on [-1, 1]
        std::vector<double> xp, wp, xc, wc;
        gl_rule(np, xp, wp);
        for (int i = 0; i < np; i++) {
            xp[i] = 5.0 * xp[i] + 5.0;
            wp[i] = 5.0 * wp[i];
        }
        gl_rule(nc, xc, wc);
        std::vector<double> nodes;
        nodes.insert(nodes.end(), xp.begin(), xp.end());
        nodes.insert(nodes.end(), wp.begin(), wp.end());
        nodes.insert(nodes.end(), xc.begin(), xc.end());
        nodes.insert(nodes.end(), wc.begin(), wc.end());
        HIPCHK(ctx, hipMalloc(&g->nodes, nodes.size() * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(g->nodes, nodes.data(), nodes.size() * sizeof(double), hipMemcpyHostToDevice));
        // capture the A and Pi kernels once; later calls replay the graph
        HIPCHK(ctx, hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(polar::grid_a_kernel, dim3(blocks_for(2 * (int64_t)npts)), dim3(256), 0, g->cap,
                           phys(ctx), npts, (const pnjl_pol_point*)g->pts, np, g->nodes, g->nodes + np, nc,
                           g->nodes + 2 * np, g->nodes + 2 * np + nc, g->A);
        hipLaunchKernelGGL(polar::grid_pi_kernel, dim3(blocks_for(nout / 2)), dim3(256), 0, g->cap, phys(ctx), tab,
                           npts, (const pnjl_pol_point*)g->pts, (const double*)g->A, nk0, (const double*)g->k0, nk,
                           (const double*)g->k, g->re, g->im);
        HIPCHK(ctx, hipStreamEndCapture(g->cap, &g->graph));
        HIPCHK(ctx, hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    }
    hipStream_t st = stream ? (hipStream_t)stream : g->cap;
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIPCHK(ctx, hipMemcpyAsync(g->pts, pts, (size_t)npts * sizeof(pnjl_pol_point), in, st));
    HIPCHK(ctx, hipMemcpyAsync(g->k0, k0, (size_t)nk0 * sizeof(double), in, st));
    HIPCHK(ctx, hipMemcpyAsync(g->k, k, (size_t)nk * sizeof(double), in, st));
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    HIPCHK(ctx, hipGraphLaunch(g->exec, st));
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    HIPCHK(ctx, hipMemcpyAsync(re, g->re, (size_t)nout * sizeof(double), outk, st));
    HIPCHK(ctx, hipMemcpyAsync(im, g->im, (size_t)nout * sizeof(double), outk, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

int pnjl_effective_couplings_batch(pnjl_ctx* ctx, int64_t n, const double* in, int from_gf, double G, double K,
                                   double* G_f, double* kc)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!in || !kc)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_effective_couplings_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int w = from_gf ? 2 : 4;
    void *a, *g, *o;
    if (int rc = stage_buf(ctx, 80, (size_t)n * w * sizeof(double), &a)) return rc;
    if (int rc = stage_buf(ctx, 81, (size_t)n * 2 * sizeof(double), &g)) return rc;
    if (int rc = stage_buf(ctx, 82, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), &o)) return rc;
    HIPCHK(ctx, hipMemcpy(a, in, (size_t)n * w * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(polar::couplings_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, n, (const double*)a,
                       from_gf ? 1 : 0, G, K, ctx->params.N_color, (double*)g, (double*)o);
    HIPCHK(ctx, hipGetLastError());
    if (G_f) HIPCHK(ctx, hipMemcpy(G_f, g, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(kc, o, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_meson_propagators_batch(pnjl_ctx* ctx, int64_t n, const double* kc, const double* pi, double* out)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!kc || !pi || !out)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_meson_propagators_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    void *k, *p, *o;
    if (int rc = stage_buf(ctx, 80, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), &k)) return rc;
    if (int rc = stage_buf(ctx, 81, (size_t)n * 12 * sizeof(double), &p)) return rc;
    if (int rc = stage_buf(ctx, 82, (size_t)n * PNJL_MESON_PROP_FIELDS * sizeof(double), &o)) return rc;
    HIPCHK(ctx, hipMemcpy(k, kc, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(p, pi, (size_t)n * 12 * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(polar::meson_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, n, (const double*)k,
                       (const double*)p, (double*)o);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpy(out, o, (size_t)n * PNJL_MESON_PROP_FIELDS * sizeof(double), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_propagator_by_channel_batch(pnjl_ctx* ctx, int64_t n, const pnjl_prop_query* q, const double* kc,
                                     double* out, int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!q || !out)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_propagator_by_channel_batch: bad argument");
    if (n == 0) return PNJL_OK;
    if (n > ((int64_t)1 << 29) || ((flags & PNJL_FLAG_ONE_PER_WAVE) && n > ((int64_t)1 << 23)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_propagator_by_channel_batch: n too large");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    if (!dev)
        for (int64_t i = 0; i < n; i++)
            if (polar::query_mesons(q[i], meson::kProcsHost) == meson::MS_NONE)
                return set_err(ctx, PNJL_ERR_ARG,
                               "pnjl_propagator_by_channel_batch: process, channel or meson mask not in the table");
    double* tab = nullptr;
    if (int rc = polar_tab(ctx, &tab)) return rc;
    void *pib, *ints, *items;
    if (int rc = stage_buf(ctx, 83, (size_t)n * 12 * sizeof(double), &pib)) return rc;
    const int nb = polar::kPropBuckets;
    if (int rc = stage_buf(ctx, 84, (size_t)(3 * nb + 1) * sizeof(unsigned), &ints)) return rc;
    if (int rc = stage_buf(ctx, 85, (size_t)3 * n * sizeof(unsigned), &items)) return rc;
    unsigned* counts = (unsigned*)ints;   // counts[nb] | cursor[nb] | offs[nb + 1]
    unsigned* cursor = counts + nb;
    unsigned* offs = cursor + nb;
    const pnjl_prop_query* dq = q;
    const double* dkc = kc;
    double* dout = out;
    if (!dev) {
        void *a, *b, *c = nullptr;
        if (int rc = stage_buf(ctx, 80, (size_t)n * sizeof(pnjl_prop_query), &a)) return rc;
        if (int rc = stage_buf(ctx, 81, (size_t)n * 4 * sizeof(double), &b)) return rc;
        if (kc)
            if (int rc = stage_buf(ctx, 82, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), &c)) return rc;
        HIPCHK(ctx, hipMemcpyAsync(a, q, (size_t)n * sizeof(pnjl_prop_query), hipMemcpyHostToDevice, st));
        if (kc)
            HIPCHK(ctx, hipMemcpyAsync(c, kc, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), hipMemcpyHostToDevice,
                                       st));
        dq = (const pnjl_prop_query*)a;
        dkc = (const double*)c;
        dout = (double*)b;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    const int stride = (flags & PNJL_FLAG_ONE_PER_WAVE) ? 64 : 1;
    hipLaunchKernelGGL(polar::zero_u32, dim3(blocks_for(2 * nb)), dim3(256), 0, st, counts, 2 * nb);
    hipLaunchKernelGGL(polar::prop_count_kernel, dim3(blocks_for(3 * n)), dim3(256), 0, st, n, dq, counts,
                       (double*)pib);
    hipLaunchKernelGGL(polar::prop_scan_kernel, dim3(1), dim3(256), 0, st, (const unsigned*)counts, offs);
    hipLaunchKernelGGL(polar::prop_scatter_kernel, dim3(blocks_for(3 * n)), dim3(256), 0, st, n, dq,
                       (const unsigned*)offs, cursor, (unsigned*)items);
    hipLaunchKernelGGL(polar::prop_pi_kernel, dim3(blocks_for(3 * n * stride)), dim3(256), 0, st, phys(ctx), tab, dq,
                       (const unsigned*)items, (const unsigned*)offs, (double*)pib, stride);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(polar::prop_compose_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, dq, dkc, ctx->params.G,
                       ctx->params.K, ctx->params.N_color, (const double*)pib, dout);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!dev) HIPCHK(ctx, hipMemcpyAsync(out, dout, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

int pnjl_meson_mass_batch(pnjl_ctx* ctx, int64_t n, const pnjl_mass_query* q, const double* kc, double* out,
                          int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!q || !out)) || n > ((int64_t)1 << 28) ||
        ((flags & PNJL_FLAG_ONE_PER_WAVE) && n > ((int64_t)1 << 22)))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_meson_mass_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    if (!dev)
        for (int64_t i = 0; i < n; i++)
            if (q[i].meson < 0 || q[i].meson >= PNJL_MESON_ID_COUNT || q[i].max_iter < 0)
                return set_err(ctx, PNJL_ERR_ARG, "pnjl_meson_mass_batch: unknown meson or negative max_iter");
    double* tab = nullptr;
    if (int rc = polar_tab(ctx, &tab)) return rc;
    const pnjl_mass_query* dq = q;
    const double* dkc = kc;
    double* dout = out;
    if (!dev) {
        void *a, *b, *c = nullptr;
        if (int rc = stage_buf(ctx, 86, (size_t)n * sizeof(pnjl_mass_query), &a)) return rc;
        if (int rc = stage_buf(ctx, 87, (size_t)n * 6 * sizeof(double), &b)) return rc;
        if (kc)
            if (int rc = stage_buf(ctx, 88, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), &c)) return rc;
        HIPCHK(ctx, hipMemcpyAsync(a, q, (size_t)n * sizeof(pnjl_mass_query), hipMemcpyHostToDevice, st));
        if (kc)
            HIPCHK(ctx, hipMemcpyAsync(c, kc, (size_t)n * PNJL_KCOEF_FIELDS * sizeof(double), hipMemcpyHostToDevice,
                                       st));
        dq = (const pnjl_mass_query*)a;
        dkc = (const double*)c;
        dout = (double*)b;
    }
    // phases of growing iteration budgets; the last one runs every remaining
    // query to its end (max_iter)
    constexpr int kPhases = 8;
    const int budgets[kPhases] = {16, 16, 32, 64, 128, 256, 512, -1};
    void *vs = nullptr, *vl0 = nullptr, *vl1 = nullptr, *vc = nullptr;
    if (int rc = stage_buf(ctx, 89, (size_t)n * sizeof(polar::MassState), &vs)) return rc;
    if (int rc = stage_buf(ctx, 90, (size_t)n * sizeof(int32_t), &vl0)) return rc;
    if (int rc = stage_buf(ctx, 91, (size_t)n * sizeof(int32_t), &vl1)) return rc;
    if (int rc = stage_buf(ctx, 92, kPhases * sizeof(unsigned), &vc)) return rc;
    unsigned* cnt = (unsigned*)vc;
    int32_t* lists[2] = {(int32_t*)vl0, (int32_t*)vl1};
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(polar::zero_u32, dim3(1), dim3(64), 0, st, cnt, kPhases);
    const int stride = (flags & PNJL_FLAG_ONE_PER_WAVE) ? 64 : 1;
    // the tail phases (from kTailPhase on: few queries, hundreds of
    // iterations each) run kMassLanes lanes per query; the
    // one-query-per-wavefront FLOP count keeps one thread per query throughout
    constexpr int kTailPhase = 2;
    for (int ph = 0; ph < kPhases; ph++) {
        const int32_t* in_list = ph == 0 ? nullptr : lists[(ph - 1) & 1];
        const unsigned* in_count = ph == 0 ? nullptr : cnt + ph - 1;
        if (ph >= kTailPhase && stride == 1)
            hipLaunchKernelGGL(polar::mass_tail_kernel, dim3(blocks_for(n * polar::kMassLanes)), dim3(256), 0, st,
                               phys(ctx), tab, dq, dkc, ctx->params.G, ctx->params.K, ctx->params.N_color, dout,
                               (polar::MassState*)vs, in_list, in_count, lists[ph & 1], cnt + ph, budgets[ph]);
        else
            hipLaunchKernelGGL(polar::mass_kernel, dim3(blocks_for(n * stride)), dim3(256), 0, st, phys(ctx), tab, n,
                               dq, dkc, ctx->params.G, ctx->params.K, ctx->params.N_color, dout, stride,
                               (polar::MassState*)vs, in_list, in_count, lists[ph & 1], cnt + ph, budgets[ph]);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!dev) HIPCHK(ctx, hipMemcpyAsync(out, dout, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

}  // extern "C"
