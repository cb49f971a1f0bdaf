// pnjl_device.h -- device-side physics of the PNJL FixedMu gap equations for
// gfx950 (wave64).  FP64 throughout.
//
// Omega(x) = chi(phi) + U(T,Phi,Phib) + E_vac(M) + L_th(M,Phi,Phib)
//   (src/pnjl/core/Thermodynamics.jl:181-195), x = (phi_u, phi_d, phi_s, Phi, Phib).
// The node integrand is the Polyakov-loop log term of calculate_log_term
// (src/pnjl/core/Integrals.jl:193-238) written for the GPU:
//   * exp(-|a|) per node; exp(-|b|) = exp(-|a|) exp(-2 mu/T) whenever both
//     exponents are non-positive (E >= |mu|), so one exp serves both signs;
//     both LSE branches are folded into a select over normalised powers
//     (n0..n3), so lanes never diverge on the branch;
//   * E and 1/E from one reciprocal square root, 1/(D+ D-) from one
//     reciprocal, each refined by Newton-type steps (about 1 ulp);
//   * ln F+ + ln F- as ONE log of the product D+ D- (both factors lie in
//     [1, 8] for physical Phi, Phib) and one reciprocal of that product;
//   * first and second derivatives in (M, Phi, Phib) from the same ratios
//     r1 = n1/D, r2 = n2/D, r3 = n3/D (the reference obtains them by nested
//     forward-mode AD; these are their closed forms);
//   * the 1e-16 floor of the reference (max(term, POLYAKOV_EPS)) is honoured
//     on a rare per-lane slow path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pnjl {

constexpr double kPi = 3.14159265358979323846;
constexpr double kPolyakovEps = 1e-16;
constexpr double kLogPolyakovEps = -36.841361487904734;   // ln(1e-16)

// Math tables at the start of every block's LDS (filled from the device node
// table, built on the host in extended precision):
//   [0, 64)    2^(j/64)                               (exp_tab)
//   [64, 192)  pairs (1/c_j, log c_j), c_j = 1 + (j + 1/2)/64   (log_tab)
constexpr int kExpTab = 0;
constexpr int kLogTab = 64;
constexpr int kMathTab = 192;

struct DevParams {
    double Lambda, G, K, m_ud0, m_s0, T0, a0, a1, a2, b3, rho0;
    double Nc;
};

// Per-flavour node sums (13 accumulators).  With C the node coefficient
// w_p w_t p^2/(2pi)^2, Q = lnF+(a) + lnF-(b), E the node energy:
//   S    = sum C Q
//   A1   = sum C Q_E / E                    -> dS/dM     = M A1
//   A2   = sum C (Q_EE - Q_E/E) / E^2       -> d2S/dM2   = A1 + M^2 A2
//   SP   = sum C Q_Phi,  SPb = sum C Q_Phib
//   B1   = sum C Q_EPhi / E                 -> d2S/dMdPhi  = M B1
//   B2   = sum C Q_EPhib / E                -> d2S/dMdPhib = M B2
//   SPP, SPPb, SPbPb = second derivatives in (Phi, Phib)
//   Sfp  = sum C f+,  Sfm = sum C f-       (distributions -> densities)
//   Sfab = sum C (f+ a + f- b)             (entropy)
struct FlavSums {
    double S, A1, A2, SP, SPb, B1, B2, SPP, SPPb, SPbPb, Sfp, Sfm, Sfab;
};

__device__ __forceinline__ void flav_zero(FlavSums& s)
{
    s.S = s.A1 = s.A2 = s.SP = s.SPb = s.B1 = s.B2 = 0.0;
    s.SPP = s.SPPb = s.SPbPb = s.Sfp = s.Sfm = s.Sfab = 0.0;
}

// 1/sqrt(x) and sqrt(x) for x in the normal range from the hardware
// estimate (v_rsq_f64, ~23 bits) and two coupled Goldschmidt steps
// (g -> sqrt(x), h -> 1/(2 sqrt(x)); Markstein, IA-64 and Elementary
// Functions, 2000): both within about 1 ulp, with no scaling or special-case
// handling -- the node energies E^2 = p^2 + M^2 are positive and normal.
__device__ __forceinline__ void rsqrt_sqrt(double x, double& rs, double& sq)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    sq = g;
    rs = 2.0 * h;
}

// 1/d for d in the normal range: v_rcp_f64 estimate and two Newton-Raphson
// steps r <- r + r (1 - d r), within about 1 ulp.
__device__ __forceinline__ double rcp_nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// exp(x) for -800 <= x <= 0 after clamping (Tang's table-driven method, ACM
// TOMS 15(2), 1989): x = (64 k + j) ln2/64 + r with |r| <= ln2/128 (Cody-Waite
// reduction with a split ln2/64), exp(x) = 2^k 2^(j/64) e^r and e^r from its
// degree-5 Taylor polynomial (truncation below 4e-17).  About 1 ulp; results
// below the double range flush through ldexp.
__device__ __forceinline__ double exp_tab(double x, const double* tab)
{
    const double k64_ln2 = 0x1.71547652b82fep+6;     // 64 / ln 2
    const double ln2_64_hi = 0x1.62e42fee00000p-7;   // ln2/64, 21 trailing zero bits
    const double ln2_64_lo = 0x1.a39ef35793c76p-39;
    x = fmax(x, -800.0);
    const double nd = rint(x * k64_ln2);
    double r = fma(nd, -ln2_64_hi, x);
    r = fma(nd, -ln2_64_lo, r);
    const int n = (int)nd;
    double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = fma(r, p, 1.0);
    return ldexp(tab[kExpTab + (n & 63)] * p, n >> 6);
}

// log(x) for normal x > 0 (Tang's table-driven method, ACM TOMS 16(4), 1990):
// x = 2^e m, m in [1, 2), c_j the centre of m's 1/64-interval, r = m/c_j - 1
// (one FMA with the tabulated 1/c_j, |r| < 1/128), log x = e ln2 + log c_j +
// log1p(r) with log1p to degree 7 (truncation below 2e-18).  Absolute
// error about 1e-16 for the arguments used here.
__device__ __forceinline__ double log_tab(double x, const double* tab)
{
    const double ln2_hi = 0x1.62e42fee00000p-1;
    const double ln2_lo = 0x1.a39ef35793c76p-33;
    int e;
    const double m = 2.0 * frexp(x, &e);           // [1, 2)
    const int j = (int)((uint32_t)(__double2hiint(m)) >> 14) & 63;
    const double2 t = *(const double2*)(tab + kLogTab + 2 * j);
    const double r = fma(m, t.x, -1.0);
    double p = fma(r, 1.0 / 7.0, -1.0 / 6.0);
    p = fma(r, p, 1.0 / 5.0);
    p = fma(r, p, -0.25);
    p = fma(r, p, 1.0 / 3.0);
    p = fma(r, p, -0.5);
    const double ed = (double)(e - 1);
    const double lo = fma(r * r, p, ed * ln2_lo);
    return fma(ed, ln2_hi, t.y) + (r + lo);
}

// Slow path for one sign when its log argument is below the 1e-16 floor (or
// not positive): returns the floored value and zero derivatives except the
// 3 from the LSE shift, exactly like log(max(term, 1e-16)) under AD.
struct Sign1 {
    double v, g, aa, ac1, ac2, r1, r2, f;
};

__device__ __forceinline__ void lnF_slow(const double* tab, double a, double c1, double c2, Sign1& o)
{
    double t = exp_tab(-fabs(a), tab);
    double t2 = t * t, t3 = t2 * t;
    bool pos = a > 0.0;
    double n0 = pos ? t3 : 1.0, n1 = pos ? t2 : t, n2 = pos ? t : t2, n3 = pos ? 1.0 : t3;
    double D = n0 + 3.0 * (c1 * n1 + c2 * n2) + n3;
    double shift = pos ? 3.0 * a : 0.0;
    if (!(D >= kPolyakovEps)) {
        o.v = shift + kLogPolyakovEps;
        o.g = pos ? 3.0 : 0.0;
        o.aa = o.ac1 = o.ac2 = o.r1 = o.r2 = 0.0;
        // distribution (used for densities) keeps its closed form
        o.f = (c1 * n1 + 2.0 * c2 * n2 + n3) / D;
        return;
    }
    double iD = 1.0 / D;
    double r1 = n1 * iD, r2 = n2 * iD, r3 = n3 * iD;
    double f = c1 * r1 + 2.0 * c2 * r2 + r3;
    o.v = shift + log_tab(D, tab);
    o.g = 3.0 * f;
    o.aa = 3.0 * (c1 * r1 + 4.0 * c2 * r2 + 3.0 * r3) - 9.0 * f * f;
    o.ac1 = 3.0 * r1 * (1.0 - 3.0 * f);
    o.ac2 = 3.0 * r2 * (2.0 - 3.0 * f);
    o.r1 = r1;
    o.r2 = r2;
    o.f = f;
}

// One node of one flavour: adds C-weighted contributions to `s`.
// K2mu = exp(-2 mu / T): for mu >= 0 and a <= 0 both exponents are
// non-positive and exp(-|b|) = exp(-|a|) K2mu, so one exp serves both signs.
// q = p^2 (1 + xi cos^2), M2 = M^2, lmu = mu/T, invT = 1/T, tmu = 2 mu / T.
__device__ __forceinline__ void node_accumulate(const double* tab, double q, double C, double M2, double mu,
                                                double invT, double K2mu, double Phi, double Phib, FlavSums& s)
{
    double E, invE;
    rsqrt_sqrt(q + M2, invE, E);
    double a = (mu - E) * invT;
    double b = -(E + mu) * invT;

    // quark: (c1,c2) = (Phi, Phib), argument a
    double ta = exp_tab(-fabs(a), tab);
    double ta2 = ta * ta, ta3 = ta2 * ta;
    bool pa = a > 0.0;
    double a0 = pa ? ta3 : 1.0, a1 = pa ? ta2 : ta, a2 = pa ? ta : ta2, a3 = pa ? 1.0 : ta3;
    double Dp = a0 + 3.0 * (Phi * a1 + Phib * a2) + a3;
    // antiquark: (c1,c2) = (Phib, Phi), argument b = a - 2 mu / T
    double tb;
    if (a <= 0.0 && K2mu <= 1.0)   // K2mu <= 1 <=> mu >= 0, then b <= 0 too
        tb = ta * K2mu;
    else
        tb = exp_tab(-fabs(b), tab);
    double tb2 = tb * tb, tb3 = tb2 * tb;
    bool pb = b > 0.0;
    double b0 = pb ? tb3 : 1.0, b1 = pb ? tb2 : tb, b2 = pb ? tb : tb2, b3 = pb ? 1.0 : tb3;
    double Dm = b0 + 3.0 * (Phib * b1 + Phi * b2) + b3;

    double Q, gp, gm, aap, aam, ac1p, ac2p, ac1m, ac2m, r1p, r2p, r1m, r2m, fp, fm;
    if (Dp >= kPolyakovEps && Dm >= kPolyakovEps) {
        double DD = Dp * Dm;
        double iDD = rcp_nr(DD);
        double iDp = Dm * iDD, iDm = Dp * iDD;
        Q = log_tab(DD, tab) + (pa ? 3.0 * a : 0.0) + (pb ? 3.0 * b : 0.0);
        r1p = a1 * iDp;
        r2p = a2 * iDp;
        double r3p = a3 * iDp;
        r1m = b1 * iDm;
        r2m = b2 * iDm;
        double r3m = b3 * iDm;
        fp = Phi * r1p + 2.0 * Phib * r2p + r3p;
        fm = Phib * r1m + 2.0 * Phi * r2m + r3m;
        gp = 3.0 * fp;
        gm = 3.0 * fm;
        aap = 3.0 * (Phi * r1p + 4.0 * Phib * r2p + 3.0 * r3p) - gp * gp;
        aam = 3.0 * (Phib * r1m + 4.0 * Phi * r2m + 3.0 * r3m) - gm * gm;
        ac1p = 3.0 * r1p * (1.0 - gp);
        ac2p = 3.0 * r2p * (2.0 - gp);
        ac1m = 3.0 * r1m * (1.0 - gm);
        ac2m = 3.0 * r2m * (2.0 - gm);
    } else {
        Sign1 sp, sm;
        lnF_slow(tab, a, Phi, Phib, sp);
        lnF_slow(tab, b, Phib, Phi, sm);
        Q = sp.v + sm.v;
        gp = sp.g;
        gm = sm.g;
        aap = sp.aa;
        aam = sm.aa;
        ac1p = sp.ac1;
        ac2p = sp.ac2;
        ac1m = sm.ac1;
        ac2m = sm.ac2;
        r1p = sp.r1;
        r2p = sp.r2;
        r1m = sm.r1;
        r2m = sm.r2;
        fp = sp.f;
        fm = sm.f;
    }
    double QE = -(gp + gm) * invT;
    double QEE = (aap + aam) * (invT * invT);
    double QEP = -(ac1p + ac2m) * invT;
    double QEPb = -(ac2p + ac1m) * invT;
    double CinvE = C * invE;
    s.S = fma(C, Q, s.S);
    s.A1 = fma(CinvE, QE, s.A1);
    s.A2 = fma(CinvE * invE, QEE - QE * invE, s.A2);
    s.SP = fma(C, 3.0 * (r1p + r2m), s.SP);
    s.SPb = fma(C, 3.0 * (r2p + r1m), s.SPb);
    s.B1 = fma(CinvE, QEP, s.B1);
    s.B2 = fma(CinvE, QEPb, s.B2);
    s.SPP = fma(C, -9.0 * (r1p * r1p + r2m * r2m), s.SPP);
    s.SPPb = fma(C, -9.0 * (r1p * r2p + r1m * r2m), s.SPPb);
    s.SPbPb = fma(C, -9.0 * (r2p * r2p + r1m * r1m), s.SPbPb);
    s.Sfp = fma(C, fp, s.Sfp);
    s.Sfm = fma(C, fm, s.Sfm);
    s.Sfab = fma(C, fp * a + fm * b, s.Sfab);
}

// vacuum_integral (src/pnjl/core/Integrals.jl:123-132) and its M-derivatives.
__device__ __forceinline__ void vacuum_I(const DevParams& P, double M, double& I, double& dI, double& d2I)
{
    double L = P.Lambda;
    double sgn = (M < 0.0) ? -1.0 : 1.0;
    double Mt = fabs(M) + 1e-12;
    double Mt2 = Mt * Mt;
    double s = sqrt(L * L + Mt2);
    double lg = log((L + s) / Mt);
    I = (L * s * (2.0 * L * L + Mt2) - Mt2 * Mt2 * lg) * (1.0 / (16.0 * kPi * kPi));
    dI = sgn * Mt * (L * s - Mt2 * lg) * (1.0 / (4.0 * kPi * kPi));
    d2I = (2.0 * L * Mt2 / s + L * s - 3.0 * Mt2 * lg) * (1.0 / (4.0 * kPi * kPi));
}

__device__ __forceinline__ void mass_vec(const DevParams& P, const double* x, double M[3])
{
    M[0] = P.m_ud0 - 4.0 * P.G * x[0] + 2.0 * P.K * x[1] * x[2];
    M[1] = P.m_ud0 - 4.0 * P.G * x[1] + 2.0 * P.K * x[0] * x[2];
    M[2] = P.m_s0 - 4.0 * P.G * x[2] + 2.0 * P.K * x[0] * x[1];
}

}  // namespace pnjl
