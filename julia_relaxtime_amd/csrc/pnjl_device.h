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

// The node functions evaluate exactly the written operations: every fused
// multiply-add is an explicit fma(), nothing is contracted by the compiler.
// Instances of the node loop compiled in different contexts (the checked and
// unchecked loops) then round identically, and results do not depend on which
// states share a wavefront.
#define NODE_MATH_EXACT _Pragma("clang fp contract(off)")

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
    NODE_MATH_EXACT
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
    NODE_MATH_EXACT
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// a * b + c with c a constant, as one VOP3 v_fma_f64 that reads c from
// scalar registers.  Left to itself the compiler keeps such a constant in one
// vector register for the whole loop and emits v_fmac (c's register tied to
// the result), so every use first copies c: a move per polynomial step.
// Same operation, same rounding as fma(a, b, c).
__device__ __forceinline__ double fma_k(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// exp(x) (Tang's table-driven method, ACM TOMS 15(2), 1989): x = (64 k + j)
// ln2/64 + r with |r| <= ln2/128 (Cody-Waite reduction with a split ln2/64),
// exp(x) = 2^k 2^(j/64) e^r and e^r from its degree-5 Taylor polynomial
// (truncation below 4e-17).  About 1 ulp; results below the double range
// flush through ldexp.  kClamp: x is first clamped to [-800, 700]; without
// it the caller guarantees x <= 700 (below -800 both forms give 0, and the
// conversion of nd to int saturates on this hardware).
template <bool kClamp = true>
__device__ __forceinline__ double exp_tab(double x, const double* tab)
{
    NODE_MATH_EXACT
    const double k64_ln2 = 0x1.71547652b82fep+6;     // 64 / ln 2
    const double ln2_64_hi = 0x1.62e42fee00000p-7;   // ln2/64, 21 trailing zero bits
    const double ln2_64_lo = 0x1.a39ef35793c76p-39;
    if (kClamp) x = fmin(fmax(x, -800.0), 700.0);
    const double nd = rint(x * k64_ln2);
    double r = fma(nd, -ln2_64_hi, x);
    r = fma(nd, -ln2_64_lo, r);
    const int n = (int)nd;
    double p = fma_k(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma_k(r, p, 1.0 / 6.0);
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
    NODE_MATH_EXACT
    const double ln2_hi = 0x1.62e42fee00000p-1;
    const double ln2_lo = 0x1.a39ef35793c76p-33;
    int e;
    const double m = 2.0 * frexp(x, &e);           // [1, 2)
    const int j = (int)((uint32_t)(__double2hiint(m)) >> 14) & 63;
    const double2 t = *(const double2*)(tab + kLogTab + 2 * j);
    const double r = fma(m, t.x, -1.0);
    double p = fma_k(r, 1.0 / 7.0, -1.0 / 6.0);
    p = fma_k(r, p, 1.0 / 5.0);
    p = fma_k(r, p, -0.25);
    p = fma_k(r, p, 1.0 / 3.0);
    p = fma(r, p, -0.5);
    const double ed = (double)(e - 1);
    const double lo = fma(r * r, p, ed * ln2_lo);
    return fma(ed, ln2_hi, t.y) + (r + lo);
}

// Per-sign terms of ln F(a) = ln(1 + 3 c1 e^a + 3 c2 e^2a + e^3a) in the
// normalisation the node loop accumulates (the factors 3 and 9 of the
// derivatives are applied once per flavour sum):
//   v = ln F,  gq = (dlnF/da) / 3,  h = (d2lnF/da2) / 3,
//   k1 = (d2lnF/da dc1) / 3,  k2 = (d2lnF/da dc2) / 3,
//   r1 = e^a / F,  r2 = e^2a / F  (dlnF/dc_i = 3 r_i),  f = distribution.
struct Sign1 {
    double v, gq, h, k1, k2, r1, r2, f;
};

// Slow path for one sign: the reference's shifted form for a > 0 and its
// 1e-16 floor on the log argument -- a floored value has zero derivatives
// except the 3 of the shift, exactly like log(max(term, 1e-16)) under AD.
__device__ __forceinline__ void lnF_slow(const double* tab, double a, double c1, double c2, Sign1& o)
{
    NODE_MATH_EXACT
    double t = exp_tab(-fabs(a), tab);
    double t2 = t * t, t3 = t2 * t;
    bool pos = a > 0.0;
    double n0 = pos ? t3 : 1.0, n1 = pos ? t2 : t, n2 = pos ? t : t2, n3 = pos ? 1.0 : t3;
    double D = n0 + 3.0 * (c1 * n1 + c2 * n2) + n3;
    double shift = pos ? 3.0 * a : 0.0;
    if (!(D >= kPolyakovEps)) {
        o.v = shift + kLogPolyakovEps;
        o.gq = pos ? 1.0 : 0.0;
        o.h = o.k1 = o.k2 = o.r1 = o.r2 = 0.0;
        // distribution (used for densities) keeps its closed form
        o.f = (c1 * n1 + 2.0 * c2 * n2 + n3) / D;
        return;
    }
    double iD = 1.0 / D;
    double r1 = n1 * iD, r2 = n2 * iD, r3 = n3 * iD;
    double f = c1 * r1 + 2.0 * c2 * r2 + r3;
    o.v = shift + log_tab(D, tab);
    o.gq = f;
    o.h = (c1 * r1 + 4.0 * c2 * r2 + 3.0 * r3) - 3.0 * f * f;
    o.k1 = r1 * (1.0 - 3.0 * f);
    o.k2 = r2 * (2.0 - 3.0 * f);
    o.r1 = r1;
    o.r2 = r2;
    o.f = f;
}

// Largest exponent taken on the unscaled fast path: e^(3a) stays below 1e131,
// so the product of both signs' arguments is finite.
constexpr double kFastExpMax = 100.0;

// True when no node of a flavour can leave the fast path.  With Phi, Phib >=
// -d the argument D = 1 + 3 Phi t + 3 Phib t^2 + t^3 is >= (1 - 6d) for t <= 1
// and D / t^3 >= (1 - 6d) for t >= 1, so at d = 0.1 every log argument is
// >= 0.4 max(1, e^3a), far from the 1e-16 floor; and a = (mu - E)/T < mu/T,
// b = -(E + mu)/T < -mu/T bound both exponents by |mu|/T.
__device__ __forceinline__ bool nodes_all_fast(double Phi, double Phib, double muT)
{
    return Phi >= -0.1 && Phib >= -0.1 && fabs(muT) <= kFastExpMax;
}

// One node of one flavour: adds C-weighted raw terms to `s` (scaled to the
// FlavSums definitions by flav_finish).
// The fast path works with the unscaled arguments D = 1 + 3 c1 t + 3 c2 t^2 +
// t^3, t = e^a (resp. e^b), for either sign of the exponents: ln F is the same
// function as the reference's shifted form 3a + ln(e^-3a + ...) for a > 0,
// and the ratios t^k / D are scale-free, so no branch on sign(a) is needed
// and e^b = e^a exp(-2 mu / T) = e^a K2mu costs one multiply.  The
// reference's 1e-16 floor applies to the argument it takes the log of, which
// for a > 0 is D e^-3a: with kChecked the fast path requires
// D >= 1e-16 max(1, t^3) for both signs and a, b <= kFastExpMax, and
// everything else goes to the per-sign shifted slow path; without kChecked
// the caller has established nodes_all_fast.
// q = p^2 (1 + xi cos^2), M2 = M^2, muT = mu/T, invT = 1/T, K2mu = exp(-2 mu/T).
template <bool kChecked>
__device__ __forceinline__ void node_accumulate(const double* tab, double q, double C, double M2, double muT,
                                                double invT, double K2mu, double Phi, double Phib, FlavSums& s)
{
    NODE_MATH_EXACT
    double E, invE;
    rsqrt_sqrt(q + M2, invE, E);
    const double a = fma(-E, invT, muT);
    const double b = fma(-E, invT, -muT);

    // quark: (c1, c2) = (Phi, Phib), t = e^a; antiquark: (Phib, Phi), e^b
    const double ta = exp_tab<kChecked>(a, tab);
    const double tb = ta * K2mu;
    const double ta2 = ta * ta, ta3 = ta2 * ta;
    const double tb2 = tb * tb, tb3 = tb2 * tb;
    const double Dp = fma(3.0, fma(Phi, ta, Phib * ta2), 1.0) + ta3;
    const double Dm = fma(3.0, fma(Phib, tb, Phi * tb2), 1.0) + tb3;

    double Q, gs, hs, k1p, k2p, k1m, k2m, r1p, r2p, r1m, r2m, fp, fm;
    bool fast = true;
    if (kChecked)
        fast = Dp >= kPolyakovEps * fmax(1.0, ta3) && Dm >= kPolyakovEps * fmax(1.0, tb3) &&
               fmax(a, b) <= kFastExpMax;
    if (fast) {
        const double DD = Dp * Dm;
        const double iDD = rcp_nr(DD);
        const double iDp = Dm * iDD, iDm = Dp * iDD;
        Q = log_tab(DD, tab);
        r1p = ta * iDp;
        r2p = ta2 * iDp;
        const double r3p = ta3 * iDp;
        r1m = tb * iDm;
        r2m = tb2 * iDm;
        const double r3m = tb3 * iDm;
        const double u1p = Phi * r1p, u1m = Phib * r1m;
        fp = fma(2.0 * Phib, r2p, u1p) + r3p;
        fm = fma(2.0 * Phi, r2m, u1m) + r3m;
        const double hp = fma(-3.0 * fp, fp, fma(3.0, r3p, fma(4.0 * Phib, r2p, u1p)));
        const double hm = fma(-3.0 * fm, fm, fma(3.0, r3m, fma(4.0 * Phi, r2m, u1m)));
        k1p = r1p * fma(-3.0, fp, 1.0);
        k2p = r2p * fma(-3.0, fp, 2.0);
        k1m = r1m * fma(-3.0, fm, 1.0);
        k2m = r2m * fma(-3.0, fm, 2.0);
        gs = fp + fm;
        hs = hp + hm;
    } else {
        Sign1 sp, sm;
        lnF_slow(tab, a, Phi, Phib, sp);
        lnF_slow(tab, b, Phib, Phi, sm);
        Q = sp.v + sm.v;
        gs = sp.gq + sm.gq;
        hs = sp.h + sm.h;
        k1p = sp.k1;
        k2p = sp.k2;
        k1m = sm.k1;
        k2m = sm.k2;
        r1p = sp.r1;
        r2p = sp.r2;
        r1m = sm.r1;
        r2m = sm.r2;
        fp = sp.f;
        fm = sm.f;
    }
    const double CinvE = C * invE;
    s.S = fma(C, Q, s.S);
    s.A1 = fma(CinvE, gs, s.A1);
    s.A2 = fma(CinvE * invE, fma(hs, invT, gs * invE), s.A2);
    s.B1 = fma(CinvE, k1p + k2m, s.B1);
    s.B2 = fma(CinvE, k2p + k1m, s.B2);
    s.SP = fma(C, r1p + r2m, s.SP);
    s.SPb = fma(C, r2p + r1m, s.SPb);
    s.SPP = fma(C, fma(r1p, r1p, r2m * r2m), s.SPP);
    s.SPPb = fma(C, fma(r1p, r2p, r1m * r2m), s.SPPb);
    s.SPbPb = fma(C, fma(r2p, r2p, r1m * r1m), s.SPbPb);
    s.Sfp = fma(C, fp, s.Sfp);
    s.Sfm = fma(C, fm, s.Sfm);
    s.Sfab = fma(C, fma(fp, a, fm * b), s.Sfab);
}

// Scales the raw node sums to the FlavSums definitions: the derivative
// factors Q_E = -3 gs/T, Q_EE = 3 hs/T^2, Q_EPhi = -3 (k1p + k2m)/T,
// Q_Phi = 3 (r1p + r2m), Q_PhiPhi = -9 (r1p^2 + r2m^2), ...
__device__ __forceinline__ void flav_finish(FlavSums& s, double invT)
{
    const double t3 = 3.0 * invT;
    s.A1 *= -t3;
    s.A2 *= t3;
    s.B1 *= -t3;
    s.B2 *= -t3;
    s.SP *= 3.0;
    s.SPb *= 3.0;
    s.SPP *= -9.0;
    s.SPPb *= -9.0;
    s.SPbPb *= -9.0;
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
