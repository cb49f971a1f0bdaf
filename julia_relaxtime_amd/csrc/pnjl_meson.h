// pnjl_meson.h -- the consumers of the polarization Pi on gfx950: effective
// couplings from the one-loop A, the simple (pi, K, sigma_pi, sigma_K) and
// mixed (eta/eta', sigma/sigma') meson propagators, and their composition
// into the D_S / D_P of a scattering process in one Mandelstam channel.
//
// Follows src/relaxtime/EffectiveCouplings.jl, src/relaxtime/MesonPropagator.jl,
// src/relaxtime/TotalPropagator.jl and the process table of
// src/Constants_PNJL.jl:168-400 (file:line at each function).  Everything is
// a handful of FP64 operations per value; the expressions keep the
// reference's association order and are compiled without FMA contraction so
// that the oracle's restatement rounds the same way.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/pnjl_hip.h"

namespace meson {

struct C {
    double r, i;
};

__device__ __forceinline__ C cadd(C a, C b) { return C{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ C csub(C a, C b) { return C{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ C cmul(C a, C b) { return C{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ C rmul(double s, C a) { return C{s * a.r, s * a.i}; }
__device__ __forceinline__ bool cfinite(C a) { return isfinite(a.r) && isfinite(a.i); }
// real / complex as s * conj(z) / |z|^2 (magnitudes here stay far from overflow)
__device__ __forceinline__ C rdiv(double s, C z)
{
    const double d = z.r * z.r + z.i * z.i;
    return C{(s * z.r) / d, -(s * z.i) / d};
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kDenomEps = 1e-12;     // PROPAGATOR_DENOM_EPS  MesonPropagator.jl:22

// the 12 effective couplings in the reference's NamedTuple order
//   (EffectiveCouplings.jl:272-277)
struct KCoef {
    double K0p, K0m, K123p, K123m, K4567p, K4567m, K8p, K8m, K08p, K08m, detKp, detKm;
};

// calculate_G_from_A  EffectiveCouplings.jl:56-60
__device__ __forceinline__ double G_from_A(double A, double m, int Nc)
{
    return -(double)Nc / (4.0 * (kPi * kPi)) * (m * A);
}

// calculate_effective_couplings  EffectiveCouplings.jl:240-278 (+ the determinants :111-115)
__device__ __forceinline__ KCoef couplings(double G, double K, double Gu, double Gs)
{
    KCoef c;
    const double t0 = (1.0 / 3.0) * K * (2.0 * Gu + Gs);
    c.K0p = G - t0;
    c.K0m = G + t0;
    const double t123 = 0.5 * K * Gs;
    c.K123p = G + t123;
    c.K123m = G - t123;
    const double t4567 = 0.5 * K * Gu;
    c.K4567p = G + t4567;
    c.K4567m = G - t4567;
    const double t8 = (1.0 / 6.0) * K * (4.0 * Gu - Gs);
    c.K8p = G + t8;
    c.K8m = G - t8;
    const double t08 = (1.0 / 6.0) * sqrt(2.0) * K * (Gu - Gs);
    c.K08p = t08;
    c.K08m = -t08;
    c.detKp = c.K0p * c.K8p - c.K08p * c.K08p;
    c.detKm = c.K0m * c.K8m - c.K08m * c.K08m;
    return c;
}

// meson_propagator_simple  MesonPropagator.jl:268-291: D = 2K / (1 - 4 K Pi)
__device__ __forceinline__ C simple(double K, C Pi)
{
    const double k4 = 4.0 * K;
    C den{1.0 - k4 * Pi.r, -(k4 * Pi.i)};
    if (!cfinite(den)) return C{0.0, 0.0};
    if (hypot(den.r, den.i) < kDenomEps) den.i += kDenomEps;
    return rdiv(2.0 * K, den);
}

// mixing_matrix_elements  EffectiveCouplings.jl:137-160 (channel P: K+, S: K-)
__device__ __forceinline__ void mixing(C Puu, C Pss, double K0, double K8, double K08, double detK, C& M00, C& M08,
                                       C& M88)
{
    const double a = (4.0 / 3.0) * detK;
    const double b = ((4.0 / 3.0) * sqrt(2.0)) * detK;
    const C x00 = rmul(a, cadd(Puu, rmul(2.0, Pss)));
    const C x08 = rmul(b, csub(Puu, Pss));
    const C x88 = rmul(a, cadd(rmul(2.0, Puu), Pss));
    M00 = C{K0 - x00.r, -x00.i};   // real - complex, real + complex as Julia forms them
    M08 = C{K08 + x08.r, x08.i};
    M88 = C{K8 - x88.r, -x88.i};
}

// J = (lambda0[row, col], lambda8[row, col]) for row = flavour of the second
// field, col = flavour of the first (MesonPropagator.jl:357-363, lambda0 and
// lambda8 of Constants_PNJL.jl:106-117); flavours 0 = u, 1 = d, 2 = s
__device__ __forceinline__ void current(int f1, int f2, double& J0, double& J8)
{
    if (f1 != f2) {
        J0 = J8 = 0.0;
        return;
    }
    J0 = sqrt(2.0 / 3.0);
    J8 = f1 == 2 ? -2.0 / sqrt(3.0) : 1.0 / sqrt(3.0);
}

// meson_propagator_mixed  MesonPropagator.jl:365-419: D = 2 detK / det M * J^T M J'
__device__ __forceinline__ C mixed(double detK, C M00, C M08, C M88, double J0, double J8, double Jp0, double Jp8)
{
    C detM = csub(cmul(M00, M88), cmul(M08, M08));
    if (!cfinite(detM)) return C{0.0, 0.0};
    if (hypot(detM.r, detM.i) < kDenomEps) detM.i += kDenomEps;
    const C r = cadd(rmul(J0, cadd(rmul(Jp0, M00), rmul(Jp8, M08))), rmul(J8, cadd(rmul(Jp0, M08), rmul(Jp8, M88))));
    if (!cfinite(r)) return C{0.0, 0.0};
    return cmul(rdiv(2.0 * detK, detM), r);
}

// ---- process table (Constants_PNJL.jl:168-400) ---------------------------------
// per process: the four legs' flavours (u, d, s = 0, 1, 2, bar ignored) and,
// for each Mandelstam channel (t, s, u), the exchanged mesons as a mask
// (bit 0 pi, 1 K, 2 sigma_pi, 3 sigma_K, 4 eta/eta', 5 sigma/sigma'), or
// 0xff where the process has no such channel.
enum : unsigned {
    MS_PI = PNJL_MESON_PI,
    MS_K = PNJL_MESON_K,
    MS_SPI = PNJL_MESON_SIGMA_PI,
    MS_SK = PNJL_MESON_SIGMA_K,
    MS_MIXP = PNJL_MESON_MIXED_P,
    MS_MIXS = PNJL_MESON_MIXED_S,
    MS_NONE = 0xff
};
constexpr unsigned kPiPair = MS_PI | MS_SPI | MS_MIXP | MS_MIXS;   // [pi, sigma_pi] + both mixed
constexpr unsigned kKPair = MS_K | MS_SK;                         // [K, sigma_K]
constexpr unsigned kMix = MS_MIXP | MS_MIXS;
constexpr unsigned kPiOnly = MS_PI | MS_SPI;

struct ProcRow {
    unsigned char f[4];
    unsigned char ch[3];   // t, s, u
};

#define PNJL_PROC_ROWS \
    {{0, 1, 0, 1}, {kPiPair, MS_NONE, kPiOnly}},   /* ud_to_ud */ \
    {{0, 0, 0, 0}, {kPiPair, MS_NONE, kPiPair}},   /* uu_to_uu */ \
    {{0, 2, 0, 2}, {kMix, MS_NONE, kKPair}},       /* us_to_us */ \
    {{2, 2, 2, 2}, {kMix, MS_NONE, kMix}},         /* ss_to_ss */ \
    {{0, 1, 0, 1}, {kPiPair, MS_NONE, kPiOnly}},   /* ubardbar_to_ubardbar */ \
    {{0, 0, 0, 0}, {kPiPair, MS_NONE, kPiPair}},   /* ubarubar_to_ubarubar */ \
    {{0, 2, 0, 2}, {kMix, MS_NONE, kKPair}},       /* ubarsbar_to_ubarsbar */ \
    {{2, 2, 2, 2}, {kMix, MS_NONE, kMix}},         /* sbarsbar_to_sbarsbar */ \
    {{0, 1, 0, 1}, {kPiPair, kPiOnly, MS_NONE}},   /* udbar_to_udbar */ \
    {{1, 0, 1, 0}, {kPiPair, kPiOnly, MS_NONE}},   /* dubar_to_dubar */ \
    {{0, 0, 0, 0}, {kPiPair, kPiPair, MS_NONE}},   /* uubar_to_uubar */ \
    {{0, 0, 1, 1}, {kPiOnly, kPiPair, MS_NONE}},   /* uubar_to_ddbar */ \
    {{0, 2, 0, 2}, {kMix, kKPair, MS_NONE}},       /* usbar_to_usbar */ \
    {{2, 0, 2, 0}, {kMix, kKPair, MS_NONE}},       /* subar_to_subar */ \
    {{0, 0, 2, 2}, {kKPair, kMix, MS_NONE}},       /* uubar_to_ssbar */ \
    {{2, 2, 0, 0}, {kKPair, kMix, MS_NONE}},       /* ssbar_to_uubar */ \
    {{2, 2, 2, 2}, {kMix, kMix, MS_NONE}},         /* ssbar_to_ssbar */

__constant__ const ProcRow kProcs[PNJL_PROC_COUNT] = {PNJL_PROC_ROWS};
// the same rows for the host's argument checks
static const ProcRow kProcsHost[PNJL_PROC_COUNT] = {PNJL_PROC_ROWS};

// the two flavour pairs a channel joins: t (1,3)(2,4), s (1,2)(3,4), u (1,4)(2,3)
//   get_flavor_factors_for_channel TotalPropagator.jl:260-275, MesonPropagator.jl:389-398
__device__ __forceinline__ void channel_pairs(const ProcRow& p, int ch, int& a1, int& b1, int& a2, int& b2)
{
    if (ch == PNJL_MANDELSTAM_T) {
        a1 = p.f[0], b1 = p.f[2], a2 = p.f[1], b2 = p.f[3];
    } else if (ch == PNJL_MANDELSTAM_S) {
        a1 = p.f[0], b1 = p.f[1], a2 = p.f[2], b2 = p.f[3];
    } else {
        a1 = p.f[0], b1 = p.f[3], a2 = p.f[1], b2 = p.f[2];
    }
}

// table 5.3 flavour factor  TotalPropagator.jl:72-84: uu 1, dd -1, ss 2, else sqrt 2
__device__ __forceinline__ double flavour_factor(int a, int b)
{
    if (a != b) return sqrt(2.0);
    return a == 0 ? 1.0 : (a == 1 ? -1.0 : 2.0);
}

// Pi slots: (pair, channel) with pair uu 0, us 1, ss 2 and channel P 0, S 1 ->
// slot = pair * 2 + channel, the pnjl_polarization_grid order
//   get_polarization_params TotalPropagator.jl:308-336, total_propagator_mixed :521-537
__device__ __forceinline__ unsigned slots_for(unsigned mesons)
{
    unsigned s = 0;
    if (mesons & (MS_PI | MS_MIXP)) s |= 1u << 0;   // uu P
    if (mesons & (MS_SPI | MS_MIXS)) s |= 1u << 1;  // uu S
    if (mesons & MS_K) s |= 1u << 2;               // us P
    if (mesons & MS_SK) s |= 1u << 3;              // us S
    if (mesons & MS_MIXP) s |= 1u << 4;            // ss P
    if (mesons & MS_MIXS) s |= 1u << 5;            // ss S
    return s;
}

// calculate_propagator_by_channel  TotalPropagator.jl:894-963 on the six Pi
// slots (pi[slot] as above); mesons is the exchanged set (the table's, or a
// caller's subset).  Returns D_S, D_P.
__device__ __forceinline__ void by_channel(const ProcRow& p, int ch, unsigned mesons, const KCoef& k, const C* pi,
                                           C& DS, C& DP)
{
    int a1, b1, a2, b2;
    channel_pairs(p, ch, a1, b1, a2, b2);
    const double T1 = flavour_factor(a1, b1), T2 = flavour_factor(a2, b2);
    DS = C{0.0, 0.0};
    DP = C{0.0, 0.0};
    if (mesons & MS_PI) DP = cadd(DP, rmul(T2, rmul(T1, simple(k.K123p, pi[0]))));
    if (mesons & MS_K) DP = cadd(DP, rmul(T2, rmul(T1, simple(k.K4567p, pi[2]))));
    if (mesons & MS_SPI) DS = cadd(DS, rmul(T2, rmul(T1, simple(k.K123m, pi[1]))));
    if (mesons & MS_SK) DS = cadd(DS, rmul(T2, rmul(T1, simple(k.K4567m, pi[3]))));
    if (mesons & (MS_MIXP | MS_MIXS)) {
        double J0, J8, Jp0, Jp8;
        current(a1, b1, J0, J8);
        current(a2, b2, Jp0, Jp8);
        C M00, M08, M88;
        if (mesons & MS_MIXP) {
            mixing(pi[0], pi[4], k.K0p, k.K8p, k.K08p, k.detKp, M00, M08, M88);
            DP = cadd(DP, mixed(k.detKp, M00, M08, M88, J0, J8, Jp0, Jp8));
        }
        if (mesons & MS_MIXS) {
            mixing(pi[1], pi[5], k.K0m, k.K8m, k.K08m, k.detKm, M00, M08, M88);
            DS = cadd(DS, mixed(k.detKm, M00, M08, M88, J0, J8, Jp0, Jp8));
        }
    }
}

}  // namespace meson
