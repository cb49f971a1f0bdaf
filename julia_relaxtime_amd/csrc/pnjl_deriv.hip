// pnjl_deriv.hip -- implicit thermodynamic derivatives at converged roots
// (ThermoDerivatives.jl:132-245 mass_derivatives order 1 and
// thermo_derivatives, :342-466 bulk_viscosity_coefficients).
//
// At a root x of F(x; T, mu) = grad_x P = 0 the implicit function theorem
// gives dx/dtheta = -(d2 Omega/dx2)^-1 d2 Omega/dx dtheta for theta = (T, mu);
// every other derivative is a chain rule through x.  The reference gets the
// partial derivatives by nested forward-mode AD; here they are closed forms
// of the node terms ln F(a), ln F(b), a = (mu - E)/T, b = -(E + mu)/T: the
// Hessian of the gap solve extended by the rows and columns of T and mu.
//
// One thread per root: this pass runs once per point after the solve, so it
// keeps the node loop simple -- every node and sign through the shifted
// per-sign evaluation of the 1e-16 floor (lnF_slow), table-driven exp/log.
#include <hip/hip_runtime.h>

#include "pnjl_ctx.h"

// Written operations only (every FMA explicit): at low T the entropy and the
// density derivatives are small differences of O(1e-9) Polyakov and node
// terms, and their last bits follow the exact operation sequence.
#pragma clang fp contract(off)

using namespace pnjl_host;

namespace pnjl_deriv {

using pnjl::DevParams;

// node sums of one flavour: derivatives of S = sum C Q in (M, Phi, Phib) and
// the mixed ones with (mu, T) at one common mu
struct FlavD {
    double S, SM, SP, SPb, SMM, SMP, SMPb, SPP, SPPb, SPbPb;
    double Smu, ST, Smumu, SmuT, STT, SMmu, SMT, SPmu, SPbmu, SPT, SPbT;
};

__device__ void flavour_d(const double* tab, int pn, int tn, double M, double mu, double T, double xi, double Phi,
                          double Phib, FlavD& o)
{
    o = FlavD{};
    const double* pp = tab + pnjl::kMathTab;
    const double* cc = pp + pn;
    const double* Cm = pp + pn + tn;
    const double invT = 1.0 / T, invT2 = invT * invT, M2 = M * M;
    for (int l = 0; l < tn; l++) {
        const double xc = xi * cc[l];
        for (int j = 0; j < pn; j++) {
            const double C = Cm[l * pn + j];
            double E, invE;
            pnjl::rsqrt_sqrt(fma(xc, pp[j], pp[j]) + M2, invE, E);
            const double a = (mu - E) * invT, b = -(E + mu) * invT;
            pnjl::Sign1 sp, sm;
            pnjl::lnF_slow(tab, a, Phi, Phib, sp);
            pnjl::lnF_slow(tab, b, Phib, Phi, sm);
            // derivatives of ln F in its own argument and loop parameters
            const double gp = 3.0 * sp.gq, gm = 3.0 * sm.gq;
            const double hp = 3.0 * sp.h, hm = 3.0 * sm.h;
            const double ac1p = 3.0 * sp.k1, ac2p = 3.0 * sp.k2, ac1m = 3.0 * sm.k1, ac2m = 3.0 * sm.k2;
            const double Q = sp.v + sm.v;
            const double QE = -(gp + gm) * invT, QEE = (hp + hm) * invT2;
            const double QP = 3.0 * (sp.r1 + sm.r2), QPb = 3.0 * (sp.r2 + sm.r1);
            const double QEP = -(ac1p + ac2m) * invT, QEPb = -(ac2p + ac1m) * invT;
            const double EM = M * invE, EMM = (1.0 - EM * EM) * invE;
            o.S += C * Q;
            o.SM += C * (QE * EM);
            o.SP += C * QP;
            o.SPb += C * QPb;
            o.SMM += C * (QEE * EM * EM + QE * EMM);
            o.SMP += C * (QEP * EM);
            o.SMPb += C * (QEPb * EM);
            o.SPP += C * (-9.0 * (sp.r1 * sp.r1 + sm.r2 * sm.r2));
            o.SPPb += C * (-9.0 * (sp.r1 * sp.r2 + sm.r1 * sm.r2));
            o.SPbPb += C * (-9.0 * (sp.r2 * sp.r2 + sm.r1 * sm.r1));
            // da/dmu = 1/T, db/dmu = -1/T, da/dT = -a/T, db/dT = -b/T, da/dE = db/dE = -1/T
            o.Smu += C * (gp - gm) * invT;
            o.ST += C * (-(gp * a + gm * b) * invT);
            o.Smumu += C * (hp + hm) * invT2;
            o.SmuT += C * (-(gp - gm) - (hp * a - hm * b)) * invT2;
            o.STT += C * (2.0 * (gp * a + gm * b) + hp * a * a + hm * b * b) * invT2;
            o.SMmu += C * (-(hp - hm) * invT2) * EM;
            o.SMT += C * ((gp + gm + hp * a + hm * b) * invT2) * EM;
            o.SPmu += C * (ac1p - ac2m) * invT;
            o.SPbmu += C * (ac2p - ac1m) * invT;
            o.SPT += C * (-(ac1p * a + ac2m * b) * invT);
            o.SPbT += C * (-(ac2p * a + ac1m * b) * invT);
        }
    }
}

// LU with partial pivoting of a 5x5 system with two right-hand sides;
// returns false on an exactly zero pivot
__device__ bool lu5_solve2(double A[5][5], double b1[5], double b2[5])
{
    for (int k = 0; k < 5; k++) {
        int p = k;
        double best = fabs(A[k][k]);
        for (int i = k + 1; i < 5; i++)
            if (fabs(A[i][k]) > best) {
                best = fabs(A[i][k]);
                p = i;
            }
        if (A[p][k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < 5; j++) {
                double t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
            double t = b1[k];
            b1[k] = b1[p];
            b1[p] = t;
            t = b2[k];
            b2[k] = b2[p];
            b2[p] = t;
        }
        for (int i = k + 1; i < 5; i++) {
            const double l = A[i][k] / A[k][k];
            for (int j = k + 1; j < 5; j++) A[i][j] -= l * A[k][j];
            b1[i] -= l * b1[k];
            b2[i] -= l * b2[k];
        }
    }
    for (int i = 4; i >= 0; i--) {
        double s1 = b1[i], s2 = b2[i];
        for (int j = i + 1; j < 5; j++) {
            s1 -= A[i][j] * b1[j];
            s2 -= A[i][j] * b2[j];
        }
        b1[i] = s1 / A[i][i];
        b2[i] = s2 / A[i][i];
    }
    return true;
}

struct DerivArgs {
    DevParams P;
    int64_t n;
    const double *T, *mu, *xi, *x;
    const double* tab;
    int pn, tn;
    double* out;   // n * PNJL_THERMO_DERIV_FIELDS
};

__global__ void __launch_bounds__(256) deriv_kernel(DerivArgs A)
{
    extern __shared__ double lds[];
    const int tab_len = pnjl::kMathTab + A.pn + A.tn + A.pn * A.tn + A.pn;
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = A.tab[i];
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n) return;
    const DevParams& P = A.P;
    const double T = A.T[g], mu = A.mu[g], xi = A.xi[g];
    double x[5];
    for (int i = 0; i < 5; i++) x[i] = A.x[g * 5 + i];
    double* out = A.out + g * PNJL_THERMO_DERIV_FIELDS;
    const double pu = x[0], pd = x[1], ps = x[2], Phi = x[3], Phib = x[4];
    const double G = P.G, K = P.K;

    // M(x) and its derivatives (Thermodynamics.jl:81-110)
    double M[3];
    pnjl::mass_vec(P, x, M);
    const double dM[3][3] = {{-4.0 * G, 2.0 * K * ps, 2.0 * K * pd},
                             {2.0 * K * ps, -4.0 * G, 2.0 * K * pu},
                             {2.0 * K * pd, 2.0 * K * pu, -4.0 * G}};
    // d2M[i][a][b]: M_u has 2K at (d, s), M_d at (u, s), M_s at (u, d)
    auto d2M = [&](int i, int a, int b) -> double {
        const int o = (i == 0) ? 0 : (i == 1 ? 1 : 2);
        return (a != b && a != o && b != o) ? 2.0 * K : 0.0;
    };
    // node sums: u, s, and d unless phi_d == phi_u
    FlavD f[3];
    flavour_d(lds, A.pn, A.tn, M[0], mu, T, xi, Phi, Phib, f[0]);
    flavour_d(lds, A.pn, A.tn, M[2], mu, T, xi, Phi, Phib, f[2]);
    if (pd == pu) f[1] = f[0];
    else flavour_d(lds, A.pn, A.tn, M[1], mu, T, xi, Phi, Phib, f[1]);

    // chiral term, vacuum term, Polyakov potential (Thermodynamics.jl:112-160)
    const double chi = 2.0 * G * (pu * pu + pd * pd + ps * ps) - 4.0 * K * pu * pd * ps;
    const double d2chi[3][3] = {{4.0 * G, -4.0 * K * ps, -4.0 * K * pd},
                                {-4.0 * K * ps, 4.0 * G, -4.0 * K * pu},
                                {-4.0 * K * pd, -4.0 * K * pu, 4.0 * G}};
    double Evac = 0.0, VM[3], VMM[3];
    for (int i = 0; i < 3; i++) {
        double I, dI, d2I;
        pnjl::vacuum_I(P, M[i], I, dI, d2I);
        Evac += I;
        VM[i] = -2.0 * P.Nc * dI;
        VMM[i] = -2.0 * P.Nc * d2I;
    }
    Evac *= -2.0 * P.Nc;
    const double Tr = P.T0 / T;
    const double Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr, Tb = P.b3 * Tr * Tr * Tr;
    const double PP = Phi * Phib;
    const double val = 1.0 - 6.0 * PP + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * PP * PP;
    const bool floored = !(val >= pnjl::kPolyakovEps);
    const double lg = floored ? pnjl::kLogPolyakovEps : log(val);
    double lP = 0, lPb = 0, lPP = 0, lPPb = 0, lPbPb = 0;
    if (!floored) {
        const double gP = -6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib;
        const double gPb = -6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib;
        lP = gP / val;
        lPb = gPb / val;
        lPP = (24.0 * Phi - 6.0 * Phib * Phib) / val - lP * lP;
        lPPb = (-6.0 - 12.0 * Phi * Phib) / val - lP * lPb;
        lPbPb = (24.0 * Phib - 6.0 * Phi * Phi) / val - lPb * lPb;
    }
    const double T4 = T * T * T * T;
    const double U = T4 * (-0.5 * Ta * PP + Tb * lg);
    const double hU[3] = {T4 * (Tb * lPP), T4 * (-0.5 * Ta + Tb * lPPb), T4 * (Tb * lPbPb)};
    // T-derivatives at fixed loops: T^4 a(T) = a0 T^4 + a1 T0 T^3 + a2 T0^2 T^2, T^4 b(T) = b3 T0^3 T
    const double A1 = 4.0 * P.a0 * T * T * T + 3.0 * P.a1 * P.T0 * T * T + 2.0 * P.a2 * P.T0 * P.T0 * T;
    const double A2 = 12.0 * P.a0 * T * T + 6.0 * P.a1 * P.T0 * T + 2.0 * P.a2 * P.T0 * P.T0;
    const double B1 = P.b3 * P.T0 * P.T0 * P.T0;
    const double UT = -0.5 * PP * A1 + B1 * lg, UTT = -0.5 * PP * A2;
    const double UPT = -0.5 * Phib * A1 + B1 * lP, UPbT = -0.5 * Phi * A1 + B1 * lPb;

    // Hessian of Omega in x
    double H[5][5];
    double Ssum = 0.0;
    for (int i = 0; i < 3; i++) Ssum += f[i].S;
    const double omega = chi + U + Evac - 2.0 * T * Ssum;
    for (int a = 0; a < 3; a++) {
        for (int b = 0; b < 3; b++) {
            double h = d2chi[a][b];
            for (int i = 0; i < 3; i++) {
                const double gM = VM[i] - 2.0 * T * f[i].SM, hM = VMM[i] - 2.0 * T * f[i].SMM;
                h += hM * dM[i][a] * dM[i][b] + gM * d2M(i, a, b);
            }
            H[a][b] = h;
        }
        double hP = 0.0, hPb = 0.0;
        for (int i = 0; i < 3; i++) {
            hP += -2.0 * T * f[i].SMP * dM[i][a];
            hPb += -2.0 * T * f[i].SMPb * dM[i][a];
        }
        H[a][3] = H[3][a] = hP;
        H[a][4] = H[4][a] = hPb;
    }
    double SPP = 0, SPPb = 0, SPbPb = 0;
    for (int i = 0; i < 3; i++) {
        SPP += f[i].SPP;
        SPPb += f[i].SPPb;
        SPbPb += f[i].SPbPb;
    }
    H[3][3] = hU[0] - 2.0 * T * SPP;
    H[3][4] = H[4][3] = hU[1] - 2.0 * T * SPPb;
    H[4][4] = hU[2] - 2.0 * T * SPbPb;

    // derivatives in theta = (T, mu)
    double Om_mu = 0, Om_T = UT, Om_mumu = 0, Om_muT = 0, Om_TT = UTT;
    double Om_xmu[5] = {0, 0, 0, 0, 0}, Om_xT[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 3; i++) {
        Om_mu += -2.0 * T * f[i].Smu;
        Om_T += -2.0 * f[i].S - 2.0 * T * f[i].ST;
        Om_mumu += -2.0 * T * f[i].Smumu;
        Om_muT += -2.0 * f[i].Smu - 2.0 * T * f[i].SmuT;
        Om_TT += -4.0 * f[i].ST - 2.0 * T * f[i].STT;
        const double LMmu = -2.0 * T * f[i].SMmu, LMT = -2.0 * f[i].SM - 2.0 * T * f[i].SMT;
        for (int a = 0; a < 3; a++) {
            Om_xmu[a] += dM[i][a] * LMmu;
            Om_xT[a] += dM[i][a] * LMT;
        }
        Om_xmu[3] += -2.0 * T * f[i].SPmu;
        Om_xmu[4] += -2.0 * T * f[i].SPbmu;
        Om_xT[3] += -2.0 * f[i].SP - 2.0 * T * f[i].SPT;
        Om_xT[4] += -2.0 * f[i].SPb - 2.0 * T * f[i].SPbT;
    }
    Om_xT[3] += UPT;
    Om_xT[4] += UPbT;

    double dxdT[5], dxdmu[5];
    for (int a = 0; a < 5; a++) {
        dxdT[a] = -Om_xT[a];
        dxdmu[a] = -Om_xmu[a];
    }
    if (!lu5_solve2(H, dxdT, dxdmu)) {
        for (int k = 0; k < PNJL_THERMO_DERIV_FIELDS; k++) out[k] = __builtin_nan("");
        return;
    }
    const double Pr = -omega, s = -Om_T, ntot = -Om_mu;
    double ds_dT = -Om_TT, ds_dmu = -Om_muT, dnt_dT = -Om_muT, dnt_dmu = -Om_mumu;
    for (int a = 0; a < 5; a++) {
        ds_dT += -Om_xT[a] * dxdT[a];
        ds_dmu += -Om_xT[a] * dxdmu[a];
        dnt_dT += -Om_xmu[a] * dxdT[a];
        dnt_dmu += -Om_xmu[a] * dxdmu[a];
    }
    const double n = ntot / 3.0, dn_dT = dnt_dT / 3.0, dn_dmu = dnt_dmu / 3.0;
    const double E_T = T * ds_dT + mu * dnt_dT, E_mu = T * ds_dmu + mu * dnt_dmu;
    const double den_e = E_T * dn_dmu - E_mu * dn_dT, den_n = dn_dT * E_mu - dn_dmu * E_T;
    out[0] = Pr;
    out[1] = -Pr + T * s + mu * ntot;
    out[2] = n;
    out[3] = ntot / (3.0 * P.rho0);
    out[4] = s;
    out[5] = s;
    out[6] = ntot;
    out[7] = E_T;
    out[8] = E_mu;
    out[9] = dn_dT;
    out[10] = dn_dmu;
    out[11] = den_e == 0.0 ? __builtin_nan("") : (s * dn_dmu - ntot * dn_dT) / den_e;
    out[12] = den_n == 0.0 ? __builtin_nan("") : (s * E_mu - ntot * E_T) / den_n;
    // masses as ThermoDerivatives.jl:111-121 writes them (bare 5.5 and 140.0 MeV)
    const double mu0 = 0.0055 / 0.197327, ms0 = 0.140 / 0.197327;
    out[13] = mu0 - 4.0 * G * pu + 2.0 * K * pd * ps;
    out[14] = mu0 - 4.0 * G * pd + 2.0 * K * pu * ps;
    out[15] = ms0 - 4.0 * G * ps + 2.0 * K * pu * pd;
    for (int i = 0; i < 3; i++) {
        double gT = 0.0, gm = 0.0;
        for (int a = 0; a < 3; a++) {
            gT += dM[i][a] * dxdT[a];
            gm += dM[i][a] * dxdmu[a];
        }
        out[16 + i] = gT;
        out[19 + i] = gm;
    }
    // bulk_viscosity_coefficients: mu_B = 3 mu_q, n_B = n
    const double ds_dmuB = ds_dmu / 3.0, dn_dmuB = dn_dmu / 3.0;
    out[22] = (s * dn_dmuB - n * dn_dT) / (T * (ds_dT * dn_dmuB - ds_dmuB * dn_dT));
    out[23] = -(n * ds_dT - s * dn_dT) / (n * ds_dmuB - s * dn_dmuB);
    for (int a = 0; a < 5; a++) {
        out[24 + a] = dxdT[a];
        out[29 + a] = dxdmu[a];
    }
    out[34] = ds_dT;
    out[35] = ds_dmu;
}

// ---------------------------------------------------------------------------
// second-order implicit derivatives (solve_with_derivatives order 2,
// ImplicitSolver.jl:660-701; mass_derivatives order 2,
// ThermoDerivatives.jl:153-172)
//
// Along the root x(T, mu), H x_a + F_a = 0 with H = d2 Omega/dx2 and F_a =
// d2 Omega/dx dtheta_a.  Differentiating once more along w_b = (x_b, e_b):
//   x_ab = -H^-1 [(D_wb H) x_a + D_wb F_a].
// H and F_a are the closed forms of deriv_kernel; written once as templates
// over the number type, they run in double (H, F_a, hence x_T, x_mu) and in
// D2, a dual number carrying the derivatives along w_T and w_mu at once
// (first-order forward mode: the products keep the eps terms, eps^2 = 0), so
// the third derivatives of Omega contracted with w are exact to rounding.
// Branches (the 1e-16 floor, the sign of a, |M|) follow the value part.
// ---------------------------------------------------------------------------
struct D2 {
    double v, t, m;   // value, derivative along w_T, along w_mu
};
__device__ inline D2 operator+(D2 a, D2 b) { return D2{a.v + b.v, a.t + b.t, a.m + b.m}; }
__device__ inline D2 operator-(D2 a, D2 b) { return D2{a.v - b.v, a.t - b.t, a.m - b.m}; }
__device__ inline D2 operator-(D2 a) { return D2{-a.v, -a.t, -a.m}; }
__device__ inline D2 operator*(D2 a, D2 b) { return D2{a.v * b.v, a.t * b.v + a.v * b.t, a.m * b.v + a.v * b.m}; }
__device__ inline D2 operator/(D2 a, D2 b)
{
    const double r = 1.0 / b.v, q = a.v * r;
    return D2{q, (a.t - q * b.t) * r, (a.m - q * b.m) * r};
}
__device__ inline D2 operator+(D2 a, double b) { return D2{a.v + b, a.t, a.m}; }
__device__ inline D2 operator+(double a, D2 b) { return D2{a + b.v, b.t, b.m}; }
__device__ inline D2 operator-(D2 a, double b) { return D2{a.v - b, a.t, a.m}; }
__device__ inline D2 operator-(double a, D2 b) { return D2{a - b.v, -b.t, -b.m}; }
__device__ inline D2 operator*(D2 a, double b) { return D2{a.v * b, a.t * b, a.m * b}; }
__device__ inline D2 operator*(double a, D2 b) { return D2{a * b.v, a * b.t, a * b.m}; }
__device__ inline D2 operator/(D2 a, double b) { return D2{a.v / b, a.t / b, a.m / b}; }
__device__ inline D2 operator/(double a, D2 b)
{
    const double r = 1.0 / b.v, q = a * r;
    return D2{q, -q * b.t * r, -q * b.m * r};
}
__device__ inline D2& operator+=(D2& a, D2 b) { return a = a + b; }

__device__ inline double val(double x) { return x; }
__device__ inline double val(const D2& x) { return x.v; }
__device__ inline bool same(double a, double b) { return a == b; }
__device__ inline bool same(const D2& a, const D2& b) { return a.v == b.v && a.t == b.t && a.m == b.m; }
template <class R> __device__ inline R cst(double c);
template <> __device__ inline double cst<double>(double c) { return c; }
template <> __device__ inline D2 cst<D2>(double c) { return D2{c, 0.0, 0.0}; }

// E = sqrt(q) and 1/E (the node energies) as deriv_kernel computes them
__device__ inline void r_rsqrt_sqrt(double q, double& rs, double& sq) { pnjl::rsqrt_sqrt(q, rs, sq); }
__device__ inline void r_rsqrt_sqrt(D2 q, D2& rs, D2& sq)
{
    double r, s;
    pnjl::rsqrt_sqrt(q.v, r, s);
    const double ds = 0.5 * r, dr = -0.5 * r * r * r;
    sq = D2{s, q.t * ds, q.m * ds};
    rs = D2{r, q.t * dr, q.m * dr};
}
__device__ inline double r_exp_tab(double x, const double* tab) { return pnjl::exp_tab(x, tab); }
__device__ inline D2 r_exp_tab(D2 x, const double* tab)
{
    const double e = pnjl::exp_tab(x.v, tab);
    return D2{e, e * x.t, e * x.m};
}
__device__ inline double r_log(double x) { return log(x); }
__device__ inline D2 r_log(D2 x) { return D2{log(x.v), x.t / x.v, x.m / x.v}; }
__device__ inline double r_sqrt(double x) { return sqrt(x); }
__device__ inline D2 r_sqrt(D2 x)
{
    const double s = sqrt(x.v);
    return D2{s, 0.5 * x.t / s, 0.5 * x.m / s};
}
__device__ inline double r_fabs(double x) { return fabs(x); }
__device__ inline D2 r_fabs(D2 x) { return x.v < 0.0 ? -x : x; }

// lnF_slow's derivative terms (pnjl_device.h) for either number type
template <class R>
struct SignR {
    R gq, h, k1, k2, r1, r2;
};

template <class R>
__device__ void lnF_r(const double* tab, R a, R c1, R c2, SignR<R>& o)
{
    const R t = r_exp_tab(-r_fabs(a), tab);
    const R t2 = t * t, t3 = t2 * t;
    const bool pos = val(a) > 0.0;
    const R one = cst<R>(1.0);
    const R n0 = pos ? t3 : one, n1 = pos ? t2 : t, n2 = pos ? t : t2, n3 = pos ? one : t3;
    const R D = n0 + 3.0 * (c1 * n1 + c2 * n2) + n3;
    if (!(val(D) >= pnjl::kPolyakovEps)) {
        o.gq = cst<R>(pos ? 1.0 : 0.0);
        o.h = o.k1 = o.k2 = o.r1 = o.r2 = cst<R>(0.0);
        return;
    }
    const R iD = 1.0 / D;
    const R r1 = n1 * iD, r2 = n2 * iD, r3 = n3 * iD;
    const R f = c1 * r1 + 2.0 * c2 * r2 + r3;
    o.gq = f;
    o.h = (c1 * r1 + 4.0 * c2 * r2 + 3.0 * r3) - 3.0 * f * f;
    o.k1 = r1 * (1.0 - 3.0 * f);
    o.k2 = r2 * (2.0 - 3.0 * f);
    o.r1 = r1;
    o.r2 = r2;
}

// the node sums of one flavour that H and F_theta need (as flavour_d)
template <class R>
struct FlavR {
    R SM, SP, SPb, SMM, SMP, SMPb, SPP, SPPb, SPbPb, SMmu, SMT, SPmu, SPbmu, SPT, SPbT;
};

template <class R>
__device__ void flavour_r(const double* tab, int pn, int tn, R M, R mu, R T, double xi, R Phi, R Phib, FlavR<R>& o)
{
    const R z = cst<R>(0.0);
    o = FlavR<R>{z, z, z, z, z, z, z, z, z, z, z, z, z, z, z};
    const double* pp = tab + pnjl::kMathTab;
    const double* cc = pp + pn;
    const double* Cm = pp + pn + tn;
    const R invT = 1.0 / T, invT2 = invT * invT, M2 = M * M;
    for (int l = 0; l < tn; l++) {
        const double xc = xi * cc[l];
        for (int j = 0; j < pn; j++) {
            const double C = Cm[l * pn + j];
            R E, invE;
            r_rsqrt_sqrt(fma(xc, pp[j], pp[j]) + M2, invE, E);
            const R a = (mu - E) * invT, b = -(E + mu) * invT;
            SignR<R> sp, sm;
            lnF_r(tab, a, Phi, Phib, sp);
            lnF_r(tab, b, Phib, Phi, sm);
            const R gp = 3.0 * sp.gq, gm = 3.0 * sm.gq;
            const R hp = 3.0 * sp.h, hm = 3.0 * sm.h;
            const R ac1p = 3.0 * sp.k1, ac2p = 3.0 * sp.k2, ac1m = 3.0 * sm.k1, ac2m = 3.0 * sm.k2;
            const R QE = -(gp + gm) * invT, QEE = (hp + hm) * invT2;
            const R QEP = -(ac1p + ac2m) * invT, QEPb = -(ac2p + ac1m) * invT;
            const R EM = M * invE, EMM = (1.0 - EM * EM) * invE;
            o.SM += C * (QE * EM);
            o.SP += C * (3.0 * (sp.r1 + sm.r2));
            o.SPb += C * (3.0 * (sp.r2 + sm.r1));
            o.SMM += C * (QEE * EM * EM + QE * EMM);
            o.SMP += C * (QEP * EM);
            o.SMPb += C * (QEPb * EM);
            o.SPP += C * (-9.0 * (sp.r1 * sp.r1 + sm.r2 * sm.r2));
            o.SPPb += C * (-9.0 * (sp.r1 * sp.r2 + sm.r1 * sm.r2));
            o.SPbPb += C * (-9.0 * (sp.r2 * sp.r2 + sm.r1 * sm.r1));
            o.SMmu += C * (-(hp - hm) * invT2) * EM;
            o.SMT += C * ((gp + gm + hp * a + hm * b) * invT2) * EM;
            o.SPmu += C * (ac1p - ac2m) * invT;
            o.SPbmu += C * (ac2p - ac1m) * invT;
            o.SPT += C * (-(ac1p * a + ac2m * b) * invT);
            o.SPbT += C * (-(ac2p * a + ac1m * b) * invT);
        }
    }
}

// dI/dM and d2I/dM2 of the vacuum term (pnjl::vacuum_I)
template <class R>
__device__ void vacuum_r(const DevParams& P, R M, R& dI, R& d2I)
{
    const double L = P.Lambda;
    const double sgn = (val(M) < 0.0) ? -1.0 : 1.0;
    const R Mt = sgn * M + 1e-12;
    const R Mt2 = Mt * Mt;
    const R s = r_sqrt(L * L + Mt2);
    const R lg = r_log((L + s) / Mt);
    dI = sgn * Mt * (L * s - Mt2 * lg) * (1.0 / (4.0 * pnjl::kPi * pnjl::kPi));
    d2I = (2.0 * L * Mt2 / s + L * s - 3.0 * Mt2 * lg) * (1.0 / (4.0 * pnjl::kPi * pnjl::kPi));
}

// H = d2 Omega/dx2, FT = d2 Omega/dx dT, Fmu = d2 Omega/dx dmu at (x, T, mu)
// (the formulas of deriv_kernel)
template <class R>
__device__ void hess_theta(const DevParams& P, const double* tab, int pn, int tn, const R* x, R T, R mu, double xi,
                           R H[5][5], R FT[5], R Fmu[5])
{
    const double G = P.G, K = P.K;
    const R pu = x[0], pd = x[1], ps = x[2], Phi = x[3], Phib = x[4];
    const R M[3] = {P.m_ud0 - 4.0 * G * pu + 2.0 * K * pd * ps, P.m_ud0 - 4.0 * G * pd + 2.0 * K * pu * ps,
                    P.m_s0 - 4.0 * G * ps + 2.0 * K * pu * pd};
    const R c4G = cst<R>(-4.0 * G);
    const R dM[3][3] = {{c4G, 2.0 * K * ps, 2.0 * K * pd}, {2.0 * K * ps, c4G, 2.0 * K * pu},
                        {2.0 * K * pd, 2.0 * K * pu, c4G}};
    const R z = cst<R>(0.0);
    for (int a = 0; a < 5; a++) {
        for (int b = 0; b < 5; b++) H[a][b] = z;
        FT[a] = Fmu[a] = z;
    }
    H[0][0] = H[1][1] = H[2][2] = cst<R>(4.0 * G);
    H[0][1] = H[1][0] = -4.0 * K * ps;
    H[0][2] = H[2][0] = -4.0 * K * pd;
    H[1][2] = H[2][1] = -4.0 * K * pu;
    R SPP = z, SPPb = z, SPbPb = z;
    FlavR<R> fu;
    for (int i = 0; i < 3; i++) {
        FlavR<R> fo;
        const FlavR<R>* f = &fo;
        if (i == 1 && same(M[1], M[0]))   // phi_d == phi_u: d aliases u
            f = &fu;
        else
            flavour_r(tab, pn, tn, M[i], mu, T, xi, Phi, Phib, fo);
        if (i == 0) fu = fo;
        R dI, d2I;
        vacuum_r(P, M[i], dI, d2I);
        const R gM = -2.0 * P.Nc * dI - 2.0 * T * f->SM, hM = -2.0 * P.Nc * d2I - 2.0 * T * f->SMM;
        for (int a = 0; a < 3; a++) {
            for (int b = 0; b < 3; b++) {
                R h = hM * dM[i][a] * dM[i][b];
                if (a != b && a != i && b != i) h += gM * (2.0 * K);   // d2M_i / dphi_a dphi_b
                H[a][b] += h;
            }
            H[a][3] += -2.0 * T * f->SMP * dM[i][a];
            H[a][4] += -2.0 * T * f->SMPb * dM[i][a];
            Fmu[a] += dM[i][a] * (-2.0 * T * f->SMmu);
            FT[a] += dM[i][a] * (-2.0 * f->SM - 2.0 * T * f->SMT);
        }
        Fmu[3] += -2.0 * T * f->SPmu;
        Fmu[4] += -2.0 * T * f->SPbmu;
        FT[3] += -2.0 * f->SP - 2.0 * T * f->SPT;
        FT[4] += -2.0 * f->SPb - 2.0 * T * f->SPbT;
        SPP += f->SPP;
        SPPb += f->SPPb;
        SPbPb += f->SPbPb;
    }
    for (int a = 0; a < 3; a++) {
        H[3][a] = H[a][3];
        H[4][a] = H[a][4];
    }
    // Polyakov potential (Thermodynamics.jl:124-160) and its mixed T terms
    const R Tr = P.T0 / T;
    const R Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr, Tb = P.b3 * Tr * Tr * Tr;
    const R PP = Phi * Phib;
    const R v = 1.0 - 6.0 * PP + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * PP * PP;
    R lP = z, lPb = z, lPP = z, lPPb = z, lPbPb = z;
    if (val(v) >= pnjl::kPolyakovEps) {
        lP = (-6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib) / v;
        lPb = (-6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib) / v;
        lPP = (24.0 * Phi - 6.0 * Phib * Phib) / v - lP * lP;
        lPPb = (-6.0 - 12.0 * Phi * Phib) / v - lP * lPb;
        lPbPb = (24.0 * Phib - 6.0 * Phi * Phi) / v - lPb * lPb;
    }
    const R T4 = T * T * T * T;
    H[3][3] = T4 * (Tb * lPP) - 2.0 * T * SPP;
    H[3][4] = H[4][3] = T4 * (-0.5 * Ta + Tb * lPPb) - 2.0 * T * SPPb;
    H[4][4] = T4 * (Tb * lPbPb) - 2.0 * T * SPbPb;
    const R A1 = 4.0 * P.a0 * T * T * T + 3.0 * P.a1 * P.T0 * T * T + 2.0 * P.a2 * P.T0 * P.T0 * T;
    const double B1 = P.b3 * P.T0 * P.T0 * P.T0;
    FT[3] += -0.5 * Phib * A1 + B1 * lP;
    FT[4] += -0.5 * Phi * A1 + B1 * lPb;
}

struct Deriv2Args {
    DevParams P;
    int64_t n;
    const double *T, *mu, *xi, *x;
    const double* tab;
    int pn, tn;
    double* out;   // n * PNJL_IMPLICIT2_FIELDS
};

__global__ void __launch_bounds__(256) deriv2_kernel(Deriv2Args A)
{
    extern __shared__ double lds[];
    const int tab_len = pnjl::kMathTab + A.pn + A.tn + A.pn * A.tn + A.pn;
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = A.tab[i];
    __syncthreads();
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n) return;
    const DevParams& P = A.P;
    const double T = A.T[g], mu = A.mu[g], xi = A.xi[g];
    double x[5];
    for (int i = 0; i < 5; i++) x[i] = A.x[g * 5 + i];
    double* out = A.out + g * PNJL_IMPLICIT2_FIELDS;
    // first order: x_T, x_mu
    double H[5][5], FT[5], Fmu[5], L[5][5], xT[5], xmu[5];
    hess_theta<double>(P, lds, A.pn, A.tn, x, T, mu, xi, H, FT, Fmu);
    for (int a = 0; a < 5; a++) {
        xT[a] = -FT[a];
        xmu[a] = -Fmu[a];
        for (int b = 0; b < 5; b++) L[a][b] = H[a][b];
    }
    bool ok = lu5_solve2(L, xT, xmu);
    // second order: D_w of H and F along w_T = (x_T, 1, 0), w_mu = (x_mu, 0, 1)
    double x2[3][5];
    if (ok) {
        D2 xd[5], Hd[5][5], FTd[5], Fmud[5];
        for (int a = 0; a < 5; a++) xd[a] = D2{x[a], xT[a], xmu[a]};
        hess_theta<D2>(P, lds, A.pn, A.tn, xd, D2{T, 1.0, 0.0}, D2{mu, 0.0, 1.0}, xi, Hd, FTd, Fmud);
        double r2[5];
        for (int a = 0; a < 5; a++) {
            double tt = FTd[a].t, tm = FTd[a].m, mm = Fmud[a].m;
            for (int b = 0; b < 5; b++) {
                tt += Hd[a][b].t * xT[b];
                tm += Hd[a][b].m * xT[b];
                mm += Hd[a][b].m * xmu[b];
            }
            x2[0][a] = -tt;
            x2[1][a] = -tm;
            x2[2][a] = -mm;
            r2[a] = 0.0;
        }
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) L[a][b] = H[a][b];
        ok = lu5_solve2(L, x2[0], x2[1]);
        for (int a = 0; a < 5; a++)
            for (int b = 0; b < 5; b++) L[a][b] = H[a][b];
        ok = ok && lu5_solve2(L, x2[2], r2);
    }
    if (!ok) {
        for (int k = 0; k < PNJL_IMPLICIT2_FIELDS; k++) out[k] = __builtin_nan("");
        return;
    }
    for (int a = 0; a < 5; a++) {
        out[a] = x[a];
        out[5 + a] = xT[a];
        out[10 + a] = xmu[a];
        out[15 + a] = x2[0][a];
        out[20 + a] = x2[1][a];
        out[25 + a] = x2[2][a];
    }
    // masses as ThermoDerivatives.jl:111-121 writes them (bare 5.5 and 140.0
    // MeV): M_i = m_i0 - 4 G phi_i + 2 K phi_j phi_k, (j, k) the other two
    const double G = P.G, K = P.K;
    for (int i = 0; i < 3; i++) {
        const int j = (i == 0) ? 1 : 0, k = (i == 2) ? 1 : 2;
        const double m0 = (i == 2) ? 0.140 / 0.197327 : 0.0055 / 0.197327;
        out[30 + i] = m0 - 4.0 * G * x[i] + 2.0 * K * x[j] * x[k];
        out[33 + i] = -4.0 * G * xT[i] + 2.0 * K * (xT[j] * x[k] + x[j] * xT[k]);
        out[36 + i] = -4.0 * G * xmu[i] + 2.0 * K * (xmu[j] * x[k] + x[j] * xmu[k]);
        for (int q = 0; q < 3; q++) {
            const double* va = (q == 2) ? xmu : xT;   // (T, T), (T, mu), (mu, mu)
            const double* vb = (q == 0) ? xT : xmu;
            out[39 + 3 * q + i] = -4.0 * G * x2[q][i] +
                                  2.0 * K * (x2[q][j] * x[k] + x[j] * x2[q][k] + va[j] * vb[k] + vb[j] * va[k]);
        }
    }
}

}  // namespace pnjl_deriv

extern "C" {

int pnjl_thermo_derivatives_batch(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi,
                                  const double* x, int p_num, int t_num, double* out, int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!T || !mu || !xi || !x || !out)) || p_num <= 0 || t_num <= 0 ||
        (int64_t)p_num * t_num > 4096)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_thermo_derivatives_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    double* tab = nullptr;
    if (int rc = get_tab(ctx, p_num, t_num, &tab)) return rc;
    const double *dT = T, *dmu = mu, *dxi = xi, *dx = x;
    double* dout = out;
    const size_t nd = (size_t)n * sizeof(double);
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    if (!dev) {
        void* b[5];
        const size_t sizes[5] = {nd, nd, nd, 5 * nd, PNJL_THERMO_DERIV_FIELDS * nd};
        for (int k = 0; k < 5; k++)
            if (int rc = stage_buf(ctx, 70 + k, sizes[k], &b[k])) return rc;
        HIPCHK(ctx, hipMemcpyAsync(b[0], T, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[1], mu, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[2], xi, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[3], x, 5 * nd, hipMemcpyHostToDevice, st));
        dT = (const double*)b[0];
        dmu = (const double*)b[1];
        dxi = (const double*)b[2];
        dx = (const double*)b[3];
        dout = (double*)b[4];
    }
    pnjl_deriv::DerivArgs A{};
    A.P = ctx->dp;
    A.n = n;
    A.T = dT;
    A.mu = dmu;
    A.xi = dxi;
    A.x = dx;
    A.tab = tab;
    A.pn = p_num;
    A.tn = t_num;
    A.out = dout;
    const size_t lds = sizeof(double) * (pnjl::kMathTab + (size_t)p_num + t_num + (size_t)p_num * t_num + p_num);
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(pnjl_deriv::deriv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), lds, st, A);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!dev)
        HIPCHK(ctx, hipMemcpyAsync(out, dout, PNJL_THERMO_DERIV_FIELDS * nd, hipMemcpyDeviceToHost, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

int pnjl_implicit_derivatives2_batch(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi,
                                     const double* x, int p_num, int t_num, double* out, int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || (n > 0 && (!T || !mu || !xi || !x || !out)) || p_num <= 0 || t_num <= 0 ||
        (int64_t)p_num * t_num > 4096)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_implicit_derivatives2_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    double* tab = nullptr;
    if (int rc = get_tab(ctx, p_num, t_num, &tab)) return rc;
    const double *dT = T, *dmu = mu, *dxi = xi, *dx = x;
    double* dout = out;
    const size_t nd = (size_t)n * sizeof(double);
    const bool dev = flags & PNJL_FLAG_DEVICE_PTRS;
    if (!dev) {
        void* b[5];
        const size_t sizes[5] = {nd, nd, nd, 5 * nd, PNJL_IMPLICIT2_FIELDS * nd};
        for (int k = 0; k < 5; k++)
            if (int rc = stage_buf(ctx, 75 + k, sizes[k], &b[k])) return rc;
        HIPCHK(ctx, hipMemcpyAsync(b[0], T, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[1], mu, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[2], xi, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[3], x, 5 * nd, hipMemcpyHostToDevice, st));
        dT = (const double*)b[0];
        dmu = (const double*)b[1];
        dxi = (const double*)b[2];
        dx = (const double*)b[3];
        dout = (double*)b[4];
    }
    pnjl_deriv::Deriv2Args A{};
    A.P = ctx->dp;
    A.n = n;
    A.T = dT;
    A.mu = dmu;
    A.xi = dxi;
    A.x = dx;
    A.tab = tab;
    A.pn = p_num;
    A.tn = t_num;
    A.out = dout;
    const size_t lds = sizeof(double) * (pnjl::kMathTab + (size_t)p_num + t_num + (size_t)p_num * t_num + p_num);
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(pnjl_deriv::deriv2_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), lds, st, A);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!dev)
        HIPCHK(ctx, hipMemcpyAsync(out, dout, PNJL_IMPLICIT2_FIELDS * nd, hipMemcpyDeviceToHost, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !dev) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

}  // extern "C"
