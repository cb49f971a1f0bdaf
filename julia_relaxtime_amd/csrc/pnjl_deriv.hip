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

}  // extern "C"
