// pnjl_scan_kernels.h -- persistent work-queue kernel that runs a whole (T, mu, xi)
// grid of PNJL FixedMu gap solves in one launch, FP64, gfx950.
//
// Geometry: a "group" of LPP lanes (LPP = 1..64, a template parameter picked
// by the autotuner) owns one work item at a time; 64/LPP groups share a
// wavefront.  Each lane integrates every LPP-th momentum node; the group's
// partial sums are combined with xor-butterfly shuffles, so every lane of a
// group holds bit-identical sums and runs the identical control logic.
// Items (points, or continuity chains) are pulled from a device work queue
// (one atomic counter, `chunk` items per pop), so points that need 3 and 30
// Newton iterations share the launch and no group idles while work is left.
//
// One trip of the loop = one node pass at the group's current state x
// (Omega, F = grad P, J = dF/dx and the thermodynamic sums, all fused) plus
// one step of the solver state machine, which follows
//   NLsolve newton_   (method=:newton, Static line search)       ImplicitSolver.jl:112
//   NLsolve trust_region_ (dogleg, autoscale)                     ImplicitSolver.jl:144
//   _nlsolve_with_tr_fallback / _choose_candidate                 ImplicitSolver.jl:66-151
//   solve(::FixedMu) with the automatic MultiSeed fallback         ImplicitSolver.jl:211-328
//   solve_multi + default_omega_selector                          ImplicitSolver.jl:532-559
//   ContinuitySeed / PhaseAwareContinuitySeed chains               SeedStrategies.jl:303-345, 762-888
//
// Register budget: the node pass is the hot loop and keeps one flavour's 13
// accumulators live; reduced sums are parked in a per-group scratch area in
// HBM ([group][field], every lane writes and re-reads only its own writes),
// and the control step (a separate non-inlined function) rebuilds Omega, F
// and the packed symmetric Jacobian from them.  Trust-region state and
// candidate records also live in that scratch, so the common Newton path
// holds only x, xold and a 5x5 LU in registers.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/pnjl_hip.h"
#include "pnjl_device.h"

namespace pnjl {

__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000ULL); }

// ---------------------------------------------------------------------------
// kernel arguments
// ---------------------------------------------------------------------------
struct ScanArgs {
    DevParams P;
    int64_t n_points;
    int64_t n_items;      // points, or chains for the chain policies
    int policy;
    int chain_len;
    const double* T;
    const double* mu;
    const double* xi;
    const double* seed;   // n*5 or null
    // outputs (device pointers; optional ones may be null)
    double* x_out;
    double* omega;
    double* pressure;
    double* rho_norm;
    double* entropy;
    double* energy;
    double* masses;
    double* n_q;
    double* n_qbar;
    double* resnorm;
    int32_t* iterations;
    int32_t* converged;
    int32_t* flags;
    int32_t* steps;
    int32_t* passes;
    // node table (staged in LDS): pp[pn] | cc[tn] | C[pn*tn] | Cs[pn]
    const double* tab;
    int pn, tn;
    // solver options
    double xtol, ftol, resmax, phi_tol;
    int max_iter, tr_fb, auto_ms, surr_cap;
    int judge_mode;                         // PNJL_JUDGE_* (SURROGATE inside the mask)
    // work queue and scratch
    unsigned long long* queue;
    int chunk;
    double* scratch;
    int64_t n_groups;
    // phase-boundary tables: nb sets of {xi, T_CEP, n, pad, T[kMaxBndRows], mu[kMaxBndRows]}
    const double* bnd;
    int nb;
    int64_t max_trips;    // hard cap on trips per point (safety exit)
    // staged pipeline for the per-point policies (stage != STAGE_NONE): each
    // item is one single-seed attempt; results go to rec_out, failures are
    // appended to next_list for the following stage
    int stage;
    int mult;                               // attempts per listed point
    const int32_t* item_list;               // points of this stage (null: all points)
    const unsigned long long* item_count;   // device count of item_list
    double* rec_out;                        // REC_SIZE doubles per record
    int32_t* next_list;
    unsigned long long* next_count;
    int32_t* mslot;                         // point -> its index in the MultiSeed list
    // learned-seed judge: a converged root inside the ambiguity mask is
    // appended to judge_list for the Omega check against the MultiSeed
    // candidates (HADRON and hint-QUARK first; where one of them fails, the
    // other four candidates via judge2_list)
    const uint8_t* judge_mask;              // [nT][nmu][nxi] bins over [jlo, jhi]
    int jdim[3];
    double jlo[3], jhi[3];
    int32_t* judge_list;
    unsigned long long* judge_count;
    int32_t* jslot;                         // point -> its index in the judge list (-1 none)
    int32_t* jb;                            // judge index -> 1 once its point is in judge2_list
    int32_t* judge2_list;
    unsigned long long* judge2_count;
};

// stages of the per-point pipeline
enum Stage {
    STAGE_NONE = 0,
    STAGE_SURR = 1,
    STAGE_PRIMARY = 2,
    STAGE_MULTI = 3,
    STAGE_MR3 = 4,
    STAGE_JUDGE = 5,    // MultiSeed candidates 0, 1 (HADRON, hint-QUARK) of a judged point
    STAGE_JUDGE2 = 6    // candidates 2..5, where candidate 0 or 1 did not converge
};
// judge records per judged point: the six MultiSeed candidates
constexpr int JUDGE_SEEDS = 6;

constexpr int kMaxBndRows = 128;
constexpr int kBndStride = 4 + 2 * kMaxBndRows;

// ---------------------------------------------------------------------------
// scratch layout (doubles, [group][field])
// ---------------------------------------------------------------------------
enum RecField {
    RF_X = 0,          // 5
    RF_OMEGA = 5,
    RF_P = 6,
    RF_RHO = 7,
    RF_S = 8,
    RF_E = 9,
    RF_M = 10,         // 3
    RF_NQ = 13,        // 3
    RF_NQB = 16,       // 3
    RF_RES = 19,
    RF_IT = 20,
    RF_FCONV = 21,
    RF_PHYS = 22,
    RF_CONV = 23,
    RF_FLAGS = 24,
    RF_EXC = 25,
    RF_XCONV = 26,
    REC_SIZE = 28
};
enum ScratchBlock {
    SB_PRIM = 0,                 // primary (Newton) candidate of the attempt
    SB_TRC = SB_PRIM + REC_SIZE, // trust-region candidate at the accepted x
    SB_SINGLE = SB_TRC + REC_SIZE,
    SB_BEST = SB_SINGLE + REC_SIZE,
    SB_TR = SB_BEST + REC_SIZE,  // r[5] Jsym[15] d[5] p[5] delta
    SB_X0 = SB_TR + 31,          // seed of the attempt
    SB_PREV = SB_X0 + 5,         // chain: previous solution
    SB_SUMS = SB_PREV + 5,       // reduced node sums of the last pass: u, s, d (13 each)
    SCRATCH_FIELDS = SB_SUMS + 39
};
enum TrField { TR_R = 0, TR_J = 5, TR_D = 20, TR_P = 25, TR_DELTA = 30 };

// a group's scratch record: SCRATCH_FIELDS contiguous doubles, so every field
// is one base address plus an immediate offset
struct Scr {
    double* p;
    __device__ __forceinline__ double& at(int f) const { return p[f]; }
};

// packed index of the symmetric 5x5 Jacobian (upper triangle, row-major)
__host__ __device__ constexpr int sidx(int i, int j)
{
    return (i <= j) ? (i * 5 - (i * (i - 1)) / 2 + (j - i)) : (j * 5 - (j * (j - 1)) / 2 + (i - j));
}

// ---------------------------------------------------------------------------
// node pass
// ---------------------------------------------------------------------------
// v from the lane selected by the DPP pattern CTRL (within its row of 16)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

enum DppCtrl {
    DPP_QUAD_XOR1 = 0xB1,       // quad_perm [1, 0, 3, 2]
    DPP_QUAD_XOR2 = 0x4E,       // quad_perm [2, 3, 0, 1]
    DPP_ROW_MIRROR = 0x140,     // lane i <- 15 - i within the row
    DPP_ROW_HALF_MIRROR = 0x141 // lane i <- 7 - i within each half row
};

// Sum over the LPP lanes of a group (aligned, LPP a power of two), returned
// in every lane of the group.  Steps across rows use lane shuffles; within a
// row, DPP: mirror (16), half mirror (8), then the quad butterfly.  Every
// step adds a value and its partner's, so all lanes of the group end with
// the same pairwise sums in the same tree, i.e. bit-identical totals.
template <int LPP>
__device__ __forceinline__ double group_sum(double v)
{
    if (LPP >= 64) v += __shfl_xor(v, 32, 64);
    if (LPP >= 32) v += __shfl_xor(v, 16, 64);
    if (LPP >= 16) v += dpp_f64<DPP_ROW_MIRROR>(v);
    if (LPP >= 8) v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
    if (LPP >= 2) v += dpp_f64<DPP_QUAD_XOR1>(v);
    if (LPP >= 4) v += dpp_f64<DPP_QUAD_XOR2>(v);
    return v;
}

template <int LPP>
__device__ __forceinline__ void reduce_flav(FlavSums& s)
{
    if (LPP > 1) {
        s.S = group_sum<LPP>(s.S);
        s.A1 = group_sum<LPP>(s.A1);
        s.A2 = group_sum<LPP>(s.A2);
        s.SP = group_sum<LPP>(s.SP);
        s.SPb = group_sum<LPP>(s.SPb);
        s.B1 = group_sum<LPP>(s.B1);
        s.B2 = group_sum<LPP>(s.B2);
        s.SPP = group_sum<LPP>(s.SPP);
        s.SPPb = group_sum<LPP>(s.SPPb);
        s.SPbPb = group_sum<LPP>(s.SPbPb);
        s.Sfp = group_sum<LPP>(s.Sfp);
        s.Sfm = group_sum<LPP>(s.Sfm);
        s.Sfab = group_sum<LPP>(s.Sfab);
    }
}

// sums slot of flavour f: 0 -> u, 1 -> s, 2 -> d
enum SumField { SF_S = 0, SF_A1, SF_A2, SF_SP, SF_SPb, SF_B1, SF_B2, SF_SPP, SF_SPPb, SF_SPbPb, SF_Sfp, SF_Sfm, SF_Sfab };

__device__ __forceinline__ void store_flav(const Scr& s, int f, const FlavSums& a)
{
    const int b = SB_SUMS + 13 * f;
    s.at(b + SF_S) = a.S;
    s.at(b + SF_A1) = a.A1;
    s.at(b + SF_A2) = a.A2;
    s.at(b + SF_SP) = a.SP;
    s.at(b + SF_SPb) = a.SPb;
    s.at(b + SF_B1) = a.B1;
    s.at(b + SF_B2) = a.B2;
    s.at(b + SF_SPP) = a.SPP;
    s.at(b + SF_SPPb) = a.SPPb;
    s.at(b + SF_SPbPb) = a.SPbPb;
    s.at(b + SF_Sfp) = a.Sfp;
    s.at(b + SF_Sfm) = a.Sfm;
    s.at(b + SF_Sfab) = a.Sfab;
}

// This lane's share of the node sums of one flavour: nodes j = gl, gl + LPP, ...
template <int LPP, bool kChecked>
__device__ __forceinline__ void node_loop(const double* lds, const double* pp, const double* cc, const double* Cm,
                                          const double* Cs, int pn, int tn, double M2, double muT, double invT,
                                          double K2mu, double xi, double Phi, double Phib, int gl, FlavSums& acc)
{
    if (xi == 0.0) {
        // E does not depend on cos(theta): the cos nodes collapse into Cs
#pragma unroll 1
        for (int j = gl; j < pn; j += LPP)
            node_accumulate<kChecked>(lds, pp[j], Cs[j], M2, muT, invT, K2mu, Phi, Phib, acc);
    } else {
#pragma unroll 1
        for (int l = 0; l < tn; l++) {
            double xc = xi * cc[l];
            const double* Cl = Cm + l * pn;
#pragma unroll 1
            for (int j = gl; j < pn; j += LPP) {
                double q = fma(xc, pp[j], pp[j]);
                node_accumulate<kChecked>(lds, q, Cl[j], M2, muT, invT, K2mu, Phi, Phib, acc);
            }
        }
    }
}

template <int LPP>
__device__ __forceinline__ void flavour_sums(const double* lds, int pn, int tn, double M, double T, double mu,
                                             double xi, double Phi, double Phib, int gl, FlavSums& acc)
{
    const double* pp = lds + kMathTab;
    const double* cc = pp + pn;
    const double* Cm = pp + pn + tn;
    const double* Cs = Cm + pn * tn;
    const double invT = 1.0 / T;
    const double muT = mu * invT;
    const double K2mu = exp(-2.0 * muT);
    const double M2 = M * M;
    flav_zero(acc);
    // one loop per wave: a state that needs the checked loop (a negative
    // Polyakov loop during an iteration, |mu|/T > 100) takes its whole wave
    // with it, instead of the wave running both loops one after the other
    if (__any(!nodes_all_fast(Phi, Phib, muT)))
        node_loop<LPP, true>(lds, pp, cc, Cm, Cs, pn, tn, M2, muT, invT, K2mu, xi, Phi, Phib, gl, acc);
    else
        node_loop<LPP, false>(lds, pp, cc, Cm, Cs, pn, tn, M2, muT, invT, K2mu, xi, Phi, Phib, gl, acc);
    reduce_flav<LPP>(acc);
    flav_finish(acc, invT);
}

// Node pass at state x: u, s, and d unless M_d == M_u (then d aliases u).
// Not inlined: the node loop gets the whole register budget, and the caller's
// live state is saved once per pass instead of competing with the loop.
template <int LPP>
__device__ __noinline__ void node_pass(const DevParams& P, const double* lds, int pn, int tn, const Scr& s,
                                          const double* x, double T, double mu, double xi, bool sym, int gl)
{
    double M[3];
    mass_vec(P, x, M);
    const int nf = sym ? 2 : 3;
#pragma unroll 1
    for (int f = 0; f < nf; f++) {
        double Mf = (f == 0) ? M[0] : ((f == 1) ? M[2] : M[1]);
        FlavSums acc;
        flavour_sums<LPP>(lds, pn, tn, Mf, T, mu, xi, x[3], x[4], gl, acc);
        store_flav(s, f, acc);
    }
}

// scratch slot holding the sums of mass index i (0 u, 1 d, 2 s)
__device__ __forceinline__ int slot_of(int i, bool sym) { return i == 0 ? 0 : (i == 2 ? 1 : (sym ? 0 : 2)); }

// ---------------------------------------------------------------------------
// Omega, F and the packed symmetric Jacobian from the sums in scratch
// ---------------------------------------------------------------------------
__device__ void assemble_core(const DevParams& P, const Scr& s, const double* x, double T, bool sym, double& omega,
                              double F[5], double Jp[15])
{
    const double G = P.G, K = P.K;
    const double pu = x[0], pd = x[1], ps = x[2], Phi = x[3], Phib = x[4];
    double M[3];
    mass_vec(P, x, M);
    // Polyakov potential (Thermodynamics.jl:124-131)
    double Tr = P.T0 / T;
    double Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr;
    double Tb = P.b3 * Tr * Tr * Tr;
    double PP = Phi * Phib;
    double val = 1.0 - 6.0 * PP + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * PP * PP;
    bool floored = !(val >= kPolyakovEps);
    double lg = floored ? log(kPolyakovEps) : log(val);
    double T2 = T * T, T4 = T2 * T2;
    double lP = 0, lPb = 0, lPP = 0, lPPb = 0, lPbPb = 0;
    if (!floored) {
        double iv = 1.0 / val;
        lP = (-6.0 * Phib + 12.0 * Phi * Phi - 6.0 * Phi * Phib * Phib) * iv;
        lPb = (-6.0 * Phi + 12.0 * Phib * Phib - 6.0 * Phi * Phi * Phib) * iv;
        lPP = (24.0 * Phi - 6.0 * Phib * Phib) * iv - lP * lP;
        lPPb = (-6.0 - 12.0 * PP) * iv - lP * lPb;
        lPbPb = (24.0 * Phib - 6.0 * Phi * Phi) * iv - lPb * lPb;
    }
    double chi = 2.0 * G * (pu * pu + pd * pd + ps * ps) - 4.0 * K * (pu * pd * ps);
    double acc = chi + T4 * (-0.5 * Ta * PP + Tb * lg);
    double gM[3], hM[3], bM1[3], bM2[3];
    double SP = 0, SPb = 0, SPP = 0, SPPb = 0, SPbPb = 0;
    const double m2T = -2.0 * T, m2Nc = -2.0 * P.Nc;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int b = SB_SUMS + 13 * slot_of(i, sym);
        double I, dI, d2I;
        vacuum_I(P, M[i], I, dI, d2I);
        double A1 = s.at(b + SF_A1);
        acc += m2Nc * I + m2T * s.at(b + SF_S);
        gM[i] = m2Nc * dI + m2T * (M[i] * A1);
        hM[i] = m2Nc * d2I + m2T * (A1 + M[i] * M[i] * s.at(b + SF_A2));
        bM1[i] = m2T * (M[i] * s.at(b + SF_B1));
        bM2[i] = m2T * (M[i] * s.at(b + SF_B2));
        SP += s.at(b + SF_SP);
        SPb += s.at(b + SF_SPb);
        SPP += s.at(b + SF_SPP);
        SPPb += s.at(b + SF_SPPb);
        SPbPb += s.at(b + SF_SPbPb);
    }
    omega = acc;
    // dM_i/dphi_a: diagonal -4G, off-diagonal 2K phi_(third index)
    const double phv[3] = {pu, pd, ps};
#define DM(i, a) ((i) == (a) ? -4.0 * G : 2.0 * K * phv[3 - (i) - (a)])
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double dchi = 4.0 * G * phv[a] - 4.0 * K * phv[(a + 1) % 3] * phv[(a + 2) % 3];
        double g = dchi, hP = 0.0, hPb = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            g += gM[i] * DM(i, a);
            hP += bM1[i] * DM(i, a);
            hPb += bM2[i] * DM(i, a);
        }
        F[a] = -g;
        Jp[sidx(a, 3)] = -hP;
        Jp[sidx(a, 4)] = -hPb;
#pragma unroll
        for (int b = a; b < 3; b++) {
            double h = (a == b) ? 4.0 * G : -4.0 * K * phv[3 - a - b];
#pragma unroll
            for (int i = 0; i < 3; i++) h += hM[i] * DM(i, a) * DM(i, b);
            if (a != b) h += gM[3 - a - b] * 2.0 * K;
            Jp[sidx(a, b)] = -h;
        }
    }
#undef DM
    F[3] = -(T4 * (-0.5 * Ta * Phib + Tb * lP) + m2T * SP);
    F[4] = -(T4 * (-0.5 * Ta * Phi + Tb * lPb) + m2T * SPb);
    Jp[sidx(3, 3)] = -(T4 * (Tb * lPP) + m2T * SPP);
    Jp[sidx(3, 4)] = -(T4 * (-0.5 * Ta + Tb * lPPb) + m2T * SPPb);
    Jp[sidx(4, 4)] = -(T4 * (Tb * lPbPb) + m2T * SPbPb);
}

// Candidate record at x: thermodynamics (calculate_thermo + densities, in
// closed form from the sums) and the physicality gate
// (_default_is_physical_solution + _all_finite_thermo, ImplicitSolver.jl:50-64).
__device__ void put_rec(const DevParams& P, const Scr& s, int blk, const double* x, double mu, double T, bool sym,
                        double phi_tol, double omega, double res, int it, bool fconv, bool xconv)
{
    double M[3];
    mass_vec(P, x, M);
    const double Phi = x[3], Phib = x[4];
    double Tr = P.T0 / T;
    double Ta = P.a0 + P.a1 * Tr + P.a2 * Tr * Tr;
    double Tb = P.b3 * Tr * Tr * Tr;
    double PP = Phi * Phib;
    double val = 1.0 - 6.0 * PP + 4.0 * (Phib * Phib * Phib + Phi * Phi * Phi) - 3.0 * PP * PP;
    double lg = (val >= kPolyakovEps) ? log(val) : log(kPolyakovEps);
    double T2 = T * T, T4 = T2 * T2;
    double dTa = -P.a1 * P.T0 / T2 - 2.0 * P.a2 * P.T0 * P.T0 / (T2 * T);
    double dTb = -3.0 * P.b3 * P.T0 * P.T0 * P.T0 / T4;
    double UT = 4.0 * T2 * T * (-0.5 * Ta * PP + Tb * lg) + T4 * (dTa * (-0.5 * PP) + dTb * lg);
    double rsum = 0.0, Sab = 0.0, Ssum = 0.0;
    const double twoNc = 2.0 * P.Nc;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int b = SB_SUMS + 13 * slot_of(i, sym);
        double sfp = s.at(b + SF_Sfp), sfm = s.at(b + SF_Sfm);
        rsum += 6.0 * (sfp - sfm);
        Sab += s.at(b + SF_Sfab);
        Ssum += s.at(b + SF_S);
        s.at(blk + RF_NQ + i) = twoNc * sfp;
        s.at(blk + RF_NQB + i) = twoNc * sfm;
        s.at(blk + RF_M + i) = M[i];
    }
    double pressure = -omega;
    double rho_norm = rsum / (3.0 * P.rho0);
    double entropy = -UT + 2.0 * Ssum - 6.0 * Sab;
    double energy = -pressure + mu * rsum + T * entropy;
    bool ph = isfinite(Phi) && isfinite(Phib) && (-phi_tol <= Phi) && (Phi <= 1.0 + phi_tol) && (-phi_tol <= Phib) &&
              (Phib <= 1.0 + phi_tol);
#pragma unroll
    for (int i = 0; i < 3; i++) ph = ph && isfinite(M[i]) && (M[i] > 0.0);
    ph = ph && isfinite(omega) && isfinite(pressure) && isfinite(rho_norm) && isfinite(entropy) && isfinite(energy);
#pragma unroll
    for (int i = 0; i < 5; i++) s.at(blk + RF_X + i) = x[i];
    s.at(blk + RF_OMEGA) = omega;
    s.at(blk + RF_P) = pressure;
    s.at(blk + RF_RHO) = rho_norm;
    s.at(blk + RF_S) = entropy;
    s.at(blk + RF_E) = energy;
    s.at(blk + RF_RES) = res;
    s.at(blk + RF_IT) = (double)it;
    s.at(blk + RF_FCONV) = fconv ? 1.0 : 0.0;
    s.at(blk + RF_XCONV) = xconv ? 1.0 : 0.0;
    s.at(blk + RF_PHYS) = ph ? 1.0 : 0.0;
    s.at(blk + RF_EXC) = 0.0;
    s.at(blk + RF_FLAGS) = 0.0;
}

__device__ void put_exc_rec(const Scr& s, int blk, const double* x0)
{
#pragma unroll
    for (int i = 0; i < 5; i++) s.at(blk + RF_X + i) = x0[i];
    for (int f = RF_OMEGA; f <= RF_RES; f++) s.at(blk + f) = qnan();
    s.at(blk + RF_IT) = 0.0;
    s.at(blk + RF_FCONV) = 0.0;
    s.at(blk + RF_XCONV) = 0.0;
    s.at(blk + RF_PHYS) = 0.0;
    s.at(blk + RF_EXC) = 1.0;
    s.at(blk + RF_FLAGS) = (double)PNJL_PT_EXCEPTION;
}

__device__ void copy_rec(const Scr& s, int dst, int src)
{
    for (int f = 0; f < REC_SIZE; f++) s.at(dst + f) = s.at(src + f);
}

__device__ __forceinline__ bool rec_good(const Scr& s, int blk, double resmax)
{
    double r = s.at(blk + RF_RES);
    return s.at(blk + RF_FCONV) != 0.0 && isfinite(r) && r <= resmax && s.at(blk + RF_PHYS) != 0.0;
}

// ---------------------------------------------------------------------------
// small dense linear algebra, fully unrolled (static register indices)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unpack5(const double Jp[15], double A[5][5])
{
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) A[i][j] = Jp[sidx(i, j)];
}

// In-place LU with partial pivoting (largest |pivot|, first on ties), then
// solves A x = b (b is overwritten by x).  Returns true when a pivot is
// exactly zero (Julia's SingularException).
__device__ __forceinline__ bool lu_solve5(double A[5][5], double b[5])
{
    bool sing = false;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        int p = k;
        double best = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < 5; i++) {
            double v = fabs(A[i][k]);
            bool gt = v > best;
            best = gt ? v : best;
            p = gt ? i : p;
        }
#pragma unroll
        for (int i = k + 1; i < 5; i++) {
            bool sw = (p == i);
#pragma unroll
            for (int j = k; j < 5; j++) {
                double t = A[k][j];
                A[k][j] = sw ? A[i][j] : t;
                A[i][j] = sw ? t : A[i][j];
            }
            double t = b[k];
            b[k] = sw ? b[i] : t;
            b[i] = sw ? t : b[i];
        }
        sing = sing || (A[k][k] == 0.0);
        double ip = 1.0 / A[k][k];
#pragma unroll
        for (int i = k + 1; i < 5; i++) {
            double l = A[i][k] * ip;
#pragma unroll
            for (int j = k + 1; j < 5; j++) A[i][j] -= l * A[k][j];
            b[i] -= l * b[k];
        }
    }
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double t = b[i];
#pragma unroll
        for (int j = i + 1; j < 5; j++) t -= A[i][j] * b[j];
        b[i] = t / A[i][i];
    }
    return sing;
}

__device__ __forceinline__ double norm_inf5(const double* v)
{
    double m = 0.0;
    bool nan = false;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double a = fabs(v[i]);
        nan = nan || (a != a);
        m = a > m ? a : m;
    }
    return nan ? qnan() : m;
}

__device__ __forceinline__ bool any_nan5(const double* v)
{
    bool r = false;
#pragma unroll
    for (int i = 0; i < 5; i++) r = r || (v[i] != v[i]);
    return r;
}

__device__ __forceinline__ bool all_finite5(const double* v)
{
    bool r = true;
#pragma unroll
    for (int i = 0; i < 5; i++) r = r && isfinite(v[i]);
    return r;
}

__device__ __forceinline__ double wnorm5(const double* d, const double* v)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) s += (d[i] * v[i]) * (d[i] * v[i]);
    return sqrt(s);
}

__device__ __forceinline__ double sumsq5(const double* v)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) s += v[i] * v[i];
    return s;
}

// Newton direction p = -J^{-1} F; on a singular J the Levenberg-modified
// direction p = -(J'J + lambda I)^{-1} J'F, lambda = 1e6 sqrt(n eps) |J'J|_1
// (NLsolve newton_; the norm is entrywise).
__device__ __noinline__ void newton_dir(const double Jp[15], const double F[5], double p[5])
{
    double A[5][5];
    unpack5(Jp, A);
#pragma unroll
    for (int i = 0; i < 5; i++) p[i] = F[i];
    if (!lu_solve5(A, p)) {
#pragma unroll
        for (int i = 0; i < 5; i++) p[i] = -p[i];
        return;
    }
    // J is symmetric: J'J = J J, J'F = J F
    double n1 = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) t += Jp[sidx(k, i)] * F[k];
        p[i] = t;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            double u = 0.0;
#pragma unroll
            for (int k = 0; k < 5; k++) u += Jp[sidx(k, i)] * Jp[sidx(k, j)];
            A[i][j] = u;
            n1 += fabs(u);
        }
    }
    double lambda = 1e6 * sqrt(5.0 * 2.220446049250313e-16) * n1;
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) A[i][j] = -(A[i][j] + (i == j ? lambda : 0.0));
    if (lu_solve5(A, p)) {
#pragma unroll
        for (int i = 0; i < 5; i++) p[i] = qnan();
    }
}

// NLsolve dogleg! with J (packed, symmetric) read from the TR block.  (An
// exactly singular J, where the reference takes a pseudo-inverse step, uses
// the scaled steepest-descent branch.)
__device__ __noinline__ void dogleg5(const Scr& s, double p[5], const double r[5], const double d[5], double delta)
{
    double pi_[5];
    {
        double A[5][5];
#pragma unroll
        for (int i = 0; i < 5; i++)
#pragma unroll
            for (int j = 0; j < 5; j++) A[i][j] = s.at(SB_TR + TR_J + sidx(i, j));
#pragma unroll
        for (int i = 0; i < 5; i++) pi_[i] = r[i];
        bool sing = lu_solve5(A, pi_);
#pragma unroll
        for (int i = 0; i < 5; i++) pi_[i] = sing ? INFINITY : -pi_[i];
    }
    if (wnorm5(d, pi_) <= delta) {
#pragma unroll
        for (int i = 0; i < 5; i++) p[i] = pi_[i];
        return;
    }
    double g[5], Jg[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) t += s.at(SB_TR + TR_J + sidx(k, j)) * r[k];
        g[j] = t / (d[j] * d[j]);
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) t += s.at(SB_TR + TR_J + sidx(i, k)) * g[k];
        Jg[i] = t;
    }
    double wg = wnorm5(d, g);
    double coef = -(wg * wg) / sumsq5(Jg);
    double pc[5];
#pragma unroll
    for (int i = 0; i < 5; i++) pc[i] = coef * g[i];
    if (wnorm5(d, pc) >= delta) {
        double sc = -delta / wg;
#pragma unroll
        for (int i = 0; i < 5; i++) p[i] = g[i] * sc;
        return;
    }
    double pdiff[5];
#pragma unroll
    for (int i = 0; i < 5; i++) pdiff[i] = pi_[i] - pc[i];
    double b = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) b += (d[i] * pc[i]) * (d[i] * pdiff[i]);
    b *= 2.0;
    double a = wnorm5(d, pdiff);
    a = a * a;
    double wpc = wnorm5(d, pc);
    double tau = (-b + sqrt(b * b - 4.0 * a * (wpc * wpc - delta * delta))) / (2.0 * a);
#pragma unroll
    for (int i = 0; i < 5; i++) p[i] = pc[i] + tau * pdiff[i];
}

// ---------------------------------------------------------------------------
// seeds (SeedStrategies.jl:56-107, 131-136, 204-284)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void seed_const(int which, double x0[5])
{
    // 0 HADRON, 1 HIGH_TEMP(=QUARK), 2 VERY_HIGH_TEMP(=HT_0p9), 3 WEAK_CHIRAL_CONF, 4 HT_0p8, 5 HT_0p95
    const double tab[6][5] = {{-1.84329, -1.84329, -2.22701, 1.0e-5, 4.0e-5},
                              {-0.73192, -0.73192, -1.79539, 0.60532, 0.60532},
                              {-0.30, -0.30, -0.90, 0.90, 0.90},
                              {-0.50, -0.50, -1.20, 1e-3, 1e-3},
                              {-0.50, -0.50, -1.20, 0.80, 0.80},
                              {-0.20, -0.20, -0.70, 0.95, 0.95}};
#pragma unroll
    for (int k = 0; k < 6; k++)
        if (which == k)
#pragma unroll
            for (int i = 0; i < 5; i++) x0[i] = tab[k][i];
}

// get_seed(DefaultSeed(phase_hint)); hint 0 auto, 1 hadron, 2 quark
__device__ __forceinline__ void default_seed(int hint, double T, double mu, double x0[5])
{
    const double hb = 197.327;
    if (hint == 0) hint = (T * hb > 150.0 || mu * hb > 300.0) ? 2 : 1;
    if (hint == 2)
        seed_const((T * hb >= 300.0) ? 2 : 1, x0);
    else
        seed_const(0, x0);
}

// MultiSeed() candidate k (get_all_seeds order)
__device__ __forceinline__ void multi_seed(int k, double T, double mu, double x0[5])
{
    if (k == 0) default_seed(1, T, mu, x0);
    else if (k == 1) default_seed(2, T, mu, x0);
    else if (k == 2) seed_const(3, x0);
    else if (k == 3) seed_const(4, x0);
    else if (k == 4) seed_const(2, x0);
    else seed_const(5, x0);
}

// phase of (T_MeV, mu_MeV) for the chain's xi: 0 unknown, 1 hadron, 2 quark, 3 crossover
__device__ int phase_of(const ScanArgs& A, double xi, double T, double mu)
{
    const double* B = nullptr;
    for (int k = 0; k < A.nb; k++) {
        const double* sb = A.bnd + (int64_t)k * kBndStride;
        if (fabs(sb[0] - xi) <= 1e-6) {
            B = sb;
            break;
        }
    }
    if (!B) return 0;
    double Tc = B[1];
    int n = (int)B[2];
    if (!isnan(Tc) && T > Tc) return 3;
    if (n == 0) return 0;
    const double* Ts = B + 4;
    const double* ms = B + 4 + kMaxBndRows;
    double muc;
    if (T <= Ts[0]) muc = ms[0];
    else if (T >= Ts[n - 1]) muc = ms[n - 1];
    else {
        muc = qnan();
        for (int i = 0; i < n - 1; i++)
            if (Ts[i] <= T && T <= Ts[i + 1]) {
                double t = (T - Ts[i]) / (Ts[i + 1] - Ts[i]);
                muc = ms[i] + t * (ms[i + 1] - ms[i]);
                break;
            }
    }
    if (isnan(muc)) return 0;
    return mu < muc ? 1 : 2;
}

// ---------------------------------------------------------------------------
// group state machine
// ---------------------------------------------------------------------------
enum Phase { PH_IDLE = 0, PH_NEWTON = 1, PH_TR = 2 };
constexpr int ATT_SURR = -2;
constexpr int ATT_PRIMARY = -1;  // 0..5: MultiSeed candidates

struct St {
    int64_t item, item_end;   // current chunk of items
    int64_t pt, pt_end;       // current point, end of its chain (exclusive)
    int64_t trips;
    double T, mu, xi;
    double x[5], xold[5];
    int phase, it, attempt, flags, steps, passes;
    int prev_phase;
    int k;                    // staged pipeline: attempt index within the listed point
    int64_t item_ix;          // staged pipeline: record index of the attempt
    bool fresh;               // next pass is the first evaluation of the phase
    bool sym;                 // seed has phi_u == phi_d: keep them equal
    bool have_best;
    bool have_prev;
    bool plan_multi;          // this point runs MultiSeed as its plan
    double prev_mass;         // branch chain: M_u of the previous point
};

__host__ __device__ inline bool branch_policy(int p)
{
    return p == PNJL_SEED_BRANCH_HADRON || p == PNJL_SEED_BRANCH_QUARK;
}

__host__ __device__ inline bool chain_policy(int p)
{
    return p == PNJL_SEED_CONTINUITY || p == PNJL_SEED_PHASE_AWARE || branch_policy(p);
}

__device__ void start_attempt(St& st, const Scr& s, int attempt, const double x0[5])
{
    st.attempt = attempt;
    st.phase = PH_NEWTON;
    st.it = 0;
    st.fresh = true;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        st.x[i] = x0[i];
        st.xold[i] = qnan();
        s.at(SB_X0 + i) = x0[i];
    }
    st.sym = (x0[0] == x0[1]);
}

__device__ void write_point(const ScanArgs& A, const Scr& s, const St& st, int blk, int gl)
{
    if (gl != 0) return;
    int64_t i = st.pt;
#pragma unroll
    for (int k = 0; k < 5; k++) A.x_out[i * 5 + k] = s.at(blk + RF_X + k);
    A.omega[i] = s.at(blk + RF_OMEGA);
    A.pressure[i] = s.at(blk + RF_P);
    A.rho_norm[i] = s.at(blk + RF_RHO);
    A.entropy[i] = s.at(blk + RF_S);
    A.energy[i] = s.at(blk + RF_E);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        A.masses[i * 3 + k] = s.at(blk + RF_M + k);
        if (A.n_q) A.n_q[i * 3 + k] = s.at(blk + RF_NQ + k);
        if (A.n_qbar) A.n_qbar[i * 3 + k] = s.at(blk + RF_NQB + k);
    }
    A.resnorm[i] = s.at(blk + RF_RES);
    A.iterations[i] = (int32_t)s.at(blk + RF_IT);
    A.converged[i] = (int32_t)s.at(blk + RF_CONV);
    if (A.flags) A.flags[i] = (int32_t)s.at(blk + RF_FLAGS) | st.flags;
    if (A.steps) A.steps[i] = st.steps;
    if (A.passes) A.passes[i] = st.passes;
}

// point finished with record `blk`: write outputs, update the chain tracker
// (update! only after a converged solve, run_gap_transport_scan.jl:439-442,
// tests/perf/pnjl/scan_perf.jl:69-72).  A branch chain ends at its first
// point that fails or jumps (_is_solution_jump, DualBranchScan.jl:410-415);
// the rest of the chain is written as skipped and not solved.
__device__ void finish_point(const ScanArgs& A, const Scr& s, St& st, int blk, int gl)
{
    bool ok = s.at(blk + RF_CONV) != 0.0 && !((int)(s.at(blk + RF_FLAGS)) & PNJL_PT_EXCEPTION) &&
              !(st.flags & PNJL_PT_NO_ROOT);
    const bool branch = branch_policy(A.policy);
    if (branch && ok && st.have_prev) {
        const double dphi = fabs(s.at(SB_PREV) - s.at(blk + RF_X));
        const double dm = fabs(st.prev_mass - s.at(blk + RF_M)) * 197.327;
        if (dphi > 0.5 || dm > 50.0) ok = false;
    }
    if (branch && !ok) st.flags |= PNJL_PT_BRANCH_END;
    write_point(A, s, st, blk, gl);
    if (chain_policy(A.policy) && ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) s.at(SB_PREV + i) = s.at(blk + RF_X + i);
        st.have_prev = true;
        st.prev_mass = s.at(blk + RF_M);
        if (A.policy == PNJL_SEED_PHASE_AWARE) st.prev_phase = phase_of(A, st.xi, st.T * 197.327, st.mu * 197.327);
    }
    if (branch && !ok) {
        if (gl == 0) {
            for (int64_t i = st.pt + 1; i < st.pt_end; i++) {
#pragma unroll
                for (int k = 0; k < 5; k++) A.x_out[i * 5 + k] = qnan();
                A.omega[i] = A.pressure[i] = A.rho_norm[i] = A.entropy[i] = A.energy[i] = qnan();
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    A.masses[i * 3 + k] = qnan();
                    if (A.n_q) A.n_q[i * 3 + k] = qnan();
                    if (A.n_qbar) A.n_qbar[i * 3 + k] = qnan();
                }
                A.resnorm[i] = qnan();
                A.iterations[i] = 0;
                A.converged[i] = 0;
                if (A.flags) A.flags[i] = PNJL_PT_SKIPPED;
                if (A.steps) A.steps[i] = 0;
                if (A.passes) A.passes[i] = 0;
            }
        }
        st.pt = st.pt_end - 1;      // next_point moves on to the next chain
    }
    st.phase = PH_IDLE;
}

// Sets up point st.pt according to the policy (first attempt).
__device__ __noinline__ void begin_point(const ScanArgs& A, const Scr& s, St& st)
{
    int64_t i = st.pt;
    st.T = A.T[i];
    st.mu = A.mu[i];
    st.xi = A.xi[i];
    st.flags = 0;
    st.steps = 0;
    st.passes = 0;
    st.trips = 0;
    st.have_best = false;
    st.plan_multi = false;
    double x0[5];
    switch (A.policy) {
    case PNJL_SEED_HADRON:
        default_seed(1, st.T, st.mu, x0);
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    case PNJL_SEED_QUARK:
        default_seed(2, st.T, st.mu, x0);
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    case PNJL_SEED_EXPLICIT:
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = A.seed[i * 5 + k];
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    case PNJL_SEED_MULTI:
        st.plan_multi = true;
        multi_seed(0, st.T, st.mu, x0);
        start_attempt(st, s, 0, x0);
        break;
    case PNJL_SEED_SURROGATE:
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = A.seed[i * 5 + k];
        start_attempt(st, s, ATT_SURR, x0);
        break;
    case PNJL_SEED_CONTINUITY:
        if (st.have_prev) {
#pragma unroll
            for (int k = 0; k < 5; k++) x0[k] = s.at(SB_PREV + k);
        } else {
            default_seed(0, st.T, st.mu, x0);
        }
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    case PNJL_SEED_BRANCH_HADRON:
    case PNJL_SEED_BRANCH_QUARK:
        if (st.have_prev) {
#pragma unroll
            for (int k = 0; k < 5; k++) x0[k] = s.at(SB_PREV + k);
        } else {
            default_seed(A.policy == PNJL_SEED_BRANCH_HADRON ? 1 : 2, st.T, st.mu, x0);
        }
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    case PNJL_SEED_PHASE_AWARE:
        if (!st.have_prev) {
            st.plan_multi = true;
            multi_seed(0, st.T, st.mu, x0);
            start_attempt(st, s, 0, x0);
        } else {
            int cur = phase_of(A, st.xi, st.T * 197.327, st.mu * 197.327);
            bool reset = (st.prev_phase == 1 && cur == 2) || (st.prev_phase == 2 && cur == 1);
            if (reset) {
                seed_const(cur == 1 ? 0 : 1, x0);
                st.flags |= PNJL_PT_CHAIN_RESET;
            } else {
#pragma unroll
                for (int k = 0; k < 5; k++) x0[k] = s.at(SB_PREV + k);
            }
            start_attempt(st, s, ATT_PRIMARY, x0);
        }
        break;
    default:  // PNJL_SEED_DEFAULT
        default_seed(0, st.T, st.mu, x0);
        start_attempt(st, s, ATT_PRIMARY, x0);
        break;
    }
}

// ambiguity mask lookup (bins over the training box; outside the box: judged)
__device__ __forceinline__ bool in_judge_region(const ScanArgs& A, double T, double mu, double xi)
{
    const double v[3] = {T, mu, xi};
    int64_t idx = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double u = (v[k] - A.jlo[k]) / (A.jhi[k] - A.jlo[k]);
        if (!(u >= 0.0 && u <= 1.0)) return true;
        int b = (int)(u * A.jdim[k]);
        b = b < 0 ? 0 : (b >= A.jdim[k] ? A.jdim[k] - 1 : b);
        idx = idx * A.jdim[k] + b;
    }
    return A.judge_mask[idx] != 0;
}

// Staged pipeline: an attempt ended -> its record goes to rec_out; a failed
// learned seed or a failed primary attempt appends the point to the next stage.
__device__ void staged_done(const ScanArgs& A, const Scr& s, St& st, int blk, int gl, bool conv, bool exc)
{
    if (gl == 0) {
        double* r = A.rec_out + st.item_ix * REC_SIZE;
        for (int f = 0; f < REC_SIZE; f++) r[f] = s.at(blk + f);
        if (A.passes) atomicAdd(&A.passes[st.pt], st.passes);
        if (A.steps) atomicAdd(&A.steps[st.pt], st.steps);
        bool append = false;
        if (A.stage == STAGE_SURR) append = !conv;
        else if (A.stage == STAGE_PRIMARY) append = !conv && A.auto_ms && !exc;
        if (append) {
            unsigned long long j = atomicAdd(A.next_count, 1ULL);
            A.next_list[j] = (int32_t)st.pt;
            if (A.stage == STAGE_PRIMARY) A.mslot[st.pt] = (int32_t)j;
        }
        if ((A.stage == STAGE_SURR || A.stage == STAGE_PRIMARY) && A.jslot) {
            // learned-seed plan: a converged root inside the ambiguity mask (the
            // learned root, or the plain root of a point whose learned seed
            // failed) is checked against the MultiSeed candidates' roots
            int32_t js = -1;
            if (conv && in_judge_region(A, st.T, st.mu, st.xi)) {
                unsigned long long j = atomicAdd(A.judge_count, 1ULL);
                A.judge_list[j] = (int32_t)st.pt;
                A.jb[j] = 0;
                js = (int32_t)j;
            }
            if (A.jslot) A.jslot[st.pt] = js;
        }
        if (A.stage == STAGE_JUDGE && !conv) {
            // HADRON or hint-QUARK failed: the other MultiSeed candidates run
            const int64_t js = st.item_ix / JUDGE_SEEDS;
            if (atomicCAS(&A.jb[js], 0, 1) == 0) {
                unsigned long long j = atomicAdd(A.judge2_count, 1ULL);
                A.judge2_list[j] = (int32_t)st.pt;
            }
        }
    }
    st.phase = PH_IDLE;
}

// An attempt (one single-seed solve) ended with its result in record `blk`.
__device__ __noinline__ void attempt_done(const ScanArgs& A, const Scr& s, St& st, int blk, int gl)
{
    bool exc = s.at(blk + RF_EXC) != 0.0;
    bool conv = !exc && rec_good(s, blk, A.resmax);
    s.at(blk + RF_CONV) = conv ? 1.0 : 0.0;
    if (A.stage != STAGE_NONE) {
        staged_done(A, s, st, blk, gl, conv, exc);
        return;
    }
    double x0[5];
    if (st.attempt == ATT_SURR) {
        if (conv) {
            finish_point(A, s, st, blk, gl);
            return;
        }
        // learned seed abandoned: the plain reference policy takes over
        st.flags |= PNJL_PT_SURR_FALLBACK;
        default_seed(0, st.T, st.mu, x0);
        start_attempt(st, s, ATT_PRIMARY, x0);
        return;
    }
    if (st.attempt == ATT_PRIMARY) {
        copy_rec(s, SB_SINGLE, blk);
        if (conv || !A.auto_ms || exc) {
            finish_point(A, s, st, SB_SINGLE, gl);
            return;
        }
        st.have_best = false;
        multi_seed(0, st.T, st.mu, x0);
        start_attempt(st, s, 0, x0);
        return;
    }
    // MultiSeed candidate k = st.attempt
    if (st.attempt == 0 && st.plan_multi) copy_rec(s, SB_SINGLE, blk);
    if (conv && (!st.have_best || s.at(blk + RF_OMEGA) < s.at(SB_BEST + RF_OMEGA))) {
        copy_rec(s, SB_BEST, blk);
        st.have_best = true;
    }
    if (st.attempt < 5) {
        int k = st.attempt + 1;
        multi_seed(k, st.T, st.mu, x0);
        start_attempt(st, s, k, x0);
        return;
    }
    if (st.have_best) {
        s.at(SB_BEST + RF_FLAGS) = (double)((int)s.at(SB_BEST + RF_FLAGS) | PNJL_PT_MULTISEED);
        finish_point(A, s, st, SB_BEST, gl);
    } else {
        if (st.plan_multi) {
            st.flags |= PNJL_PT_NO_ROOT;
            s.at(SB_SINGLE + RF_CONV) = 0.0;
        }
        finish_point(A, s, st, SB_SINGLE, gl);
    }
}

// _choose_candidate (ImplicitSolver.jl:72-101) between PRIM and TRC -> block
__device__ int choose_candidate(const ScanArgs& A, const Scr& s)
{
    bool pg = rec_good(s, SB_PRIM, A.resmax);
    bool fg = rec_good(s, SB_TRC, A.resmax);
    double pres = s.at(SB_PRIM + RF_RES), fres = s.at(SB_TRC + RF_RES);
    bool pick_fb;
    if (fg && !pg) pick_fb = true;
    else if (pg && !fg) pick_fb = false;
    else if (fg && pg) {
        double po = s.at(SB_PRIM + RF_OMEGA), fo = s.at(SB_TRC + RF_OMEGA);
        if (fo < po) pick_fb = true;
        else if (fo > po) pick_fb = false;
        else pick_fb = fres < pres;
    } else {
        bool pf = s.at(SB_PRIM + RF_FCONV) != 0.0, ff = s.at(SB_TRC + RF_FCONV) != 0.0;
        if (ff && !pf) pick_fb = true;
        else if (pf && !ff) pick_fb = false;
        else if (isfinite(fres) && isfinite(pres)) pick_fb = fres < pres;
        else pick_fb = false;
    }
    if (pick_fb) {
        s.at(SB_TRC + RF_FLAGS) = (double)PNJL_PT_TR_USED;
        return SB_TRC;
    }
    return SB_PRIM;
}

// trust-region step bookkeeping (NLsolve trust_region_, autoscale, factor 1,
// eta 1e-4), after an evaluation at st.x.  Returns true if another pass is
// needed; otherwise the TR candidate is final in SB_TRC.
__device__ __noinline__ bool tr_step(const ScanArgs& A, const Scr& s, St& st, double omega, const double F[5],
                                     const double Jp[15])
{
    double r[5], d[5], p[5], delta;
    bool converged = false, stopped;
    if (st.fresh) {
        st.fresh = false;
        bool fconv = norm_inf5(F) < A.ftol;
        put_rec(A.P, s, SB_TRC, st.x, st.mu, st.T, st.sym, A.phi_tol, omega, norm_inf5(F), 0, fconv, false);
#pragma unroll
        for (int i = 0; i < 5; i++) r[i] = F[i];
#pragma unroll
        for (int k = 0; k < 15; k++) s.at(SB_TR + TR_J + k) = Jp[k];
        stopped = any_nan5(st.x) || any_nan5(r);
        converged = fconv;
        delta = 1.0;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            double cn = 0.0;
#pragma unroll
            for (int i = 0; i < 5; i++) cn += Jp[sidx(i, j)] * Jp[sidx(i, j)];
            cn = sqrt(cn);
            d[j] = (cn == 0.0) ? 1.0 : cn;
        }
        if (!converged) {
            delta = wnorm5(d, st.x);
            if (delta == 0.0) delta = 1.0;
        }
    } else {
        // evaluation at the trial point x = xold + p
        double rpred[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            r[i] = s.at(SB_TR + TR_R + i);
            d[i] = s.at(SB_TR + TR_D + i);
            p[i] = s.at(SB_TR + TR_P + i);
        }
        delta = s.at(SB_TR + TR_DELTA);
#pragma unroll
        for (int i = 0; i < 5; i++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < 5; k++) t += s.at(SB_TR + TR_J + sidx(i, k)) * p[k];
            rpred[i] = t + r[i];
        }
        double ssr = sumsq5(r);
        double rho = (ssr - sumsq5(F)) / (ssr - sumsq5(rpred));
        if (rho > 1e-4) {
#pragma unroll
            for (int i = 0; i < 5; i++) r[i] = F[i];
#pragma unroll
            for (int k = 0; k < 15; k++) s.at(SB_TR + TR_J + k) = Jp[k];
#pragma unroll
            for (int j = 0; j < 5; j++) {
                double cn = 0.0;
#pragma unroll
                for (int i = 0; i < 5; i++) cn += Jp[sidx(i, j)] * Jp[sidx(i, j)];
                cn = sqrt(cn);
                double t = 0.1 * d[j];
                d[j] = t > cn ? t : cn;
            }
            double cheb = 0.0;
            bool xnan = false;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                xnan = xnan || (st.xold[i] != st.xold[i]);
                double dd = fabs(st.x[i] - st.xold[i]);
                cheb = (dd > cheb || dd != dd) ? dd : cheb;
            }
            bool xconv = !xnan && cheb < A.xtol;
            bool fconv = norm_inf5(r) < A.ftol;
            converged = xconv || fconv;
            put_rec(A.P, s, SB_TRC, st.x, st.mu, st.T, st.sym, A.phi_tol, omega, norm_inf5(r), st.it, fconv, xconv);
        } else {
#pragma unroll
            for (int i = 0; i < 5; i++) st.x[i] = st.xold[i];
            s.at(SB_TRC + RF_XCONV) = 0.0;
            converged = false;
        }
        if (rho < 0.1) delta = delta / 2.0;
        else if (rho >= 0.9) delta = 2.0 * wnorm5(d, p);
        else if (rho >= 0.5) {
            double t = 2.0 * wnorm5(d, p);
            delta = delta > t ? delta : t;
        }
        stopped = any_nan5(st.x) || any_nan5(F);
    }
    if (!stopped && !converged && st.it < A.max_iter) {
        st.it++;
        st.steps++;
        dogleg5(s, p, r, d, delta);
#pragma unroll
        for (int i = 0; i < 5; i++) {
            s.at(SB_TR + TR_R + i) = r[i];
            s.at(SB_TR + TR_D + i) = d[i];
            s.at(SB_TR + TR_P + i) = p[i];
            st.xold[i] = st.x[i];
            st.x[i] = st.xold[i] + p[i];
        }
        s.at(SB_TR + TR_DELTA) = delta;
        if (st.sym) st.x[1] = st.x[0];
        return true;
    }
    s.at(SB_TRC + RF_IT) = (double)st.it;
    return false;
}

// One control step after a pass at st.x (sums in scratch).
__device__ __noinline__ void control_step(const ScanArgs& A, const Scr& s, St& st, int gl)
{
    double omega, F[5], Jp[15];
    assemble_core(A.P, s, st.x, st.T, st.sym, omega, F, Jp);
    if (st.phase == PH_NEWTON) {
        bool fconv, xconv;
        if (st.fresh) {
            st.fresh = false;
            if (!all_finite5(F)) {
                // NLsolve check_isfinite throws on the seed
                double x0[5];
#pragma unroll
                for (int i = 0; i < 5; i++) x0[i] = s.at(SB_X0 + i);
                put_exc_rec(s, SB_PRIM, x0);
                attempt_done(A, s, st, SB_PRIM, gl);
                return;
            }
            fconv = norm_inf5(F) < A.ftol;
            xconv = false;
        } else {
            double cheb = 0.0;
            bool xnan = false;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                xnan = xnan || (st.xold[i] != st.xold[i]);
                double dd = fabs(st.x[i] - st.xold[i]);
                cheb = (dd > cheb || dd != dd) ? dd : cheb;
            }
            xconv = !xnan && cheb < A.xtol;
            fconv = norm_inf5(F) < A.ftol;
        }
        bool stopped = any_nan5(st.x) || any_nan5(F);
        int cap = (st.attempt == ATT_SURR && A.surr_cap < A.max_iter) ? A.surr_cap : A.max_iter;
        if (!stopped && !(fconv || xconv) && st.it < cap) {
            st.it++;
            st.steps++;
            double p[5];
            newton_dir(Jp, F, p);
#pragma unroll
            for (int i = 0; i < 5; i++) {
                st.xold[i] = st.x[i];
                st.x[i] = st.xold[i] + p[i];
            }
            if (st.sym) st.x[1] = st.x[0];
            return;
        }
        // Newton finished: candidate at x
        double res = norm_inf5(F);
        put_rec(A.P, s, SB_PRIM, st.x, st.mu, st.T, st.sym, A.phi_tol, omega, res, st.it, fconv, xconv);
        bool good = rec_good(s, SB_PRIM, A.resmax);
        if (A.tr_fb && !good && st.attempt != ATT_SURR) {
            st.phase = PH_TR;
            st.it = 0;
            st.fresh = true;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                st.x[i] = s.at(SB_X0 + i);
                st.xold[i] = qnan();
            }
            st.sym = (st.x[0] == st.x[1]);
            return;
        }
        attempt_done(A, s, st, SB_PRIM, gl);
        return;
    }
    // trust region
    if (st.fresh && !all_finite5(F)) {
        // the fallback throws: the primary candidate is returned
        st.fresh = false;
        attempt_done(A, s, st, SB_PRIM, gl);
        return;
    }
    if (tr_step(A, s, st, omega, F, Jp)) return;
    attempt_done(A, s, st, choose_candidate(A, s), gl);
}

// Staged pipeline: sets up attempt st.k of point st.pt for the kernel's stage.
__device__ __noinline__ void begin_item(const ScanArgs& A, const Scr& s, St& st, int gl)
{
    int64_t i = st.pt;
    st.T = A.T[i];
    st.mu = A.mu[i];
    st.xi = A.xi[i];
    st.flags = 0;
    st.steps = 0;
    st.passes = 0;
    st.trips = 0;
    st.have_best = false;
    st.plan_multi = false;
    double x0[5];
    if (A.stage == STAGE_SURR) {
        if (A.judge_mode == PNJL_JUDGE_REFERENCE && A.judge_mask && in_judge_region(A, st.T, st.mu, st.xi)) {
            // ambiguity region: the reference policy takes the point unseen
            if (gl == 0) {
                double* r = A.rec_out + st.item_ix * REC_SIZE;
                r[RF_CONV] = 0.0;
                r[RF_FLAGS] = (double)PNJL_PT_SURR_MASKED;
                unsigned long long j = atomicAdd(A.next_count, 1ULL);
                A.next_list[j] = (int32_t)st.pt;
            }
            st.phase = PH_IDLE;
            return;
        }
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = A.seed[i * 5 + k];
        start_attempt(st, s, ATT_SURR, x0);
        return;
    }
    if (A.stage == STAGE_MULTI) {
        multi_seed(st.k, st.T, st.mu, x0);
        start_attempt(st, s, st.k, x0);
        return;
    }
    if (A.stage == STAGE_JUDGE || A.stage == STAGE_JUDGE2) {
        // the Omega check of a learned-seed root: the MultiSeed candidates
        const int c = (A.stage == STAGE_JUDGE) ? st.k : 2 + st.k;
        multi_seed(c, st.T, st.mu, x0);
        start_attempt(st, s, c, x0);
        return;
    }
    if (A.stage == STAGE_MR3) {
        // learned seed, then the chirally broken and restored reference seeds
        if (st.k == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) x0[k] = A.seed[i * 5 + k];
        } else {
            seed_const(st.k == 1 ? 0 : 1, x0);
        }
        start_attempt(st, s, st.k, x0);
        return;
    }
    // STAGE_PRIMARY
    switch (A.policy) {
    case PNJL_SEED_HADRON: default_seed(1, st.T, st.mu, x0); break;
    case PNJL_SEED_QUARK: default_seed(2, st.T, st.mu, x0); break;
    case PNJL_SEED_EXPLICIT:
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = A.seed[i * 5 + k];
        break;
    default: default_seed(0, st.T, st.mu, x0); break;
    }
    start_attempt(st, s, ATT_PRIMARY, x0);
}

// ---------------------------------------------------------------------------
// the persistent kernel
// ---------------------------------------------------------------------------
// next work point for the group; false when the queue is drained
template <int LPP>
__device__ __forceinline__ bool next_point(const ScanArgs& A, St& st, int gl, bool chains, int64_t n_items)
{
    if (A.stage != STAGE_NONE) {
        if (st.item >= st.item_end) {
            unsigned long long base = 0;
            if (gl == 0) base = atomicAdd(A.queue, (unsigned long long)A.chunk);
            base = __shfl(base, 0, LPP);
            st.item = (int64_t)base;
            st.item_end = st.item + A.chunk;
            if (st.item_end > n_items) st.item_end = n_items;
        }
        if (st.item >= st.item_end) return false;
        int64_t it = st.item++;
        int64_t j = it / A.mult;
        st.k = (int)(it - j * A.mult);
        st.pt = A.item_list ? (int64_t)A.item_list[j] : j;
        // records: per point for single-attempt stages, per item otherwise;
        // judge records: JUDGE_SEEDS per judge index
        if (A.stage == STAGE_JUDGE) st.item_ix = j * JUDGE_SEEDS + st.k;
        else if (A.stage == STAGE_JUDGE2) st.item_ix = (int64_t)A.jslot[st.pt] * JUDGE_SEEDS + 2 + st.k;
        else st.item_ix = (A.mult == 1) ? st.pt : it;
        return true;
    }
    if (chains && st.pt + 1 < st.pt_end) {
        st.pt++;
        return true;
    }
    if (st.item >= st.item_end) {
        unsigned long long base = 0;
        if (gl == 0) base = atomicAdd(A.queue, (unsigned long long)A.chunk);
        base = __shfl(base, 0, LPP);
        st.item = (int64_t)base;
        st.item_end = st.item + A.chunk;
        if (st.item_end > A.n_items) st.item_end = A.n_items;
    }
    if (st.item >= st.item_end) return false;
    int64_t it = st.item++;
    if (chains) {
        st.pt = it * A.chain_len;
        st.pt_end = st.pt + A.chain_len;
        if (st.pt_end > A.n_points) st.pt_end = A.n_points;
        st.have_prev = false;
        st.prev_phase = 0;
        st.prev_mass = 0.0;
    } else {
        st.pt = it;
        st.pt_end = it + 1;
    }
    return true;
}

// 3 waves per SIMD: every function fits 168 VGPRs; the node pass is a
// separate function so that its loop keeps all of them (no spills), the
// control path spills a little once per trip.
template <int LPP>
__global__ void __launch_bounds__(256, 3) scan_kernel(ScanArgs A)
{
    extern __shared__ double lds[];
    const int tab_len = kMathTab + A.pn + A.tn + A.pn * A.tn + A.pn;
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = A.tab[i];
    __syncthreads();

    const int gl = threadIdx.x % LPP;
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPP;
    if (g >= A.n_groups) return;  // beyond the scratch the host allocated
    const Scr s{A.scratch + g * SCRATCH_FIELDS};
    const bool chains = (A.stage == STAGE_NONE) && chain_policy(A.policy);
    const int64_t n_items = (A.stage == STAGE_NONE) ? A.n_items
                            : (A.item_list ? (int64_t)(*A.item_count) : A.n_points) * A.mult;
    St st;
    st.phase = PH_IDLE;
    st.item = 0;
    st.item_end = 0;
    st.pt = 0;
    st.pt_end = 0;
    st.have_prev = false;
    st.prev_phase = 0;
    st.prev_mass = 0.0;
    st.k = 0;
    st.item_ix = 0;

    while (true) {
        if (st.phase == PH_IDLE) {
            if (!next_point<LPP>(A, st, gl, chains, n_items)) break;
            if (A.stage == STAGE_NONE)
                begin_point(A, s, st);
            else
                begin_item(A, s, st, gl);
            if (st.phase == PH_IDLE) continue;   // handed on without an attempt
        }
        {
            double x[5];
#pragma unroll
            for (int i = 0; i < 5; i++) x[i] = st.x[i];
            node_pass<LPP>(A.P, lds, A.pn, A.tn, s, x, st.T, st.mu, st.xi, st.sym, gl);
            st.passes += st.sym ? 2 : 3;
        }
        control_step(A, s, st, gl);
        if (st.phase != PH_IDLE && ++st.trips > A.max_trips) {
            // safety exit: report the point as failed and move on
            double xn[5] = {qnan(), qnan(), qnan(), qnan(), qnan()};
            put_exc_rec(s, SB_SINGLE, xn);
            s.at(SB_SINGLE + RF_CONV) = 0.0;
            if (A.stage == STAGE_NONE)
                finish_point(A, s, st, SB_SINGLE, gl);
            else
                staged_done(A, s, st, SB_SINGLE, gl, false, true);
        }
    }
}

// ---------------------------------------------------------------------------
// one-pass residual kernel (pnjl_gap_residual_batch): one group per state
// ---------------------------------------------------------------------------
struct ResArgs {
    DevParams P;
    int64_t n;
    const double *x, *T, *mu, *xi;
    double *F, *J, *omega;
    const double* tab;
    int pn, tn;
    double* scratch;   // n * SCRATCH_FIELDS doubles
};

template <int LPP>
__global__ void __launch_bounds__(256, 3) residual_kernel(ResArgs R)
{
    extern __shared__ double lds[];
    const int tab_len = kMathTab + R.pn + R.tn + R.pn * R.tn + R.pn;
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = R.tab[i];
    __syncthreads();
    const int gl = threadIdx.x % LPP;
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPP;
    if (g >= R.n) return;
    const Scr s{R.scratch + g * SCRATCH_FIELDS};
    double x[5];
#pragma unroll
    for (int i = 0; i < 5; i++) x[i] = R.x[g * 5 + i];
    bool sym = (x[0] == x[1]);
    node_pass<LPP>(R.P, lds, R.pn, R.tn, s, x, R.T[g], R.mu[g], R.xi[g], sym, gl);
    double omega, F[5], Jp[15];
    assemble_core(R.P, s, x, R.T[g], sym, omega, F, Jp);
    if (gl == 0) {
#pragma unroll
        for (int i = 0; i < 5; i++) {
            R.F[g * 5 + i] = F[i];
#pragma unroll
            for (int j = 0; j < 5; j++) R.J[g * 25 + i * 5 + j] = Jp[sidx(i, j)];
        }
        R.omega[g] = omega;
    }
}

__global__ void zero_u64(unsigned long long* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ULL;
}

__global__ void zero_i32(int32_t* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p && i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// staged pipeline: per-point selection of the final result
// ---------------------------------------------------------------------------
enum Plan { PLAN_DEFAULT = 0, PLAN_SURR = 1, PLAN_MULTI = 2, PLAN_MR3 = 3 };

struct FinArgs {
    int64_t n;
    int plan;
    int auto_ms;
    const double* recS;   // per point (PLAN_SURR)
    const double* recP;   // per point (PLAN_DEFAULT, PLAN_SURR)
    const double* recM;   // 6 per MultiSeed point, or 6 / 3 per point (PLAN_MULTI / PLAN_MR3)
    const int32_t* mslot;
    const double* recJ;   // JUDGE_SEEDS per judged point (the MultiSeed candidates)
    const int32_t* jslot;
    const int32_t* jb;    // judge index -> candidates 2..5 were run
    double *x_out, *omega, *pressure, *rho_norm, *entropy, *energy, *masses, *n_q, *n_qbar, *resnorm;
    int32_t *iterations, *converged, *flags;
};

// sameness of two roots (DualBranchScan._is_same_solution, DualBranchScan.jl:409-419)
__device__ __forceinline__ bool same_root(const double* a, const double* b)
{
    return fabs(a[RF_X + 0] - b[RF_X + 0]) < 0.01 && fabs(a[RF_X + 3] - b[RF_X + 3]) < 0.01 &&
           fabs(a[RF_M + 0] - b[RF_M + 0]) * 197.327 < 0.01 * 197.327;
}

// argmin Omega over the converged records (first on ties), -1 if none
__device__ __forceinline__ int best_of(const double* base, int m)
{
    int best = -1;
    for (int k = 0; k < m; k++) {
        const double* r = base + (int64_t)k * REC_SIZE;
        if (r[RF_CONV] != 0.0 && (best < 0 || r[RF_OMEGA] < base[(int64_t)best * REC_SIZE + RF_OMEGA])) best = k;
    }
    return best;
}

__global__ void finalize_kernel(FinArgs F)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= F.n) return;
    const double* rec = nullptr;
    int flags = 0;
    bool conv_override = false;
    if (F.plan == PLAN_MULTI || F.plan == PLAN_MR3) {
        const int m = (F.plan == PLAN_MULTI) ? 6 : 3;
        const double* base = F.recM + p * m * REC_SIZE;
        int b = best_of(base, m);
        if (b >= 0) {
            rec = base + (int64_t)b * REC_SIZE;
            if (F.plan == PLAN_MULTI) flags |= PNJL_PT_MULTISEED;
            if (F.plan == PLAN_MR3) {
                // first-order flag: two distinct converged roots
                for (int i = 0; i < m; i++)
                    for (int j = i + 1; j < m; j++) {
                        const double* ri = base + (int64_t)i * REC_SIZE;
                        const double* rj = base + (int64_t)j * REC_SIZE;
                        if (ri[RF_CONV] != 0.0 && rj[RF_CONV] != 0.0 && !same_root(ri, rj)) flags |= PNJL_PT_FIRST_ORDER;
                    }
                // the learned seed's root lost the Omega check / did not converge
                if (base[RF_CONV] == 0.0) flags |= PNJL_PT_SURR_FALLBACK;
                else if (b != 0 && !same_root(base, rec)) flags |= PNJL_PT_SURR_OMEGA;
            }
        } else {
            rec = base;
            flags |= PNJL_PT_NO_ROOT;
            conv_override = true;
        }
    } else {
        if (F.plan == PLAN_SURR) {
            const double* rs = F.recS + p * REC_SIZE;
            if (rs[RF_CONV] != 0.0) rec = rs;
            else flags |= ((int)rs[RF_FLAGS] & PNJL_PT_SURR_MASKED) ? PNJL_PT_SURR_MASKED : PNJL_PT_SURR_FALLBACK;
        }
        if (!rec) {
            const double* single = F.recP + p * REC_SIZE;
            bool conv = single[RF_CONV] != 0.0, exc = single[RF_EXC] != 0.0;
            rec = single;
            if (!conv && F.auto_ms && !exc) {
                const double* base = F.recM + (int64_t)F.mslot[p] * 6 * REC_SIZE;
                int b = best_of(base, 6);
                if (b >= 0) {
                    rec = base + (int64_t)b * REC_SIZE;
                    flags |= PNJL_PT_MULTISEED;
                }
            }
        }
        const int32_t js = (F.plan == PLAN_SURR && F.jslot) ? F.jslot[p] : -1;
        if (js >= 0) {
            // Omega check: minimum Omega among the learned (or fallback) root and
            // the roots of the six MultiSeed candidates
            const double* r0 = rec;
            bool distinct = false;
            const int nk = F.jb[js] ? JUDGE_SEEDS : 2;
            for (int k = 0; k < nk; k++) {
                const double* rj = F.recJ + ((int64_t)js * JUDGE_SEEDS + k) * REC_SIZE;
                if (rj[RF_CONV] == 0.0) continue;
                if (!same_root(rj, r0)) distinct = true;
                if (rj[RF_OMEGA] < rec[RF_OMEGA]) rec = rj;
            }
            if (distinct) flags |= PNJL_PT_FIRST_ORDER;
            if (rec != r0 && !same_root(rec, r0)) flags |= PNJL_PT_SURR_OMEGA;
        }
    }
    for (int k = 0; k < 5; k++) F.x_out[p * 5 + k] = rec[RF_X + k];
    F.omega[p] = rec[RF_OMEGA];
    F.pressure[p] = rec[RF_P];
    F.rho_norm[p] = rec[RF_RHO];
    F.entropy[p] = rec[RF_S];
    F.energy[p] = rec[RF_E];
    for (int k = 0; k < 3; k++) {
        F.masses[p * 3 + k] = rec[RF_M + k];
        if (F.n_q) F.n_q[p * 3 + k] = rec[RF_NQ + k];
        if (F.n_qbar) F.n_qbar[p * 3 + k] = rec[RF_NQB + k];
    }
    F.resnorm[p] = rec[RF_RES];
    F.iterations[p] = (int32_t)rec[RF_IT];
    F.converged[p] = conv_override ? 0 : (int32_t)rec[RF_CONV];
    if (F.flags) F.flags[p] = (int32_t)rec[RF_FLAGS] | flags;
}

}  // namespace pnjl
