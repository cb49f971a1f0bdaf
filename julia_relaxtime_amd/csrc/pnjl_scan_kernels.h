// pnjl_scan_kernels.h -- persistent work-queue kernel that runs a whole (T, mu, xi)
// grid of PNJL FixedMu gap solves in one launch, FP64, gfx950.
//
// Geometry: a "group" of LPP lanes (LPP = 1..64, a template parameter picked
// by the autotuner) owns one work item at a time; 64/LPP groups share a
// wavefront.  Each lane integrates every LPP-th momentum node; the group's
// partial sums are combined with xor-butterfly shuffles, so every lane of a
// group holds bit-identical sums and runs the identical control logic.
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
// Per-point policies run as a few persistent launches ("stages"), each over
// a queue of items pulled with one atomic counter: stage 1 takes every point
// -- the learned-seed plan's learned attempt continues in the same group
// with the plain (DefaultSeed) attempt where the seed fails or the judge
// sends the point to the reference policy, and those judged points are
// queued first, since they run longest; a point whose plain attempt fails is
// appended to the MultiSeed list, a judged root to the Omega judge's list.
// Stage 2 runs the six MultiSeed candidates of every listed point, and the
// Omega judge's stages its candidates.  A group leaves as soon as its
// stage's queue is drained.  Chain policies (continuity, phase-aware and
// branch chains) run one chain per group with the sequential fallback logic
// of solve().
//
// Storage: the group's hot state -- the reduced node sums of the last pass,
// the trust-region state, the attempt's seed, the previous chain solution
// and x_old -- lives in LDS next to the node tables (STATE_STRIDE doubles
// per group).  The node pass and the Newton step are inlined into the
// kernel; the cold paths (an item's start and end, candidate records, the
// candidate choice, the trust-region step) are out-of-line functions that
// take the group's storage and read the launch arguments from an LDS copy
// (cold_args), so the kernel keeps 2 waves per SIMD without spilling vector
// registers.  The group's candidate records (x, Omega, thermodynamics,
// flags) sit in LDS behind its hot state: an attempt's end reads back the
// record it has just written (the convergence gate, the candidate choice,
// the copy to the item's output slot), and with the records in HBM each of
// those reads was a round trip the whole wavefront waited for.  The storage
// is handed to functions as an offset into the kernels' LDS array (Scr), so
// that out-of-line functions address it as LDS, and they cast the launch
// arguments' pointers to global ones (gmem) where they touch HBM on every
// item: generic pointers would make both flat accesses.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/pnjl_hip.h"
#include "pnjl_device.h"

// Every function below evaluates exactly the written operations (no
// compiler contraction of a*b+c into an FMA): the same arithmetic rounds the
// same way in every kernel it is inlined into, whatever the launch geometry.
#pragma clang fp contract(off)

namespace pnjl {

__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000ULL); }

// ---------------------------------------------------------------------------
// kernel arguments
// ---------------------------------------------------------------------------
struct ScanArgs {
    DevParams P;
    int64_t n_points;
    int64_t n_items;      // chain policies: chains
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
    // node table (staged in LDS): math tables | pp[pn] | cc[tn] | C[pn*tn] | Cs[pn]
    const double* tab;
    int pn, tn;
    // solver options
    double xtol, ftol, resmax, phi_tol;
    int max_iter, tr_fb, auto_ms, surr_cap;
    int judge_mode;                         // PNJL_JUDGE_* (SURROGATE inside the mask)
    // queue counters (Q_* words)
    unsigned long long* queue;
    int chunk;
    // phase-boundary tables: nb sets of {xi, T_CEP, n, pad, T[kMaxBndRows], mu[kMaxBndRows]}
    const double* bnd;
    int nb;
    int64_t max_trips;    // hard cap on trips per point (safety exit)
    // ---- per-point pipeline (plan != PLAN_CHAINS) ----
    int plan;                               // Plan
    int kind;                               // ItemKind of this stage's items
    int mult;                               // items per listed point (1, or 2 / 3 / 4 / 6 seeds)
    int k_base;                             // seed index of a point's first item
    const int32_t* order;                   // stage 1: item -> point (null: identity)
    const int32_t* item_list;               // later stages: list index -> point
    const unsigned long long* item_count;   // later stages: list length (on the device)
    double* recS;                           // learned-seed attempt, per point
    double* recP;                           // primary (DefaultSeed) attempt, per point
    double* recM;                           // MultiSeed candidates: 6 per listed point (or 6 / 3 per point)
    int32_t* mlist;                         // MultiSeed list index -> point
    int32_t* mslot;                         // point -> its MultiSeed list index
    // learned-seed judge: a converged root inside the ambiguity mask gets the
    // Omega check against the MultiSeed candidates (HADRON and hint-QUARK
    // first; where one of them fails, the other four)
    const uint8_t* judge_mask;              // [nT][nmu][nxi] bins over [jlo, jhi]
    int jdim[3];
    double jlo[3], jhi[3];
    double* recJ;                           // JUDGE_SEEDS records per judged point
    int32_t* jlist;                         // judge index -> point
    int32_t* jslot;                         // point -> its judge index (-1 none)
    int32_t* jb;                            // judge index -> 1 once it is in the second judge list
    int32_t* j2list;                        // second judge list: points whose HADRON or hint-QUARK run failed
};

// plans of the per-point pipeline
enum Plan {
    PLAN_DEFAULT = 0,   // primary attempt (DefaultSeed / HADRON / QUARK / EXPLICIT), MultiSeed on failure
    PLAN_SURR = 1,      // learned seed, primary on failure or inside the mask, MultiSeed, [Omega judge]
    PLAN_MULTI = 2,     // the six MultiSeed candidates of every point
    PLAN_MR3 = 3,       // learned, HADRON and QUARK seeds of every point
    PLAN_CHAINS = 4     // chain policies (one chain per group, sequential fallbacks)
};

// item kinds of the stages
enum ItemKind {
    K_PRIMARY = 1,      // the point's primary attempt
    K_SURR = 2,         // learned seed (then the primary attempt in the same group)
    K_MULTI = 3,        // MultiSeed candidate k of a listed point
    K_JUDGE = 4,        // Omega judge: MultiSeed candidate k of a judged point (0, 1; then 2..5)
    K_MPOL = 5,         // MultiSeed candidate k of every point (PLAN_MULTI)
    K_MR3 = 6           // k = 0 learned, 1 HADRON, 2 QUARK (PLAN_MR3)
};
// judge records per judged point: the six MultiSeed candidates
constexpr int JUDGE_SEEDS = 6;

// queue words
enum QueueWord {
    Q_HEAD = 0,         // next item of the running stage
    Q_MCOUNT = 4,       // MultiSeed list length
    Q_JCOUNT = 5,       // judge list length
    Q_J2COUNT = 6       // second judge list length
};

constexpr int kMaxBndRows = 128;
constexpr int kBndStride = 4 + 2 * kMaxBndRows;

// ---------------------------------------------------------------------------
// candidate records (doubles) and the group's state
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
// the group's record area (offsets of its records within SB_REC)
enum RecBlock {
    SB_PRIM = 0,                 // primary (Newton) candidate of the attempt
    SB_TRC = SB_PRIM + REC_SIZE, // trust-region candidate at the accepted x
    SB_SINGLE = SB_TRC + REC_SIZE,
    SB_BEST = SB_SINGLE + REC_SIZE,
    REC_AREA = SB_BEST + REC_SIZE
};
// the group's hot state in LDS
enum StateField {
    SB_SUMS = 0,                 // reduced node sums of the last pass: u, s, d (13 each)
    SB_TR = SB_SUMS + 39,        // r[5] Jsym[15] d[5] p[5] delta
    SB_X0 = SB_TR + 31,          // seed of the attempt
    SB_PREV = SB_X0 + 5,         // chain: previous solution
    SB_XOLD = SB_PREV + 5,       // x of the previous Newton / trust-region step
    SB_ST = SB_XOLD + 5,         // the group's solver state (struct St, 24 doubles)
    SB_FJ = SB_ST + 24,          // F[5] and packed J[15] of the last pass, for the trust-region step
    SB_REC = SB_FJ + 20,         // the candidate records (REC_AREA doubles)
    // an odd number of doubles per group: the groups of a wavefront hit
    // distinct LDS banks when they access the same field
    STATE_STRIDE = SB_REC + REC_AREA + ((SB_REC + REC_AREA) % 2 == 0 ? 1 : 0)
};
enum TrField { TR_R = 0, TR_J = 5, TR_D = 20, TR_P = 25, TR_DELTA = 30 };

// The launch arguments as the out-of-line (cold) functions see them: a copy
// in LDS that every block writes at its start.  The kernel's inlined code
// reads its by-value parameter from kernarg memory into scalar registers;
// handing that parameter's address to a call instead would make the
// compiler copy the whole struct to per-lane private memory and read every
// field from there on every trip.
__shared__ ScanArgs cold_args;

// The kernels' dynamic LDS: node tables, then the groups' storage.
extern __shared__ double scan_lds[];

// The group's solver state.  It lives in the group's LDS block (every lane
// of the group reads and writes the same values), which keeps it out of the
// registers the node loop needs.
struct St {
    int64_t item, item_end;   // current chunk of items (chains: of chains)
    int64_t pt, pt_end;       // current point, end of its chain (exclusive)
    int64_t trips;
    int64_t idx;              // pipeline: list index of the item's point
    double T, mu, xi;
    double x[5];
    double prev_mass;         // branch chain: M_u of the previous point
    int phase, it, attempt, flags, steps, passes;
    int prev_phase;
    int kind, k;              // pipeline: item kind and seed index
    bool fresh;               // next pass is the first evaluation of the phase
    bool sym;                 // seed has phi_u == phi_d: keep them equal
    bool have_best;
    bool have_prev;
    bool plan_multi;          // this point runs MultiSeed as its plan
    bool judged;              // stage 1: the point is in the learned seed's reference-judge mask
};
static_assert(sizeof(St) <= 24 * sizeof(double), "St must fit its LDS slot");

// a group's storage in LDS: hot state (at), candidate records (rec) and the
// solver state (st), addressed from the LDS array itself (an offset, not a
// pointer), so that out-of-line functions access it as LDS rather than
// through generic (flat) addresses
struct Scr {
    int off;
    __device__ __forceinline__ double& at(int f) const { return scan_lds[off + f]; }
    __device__ __forceinline__ double& rec(int f) const { return scan_lds[off + SB_REC + f]; }
    __device__ __forceinline__ double* ptr(int f) const { return scan_lds + off + f; }
    __device__ __forceinline__ St& st() const { return *reinterpret_cast<St*>(scan_lds + off + SB_ST); }
};

// p as a global-memory pointer.  Out-of-line functions read the launch
// arguments from LDS (cold_args), so their pointers are generic and every
// access would be a flat instruction, which also holds up the function's
// LDS accesses (flat operations count against the LDS wait counter too).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

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
__device__ __forceinline__ void flavour_sums(const double* lds, int pn, int tn, double M, double invT, double muT,
                                             double K2mu, double xi, double Phi, double Phib, int gl, FlavSums& acc)
{
    const double* pp = lds + kMathTab;
    const double* cc = pp + pn;
    const double* Cm = pp + pn + tn;
    const double* Cs = Cm + pn * tn;
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

// Node pass at state x: u, s, and d unless M_d == M_u (then d aliases u);
// the reduced sums go to the group's LDS state.
template <int LPP>
__device__ __forceinline__ void node_pass(const DevParams& P, const double* lds, int pn, int tn, const Scr& s,
                                          const double* x, double T, double mu, double xi, bool sym, int gl)
{
    double M[3];
    mass_vec(P, x, M);
    // 1/T, mu/T and exp(-2 mu/T): once per pass, not per flavour
    const double invT = 1.0 / T;
    const double muT = mu * invT;
    const double K2mu = exp(-2.0 * muT);
    const int nf = sym ? 2 : 3;
#pragma unroll 1
    for (int f = 0; f < nf; f++) {
        double Mf = (f == 0) ? M[0] : ((f == 1) ? M[2] : M[1]);
        FlavSums acc;
        flavour_sums<LPP>(lds, pn, tn, Mf, invT, muT, K2mu, xi, x[3], x[4], gl, acc);
        store_flav(s, f, acc);
    }
}

// scratch slot holding the sums of mass index i (0 u, 1 d, 2 s)
__device__ __forceinline__ int slot_of(int i, bool sym) { return i == 0 ? 0 : (i == 2 ? 1 : (sym ? 0 : 2)); }

// ---------------------------------------------------------------------------
// Omega, F and the packed symmetric Jacobian from the sums in scratch
// ---------------------------------------------------------------------------
// vacuum_I of the three masses.  With LPP >= 4 lanes per group, lane gl
// evaluates mass gl % 3 and the group takes the three results from its lanes
// 0, 1, 2 (the same arithmetic on the same inputs, so the same bits): one
// evaluation's instructions per wavefront instead of three.  Every lane of
// the group must be active.
template <int LPP>
__device__ __forceinline__ void vacuum_I3(const DevParams& P, const double M[3], int gl, double I[3], double dI[3],
                                          double d2I[3])
{
    if constexpr (LPP >= 4) {
        const int f = gl % 3;
        double a, b, c;
        vacuum_I(P, f == 0 ? M[0] : (f == 1 ? M[1] : M[2]), a, b, c);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            I[i] = __shfl(a, i, LPP);
            dI[i] = __shfl(b, i, LPP);
            d2I[i] = __shfl(c, i, LPP);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) vacuum_I(P, M[i], I[i], dI[i], d2I[i]);
    }
}

template <int LPP>
__device__ __forceinline__ void assemble_core(const DevParams& P, const Scr& s, const double* x, double T, bool sym, double& omega,
                              double F[5], double Jp[15], int gl)
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
    double vI[3], vdI[3], vd2I[3];
    vacuum_I3<LPP>(P, M, gl, vI, vdI, vd2I);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int b = SB_SUMS + 13 * slot_of(i, sym);
        const double I = vI[i], dI = vdI[i], d2I = vd2I[i];
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
// x is the group's current state st.x, read here from LDS.
__device__ __noinline__ void put_rec(Scr s, int blk, double mu, double T, bool sym,
                        double phi_tol, double omega, double res, int it, bool fconv, bool xconv)
{
    const double* x = s.st().x;
    const DevParams& P = cold_args.P;
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
        s.rec(blk + RF_NQ + i) = twoNc * sfp;
        s.rec(blk + RF_NQB + i) = twoNc * sfm;
        s.rec(blk + RF_M + i) = M[i];
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
    for (int i = 0; i < 5; i++) s.rec(blk + RF_X + i) = x[i];
    s.rec(blk + RF_OMEGA) = omega;
    s.rec(blk + RF_P) = pressure;
    s.rec(blk + RF_RHO) = rho_norm;
    s.rec(blk + RF_S) = entropy;
    s.rec(blk + RF_E) = energy;
    s.rec(blk + RF_RES) = res;
    s.rec(blk + RF_IT) = (double)it;
    s.rec(blk + RF_FCONV) = fconv ? 1.0 : 0.0;
    s.rec(blk + RF_XCONV) = xconv ? 1.0 : 0.0;
    s.rec(blk + RF_PHYS) = ph ? 1.0 : 0.0;
    s.rec(blk + RF_EXC) = 0.0;
    s.rec(blk + RF_FLAGS) = 0.0;
}

// x0: the point's seed (null: NaN)
__device__ __noinline__ void put_exc_rec(Scr s, int blk, const double* x0)
{
#pragma unroll
    for (int i = 0; i < 5; i++) s.rec(blk + RF_X + i) = x0 ? x0[i] : qnan();
    for (int f = RF_OMEGA; f <= RF_RES; f++) s.rec(blk + f) = qnan();
    s.rec(blk + RF_IT) = 0.0;
    s.rec(blk + RF_FCONV) = 0.0;
    s.rec(blk + RF_XCONV) = 0.0;
    s.rec(blk + RF_PHYS) = 0.0;
    s.rec(blk + RF_EXC) = 1.0;
    s.rec(blk + RF_FLAGS) = (double)PNJL_PT_EXCEPTION;
}

__device__ __forceinline__ void copy_rec(const Scr& s, int dst, int src)
{
    for (int f = 0; f < REC_SIZE; f++) s.rec(dst + f) = s.rec(src + f);
}

__device__ __forceinline__ bool rec_good(const Scr& s, int blk, double resmax)
{
    double r = s.rec(blk + RF_RES);
    return s.rec(blk + RF_FCONV) != 0.0 && isfinite(r) && r <= resmax && s.rec(blk + RF_PHYS) != 0.0;
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
__device__ __forceinline__ void newton_dir(const double Jp[15], const double F[5], double p[5])
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
__device__ __forceinline__ void dogleg5(const Scr& s, double p[5], const double r[5], const double d[5], double delta)
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
__device__ __forceinline__ int phase_of(const ScanArgs& A, double xi, double T, double mu)
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

static_assert(SB_ST + 24 <= SB_FJ && SB_REC + REC_AREA <= STATE_STRIDE && STATE_STRIDE % 2 == 1, "state layout");

__host__ __device__ inline bool branch_policy(int p)
{
    return p == PNJL_SEED_BRANCH_HADRON || p == PNJL_SEED_BRANCH_QUARK;
}

__host__ __device__ inline bool chain_policy(int p)
{
    return p == PNJL_SEED_CONTINUITY || p == PNJL_SEED_PHASE_AWARE || branch_policy(p);
}

__device__ __forceinline__ void start_attempt(St& st, const Scr& s, int attempt, const double x0[5])
{
    st.attempt = attempt;
    st.phase = PH_NEWTON;
    st.it = 0;
    st.fresh = true;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        st.x[i] = x0[i];
        s.at(SB_XOLD + i) = qnan();
        s.at(SB_X0 + i) = x0[i];
    }
    st.sym = (x0[0] == x0[1]);
}

// ---------------------------------------------------------------------------
// chains: one chain per group, solve()'s sequential fallbacks
// ---------------------------------------------------------------------------
__device__ __forceinline__ void write_point(const ScanArgs& A, const Scr& s, const St& st, int blk, int gl)
{
    if (gl != 0) return;
    int64_t i = st.pt;
#pragma unroll
    for (int k = 0; k < 5; k++) gmem(A.x_out)[i * 5 + k] = s.rec(blk + RF_X + k);
    gmem(A.omega)[i] = s.rec(blk + RF_OMEGA);
    gmem(A.pressure)[i] = s.rec(blk + RF_P);
    gmem(A.rho_norm)[i] = s.rec(blk + RF_RHO);
    gmem(A.entropy)[i] = s.rec(blk + RF_S);
    gmem(A.energy)[i] = s.rec(blk + RF_E);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        gmem(A.masses)[i * 3 + k] = s.rec(blk + RF_M + k);
        if (A.n_q) gmem(A.n_q)[i * 3 + k] = s.rec(blk + RF_NQ + k);
        if (A.n_qbar) gmem(A.n_qbar)[i * 3 + k] = s.rec(blk + RF_NQB + k);
    }
    gmem(A.resnorm)[i] = s.rec(blk + RF_RES);
    gmem(A.iterations)[i] = (int32_t)s.rec(blk + RF_IT);
    gmem(A.converged)[i] = (int32_t)s.rec(blk + RF_CONV);
    if (A.flags) gmem(A.flags)[i] = (int32_t)s.rec(blk + RF_FLAGS) | st.flags;
    if (A.steps) gmem(A.steps)[i] = st.steps;
    if (A.passes) gmem(A.passes)[i] = st.passes;
}

// point finished with record `blk`: write outputs, update the chain tracker
// (update! only after a converged solve, run_gap_transport_scan.jl:439-442,
// tests/perf/pnjl/scan_perf.jl:69-72).  A branch chain ends at its first
// point that fails or jumps (_is_solution_jump, DualBranchScan.jl:410-415);
// the rest of the chain is written as skipped (by all lanes of the group)
// and not solved.
template <int LPP>
__device__ __forceinline__ void finish_point(const ScanArgs& A, const Scr& s, St& st, int blk, int gl)
{
    bool ok = s.rec(blk + RF_CONV) != 0.0 && !((int)(s.rec(blk + RF_FLAGS)) & PNJL_PT_EXCEPTION) &&
              !(st.flags & PNJL_PT_NO_ROOT);
    const bool branch = branch_policy(A.policy);
    if (branch && ok && st.have_prev) {
        const double dphi = fabs(s.at(SB_PREV) - s.rec(blk + RF_X));
        const double dm = fabs(st.prev_mass - s.rec(blk + RF_M)) * 197.327;
        if (dphi > 0.5 || dm > 50.0) ok = false;
    }
    if (branch && !ok) st.flags |= PNJL_PT_BRANCH_END;
    write_point(A, s, st, blk, gl);
    if (chain_policy(A.policy) && ok) {
#pragma unroll
        for (int i = 0; i < 5; i++) s.at(SB_PREV + i) = s.rec(blk + RF_X + i);
        st.have_prev = true;
        st.prev_mass = s.rec(blk + RF_M);
        if (A.policy == PNJL_SEED_PHASE_AWARE) st.prev_phase = phase_of(A, st.xi, st.T * 197.327, st.mu * 197.327);
    }
    if (branch && !ok) {
        for (int64_t i = st.pt + 1 + gl; i < st.pt_end; i += LPP) {
#pragma unroll
            for (int k = 0; k < 5; k++) gmem(A.x_out)[i * 5 + k] = qnan();
            gmem(A.omega)[i] = gmem(A.pressure)[i] = gmem(A.rho_norm)[i] = gmem(A.entropy)[i] = gmem(A.energy)[i] = qnan();
#pragma unroll
            for (int k = 0; k < 3; k++) {
                gmem(A.masses)[i * 3 + k] = qnan();
                if (A.n_q) gmem(A.n_q)[i * 3 + k] = qnan();
                if (A.n_qbar) gmem(A.n_qbar)[i * 3 + k] = qnan();
            }
            gmem(A.resnorm)[i] = qnan();
            gmem(A.iterations)[i] = 0;
            gmem(A.converged)[i] = 0;
            if (A.flags) gmem(A.flags)[i] = PNJL_PT_SKIPPED;
            if (A.steps) gmem(A.steps)[i] = 0;
            if (A.passes) gmem(A.passes)[i] = 0;
        }
        st.pt = st.pt_end - 1;      // next_chain_point moves on to the next chain
    }
    st.phase = PH_IDLE;
}

// Sets up point st.pt of a chain according to the policy (first attempt).
__device__ __noinline__ void begin_chain_point(Scr s, St&)
{
    St& st = s.st();   // the caller's state, addressed as LDS
    const ScanArgs& A = cold_args;
    int64_t i = st.pt;
    st.T = gmem(A.T)[i];
    st.mu = gmem(A.mu)[i];
    st.xi = gmem(A.xi)[i];
    st.flags = 0;
    st.steps = 0;
    st.passes = 0;
    st.trips = 0;
    st.have_best = false;
    st.plan_multi = false;
    double x0[5];
    switch (A.policy) {
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
    default:  // PNJL_SEED_PHASE_AWARE
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
    }
}

// A chain point's attempt ended with its result in record `blk`: the
// automatic MultiSeed fallback runs its six candidates one after another.
template <int LPP>
__device__ __noinline__ void chain_attempt_done(Scr s, St&, int blk, int gl, bool conv, bool exc)
{
    St& st = s.st();   // the caller's state, addressed as LDS
    const ScanArgs& A = cold_args;
    double x0[5];
    if (st.attempt == ATT_PRIMARY) {
        copy_rec(s, SB_SINGLE, blk);
        if (conv || !A.auto_ms || exc) {
            finish_point<LPP>(A, s, st, SB_SINGLE, gl);
            return;
        }
        st.have_best = false;
        multi_seed(0, st.T, st.mu, x0);
        start_attempt(st, s, 0, x0);
        return;
    }
    // MultiSeed candidate k = st.attempt
    if (st.attempt == 0 && st.plan_multi) copy_rec(s, SB_SINGLE, blk);
    if (conv && (!st.have_best || s.rec(blk + RF_OMEGA) < s.rec(SB_BEST + RF_OMEGA))) {
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
        s.rec(SB_BEST + RF_FLAGS) = (double)((int)s.rec(SB_BEST + RF_FLAGS) | PNJL_PT_MULTISEED);
        finish_point<LPP>(A, s, st, SB_BEST, gl);
    } else {
        if (st.plan_multi) {
            st.flags |= PNJL_PT_NO_ROOT;
            s.rec(SB_SINGLE + RF_CONV) = 0.0;
        }
        finish_point<LPP>(A, s, st, SB_SINGLE, gl);
    }
}

// next chain point for the group; false when the queue is drained
template <int LPP>
__device__ __forceinline__ bool next_chain_point(const ScanArgs& A, St& st, int gl)
{
    if (st.pt + 1 < st.pt_end) {
        st.pt++;
        return true;
    }
    if (st.item >= st.item_end) {
        unsigned long long base = 0;
        if (gl == 0) base = atomicAdd(A.queue + Q_HEAD, (unsigned long long)A.chunk);
        base = __shfl(base, 0, LPP);
        st.item = (int64_t)base;
        st.item_end = st.item + A.chunk;
        if (st.item_end > A.n_items) st.item_end = A.n_items;
    }
    if (st.item >= st.item_end) return false;
    int64_t it = st.item++;
    st.pt = it * A.chain_len;
    st.pt_end = st.pt + A.chain_len;
    if (st.pt_end > A.n_points) st.pt_end = A.n_points;
    st.have_prev = false;
    st.prev_phase = 0;
    st.prev_mass = 0.0;
    return true;
}

// ---------------------------------------------------------------------------
// per-point pipeline: the two-level work queue
// ---------------------------------------------------------------------------
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

// Next item of the stage for the group (st.kind, st.k, st.idx, st.pt set);
// false when the stage's queue is drained.
template <int LPP>
__device__ __forceinline__ bool pipeline_next(const ScanArgs& A, St& st, int gl, int64_t n_items)
{
    if (st.item >= st.item_end) {
        unsigned long long base = 0;
        if (gl == 0) base = atomicAdd(A.queue + Q_HEAD, (unsigned long long)A.chunk);
        base = __shfl(base, 0, LPP);
        st.item = (int64_t)base;
        st.item_end = st.item + A.chunk;
        if (st.item_end > n_items) st.item_end = n_items;
    }
    if (st.item >= st.item_end) return false;
    const int64_t i = st.item++;
    const int64_t j = i / A.mult;
    st.k = A.k_base + (int)(i - j * A.mult);
    st.idx = j;
    st.kind = A.kind;
    st.judged = false;
    if (A.item_list) {
        st.pt = (int64_t)A.item_list[j];
    } else if (A.order) {
        // the ordering pass marks the judged points by complementing their index
        const int32_t v = A.order[j];
        st.judged = v < 0;
        st.pt = (int64_t)(v < 0 ? ~v : v);
    } else {
        st.pt = j;
    }
    return true;
}

__device__ __forceinline__ void reset_point(const ScanArgs& A, St& st)
{
    const int64_t i = st.pt;
    st.T = gmem(A.T)[i];
    st.mu = gmem(A.mu)[i];
    st.xi = gmem(A.xi)[i];
    st.flags = 0;
    st.steps = 0;
    st.passes = 0;
    st.trips = 0;
}

// the primary attempt of point st.pt (DefaultSeed with the policy's hint, or the given seed)
__device__ __forceinline__ void begin_primary(const ScanArgs& A, const Scr& s, St& st)
{
    double x0[5];
    st.kind = K_PRIMARY;
    switch (A.policy) {
    case PNJL_SEED_HADRON: default_seed(1, st.T, st.mu, x0); break;
    case PNJL_SEED_QUARK: default_seed(2, st.T, st.mu, x0); break;
    case PNJL_SEED_EXPLICIT:
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = A.seed[st.pt * 5 + k];
        break;
    default: default_seed(0, st.T, st.mu, x0); break;
    }
    start_attempt(st, s, ATT_PRIMARY, x0);
}

// Sets up the attempt of the item the group just took.
__device__ __noinline__ void begin_item(Scr s, St&, int gl)
{
    St& st = s.st();   // the caller's state, addressed as LDS
    const ScanArgs& A = cold_args;
    const int64_t p = st.pt;
    double x0[5];
    if (st.kind == K_SURR) {
        // the learned seed is read with the point's (T, mu, xi), in one round trip
#pragma unroll
        for (int k = 0; k < 5; k++) x0[k] = gmem(A.seed)[p * 5 + k];
    }
    reset_point(A, st);
    switch (st.kind) {
    case K_SURR:
        // judged: marked by the ordering pass (the same mask test on the same
        // inputs), else tested here
        if (A.order ? st.judged
                    : (A.judge_mode == PNJL_JUDGE_REFERENCE && A.judge_mask && in_judge_region(A, st.T, st.mu, st.xi))) {
            // ambiguity region: the reference policy takes the point unseen
            if (gl == 0) {
                double* r = A.recS + p * REC_SIZE;
                r[RF_CONV] = 0.0;
                r[RF_FLAGS] = (double)PNJL_PT_SURR_MASKED;
            }
            begin_primary(A, s, st);
            return;
        }
        start_attempt(st, s, ATT_SURR, x0);
        return;
    case K_MULTI:
    case K_JUDGE:
    case K_MPOL:
        multi_seed(st.k, st.T, st.mu, x0);
        start_attempt(st, s, st.k, x0);
        return;
    case K_MR3:
        // learned seed, then the chirally broken and restored reference seeds
        if (st.k == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) x0[k] = A.seed[p * 5 + k];
        } else {
            seed_const(st.k == 1 ? 0 : 1, x0);
        }
        start_attempt(st, s, st.k, x0);
        return;
    default:  // K_PRIMARY
        begin_primary(A, s, st);
        return;
    }
}

// An attempt of a pipeline item ended with its record in `blk`: the record
// goes to the item's slot; a failed primary attempt appends its point to
// the MultiSeed list, a judged root to the Omega judge's list, a failed
// judge candidate (HADRON, hint-QUARK) to the second judge list; a failed
// learned-seed attempt continues with the primary attempt in the same group.
__device__ __noinline__ void item_done(Scr s, St&, int blk, int gl, bool conv, bool exc)
{
    St& st = s.st();   // the caller's state, addressed as LDS
    const ScanArgs& A = cold_args;
    const int64_t p = st.pt;
    const bool cont = (st.kind == K_SURR) && !conv;
    if (gl == 0) {
        double* dst;
        switch (st.kind) {
        case K_SURR: dst = A.recS + p * REC_SIZE; break;
        case K_PRIMARY: dst = A.recP + p * REC_SIZE; break;
        case K_MULTI: dst = A.recM + (st.idx * 6 + st.k) * REC_SIZE; break;
        case K_JUDGE: dst = A.recJ + ((int64_t)A.jslot[p] * JUDGE_SEEDS + st.k) * REC_SIZE; break;
        case K_MPOL: dst = A.recM + (p * 6 + st.k) * REC_SIZE; break;
        default: dst = A.recM + (p * 3 + st.k) * REC_SIZE; break;   // K_MR3
        }
        for (int f = 0; f < REC_SIZE; f++) gmem(dst)[f] = s.rec(blk + f);
        // atomicAdd's relaxed, device-scope add on the global address
        if (A.passes) __hip_atomic_fetch_add(gmem(A.passes) + p, st.passes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (A.steps) __hip_atomic_fetch_add(gmem(A.steps) + p, st.steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (st.kind == K_PRIMARY && !conv && A.auto_ms && !exc) {
            const unsigned long long j = atomicAdd(A.queue + Q_MCOUNT, 1ULL);
            A.mlist[j] = (int32_t)p;
            A.mslot[p] = (int32_t)j;
        }
        if ((st.kind == K_SURR || st.kind == K_PRIMARY) && A.jslot && !cont) {
            // learned-seed plan: a converged root inside the ambiguity mask (the
            // learned root, or the plain root of a point whose learned seed
            // failed) is checked against the MultiSeed candidates' roots
            int32_t js = -1;
            if (conv && in_judge_region(A, st.T, st.mu, st.xi)) {
                const unsigned long long j = atomicAdd(A.queue + Q_JCOUNT, 1ULL);
                A.jlist[j] = (int32_t)p;
                A.jb[j] = 0;
                js = (int32_t)j;
            }
            A.jslot[p] = js;
        }
        if (st.kind == K_JUDGE && st.k < 2 && !conv) {
            // HADRON or hint-QUARK failed: the other MultiSeed candidates run
            const int32_t js = A.jslot[p];
            if (atomicCAS(&A.jb[js], 0, 1) == 0) {
                const unsigned long long j = atomicAdd(A.queue + Q_J2COUNT, 1ULL);
                A.j2list[j] = (int32_t)p;
            }
        }
    }
    if (cont) {
        st.flags = 0;
        st.steps = 0;
        st.passes = 0;
        st.trips = 0;
        begin_primary(A, s, st);
        return;
    }
    st.phase = PH_IDLE;
}

// A learned-seed run that lands farther from its seed (in the group's X0)
// than DualBranchScan's jump test allows a continuation seed (|d phi_u| >
// 0.5 or |d M_u| > 50 MeV, DualBranchScan.jl:410-415) has left the seed's
// basin: its root is not trusted and the plain policy takes the point.
__device__ __forceinline__ bool learned_jump(const ScanArgs& A, const Scr& s, int blk)
{
    double x0[5], M0[3];
#pragma unroll
    for (int i = 0; i < 5; i++) x0[i] = s.at(SB_X0 + i);
    mass_vec(A.P, x0, M0);
    return fabs(s.rec(blk + RF_X) - x0[0]) > 0.5 || fabs(s.rec(blk + RF_M) - M0[0]) * 197.327 > 50.0;
}

template <int LPP>
__device__ __forceinline__ void attempt_done(const ScanArgs& A, const Scr& s, St& st, int blk, int gl)
{
    const bool exc = s.rec(blk + RF_EXC) != 0.0;
    bool conv = !exc && rec_good(s, blk, A.resmax);
    if (conv && A.plan == PLAN_SURR && st.kind == K_SURR && learned_jump(A, s, blk)) conv = false;
    s.rec(blk + RF_CONV) = conv ? 1.0 : 0.0;
    if (A.plan == PLAN_CHAINS)
        chain_attempt_done<LPP>(s, st, blk, gl, conv, exc);
    else
        item_done(s, st, blk, gl, conv, exc);
}

// _choose_candidate (ImplicitSolver.jl:72-101) between PRIM and TRC -> block
__device__ __noinline__ int choose_candidate(Scr s)
{
    const ScanArgs& A = cold_args;
    bool pg = rec_good(s, SB_PRIM, A.resmax);
    bool fg = rec_good(s, SB_TRC, A.resmax);
    double pres = s.rec(SB_PRIM + RF_RES), fres = s.rec(SB_TRC + RF_RES);
    bool pick_fb;
    if (fg && !pg) pick_fb = true;
    else if (pg && !fg) pick_fb = false;
    else if (fg && pg) {
        double po = s.rec(SB_PRIM + RF_OMEGA), fo = s.rec(SB_TRC + RF_OMEGA);
        if (fo < po) pick_fb = true;
        else if (fo > po) pick_fb = false;
        else pick_fb = fres < pres;
    } else {
        bool pf = s.rec(SB_PRIM + RF_FCONV) != 0.0, ff = s.rec(SB_TRC + RF_FCONV) != 0.0;
        if (ff && !pf) pick_fb = true;
        else if (pf && !ff) pick_fb = false;
        else if (isfinite(fres) && isfinite(pres)) pick_fb = fres < pres;
        else pick_fb = false;
    }
    if (pick_fb) {
        s.rec(SB_TRC + RF_FLAGS) = (double)PNJL_PT_TR_USED;
        return SB_TRC;
    }
    return SB_PRIM;
}

// trust-region step bookkeeping (NLsolve trust_region_, autoscale, factor 1,
// eta 1e-4), after an evaluation at st.x.  Returns true if another pass is
// needed; otherwise the TR candidate is final in SB_TRC.
// F and the packed J of the pass are in the group's SB_FJ state.
__device__ __noinline__ bool tr_step(Scr s, St&, double omega)
{
    St& st = s.st();   // the caller's state, addressed as LDS
    const ScanArgs& A = cold_args;
    const double* F = s.ptr(SB_FJ);
    const double* Jp = s.ptr(SB_FJ + 5);
    double r[5], d[5], p[5], delta;
    bool converged = false, stopped;
    if (st.fresh) {
        st.fresh = false;
        bool fconv = norm_inf5(F) < A.ftol;
        put_rec(s, SB_TRC, st.mu, st.T, st.sym, A.phi_tol, omega, norm_inf5(F), 0, fconv, false);
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
                const double xo = s.at(SB_XOLD + i);
                xnan = xnan || (xo != xo);
                double dd = fabs(st.x[i] - xo);
                cheb = (dd > cheb || dd != dd) ? dd : cheb;
            }
            bool xconv = !xnan && cheb < A.xtol;
            bool fconv = norm_inf5(r) < A.ftol;
            converged = xconv || fconv;
            put_rec(s, SB_TRC, st.mu, st.T, st.sym, A.phi_tol, omega, norm_inf5(r), st.it, fconv, xconv);
        } else {
#pragma unroll
            for (int i = 0; i < 5; i++) st.x[i] = s.at(SB_XOLD + i);
            s.rec(SB_TRC + RF_XCONV) = 0.0;
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
            s.at(SB_XOLD + i) = st.x[i];
            st.x[i] = st.x[i] + p[i];
        }
        s.at(SB_TR + TR_DELTA) = delta;
        if (st.sym) st.x[1] = st.x[0];
        return true;
    }
    s.rec(SB_TRC + RF_IT) = (double)st.it;
    return false;
}

// One control step after a pass at st.x (sums in the group's LDS state).
template <int LPP>
__device__ __forceinline__ void control_step(const ScanArgs& A, const Scr& s, St& st, int gl)
{
    double omega, F[5], Jp[15];
    assemble_core<LPP>(A.P, s, st.x, st.T, st.sym, omega, F, Jp, gl);
    if (st.phase == PH_NEWTON) {
        bool fconv, xconv;
        if (st.fresh) {
            st.fresh = false;
            if (!all_finite5(F)) {
                // NLsolve check_isfinite throws on the seed
                put_exc_rec(s, SB_PRIM, s.ptr(SB_X0));
                attempt_done<LPP>(A, s, st, SB_PRIM, gl);
                return;
            }
            fconv = norm_inf5(F) < A.ftol;
            xconv = false;
        } else {
            double cheb = 0.0;
            bool xnan = false;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                const double xo = s.at(SB_XOLD + i);
                xnan = xnan || (xo != xo);
                double dd = fabs(st.x[i] - xo);
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
                s.at(SB_XOLD + i) = st.x[i];
                st.x[i] = st.x[i] + p[i];
            }
            if (st.sym) st.x[1] = st.x[0];
            return;
        }
        // Newton finished: candidate at x
        double res = norm_inf5(F);
        put_rec(s, SB_PRIM, st.mu, st.T, st.sym, A.phi_tol, omega, res, st.it, fconv, xconv);
        bool good = rec_good(s, SB_PRIM, A.resmax);
        if (A.tr_fb && !good && st.attempt != ATT_SURR) {
            st.phase = PH_TR;
            st.it = 0;
            st.fresh = true;
#pragma unroll
            for (int i = 0; i < 5; i++) {
                st.x[i] = s.at(SB_X0 + i);
                s.at(SB_XOLD + i) = qnan();
            }
            st.sym = (st.x[0] == st.x[1]);
            return;
        }
        attempt_done<LPP>(A, s, st, SB_PRIM, gl);
        return;
    }
    // trust region
    if (st.fresh && !all_finite5(F)) {
        // the fallback throws: the primary candidate is returned
        st.fresh = false;
        attempt_done<LPP>(A, s, st, SB_PRIM, gl);
        return;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) s.at(SB_FJ + i) = F[i];
#pragma unroll
    for (int k = 0; k < 15; k++) s.at(SB_FJ + 5 + k) = Jp[k];
    if (tr_step(s, st, omega)) return;
    attempt_done<LPP>(A, s, st, choose_candidate(s), gl);
}

// ---------------------------------------------------------------------------
// the persistent kernel
// ---------------------------------------------------------------------------
__host__ __device__ inline int table_len(int pn, int tn) { return kMathTab + pn + tn + pn * tn + pn; }

// 2 waves per SIMD, so the loop has the VGPRs (~230) to run without spills
// (at 3 waves it spills ~55 loop-invariant values and reloads them every
// trip; the learned-seed scan is ~4 % faster at 2).  The node pass and the
// Newton step are inlined; the cold bookkeeping (attempt and item ends, the
// trust-region step) is out of line and reads the block's LDS copy of the
// arguments, and no local array's address reaches a call, so the loop's
// state stays in registers and LDS.
template <int LPP>
__global__ void __launch_bounds__(256, 2) scan_kernel(ScanArgs A)
{
    double* const lds = scan_lds;
    const int tab_len = table_len(A.pn, A.tn);
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = A.tab[i];
    if (threadIdx.x == 0) cold_args = A;
    __syncthreads();

    const int gl = threadIdx.x % LPP;
    const Scr s{(int)(tab_len + (threadIdx.x / LPP) * STATE_STRIDE)};
    const bool chains = (A.plan == PLAN_CHAINS);
    St& st = s.st();
    st.phase = PH_IDLE;
    st.item = 0;
    st.item_end = 0;
    st.pt = 0;
    st.pt_end = 0;
    st.have_prev = false;
    st.prev_phase = 0;
    st.prev_mass = 0.0;
    st.k = 0;
    st.kind = 0;
    st.idx = 0;
    const int64_t n_items = chains ? A.n_items
                                   : (A.item_list ? (int64_t)(*A.item_count) : A.n_points) * (int64_t)A.mult;

    while (true) {
        if (st.phase == PH_IDLE) {
            if (chains) {
                if (!next_chain_point<LPP>(A, st, gl)) break;
                begin_chain_point(s, st);
            } else {
                if (!pipeline_next<LPP>(A, st, gl, n_items)) break;
                begin_item(s, st, gl);
            }
            if (st.phase == PH_IDLE) continue;
        }
        {
            double x[5];
#pragma unroll
            for (int i = 0; i < 5; i++) x[i] = st.x[i];
            node_pass<LPP>(A.P, lds, A.pn, A.tn, s, x, st.T, st.mu, st.xi, st.sym, gl);
            st.passes += st.sym ? 2 : 3;
        }
        control_step<LPP>(A, s, st, gl);
        if (st.phase != PH_IDLE && ++st.trips > A.max_trips) {
            // safety exit: report the point as failed and move on
            put_exc_rec(s, SB_SINGLE, nullptr);
            s.rec(SB_SINGLE + RF_CONV) = 0.0;
            if (chains)
                finish_point<LPP>(A, s, st, SB_SINGLE, gl);
            else
                item_done(s, st, SB_SINGLE, gl, false, true);
        }
    }
}

// Stage-1 item order: four classes, each contiguous, in this order -- points
// the learned seed's judge sends to the plain policy (they run longest), then
// the rest; within each, points at xi != 0 (four angle nodes per momentum
// node) before points at xi = 0 (one).  The groups of a wavefront then mostly
// hold items with the same node count, and the short items run last.
__device__ __forceinline__ int order_class(const ScanArgs& A, int64_t p)
{
    const bool judged = A.plan == PLAN_SURR && A.judge_mask && A.judge_mode == PNJL_JUDGE_REFERENCE &&
                        in_judge_region(A, A.T[p], A.mu[p], A.xi[p]);
    return (judged ? 0 : 2) + (A.xi[p] == 0.0 ? 1 : 0);
}

// counters[0..3]: class sizes (COUNT pass); counters[4..7]: fill cursors of
// the classes (PLACE pass).  One atomic per class and block of 1024 points
// (wave counts by ballot, block totals in LDS): same-address atomics
// serialise, so fewer and wider ones.
constexpr int kOrderBlock = 1024;
template <bool PLACE>
__global__ void __launch_bounds__(kOrderBlock) order_kernel(ScanArgs A, int32_t* order, unsigned long long* counters)
{
    __shared__ unsigned int wcount[kOrderBlock / 64][4];
    __shared__ unsigned long long bbase[4];
    const int64_t p = (int64_t)blockIdx.x * kOrderBlock + threadIdx.x;
    const int c = p < A.n_points ? order_class(A, p) : -1;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    unsigned int rank = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned long long m = __ballot(c == k);
        if (lane == 0) wcount[wave][k] = (unsigned int)__popcll(m);
        if (c == k) rank = (unsigned int)__popcll(m & ((1ULL << lane) - 1ULL));
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = (int)threadIdx.x;
        unsigned long long total = 0;
        for (int w = 0; w < kOrderBlock / 64; w++) total += wcount[w][k];
        bbase[k] = total ? atomicAdd(counters + (PLACE ? 4 : 0) + k, total) : 0ULL;
    }
    __syncthreads();
    if (PLACE && c >= 0) {
        unsigned long long pos = bbase[c] + rank;
        for (int w = 0; w < wave; w++) pos += wcount[w][c];
        for (int k = 0; k < c; k++) pos += counters[k];
        // classes 0 and 1 are the judged points: their index complemented
        order[pos] = c < 2 ? ~(int32_t)p : (int32_t)p;
    }
}

__global__ void zero_u64(unsigned long long* q, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) q[i] = 0ULL;
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
};

template <int LPP>
__global__ void __launch_bounds__(256, 3) residual_kernel(ResArgs R)
{
    double* const lds = scan_lds;
    const int tab_len = table_len(R.pn, R.tn);
    for (int i = threadIdx.x; i < tab_len; i += blockDim.x) lds[i] = R.tab[i];
    __syncthreads();
    const int gl = threadIdx.x % LPP;
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPP;
    if (g >= R.n) return;
    const Scr s{(int)(tab_len + (threadIdx.x / LPP) * STATE_STRIDE)};
    double x[5];
#pragma unroll
    for (int i = 0; i < 5; i++) x[i] = R.x[g * 5 + i];
    bool sym = (x[0] == x[1]);
    node_pass<LPP>(R.P, lds, R.pn, R.tn, s, x, R.T[g], R.mu[g], R.xi[g], sym, gl);
    double omega, F[5], Jp[15];
    assemble_core<LPP>(R.P, s, x, R.T[g], sym, omega, F, Jp, gl);
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

__global__ void zero_i32(int32_t* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p && i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// per-point selection of the final result
// ---------------------------------------------------------------------------
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
