/*
 * pnjl_hip.h -- C ABI of libpnjl_hip.so, the MI355X (gfx950) backend of the PNJL
 * equilibrium-scan hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int
 * status (PNJL_OK = 0) and never throws or longjmps across the boundary.  The last
 * error message of a context is available from pnjl_last_error().
 *
 * Pointer convention: unless a call says otherwise, array arguments are HOST
 * pointers owned by the caller; the library stages them to HBM, runs, and copies
 * back before returning.  With PNJL_FLAG_DEVICE_PTRS every array argument of
 * that call (inputs and the output struct's arrays) is instead a device pointer
 * into HBM of the context's device (e.g. a torch.cuda tensor's data_ptr()); the
 * call then launches on `stream` (NULL = the null stream) and synchronises it
 * before returning unless PNJL_FLAG_ASYNC is also given.
 *
 * Units: fm^-1 for T, mu, masses and momenta (MeV / 197.327), as in the
 * reference (src/Constants_PNJL.jl:80-97).
 */
#ifndef PNJL_HIP_H
#define PNJL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* status codes                                                              */
/* ------------------------------------------------------------------------ */
#define PNJL_OK               0
#define PNJL_ERR_ARG          1   /* bad argument (size, policy, NULL pointer) */
#define PNJL_ERR_HIP          2   /* HIP runtime error (message in pnjl_last_error) */
#define PNJL_ERR_NO_DEVICE    3   /* no gfx950 device visible */
#define PNJL_ERR_CHANNEL      4   /* polarization channel other than P/S */
#define PNJL_ERR_STATE        5   /* call order (e.g. surrogate used before trained) */

/* flags */
#define PNJL_FLAG_DEVICE_PTRS 0x1
#define PNJL_FLAG_ASYNC       0x2
/* pnjl_polarization_batch, pnjl_propagator_by_channel_batch and
 * pnjl_meson_mass_batch only: one query (one Pi pair of a query) per
 * wavefront (63 lanes idle), so that
 * hardware FP64 counters see each query's own operations without the
 * inactive lanes of divergent waves -- for counting algorithmic FLOPs */
#define PNJL_FLAG_ONE_PER_WAVE 0x4

/* ------------------------------------------------------------------------ */
/* model parameters                                                          */
/* replaces: src/Constants_PNJL.jl:17-103 + config/pnjl/default.toml:1-21    */
/* ------------------------------------------------------------------------ */
typedef struct pnjl_model_params {
    double hbarc;          /* MeV fm                          (197.327)        */
    double Lambda;         /* cutoff, fm^-1                   (602.3/hbarc)     */
    double G;              /* fm^2                            (1.835/Lambda^2)  */
    double K;              /* fm^5                            (12.36/Lambda^5)  */
    double m_ud0;          /* fm^-1                           (5.5/hbarc)       */
    double m_s0;           /* fm^-1                           (140.7/hbarc)     */
    double T0;             /* fm^-1                           (210/hbarc)       */
    double a0, a1, a2, b3; /* Polyakov potential             (3.51,-2.47,15.2,-1.75) */
    double rho0;           /* fm^-3                           (0.16)            */
    int    N_color;        /* 3                                                  */
} pnjl_model_params;

/* Fills the default profile (config/pnjl/default.toml). */
void pnjl_default_params(pnjl_model_params* out);

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct pnjl_ctx pnjl_ctx;

/* Creates a context on HIP device `device`.  params == NULL -> defaults.   */
int  pnjl_ctx_create(const pnjl_model_params* params, int device, pnjl_ctx** out);
void pnjl_ctx_destroy(pnjl_ctx* ctx);
const char* pnjl_last_error(const pnjl_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* quadrature nodes                                                          */
/* replaces: cached_nodes(p_num, t_num)  src/pnjl/core/Integrals.jl:87-109   */
/* p nodes = GL on [0,10] fm^-1, cos(theta) nodes = GL on [0,1], weights x2. */
/* Writes the column-major p_num x t_num meshes (index j + l*p_num) when the */
/* output pointers are non-NULL (host pointers).                            */
/* ------------------------------------------------------------------------ */
int pnjl_nodes(pnjl_ctx* ctx, int p_num, int t_num,
               double* p_mesh, double* cos_mesh, double* coef_mesh);

/* Gauss-Legendre nodes/weights on [a,b], ascending.
 * replaces: gauleg  src/integration/GaussLegendre.jl:94-119                 */
int pnjl_gauleg(double a, double b, int n, double* nodes, double* weights);

/* ------------------------------------------------------------------------ */
/* gap solver                                                                */
/* replaces: PNJL.solve(::FixedMu, T_fm, mu_fm; xi, seed_strategy, p_num,     */
/*           t_num, iterations, trust_region_fallback,                      */
/*           auto_multiseed_fallback, residual_norm_max)                     */
/*           src/pnjl/solver/ImplicitSolver.jl:211-328, solve_multi :532-559 */
/* ------------------------------------------------------------------------ */

/* seed policies (src/pnjl/solver/SeedStrategies.jl) */
#define PNJL_SEED_DEFAULT        0  /* DefaultSeed(:auto)            :200-225 */
#define PNJL_SEED_HADRON         1  /* DefaultSeed(phase_hint=:hadron)         */
#define PNJL_SEED_QUARK          2  /* DefaultSeed(phase_hint=:quark)          */
#define PNJL_SEED_MULTI          3  /* MultiSeed() -> solve_multi   :257-284 */
#define PNJL_SEED_EXPLICIT       4  /* per-point seed (ContinuitySeed with a
                                       previous solution)            :313-325 */
#define PNJL_SEED_CONTINUITY     5  /* ContinuitySeed chains        :303-345 */
#define PNJL_SEED_PHASE_AWARE    6  /* PhaseAwareContinuitySeed chains with the
                                       driver's MultiSeed first point
                                       (scripts/relaxtime/run_gap_transport_scan.jl:410-442) */
#define PNJL_SEED_SURROGATE      7  /* learned seed; plain seeds as fallback  */
#define PNJL_SEED_MULTIROOT3     8  /* surrogate + HADRON + QUARK, min Omega  */
/* DualBranchScan branches as chains (DualBranchScan.jl:104-169, :330-445):
 * the first point of each chain is solved from DefaultSeed(:hadron) (resp.
 * :quark), every next one from the previous solution; a chain ENDS at its
 * first point that does not converge, raises (PNJL_PT_EXCEPTION /
 * PNJL_PT_NO_ROOT) or jumps (|phi_u - previous phi_u| > 0.5 or |M_u -
 * previous M_u| > 50 MeV): that point is reported with PNJL_PT_BRANCH_END
 * and the points after it are not solved (PNJL_PT_SKIPPED, converged 0,
 * NaN values).  Callers lay out each chain in the branch's mu order
 * (ascending for the hadron branch, descending for the quark branch) and set
 * pnjl_solve_opts.chain_len.                                                */
#define PNJL_SEED_BRANCH_HADRON  9
#define PNJL_SEED_BRANCH_QUARK  10

typedef struct pnjl_solve_opts {
    double xtol;                  /* NLsolve xtol (reference 1e-9)           */
    double ftol;                  /* NLsolve ftol (reference 1e-9)           */
    double residual_norm_max;     /* "converged" gate (reference 1e-6)       */
    double phi_tol;               /* physicality tolerance (1e-8)            */
    int    max_iter;              /* NLsolve `iterations` (1000; driver 40)  */
    int    trust_region_fallback; /* 1                                       */
    int    auto_multiseed_fallback;/* 1                                      */
    int    surrogate_max_steps;   /* Newton steps before a learned seed is
                                     abandoned (8)                           */
    int    chain_len;             /* CONTINUITY / PHASE_AWARE: points per
                                     chain (consecutive in the batch)       */
    int    surrogate_judge;       /* SURROGATE inside the judge mask:
                                     PNJL_JUDGE_REFERENCE (0) or
                                     PNJL_JUDGE_OMEGA (1)                    */
} pnjl_solve_opts;

/* judge modes of PNJL_SEED_SURROGATE inside the ambiguity mask
 * (pnjl_set_judge_mask):
 *   REFERENCE  the learned seed is not used there: the reference DefaultSeed
 *              policy (hint seed, automatic MultiSeed fallback) solves the
 *              point, so the scan returns the reference's own root branch;
 *   OMEGA      the learned (or fallback) root is compared with the roots of
 *              the MultiSeed candidates -- HADRON and hint-QUARK, plus the
 *              other four where either of those fails -- and the
 *              minimum-Omega root is kept (the thermodynamically stable
 *              branch).                                                    */
#define PNJL_JUDGE_REFERENCE 0
#define PNJL_JUDGE_OMEGA     1

void pnjl_default_solve_opts(pnjl_solve_opts* out);

/* Structure-of-arrays result; each array holds n entries (x, M, n_q, n_qbar:
 * n*5 / n*3, point-major).  Maps 1:1 onto SolverResult
 * (src/pnjl/solver/ImplicitSolver.jl:178-193) plus the number densities of
 * calculate_number_densities (src/pnjl/core/Thermodynamics.jl:255-281).     */
typedef struct pnjl_solve_out {
    double*  x;          /* n*5: phi_u, phi_d, phi_s, Phi, Phibar            */
    double*  omega;      /* n                                                */
    double*  pressure;   /* n                                                */
    double*  rho_norm;   /* n                                                */
    double*  entropy;    /* n                                                */
    double*  energy;     /* n                                                */
    double*  masses;     /* n*3                                              */
    double*  n_q;        /* n*3  quark number densities  (may be NULL)       */
    double*  n_qbar;     /* n*3  antiquark densities     (may be NULL)       */
    double*  residual_norm; /* n                                             */
    int32_t* iterations; /* n                                                */
    int32_t* converged;  /* n   reference `converged` (res <= residual_norm_max,
                                f-converged, physical)                       */
    int32_t* flags;      /* n   PNJL_FLAG_* bits below (may be NULL)         */
    int32_t* newton_steps; /* n total Newton/TR steps spent on the point,
                              all attempts (may be NULL)                     */
    int32_t* passes;     /* n   flavour node passes spent on the point (each
                                covers p_num x t_eff nodes; t_eff = 1 at
                                xi = 0, else t_num) (may be NULL)            */
} pnjl_solve_out;

/* per-point flag bits */
#define PNJL_PT_TR_USED        0x01  /* trust-region candidate chosen          */
#define PNJL_PT_MULTISEED      0x02  /* automatic MultiSeed result returned    */
#define PNJL_PT_EXCEPTION      0x04  /* non-finite residual at the seed (the
                                        reference throws IsFiniteException)   */
#define PNJL_PT_NO_ROOT        0x08  /* MultiSeed: no seed converged (reference
                                        throws ErrorException)                */
#define PNJL_PT_SURR_FALLBACK  0x10  /* learned seed abandoned -> plain seeds  */
#define PNJL_PT_SURR_OMEGA     0x20  /* learned root lost the Omega check      */
#define PNJL_PT_FIRST_ORDER    0x40  /* >= 2 distinct physical roots (MULTIROOT3,
                                        SURROGATE with PNJL_JUDGE_OMEGA)      */
#define PNJL_PT_CHAIN_RESET    0x80  /* chain: phase switch re-seeded the point */
#define PNJL_PT_SURR_MASKED    0x100 /* in the judge mask: the reference policy
                                        solved the point (PNJL_JUDGE_REFERENCE) */
#define PNJL_PT_BRANCH_END     0x200 /* branch chain: this point ended it      */
#define PNJL_PT_SKIPPED        0x400 /* branch chain: after the end, not solved */

/* Solves n FixedMu points in one persistent-kernel launch.
 * T, mu, xi: n values each.  seed: n*5 (EXPLICIT) or NULL.
 * p_num/t_num: quadrature sizes (reference defaults 64 and 8).             */
int pnjl_solve_fixedmu_batch(pnjl_ctx* ctx, int64_t n,
                             const double* T, const double* mu, const double* xi,
                             const double* seed, int seed_policy,
                             int p_num, int t_num,
                             const pnjl_solve_opts* opts,
                             const pnjl_solve_out* out,
                             int flags, void* stream);

/* pnjl_solve_fixedmu_batch without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, policies and results as pnjl_solve_fixedmu_batch, but
 * every array is a device pointer (flags must hold PNJL_FLAG_DEVICE_PTRS)
 * and n <= 8192 points per call (larger n: PNJL_ERR_ARG).  The node table,
 * work-queue counters, candidate records and per-point lists live in one
 * fixed scratch array of ~26 MB per GPU, held from the library's load until
 * the process ends; its address is looked up once per device by
 * pnjl_ctx_create.  Every call on a GPU shares that GPU's array, so calls
 * that overlap in time on one GPU race (serialise them on one stream).  A
 * call whose p_num/t_num differ from the previous call's on that GPU
 * rebuilds the node table in the array and first waits for the device.
 * The phase-boundary tables and judge mask are the context's (set up
 * beforehand with pnjl_set_phase_boundary / pnjl_set_judge_mask).          */
int pnjl_solve_fixedmu_batch_noalloc(pnjl_ctx* ctx, int64_t n,
                                     const double* T, const double* mu, const double* xi,
                                     const double* seed, int seed_policy,
                                     int p_num, int t_num,
                                     const pnjl_solve_opts* opts,
                                     const pnjl_solve_out* out,
                                     int flags, void* stream);

/* Implicit thermodynamic derivatives at given roots x[n*5] of the FixedMu gap
 * equations (solve first; xi and the node counts as in the solve): out[n*36],
 * per point, in this order --
 *   0 pressure, 1 energy, 2 rho (sum_i rho_i / 3), 3 rho_norm, 4 entropy,
 *   5 dP_dT, 6 dP_dmu, 7 dEpsilon_dT, 8 dEpsilon_dmu, 9 dn_dT, 10 dn_dmu,
 *   11 dP_depsilon_n, 12 dP_dn_epsilon, 13-15 masses (u, d, s; bare masses
 *   5.5 and 140.0 MeV as the reference's derivative module writes them),
 *   16-18 dM_dT, 19-21 dM_dmu (quark mu), 22 v_n_sq, 23 dmuB_dT_sigma,
 *   24-28 dx_dT, 29-33 dx_dmu, 34 ds_dT, 35 ds_dmu.
 * dx/dtheta = -(d2 Omega/dx2)^-1 d2 Omega/dx dtheta (implicit function
 * theorem); all-NaN row where that Hessian is singular.  Host pointers unless
 * PNJL_FLAG_DEVICE_PTRS.
 * replaces: mass_derivatives (order 1), thermo_derivatives, dP_dT / dP_dmu
 *           (order 1, 2) and bulk_viscosity_coefficients
 *           src/pnjl/derivatives/ThermoDerivatives.jl:132-466              */
#define PNJL_THERMO_DERIV_FIELDS 36
int pnjl_thermo_derivatives_batch(pnjl_ctx* ctx, int64_t n,
                                  const double* T, const double* mu, const double* xi,
                                  const double* x, int p_num, int t_num,
                                  double* out, int flags, void* stream);

/* Second-order implicit derivatives at given roots x[n*5] (solve first; xi
 * and the node counts as in the solve): out[n*48], per point --
 *   0-4 x, 5-9 dx_dT, 10-14 dx_dmu, 15-19 d2x_dT2, 20-24 d2x_dTdmu,
 *   25-29 d2x_dmu2, 30-32 masses (u, d, s; bare masses 5.5 and 140.0 MeV as
 *   the reference's derivative module writes them), 33-35 dM_dT,
 *   36-38 dM_dmu, 39-41 d2M_dT2, 42-44 d2M_dTdmu, 45-47 d2M_dmu2.
 * Along the root x(T, mu), differentiating H x_a + F_a = 0 (H = d2 Omega/dx2,
 * F_a = d2 Omega/dx dtheta_a) once more along w_b = (x_b, e_b) gives
 * x_ab = -H^-1 [(D_wb H) x_a + D_wb F_a]; the directional derivatives are the
 * closed forms of H and F_a evaluated in dual-number arithmetic, exact to
 * rounding like the reference's nested forward-mode AD.  All-NaN row where
 * H is singular.  Host pointers unless PNJL_FLAG_DEVICE_PTRS.
 * replaces: solve_with_derivatives (order 1, 2)
 *             src/pnjl/solver/ImplicitSolver.jl:660-701
 *           mass_derivatives (order 2)
 *             src/pnjl/derivatives/ThermoDerivatives.jl:153-172            */
#define PNJL_IMPLICIT2_FIELDS 48
int pnjl_implicit_derivatives2_batch(pnjl_ctx* ctx, int64_t n,
                                     const double* T, const double* mu, const double* xi,
                                     const double* x, int p_num, int t_num,
                                     double* out, int flags, void* stream);

/* Residual F = grad_x P (5), Jacobian J = dF/dx (5x5, row-major) and Omega at
 * given states -- the building block of the Newton loop.
 * replaces: gap_conditions  src/pnjl/solver/Conditions.jl:72-81 and the
 *           ForwardDiff Jacobian NLsolve builds from it.                    */
int pnjl_gap_residual_batch(pnjl_ctx* ctx, int64_t n,
                            const double* x, const double* T, const double* mu,
                            const double* xi, int p_num, int t_num,
                            double* F, double* J, double* omega,
                            int flags, void* stream);

/* Phase-boundary tables for PNJL_SEED_PHASE_AWARE chains: rows
 * (xi, T_MeV, mu_MeV) of boundary.csv and (xi, T_CEP_MeV, mu_CEP_MeV) of
 * cep.csv (host pointers).  For a chain of a given xi the rows with
 * |xi_row - xi| <= 1e-6 are used, sorted by T, with the first matching CEP.
 * replaces: load_phase_boundary / interpolate_mu_c / _get_current_phase
 *           src/pnjl/solver/SeedStrategies.jl:388-475, 762-782           */
int pnjl_set_phase_boundary(pnjl_ctx* ctx, int nrows, const double* xi, const double* T_MeV,
                            const double* mu_MeV, int ncep, const double* cep_xi,
                            const double* cep_T_MeV, const double* cep_mu_MeV);

/* Ambiguity mask of the learned seed's judge (no counterpart in the
 * reference).  A point of a PNJL_SEED_SURROGATE batch whose (T, mu, xi)
 * falls in a set bin of this nT x nmu x nxi grid over [lo, hi] (row-major,
 * T slowest), or outside the box, is judged as opts->surrogate_judge says
 * (PNJL_JUDGE_*).  mask == NULL clears it (no judge).                       */
int pnjl_set_judge_mask(pnjl_ctx* ctx, int nT, int nmu, int nxi, const double* lo3, const double* hi3,
                        const uint8_t* mask);

/* Launch geometry of the persistent scan kernel; chosen per config by the
 * autotuner (pnjl_autotune) or set explicitly.                             */
typedef struct pnjl_launch_cfg {
    int lanes_per_point;  /* 1,2,4,8,16,32,64 -> 64/lanes points per wavefront */
    int chunk;            /* points taken from the work queue per pop        */
    int waves_per_cu;     /* persistent grid size = 256 CUs x waves_per_cu   */
    int tail_lanes_per_point; /* lanes per point of the later pipeline stages
                             (MultiSeed candidates, Omega-judge candidates:
                             few, long-running items; they run on the full
                             persistent grid); 0 = same as lanes_per_point */
} pnjl_launch_cfg;

int pnjl_set_launch_cfg(pnjl_ctx* ctx, const pnjl_launch_cfg* cfg);
int pnjl_get_launch_cfg(const pnjl_ctx* ctx, pnjl_launch_cfg* cfg);

/* Autotuner: times launch geometries on the first `sample` points of the
 * batch (host pointers as in pnjl_solve_fixedmu_batch; each candidate warmed
 * once, best of two timed runs) -- first every lanes_per_point (points per
 * wavefront and nodes per lane) at chunk 1 and 12 waves/CU, then chunk
 * {1,2,4} x waves/CU {8,12,16} for the two best lane counts, then the later
 * stages' lanes {16,32,64} -- keeps the fastest in the context and returns it
 * in *best (and its pipeline time per point in *ns_per_point).  No
 * counterpart in the reference.                                            */
int pnjl_autotune(pnjl_ctx* ctx, int64_t sample, const double* T, const double* mu, const double* xi,
                  const double* seed, int seed_policy, int p_num, int t_num,
                  const pnjl_solve_opts* opts, pnjl_launch_cfg* best, double* ns_per_point);

/* Kernel time of the last solve/residual launch in milliseconds (HIP events
 * recorded around the kernel on the call's stream).                        */
double pnjl_last_kernel_ms(const pnjl_ctx* ctx);

/* Durations (ms, HIP events on the call's stream) of each persistent-kernel
 * launch of the last solve -- one per pipeline stage that ran -- into
 * ms[0..max-1]; returns the number of launches (waits for the last one).
 * No counterpart in the reference: measurement of the GPU pipeline.        */
int pnjl_last_scan_ms(pnjl_ctx* ctx, double* ms, int max);

/* ------------------------------------------------------------------------ */
/* learned seed (no counterpart in the reference: it feeds                  */
/* PNJL_SEED_SURROGATE / PNJL_SEED_MULTIROOT3, whose fallback and judge are  */
/* the reference seeds of SeedStrategies.jl)                                 */
/* MLP 3 -> 128 -> 128 -> 128 -> 4 (SiLU) on normalised (T, mu, xi) giving   */
/* normalised (phi_u, phi_s, Phi, Phibar); training: BF16 weights and      */
/* activations, FP32 MFMA accumulation, FP32 master weights trained by Adam */
/* on the GPU; inference: the FP32 masters at FP32-class precision (split   */
/* BF16 products on the matrix cores).                                      */
/* Parameters are one packed FP32 vector, row-major:                        */
/*   W1[128][3] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]           */
/*   W4[4][128] b4[4]   (34052 values)                                      */
/* ------------------------------------------------------------------------ */
typedef struct pnjl_surrogate pnjl_surrogate;

int  pnjl_surrogate_param_count(void);
int  pnjl_surrogate_create(int device, uint64_t init_seed, pnjl_surrogate** out);
void pnjl_surrogate_destroy(pnjl_surrogate* s);
const char* pnjl_surrogate_last_error(const pnjl_surrogate* s);
double pnjl_surrogate_last_ms(const pnjl_surrogate* s);
/* inputs map [in_lo, in_hi] -> [-1, 1]; outputs y = mean + std * net      */
int  pnjl_surrogate_set_scaling(pnjl_surrogate* s, const double* in_lo3, const double* in_hi3,
                                const double* out_mean4, const double* out_std4);
int  pnjl_surrogate_get_params(pnjl_surrogate* s, float* params);
int  pnjl_surrogate_set_params(pnjl_surrogate* s, const float* params);
/* Adam (beta 0.9/0.999, eps 1e-8) on the MSE of the normalised outputs;
 * each step draws `batch` rows of (T, mu, xi, roots[n*5]) at random.
 * loss_out: `steps` floats or NULL.                                        */
int  pnjl_surrogate_train(pnjl_surrogate* s, int64_t n, const double* T, const double* mu,
                          const double* xi, const double* roots, int steps, int batch, float lr,
                          uint64_t seed, float* loss_out, int flags, void* stream);
/* loss and gradient of the MSE over rows 0..n-1 (no update; host pointers) */
int  pnjl_surrogate_loss_grad(pnjl_surrogate* s, int64_t n, const double* T, const double* mu,
                              const double* xi, const double* roots, float* loss, float* grad);
/* seeds[n*5] = (phi_u, phi_u, phi_s, Phi, Phibar) predicted at (T, mu, xi)  */
int  pnjl_surrogate_infer(pnjl_surrogate* s, int64_t n, const double* T, const double* mu,
                          const double* xi, double* seeds, int flags, void* stream);

/* ------------------------------------------------------------------------ */
/* anisotropic one-loop polarization (src/relaxtime)                        */
/* All momenta, energies, masses, chemical potentials and T in fm^-1.        */
/* Results reproduce the reference's fixed quadrature (same segmentation,   */
/* nodes and skip-non-finite rule), not the exact integrals.                 */
/* ------------------------------------------------------------------------ */
#define PNJL_CHANNEL_P 1          /* pseudoscalar (PolarizationCache.jl:32-35 codes) */
#define PNJL_CHANNEL_S 2          /* scalar                                          */

/* One polarization query: the arguments of
 *   polarization_aniso(channel, k0, k_norm, m1, m2, mu1, mu2, T, Phi, Phibar,
 *                      xi, A1_value, A2_value, num_s_quark)
 *     src/relaxtime/PolarizationAniso.jl:22-24
 * and, with width = 1, of polarization_with_width (:88-91) / the complex
 * p0 = k0 + i gamma/2 of polarization_complex (:127-134).                  */
typedef struct pnjl_pol_query {
    int    channel;               /* PNJL_CHANNEL_P or PNJL_CHANNEL_S        */
    int    num_s;                 /* num_s_quark (1 -> k0-symmetrised B0)     */
    int    width;                 /* 0: polarization_aniso, 1: _with_width   */
    int    reserved;
    double k0, gamma, k, m1, m2, mu1, mu2, T, Phi, Phib, xi, A1, A2;
} pnjl_pol_query;

/* Pi for n queries -> re[n], im[n].  Host pointers unless flags has
 * PNJL_FLAG_DEVICE_PTRS.  A bad channel fails the call with PNJL_ERR_ARG
 * (the reference throws ArgumentError, PolarizationAniso.jl:62).
 * replaces: polarization_aniso / polarization_with_width / polarization_complex
 *           src/relaxtime/PolarizationAniso.jl:22-134                      */
int pnjl_polarization_batch(pnjl_ctx* ctx, int64_t n, const pnjl_pol_query* q, double* re, double* im,
                            int flags, void* stream);

/* B0(lambda, k, m1, mu1, m2, mu2, T; Phi, Phibar) for n rows of
 * args[n*9] = (lambda, k, m1, mu1, m2, mu2, T, Phi, Phibar) (host pointers).
 * replaces: B0  src/relaxtime/OneLoopIntegrals.jl:489-503                  */
int pnjl_B0_batch(pnjl_ctx* ctx, int64_t n, const double* args, double* re, double* im);

/* B0_correction(lambda, k, m1, m2, mu1, mu2, T, Phi, Phibar, xi) for n rows
 * of args[n*10] in that order (host pointers).
 * replaces: B0_correction  src/relaxtime/OneLoopIntegralsAniso.jl:396-413  */
int pnjl_B0_correction_batch(pnjl_ctx* ctx, int64_t n, const double* args, double* re, double* im);

/* A(m, mu, T, Phi, Phibar, nodes_p, weights_p) for n rows of
 * args[n*5] = (m, mu, T, Phi, Phibar); with nc > 0 instead
 * A_aniso(m, mu, T, Phi, Phibar, xi, nodes_p, weights_p, nodes_cos,
 * weights_cos) with args[n*6] = (m, mu, T, Phi, Phibar, xi).  Host pointers.
 * replaces: A  src/relaxtime/OneLoopIntegrals.jl:531-543
 *           A_aniso  src/relaxtime/OneLoopIntegralsAniso.jl:446-467       */
int pnjl_A_batch(pnjl_ctx* ctx, int64_t n, const double* args, int np, const double* nodes_p,
                 const double* weights_p, int nc, const double* nodes_cos, const double* weights_cos,
                 double* out);

/* One (T, mu, xi) state of the polarization grid: the converged gap
 * solution (masses and Polyakov loops) at that point.                      */
typedef struct pnjl_pol_point {
    double T, mu, xi, m_u, m_s, Phi, Phib;
} pnjl_pol_point;

/* The Pi(k0, k) grid of n points: for each point A_aniso(m_u) and
 * A_aniso(m_s) on np momentum nodes (GL on [0, 10] fm^-1) x nc cos nodes
 * (GL on [-1, 1]), then Pi for the pairs (u ubar, m_u m_u, num_s 0),
 * (u sbar, m_u m_s, num_s 1), (s sbar, m_s m_s, num_s 2), both channels
 * (P, S), mu1 = mu2 = mu, at every k0[i] x k[j]:
 *   re/im[((pt*3 + pair)*2 + channel)*nk0*nk + i*nk + j]   (channel 0 = P).
 * Host pointers unless PNJL_FLAG_DEVICE_PTRS.  The A and Pi kernels are
 * captured once into a hipGraph per grid shape and replayed on later calls
 * (inputs are copied into graph-owned buffers first).  No counterpart in
 * the reference as one call: it batches the polarization_aniso calls of
 * TotalPropagator.jl:440-535 / MesonMass.jl:50-216 over a whole grid.     */
int pnjl_polarization_grid(pnjl_ctx* ctx, int npts, const pnjl_pol_point* pts, int nk0, const double* k0,
                           int nk, const double* k, int np, int nc, double* re, double* im, int flags,
                           void* stream);

/* ------------------------------------------------------------------------ */
/* consumers of Pi: effective couplings, meson propagators, D_S / D_P        */
/* ------------------------------------------------------------------------ */
/* The 12 effective couplings, in the reference's NamedTuple order
 * (EffectiveCouplings.jl:272-277):
 *   K0+, K0-, K123+, K123-, K4567+, K4567-, K8+, K8-, K08+, K08-, detK+, detK-   */
#define PNJL_KCOEF_FIELDS 12

/* G_f = calculate_G_from_A(A_f, m_f; Nc) for f = u, s and the couplings
 * calculate_effective_couplings(G, K, G_u, G_s), for n rows of
 * in[n*4] = (A_u, A_s, m_u, m_s), or with from_gf = 1 of in[n*2] =
 * (G_u, G_s) directly.  G_f (n*2, may be NULL) and kc (n*12).  Host
 * pointers.  G and K are arguments (the model's G_fm2 and K_fm5 are the
 * context's params.G and params.K), as in the reference's signature.
 * replaces: calculate_G_from_A  src/relaxtime/EffectiveCouplings.jl:56-60
 *           calculate_effective_couplings, coupling_matrix_determinant
 *           src/relaxtime/EffectiveCouplings.jl:240-278, :111-115           */
int pnjl_effective_couplings_batch(pnjl_ctx* ctx, int64_t n, const double* in, int from_gf, double G, double K,
                                   double* G_f, double* kc);

/* Meson propagators from given Pi: for n rows of kc[n*12] (couplings as
 * above) and pi[n*12] = Pi (re, im) in the pnjl_polarization_grid order
 * (u ubar P, u ubar S, u sbar P, u sbar S, s sbar P, s sbar S), out[n*36],
 * complex values as (re, im):
 *    0 D_pi        2 D_K         4 D_sigma_pi    6 D_sigma_K
 *    8 M00 (P)    10 M08 (P)    12 M88 (P)     14 M00 (S)   16 M08 (S)   18 M88 (S)
 *   20 eta/eta' D for currents (J, J') = (a, a), 22 (a, b), 24 (b, a), 26 (b, b)
 *   28 sigma/sigma' D for the same four current pairs
 * where a is the current of a u ubar or d dbar pair (lambda0, lambda8 at u)
 * and b that of s sbar (lambda8 at s); pairs of different flavour carry no
 * current and give D = 0.  Host pointers.
 * replaces: meson_propagator_simple, meson_propagator_mixed,
 *           calculate_coupling_matrix / calculate_coupling_elements
 *           src/relaxtime/MesonPropagator.jl:215-419
 *           mixing_matrix_elements  src/relaxtime/EffectiveCouplings.jl:137-160 */
#define PNJL_MESON_PROP_FIELDS 36
int pnjl_meson_propagators_batch(pnjl_ctx* ctx, int64_t n, const double* kc, const double* pi, double* out);

/* Mandelstam channel of the exchange (TotalPropagator.jl channel symbols) */
#define PNJL_MANDELSTAM_T 0
#define PNJL_MANDELSTAM_S 1
#define PNJL_MANDELSTAM_U 2

/* Scattering processes, in the order of SCATTERING_MESON_MAP
 * (src/Constants_PNJL.jl:168-400).                                          */
enum {
    PNJL_PROC_UD_TO_UD = 0, PNJL_PROC_UU_TO_UU, PNJL_PROC_US_TO_US, PNJL_PROC_SS_TO_SS,
    PNJL_PROC_UBARDBAR_TO_UBARDBAR, PNJL_PROC_UBARUBAR_TO_UBARUBAR, PNJL_PROC_UBARSBAR_TO_UBARSBAR,
    PNJL_PROC_SBARSBAR_TO_SBARSBAR, PNJL_PROC_UDBAR_TO_UDBAR, PNJL_PROC_DUBAR_TO_DUBAR,
    PNJL_PROC_UUBAR_TO_UUBAR, PNJL_PROC_UUBAR_TO_DDBAR, PNJL_PROC_USBAR_TO_USBAR, PNJL_PROC_SUBAR_TO_SUBAR,
    PNJL_PROC_UUBAR_TO_SSBAR, PNJL_PROC_SSBAR_TO_UUBAR, PNJL_PROC_SSBAR_TO_SSBAR,
    PNJL_PROC_COUNT
};

/* exchanged-meson masks */
#define PNJL_MESON_PI       0x01
#define PNJL_MESON_K        0x02
#define PNJL_MESON_SIGMA_PI 0x04
#define PNJL_MESON_SIGMA_K  0x08
#define PNJL_MESON_MIXED_P  0x10   /* eta / eta'     */
#define PNJL_MESON_MIXED_S  0x20   /* sigma / sigma' */

/* One propagator query: the arguments of
 *   calculate_propagator_by_channel(process, channel, k0, k_norm,
 *                                   quark_params, thermo_params, K_coeffs)
 *     src/relaxtime/TotalPropagator.jl:894-963
 * with quark_params = (m, mu, A) of u (= d) and s and thermo_params =
 * (T, Phi, Phibar, xi).  mesons = 0 takes the exchanged mesons from the
 * process table (the reference's call); a PNJL_MESON_* mask instead gives
 * total_propagator_simple's meson list (:423-470) and / or
 * total_propagator_mixed (:514-560) for any process and channel.          */
typedef struct pnjl_prop_query {
    int    process;               /* PNJL_PROC_*                             */
    int    channel;               /* PNJL_MANDELSTAM_*                       */
    int    mesons;                /* 0 or a PNJL_MESON_* mask                */
    int    reserved;
    double k0, k;                 /* exchanged four-momentum, fm^-1          */
    double T, Phi, Phib, xi;
    double m_u, m_s, mu_u, mu_s, A_u, A_s;
} pnjl_prop_query;

/* (D_S, D_P) for n queries -> out[n*4] = (D_S re, im, D_P re, im).  One
 * kernel evaluates the Pi values the exchanged mesons need (each with the
 * pnjl_polarization_batch arithmetic), a second composes the propagators.
 * kc = NULL derives the couplings from (A_u, A_s, m_u, m_s) with the
 * context's G and K, as the reference's callers do; otherwise kc[n*12]
 * gives them per query.  Host pointers unless PNJL_FLAG_DEVICE_PTRS; a query
 * naming a process, channel or mask outside the table fails the call with
 * PNJL_ERR_ARG (device pointers: its row is NaN).  No host-side memo: every
 * Pi is evaluated (replaces polarization_aniso_cached's Dict,
 * src/relaxtime/PolarizationCache.jl:203-256, for this consumer).
 * replaces: calculate_propagator_by_channel, total_propagator_simple,
 *           total_propagator_mixed, get_polarization_params,
 *           get_flavor_factors_for_channel
 *           src/relaxtime/TotalPropagator.jl:239-560, :894-963              */
int pnjl_propagator_by_channel_batch(pnjl_ctx* ctx, int64_t n, const pnjl_prop_query* q, const double* kc,
                                     double* out, int flags, void* stream);

/* Meson pole masses: the mesons of MesonMass.jl (pi, K, sigma_pi, sigma_K
 * and the mixed eta / eta', sigma / sigma')                                 */
enum {
    PNJL_MESON_ID_PI = 0, PNJL_MESON_ID_K, PNJL_MESON_ID_SIGMA_PI, PNJL_MESON_ID_SIGMA_K, PNJL_MESON_ID_ETA,
    PNJL_MESON_ID_ETA_PRIME, PNJL_MESON_ID_SIGMA, PNJL_MESON_ID_SIGMA_PRIME,
    PNJL_MESON_ID_COUNT
};

/* One pole search: solve_meson_mass(meson, quark_params, thermo_params;
 * K_coeffs, k_norm, initial_mass, initial_gamma, nlsolve_kwargs)
 * (src/relaxtime/MesonMass.jl:241-276).  mass0 = NaN takes
 * default_meson_mass_guess (:183-204).  NLsolve's defaults are ftol 1e-8,
 * xtol 0, max_iter 1000.                                                   */
typedef struct pnjl_mass_query {
    int    meson;                 /* PNJL_MESON_ID_*                         */
    int    max_iter;              /* NLsolve iterations                      */
    int    reserved[2];
    double k;                     /* k_norm, fm^-1                           */
    double mass0, gamma0;         /* initial (mass, gamma), fm^-1            */
    double ftol, xtol;
    double T, Phi, Phib, xi;
    double m_u, m_s, mu_u, mu_s, A_u, A_s;
} pnjl_mass_query;

/* Pole masses for n queries -> out[n*6] = (mass, gamma, residual_norm,
 * converged, iterations, exception) with residual_norm the inf-norm of
 * (Re f, Im f) and exception = 1 where the residual at the initial point is
 * not finite (the reference's NLsolve throws there).  One GPU thread runs
 * one query's trust-region iteration (dogleg, autoscaling, central
 * finite-difference Jacobian, as NLsolve's defaults); every residual is
 * 1 - 4 K Pi(k0 + i gamma/2) with Pi from polarization_with_width, or the
 * light / heavy eigenvalue of 2M for the mixed mesons.  kc = NULL derives
 * the couplings from A as _ensure_K_coeffs does.  Host pointers unless
 * PNJL_FLAG_DEVICE_PTRS.
 * replaces: solve_meson_mass, meson_mass_equation, default_meson_mass_guess
 *           src/relaxtime/MesonMass.jl:183-276                             */
int pnjl_meson_mass_batch(pnjl_ctx* ctx, int64_t n, const pnjl_mass_query* q, const double* kc, double* out,
                          int flags, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PNJL_HIP_H */
