// pnjl_lib.hip -- host side of libpnjl_hip.so: the extern "C" ABI declared in
// include/pnjl_hip.h.  Owns the device buffers, the node tables, the work
// queue counter, the per-group scratch, the phase-boundary tables and the
// launch geometry of a context; launches the kernels of pnjl_scan_kernels.h.
#include <hip/hip_runtime.h>
#include <cstring>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pnjl_hip.h"
#include "pnjl_ctx.h"
#include "pnjl_scan_kernels.h"

using namespace pnjl_host;

namespace pnjl_host {

// n-point Gauss-Legendre rule on [-1,1], ascending (Newton iteration on the
// Legendre three-term recurrence from the Tricomi-type initial guesses), into
// x[n], w[n].
void gl_rule_raw(int n, double* x, double* w)
{
    for (int i = 0; i < n; i++) x[i] = w[i] = 0.0;
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1.0; p1 = z; }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) <= 1e-17 * (1.0 + fabs(z)) && it > 2) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        if (n == 1) { p0 = 1.0; p1 = z; }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

void gl_rule(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    gl_rule_raw(n, x.data(), w.data());
}

// Gauss-Legendre rule on [a, b] into x[n], w[n]
void gauleg_raw(double a, double b, int n, double* x, double* w)
{
    gl_rule_raw(n, x, w);
    double sc = (b - a) / 2.0, sh = (b + a) / 2.0;
    for (int i = 0; i < n; i++) {
        x[i] = sc * x[i] + sh;
        w[i] = sc * w[i];
    }
}

void gauleg_host(double a, double b, int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    gauleg_raw(a, b, n, x.data(), w.data());
}

// meshes of cached_nodes (Integrals.jl:87-97)
void build_meshes(int pn, int tn, std::vector<double>& p, std::vector<double>& c, std::vector<double>& C)
{
    std::vector<double> px, pw, tx, tw;
    gauleg_host(0.0, 10.0, pn, px, pw);
    gauleg_host(0.0, 1.0, tn, tx, tw);
    const double pi = 3.14159265358979323846;
    double twopi2 = (2.0 * pi) * (2.0 * pi);
    p.resize((size_t)pn * tn);
    c.resize((size_t)pn * tn);
    C.resize((size_t)pn * tn);
    for (int l = 0; l < tn; l++)
        for (int j = 0; j < pn; j++) {
            size_t k = (size_t)j + (size_t)l * pn;
            p[k] = px[j];
            c[k] = tx[l];
            C[k] = (pw[j] * (tw[l] * 2.0)) * (px[j] * px[j]) / twopi2;
        }
}

// exp_tab / log_tab tables (pnjl_device.h), in x87 extended precision and
// rounded once to double
void math_tables(double* t)
{
    for (int j = 0; j < 64; j++) t[pnjl::kExpTab + j] = (double)exp2l((long double)j / 64.0L);
    for (int j = 0; j < 64; j++) {
        const long double c = 1.0L + ((long double)j + 0.5L) / 64.0L;
        const double invc = (double)(1.0L / c);
        t[pnjl::kLogTab + 2 * j] = invc;
        t[pnjl::kLogTab + 2 * j + 1] = (double)(-logl((long double)invc));
    }
}

// Node table of (pn, tn) into t[table_len(pn, tn)]: math tables | p^2 [pn] |
// cos^2 [tn] | C [pn*tn] | Cs [pn], with C the coefficients of cached_nodes
// (Integrals.jl:87-97) and Cs their sums over the angle nodes.  work holds
// 2 pn + 2 tn doubles.  Allocates nothing.
void fill_tab(int pn, int tn, double* t, double* work)
{
    math_tables(t);
    double *px = work, *pw = work + pn, *tx = work + 2 * pn, *tw = work + 2 * pn + tn;
    gauleg_raw(0.0, 10.0, pn, px, pw);
    gauleg_raw(0.0, 1.0, tn, tx, tw);
    const double pi = 3.14159265358979323846;
    const double twopi2 = (2.0 * pi) * (2.0 * pi);
    double* nt = t + pnjl::kMathTab;
    double* C = nt + pn + tn;
    for (int j = 0; j < pn; j++) nt[j] = px[j] * px[j];
    for (int l = 0; l < tn; l++) nt[pn + l] = tx[l] * tx[l];
    for (int l = 0; l < tn; l++)
        for (int j = 0; j < pn; j++)
            C[(size_t)j + (size_t)l * pn] = (pw[j] * (tw[l] * 2.0)) * (px[j] * px[j]) / twopi2;
    for (int j = 0; j < pn; j++) {
        double s = 0.0;
        for (int l = 0; l < tn; l++) s += C[(size_t)j + (size_t)l * pn];
        C[(size_t)pn * tn + j] = s;
    }
}

int get_tab(pnjl_ctx* ctx, int pn, int tn, double** out)
{
    auto key = std::make_pair(pn, tn);
    auto it = ctx->tabs.find(key);
    if (it != ctx->tabs.end()) {
        *out = it->second;
        return PNJL_OK;
    }
    std::vector<double> tab((size_t)pnjl::table_len(pn, tn)), work(2 * (size_t)pn + 2 * (size_t)tn);
    fill_tab(pn, tn, tab.data(), work.data());
    double* d = nullptr;
    HIPCHK(ctx, hipMalloc(&d, tab.size() * sizeof(double)));
    HIPCHK(ctx, hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->tabs[key] = d;
    *out = d;
    return PNJL_OK;
}

// staging slot k of at least `bytes`
int stage_buf(pnjl_ctx* ctx, size_t k, size_t bytes, void** out)
{
    if (ctx->stage.size() <= k) {
        ctx->stage.resize(k + 1, nullptr);
        ctx->stage_bytes.resize(k + 1, 0);
    }
    if (ctx->stage_bytes[k] < bytes) {
        if (ctx->stage[k]) HIPCHK(ctx, hipFree(ctx->stage[k]));
        ctx->stage[k] = nullptr;
        ctx->stage_bytes[k] = 0;
        HIPCHK(ctx, hipMalloc(&ctx->stage[k], bytes));
        ctx->stage_bytes[k] = bytes;
    }
    *out = ctx->stage[k];
    return PNJL_OK;
}

// page-locked host staging, grown on demand (hipHostMalloc: allocated, never registered)
int pinned_buf(pnjl_ctx* ctx, size_t bytes, void** out)
{
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) HIPCHK(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        HIPCHK(ctx, hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
        ctx->pinned_bytes = bytes;
    }
    *out = ctx->pinned;
    return PNJL_OK;
}

bool valid_lpp(int l) { return l == 1 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64; }

// Block geometry of the scan kernel for lanes per point `lpp` and node table
// (pn, tn): blocks of whole wavefronts holding at most 64 groups, with the
// node table and every group's LDS state (hot state and candidate records)
// within the block's budget lds_block (80 KB on a 160 KB-LDS CU); a lane
// count whose groups do not fit next to a large table is raised to the next
// power of two (a different summation order, equal to rounding).
struct ScanGeom {
    int lpp, threads, groups_per_block;
    size_t lds;
};

ScanGeom scan_geom(int lpp, int pn, int tn, size_t lds_block)
{
    const size_t tab = (size_t)pnjl::table_len(pn, tn) * sizeof(double);
    const size_t group = (size_t)pnjl::STATE_STRIDE * sizeof(double);
    // the block's LDS budget, less its copy of the launch arguments (static LDS)
    const size_t limit = lds_block - (sizeof(pnjl::ScanArgs) + 255) / 256 * 256;
    for (int l = lpp; l <= 64; l *= 2) {
        for (int waves = 4; waves >= 1; waves--) {
            const int threads = 64 * waves;
            const int gpb = threads / l;
            if (gpb > 64) continue;
            const size_t lds = tab + (size_t)gpb * group;
            if (lds <= limit) return ScanGeom{l, threads, gpb, lds};
        }
    }
    return ScanGeom{64, 64, 1, tab + group};
}

// Lets every scan_kernel instantiation take `bytes` of dynamic LDS.
bool allow_scan_lds(int bytes)
{
    using namespace pnjl;
    const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
    return hipFuncSetAttribute((const void*)scan_kernel<1>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<2>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<4>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<8>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<16>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<32>, a, bytes) == hipSuccess &&
           hipFuncSetAttribute((const void*)scan_kernel<64>, a, bytes) == hipSuccess;
}

void launch_scan(pnjl_ctx* ctx, const pnjl::ScanArgs& A, const ScanGeom& G, int blocks, hipStream_t st)
{
    using namespace pnjl;
    const int k = ctx->n_scans < kMaxTimedScans ? ctx->n_scans++ : -1;
    if (k >= 0) (void)hipEventRecord(ctx->ev_scan[2 * k], st);
    const dim3 grid(blocks), block(G.threads);
    switch (G.lpp) {
    case 1: hipLaunchKernelGGL(scan_kernel<1>, grid, block, G.lds, st, A); break;
    case 2: hipLaunchKernelGGL(scan_kernel<2>, grid, block, G.lds, st, A); break;
    case 4: hipLaunchKernelGGL(scan_kernel<4>, grid, block, G.lds, st, A); break;
    case 8: hipLaunchKernelGGL(scan_kernel<8>, grid, block, G.lds, st, A); break;
    case 16: hipLaunchKernelGGL(scan_kernel<16>, grid, block, G.lds, st, A); break;
    case 32: hipLaunchKernelGGL(scan_kernel<32>, grid, block, G.lds, st, A); break;
    default: hipLaunchKernelGGL(scan_kernel<64>, grid, block, G.lds, st, A); break;
    }
    if (k >= 0) (void)hipEventRecord(ctx->ev_scan[2 * k + 1], st);
}

// ---------------------------------------------------------------------------
// Fixed scratch of pnjl_solve_fixedmu_batch_noalloc: one file-scope device
// array per GPU, carved into the regions a solve needs beyond its inputs and
// outputs.  Calls of the no-alloc entry point on one GPU share it: calls that
// overlap in time on one GPU race.
constexpr int64_t kNaPoints = 8192;    // points per call
constexpr size_t kNaTabLen = 12544;    // doubles >= table_len(pn, tn) for pn * tn <= 4096, pn <= 4096
constexpr size_t na_round(size_t d) { return (d + 31) / 32 * 32; }   // 256-byte boundaries
constexpr size_t kNaQueue = 0;                                                      // kQueueWords u64
constexpr size_t kNaTab = na_round(kNaQueue + kQueueWords);
constexpr size_t kNaRecS = na_round(kNaTab + kNaTabLen);
constexpr size_t kNaRecP = na_round(kNaRecS + (size_t)kNaPoints * pnjl::REC_SIZE);
constexpr size_t kNaRecM = na_round(kNaRecP + (size_t)kNaPoints * pnjl::REC_SIZE);
constexpr size_t kNaRecJ = na_round(kNaRecM + (size_t)kNaPoints * 6 * pnjl::REC_SIZE);
constexpr size_t kNaLists = na_round(kNaRecJ + (size_t)kNaPoints * pnjl::JUDGE_SEEDS * pnjl::REC_SIZE);
constexpr int kNaNumLists = 7;         // mlist, mslot, order, jlist, jslot, jb, j2list (int32, kNaPoints each)
constexpr size_t kNaDoubles = na_round(kNaLists + (size_t)kNaNumLists * kNaPoints / 2);
static_assert(kNaTabLen >= (size_t)pnjl::kMathTab + 4096 + 1 + 4096 + 4096, "node table region");

}  // namespace pnjl_host

static __device__ __align__(256) double scratch[pnjl_host::kNaDoubles];

namespace pnjl_host {

constexpr int kNaMaxDevices = 64;
struct NaDevice {
    double* base = nullptr;   // device address of `scratch` on this GPU
    int pn = 0, tn = 0;       // node table held in its table region
};
static NaDevice g_na[kNaMaxDevices];
static std::mutex g_na_mu;
static double g_na_host[kNaTabLen + 2 * 4096 + 2 * 64];   // host image of a table + fill_tab work

// once per device, with that device current: the device address of `scratch`
int na_init(pnjl_ctx* ctx)
{
    if (ctx->device < 0 || ctx->device >= kNaMaxDevices) return PNJL_OK;   // the entry point reports it
    std::lock_guard<std::mutex> lk(g_na_mu);
    NaDevice& d = g_na[ctx->device];
    if (d.base) return PNJL_OK;
    double* ptr = nullptr;
    HIPCHK(ctx, hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_err(ctx, PNJL_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    d.base = ptr;
    return PNJL_OK;
}

// the node table of (pn, tn) in the device's scratch, rebuilt when the node
// counts change (after the device has finished every earlier call)
int na_table(pnjl_ctx* ctx, int pn, int tn, hipStream_t st, double** out)
{
    std::lock_guard<std::mutex> lk(g_na_mu);
    NaDevice& d = g_na[ctx->device];
    double* dev = d.base + kNaTab;
    if (d.pn != pn || d.tn != tn) {
        HIPCHK(ctx, hipDeviceSynchronize());
        fill_tab(pn, tn, g_na_host, g_na_host + kNaTabLen);
        HIPCHK(ctx, hipMemcpyAsync(dev, g_na_host, (size_t)pnjl::table_len(pn, tn) * sizeof(double),
                                   hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        d.pn = pn;
        d.tn = tn;
    }
    *out = dev;
    return PNJL_OK;
}

int check_policy(int p)
{
    return (p >= PNJL_SEED_DEFAULT && p <= PNJL_SEED_BRANCH_QUARK) ? PNJL_OK : PNJL_ERR_ARG;
}

// core of pnjl_solve_fixedmu_batch with device pointers.  fixed: the node
// table, queue words, the items' candidate records and the per-point lists
// live in the device's no-alloc scratch (n <= kNaPoints) instead of buffers
// the context allocates; the launch, and with it every point's result, is the
// same either way.
int solve_dev(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi, const double* seed,
              int policy, int pn, int tn, const pnjl_solve_opts* o, const pnjl_solve_out* out,
              const pnjl_launch_cfg& cfg, hipStream_t stream, bool fixed = false)
{
    double* tab = nullptr;
    int rc = fixed ? na_table(ctx, pn, tn, stream, &tab) : get_tab(ctx, pn, tn, &tab);
    if (rc) return rc;
    double* const na = fixed ? g_na[ctx->device].base : nullptr;
    ctx->n_scans = 0;
    pnjl::ScanArgs A{};
    A.P = ctx->dp;
    A.n_points = n;
    A.policy = policy;
    const bool chains = pnjl::chain_policy(policy);
    const bool surr = (policy == PNJL_SEED_SURROGATE), multi = (policy == PNJL_SEED_MULTI);
    const bool mr3 = (policy == PNJL_SEED_MULTIROOT3);
    A.plan = chains ? pnjl::PLAN_CHAINS
                    : (surr ? pnjl::PLAN_SURR : (multi ? pnjl::PLAN_MULTI : (mr3 ? pnjl::PLAN_MR3 : pnjl::PLAN_DEFAULT)));
    A.chain_len = chains ? (o->chain_len > 0 ? o->chain_len : (int)std::min<int64_t>(n, 1 << 30)) : 1;
    A.n_items = chains ? (n + A.chain_len - 1) / A.chain_len : n;
    A.mult = multi ? 6 : (mr3 ? 3 : 1);
    A.T = T;
    A.mu = mu;
    A.xi = xi;
    A.seed = seed;
    A.x_out = out->x;
    A.omega = out->omega;
    A.pressure = out->pressure;
    A.rho_norm = out->rho_norm;
    A.entropy = out->entropy;
    A.energy = out->energy;
    A.masses = out->masses;
    A.n_q = out->n_q;
    A.n_qbar = out->n_qbar;
    A.resnorm = out->residual_norm;
    A.iterations = out->iterations;
    A.converged = out->converged;
    A.flags = out->flags;
    A.steps = out->newton_steps;
    A.passes = out->passes;
    A.tab = tab;
    A.pn = pn;
    A.tn = tn;
    A.xtol = o->xtol;
    A.ftol = o->ftol;
    A.resmax = o->residual_norm_max;
    A.phi_tol = o->phi_tol;
    A.max_iter = o->max_iter;
    A.tr_fb = o->trust_region_fallback;
    A.auto_ms = o->auto_multiseed_fallback;
    A.surr_cap = o->surrogate_max_steps;
    A.judge_mode = o->surrogate_judge;
    A.queue = fixed ? (unsigned long long*)(na + kNaQueue) : ctx->queue;
    A.chunk = std::max(1, cfg.chunk);
    A.bnd = ctx->bnd;
    A.nb = ctx->nb;
    A.max_trips = 9LL * (2LL * o->max_iter + 4) + 16;
    const ScanGeom G = scan_geom(cfg.lanes_per_point, pn, tn, ctx->lds_block);
    const int waves_per_block = G.threads / 64;
    int blocks = std::max(1, ctx->num_cu * std::max(1, cfg.waves_per_cu) / waves_per_block);
    // never more groups than work items: a small batch gets a small grid
    const int64_t work = chains ? A.n_items : n * A.mult;
    const int64_t need_groups = std::max<int64_t>(1, (work + A.chunk - 1) / A.chunk);
    if ((int64_t)blocks * G.groups_per_block > need_groups)
        blocks = (int)std::max<int64_t>(1, (need_groups + G.groups_per_block - 1) / G.groups_per_block);
    // the later pipeline stages (few, long items) run the full persistent grid
    // at their own lanes per point
    const ScanGeom G2 = scan_geom(cfg.tail_lanes_per_point > 0 ? cfg.tail_lanes_per_point : G.lpp, pn, tn,
                                  ctx->lds_block);
    const int fullblocks = std::max(1, ctx->num_cu * std::max(1, cfg.waves_per_cu) / (G2.threads / 64));
    unsigned long long* q = A.queue;
    if (chains) {
        hipLaunchKernelGGL(pnjl::zero_u64, dim3(1), dim3(64), 0, stream, q, kQueueWords);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ctx->ev0, stream));
        launch_scan(ctx, A, G, blocks, stream);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ctx->ev1, stream));
        return PNJL_OK;
    }
    // ---- per-point pipeline: stage launches over device lists, then the per-point selection
    const size_t rec = pnjl::REC_SIZE * sizeof(double);
    const bool omega_judge = surr && ctx->jmask && o->surrogate_judge == PNJL_JUDGE_OMEGA;
    void *recS = nullptr, *recP = nullptr, *recM = nullptr, *mlist = nullptr, *mslot = nullptr, *order = nullptr;
    void *recJ = nullptr, *jlist = nullptr, *jslot = nullptr, *jb = nullptr, *j2list = nullptr;
    // staging slot k of `bytes`, or its fixed region (offset in doubles) of the no-alloc scratch
    auto buf = [&](size_t k, size_t bytes, size_t region, void** p) -> int {
        if (!fixed) return stage_buf(ctx, k, bytes, p);
        *p = na + region;
        return PNJL_OK;
    };
    auto list = [](int i) { return kNaLists + (size_t)i * kNaPoints / 2; };
    if ((rc = buf(40, (size_t)n * rec * 6, kNaRecM, &recM))) return rc;
    if (!multi && !mr3) {
        if ((rc = buf(41, (size_t)n * rec, kNaRecP, &recP))) return rc;
        if ((rc = buf(42, (size_t)n * 4, list(0), &mlist))) return rc;
        if ((rc = buf(43, (size_t)n * 4, list(1), &mslot))) return rc;
    }
    if ((rc = buf(45, (size_t)n * 4, list(2), &order))) return rc;
    if (surr) {
        if ((rc = buf(44, (size_t)n * rec, kNaRecS, &recS))) return rc;
        if (omega_judge) {
            if ((rc = buf(46, (size_t)n * pnjl::JUDGE_SEEDS * rec, kNaRecJ, &recJ))) return rc;
            if ((rc = buf(47, (size_t)n * 4, list(3), &jlist))) return rc;
            if ((rc = buf(48, (size_t)n * 4, list(4), &jslot))) return rc;
            if ((rc = buf(49, (size_t)n * 4, list(5), &jb))) return rc;
            if ((rc = buf(50, (size_t)n * 4, list(6), &j2list))) return rc;
        }
    }
    A.recS = (double*)recS;
    A.recP = (double*)recP;
    A.recM = (double*)recM;
    A.mlist = (int32_t*)mlist;
    A.mslot = (int32_t*)mslot;
    if (surr && ctx->jmask) {
        A.judge_mask = ctx->jmask;
        for (int k = 0; k < 3; k++) {
            A.jdim[k] = ctx->jdim[k];
            A.jlo[k] = ctx->jlo[k];
            A.jhi[k] = ctx->jhi[k];
        }
    }
    if (omega_judge) {
        A.recJ = (double*)recJ;
        A.jlist = (int32_t*)jlist;
        A.jslot = (int32_t*)jslot;
        A.jb = (int32_t*)jb;
        A.j2list = (int32_t*)j2list;
    }
    // queue words: q[0] the running stage's item counter (cleared before each
    // stage), q[Q_MCOUNT..Q_J2COUNT] the list lengths, q[kOrderWords..] the
    // ordering pass's class sizes and cursors
    HIPCHK(ctx, hipEventRecord(ctx->ev0, stream));
    hipLaunchKernelGGL(pnjl::zero_u64, dim3(1), dim3(64), 0, stream, q, kQueueWords);
    if (A.passes)
        hipLaunchKernelGGL(pnjl::zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, A.passes, n);
    if (A.steps)
        hipLaunchKernelGGL(pnjl::zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, A.steps, n);
    // stage-1 order: judged points first, xi = 0 points after the others
    const dim3 ogrid((unsigned)((n + pnjl::kOrderBlock - 1) / pnjl::kOrderBlock)), oblock(pnjl::kOrderBlock);
    hipLaunchKernelGGL(pnjl::order_kernel<false>, ogrid, oblock, 0, stream, A, (int32_t*)order, q + kOrderWords);
    hipLaunchKernelGGL(pnjl::order_kernel<true>, ogrid, oblock, 0, stream, A, (int32_t*)order, q + kOrderWords);
    A.order = (const int32_t*)order;
    HIPCHK(ctx, hipGetLastError());
    // stage 1: every point (or every (point, seed) of the multi-seed plans)
    pnjl::ScanArgs S = A;
    S.kind = surr ? pnjl::K_SURR : (multi ? pnjl::K_MPOL : (mr3 ? pnjl::K_MR3 : pnjl::K_PRIMARY));
    launch_scan(ctx, S, G, blocks, stream);
    HIPCHK(ctx, hipGetLastError());
    // the later stages run on the full persistent grid
    auto next_stage = [&](int kind, int mult, int k_base, const void* list, int count_word) -> int {
        hipLaunchKernelGGL(pnjl::zero_u64, dim3(1), dim3(64), 0, stream, q + pnjl::Q_HEAD, 1);
        pnjl::ScanArgs L = A;
        L.kind = kind;
        L.mult = mult;
        L.k_base = k_base;
        L.order = nullptr;
        L.item_list = (const int32_t*)list;
        L.item_count = q + count_word;
        launch_scan(ctx, L, G2, fullblocks, stream);
        HIPCHK(ctx, hipGetLastError());
        return PNJL_OK;
    };
    if (!multi && !mr3 && o->auto_multiseed_fallback)
        if ((rc = next_stage(pnjl::K_MULTI, 6, 0, mlist, pnjl::Q_MCOUNT))) return rc;
    if (omega_judge) {
        if ((rc = next_stage(pnjl::K_JUDGE, 2, 0, jlist, pnjl::Q_JCOUNT))) return rc;
        if ((rc = next_stage(pnjl::K_JUDGE, pnjl::JUDGE_SEEDS - 2, 2, j2list, pnjl::Q_J2COUNT))) return rc;
    }
    pnjl::FinArgs F{};
    F.n = n;
    F.plan = A.plan;
    F.auto_ms = o->auto_multiseed_fallback;
    F.recS = (const double*)recS;
    F.recP = (const double*)recP;
    F.recM = (const double*)recM;
    F.mslot = (const int32_t*)mslot;
    F.recJ = (const double*)recJ;
    F.jslot = (const int32_t*)jslot;
    F.jb = (const int32_t*)jb;
    F.x_out = out->x;
    F.omega = out->omega;
    F.pressure = out->pressure;
    F.rho_norm = out->rho_norm;
    F.entropy = out->entropy;
    F.energy = out->energy;
    F.masses = out->masses;
    F.n_q = out->n_q;
    F.n_qbar = out->n_qbar;
    F.resnorm = out->residual_norm;
    F.iterations = out->iterations;
    F.converged = out->converged;
    F.flags = out->flags;
    hipLaunchKernelGGL(pnjl::finalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, F);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, stream));
    return PNJL_OK;
}

}  // namespace pnjl_host

extern "C" {

void pnjl_default_params(pnjl_model_params* o)
{
    o->hbarc = 197.327;
    o->Lambda = 602.3 / o->hbarc;
    double L2 = o->Lambda * o->Lambda;
    o->G = 1.835 / L2;
    o->K = 12.36 / (L2 * L2 * o->Lambda);
    o->m_ud0 = 5.5 / o->hbarc;
    o->m_s0 = 140.7 / o->hbarc;
    o->T0 = 210.0 / o->hbarc;
    o->a0 = 3.51;
    o->a1 = -2.47;
    o->a2 = 15.2;
    o->b3 = -1.75;
    o->rho0 = 0.16;
    o->N_color = 3;
}

void pnjl_default_solve_opts(pnjl_solve_opts* o)
{
    o->xtol = 1e-9;
    o->ftol = 1e-9;
    o->residual_norm_max = 1e-6;
    o->phi_tol = 1e-8;
    o->max_iter = 1000;
    o->trust_region_fallback = 1;
    o->auto_multiseed_fallback = 1;
    o->surrogate_max_steps = 8;
    o->chain_len = 0;
    o->surrogate_judge = PNJL_JUDGE_REFERENCE;
}

int pnjl_ctx_create(const pnjl_model_params* params, int device, pnjl_ctx** out)
{
    if (!out) return PNJL_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PNJL_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return PNJL_ERR_ARG;
    pnjl_ctx* c = new pnjl_ctx();
    c->device = device;
    if (params)
        c->params = *params;
    else
        pnjl_default_params(&c->params);
    const pnjl_model_params& p = c->params;
    c->dp = pnjl::DevParams{p.Lambda, p.G, p.K, p.m_ud0, p.m_s0, p.T0, p.a0, p.a1, p.a2, p.b3, p.rho0,
                            (double)p.N_color};
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return PNJL_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) {
        c->num_cu = prop.multiProcessorCount;
        // two blocks of four wavefronts per CU (2 waves per SIMD) share its
        // LDS: at most half of it per block, and within the per-block limit
        const size_t half = prop.maxSharedMemoryPerMultiProcessor / 2;
        const size_t want = std::min<size_t>(prop.sharedMemPerBlock, half);
        if (want > c->lds_block && allow_scan_lds((int)want)) c->lds_block = want;
    }
    bool ev_ok = true;
    for (int i = 0; i < 2 * kMaxTimedScans; i++) ev_ok = ev_ok && hipEventCreate(&c->ev_scan[i]) == hipSuccess;
    if (!ev_ok || hipMalloc(&c->queue, kQueueWords * sizeof(unsigned long long)) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return PNJL_ERR_HIP;
    }
    // the device address of the no-alloc entry point's scratch, once per device
    if (na_init(c) != PNJL_OK) {
        pnjl_ctx_destroy(c);
        return PNJL_ERR_HIP;
    }
    *out = c;
    return PNJL_OK;
}

void pnjl_ctx_destroy(pnjl_ctx* c)
{
    if (!c) return;
    // teardown: errors cannot be reported any more and are ignored
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto& kv : c->tabs) (void)hipFree(kv.second);
    for (void* b : c->stage)
        if (b) (void)hipFree(b);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->queue) (void)hipFree(c->queue);
    if (c->bnd) (void)hipFree(c->bnd);
    if (c->jmask) (void)hipFree(c->jmask);
    polar_release(c);
    for (int i = 0; i < 2 * kMaxTimedScans; i++)
        if (c->ev_scan[i]) (void)hipEventDestroy(c->ev_scan[i]);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c;
}

const char* pnjl_last_error(const pnjl_ctx* c) { return c ? c->err.c_str() : "no context"; }

double pnjl_last_kernel_ms(const pnjl_ctx* c) { return c ? c->last_ms : 0.0; }

int pnjl_last_scan_ms(pnjl_ctx* c, double* ms, int max)
{
    if (!c) return PNJL_ERR_ARG;
    const int n = c->n_scans;
    if (n > 0) HIPCHK(c, hipEventSynchronize(c->ev_scan[2 * n - 1]));
    for (int i = 0; i < n && i < max && ms; i++) {
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev_scan[2 * i], c->ev_scan[2 * i + 1]));
        ms[i] = t;
    }
    return n;
}

int pnjl_gauleg(double a, double b, int n, double* nodes, double* weights)
{
    if (n <= 0 || !(a < b) || !nodes || !weights) return PNJL_ERR_ARG;
    std::vector<double> x, w;
    gauleg_host(a, b, n, x, w);
    memcpy(nodes, x.data(), sizeof(double) * n);
    memcpy(weights, w.data(), sizeof(double) * n);
    return PNJL_OK;
}

int pnjl_nodes(pnjl_ctx* ctx, int p_num, int t_num, double* p_mesh, double* cos_mesh, double* coef_mesh)
{
    if (p_num <= 0 || t_num <= 0 || p_num > 4096 || t_num > 64 || (int64_t)p_num * t_num > 4096)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_nodes: p_num/t_num out of range");
    std::vector<double> p, c, C;
    build_meshes(p_num, t_num, p, c, C);
    if (p_mesh) memcpy(p_mesh, p.data(), sizeof(double) * p.size());
    if (cos_mesh) memcpy(cos_mesh, c.data(), sizeof(double) * c.size());
    if (coef_mesh) memcpy(coef_mesh, C.data(), sizeof(double) * C.size());
    if (ctx) {
        double* d;
        HIPCHK(ctx, hipSetDevice(ctx->device));
        return get_tab(ctx, p_num, t_num, &d);
    }
    return PNJL_OK;
}

int pnjl_set_phase_boundary(pnjl_ctx* ctx, int nrows, const double* xi, const double* T, const double* mu,
                            int ncep, const double* cxi, const double* cT, const double* cmu)
{
    if (!ctx || nrows < 0 || ncep < 0) return PNJL_ERR_ARG;
    using pnjl::kBndStride;
    using pnjl::kMaxBndRows;
    std::vector<double> xs;
    for (int i = 0; i < nrows; i++)
        if (std::find(xs.begin(), xs.end(), xi[i]) == xs.end()) xs.push_back(xi[i]);
    std::vector<double> h(xs.size() * kBndStride, 0.0);
    for (size_t k = 0; k < xs.size(); k++) {
        double* s = &h[k * kBndStride];
        std::vector<std::pair<double, double>> rows;
        for (int i = 0; i < nrows; i++)
            if (xi[i] == xs[k]) rows.push_back({T[i], mu[i]});
        std::stable_sort(rows.begin(), rows.end(),
                         [](const std::pair<double, double>& a, const std::pair<double, double>& b) {
                             return a.first < b.first;
                         });
        if ((int)rows.size() > kMaxBndRows)
            return set_err(ctx, PNJL_ERR_ARG, "pnjl_set_phase_boundary: too many rows for one xi");
        s[0] = xs[k];
        s[1] = NAN;
        for (int c = 0; c < ncep; c++)
            if (fabs(cxi[c] - xs[k]) <= 1e-6) {
                s[1] = cT[c];
                break;
            }
        s[2] = (double)rows.size();
        for (size_t i = 0; i < rows.size(); i++) {
            s[4 + i] = rows[i].first;
            s[4 + kMaxBndRows + i] = rows[i].second;
        }
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->bnd) HIPCHK(ctx, hipFree(ctx->bnd));
    ctx->bnd = nullptr;
    ctx->nb = 0;
    if (!h.empty()) {
        HIPCHK(ctx, hipMalloc(&ctx->bnd, h.size() * sizeof(double)));
        HIPCHK(ctx, hipMemcpy(ctx->bnd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        ctx->nb = (int)xs.size();
    }
    return PNJL_OK;
}

int pnjl_set_judge_mask(pnjl_ctx* ctx, int nT, int nmu, int nxi, const double* lo, const double* hi,
                        const uint8_t* mask)
{
    if (!ctx) return PNJL_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (ctx->jmask) HIPCHK(ctx, hipFree(ctx->jmask));
    ctx->jmask = nullptr;
    if (!mask) return PNJL_OK;
    if (nT <= 0 || nmu <= 0 || nxi <= 0 || (int64_t)nT * nmu * nxi > (1 << 26) || !lo || !hi)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_set_judge_mask: bad geometry");
    for (int k = 0; k < 3; k++)
        if (!(hi[k] > lo[k])) return set_err(ctx, PNJL_ERR_ARG, "pnjl_set_judge_mask: empty box");
    size_t bytes = (size_t)nT * nmu * nxi;
    HIPCHK(ctx, hipMalloc(&ctx->jmask, bytes));
    HIPCHK(ctx, hipMemcpy(ctx->jmask, mask, bytes, hipMemcpyHostToDevice));
    ctx->jdim[0] = nT;
    ctx->jdim[1] = nmu;
    ctx->jdim[2] = nxi;
    for (int k = 0; k < 3; k++) {
        ctx->jlo[k] = lo[k];
        ctx->jhi[k] = hi[k];
    }
    return PNJL_OK;
}

int pnjl_set_launch_cfg(pnjl_ctx* ctx, const pnjl_launch_cfg* cfg)
{
    if (!ctx || !cfg || !valid_lpp(cfg->lanes_per_point) || cfg->chunk < 1 || cfg->waves_per_cu < 1 ||
        (cfg->tail_lanes_per_point != 0 && !valid_lpp(cfg->tail_lanes_per_point)) ||
        cfg->waves_per_cu > 64)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_set_launch_cfg: invalid geometry");
    ctx->cfg = *cfg;
    return PNJL_OK;
}

int pnjl_get_launch_cfg(const pnjl_ctx* ctx, pnjl_launch_cfg* cfg)
{
    if (!ctx || !cfg) return PNJL_ERR_ARG;
    *cfg = ctx->cfg;
    return PNJL_OK;
}

static int solve_impl(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi,
                      const double* seed, int policy, int pn, int tn, const pnjl_solve_opts* opts,
                      const pnjl_solve_out* out, int flags, void* stream, const pnjl_launch_cfg& cfg,
                      bool fixed = false)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || !T || !mu || !xi || !out || check_policy(policy))
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_solve_fixedmu_batch: bad argument");
    if ((policy == PNJL_SEED_EXPLICIT || policy == PNJL_SEED_SURROGATE || policy == PNJL_SEED_MULTIROOT3) && !seed)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_solve_fixedmu_batch: this policy needs seeds");
    if (!out->x || !out->omega || !out->pressure || !out->rho_norm || !out->entropy || !out->energy ||
        !out->masses || !out->residual_norm || !out->iterations || !out->converged)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_solve_fixedmu_batch: missing output array");
    if (pn <= 0 || tn <= 0 || (int64_t)pn * tn > 4096)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_solve_fixedmu_batch: p_num/t_num out of range");
    pnjl_solve_opts o;
    if (opts)
        o = *opts;
    else
        pnjl_default_solve_opts(&o);
    if (o.max_iter < 0 || o.max_iter > 100000) return set_err(ctx, PNJL_ERR_ARG, "max_iter out of range");
    if (o.surrogate_judge != PNJL_JUDGE_REFERENCE && o.surrogate_judge != PNJL_JUDGE_OMEGA)
        return set_err(ctx, PNJL_ERR_ARG, "surrogate_judge must be PNJL_JUDGE_REFERENCE or PNJL_JUDGE_OMEGA");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (flags & PNJL_FLAG_DEVICE_PTRS) {
        rc = solve_dev(ctx, n, T, mu, xi, seed, policy, pn, tn, &o, out, cfg, st, fixed);
        if (rc) return rc;
    } else {
        // Host arrays travel packed: the inputs are gathered into one page-locked
        // buffer and go to HBM in one copy; the kernels write the outputs into
        // one device block, which comes back in one copy and is scattered into
        // the caller's arrays (two DMA transfers per call instead of ~19
        // pageable ones).  Layout (in units of n values): inputs T, mu, xi,
        // seed(5); outputs x(5), omega, pressure, rho_norm, entropy, energy,
        // masses(3), n_q(3), n_qbar(3), residual_norm as doubles, then
        // iterations, converged, flags, newton_steps, passes as int32.
        const size_t nd = (size_t)n * sizeof(double), ni = (size_t)n * 4;
        const size_t in_bytes = (seed ? 8 : 3) * nd;
        const size_t out_d = 20 * nd, out_bytes = out_d + 5 * ni;
        void *din = nullptr, *dout = nullptr, *hp = nullptr;
        if ((rc = stage_buf(ctx, 0, in_bytes, &din))) return rc;
        if ((rc = stage_buf(ctx, 1, out_bytes, &dout))) return rc;
        if ((rc = pinned_buf(ctx, std::max(in_bytes, out_bytes), &hp))) return rc;
        // the pinned buffer may still be read by the previous call's copy only
        // if that call was asynchronous, which host-pointer calls never are
        char* h = (char*)hp;
        std::memcpy(h, T, nd);
        std::memcpy(h + nd, mu, nd);
        std::memcpy(h + 2 * nd, xi, nd);
        if (seed) std::memcpy(h + 3 * nd, seed, 5 * nd);
        HIPCHK(ctx, hipMemcpyAsync(din, h, in_bytes, hipMemcpyHostToDevice, st));
        double* di = (double*)din;
        double* dd = (double*)dout;
        int32_t* dj = (int32_t*)((char*)dout + out_d);
        pnjl_solve_out d;
        d.x = dd;
        d.omega = dd + 5 * n;
        d.pressure = dd + 6 * n;
        d.rho_norm = dd + 7 * n;
        d.entropy = dd + 8 * n;
        d.energy = dd + 9 * n;
        d.masses = dd + 10 * n;
        d.n_q = out->n_q ? dd + 13 * n : nullptr;
        d.n_qbar = out->n_qbar ? dd + 16 * n : nullptr;
        d.residual_norm = dd + 19 * n;
        d.iterations = dj;
        d.converged = dj + n;
        d.flags = out->flags ? dj + 2 * n : nullptr;
        d.newton_steps = out->newton_steps ? dj + 3 * n : nullptr;
        d.passes = out->passes ? dj + 4 * n : nullptr;
        rc = solve_dev(ctx, n, di, di + n, di + 2 * n, seed ? di + 3 * n : nullptr, policy, pn, tn, &o, &d, cfg,
                       st);
        if (rc) return rc;
        HIPCHK(ctx, hipMemcpyAsync(h, dout, out_bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        const double* hd = (const double*)h;
        const int32_t* hj = (const int32_t*)(h + out_d);
        std::memcpy(out->x, hd, 5 * nd);
        std::memcpy(out->omega, hd + 5 * n, nd);
        std::memcpy(out->pressure, hd + 6 * n, nd);
        std::memcpy(out->rho_norm, hd + 7 * n, nd);
        std::memcpy(out->entropy, hd + 8 * n, nd);
        std::memcpy(out->energy, hd + 9 * n, nd);
        std::memcpy(out->masses, hd + 10 * n, 3 * nd);
        if (out->n_q) std::memcpy(out->n_q, hd + 13 * n, 3 * nd);
        if (out->n_qbar) std::memcpy(out->n_qbar, hd + 16 * n, 3 * nd);
        std::memcpy(out->residual_norm, hd + 19 * n, nd);
        std::memcpy(out->iterations, hj, ni);
        std::memcpy(out->converged, hj + n, ni);
        if (out->flags) std::memcpy(out->flags, hj + 2 * n, ni);
        if (out->newton_steps) std::memcpy(out->newton_steps, hj + 3 * n, ni);
        if (out->passes) std::memcpy(out->passes, hj + 4 * n, ni);
    }
    if (!(flags & PNJL_FLAG_ASYNC) || !(flags & PNJL_FLAG_DEVICE_PTRS)) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

int pnjl_solve_fixedmu_batch(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi,
                             const double* seed, int seed_policy, int p_num, int t_num, const pnjl_solve_opts* opts,
                             const pnjl_solve_out* out, int flags, void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    return solve_impl(ctx, n, T, mu, xi, seed, seed_policy, p_num, t_num, opts, out, flags, stream, ctx->cfg);
}

int pnjl_solve_fixedmu_batch_noalloc(pnjl_ctx* ctx, int64_t n, const double* T, const double* mu, const double* xi,
                                     const double* seed, int seed_policy, int p_num, int t_num,
                                     const pnjl_solve_opts* opts, const pnjl_solve_out* out, int flags,
                                     void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (!(flags & PNJL_FLAG_DEVICE_PTRS))
        return set_err(ctx, PNJL_ERR_ARG,
                       "pnjl_solve_fixedmu_batch_noalloc: device pointers only (PNJL_FLAG_DEVICE_PTRS); host "
                       "arrays would need staging buffers");
    if (n > kNaPoints)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_solve_fixedmu_batch_noalloc: n = " + std::to_string(n) +
                                              " exceeds the maximum of " + std::to_string(kNaPoints) +
                                              " points per call");
    if (ctx->device < 0 || ctx->device >= kNaMaxDevices || !g_na[ctx->device].base)
        return set_err(ctx, PNJL_ERR_STATE, "pnjl_solve_fixedmu_batch_noalloc: no scratch on this device");
    return solve_impl(ctx, n, T, mu, xi, seed, seed_policy, p_num, t_num, opts, out, flags, stream, ctx->cfg,
                      true);
}

int pnjl_gap_residual_batch(pnjl_ctx* ctx, int64_t n, const double* x, const double* T, const double* mu,
                            const double* xi, int pn, int tn, double* F, double* J, double* omega, int flags,
                            void* stream)
{
    if (!ctx) return PNJL_ERR_ARG;
    if (n < 0 || !x || !T || !mu || !xi || !F || !J || !omega || pn <= 0 || tn <= 0 || (int64_t)pn * tn > 4096)
        return set_err(ctx, PNJL_ERR_ARG, "pnjl_gap_residual_batch: bad argument");
    if (n == 0) return PNJL_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    double* tab = nullptr;
    int rc = get_tab(ctx, pn, tn, &tab);
    if (rc) return rc;
    const double *dx = x, *dT = T, *dmu = mu, *dxi = xi;
    double *dF = F, *dJ = J, *dom = omega;
    size_t nd = (size_t)n * sizeof(double);
    if (!(flags & PNJL_FLAG_DEVICE_PTRS)) {
        void* b[7];
        size_t sizes[] = {5 * nd, nd, nd, nd, 5 * nd, 25 * nd, nd};
        for (int k = 0; k < 7; k++) {
            rc = stage_buf(ctx, 20 + k, sizes[k], &b[k]);
            if (rc) return rc;
        }
        HIPCHK(ctx, hipMemcpyAsync(b[0], x, 5 * nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[1], T, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[2], mu, nd, hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(b[3], xi, nd, hipMemcpyHostToDevice, st));
        dx = (double*)b[0];
        dT = (double*)b[1];
        dmu = (double*)b[2];
        dxi = (double*)b[3];
        dF = (double*)b[4];
        dJ = (double*)b[5];
        dom = (double*)b[6];
    }
    pnjl::ResArgs R{};
    R.P = ctx->dp;
    R.n = n;
    R.x = dx;
    R.T = dT;
    R.mu = dmu;
    R.xi = dxi;
    R.F = dF;
    R.J = dJ;
    R.omega = dom;
    R.tab = tab;
    R.pn = pn;
    R.tn = tn;
    const int LPP = 16;
    int64_t blocks = (n * LPP + 255) / 256;
    const size_t lds = sizeof(double) * ((size_t)pnjl::table_len(pn, tn) + (256 / LPP) * pnjl::STATE_STRIDE);
    HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(pnjl::residual_kernel<LPP>, dim3((unsigned)blocks), dim3(256), lds, st, R);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
    if (!(flags & PNJL_FLAG_DEVICE_PTRS)) {
        HIPCHK(ctx, hipMemcpyAsync(F, dF, 5 * nd, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(J, dJ, 25 * nd, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(omega, dom, nd, hipMemcpyDeviceToHost, st));
    }
    if (!(flags & PNJL_FLAG_ASYNC) || !(flags & PNJL_FLAG_DEVICE_PTRS)) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    }
    return PNJL_OK;
}

int pnjl_autotune(pnjl_ctx* ctx, int64_t sample, const double* T, const double* mu, const double* xi,
                  const double* seed, int policy, int pn, int tn, const pnjl_solve_opts* opts, pnjl_launch_cfg* best,
                  double* ns_per_point)
{
    if (!ctx || sample <= 0 || !T || !mu || !xi) return PNJL_ERR_ARG;
    std::vector<double> x(5 * sample), om(sample), pr(sample), rh(sample), en(sample), ee(sample), ms(3 * sample),
        rs(sample);
    std::vector<int32_t> it(sample), cv(sample);
    pnjl_solve_out out{};
    out.x = x.data();
    out.omega = om.data();
    out.pressure = pr.data();
    out.rho_norm = rh.data();
    out.entropy = en.data();
    out.energy = ee.data();
    out.masses = ms.data();
    out.residual_norm = rs.data();
    out.iterations = it.data();
    out.converged = cv.data();
    pnjl_launch_cfg bcfg = ctx->cfg;
    double bms = 1e300;
    auto trial = [&](const pnjl_launch_cfg& cfg, double* t_out) -> int {
        // warm once, then keep the better of two timed runs (HIP events
        // around the pipeline)
        double t = 1e300;
        for (int rep = 0; rep < 3; rep++) {
            int rc = solve_impl(ctx, sample, T, mu, xi, seed, policy, pn, tn, opts, &out, 0, nullptr, cfg);
            if (rc) return rc;
            if (rep > 0) t = std::min(t, ctx->last_ms);
        }
        if (t < bms) {
            bms = t;
            bcfg = cfg;
        }
        if (t_out) *t_out = t;
        return PNJL_OK;
    };
    // 1) lanes per point (points per wavefront = 64 / lanes; nodes per lane =
    //    p_num * t_eff / lanes) at queue chunk 1 and 12 waves/CU
    const int lanes[] = {1, 2, 4, 8, 16, 32, 64};
    double tl[7];
    for (int i = 0; i < 7; i++)
        if (int rc = trial(pnjl_launch_cfg{lanes[i], 1, 12, 0}, &tl[i])) return rc;
    // 2) queue chunk and resident waves per CU for the two best lane counts
    int i1 = 0, i2 = -1;
    for (int i = 1; i < 7; i++)
        if (tl[i] < tl[i1]) i1 = i;
    for (int i = 0; i < 7; i++)
        if (i != i1 && (i2 < 0 || tl[i] < tl[i2])) i2 = i;
    for (int i : {i1, i2})
        for (int c : {1, 2, 4})
            for (int w : {8, 12, 16})
                if (!(c == 1 && w == 12))
                    if (int rc = trial(pnjl_launch_cfg{lanes[i], c, w, 0}, nullptr)) return rc;
    // 3) lanes of the later stages (MultiSeed and judge candidates: few, long items)
    const pnjl_launch_cfg b2 = bcfg;
    for (int t : {16, 32, 64})
        if (t != b2.lanes_per_point)
            if (int rc = trial(pnjl_launch_cfg{b2.lanes_per_point, b2.chunk, b2.waves_per_cu, t}, nullptr)) return rc;
    ctx->cfg = bcfg;
    if (best) *best = bcfg;
    if (ns_per_point) *ns_per_point = bms * 1e6 / (double)sample;
    return PNJL_OK;
}

}  // extern "C"
