// pnjl_ctx.h -- the library context (opaque pnjl_ctx of include/pnjl_hip.h)
// and the host helpers shared by the translation units of libpnjl_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pnjl_hip.h"
#include "pnjl_device.h"

// device work-queue counters of a solve (u64 words: the Q_* words of the
// scan kernel, then two counters of the learned-seed plan's ordering pass)
constexpr int kQueueWords = 16;
constexpr int kOrderWords = 8;
// persistent-kernel launches timed per solve
constexpr int kMaxTimedScans = 8;

struct pnjl_ctx {
    int device = 0;
    int num_cu = 256;
    pnjl_model_params params{};
    pnjl::DevParams dp{};
    std::string err;
    pnjl_launch_cfg cfg{4, 4, 8, 0};
    std::map<std::pair<int, int>, double*> tabs;
    size_t lds_block = 64 * 1024;       // LDS budget of one scan-kernel block (bytes)
    unsigned long long* queue = nullptr;
    double* bnd = nullptr;
    int nb = 0;
    uint8_t* jmask = nullptr;
    int jdim[3] = {0, 0, 0};
    double jlo[3] = {0, 0, 0}, jhi[3] = {1, 1, 1};
    std::vector<void*> stage;   // device staging buffers for host-pointer calls
    std::vector<size_t> stage_bytes;
    void* pinned = nullptr;     // page-locked host staging of the packed host-pointer solves
    size_t pinned_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // begin/end events of each persistent-kernel launch of the last solve
    hipEvent_t ev_scan[2 * kMaxTimedScans] = {};
    int n_scans = 0;
    // polarization (pnjl_polar.hip): quadrature tables and the captured grid graph
    double* polar_tab = nullptr;
    struct PolarGraph* pgraph = nullptr;
};


namespace pnjl_host {

inline int set_err(pnjl_ctx* c, int code, const std::string& m)
{
    if (c) c->err = m;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return pnjl_host::set_err(ctx, PNJL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// node table of (p_num, t_num) on the device (math tables | p^2 | cos^2 | C | Cs), cached per context
int get_tab(pnjl_ctx* ctx, int pn, int tn, double** out);
// n-point Gauss-Legendre rule on [-1,1], ascending
void gl_rule(int n, std::vector<double>& x, std::vector<double>& w);
// device staging buffer k of at least `bytes` (grown on demand, owned by the context)
int stage_buf(pnjl_ctx* ctx, size_t k, size_t bytes, void** out);
// frees the polarization resources of a context (pnjl_polar.hip)
void polar_release(pnjl_ctx* ctx);

}  // namespace pnjl_host
