// pnjl_mlp.hip -- the learned seed: MLP 3 -> 128 -> 128 -> 128 -> 4 (SiLU) mapping
// normalised (T, mu, xi) to normalised (phi_u, phi_s, Phi, Phibar), trained on
// roots the plain solver produced.  BF16 weights and activations, FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), FP32
// master weights updated by Adam.
//
// Kernels
//   train_tile_kernel   one 32-sample tile per block (4 waves): forward, MSE
//                       loss, backward; writes the block's partial gradient
//   adam_kernel         sums the block partials and applies Adam; refreshes
//                       the BF16 (and transposed BF16) weight copies and the
//                       BF16 residuals W - bf16(W) of the hidden layers
//   infer_kernel        forward only, at FP32-class precision: hidden-layer
//                       products as split BF16 (x = hi + lo, three MFMAs:
//                       hi*hi + hi*lo + lo*hi), first and last layers on the
//                       FP32 masters; writes FP64 seeds (phi_u, phi_u, phi_s,
//                       Phi, Phibar).  Training runs the plain BF16 forward;
//                       the seeds come from the FP32 weights it maintains,
//                       whose exact forward lands closer to the roots (fewer
//                       Newton steps) than its BF16 rounding.
//
// MFMA lane layout (wave64, 16x16x32): lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16] (i = 0..7) and receives D[4(l/16)+r][l%16] (r = 0..3).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/pnjl_hip.h"

namespace mlp {

constexpr int IN = 3, H = 128, OUT = 4, TILE = 32, THREADS = 256;
// packed FP32 parameter vector (row-major): W1[H][IN] b1[H] W2[H][H] b2[H] W3[H][H] b3[H] W4[OUT][H] b4[OUT]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H, OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + OUT * H,
              NPARAM = OFF_B4 + OUT;
static_assert(NPARAM == 34052, "parameter count");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

// BF16 weight copies used by the kernels (refreshed from the FP32 masters)
struct WB {
    unsigned short* W1;   // [H][IN]
    unsigned short* W2;   // [H][H]  (o, i)
    unsigned short* W2T;  // [H][H]  (i, o)
    unsigned short* W3;
    unsigned short* W3T;
    unsigned short* W4;   // [OUT][H]
    unsigned short* W2lo; // [H][H]  (o, i): bf16(W2 - bf16(W2))
    unsigned short* W3lo;
};

struct TrainArgs {
    const float* xin;     // [n][IN] normalised inputs
    const float* tgt;     // [n][OUT] normalised targets
    int64_t n;
    int64_t batch;        // samples in this step
    int explicit_rows;    // 1: rows 0..batch-1 in order (gradient check), 0: random draw
    uint64_t seed, step;
    const float* P;       // FP32 master parameters (biases are read from here)
    WB w;
    float* partial;       // [nblocks][NPARAM]
    float* loss_partial;  // [nblocks]
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// ---- 16x16 output tile of C = A(MxK) B(KxN) with K a multiple of 32 ----
// Fragment loaders are passed as lambdas returning the 8 bf16 values.
template <typename FA, typename FB>
__device__ __forceinline__ float4v mma_tile(int K, FA fa, FB fb)
{
    float4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < K; kk += 32) {
        bf16x8 a = __builtin_bit_cast(bf16x8, fa(kk));
        bf16x8 b = __builtin_bit_cast(bf16x8, fb(kk));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ us8 ld8(const unsigned short* p) { return *(const us8*)p; }

__device__ __forceinline__ us8 gather8(const unsigned short* p, int stride)
{
    us8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = p[i * stride];
    return v;
}

// forward hidden layer: out[s][o] = silu(sum_i in[s][i] W[o][i] + b[o]); also silu'(z)
__device__ void fwd_hidden(const unsigned short* in, const unsigned short* W, const float* b, unsigned short* out,
                           unsigned short* gout)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (int t = wave; t < (TILE / 16) * (H / 16); t += THREADS / 64) {
        const int ts = t / (H / 16), to = t % (H / 16);
        float4v acc = mma_tile(
            H, [&](int kk) { return ld8(in + (ts * 16 + r) * H + kk + 8 * q); },
            [&](int kk) { return ld8(W + (to * 16 + r) * H + kk + 8 * q); });
        const int o = to * 16 + r;
        const float bo = b[o];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int srow = ts * 16 + 4 * q + i;
            float z = acc[i] + bo;
            float sg = sigmoidf_(z);
            out[srow * H + o] = f2bf(z * sg);
            if (gout) gout[srow * H + o] = f2bf(sg * (1.0f + z * (1.0f - sg)));
        }
    }
}

// backward through a hidden layer's weights: dh[s][i] = sum_o dz[s][o] W[o][i],
// then dzp[s][i] = dh * g[s][i] (BF16)
__device__ void bwd_data(const unsigned short* dz, const unsigned short* WT, const unsigned short* g,
                         unsigned short* dzp)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (int t = wave; t < (TILE / 16) * (H / 16); t += THREADS / 64) {
        const int ts = t / (H / 16), ti = t % (H / 16);
        float4v acc = mma_tile(
            H, [&](int kk) { return ld8(dz + (ts * 16 + r) * H + kk + 8 * q); },
            [&](int kk) { return ld8(WT + (ti * 16 + r) * H + kk + 8 * q); });
        const int i0 = ti * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int srow = ts * 16 + 4 * q + i;
            dzp[srow * H + i0] = f2bf(acc[i] * bf2f(g[srow * H + i0]));
        }
    }
}

// weight gradient of a hidden layer: dW[o][i] = sum_s dz[s][o] in[s][i]
__device__ void bwd_weight(const unsigned short* dz, const unsigned short* in, float* dW)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (int t = wave; t < (H / 16) * (H / 16); t += THREADS / 64) {
        const int to = t / (H / 16), ti = t % (H / 16);
        float4v acc = mma_tile(
            TILE, [&](int kk) { return gather8(dz + (kk + 8 * q) * H + to * 16 + r, H); },
            [&](int kk) { return gather8(in + (kk + 8 * q) * H + ti * 16 + r, H); });
        const int i0 = ti * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; i++) dW[(to * 16 + 4 * q + i) * H + i0] = acc[i];
    }
}

struct Smem {
    float xs[TILE][IN];
    float tg[TILE][OUT];
    float y[TILE][OUT];
    float dy[TILE][OUT];
    float red[THREADS];
    int valid[TILE];
    unsigned short h1[TILE * H], h2[TILE * H], h3[TILE * H];
    // silu'(z) of each hidden layer; the backward pass overwrites g_l in place
    // with dL/dz_l (each element is read once, then written by the same lane)
    unsigned short g1[TILE * H], g2[TILE * H], g3[TILE * H];
};
static_assert(sizeof(Smem) <= 64 * 1024, "LDS budget");

// first layer (K = 3) with scalar FMAs
__device__ void fwd_first(Smem& S, const unsigned short* W1, const float* b1)
{
    for (int e = threadIdx.x; e < TILE * H; e += THREADS) {
        const int s = e / H, o = e % H;
        float z = b1[o];
#pragma unroll
        for (int k = 0; k < IN; k++) z += bf2f(W1[o * IN + k]) * S.xs[s][k];
        float sg = sigmoidf_(z);
        S.h1[e] = f2bf(z * sg);
        S.g1[e] = f2bf(sg * (1.0f + z * (1.0f - sg)));
    }
}

// last layer (N = 4)
__device__ void fwd_last(Smem& S, const unsigned short* W4, const float* b4)
{
    for (int e = threadIdx.x; e < TILE * OUT; e += THREADS) {
        const int s = e / OUT, j = e % OUT;
        float z = b4[j];
        for (int o = 0; o < H; o++) z += bf2f(W4[j * H + o]) * bf2f(S.h3[s * H + o]);
        S.y[s][j] = z;
    }
}

__global__ void __launch_bounds__(THREADS) train_tile_kernel(TrainArgs A)
{
    extern __shared__ unsigned char smem_raw[];
    Smem& S = *reinterpret_cast<Smem*>(smem_raw);
    const int64_t row0 = (int64_t)blockIdx.x * TILE;
    float* G = A.partial + (int64_t)blockIdx.x * NPARAM;
    // gather the tile
    for (int s = threadIdx.x; s < TILE; s += THREADS) {
        int64_t row = row0 + s;
        bool ok = row < A.batch;
        int64_t idx = 0;
        if (ok) idx = A.explicit_rows ? row : (int64_t)(splitmix64(A.seed ^ (A.step * 0x100000001B3ULL) ^
                                                                   (uint64_t)row * 0x9E3779B97F4A7C15ULL) %
                                                        (uint64_t)A.n);
        S.valid[s] = ok ? 1 : 0;
        for (int k = 0; k < IN; k++) S.xs[s][k] = ok ? A.xin[idx * IN + k] : 0.f;
        for (int j = 0; j < OUT; j++) S.tg[s][j] = ok ? A.tgt[idx * OUT + j] : 0.f;
    }
    __syncthreads();
    // forward
    fwd_first(S, A.w.W1, A.P + OFF_B1);
    __syncthreads();
    fwd_hidden(S.h1, A.w.W2, A.P + OFF_B2, S.h2, S.g2);
    __syncthreads();
    fwd_hidden(S.h2, A.w.W3, A.P + OFF_B3, S.h3, S.g3);
    __syncthreads();
    fwd_last(S, A.w.W4, A.P + OFF_B4);
    __syncthreads();
    // MSE over the batch and the four outputs: L = sum (y - t)^2 / (4 B)
    const float inv = 1.0f / (4.0f * (float)A.batch);
    float lsum = 0.f;
    for (int e = threadIdx.x; e < TILE * OUT; e += THREADS) {
        const int s = e / OUT, j = e % OUT;
        float d = S.valid[s] ? (S.y[s][j] - S.tg[s][j]) : 0.f;
        S.dy[s][j] = 2.0f * d * inv;
        lsum += d * d * inv;
    }
    S.red[threadIdx.x] = lsum;
    __syncthreads();
    for (int off = THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) S.red[threadIdx.x] += S.red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) A.loss_partial[blockIdx.x] = S.red[0];
    // layer 4 gradients and dz3
    for (int e = threadIdx.x; e < OUT * H; e += THREADS) {
        const int j = e / H, o = e % H;
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += S.dy[s][j] * bf2f(S.h3[s * H + o]);
        G[OFF_W4 + e] = acc;
    }
    if (threadIdx.x < OUT) {
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += S.dy[s][threadIdx.x];
        G[OFF_B4 + threadIdx.x] = acc;
    }
    __syncthreads();   // dW4 has read h3
    for (int e = threadIdx.x; e < TILE * H; e += THREADS) {
        const int s = e / H, o = e % H;
        float dh = 0.f;
#pragma unroll
        for (int j = 0; j < OUT; j++) dh += S.dy[s][j] * bf2f(A.w.W4[j * H + o]);
        S.g3[e] = f2bf(dh * bf2f(S.g3[e]));   // dz3
    }
    __syncthreads();
    // layer 3
    bwd_weight(S.g3, S.h2, G + OFF_W3);
    if (threadIdx.x < H) {
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += bf2f(S.g3[s * H + threadIdx.x]);
        G[OFF_B3 + threadIdx.x] = acc;
    }
    bwd_data(S.g3, A.w.W3T, S.g2, S.g2);  // dz2 in g2
    __syncthreads();
    // layer 2
    bwd_weight(S.g2, S.h1, G + OFF_W2);
    if (threadIdx.x < H) {
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += bf2f(S.g2[s * H + threadIdx.x]);
        G[OFF_B2 + threadIdx.x] = acc;
    }
    bwd_data(S.g2, A.w.W2T, S.g1, S.g1);  // dz1 in g1
    __syncthreads();
    // layer 1 (K = 3)
    for (int e = threadIdx.x; e < H * IN; e += THREADS) {
        const int o = e / IN, k = e % IN;
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += bf2f(S.g1[s * H + o]) * S.xs[s][k];
        G[OFF_W1 + e] = acc;
    }
    if (threadIdx.x < H) {
        float acc = 0.f;
        for (int s = 0; s < TILE; s++) acc += bf2f(S.g1[s * H + threadIdx.x]);
        G[OFF_B1 + threadIdx.x] = acc;
    }
}

struct AdamArgs {
    float* P;             // master
    float* m;
    float* v;
    const float* partial; // [nblocks][NPARAM]
    int nblocks;
    float* grad_out;      // optional: summed gradient
    float lr, b1, b2, eps;
    int64_t t;            // step number (1-based) for bias correction; 0 -> no update
    WB w;
    const float* loss_partial;
    float* loss_out;      // optional: loss of this step
};

// ADAM_PARAMS parameters per block; the THREADS / ADAM_PARAMS threads of a
// parameter each sum every ADAM_SLICES-th block partial (independent loads in
// flight), then the slices add in LDS in a fixed order
constexpr int ADAM_PARAMS = 64, ADAM_SLICES = THREADS / ADAM_PARAMS;

__global__ void __launch_bounds__(THREADS) adam_kernel(AdamArgs A)
{
    __shared__ float red[THREADS];
    const int lane = threadIdx.x % ADAM_PARAMS, slice = threadIdx.x / ADAM_PARAMS;
    const int p = blockIdx.x * ADAM_PARAMS + lane;
    if (A.loss_out && blockIdx.x == 0) {
        float l = 0.f;
        for (int b = threadIdx.x; b < A.nblocks; b += THREADS) l += A.loss_partial[b];
        red[threadIdx.x] = l;
        __syncthreads();
        for (int w = THREADS / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) *A.loss_out = red[0];
        __syncthreads();
    }
    float g = 0.f;
    if (p < NPARAM) {
#pragma unroll 8
        for (int b = slice; b < A.nblocks; b += ADAM_SLICES) g += A.partial[(int64_t)b * NPARAM + p];
    }
    red[threadIdx.x] = g;
    __syncthreads();
    if (slice != 0 || p >= NPARAM) return;
    g = (red[lane] + red[lane + ADAM_PARAMS]) + (red[lane + 2 * ADAM_PARAMS] + red[lane + 3 * ADAM_PARAMS]);
    if (A.grad_out) A.grad_out[p] = g;
    float w = A.P[p];
    if (A.t > 0) {
        float m = A.b1 * A.m[p] + (1.f - A.b1) * g;
        float v = A.b2 * A.v[p] + (1.f - A.b2) * g * g;
        A.m[p] = m;
        A.v[p] = v;
        float mh = m / (1.f - powf(A.b1, (float)A.t));
        float vh = v / (1.f - powf(A.b2, (float)A.t));
        w -= A.lr * mh / (sqrtf(vh) + A.eps);
        A.P[p] = w;
    }
    // refresh the BF16 copies
    unsigned short wb = f2bf(w);
    if (p < OFF_B1) {
        A.w.W1[p - OFF_W1] = wb;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        int e = p - OFF_W2, o = e / H, i = e % H;
        A.w.W2[e] = wb;
        A.w.W2T[i * H + o] = wb;
        A.w.W2lo[e] = f2bf(w - bf2f(wb));
    } else if (p >= OFF_W3 && p < OFF_B3) {
        int e = p - OFF_W3, o = e / H, i = e % H;
        A.w.W3[e] = wb;
        A.w.W3T[i * H + o] = wb;
        A.w.W3lo[e] = f2bf(w - bf2f(wb));
    } else if (p >= OFF_W4 && p < OFF_B4) {
        A.w.W4[p - OFF_W4] = wb;
    }
}

struct InferArgs {
    const double *T, *mu, *xi;
    int64_t n;
    float in_lo[IN], in_scale[IN];   // x_norm = (x - lo) * scale - 1
    float out_mean[OUT], out_std[OUT];
    const float* P;
    WB w;
    double* seeds;        // [n][5]
};

// h (FP32) as hi + lo, both BF16: hi = bf16(h), lo = bf16(h - hi)
__device__ __forceinline__ void split_store(float h, unsigned short* hi, unsigned short* lo, int e)
{
    const unsigned short b = f2bf(h);
    hi[e] = b;
    lo[e] = f2bf(h - bf2f(b));
}

// first layer on the FP32 masters, output split
__device__ void fwd_first_split(Smem& S, const float* W1, const float* b1)
{
    for (int e = threadIdx.x; e < TILE * H; e += THREADS) {
        const int s = e / H, o = e % H;
        float z = b1[o];
#pragma unroll
        for (int k = 0; k < IN; k++) z = fmaf(W1[o * IN + k], S.xs[s][k], z);
        split_store(z * sigmoidf_(z), S.h1, S.g1, e);
    }
}

// hidden layer with split operands: z = in W^T + b as hi*hi + hi*lo + lo*hi
// (the lo*lo term is below FP32 rounding of the sum), output split
__device__ void fwd_hidden_split(const unsigned short* in_hi, const unsigned short* in_lo, const unsigned short* W,
                                 const unsigned short* Wlo, const float* b, unsigned short* out_hi,
                                 unsigned short* out_lo)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (int t = wave; t < (TILE / 16) * (H / 16); t += THREADS / 64) {
        const int ts = t / (H / 16), to = t % (H / 16);
        float4v acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < H; kk += 32) {
            const int ia = (ts * 16 + r) * H + kk + 8 * q, iw = (to * 16 + r) * H + kk + 8 * q;
            const bf16x8 ah = __builtin_bit_cast(bf16x8, ld8(in_hi + ia));
            const bf16x8 al = __builtin_bit_cast(bf16x8, ld8(in_lo + ia));
            const bf16x8 wh = __builtin_bit_cast(bf16x8, ld8(W + iw));
            const bf16x8 wl = __builtin_bit_cast(bf16x8, ld8(Wlo + iw));
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, wh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, wl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, wh, acc, 0, 0, 0);
        }
        const int o = to * 16 + r;
        const float bo = b[o];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = (ts * 16 + 4 * q + i) * H + o;
            const float z = acc[i] + bo;
            split_store(z * sigmoidf_(z), out_hi, out_lo, e);
        }
    }
}

// last layer on the FP32 masters over the split activations
__device__ void fwd_last_split(Smem& S, const float* W4, const float* b4)
{
    for (int e = threadIdx.x; e < TILE * OUT; e += THREADS) {
        const int s = e / OUT, j = e % OUT;
        float z = b4[j];
        for (int o = 0; o < H; o++)
            z = fmaf(W4[j * H + o], bf2f(S.h3[s * H + o]) + bf2f(S.g3[s * H + o]), z);
        S.y[s][j] = z;
    }
}

__global__ void __launch_bounds__(THREADS) infer_kernel(InferArgs A)
{
    extern __shared__ unsigned char smem_raw[];
    Smem& S = *reinterpret_cast<Smem*>(smem_raw);
    const int64_t row0 = (int64_t)blockIdx.x * TILE;
    for (int s = threadIdx.x; s < TILE; s += THREADS) {
        int64_t row = row0 + s;
        bool ok = row < A.n;
        double in[IN] = {ok ? A.T[row] : 0.0, ok ? A.mu[row] : 0.0, ok ? A.xi[row] : 0.0};
        for (int k = 0; k < IN; k++) S.xs[s][k] = ok ? (float)((in[k] - A.in_lo[k]) * A.in_scale[k] - 1.0) : 0.f;
    }
    __syncthreads();
    // split activations: hi in h_l, lo in g_l (unused by the forward)
    fwd_first_split(S, A.P + OFF_W1, A.P + OFF_B1);
    __syncthreads();
    fwd_hidden_split(S.h1, S.g1, A.w.W2, A.w.W2lo, A.P + OFF_B2, S.h2, S.g2);
    __syncthreads();
    fwd_hidden_split(S.h2, S.g2, A.w.W3, A.w.W3lo, A.P + OFF_B3, S.h3, S.g3);
    __syncthreads();
    fwd_last_split(S, A.P + OFF_W4, A.P + OFF_B4);
    __syncthreads();
    for (int s = threadIdx.x; s < TILE; s += THREADS) {
        int64_t row = row0 + s;
        if (row >= A.n) continue;
        double y[OUT];
        for (int j = 0; j < OUT; j++) y[j] = (double)S.y[s][j] * A.out_std[j] + A.out_mean[j];
        double* o = A.seeds + row * 5;
        o[0] = y[0];
        o[1] = y[0];
        o[2] = y[1];
        o[3] = y[2];
        o[4] = y[3];
    }
}

// normalised training data from (T, mu, xi, roots)
__global__ void prep_kernel(const double* T, const double* mu, const double* xi, const double* roots, int64_t n,
                            InferArgs sc, float* xin, float* tgt)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double in[IN] = {T[i], mu[i], xi[i]};
    for (int k = 0; k < IN; k++) xin[i * IN + k] = (float)((in[k] - sc.in_lo[k]) * sc.in_scale[k] - 1.0);
    const double y[OUT] = {roots[i * 5 + 0], roots[i * 5 + 2], roots[i * 5 + 3], roots[i * 5 + 4]};
    for (int j = 0; j < OUT; j++) tgt[i * OUT + j] = (float)((y[j] - sc.out_mean[j]) / sc.out_std[j]);
}

}  // namespace mlp

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct pnjl_surrogate {
    int device = 0;
    std::string err;
    float* P = nullptr;   // master
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    unsigned short* wbuf = nullptr;
    mlp::WB w{};
    float* partial = nullptr;
    int partial_blocks = 0;
    float* loss_partial = nullptr;
    float* loss_dev = nullptr;
    float* data = nullptr;   // xin + tgt
    int64_t data_n = 0;
    double* stage = nullptr; // host-pointer staging (T, mu, xi, roots / seeds)
    size_t stage_bytes = 0;
    int64_t t = 0;           // Adam step counter
    double in_lo[3] = {20.0 / 197.327, 0.0, -0.4}, in_hi[3] = {300.0 / 197.327, 450.0 / 197.327, 0.4};
    double out_mean[4] = {0, 0, 0, 0}, out_std[4] = {1, 1, 1, 1};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
};

namespace {

int serr(pnjl_surrogate* s, int code, const std::string& m)
{
    if (s) s->err = m;
    return code;
}

#define SCHK(s, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return serr(s, PNJL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

mlp::InferArgs scaling(const pnjl_surrogate* s)
{
    mlp::InferArgs a{};
    for (int k = 0; k < 3; k++) {
        a.in_lo[k] = (float)s->in_lo[k];
        a.in_scale[k] = (float)(2.0 / (s->in_hi[k] - s->in_lo[k]));
    }
    for (int j = 0; j < 4; j++) {
        a.out_mean[j] = (float)s->out_mean[j];
        a.out_std[j] = (float)s->out_std[j];
    }
    return a;
}

int refresh_bf16(pnjl_surrogate* s, hipStream_t st)
{
    mlp::AdamArgs a{};
    a.P = s->P;
    a.m = s->m;
    a.v = s->v;
    a.partial = s->partial;
    a.nblocks = 0;
    a.t = 0;
    a.w = s->w;
    hipLaunchKernelGGL(mlp::adam_kernel, dim3((mlp::NPARAM + mlp::ADAM_PARAMS - 1) / mlp::ADAM_PARAMS),
                       dim3(mlp::THREADS), 0, st, a);
    SCHK(s, hipGetLastError());
    return PNJL_OK;
}

int ensure_partial(pnjl_surrogate* s, int nblocks)
{
    if (s->partial_blocks >= nblocks) return PNJL_OK;
    if (s->partial) SCHK(s, hipFree(s->partial));
    if (s->loss_partial) SCHK(s, hipFree(s->loss_partial));
    SCHK(s, hipMalloc(&s->partial, (size_t)nblocks * mlp::NPARAM * sizeof(float)));
    SCHK(s, hipMalloc(&s->loss_partial, (size_t)nblocks * sizeof(float)));
    s->partial_blocks = nblocks;
    return PNJL_OK;
}

int ensure_stage(pnjl_surrogate* s, size_t bytes)
{
    if (s->stage_bytes >= bytes) return PNJL_OK;
    if (s->stage) SCHK(s, hipFree(s->stage));
    s->stage = nullptr;
    SCHK(s, hipMalloc(&s->stage, bytes));
    s->stage_bytes = bytes;
    return PNJL_OK;
}

// runs one forward/backward over `batch` samples and the Adam kernel
int step(pnjl_surrogate* s, int64_t batch, int explicit_rows, uint64_t seed, float lr, bool update, float* loss_dev,
         float* grad_out, hipStream_t st)
{
    const int nblocks = (int)((batch + mlp::TILE - 1) / mlp::TILE);
    int rc = ensure_partial(s, nblocks);
    if (rc) return rc;
    mlp::TrainArgs T{};
    T.xin = s->data;
    T.tgt = s->data + s->data_n * mlp::IN;
    T.n = s->data_n;
    T.batch = batch;
    T.explicit_rows = explicit_rows;
    T.seed = seed;
    T.step = (uint64_t)s->t;
    T.P = s->P;
    T.w = s->w;
    T.partial = s->partial;
    T.loss_partial = s->loss_partial;
    hipLaunchKernelGGL(mlp::train_tile_kernel, dim3(nblocks), dim3(mlp::THREADS), sizeof(mlp::Smem), st, T);
    SCHK(s, hipGetLastError());
    mlp::AdamArgs a{};
    a.P = s->P;
    a.m = s->m;
    a.v = s->v;
    a.partial = s->partial;
    a.nblocks = nblocks;
    a.grad_out = grad_out;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.t = update ? ++s->t : 0;
    a.w = s->w;
    a.loss_partial = s->loss_partial;
    a.loss_out = loss_dev;
    hipLaunchKernelGGL(mlp::adam_kernel, dim3((mlp::NPARAM + mlp::ADAM_PARAMS - 1) / mlp::ADAM_PARAMS),
                       dim3(mlp::THREADS), 0, st, a);
    SCHK(s, hipGetLastError());
    return PNJL_OK;
}

int load_data(pnjl_surrogate* s, int64_t n, const double* T, const double* mu, const double* xi, const double* roots,
              int flags, hipStream_t st)
{
    if (s->data_n < n) {
        if (s->data) SCHK(s, hipFree(s->data));
        s->data = nullptr;
        SCHK(s, hipMalloc(&s->data, (size_t)n * (mlp::IN + mlp::OUT) * sizeof(float)));
    }
    s->data_n = n;
    const double *dT = T, *dmu = mu, *dxi = xi, *dr = roots;
    if (!(flags & PNJL_FLAG_DEVICE_PTRS)) {
        int rc = ensure_stage(s, (size_t)n * 8 * sizeof(double));
        if (rc) return rc;
        double* b = s->stage;
        SCHK(s, hipMemcpyAsync(b, T, n * 8, hipMemcpyHostToDevice, st));
        SCHK(s, hipMemcpyAsync(b + n, mu, n * 8, hipMemcpyHostToDevice, st));
        SCHK(s, hipMemcpyAsync(b + 2 * n, xi, n * 8, hipMemcpyHostToDevice, st));
        SCHK(s, hipMemcpyAsync(b + 3 * n, roots, n * 40, hipMemcpyHostToDevice, st));
        dT = b;
        dmu = b + n;
        dxi = b + 2 * n;
        dr = b + 3 * n;
    }
    mlp::InferArgs sc = scaling(s);
    hipLaunchKernelGGL(mlp::prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dT, dmu, dxi, dr, n, sc,
                       s->data, s->data + n * mlp::IN);
    SCHK(s, hipGetLastError());
    return PNJL_OK;
}

}  // namespace

extern "C" {

int pnjl_surrogate_param_count(void) { return mlp::NPARAM; }

int pnjl_surrogate_create(int device, uint64_t init_seed, pnjl_surrogate** out)
{
    if (!out) return PNJL_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PNJL_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return PNJL_ERR_ARG;
    pnjl_surrogate* s = new pnjl_surrogate();
    s->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete s;
        return PNJL_ERR_HIP;
    }
    const size_t pb = mlp::NPARAM * sizeof(float);
    const size_t nb = (size_t)(mlp::H * mlp::IN + 6 * mlp::H * mlp::H + mlp::OUT * mlp::H);
    if (hipMalloc(&s->P, pb) != hipSuccess || hipMalloc(&s->m, pb) != hipSuccess || hipMalloc(&s->v, pb) != hipSuccess ||
        hipMalloc(&s->grad, pb) != hipSuccess || hipMalloc(&s->wbuf, nb * 2) != hipSuccess ||
        hipMalloc(&s->loss_dev, sizeof(float)) != hipSuccess || hipEventCreate(&s->e0) != hipSuccess ||
        hipEventCreate(&s->e1) != hipSuccess) {
        pnjl_surrogate_destroy(s);
        return PNJL_ERR_HIP;
    }
    s->w.W1 = s->wbuf;
    s->w.W2 = s->w.W1 + mlp::H * mlp::IN;
    s->w.W2T = s->w.W2 + mlp::H * mlp::H;
    s->w.W3 = s->w.W2T + mlp::H * mlp::H;
    s->w.W3T = s->w.W3 + mlp::H * mlp::H;
    s->w.W4 = s->w.W3T + mlp::H * mlp::H;
    s->w.W2lo = s->w.W4 + mlp::OUT * mlp::H;
    s->w.W3lo = s->w.W2lo + mlp::H * mlp::H;
    // He-uniform initialisation of the weights, zero biases (splitmix64 stream)
    std::vector<float> h(mlp::NPARAM, 0.f);
    uint64_t st = init_seed;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    auto fill = [&](int off, int fan_out, int fan_in) {
        double a = sqrt(6.0 / fan_in);
        for (int i = 0; i < fan_out * fan_in; i++) h[off + i] = (float)(a * uni());
    };
    fill(mlp::OFF_W1, mlp::H, mlp::IN);
    fill(mlp::OFF_W2, mlp::H, mlp::H);
    fill(mlp::OFF_W3, mlp::H, mlp::H);
    fill(mlp::OFF_W4, mlp::OUT, mlp::H);
    if (hipMemcpy(s->P, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess || hipMemset(s->m, 0, pb) != hipSuccess ||
        hipMemset(s->v, 0, pb) != hipSuccess || refresh_bf16(s, nullptr) != PNJL_OK ||
        hipDeviceSynchronize() != hipSuccess) {
        pnjl_surrogate_destroy(s);
        return PNJL_ERR_HIP;
    }
    *out = s;
    return PNJL_OK;
}

void pnjl_surrogate_destroy(pnjl_surrogate* s)
{
    if (!s) return;
    // teardown: errors cannot be reported any more and are ignored
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    for (void* p : {(void*)s->P, (void*)s->m, (void*)s->v, (void*)s->grad, (void*)s->wbuf, (void*)s->partial,
                    (void*)s->loss_partial, (void*)s->loss_dev, (void*)s->data, (void*)s->stage})
        if (p) (void)hipFree(p);
    if (s->e0) (void)hipEventDestroy(s->e0);
    if (s->e1) (void)hipEventDestroy(s->e1);
    delete s;
}

const char* pnjl_surrogate_last_error(const pnjl_surrogate* s) { return s ? s->err.c_str() : "no surrogate"; }

double pnjl_surrogate_last_ms(const pnjl_surrogate* s) { return s ? s->last_ms : 0.0; }

int pnjl_surrogate_set_scaling(pnjl_surrogate* s, const double* in_lo, const double* in_hi, const double* out_mean,
                               const double* out_std)
{
    if (!s || !in_lo || !in_hi || !out_mean || !out_std) return PNJL_ERR_ARG;
    for (int k = 0; k < 3; k++) {
        if (!(in_hi[k] > in_lo[k])) return serr(s, PNJL_ERR_ARG, "set_scaling: empty input range");
        s->in_lo[k] = in_lo[k];
        s->in_hi[k] = in_hi[k];
    }
    for (int j = 0; j < 4; j++) {
        if (!(out_std[j] > 0)) return serr(s, PNJL_ERR_ARG, "set_scaling: non-positive output scale");
        s->out_mean[j] = out_mean[j];
        s->out_std[j] = out_std[j];
    }
    return PNJL_OK;
}

int pnjl_surrogate_get_params(pnjl_surrogate* s, float* out)
{
    if (!s || !out) return PNJL_ERR_ARG;
    SCHK(s, hipSetDevice(s->device));
    SCHK(s, hipMemcpy(out, s->P, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_surrogate_set_params(pnjl_surrogate* s, const float* in)
{
    if (!s || !in) return PNJL_ERR_ARG;
    SCHK(s, hipSetDevice(s->device));
    SCHK(s, hipMemcpy(s->P, in, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice));
    int rc = refresh_bf16(s, nullptr);
    if (rc) return rc;
    SCHK(s, hipDeviceSynchronize());
    return PNJL_OK;
}

int pnjl_surrogate_train(pnjl_surrogate* s, int64_t n, const double* T, const double* mu, const double* xi,
                         const double* roots, int steps, int batch, float lr, uint64_t seed, float* loss_out, int flags,
                         void* stream)
{
    if (!s || n <= 0 || !T || !mu || !xi || !roots || steps < 0 || batch <= 0 || batch > (1 << 22))
        return serr(s, PNJL_ERR_ARG, "pnjl_surrogate_train: bad argument");
    SCHK(s, hipSetDevice(s->device));
    hipStream_t st = (hipStream_t)stream;
    int rc = load_data(s, n, T, mu, xi, roots, flags, st);
    if (rc) return rc;
    float* losses = nullptr;
    if (loss_out) SCHK(s, hipMalloc(&losses, (size_t)std::max(steps, 1) * sizeof(float)));
    SCHK(s, hipEventRecord(s->e0, st));
    for (int k = 0; k < steps; k++) {
        rc = step(s, batch, 0, seed, lr, true, losses ? losses + k : nullptr, nullptr, st);
        if (rc) return rc;
    }
    SCHK(s, hipEventRecord(s->e1, st));
    SCHK(s, hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s->e0, s->e1) == hipSuccess) s->last_ms = ms;
    if (losses) {
        SCHK(s, hipMemcpy(loss_out, losses, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
        SCHK(s, hipFree(losses));
    }
    return PNJL_OK;
}

int pnjl_surrogate_loss_grad(pnjl_surrogate* s, int64_t n, const double* T, const double* mu, const double* xi,
                             const double* roots, float* loss, float* grad)
{
    if (!s || n <= 0 || n > (1 << 20) || !T || !mu || !xi || !roots || !loss || !grad)
        return serr(s, PNJL_ERR_ARG, "pnjl_surrogate_loss_grad: bad argument");
    SCHK(s, hipSetDevice(s->device));
    int rc = load_data(s, n, T, mu, xi, roots, 0, nullptr);
    if (rc) return rc;
    rc = step(s, n, 1, 0, 0.f, false, s->loss_dev, s->grad, nullptr);
    if (rc) return rc;
    SCHK(s, hipDeviceSynchronize());
    SCHK(s, hipMemcpy(loss, s->loss_dev, sizeof(float), hipMemcpyDeviceToHost));
    SCHK(s, hipMemcpy(grad, s->grad, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return PNJL_OK;
}

int pnjl_surrogate_infer(pnjl_surrogate* s, int64_t n, const double* T, const double* mu, const double* xi,
                         double* seeds, int flags, void* stream)
{
    if (!s || n < 0 || !T || !mu || !xi || !seeds) return serr(s, PNJL_ERR_ARG, "pnjl_surrogate_infer: bad argument");
    if (n == 0) return PNJL_OK;
    SCHK(s, hipSetDevice(s->device));
    hipStream_t st = (hipStream_t)stream;
    const double *dT = T, *dmu = mu, *dxi = xi;
    double* dseed = seeds;
    if (!(flags & PNJL_FLAG_DEVICE_PTRS)) {
        int rc = ensure_stage(s, (size_t)n * 8 * sizeof(double));
        if (rc) return rc;
        double* b = s->stage;
        SCHK(s, hipMemcpyAsync(b, T, n * 8, hipMemcpyHostToDevice, st));
        SCHK(s, hipMemcpyAsync(b + n, mu, n * 8, hipMemcpyHostToDevice, st));
        SCHK(s, hipMemcpyAsync(b + 2 * n, xi, n * 8, hipMemcpyHostToDevice, st));
        dT = b;
        dmu = b + n;
        dxi = b + 2 * n;
        dseed = b + 3 * n;
    }
    mlp::InferArgs a = scaling(s);
    a.T = dT;
    a.mu = dmu;
    a.xi = dxi;
    a.n = n;
    a.P = s->P;
    a.w = s->w;
    a.seeds = dseed;
    SCHK(s, hipEventRecord(s->e0, st));
    hipLaunchKernelGGL(mlp::infer_kernel, dim3((unsigned)((n + mlp::TILE - 1) / mlp::TILE)), dim3(mlp::THREADS),
                       sizeof(mlp::Smem), st, a);
    SCHK(s, hipGetLastError());
    SCHK(s, hipEventRecord(s->e1, st));
    if (!(flags & PNJL_FLAG_DEVICE_PTRS)) SCHK(s, hipMemcpyAsync(seeds, dseed, n * 40, hipMemcpyDeviceToHost, st));
    if (!(flags & PNJL_FLAG_ASYNC) || !(flags & PNJL_FLAG_DEVICE_PTRS)) {
        SCHK(s, hipStreamSynchronize(st));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->e0, s->e1) == hipSuccess) s->last_ms = ms;
    }
    return PNJL_OK;
}

}  // extern "C"
