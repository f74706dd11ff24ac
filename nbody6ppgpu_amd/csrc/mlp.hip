// mlp.hip -- the learned neighbour-radius predictor: a small MLP
// (8 -> 64 -> 64 -> 1, SiLU) that maps a particle's local state to log RS.
// Inference and training run on the GPU with these kernels; the plain
// Ahmad-Cohen radius update stays the fallback in the caller.
//
// Arithmetic: weights, activations and back-propagated deltas enter the
// products as BF16 (round to nearest even); products are accumulated in
// FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16).  Bias additions, SiLU
// and the bias / output-layer gradients are FP32 vector work.  Training keeps
// FP32 master weights, FP32 gradients and an FP32 Adam state; the BF16
// weights are re-derived from the masters after every step.
//
// MFMA mapping.  A wavefront evaluates 16 samples.  Every layer is computed
// transposed, Z^T = W H^T: the weights are the A operand (16 output units x
// 16 inputs per instruction), the activations the B operand (16 inputs x 16
// samples).  The 16x16 FP32 result tile leaves lane l holding sample l%16 and
// units 4(l/16)..4(l/16)+3 of the tile, which is exactly the B-operand layout
// of the next layer's k-step, so activations never leave registers.  The
// backward delta d1^T = W2^T d2^T reuses the same layout with W2^T as A.
// Weight gradients (sums over samples) need the samples on the k dimension:
// the wave stages its deltas and activations through LDS and reads them back
// transposed.
//
// Kernels
//   k_mlp_forward   inference; 4 waves x 16 samples per workgroup, loop.
//   k_mlp_fwdbwd    forward + backward, per-wave gradient sums accumulated
//                   in registers over a workgroup-strided loop, one partial
//                   row per wave.
//   k_mlp_reduce    sums the partial rows in a fixed order (deterministic)
//                   and applies Adam (or only stores the sum).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpunb.h"
#include "gpunb_internal.h"

namespace gpunb_mlp {

constexpr int NI = 8, NH = 64;
// Parameter layout (FP32 master, BF16 copy): W1[NH][NI] b1[NH] W2[NH][NH]
// b2[NH] W3[NH] b3.
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH,
              OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH,
              NP = OB3 + 1;
constexpr int kWaves = 4;               // waves per workgroup
constexpr int kGroup = 16 * kWaves;     // samples per workgroup iteration
constexpr int kMaxBlocks = 256;         // training workgroups (partial rows / 4)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t bf16_bits(float x) {
    // Round to nearest even; upper 16 bits of the FP32 pattern.
    uint32_t u;
    memcpy(&u, &x, 4);
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf16_val(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

struct Norm {
    float xm[NI], xs[NI], ym, ys;
};

// Per-lane weights of one wave.  Lane l (r = l%16, q = l/16):
//   w1[T]      A operand of layer 1, output tile T: W1[16T+r][4q..4q+3]
//              (inputs 8..15 are zero padding, q >= 2);
//   w2[T][t]   A operand of layer 2: W2[16T+r][16t+4q..+3];
//   b1/b2/w3   FP32 values for this lane's units u = 16T+4q+v (result layout).
struct FwdWeights {
    bf16x4 w1[4];
    bf16x4 w2[4][4];
    float b1[4][4], b2[4][4], w3[4][4], b3;
};

__device__ inline void load_fwd(const uint16_t *__restrict__ wh, int l,
                                FwdWeights &W) {
    const int r = l & 15, q = l >> 4;
#pragma unroll
    for (int T = 0; T < 4; ++T) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            W.w1[T][j] = q < 2 ? (short)wh[OW1 + (16 * T + r) * NI + 4 * q + j] : 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                W.w2[T][t][j] = (short)wh[OW2 + (16 * T + r) * NH + 16 * t + 4 * q + j];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int u = 16 * T + 4 * q + v;
            W.b1[T][v] = bf16_val(wh[OB1 + u]);
            W.b2[T][v] = bf16_val(wh[OB2 + u]);
            W.w3[T][v] = bf16_val(wh[OW3 + u]);
        }
    }
    W.b3 = bf16_val(wh[OB3]);
}

// Layer-1 B operand of sample `sidx` (lanes q < 2 hold inputs 4q..4q+3).
__device__ inline bf16x4 input_operand(const Norm &nm, const float *__restrict__ x,
                                       int sidx, bool valid, int q, float *xh) {
    bf16x4 b = {0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 4; ++v) xh[v] = 0.0f;
    if (valid && q < 2) {
        const float4 xv = *reinterpret_cast<const float4 *>(x + (size_t)NI * sidx + 4 * q);
        const float xx[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint16_t h = bf16_bits((xx[v] - nm.xm[4 * q + v]) / nm.xs[4 * q + v]);
            b[v] = (short)h;
            xh[v] = bf16_val(h);
        }
    }
    return b;
}

// Forward of the wave's 16 samples.  Outputs (result layout, lane = sample
// l%16, units 16T+4q+v): h1 as the layer-2 B operands, silu'(z1), h2,
// silu'(z2); returns the network output for sample l%16 (in every lane).
__device__ inline float forward16(const FwdWeights &W, bf16x4 xb, bf16x4 *h1b,
                                  float (*ds1)[4], float (*h2)[4],
                                  float (*ds2)[4]) {
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        const f32x4 z = mfma16(W.w1[T], xb, zero);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float zz = z[v] + W.b1[T][v];
            h1b[T][v] = (short)bf16_bits(silu(zz));
            if (ds1) ds1[T][v] = dsilu(zz);
        }
    }
    float part = 0.0f;
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        f32x4 z = zero;
#pragma unroll
        for (int t = 0; t < 4; ++t) z = mfma16(W.w2[T][t], h1b[t], z);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float zz = z[v] + W.b2[T][v];
            const float h = bf16_val(bf16_bits(silu(zz)));
            if (h2) h2[T][v] = h;
            if (ds2) ds2[T][v] = dsilu(zz);
            part = fmaf(W.w3[T][v], h, part);
        }
    }
    // Sum over the four lanes holding this sample (q = 0..3).
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + W.b3;
}

__global__ __launch_bounds__(256) void k_mlp_forward(
    int n, const float *__restrict__ x, const uint16_t *__restrict__ wh,
    Norm nm, float *__restrict__ y) {
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = l & 15, q = l >> 4;
    FwdWeights W;
    load_fwd(wh, l, W);
    const int ngroups = (n + 15) / 16;
    for (int g = blockIdx.x * kWaves + wave; g < ngroups; g += gridDim.x * kWaves) {
        const int sidx = 16 * g + r;
        const bool valid = sidx < n;
        float xh[4];
        const bf16x4 xb = input_operand(nm, x, sidx, valid, q, xh);
        bf16x4 h1b[4];
        const float out = forward16(W, xb, h1b, nullptr, nullptr, nullptr);
        if (valid && q == 0) y[sidx] = out * nm.ys + nm.ym;
    }
}

// Forward + backward of the samples idx[base .. base+count) (idx may be null:
// identity).  Loss = mean over the batch of (yhat - t_normalised)^2.  The
// derivative of the BF16 rounding is taken as 1 (straight-through).  Deltas
// d2 and d1 enter the MFMA products (d1 = W2^T d2, the weight gradients) as
// BF16; the bias and output-layer gradients use them in FP32.  Each wave
// accumulates its gradient sums over its samples and writes them, with its
// summed squared error, to row blockIdx.x*kWaves + wave of gpart / lpart.
constexpr int LDS_ROW = NH + 8;   // bf16 per sample row (16-byte padding)
__global__ __launch_bounds__(256) void k_mlp_fwdbwd(
    int count, int base, const int *__restrict__ idx,
    const float *__restrict__ x, const float *__restrict__ t,
    const uint16_t *__restrict__ wh, Norm nm, float inv_batch,
    float *__restrict__ gpart, float *__restrict__ lpart) {
    __shared__ uint16_t sH1[kWaves][16 * LDS_ROW];
    __shared__ uint16_t sD2[kWaves][16 * LDS_ROW];
    __shared__ uint16_t sD1[kWaves][16 * LDS_ROW];
    __shared__ uint16_t sXH[kWaves][16 * 16];
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = l & 15, q = l >> 4;
    FwdWeights W;
    load_fwd(wh, l, W);
    // A operand of d1^T = W2^T d2^T: lane holds W2[16t+4q+j][16T+r].
    bf16x4 w2t[4][4];
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                w2t[T][tt][j] = (short)wh[OW2 + (16 * tt + 4 * q + j) * NH + 16 * T + r];
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 gw2[4][4], gw1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.0f, loss = 0.0f;
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        gw1[T] = zero;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            gw2[T][v] = zero;
            gb1[T][v] = gb2[T][v] = gw3[T][v] = 0.0f;
        }
    }
    uint16_t *h1s = sH1[wave], *d2s = sD2[wave], *d1s = sD1[wave], *xs = sXH[wave];
    const int nbg = (count + kGroup - 1) / kGroup;
    // Uniform trip count across the workgroup (barriers inside the loop).
    for (int bg = blockIdx.x; bg < nbg; bg += gridDim.x) {
        const int qi = bg * kGroup + 16 * wave + r;
        const bool valid = qi < count;
        const int sidx = valid ? (idx ? idx[base + qi] : base + qi) : 0;
        float xh[4];
        const bf16x4 xb = input_operand(nm, x, sidx, valid, q, xh);
        bf16x4 h1b[4];
        float ds1[4][4], h2[4][4], ds2[4][4];
        const float out = forward16(W, xb, h1b, ds1, h2, ds2);
        float d3 = 0.0f;
        if (valid) {
            const float e = out - (t[sidx] - nm.ym) / nm.ys;
            d3 = 2.0f * e * inv_batch;
            if (q == 0) loss += e * e;
        }
        if (q == 0) gb3 += d3;
        // d2 (result layout of layer 2 = B operand of the d1 product).
        bf16x4 d2b[4];
#pragma unroll
        for (int T = 0; T < 4; ++T)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float d2 = d3 * W.w3[T][v] * ds2[T][v];
                d2b[T][v] = (short)bf16_bits(d2);
                gb2[T][v] += d2;
                gw3[T][v] = fmaf(d3, h2[T][v], gw3[T][v]);
            }
        // d1^T = W2^T d2^T, then times silu'(z1).
#pragma unroll
        for (int T = 0; T < 4; ++T) {
            f32x4 z = zero;
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) z = mfma16(w2t[T][tt], d2b[tt], z);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float d1 = z[v] * ds1[T][v];
                gb1[T][v] += d1;
                const int u = 16 * T + 4 * q + v;
                d1s[r * LDS_ROW + u] = bf16_bits(d1);
                d2s[r * LDS_ROW + u] = (uint16_t)d2b[T][v];
                h1s[r * LDS_ROW + u] = (uint16_t)h1b[T][v];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
            xs[r * 16 + 4 * q + v] = q < 2 ? bf16_bits(xh[v]) : 0;
        __syncthreads();
        // Weight gradients: samples on the k dimension.
        //   gW2[b][a] += sum_s d2[s][b] h1[s][a]:  A[b][s], B[s][a]
        //   gW1[a][k] += sum_s d1[s][a] xh[s][k]:  A[a][s], B[s][k]
        bf16x4 ad2[4], ad1[4], bh1[4], bx;
#pragma unroll
        for (int T = 0; T < 4; ++T)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = (4 * q + j) * LDS_ROW + 16 * T + r;
                ad2[T][j] = (short)d2s[row];
                ad1[T][j] = (short)d1s[row];
                bh1[T][j] = (short)h1s[row];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) bx[j] = (short)xs[(4 * q + j) * 16 + r];
#pragma unroll
        for (int T = 0; T < 4; ++T) {
#pragma unroll
            for (int T2 = 0; T2 < 4; ++T2) gw2[T][T2] = mfma16(ad2[T], bh1[T2], gw2[T][T2]);
            gw1[T] = mfma16(ad1[T], bx, gw1[T]);
        }
        __syncthreads();
    }
    // Partial row of this wave.
    float *g = gpart + (size_t)(blockIdx.x * kWaves + wave) * NP;
#pragma unroll
    for (int T = 0; T < 4; ++T)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int u = 16 * T + 4 * q + v;   // row unit of the result tiles
#pragma unroll
            for (int T2 = 0; T2 < 4; ++T2) g[OW2 + u * NH + 16 * T2 + r] = gw2[T][T2][v];
            if (r < NI) g[OW1 + u * NI + r] = gw1[T][v];
            // Bias / output-layer sums over the 16 samples of this lane group.
            float a1 = gb1[T][v], a2 = gb2[T][v], a3 = gw3[T][v];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a1 += __shfl_xor(a1, o);
                a2 += __shfl_xor(a2, o);
                a3 += __shfl_xor(a3, o);
            }
            if (r == 0) {
                g[OB1 + u] = a1;
                g[OB2 + u] = a2;
                g[OW3 + u] = a3;
            }
        }
    for (int o = 32; o >= 1; o >>= 1) {
        gb3 += __shfl_xor(gb3, o);
        loss += __shfl_xor(loss, o);
    }
    if (l == 0) {
        g[OB3] = gb3;
        lpart[blockIdx.x * kWaves + wave] = loss;
    }
}

// Sum the nrow partial rows in row order; with adam != 0 apply one Adam step
// to the FP32 masters and refresh the BF16 copy.
__global__ __launch_bounds__(256) void k_mlp_reduce(
    int nrow, const float *__restrict__ gpart, float *__restrict__ gsum,
    int adam, float lr, float b1, float b2, float eps, float c1, float c2,
    float *__restrict__ w, uint16_t *__restrict__ wh, float *__restrict__ m,
    float *__restrict__ v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NP) return;
    float g = 0.0f;
    for (int b = 0; b < nrow; ++b) g += gpart[(size_t)b * NP + k];
    gsum[k] = g;
    if (adam) {
        const float mk = b1 * m[k] + (1.0f - b1) * g;
        const float vk = b2 * v[k] + (1.0f - b2) * g * g;
        m[k] = mk;
        v[k] = vk;
        const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        w[k] = wk;
        wh[k] = bf16_bits(wk);
    }
}

__global__ void k_bf16_copy(const float *__restrict__ w, uint16_t *__restrict__ wh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NP) wh[k] = bf16_bits(w[k]);
}

// ---------------------------------------------------------------------------
// Host state and C-ABI
// ---------------------------------------------------------------------------

struct State {
    bool init = false;
    hipStream_t st = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wh = nullptr;      // BF16 copy of the parameters
    float *gsum = nullptr;
    Norm nm{};
    long long adam_t = 0;
    // device copies of data
    float *dx = nullptr, *dt = nullptr, *dy = nullptr, *gpart = nullptr,
          *lpart = nullptr;
    int *didx = nullptr;
    size_t cap_n = 0, cap_rows = 0;
};
State M;

#define MCHK(e)                                                               \
    do {                                                                      \
        hipError_t r_ = (e);                                                  \
        if (r_ != hipSuccess)                                                 \
            throw std::runtime_error(std::string(#e) + ": " +                 \
                                     hipGetErrorString(r_));                  \
    } while (0)

void ensure_init() {
    if (M.init) return;
    MCHK(hipStreamCreateWithFlags(&M.st, hipStreamNonBlocking));
    MCHK(hipMalloc((void **)&M.w, NP * sizeof(float)));
    MCHK(hipMalloc((void **)&M.wh, NP * sizeof(uint16_t)));
    MCHK(hipMalloc((void **)&M.m, NP * sizeof(float)));
    MCHK(hipMalloc((void **)&M.v, NP * sizeof(float)));
    MCHK(hipMalloc((void **)&M.gsum, NP * sizeof(float)));
    for (int k = 0; k < NI; ++k) {
        M.nm.xm[k] = 0.0f;
        M.nm.xs[k] = 1.0f;
    }
    M.nm.ym = 0.0f;
    M.nm.ys = 1.0f;
    M.init = true;
}

void ensure_data(size_t n, size_t nrows) {
    if (n > M.cap_n) {
        if (M.dx) (void)hipFree(M.dx);
        if (M.dt) (void)hipFree(M.dt);
        if (M.dy) (void)hipFree(M.dy);
        if (M.didx) (void)hipFree(M.didx);
        MCHK(hipMalloc((void **)&M.dx, n * NI * sizeof(float)));
        MCHK(hipMalloc((void **)&M.dt, n * sizeof(float)));
        MCHK(hipMalloc((void **)&M.dy, n * sizeof(float)));
        MCHK(hipMalloc((void **)&M.didx, n * sizeof(int)));
        M.cap_n = n;
    }
    if (nrows > M.cap_rows) {
        if (M.gpart) (void)hipFree(M.gpart);
        if (M.lpart) (void)hipFree(M.lpart);
        MCHK(hipMalloc((void **)&M.gpart, nrows * NP * sizeof(float)));
        MCHK(hipMalloc((void **)&M.lpart, nrows * sizeof(float)));
        M.cap_rows = nrows;
    }
}

void set_params(const float *p) {
    ensure_init();
    MCHK(hipMemcpyAsync(M.w, p, NP * sizeof(float), hipMemcpyHostToDevice, M.st));
    hipLaunchKernelGGL(k_bf16_copy, dim3((NP + 255) / 256), dim3(256), 0, M.st,
                       M.w, M.wh);
    MCHK(hipMemsetAsync(M.m, 0, NP * sizeof(float), M.st));
    MCHK(hipMemsetAsync(M.v, 0, NP * sizeof(float), M.st));
    MCHK(hipStreamSynchronize(M.st));
    M.adam_t = 0;
}

// One pass of forward+backward over count samples starting at base (through
// idx if given); gradient (mean-square loss) summed into M.gsum, optionally
// followed by an Adam step.  Returns the summed squared error.
double grad_step(int count, int base, const int *didx, float lr, int adam) {
    const int nblk = std::min((count + kGroup - 1) / kGroup, kMaxBlocks);
    const int nrows = nblk * kWaves;
    ensure_data(M.cap_n, (size_t)nrows);
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(nblk), dim3(64 * kWaves), 0, M.st,
                       count, base, didx, M.dx, M.dt, M.wh, M.nm,
                       1.0f / count, M.gpart, M.lpart);
    float c1 = 1.0f, c2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        M.adam_t++;
        c1 = 1.0f - std::pow(b1, (float)M.adam_t);
        c2 = 1.0f - std::pow(b2, (float)M.adam_t);
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0,
                       M.st, nrows, M.gpart, M.gsum, adam, lr, b1, b2, 1e-8f,
                       c1, c2, M.w, M.wh, M.m, M.v);
    MCHK(hipGetLastError());
    std::vector<float> lp(nrows);
    MCHK(hipMemcpyAsync(lp.data(), M.lpart, nrows * sizeof(float),
                        hipMemcpyDeviceToHost, M.st));
    MCHK(hipStreamSynchronize(M.st));
    double l = 0.0;
    for (float q : lp) l += q;
    return l;
}

void upload_data(int n, const float *x, const float *t) {
    ensure_init();
    ensure_data((size_t)n, 1);
    MCHK(hipMemcpyAsync(M.dx, x, (size_t)n * NI * sizeof(float),
                        hipMemcpyHostToDevice, M.st));
    if (t)
        MCHK(hipMemcpyAsync(M.dt, t, (size_t)n * sizeof(float),
                            hipMemcpyHostToDevice, M.st));
}

}  // namespace gpunb_mlp

using namespace gpunb_mlp;

#define MLP_EXPORT __attribute__((visibility("default")))

extern "C" {

MLP_EXPORT int gpunb_mlp_nparams(void) { return NP; }

MLP_EXPORT int gpunb_mlp_init(unsigned seed) {
    return gpunb::run_guarded("gpunb_mlp_init", [&] {
        std::mt19937 rng(seed);
        std::vector<float> p(NP, 0.0f);
        auto fill = [&](int off, int fan_in, int fan_out, int count) {
            const float a = std::sqrt(6.0f / (fan_in + fan_out));
            std::uniform_real_distribution<float> u(-a, a);
            for (int k = 0; k < count; ++k) p[off + k] = u(rng);
        };
        fill(OW1, NI, NH, NH * NI);
        fill(OW2, NH, NH, NH * NH);
        fill(OW3, NH, 1, NH);
        set_params(p.data());
    });
}

MLP_EXPORT int gpunb_mlp_set_norm(const float *xmean, const float *xstd,
                                  float ymean, float ystd) {
    return gpunb::run_guarded("gpunb_mlp_set_norm", [&] {
        ensure_init();
        for (int k = 0; k < NI; ++k) {
            if (!(xstd[k] > 0.0f)) throw std::runtime_error("xstd must be > 0");
            M.nm.xm[k] = xmean[k];
            M.nm.xs[k] = xstd[k];
        }
        if (!(ystd > 0.0f)) throw std::runtime_error("ystd must be > 0");
        M.nm.ym = ymean;
        M.nm.ys = ystd;
    });
}

MLP_EXPORT int gpunb_mlp_get_params(float *p) {
    return gpunb::run_guarded("gpunb_mlp_get_params", [&] {
        ensure_init();
        MCHK(hipMemcpyAsync(p, M.w, NP * sizeof(float), hipMemcpyDeviceToHost,
                            M.st));
        MCHK(hipStreamSynchronize(M.st));
    });
}

MLP_EXPORT int gpunb_mlp_set_params(const float *p) {
    return gpunb::run_guarded("gpunb_mlp_set_params", [&] { set_params(p); });
}

MLP_EXPORT int gpunb_mlp_predict(int n, const float *x, float *y) {
    return gpunb::run_guarded("gpunb_mlp_predict", [&] {
        if (n <= 0) return;
        upload_data(n, x, nullptr);
        const int ngroups = (n + 15) / 16;
        const int nblk = std::min((ngroups + kWaves - 1) / kWaves, 2048);
        hipLaunchKernelGGL(k_mlp_forward, dim3(nblk), dim3(64 * kWaves), 0,
                           M.st, n, M.dx, M.wh, M.nm, M.dy);
        MCHK(hipGetLastError());
        MCHK(hipMemcpyAsync(y, M.dy, (size_t)n * sizeof(float),
                            hipMemcpyDeviceToHost, M.st));
        MCHK(hipStreamSynchronize(M.st));
    });
}

MLP_EXPORT int gpunb_mlp_grad(int n, const float *x, const float *t,
                              float *grad, double *loss) {
    return gpunb::run_guarded("gpunb_mlp_grad", [&] {
        if (n <= 0) throw std::runtime_error("n must be > 0");
        upload_data(n, x, t);
        const double l = grad_step(n, 0, nullptr, 0.0f, 0);
        MCHK(hipMemcpyAsync(grad, M.gsum, NP * sizeof(float),
                            hipMemcpyDeviceToHost, M.st));
        MCHK(hipStreamSynchronize(M.st));
        if (loss) *loss = l / n;
    });
}

MLP_EXPORT int gpunb_mlp_train(int n, const float *x, const float *t,
                               int epochs, int batch, float lr, unsigned seed,
                               double *loss_out) {
    return gpunb::run_guarded("gpunb_mlp_train", [&] {
        if (n <= 0 || epochs <= 0 || batch <= 0)
            throw std::runtime_error("n, epochs and batch must be > 0");
        upload_data(n, x, t);
        std::vector<int> idx(n);
        for (int k = 0; k < n; ++k) idx[k] = k;
        std::mt19937 rng(seed);
        double last = 0.0;
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(idx.begin(), idx.end(), rng);
            MCHK(hipMemcpyAsync(M.didx, idx.data(), (size_t)n * sizeof(int),
                                hipMemcpyHostToDevice, M.st));
            double sum = 0.0;
            for (int b0 = 0; b0 < n; b0 += batch) {
                const int cnt = std::min(batch, n - b0);
                sum += grad_step(cnt, b0, M.didx, lr, 1);
            }
            last = sum / n;
        }
        if (loss_out) *loss_out = last;
    });
}

}  // extern "C"
