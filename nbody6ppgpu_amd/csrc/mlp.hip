// mlp.hip -- the learned neighbour-radius predictor: a small MLP
// (8 -> 64 -> 64 -> 1, SiLU) that maps a particle's local state to log RS.
// Inference and training run on the GPU with these kernels; the plain
// Ahmad-Cohen radius update stays the fallback in the caller.
//
// Arithmetic: weights and activations are rounded to BF16 (round to nearest
// even) before every multiply; products are accumulated in FP32.  Training
// keeps FP32 master weights, FP32 gradients and an FP32 Adam state; the BF16
// weights are re-derived from the masters after every step.
//
// Kernels
//   k_mlp_forward   one sample per thread, weights in LDS (BF16).
//   k_mlp_fwdbwd    forward + backward of 64 samples per workgroup, per-block
//                   gradient sums (activations and deltas staged in LDS).
//   k_mlp_reduce    sums the per-block gradients in a fixed order
//                   (deterministic) and applies Adam (or only stores the sum).
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
constexpr int BS = 64;  // samples per training workgroup

__device__ __host__ inline float bf16r(float x) {
    // Round to nearest even to BF16 and back to FP32.
    uint32_t u;
    memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

struct Norm {
    float xm[NI], xs[NI], ym, ys;
};
static_assert(BS == NH, "k_mlp_fwdbwd maps one lane to one hidden unit");

// Weights as BF16-rounded FP32 in LDS (exact BF16 values, FP32 storage).
__device__ inline void load_weights(const float *__restrict__ wb, float *sw) {
    for (int k = threadIdx.x; k < NP; k += blockDim.x) sw[k] = wb[k];
    __syncthreads();
}

// Layer 1 of one sample: h1 = bf16(silu(W1 xh + b1)); ds1 = silu'(z1).
__device__ inline void layer1(const float *sw, const float *xh, float *h1,
                              float *ds1) {
#pragma unroll
    for (int a = 0; a < NH; ++a) {
        float acc = sw[OB1 + a];
#pragma unroll
        for (int k = 0; k < NI; ++k) acc = fmaf(sw[OW1 + a * NI + k], xh[k], acc);
        h1[a] = bf16r(silu(acc));
        if (ds1) ds1[a] = dsilu(acc);
    }
}

// Layer 2 unit b of one sample: z2 = W2[b,:] h1 + b2[b].
__device__ inline float layer2(const float *sw, const float *h1, int b) {
    float acc = sw[OB2 + b];
#pragma unroll
    for (int a = 0; a < NH; ++a) acc = fmaf(sw[OW2 + b * NH + a], h1[a], acc);
    return acc;
}

__device__ inline void norm_in(const Norm &nm, const float *x, float *xh) {
#pragma unroll
    for (int k = 0; k < NI; ++k) xh[k] = bf16r((x[k] - nm.xm[k]) / nm.xs[k]);
}

__global__ __launch_bounds__(256) void k_mlp_forward(
    int n, const float *__restrict__ x, const float *__restrict__ wb, Norm nm,
    float *__restrict__ y) {
    __shared__ float sw[NP];
    load_weights(wb, sw);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xh[NI], h1[NH];
    norm_in(nm, x + (size_t)NI * s, xh);
    layer1(sw, xh, h1, nullptr);
    float out = sw[OB3];
#pragma unroll
    for (int b = 0; b < NH; ++b)
        out = fmaf(sw[OW3 + b], bf16r(silu(layer2(sw, h1, b))), out);
    y[s] = out * nm.ys + nm.ym;
}

// Forward + backward of the samples idx[base .. base+count) (idx may be null:
// identity).  Loss = mean over the batch of (yhat - t_normalised)^2, yhat and
// the target in normalised units.  The derivative of the BF16 rounding is
// taken as 1 (straight-through).  Writes this block's gradient sums to
// gpart[blockIdx.x][NP] and its summed squared error to lpart[blockIdx.x].
constexpr int LP = NH + 1;  // padded LDS row (no bank conflicts per column)
__global__ __launch_bounds__(BS) void k_mlp_fwdbwd(
    int count, int base, const int *__restrict__ idx,
    const float *__restrict__ x, const float *__restrict__ t,
    const float *__restrict__ wb, Norm nm, float inv_batch,
    float *__restrict__ gpart, float *__restrict__ lpart) {
    __shared__ float sw[NP];
    __shared__ float sh1[BS * LP], sh2[BS * LP], sd1[BS * LP], sd2[BS * LP];
    __shared__ float sxh[BS * (NI + 1)];
    __shared__ float sd3[BS];
    load_weights(wb, sw);
    const int r = threadIdx.x;
    const int q = blockIdx.x * BS + r;
    float loss = 0.0f;
    if (q < count) {
        // Per-sample vectors live in this lane's LDS rows; only h1 and d2
        // are held in registers (fully unrolled inner loops), so the outer
        // loops stay rolled and the weights are not hoisted into registers.
        const int s = idx ? idx[base + q] : base + q;
        float xh[NI], h1[NH];
        norm_in(nm, x + (size_t)NI * s, xh);
#pragma unroll
        for (int a = 0; a < NH; ++a) {
            float acc = sw[OB1 + a];
#pragma unroll
            for (int k = 0; k < NI; ++k) acc = fmaf(sw[OW1 + a * NI + k], xh[k], acc);
            h1[a] = bf16r(silu(acc));
            sh1[r * LP + a] = h1[a];
            sd1[r * LP + a] = dsilu(acc);      // silu'(z1), replaced by d1
        }
        float out = sw[OB3];
#pragma unroll 1
        for (int b = 0; b < NH; ++b) {
            const float z = layer2(sw, h1, b);
            const float h = bf16r(silu(z));
            sd2[r * LP + b] = dsilu(z);        // silu'(z2), replaced by d2
            sh2[r * LP + b] = h;
            out = fmaf(sw[OW3 + b], h, out);
        }
        const float e = out - (t[s] - nm.ym) / nm.ys;
        loss = e * e;
        const float d3 = 2.0f * e * inv_batch;
        sd3[r] = d3;
        float d2[NH];
#pragma unroll
        for (int b = 0; b < NH; ++b) {
            d2[b] = d3 * sw[OW3 + b] * sd2[r * LP + b];
            sd2[r * LP + b] = d2[b];
        }
#pragma unroll 1
        for (int a = 0; a < NH; ++a) {
            float acc = 0.0f;
#pragma unroll
            for (int b = 0; b < NH; ++b) acc = fmaf(sw[OW2 + b * NH + a], d2[b], acc);
            sd1[r * LP + a] = acc * sd1[r * LP + a];
        }
#pragma unroll
        for (int k = 0; k < NI; ++k) sxh[r * (NI + 1) + k] = xh[k];
    } else {
        sd3[r] = 0.0f;
        for (int b = 0; b < NH; ++b) {
            sd2[r * LP + b] = 0.0f;
            sh2[r * LP + b] = 0.0f;
            sd1[r * LP + b] = 0.0f;
            sh1[r * LP + b] = 0.0f;
        }
        for (int k = 0; k < NI; ++k) sxh[r * (NI + 1) + k] = 0.0f;
    }
    __syncthreads();
    // Per-block sums over the BS samples; thread r owns unit r of both
    // hidden layers (BS == NH).
    float *g = gpart + (size_t)blockIdx.x * NP;
    const int u = r;
    {
        float gb2 = 0.0f, gw3 = 0.0f, gb1 = 0.0f;
        for (int s2 = 0; s2 < BS; ++s2) {
            gb2 += sd2[s2 * LP + u];
            gw3 = fmaf(sd3[s2], sh2[s2 * LP + u], gw3);
            gb1 += sd1[s2 * LP + u];
        }
        g[OB2 + u] = gb2;
        g[OW3 + u] = gw3;
        g[OB1 + u] = gb1;
    }
    for (int a = 0; a < NH; ++a) {
        float acc = 0.0f;
        for (int s2 = 0; s2 < BS; ++s2)
            acc = fmaf(sd2[s2 * LP + u], sh1[s2 * LP + a], acc);
        g[OW2 + u * NH + a] = acc;
    }
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        float acc = 0.0f;
        for (int s2 = 0; s2 < BS; ++s2)
            acc = fmaf(sd1[s2 * LP + u], sxh[s2 * (NI + 1) + k], acc);
        g[OW1 + u * NI + k] = acc;
    }
    // b3 and the loss: a reduction over the 64 lanes of the wave.
    float gb3 = sd3[r];
    for (int o = 32; o >= 1; o >>= 1) {
        gb3 += __shfl_xor(gb3, o);
        loss += __shfl_xor(loss, o);
    }
    if (r == 0) {
        g[OB3] = gb3;
        lpart[blockIdx.x] = loss;
    }
}

// Sum the nblk per-block gradients in block order; with adam != 0 apply one
// Adam step to the FP32 masters and refresh the BF16 copy.
__global__ __launch_bounds__(256) void k_mlp_reduce(
    int nblk, const float *__restrict__ gpart, float *__restrict__ gsum,
    int adam, float lr, float b1, float b2, float eps, float c1, float c2,
    float *__restrict__ w, float *__restrict__ wb, float *__restrict__ m,
    float *__restrict__ v) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NP) return;
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += gpart[(size_t)b * NP + k];
    gsum[k] = g;
    if (adam) {
        const float mk = b1 * m[k] + (1.0f - b1) * g;
        const float vk = b2 * v[k] + (1.0f - b2) * g * g;
        m[k] = mk;
        v[k] = vk;
        const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        w[k] = wk;
        wb[k] = bf16r(wk);
    }
}

__global__ void k_bf16_copy(const float *__restrict__ w, float *__restrict__ wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < NP) wb[k] = bf16r(w[k]);
}

// ---------------------------------------------------------------------------
// Host state and C-ABI
// ---------------------------------------------------------------------------

struct State {
    bool init = false;
    hipStream_t st = nullptr;
    float *w = nullptr, *wb = nullptr, *m = nullptr, *v = nullptr;
    float *gsum = nullptr;
    Norm nm{};
    long long adam_t = 0;
    // device copies of data
    float *dx = nullptr, *dt = nullptr, *dy = nullptr, *gpart = nullptr,
          *lpart = nullptr;
    int *didx = nullptr;
    size_t cap_n = 0, cap_blk = 0;
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
    MCHK(hipMalloc((void **)&M.wb, NP * sizeof(float)));
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

void ensure_data(size_t n, size_t nblk) {
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
    if (nblk > M.cap_blk) {
        if (M.gpart) (void)hipFree(M.gpart);
        if (M.lpart) (void)hipFree(M.lpart);
        MCHK(hipMalloc((void **)&M.gpart, nblk * NP * sizeof(float)));
        MCHK(hipMalloc((void **)&M.lpart, nblk * sizeof(float)));
        M.cap_blk = nblk;
    }
}

void set_params(const float *p) {
    ensure_init();
    MCHK(hipMemcpyAsync(M.w, p, NP * sizeof(float), hipMemcpyHostToDevice, M.st));
    hipLaunchKernelGGL(k_bf16_copy, dim3((NP + 255) / 256), dim3(256), 0, M.st,
                       M.w, M.wb);
    MCHK(hipMemsetAsync(M.m, 0, NP * sizeof(float), M.st));
    MCHK(hipMemsetAsync(M.v, 0, NP * sizeof(float), M.st));
    MCHK(hipStreamSynchronize(M.st));
    M.adam_t = 0;
}

// One pass of forward+backward over count samples starting at base (through
// idx if given); gradient (mean-square loss) summed into M.gsum, optionally
// followed by an Adam step.  Returns the summed squared error.
double grad_step(int count, int base, const int *didx, float lr, int adam) {
    const int nblk = (count + BS - 1) / BS;
    ensure_data(M.cap_n, (size_t)nblk);
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(nblk), dim3(BS), 0, M.st, count,
                       base, didx, M.dx, M.dt, M.wb, M.nm, 1.0f / count,
                       M.gpart, M.lpart);
    float c1 = 1.0f, c2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        M.adam_t++;
        c1 = 1.0f - std::pow(b1, (float)M.adam_t);
        c2 = 1.0f - std::pow(b2, (float)M.adam_t);
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0,
                       M.st, nblk, M.gpart, M.gsum, adam, lr, b1, b2, 1e-8f,
                       c1, c2, M.w, M.wb, M.m, M.v);
    MCHK(hipGetLastError());
    std::vector<float> lp(nblk);
    MCHK(hipMemcpyAsync(lp.data(), M.lpart, nblk * sizeof(float),
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
        hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0,
                           M.st, n, M.dx, M.wb, M.nm, M.dy);
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
