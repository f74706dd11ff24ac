// energy.hip -- the potential-energy sum of NBODY6++GPU's energy check on the
// device: the GPU branch of ENERGY (src/Main/energy.F:37-79) with the
// differential corrections of PHICOR (src/Main/phicor.f:1-112) and the
// kinetic energy (energy.F:141-146).
//
//   1. PHIDBL(IFIRST..NTOT) from the O(N^2) potential kernels of gpupot_
//      (optional: a caller may pass its own PHIDBL);
//   2. per body I = IFIRST..NTOT:
//        single (I <= N, one thread each): the c.m. neighbours at the end
//          of LIST(.,I) inside CMSEP2*RMIN2 are replaced by their two KS
//          components;
//        c.m. (I > N, one wavefront each): every perturber of the pair
//          (LIST(.,I1)) is resolved against each component; PHIDBL(I1),
//          PHIDBL(I1+1) are set to the c.m. value plus the component's
//          correction;
//   3. POT = 1/2 sum BODY * PHIDBL over singles and KS components and
//      ZKIN = 1/2 sum BODY |XDOT|^2 over IFIRST..NTOT, summed in a fixed
//      two-level order (deterministic run to run; the reference's serial
//      sum differs only in association).
//
// The corrections use the reference's operations in its order (the library
// is built without FMA contraction, division and sqrt are correctly
// rounded), so PHIDBL matches a serial restatement bit for bit.

#include "gpunb_internal.h"

namespace gpunb {

namespace {

__device__ inline double d2(const double *x, int a, int b) {
    // RIJ2 = 0; RIJ2 = RIJ2 + (X(K,a) - X(K,b))**2, K = 1..3
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double d = x[3 * (size_t)(a - 1) + k] - x[3 * (size_t)(b - 1) + k];
        s = s + d * d;
    }
    return s;
}

__global__ void k_energy_corr(EnergyLaunch L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = L.ifirst + t;
    if (I > L.ntot) return;
    const int N = L.n;
    const double *X = L.x, *B = L.body;
    double *PHI = L.phidbl;
    double dphi1 = 0.0;
    bool ok = true;
    auto valid = [&](int J) { return J >= 1 && J <= L.ntot; };
    if (I <= N) {
        const int *row = L.list + (size_t)(I - 1) * L.lmax;
        const int nnb = row[0];
        ok = nnb >= 0 && nnb + 1 <= L.lmax;
        // Backwards until the last c.m. body (phicor.f:14-39).
        for (int l = ok ? nnb + 1 : 1; l >= 2; --l) {
            const int J = row[l - 1];
            if (!valid(J)) {
                ok = false;
                break;
            }
            if (J <= N) break;
            const double rij2 = d2(X, I, J);
            if (rij2 < L.cmsep2_rmin2) {
                const int j1 = 2 * (J - N) - 1, j2 = j1 + 1;
                const double rj1 = d2(X, I, j1);
                const double rj2 = d2(X, I, j2);
                dphi1 = dphi1 - B[J - 1] / sqrt(rij2) + B[j1 - 1] / sqrt(rj1) +
                        B[j2 - 1] / sqrt(rj2);
            }
        }
        double c = 0.0;
        if (ok) {
            const double phi = PHI[I - 1] + dphi1;
            PHI[I - 1] = phi;
            c = B[I - 1] * phi;
        }
        L.cpot[t] = c;
    }
    // KS pair c.m. bodies (I > N) are corrected by k_energy_cm.
    if (!ok) atomicOr(L.err, ERR_ENERGY);
    const double *v = L.xdot + 3 * (size_t)(I - 1);
    L.ckin[t] = B[I - 1] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// KS pair c.m. bodies (phicor.f:40-108), one wavefront per pair: the pair's
// perturbers (up to several hundred in a dense core) are gathered and their
// potential terms evaluated 64 at a time, one per lane, with the same
// operations as the serial loop; lane 0 then adds the terms in list order,
// so DPHI1 / DPHI2 are the serial loop's values bit for bit.
__global__ __launch_bounds__(64) void k_energy_cm(EnergyLaunch L) {
    const int N = L.n;
    const int I = N + 1 + (int)blockIdx.x;
    if (I > L.ntot) return;
    const int lane = threadIdx.x;
    const double *X = L.x, *B = L.body;
    const int i1 = 2 * (I - N) - 1, i2 = i1 + 1;
    const int *row = L.list + (size_t)(i1 - 1) * L.lmax;
    const int np = row[0];
    __shared__ double ta[64], tb1[64], tb2[64], tc1[64], tc2[64];
    __shared__ int kind[64];
    __shared__ int bad;
    if (lane == 0) bad = (np < 0 || np + 1 > L.lmax) ? 1 : 0;
    __syncthreads();
    double dphi1 = 0.0, dphi2 = 0.0;
    for (int base = 0; !bad && base < np; base += 64) {
        const int l = base + lane;
        int kd = 2;
        if (l < np) {
            const int J = row[l + 1];          // LIST(l + 2, I1)
            if (J < 1 || J > L.ntot) {
                kd = -1;
            } else {
                const double a = B[J - 1] / sqrt(d2(X, I, J));
                ta[lane] = a;
                if (J <= N) {
                    kd = 0;
                    tb1[lane] = B[J - 1] / sqrt(d2(X, i1, J));
                    tb2[lane] = B[J - 1] / sqrt(d2(X, i2, J));
                } else {
                    kd = 1;
                    const int j1 = 2 * (J - N) - 1;
                    tb1[lane] = B[j1 - 1] / sqrt(d2(X, i1, j1));
                    tb2[lane] = B[j1 - 1] / sqrt(d2(X, i2, j1));
                    tc1[lane] = B[j1] / sqrt(d2(X, i1, j1 + 1));
                    tc2[lane] = B[j1] / sqrt(d2(X, i2, j1 + 1));
                }
            }
        }
        kind[lane] = kd;
        __syncthreads();
        if (lane == 0) {
            const int cnt = np - base < 64 ? np - base : 64;
            for (int k = 0; k < cnt; ++k) {
                if (kind[k] < 0) {
                    bad = 1;
                    break;
                }
                if (kind[k] == 0) {
                    dphi1 = dphi1 - ta[k] + tb1[k];
                    dphi2 = dphi2 - ta[k] + tb2[k];
                } else {
                    dphi1 = dphi1 - ta[k];
                    dphi2 = dphi2 - ta[k];
                    dphi1 = dphi1 + tb1[k];
                    dphi2 = dphi2 + tb2[k];
                    dphi1 = dphi1 + tc1[k];
                    dphi2 = dphi2 + tc2[k];
                }
            }
        }
        __syncthreads();
    }
    if (lane != 0) return;
    const int t = I - L.ifirst;
    double c = 0.0;
    if (!bad) {
        double *PHI = L.phidbl;
        const double pcm = PHI[I - 1];
        const double p1 = pcm + dphi1, p2 = pcm + dphi2;
        PHI[i1 - 1] = p1;
        PHI[i2 - 1] = p2;
        c = B[i1 - 1] * p1 + B[i2 - 1] * p2;
    } else {
        atomicOr(L.err, ERR_ENERGY);
    }
    L.cpot[t] = c;
}

// Deterministic two-level sum.  Level 1: block b of 256 threads sums items
// [2048 b, 2048 (b + 1)), eight per thread in order, then a fixed tree in
// LDS.  Level 2: one block sums the level-1 partials in the same way and
// writes out[0] = POT / 2, out[1] = ZKIN / 2.
constexpr int kSumItems = 2048;

__device__ inline void block_tree(double *sp, double *sk, int t) {
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            sp[t] += sp[t + s];
            sk[t] += sk[t + s];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_energy_sum1(const double *cpot,
                                                     const double *ckin,
                                                     int n, double *ppart,
                                                     double *kpart) {
    __shared__ double sp[256], sk[256];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kSumItems;
    double vp[8], vk[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int i = base + u * 256 + t;
        vp[u] = i < n ? cpot[i] : 0.0;
        vk[u] = i < n ? ckin[i] : 0.0;
    }
    double p = 0.0, k = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        p += vp[u];
        k += vk[u];
    }
    sp[t] = p;
    sk[t] = k;
    block_tree(sp, sk, t);
    if (t == 0) {
        ppart[blockIdx.x] = sp[0];
        kpart[blockIdx.x] = sk[0];
    }
}

__global__ __launch_bounds__(256) void k_energy_sum2(const double *ppart,
                                                     const double *kpart,
                                                     int nb, double *out) {
    __shared__ double sp[256], sk[256];
    const int t = threadIdx.x;
    double p = 0.0, k = 0.0;
    for (int i = t; i < nb; i += 256) {
        p += ppart[i];
        k += kpart[i];
    }
    sp[t] = p;
    sk[t] = k;
    block_tree(sp, sk, t);
    if (t == 0) {
        out[0] = 0.5 * sp[0];
        out[1] = 0.5 * sk[0];
    }
}

}  // namespace

void launch_energy(hipStream_t st, const EnergyLaunch &L) {
    const int nb = L.ntot - L.ifirst + 1;
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_energy_corr, dim3((unsigned)((nb + 255) / 256)),
                       dim3(256), 0, st, L);
    const int npairs = L.ntot - L.n;
    if (npairs > 0)
        hipLaunchKernelGGL(k_energy_cm, dim3((unsigned)npairs), dim3(64), 0,
                           st, L);
    const int nsum = (nb + kSumItems - 1) / kSumItems;
    hipLaunchKernelGGL(k_energy_sum1, dim3((unsigned)nsum), dim3(256), 0, st,
                       L.cpot, L.ckin, nb, L.ppart, L.kpart);
    hipLaunchKernelGGL(k_energy_sum2, dim3(1), dim3(256), 0, st, L.ppart,
                       L.kpart, nsum, L.out);
}

}  // namespace gpunb
