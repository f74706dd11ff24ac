// kernels.hip -- CDNA4 (gfx950) kernels of the regular-force engine.
//
// Compiled with -ffp-contract=off: every expression is evaluated as written.
// The FP32 force arithmetic uses explicit fmaf(); the FP64 neighbour
// predicate and the predictor contain no fused operation, so they reproduce
// the FP64 restatement of the reference bit for bit.
//
// Kernels
//   k_pack_j       FP64 AoS (m, x, v) -> 64-byte JRec records; NaN check;
//                  max |v| and min positive m of the j set.
//   k_predict      Hermite predictor of every j (xbpredall.f:18-26) -> JRec.
//   k_jpset        scatter of corrected Hermite states into the predictor SoA.
//   k_regf_pass1   mixed kernel: pair sums + neighbour capture over one j
//                  slice per workgroup of four wavefronts; tiles far from
//                  every lane in FP32, near tiles in FP64 (_q: persistent
//                  form taking tasks from per-XCD queues).
//   k_regf_pass1_f64  the same in FP64 throughout (_q: persistent form).
//   k_regf_pass2   per-i FP64 reduction of the slice partials, ordered list
//                  compaction and (when asked for) the nearest neighbour.
//   k_pot_pass1/2  O(N^2) potential (gpupot.gpu.cu:32-57).
#include "gpunb_internal.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

namespace gpunb {

namespace {

constexpr float kU32 = 5.9604644775390625e-08f;  // 2^-24, FP32 unit roundoff

__device__ inline unsigned long long dbits(double v) {
    return (unsigned long long)__double_as_longlong(v);
}
__device__ inline double bitsd(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

// Squared norm with the association ((a*a + b*b) + c*c) and no FMA.
__device__ inline double sq3d(double a, double b, double c) {
    double s = a * a;
    s = s + b * b;
    s = s + c * c;
    return s;
}

__device__ inline double wave_max(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_min(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ inline int wave_sum(int v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline void set_pad(JRec &r) {
    r.x = kPadPos;
    r.y = kPadPos;
    r.z = kPadPos;
    r.vx = 0.0;
    r.vy = 0.0;
    r.vz = 0.0;
    r.m = 0.0;
    r.m32 = 0.0f;
    r.pad = 0.0f;
}

// Fold the per-thread speed and mass extrema of one wave into JStats.
// Workgroup-wide max |v| / min positive m folded into *jst.  Called by every
// thread of a 256-thread workgroup (it synchronises).  The bounds only grow
// (vmax) or shrink (mmin), so an update that the stored value already covers
// is skipped: at most one atomic per workgroup and quantity, and after the
// first workgroups almost none (same-address atomics serialise in L2).
__device__ inline void stats_fold(JStats *jst, double vmag, double mpos) {
    __shared__ double s_v[4], s_m[4];
    vmag = wave_max(vmag);
    mpos = wave_min(mpos);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_v[w] = vmag;
        s_m[w] = mpos;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int k = 1; k < nw; ++k) {
            vmag = fmax(vmag, s_v[k]);
            mpos = fmin(mpos, s_m[k]);
        }
        if (vmag > 0.0 &&
            dbits(vmag) > __hip_atomic_load(&jst->vmax_bits, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&jst->vmax_bits, dbits(vmag));
        if (mpos < INFINITY &&
            dbits(mpos) < __hip_atomic_load(&jst->mmin_bits, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&jst->mmin_bits, dbits(mpos));
    }
}

// |v| rounded up by a few ulps, so the bound built on it is safe.
__device__ inline double speed_up(double vx, double vy, double vz) {
    return sqrt(sq3d(vx, vy, vz)) * (1.0 + 1e-15);
}

// Cluster of a pass-1 task / pass-2 workgroup: last c with key(c) <= k.
__device__ inline int find_task(const ClusterDesc *__restrict__ d, int ncl,
                                long long k) {
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].task0 <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__device__ inline int find_blk(const ClusterDesc *__restrict__ d, int ncl,
                               long long k) {
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].blk0 <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// ---------------------------------------------------------------------------
// j-set preparation
// ---------------------------------------------------------------------------

__global__ void k_stats_reset(JStats *jst, int *total) {
    if (threadIdx.x == 0) {
        jst->vmax_bits = 0ull;
        jst->mmin_bits = dbits(INFINITY);
        if (total) *total = 0;
    }
}

__global__ __launch_bounds__(256) void k_pack_j(
    const double *__restrict__ m, const double *__restrict__ x,
    const double *__restrict__ v, int nj, int nj_pad, JRec *__restrict__ jp,
    JStats *__restrict__ jst, unsigned *__restrict__ err) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double vmag = 0.0, mpos = INFINITY;
    if (j < nj) {
        JRec r;
        r.x = x[3 * (size_t)j + 0];
        r.y = x[3 * (size_t)j + 1];
        r.z = x[3 * (size_t)j + 2];
        r.vx = v[3 * (size_t)j + 0];
        r.vy = v[3 * (size_t)j + 1];
        r.vz = v[3 * (size_t)j + 2];
        r.m = m[j];
        r.m32 = (float)r.m;
        r.pad = 0.0f;
        const double chk = r.x + r.y + r.z + r.vx + r.vy + r.vz + r.m;
        if (chk != chk) atomicOr(err, ERR_NAN_J);
        jp[j] = r;
        vmag = speed_up(r.vx, r.vy, r.vz);
        if (r.m > 0.0) mpos = r.m;
    } else if (j < nj_pad) {
        JRec r;
        set_pad(r);
        jp[j] = r;
    }
    stats_fold(jst, vmag, mpos);
}

__global__ __launch_bounds__(256) void k_predict(
    double t, int nj, int nj_pad, const double *__restrict__ x0,
    const double *__restrict__ v0, const double *__restrict__ f,
    const double *__restrict__ fd, const double *__restrict__ t0,
    const double *__restrict__ m, JRec *__restrict__ jp,
    JStats *__restrict__ jst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    // The block's slices of the four (3,N) arrays are staged through LDS
    // with unit-stride loads (a thread reading its own three components
    // would touch each cache line in thirds).
    __shared__ double sx[768], sv[768], sf[768], sfd[768];
    const double tj = j < nj ? t0[j] : 0.0;
    const double mj = j < nj ? m[j] : 0.0;
    {
        const size_t base = 3 * (size_t)blockIdx.x * blockDim.x;
        const size_t lim = 3 * (size_t)nj;
        for (int e = threadIdx.x; e < 3 * (int)blockDim.x; e += blockDim.x) {
            const size_t q = base + (size_t)e;
            if (q < lim) {
                sx[e] = x0[q];
                sv[e] = v0[q];
                sf[e] = f[q];
                sfd[e] = fd[q];
            }
        }
    }
    __syncthreads();
    double vmag = 0.0, mpos = INFINITY;
    if (j < nj) {
        // xbpredall.f:18-26, evaluated exactly as written (no FMA).
        const double s = t - tj;
        const double s1 = 1.5 * s;
        const double s2 = 2.0 * s;
        double xo[3], vo[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int e = 3 * (int)threadIdx.x + k;
            xo[k] = ((sfd[e] * s + sf[e]) * s + sv[e]) * s + sx[e];
            vo[k] = (sfd[e] * s1 + sf[e]) * s2 + sv[e];
        }
        JRec r;
        r.x = xo[0];
        r.y = xo[1];
        r.z = xo[2];
        r.vx = vo[0];
        r.vy = vo[1];
        r.vz = vo[2];
        r.m = mj;
        r.m32 = (float)r.m;
        r.pad = 0.0f;
        jp[j] = r;
        vmag = speed_up(r.vx, r.vy, r.vz);
        if (r.m > 0.0) mpos = r.m;
    } else if (j < nj_pad) {
        JRec r;
        set_pad(r);
        jp[j] = r;
    }
    stats_fold(jst, vmag, mpos);
}

// src holds n records of 14 doubles: x0[3] v0[3] f[3] fdot[3] m t0.
__global__ __launch_bounds__(256) void k_jpset(
    int n, const int *__restrict__ addr, const double *__restrict__ src,
    double *__restrict__ x0, double *__restrict__ v0, double *__restrict__ f,
    double *__restrict__ fd, double *__restrict__ mass,
    double *__restrict__ t0, unsigned *__restrict__ err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = addr[k] - 1;
    const double *s = src + 14 * (size_t)k;
    double chk = 0.0;
#pragma unroll
    for (int c = 0; c < 14; ++c) chk += s[c];
    if (chk != chk) atomicOr(err, ERR_NAN_J);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x0[3 * (size_t)j + c] = s[c];
        v0[3 * (size_t)j + c] = s[3 + c];
        f[3 * (size_t)j + c] = s[6 + c];
        fd[3 * (size_t)j + c] = s[9 + c];
    }
    mass[j] = s[12];
    t0[j] = s[13];
}

__global__ __launch_bounds__(256) void k_jget(int n, const JRec *__restrict__ jp,
                                              double *__restrict__ x,
                                              double *__restrict__ v) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const JRec r = jp[j];
    x[3 * (size_t)j + 0] = r.x;
    x[3 * (size_t)j + 1] = r.y;
    x[3 * (size_t)j + 2] = r.z;
    v[3 * (size_t)j + 0] = r.vx;
    v[3 * (size_t)j + 1] = r.vy;
    v[3 * (size_t)j + 2] = r.vz;
}

// Device buffer -> mapped pinned host buffer, 16 bytes per lane (each
// wavefront writes whole 1 KiB runs across the host link), plus the error
// word: the results of a host-array call come back without a copy-engine
// transfer, in stream order.
__global__ __launch_bounds__(256) void k_copy_to_host(
    const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16,
    const unsigned *__restrict__ err, unsigned *__restrict__ herr) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16;
         k += stride)
        dst[k] = src[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) *herr = *err;
}

// Mapped pinned host buffer -> device buffer, 16 bytes per lane, and
// *zero = 0 (the list cursor of the call): the inputs of a host-array call
// go up in stream order, without a copy-engine transfer.
__global__ __launch_bounds__(256) void k_copy_from_host(
    const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16,
    int *__restrict__ zero) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16;
         k += stride)
        dst[k] = src[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
}

__global__ void k_set_desc(ClusterDesc *d, ClusterDesc v) { *d = v; }

__global__ void k_zero_int(int *p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// The kP1W wavefronts of a pass-1 workgroup share one (i-wavefront, j
// slice) task, each over a contiguous part of the slice's tiles; their FP64
// sums are added in wavefront order through LDS and the first wavefront
// writes the slice partial of its lanes (a fixed order: the results do not
// depend on which wavefront finishes first).
__device__ inline void slice_partial_store(const double (&v)[7], int wv,
                                           int lane, bool act, double *p) {
    __shared__ double red[kP1W - 1][7][64];
    if (wv > 0) {
#pragma unroll
        for (int c = 0; c < 7; ++c) red[wv - 1][c][lane] = v[c];
    }
    __syncthreads();
    if (wv == 0 && act) {
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            double t = v[c];
#pragma unroll
            for (int w = 0; w < kP1W - 1; ++w) t += red[w][c][lane];
            p[c] = t;
        }
        p[7] = 0.0;
    }
}

// Threshold of the FP64 screen min(r2, r2 + 2 dtr dx.dv) < thr (times m_j
// with mass-weighted lists) for a lane with h2 > 0: pairs below it are
// decided by the oracle's exact predicate.  The FP32-form screen's value in
// FP64 is within 4.1u S <= 4.1u H (1 + 2 rho)^2 of its exact value (u =
// 2^-53, rho = |dtr| (|v_i| + max|v_j|) / sqrt(H_min)), and the oracle's r2p
// (separate multiply and add in the prediction) within 3.1u H + 4.1u H (1 +
// 2 rho) of the exact one near the threshold; the threshold allows twice
// their sum.
template <bool MFLAG>
__device__ inline double f64_screen_thr(double h2d, double dtrd, double u0,
                                        double w0, double q0,
                                        const JStats *__restrict__ jst) {
    const double vmax = bitsd(jst->vmax_bits);
    double hmin = h2d;
    if (MFLAG) {
        const double mmin = bitsd(jst->mmin_bits);
        hmin = (mmin < INFINITY) ? h2d * mmin : h2d;
    }
    const double u = 1.1102230246251565e-16;  // 2^-53
    const double rho = fabs(dtrd) * (speed_up(u0, w0, q0) + vmax) / sqrt(hmin);
    const double g = 1.0 + 2.0 * rho;
    const double eps =
        2.0 * (8.0 * u * g * g + 4.0 * u) + (MFLAG ? 24.0 : 16.0) * u;
    const double tt = h2d + h2d * eps;
    return fmax(nextafter(nextafter(tt, INFINITY), INFINITY), 1e-30);
}

// ---------------------------------------------------------------------------
// Regular force, pass 1: one wavefront = 64 i-particles (one per lane) x one
// contiguous slice of the Hilbert-sorted j set, tile by tile (64 j).  j
// records are wave-uniform, so they arrive through the scalar cache as SGPR
// operands.  Each tile is run in one of two ways, chosen per wave:
//
// Far tiles: FP32 pair terms on tile-relative differences.  Every j is
// stored as an FP32 offset o_j from the FP64 centre C_t of its tile, so per
// (lane, tile) the FP64 difference C_t - x_i is formed once and rounded to
// FP32 (D), and per pair dx = D + o_j is one FP32 add per component
// (velocities alike).  With |o_j| <= e_t (the tile's offset bound) the
// rounding error is, per component, at most 2u (|dx_c| + |o_c|) (u = 2^-24).
// The FP32 sums of a tile are folded into FP64 accumulators at its end.
//
// Near tiles: the FP64 pair loop of k_regf_pass1_f64 (FP64 differences of
// the sorted FP64 records, 1/r by v_rsq_f64 and one Newton step, the
// oracle's neighbour predicate bit for bit on the pairs its screen cannot
// decide), accumulating into the FP64 sums directly.
//
// The split.  FP32 rounding of a pair term is harmless unless the pair is
// close: relative to the neighbour sphere, and more so for heavy j whose
// terms are large.  A pair is in the near band when min(r2, q) < B w_j,
//   q   = r2 + 2 dtr dx.dv (a lower bound of the predicted distance
//         r2p = q + dtr^2 |dv|^2, so min(r2, q) <= min(r2, r2p)),
//   B   = K h2 (K h2 <m> with mass-weighted lists), K = kBand, per lane,
//         raised per tile to (e_t / 2)^2, so that no FP32 pair has |dx|
//         below e_t / 2 (where the tile-relative difference would carry a
//         large relative error),
//   w_j = max(1, x^(2/3), 1.4 x / K), x = m_j / <m>, per j.
// Per (lane, tile), with d = |D|, R = d - e_t and A = |dtr| (|DV| + ev_t)
// (|DV| as d, ev_t the tile's velocity-offset bound), every pair of the
// tile has r >= R and |dtr| |dv| <= A, so min(r2, r2p) >= (R - A)^2 when
// R > A; a lane finds the tile far when R > A + sqrt(A^2 + B_t wmax_t)
// (wmax_t the tile's largest w, 1e-5 relative margins for the FP32
// evaluation of the test), i.e. (R - A)^2 > B_t wmax_t: no pair of the tile
// is in the band.  A wave runs a tile in FP32 only when the tile is far for
// every lane, else in FP64.  Since B_t w_j >= 1.4 H (H = h2, or m_j h2 with
// mass-weighted lists: K w_j >= 1.4 x), a far tile holds no neighbour, so
// every list entry comes from the FP64 predicate.  Heavy j (m_j > kHeavy
// <m>) sit in tiles of their own, so light tiles keep wmax_t <= ~2.
// ---------------------------------------------------------------------------

// Opaque copy of a record held in SGPRs: the look-ahead loads of the pair
// loops are loop-carried values, and without this LLVM folds the phi of the
// pre-loop load and the in-loop load into one load at the loop top (waited
// on at once).  Emits no instruction.
__device__ inline void opaque_sgpr(JRec32 &r) {
    asm volatile("" : "+s"(r.ox), "+s"(r.oy), "+s"(r.oz), "+s"(r.m),
                      "+s"(r.ovx), "+s"(r.ovy), "+s"(r.ovz), "+s"(r.w));
}
__device__ inline void opaque_sgpr(JRec &r) {
    asm volatile("" : "+s"(r.x), "+s"(r.y), "+s"(r.z), "+s"(r.vx),
                      "+s"(r.vy), "+s"(r.vz), "+s"(r.m));
}

// Task of a cluster -> (j slice s, i wavefront iw).  XCD x runs the x-th
// eighth of a launch's tasks (xcd_task).  Slice-major order gives it a share
// of the slices over most of the i.  In a one-cluster launch of at least
// kIBlockMin i-wavefronts, the wavefronts are instead cut into eight blocks of
// consecutive ones, taken block by block and slice by slice within a block:
// XCD x then runs every slice for one eighth of the i, in Hilbert order.
// The King N = 131072 launch takes 13.8 ms that way against 15.0 ms
// slice-major, with the same results; ensembles and small calls (1024 i)
// measured slower, so they keep slice-major order.
constexpr int kIBlocks = 8;
constexpr int kIBlockMin = 256;
__device__ inline void task_slice(int task, int n_iw, int nslice, bool one_cluster,
                                  int &s, int &iw) {
    const int B = one_cluster && n_iw >= kIBlockMin
                      ? (n_iw + kIBlocks - 1) / kIBlocks : n_iw;
    const int nb = (n_iw + B - 1) / B;
    const int b = min(task / (B * nslice), nb - 1);
    const int r = task - b * B * nslice;
    const int Bb = b == nb - 1 ? n_iw - b * B : B;
    s = r / Bb;
    iw = b * B + (r - s * Bb);
}

// One pass-1 task of the mixed kernel (a workgroup: one i-wavefront x one j
// slice, task number gtask < ntask).
template <bool MFLAG, int G>
__device__ __forceinline__ void regf_pass1_task(
    long long gtask,
    const JRec32 *__restrict__ j32_all, const JTile *__restrict__ tile_all,
    const JRec *__restrict__ jps_all, const JClus *__restrict__ jcl_all,
    const ClusterDesc *__restrict__ desc, int ncl, long long ntask,
    const JStats *__restrict__ jst, const double *__restrict__ xi_all,
    const double *__restrict__ vi_all, const double *__restrict__ h2_all,
    const double *__restrict__ dtr_all, double *__restrict__ part_all,
    int *__restrict__ nbc_all, int *__restrict__ nbl_all,
    unsigned *__restrict__ err) {
    const ClusterDesc &D = desc[find_task(desc, ncl, gtask)];
    const int nj = D.nj, ntile = D.ntile, ni = D.ni, n_iw = D.n_iw;
    const int nslice = D.nslice, cap = D.cap;
    const JRec32 *__restrict__ j32 = j32_all + D.j0;
    const JTile *__restrict__ tiles = tile_all + D.j0 / kVTile;
    const JRec *__restrict__ jps = jps_all + D.j0;
    const double mbar = jcl_all[D.jc].mbar;
    const double *__restrict__ xi = xi_all + 3 * (size_t)D.i0;
    const double *__restrict__ vi = vi_all + 3 * (size_t)D.i0;
    const double *__restrict__ h2 = h2_all + D.i0;
    const double *__restrict__ dtr = dtr_all + D.i0;
    double *__restrict__ part = part_all + 8 * D.pbase;
    int *__restrict__ nbc = nbc_all + D.ibase;
    int *__restrict__ nbl = nbl_all + D.sbase;
    const int task = (int)(gtask - D.task0);
    int s, iw;
    task_slice(task, n_iw, nslice, ncl == 1, s, iw);
    // Wavefront index (uniform within the wavefront: keeps the j record
    // addresses scalar).
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int i = iw * 64 + lane;
    const bool act = i < ni;

    // Slice s of the j tiles, and this wavefront's contiguous share of it.
    const int sb = (int)(((long long)ntile * s) / nslice);
    const int se = (int)(((long long)ntile * (s + 1)) / nslice);
    const int tb = sb + (int)(((long long)(se - sb) * wv) / kP1W);
    const int te = sb + (int)(((long long)(se - sb) * (wv + 1)) / kP1W);

    double x0 = 0.0, y0 = 0.0, z0 = 0.0, u0 = 0.0, w0 = 0.0, q0 = 0.0;
    double h2d = 0.0, dtrd = 0.0, dtr2x = 0.0;   // dtr2x = 2 dtr
    double thr = -1.0;          // FP64 screen threshold (k_regf_pass1_f64)
    float adtr = 0.0f;          // >= |dtr|
    float B = 0.0f;             // band base
    if (act) {
        x0 = xi[3 * (size_t)i + 0];
        y0 = xi[3 * (size_t)i + 1];
        z0 = xi[3 * (size_t)i + 2];
        u0 = vi[3 * (size_t)i + 0];
        w0 = vi[3 * (size_t)i + 1];
        q0 = vi[3 * (size_t)i + 2];
        h2d = h2[i];
        dtrd = dtr[i];
        dtr2x = 2.0 * dtrd;
        const double chk = x0 + y0 + z0 + u0 + w0 + q0 + h2d + dtrd;
        if (chk != chk) atomicOr(err, ERR_NAN_I);
        adtr = fabsf((float)dtrd) * (1.0f + 2.0f * kU32);
        if (h2d > 0.0) {
            const double bb = (double)kBand * h2d * (MFLAG ? mbar : 1.0) *
                              (1.0 + 8.0 * (double)kU32);
            B = (float)bb * (1.0f + 2.0f * kU32);
            thr = f64_screen_thr<MFLAG>(h2d, dtrd, u0, w0, q0, jst);
        } else {
            thr = 1e-30;
        }
    }

    double Ax = 0.0, Ay = 0.0, Az = 0.0, Jx = 0.0, Jy = 0.0, Jz = 0.0, Ph = 0.0;

    // FP64 pair of a near tile: the screen of k_regf_pass1_f64, the exact
    // predicate for the pairs it flags.
    auto resolve = [&](const JRec &J, int j) {
        const double dx = J.x - x0, dy = J.y - y0, dz = J.z - z0;
        const double dvx = J.vx - u0, dvy = J.vy - w0, dvz = J.vz - q0;
        const double r2d = sq3d(dx, dy, dz);
        const double pxd = dx + dtrd * dvx;
        const double pyd = dy + dtrd * dvy;
        const double pzd = dz + dtrd * dvz;
        const double r2pd = sq3d(pxd, pyd, pzd);
        const double mnd = r2d < r2pd ? r2d : r2pd;
        const double H = MFLAG ? J.m * h2d : h2d;
        if ((mnd < H) && (j < nj)) {
            // Slices append in any order (pass 2 sorts); entries beyond cap
            // only count (overflow).
            const int k2 = atomicAdd(&nbc[i], 1);
            if (k2 < cap) nbl[(size_t)i * cap + k2] = j;
        } else if (r2d != 0.0) {
            // Rare (band pairs, r2 < 1e-30): exact FP64 1/sqrt.
            const double rinv = 1.0 / sqrt(r2d);
            const double rinv2 = rinv * rinv;
            const double mr = J.m * rinv;
            const double mr3 = mr * rinv2;
            const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
            const double al = (-3.0 * rv) * rinv2;
            Ph += mr;
            Ax = fma(mr3, dx, Ax);
            Ay = fma(mr3, dy, Ay);
            Az = fma(mr3, dz, Az);
            Jx = fma(mr3, fma(al, dx, dvx), Jx);
            Jy = fma(mr3, fma(al, dy, dvy), Jy);
            Jz = fma(mr3, fma(al, dz, dvz), Jz);
        }
    };

    // The records of the next group (and the next tile record) are loaded
    // while the current ones are evaluated; the look-ahead restarts in every
    // tile, so the prefetched records stay in SGPRs (carried across the
    // tile's branches the compiler would copy them to VGPRs every group).
    // The record arrays carry a spare tile past their end, so the last
    // look-ahead stays in bounds.
    // The wavefront's tiles [tb, te) in view tiles of kVTile j.
    constexpr int kR = kTile / kVTile;
    JTile Tn = tiles[tb * kR];
    for (int t = tb * kR; t < te * kR; ++t) {
        const JTile T = Tn;
        Tn = tiles[t + 1];
        const int jt = t * kVTile;
        const float Dx = (float)(T.cx - x0), Dy = (float)(T.cy - y0),
                    Dz = (float)(T.cz - z0);
        const float Vx = (float)(T.cu - u0), Vy = (float)(T.cv - w0),
                    Vz = (float)(T.cw - q0);
        const float Bt = fmaxf(B, T.eb);
        const float d = sqrtf(fmaf(Dz, Dz, fmaf(Dy, Dy, Dx * Dx)));
        const float dvn = sqrtf(fmaf(Vz, Vz, fmaf(Vy, Vy, Vx * Vx)));
        const float R = fmaf(d, 1.0f - 1e-5f, -T.e);
        const float A = adtr * (dvn + T.ev) * (1.0f + 1e-5f);
        const float Tmax = Bt * T.wmax * (1.0f + 1e-5f);
        const bool far =
            !act || (R > fmaf(sqrtf(fmaf(A, A, Tmax)), 1.0f + 1e-5f, A));
        if (__ballot(!far) == 0) {
            float ax = 0.f, ay = 0.f, az = 0.f, jx = 0.f, jy = 0.f, jz = 0.f,
                  ph = 0.f;
            auto pair32 = [&](const JRec32 &J) {
                const float fx = Dx + J.ox, fy = Dy + J.oy, fz = Dz + J.oz;
                const float gx = Vx + J.ovx, gy = Vy + J.ovy, gz = Vz + J.ovz;
                const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                const float rv = fmaf(fz, gz, fmaf(fy, gy, fx * gx));
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float rinv2 = rinv * rinv;
                const float mr = J.m * rinv;
                const float mr3 = mr * rinv2;
                const float al = (-3.0f * rv) * rinv2;
                ph += mr;
                ax = fmaf(mr3, fx, ax);
                ay = fmaf(mr3, fy, ay);
                az = fmaf(mr3, fz, az);
                jx = fmaf(mr3, fmaf(al, fx, gx), jx);
                jy = fmaf(mr3, fmaf(al, fy, gy), jy);
                jz = fmaf(mr3, fmaf(al, fz, gz), jz);
            };
            // Two record buffers in turn: the next group's records load
            // while the current group is evaluated, with no register copies.
            JRec32 Pa[G], Pb[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                Pa[k] = j32[jt + k];
                opaque_sgpr(Pa[k]);
            }
            for (int jg = 0; jg < kVTile; jg += 2 * G) {
#pragma unroll
                for (int k = 0; k < G; ++k) Pb[k] = j32[jt + jg + G + k];
#pragma unroll
                for (int k = 0; k < G; ++k) pair32(Pa[k]);
#pragma unroll
                for (int k = 0; k < G; ++k) Pa[k] = j32[jt + jg + 2 * G + k];
#pragma unroll
                for (int k = 0; k < G; ++k) pair32(Pb[k]);
            }
            Ax += (double)ax;
            Ay += (double)ay;
            Az += (double)az;
            Jx += (double)jx;
            Jy += (double)jy;
            Jz += (double)jz;
            Ph += (double)ph;
        } else {
            // FP64 group of G records at sorted position jb: the screen
            // flags the pairs the exact predicate decides.
            auto group64 = [&](const JRec (&J)[G], int jb) {
                bool fl[G];
                bool any = false;
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const double dx = J[k].x - x0;
                    const double dy = J[k].y - y0;
                    const double dz = J[k].z - z0;
                    const double dvx = J[k].vx - u0;
                    const double dvy = J[k].vy - w0;
                    const double dvz = J[k].vz - q0;
                    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
                    const double q = fma(dtr2x, rv, r2);
                    double tj = thr;
                    if (MFLAG) tj = fmax(thr * J[k].m, 1e-30);
                    fl[k] = fmin(r2, q) < tj;
                    any = any || fl[k];
                    double y = __builtin_amdgcn_rsq(r2);
                    y = fma(y, 0.5 * fma(-r2 * y, y, 1.0), y);
                    const double rinv = fl[k] ? 0.0 : y;
                    const double rinv2 = rinv * rinv;
                    const double mr = J[k].m * rinv;
                    const double mr3 = mr * rinv2;
                    const double al = (-3.0 * rv) * rinv2;
                    Ph += mr;
                    Ax = fma(mr3, dx, Ax);
                    Ay = fma(mr3, dy, Ay);
                    Az = fma(mr3, dz, Az);
                    Jx = fma(mr3, fma(al, dx, dvx), Jx);
                    Jy = fma(mr3, fma(al, dy, dvy), Jy);
                    Jz = fma(mr3, fma(al, dz, dvz), Jz);
                }
                if (any) {
#pragma unroll
                    for (int k = 0; k < G; ++k)
                        if (fl[k]) resolve(J[k], jb + k);
                }
            };
            JRec Pa[G], Pb[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                Pa[k] = jps[jt + k];
                opaque_sgpr(Pa[k]);
            }
            for (int jg = 0; jg < kVTile; jg += 2 * G) {
#pragma unroll
                for (int k = 0; k < G; ++k) Pb[k] = jps[jt + jg + G + k];
                group64(Pa, jt + jg);
#pragma unroll
                for (int k = 0; k < G; ++k) Pa[k] = jps[jt + jg + 2 * G + k];
                group64(Pb, jt + jg + G);
            }
        }
    }

    double sums[7] = {Ax, Ay, Az, Jx, Jy, Jz, Ph};
    slice_partial_store(sums, wv, lane, act,
                        part + ((size_t)s * n_iw * 64 + i) * 8);
}

// ---------------------------------------------------------------------------
// Regular force, pass 1, FP64 pair arithmetic.  On CDNA4 an FP64 FMA issues
// at the rate of an unpacked FP32 FMA, so the whole pair can be evaluated in
// FP64 for about the instruction count of the FP32 form: no conversions and
// no FP32 -> FP64 folds (the sums accumulate in FP64 directly).  The screen
// is the FP32 kernel's min(r2, r2 + 2 dtr dx.dv) in FP64 with a band of a
// few 2^-53 around the threshold; pairs inside it are decided by the
// oracle's exact predicate as in the FP32 kernel.  1/r: v_rsq_f64 refined by
// NR Newton steps.  Measured on MI355X (King W0=6, N = 131072, one regular
// block) this form takes 1.45x the FP32 kernel's time at NR = 0: FP64 FMAs
// issue at the rate of unpacked FP32 ones, but FP64 operands occupy register
// pairs and the FP32 kernel's mixed FP32/FP64/convert stream co-issues better.
// Jerk is accumulated as J1 = sum mr3 dv and J2 = sum mr3 (dx.dv/r2) dx and
// combined as J1 - 3 J2 at the end (one multiply per pair less).
// ---------------------------------------------------------------------------

// One pass-1 task of the FP64 kernel.
template <bool MFLAG, int G, int NR>
__device__ __forceinline__ void regf_pass1_f64_task(
    long long gtask,
    const JRec *__restrict__ jp_all, const ClusterDesc *__restrict__ desc,
    int ncl, long long ntask, const JStats *__restrict__ jst,
    const double *__restrict__ xi_all, const double *__restrict__ vi_all,
    const double *__restrict__ h2_all, const double *__restrict__ dtr_all,
    double *__restrict__ part_all, int *__restrict__ nbc_all,
    int *__restrict__ nbl_all, unsigned *__restrict__ err) {
    const ClusterDesc &D = desc[find_task(desc, ncl, gtask)];
    const int nj = D.nj, ntile = D.ntile, ni = D.ni, n_iw = D.n_iw;
    const int nslice = D.nslice, cap = D.cap;
    const JRec *__restrict__ jp = jp_all + D.j0;
    const double *__restrict__ xi = xi_all + 3 * (size_t)D.i0;
    const double *__restrict__ vi = vi_all + 3 * (size_t)D.i0;
    const double *__restrict__ h2 = h2_all + D.i0;
    const double *__restrict__ dtr = dtr_all + D.i0;
    double *__restrict__ part = part_all + 8 * D.pbase;
    int *__restrict__ nbc = nbc_all + D.ibase;
    int *__restrict__ nbl = nbl_all + D.sbase;
    const int task = (int)(gtask - D.task0);
    int s, iw;
    task_slice(task, n_iw, nslice, ncl == 1, s, iw);
    // Wavefront index (uniform within the wavefront: keeps the j record
    // addresses scalar).
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int i = iw * 64 + lane;
    const bool act = i < ni;

    // Slice s of the j tiles, and this wavefront's contiguous share of it.
    const int sb = (int)(((long long)ntile * s) / nslice);
    const int se = (int)(((long long)ntile * (s + 1)) / nslice);
    const int tb = sb + (int)(((long long)(se - sb) * wv) / kP1W);
    const int te = sb + (int)(((long long)(se - sb) * (wv + 1)) / kP1W);

    double x0 = 0.0, y0 = 0.0, z0 = 0.0, u0 = 0.0, w0 = 0.0, q0 = 0.0;
    double h2d = 0.0, dtrd = 0.0, dtr2x = 0.0;   // dtr2x = 2 dtr
    double thr = -1.0;
    if (act) {
        x0 = xi[3 * (size_t)i + 0];
        y0 = xi[3 * (size_t)i + 1];
        z0 = xi[3 * (size_t)i + 2];
        u0 = vi[3 * (size_t)i + 0];
        w0 = vi[3 * (size_t)i + 1];
        q0 = vi[3 * (size_t)i + 2];
        h2d = h2[i];
        dtrd = dtr[i];
        dtr2x = 2.0 * dtrd;
        const double chk = x0 + y0 + z0 + u0 + w0 + q0 + h2d + dtrd;
        if (chk != chk) atomicOr(err, ERR_NAN_I);
        if (h2d > 0.0) {
            thr = f64_screen_thr<MFLAG>(h2d, dtrd, u0, w0, q0, jst);
        } else {
            thr = 1e-30;
        }
    }

    double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
    double j1x = 0.0, j1y = 0.0, j1z = 0.0, j2x = 0.0, j2y = 0.0, j2z = 0.0;

    // 1/sqrt(r2): v_rsq_f64 (about FP32-level accuracy) refined by NR Newton
    // steps in FP64 (one step reaches ~1e-14 relative).
    auto rsq64 = [&](double r2) {
        double y = __builtin_amdgcn_rsq(r2);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double t = r2 * y;
            const double e = fma(-t, y, 1.0);
            y = fma(y, 0.5 * e, y);
        }
        return y;
    };

    auto pair_terms = [&](const JRec &J, double dx, double dy, double dz,
                          double dvx, double dvy, double dvz, double rv,
                          double rinv) {
        const double rinv2 = rinv * rinv;
        const double mr = J.m * rinv;
        const double mr3 = mr * rinv2;
        const double b = mr3 * (rv * rinv2);
        ph += mr;
        ax = fma(mr3, dx, ax);
        ay = fma(mr3, dy, ay);
        az = fma(mr3, dz, az);
        j1x = fma(mr3, dvx, j1x);
        j1y = fma(mr3, dvy, j1y);
        j1z = fma(mr3, dvz, j1z);
        j2x = fma(b, dx, j2x);
        j2y = fma(b, dy, j2y);
        j2z = fma(b, dz, j2z);
    };

    auto resolve = [&](const JRec &J, int j) {
        const double dx = J.x - x0, dy = J.y - y0, dz = J.z - z0;
        const double dvx = J.vx - u0, dvy = J.vy - w0, dvz = J.vz - q0;
        const double r2d = sq3d(dx, dy, dz);
        const double pxd = dx + dtrd * dvx;
        const double pyd = dy + dtrd * dvy;
        const double pzd = dz + dtrd * dvz;
        const double r2pd = sq3d(pxd, pyd, pzd);
        const double mnd = r2d < r2pd ? r2d : r2pd;
        const double H = MFLAG ? J.m * h2d : h2d;
        if ((mnd < H) && (j < nj)) {
            const int k2 = atomicAdd(&nbc[i], 1);
            if (k2 < cap) nbl[(size_t)i * cap + k2] = j;
        } else if (r2d != 0.0) {
            // Rare (band pairs, r2 < 1e-30): exact FP64 1/sqrt.
            const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
            pair_terms(J, dx, dy, dz, dvx, dvy, dvz, rv, 1.0 / sqrt(r2d));
        }
    };

    // The records of the next group are loaded while the current group is
    // evaluated, hiding the scalar-load latency inside the wave (5% faster
    // per launch).  The j array carries 64 spare records past its last tile,
    // so the look-ahead of the last group stays in bounds.
    JRec Jn[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        Jn[k] = jp[tb * kTile + k];
        opaque_sgpr(Jn[k]);
    }
    for (int t = tb; t < te; ++t) {
        const int jt = t * kTile;
        for (int jg = 0; jg < kTile; jg += G) {
            JRec J[G];
            bool fl[G];
            bool any = false;
#pragma unroll
            for (int k = 0; k < G; ++k) J[k] = Jn[k];
#pragma unroll
            for (int k = 0; k < G; ++k) Jn[k] = jp[jt + jg + G + k];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const double dx = J[k].x - x0;
                const double dy = J[k].y - y0;
                const double dz = J[k].z - z0;
                const double dvx = J[k].vx - u0;
                const double dvy = J[k].vy - w0;
                const double dvz = J[k].vz - q0;
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
                const double q = fma(dtr2x, rv, r2);
                double tj = thr;
                if (MFLAG) tj = fmax(thr * J[k].m, 1e-30);
                fl[k] = fmin(r2, q) < tj;
                any = any || fl[k];
                const double rinv = fl[k] ? 0.0 : rsq64(r2);
                pair_terms(J[k], dx, dy, dz, dvx, dvy, dvz, rv, rinv);
            }
            if (any) {
#pragma unroll
                for (int k = 0; k < G; ++k)
                    if (fl[k]) resolve(J[k], jt + jg + k);
            }
        }
    }

    double sums[7] = {ax, ay, az, fma(-3.0, j2x, j1x), fma(-3.0, j2y, j1y),
                      fma(-3.0, j2z, j1z), ph};
    slice_partial_store(sums, wv, lane, act,
                        part + ((size_t)s * n_iw * 64 + i) * 8);
}


// Block-to-task map of a plain launch: workgroups are dispatched round-robin
// over the 8 XCDs, so each XCD is given a contiguous range of tasks (what a
// range covers: task_slice).  -1: no task for this workgroup.
__device__ inline long long xcd_task(long long ntask) {
    const long long per = (ntask + 7) >> 3;
    const long long t = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    return t < ntask ? t : -1;
}

// Next task of a persistent workgroup.  Tasks [xr[x], xr[x + 1]) belong to
// XCD x (the workgroup's XCD is blockIdx.x & 7: workgroups of a grid that
// fits the machine are dispatched round-robin over the XCDs) and are taken in
// order through the counter q[x]; once its own range is used up a workgroup
// takes tasks from the following XCDs' ranges.  Returns -1 when every range
// is used up, so every workgroup ends.  Thread 0 takes the task; x and tries
// are its state across calls.
__device__ inline long long next_task(unsigned *__restrict__ q,
                                      const long long *__restrict__ xr, int &x,
                                      int &tries) {
    __shared__ long long s_task;
    __syncthreads();   // every thread has read the previous task
    if (threadIdx.x == 0) {
        long long t = -1;
        while (tries < 8) {
            const long long lo = xr[x], hi = xr[x + 1];
            if (lo < hi) {
                const long long k = (long long)atomicAdd(&q[x], 1u);
                if (lo + k < hi) {
                    t = lo + k;
                    break;
                }
            }
            x = (x + 1) & 7;
            ++tries;
        }
        s_task = t;
    }
    __syncthreads();
    return s_task;
}

template <bool MFLAG, int G>
__global__ __launch_bounds__(64 * kP1W) void k_regf_pass1(
    const JRec32 *__restrict__ j32_all, const JTile *__restrict__ tile_all,
    const JRec *__restrict__ jps_all, const JClus *__restrict__ jcl_all,
    const ClusterDesc *__restrict__ desc, int ncl, long long ntask,
    const JStats *__restrict__ jst, const double *__restrict__ xi_all,
    const double *__restrict__ vi_all, const double *__restrict__ h2_all,
    const double *__restrict__ dtr_all, double *__restrict__ part_all,
    int *__restrict__ nbc_all, int *__restrict__ nbl_all,
    unsigned *__restrict__ err) {
    const long long t = xcd_task(ntask);
    if (t < 0) return;
    regf_pass1_task<MFLAG, G>(t, j32_all, tile_all, jps_all, jcl_all, desc, ncl, ntask, jst, xi_all, vi_all,
        h2_all, dtr_all, part_all, nbc_all, nbl_all, err);
}

// Persistent form: a grid of resident workgroups takes tasks from the XCD
// queues until all are done (the ensemble's uneven tasks, largest first).
template <bool MFLAG, int G>
__global__ __launch_bounds__(64 * kP1W) void k_regf_pass1_q(
    const JRec32 *__restrict__ j32_all, const JTile *__restrict__ tile_all,
    const JRec *__restrict__ jps_all, const JClus *__restrict__ jcl_all,
    const ClusterDesc *__restrict__ desc, int ncl, long long ntask,
    const JStats *__restrict__ jst, const double *__restrict__ xi_all,
    const double *__restrict__ vi_all, const double *__restrict__ h2_all,
    const double *__restrict__ dtr_all, double *__restrict__ part_all,
    int *__restrict__ nbc_all, int *__restrict__ nbl_all,
    unsigned *__restrict__ err, unsigned *__restrict__ q, const long long *__restrict__ xr) {
    int x = blockIdx.x & 7, tries = 0;
    for (;;) {
        const long long t = next_task(q, xr, x, tries);
        if (t < 0) return;
        regf_pass1_task<MFLAG, G>(t, j32_all, tile_all, jps_all, jcl_all, desc, ncl, ntask, jst, xi_all, vi_all,
        h2_all, dtr_all, part_all, nbc_all, nbl_all, err);
    }
}

template <bool MFLAG, int G, int NR>
__global__ __launch_bounds__(64 * kP1W) void k_regf_pass1_f64(
    const JRec *__restrict__ jp_all, const ClusterDesc *__restrict__ desc,
    int ncl, long long ntask, const JStats *__restrict__ jst,
    const double *__restrict__ xi_all, const double *__restrict__ vi_all,
    const double *__restrict__ h2_all, const double *__restrict__ dtr_all,
    double *__restrict__ part_all, int *__restrict__ nbc_all,
    int *__restrict__ nbl_all, unsigned *__restrict__ err) {
    const long long t = xcd_task(ntask);
    if (t < 0) return;
    regf_pass1_f64_task<MFLAG, G, NR>(t, jp_all, desc, ncl, ntask, jst, xi_all, vi_all, h2_all, dtr_all,
        part_all, nbc_all, nbl_all, err);
}

template <bool MFLAG, int G, int NR>
__global__ __launch_bounds__(64 * kP1W) void k_regf_pass1_f64_q(
    const JRec *__restrict__ jp_all, const ClusterDesc *__restrict__ desc,
    int ncl, long long ntask, const JStats *__restrict__ jst,
    const double *__restrict__ xi_all, const double *__restrict__ vi_all,
    const double *__restrict__ h2_all, const double *__restrict__ dtr_all,
    double *__restrict__ part_all, int *__restrict__ nbc_all,
    int *__restrict__ nbl_all, unsigned *__restrict__ err, unsigned *__restrict__ q, const long long *__restrict__ xr) {
    int x = blockIdx.x & 7, tries = 0;
    for (;;) {
        const long long t = next_task(q, xr, x, tries);
        if (t < 0) return;
        regf_pass1_f64_task<MFLAG, G, NR>(t, jp_all, desc, ncl, ntask, jst, xi_all, vi_all, h2_all, dtr_all,
        part_all, nbc_all, nbl_all, err);
    }
}

// ---------------------------------------------------------------------------
// Regular force, pass 2: a workgroup of 4 wavefronts serves ipb <= kP2I
// i-particles (ipb chosen per launch so that small calls still spread over
// the machine), each wavefront every 4th of them.  Per i: the slice partials
// are summed in a fixed order (deterministic FP64) and the forces written,
// and the list count read (and reset for the next launch on this scratch);
// then one thread scans the block's list lengths and claims the block's list
// space with a single atomic; then every i's list entries (sorted positions,
// appended by the slices in any order) are mapped to original j indices in
// the wavefront's LDS row and each written at its rank among them, so the
// list comes out ascending whatever the append order was; the nearest
// neighbour (smallest FP64 r2 > 0, ties to the smaller j) is found on the
// way.
// ---------------------------------------------------------------------------

constexpr int kP2I = 32;
constexpr int kP2U = 4;   // 64-slice chunks per pass-2 load batch

__global__ __launch_bounds__(256) void k_regf_pass2(
    const JRec *__restrict__ jps_all, const int *__restrict__ jidx_all,
    const ClusterDesc *__restrict__ desc, int ncl,
    const double *__restrict__ xi_all, const double *__restrict__ part_all,
    int *__restrict__ nbc_all, const int *__restrict__ nbl_all,
    double *__restrict__ acc_all, double *__restrict__ jrk_all,
    double *__restrict__ pot_all, int *__restrict__ nnb_all,
    int *__restrict__ nnid_all, int *__restrict__ list,
    int *__restrict__ off_all, int *__restrict__ total, long long list_cap,
    int nnbmax, int sortcap, const int *__restrict__ perm,
    unsigned *__restrict__ err) {
    __shared__ int s_tot[kP2I];
    __shared__ int s_off[kP2I];
    extern __shared__ int s_rows[];   // [4][sortcap]: one list row per wave
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const ClusterDesc &D = desc[find_blk(desc, ncl, blockIdx.x)];
    const int ni = D.ni, nslice = D.nslice, cap = D.cap;
    const int ipb = D.ipb;
    const int i0 = (int)(blockIdx.x - D.blk0) * ipb;
    const JRec *__restrict__ jps = jps_all + D.j0;
    // NULL: the j records are in the caller's order (position = index).
    const int *__restrict__ jidx = jidx_all ? jidx_all + D.j0 : nullptr;
    const double *__restrict__ xi = xi_all + 3 * (size_t)D.i0;
    const double *__restrict__ part = part_all + 8 * D.pbase;
    int *__restrict__ nbc = nbc_all + D.ibase;
    const int *__restrict__ nbl = nbl_all + D.sbase;
    // Output row of local i: the caller's index (through perm when the i
    // were sorted).
    auto orow = [&](int i) -> size_t {
        const int g = D.i0 + i;
        return (size_t)(perm ? perm[g] : g);
    };

    for (int r = wave; r < ipb; r += 4) {
        const int i = i0 + r;
        int tot = 0;
        if (i < ni) {
            // kP2U chunks of 64 slices per batch: their partials are loaded
            // together, then added in slice order per lane.
            double sum[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int b0 = 0; b0 < nslice; b0 += 64 * kP2U) {
                double v[kP2U][7];
#pragma unroll
                for (int u = 0; u < kP2U; ++u) {
                    const int q = b0 + 64 * u + lane;
                    if (q < nslice) {
                        const double *p =
                            part + ((size_t)q * D.n_iw * 64 + i) * 8;
#pragma unroll
                        for (int k = 0; k < 7; ++k) v[u][k] = p[k];
                    } else {
#pragma unroll
                        for (int k = 0; k < 7; ++k) v[u][k] = 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < kP2U; ++u) {
                    if (b0 + 64 * u + lane < nslice) {
#pragma unroll
                        for (int k = 0; k < 7; ++k) sum[k] += v[u][k];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 7; ++k)
                for (int o = 32; o >= 1; o >>= 1)
                    sum[k] += __shfl_xor(sum[k], o);
            tot = nbc[i];
            if (lane == 0) {
                const size_t q = orow(i);
                acc_all[3 * q + 0] = sum[0];
                acc_all[3 * q + 1] = sum[1];
                acc_all[3 * q + 2] = sum[2];
                jrk_all[3 * q + 0] = sum[3];
                jrk_all[3 * q + 1] = sum[4];
                jrk_all[3 * q + 2] = sum[5];
                pot_all[q] = sum[6];
                nnb_all[q] = (tot <= nnbmax) ? tot : -tot;
                nbc[i] = 0;
            }
        }
        if (lane == 0) s_tot[r] = tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int r = 0; r < ipb; ++r) {
            const int t = s_tot[r];
            s_off[r] = run;
            if (t <= nnbmax) run += t;
        }
        int base = run > 0 ? atomicAdd(total, run) : 0;
        if ((long long)base + run > list_cap) {
            atomicOr(err, ERR_LIST_CAP);
            base = -1;
        }
        for (int r = 0; r < ipb; ++r)
            s_off[r] = (base < 0 || s_tot[r] > nnbmax) ? -1 : base + s_off[r];
    }
    __syncthreads();

    int *__restrict__ row = s_rows + wave * sortcap;
    for (int r = wave; r < ipb; r += 4) {
        const int i = i0 + r;
        if (i >= ni) break;
        const int o = s_off[r];
        int bestj = -1;
        if (o >= 0) {
            const int n = s_tot[r];
            const int *__restrict__ ent = nbl + (size_t)i * cap;
            const double xa = xi[3 * (size_t)i + 0];
            const double ya = xi[3 * (size_t)i + 1];
            const double za = xi[3 * (size_t)i + 2];
            double br2 = INFINITY;
            int bj = INT_MAX;
            for (int e = lane; e < n; e += 64) {
                const int p = ent[e];
                const int j = jidx ? jidx[p] : p;
                row[e] = j;
                // The nearest neighbour only when the caller asks for it:
                // it costs a record gather per list entry.
                if (nnid_all) {
                    const JRec &J = jps[p];
                    const double r2 = sq3d(J.x - xa, J.y - ya, J.z - za);
                    if (r2 > 0.0 && (r2 < br2 || (r2 == br2 && j < bj))) {
                        br2 = r2;
                        bj = j;
                    }
                }
            }
            // Every lane of this wave reads the whole row below: the row's
            // LDS writes complete first (the lanes run in step).
            __threadfence_block();
            for (int e = lane; e < n; e += 64) {
                const int v = row[e];
                int rank = 0;
                for (int k = 0; k < n; ++k) rank += row[k] < v ? 1 : 0;
                list[o + rank] = v;
            }
            __threadfence_block();
            // Lexicographic (r2, j) minimum over the wave.
            for (int w = 32; w >= 1; w >>= 1) {
                const double r2o = __shfl_xor(br2, w);
                const int jo = __shfl_xor(bj, w);
                if (r2o < br2 || (r2o == br2 && jo < bj)) {
                    br2 = r2o;
                    bj = jo;
                }
            }
            bestj = (bj == INT_MAX) ? -1 : bj;
        }
        if (lane == 0) {
            const size_t q = orow(i);
            off_all[q] = o;
            if (nnid_all) nnid_all[q] = bestj;
        }
    }
}

// ---------------------------------------------------------------------------
// Potential: phi_i = sum_{j: r2 > 0} m_j / r_ij.  FP64 differences, FP32
// r2 / rsqrt / products, FP32 tile sums folded into FP64.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_pack_p(const double *__restrict__ m,
                                                const double *__restrict__ x,
                                                int n, int n_pad,
                                                PRec *__restrict__ p,
                                                unsigned *__restrict__ err) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        PRec r;
        r.x = x[3 * (size_t)j + 0];
        r.y = x[3 * (size_t)j + 1];
        r.z = x[3 * (size_t)j + 2];
        const double mj = m[j];
        r.m32 = (float)mj;
        r.pad = 0.0f;
        const double chk = r.x + r.y + r.z + mj;
        if (chk != chk) atomicOr(err, ERR_NAN_J);
        p[j] = r;
    } else if (j < n_pad) {
        PRec r;
        r.x = r.y = r.z = kPadPos;
        r.m32 = 0.0f;
        r.pad = 0.0f;
        p[j] = r;
    }
}

__global__ __launch_bounds__(64) void k_pot_pass1(const PRec *__restrict__ p,
                                                  int ntile, int i0, int ni,
                                                  int n_iw, int nslice,
                                                  double *__restrict__ part) {
    const int ntask = n_iw * nslice;
    const int per = (ntask + 7) >> 3;
    const int task = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (task >= ntask) return;
    const int s = task / n_iw;
    const int iw = task - s * n_iw;
    const int q = iw * 64 + threadIdx.x;
    const bool act = q < ni;
    const int tb = (int)(((long long)ntile * s) / nslice);
    const int te = (int)(((long long)ntile * (s + 1)) / nslice);
    double x0 = 0.0, y0 = 0.0, z0 = 0.0;
    if (act) {
        const PRec r = p[i0 + q];
        x0 = r.x;
        y0 = r.y;
        z0 = r.z;
    }
    float a = 0.f;
    double A = 0.0;
    // Groups of four records; the next group is loaded while the current one
    // is evaluated (the record array carries a spare tile past its end, so
    // the last look-ahead stays in bounds).  Terms are added in j order.
    constexpr int GP = 4;
    PRec Jn[GP];
#pragma unroll
    for (int k = 0; k < GP; ++k) Jn[k] = p[tb * kTile + k];
    for (int t = tb; t < te; ++t) {
        const int jt = t * kTile;
        for (int jg = 0; jg < kTile; jg += GP) {
            PRec J[GP];
#pragma unroll
            for (int k = 0; k < GP; ++k) J[k] = Jn[k];
#pragma unroll
            for (int k = 0; k < GP; ++k) Jn[k] = p[jt + jg + GP + k];
#pragma unroll
            for (int k = 0; k < GP; ++k) {
                const float fx = (float)(J[k].x - x0);
                const float fy = (float)(J[k].y - y0);
                const float fz = (float)(J[k].z - z0);
                const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                const float term = J[k].m32 * __builtin_amdgcn_rsqf(r2);
                a += (r2 > 0.0f) ? term : 0.0f;
            }
        }
        A += (double)a;
        a = 0.f;
    }
    if (act) part[(size_t)q * nslice + s] = A;
}

__global__ __launch_bounds__(256) void k_pot_pass2(const double *__restrict__ part,
                                                   int ni, int nslice,
                                                   double *__restrict__ pot) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ni) return;
    double s = 0.0;
    for (int k = 0; k < nslice; ++k) s += part[(size_t)q * nslice + k];
    pot[q] = s;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

void launch_stats_reset(hipStream_t st, JStats *jst, int *total) {
    hipLaunchKernelGGL(k_stats_reset, dim3(1), dim3(64), 0, st, jst, total);
}

void launch_pack_j(hipStream_t st, const double *m, const double *x,
                   const double *v, int nj, int nj_pad, JRec *jp, JStats *jst,
                   unsigned *err) {
    if (nj_pad <= 0) return;
    hipLaunchKernelGGL(k_pack_j, dim3(cdiv(nj_pad, 256)), dim3(256), 0, st, m,
                       x, v, nj, nj_pad, jp, jst, err);
}

void launch_predict(hipStream_t st, double t, int nj, int nj_pad,
                    const double *x0, const double *v0, const double *f,
                    const double *fd, const double *t0, const double *m,
                    JRec *jp, JStats *jst) {
    if (nj_pad <= 0) return;
    hipLaunchKernelGGL(k_predict, dim3(cdiv(nj_pad, 256)), dim3(256), 0, st, t,
                       nj, nj_pad, x0, v0, f, fd, t0, m, jp, jst);
}

void launch_jpset(hipStream_t st, int n, const int *addr, const double *src,
                  double *x0, double *v0, double *f, double *fd, double *mass,
                  double *t0, unsigned *err) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jpset, dim3(cdiv(n, 256)), dim3(256), 0, st, n, addr,
                       src, x0, v0, f, fd, mass, t0, err);
}

void launch_jget(hipStream_t st, int n, const JRec *jp, double *x, double *v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jget, dim3(cdiv(n, 256)), dim3(256), 0, st, n, jp, x,
                       v);
}

void launch_copy_to_host(hipStream_t st, const void *src, void *dst,
                         size_t bytes, const unsigned *err, unsigned *herr) {
    const size_t n16 = bytes / 16;
    const int nb = (int)std::min<size_t>(1024, std::max<size_t>(1, (n16 + 255) / 256));
    hipLaunchKernelGGL(k_copy_to_host, dim3(nb), dim3(256), 0, st,
                       (const uint4 *)src, (uint4 *)dst, n16, err, herr);
}

void launch_copy_from_host(hipStream_t st, const void *src, void *dst,
                           size_t bytes, int *zero) {
    const size_t n16 = bytes / 16;
    const int nb = (int)std::min<size_t>(1024, std::max<size_t>(1, (n16 + 255) / 256));
    hipLaunchKernelGGL(k_copy_from_host, dim3(nb), dim3(256), 0, st,
                       (const uint4 *)src, (uint4 *)dst, n16, zero);
}

void launch_set_desc(hipStream_t st, ClusterDesc *d, const ClusterDesc &v) {
    hipLaunchKernelGGL(k_set_desc, dim3(1), dim3(1), 0, st, d, v);
}

void launch_zero_int(hipStream_t st, int *p, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_zero_int, dim3(cdiv(n, 256)), dim3(256), 0, st, p, n);
}

template <bool MFLAG, int G>
static void launch_pass1(hipStream_t st, const RegfLaunch &L, long long nblk) {
    if (L.queue) {
        hipLaunchKernelGGL((k_regf_pass1_q<MFLAG, G>), dim3((unsigned)L.pgrid),
                           dim3(64 * kP1W), 0, st, L.j32, L.tiles, L.jps, L.jcl,
                           L.desc, L.ncl, L.ntask, L.jst, L.xi, L.vi, L.h2,
                           L.dtr, L.part, L.nbc, L.nbl, L.err, L.queue,
                           L.xrange);
        return;
    }
    hipLaunchKernelGGL((k_regf_pass1<MFLAG, G>), dim3((unsigned)nblk),
                       dim3(64 * kP1W),
                       0, st, L.j32, L.tiles, L.jps, L.jcl, L.desc, L.ncl,
                       L.ntask, L.jst, L.xi, L.vi, L.h2, L.dtr, L.part, L.nbc,
                       L.nbl, L.err);
}

template <bool MFLAG, int G, int NR>
static void launch_pass1_f64(hipStream_t st, const RegfLaunch &L,
                             long long nblk) {
    if (L.queue) {
        hipLaunchKernelGGL((k_regf_pass1_f64_q<MFLAG, G, NR>),
                           dim3((unsigned)L.pgrid), dim3(64 * kP1W), 0, st,
                           L.jps, L.desc, L.ncl, L.ntask, L.jst, L.xi, L.vi,
                           L.h2, L.dtr, L.part, L.nbc, L.nbl, L.err, L.queue,
                           L.xrange);
        return;
    }
    hipLaunchKernelGGL((k_regf_pass1_f64<MFLAG, G, NR>), dim3((unsigned)nblk),
                       dim3(64 * kP1W), 0, st, L.jps, L.desc, L.ncl, L.ntask, L.jst,
                       L.xi, L.vi, L.h2, L.dtr, L.part, L.nbc, L.nbl, L.err);
}

void launch_regf_pass1(hipStream_t st, const RegfLaunch &L) {
    if (L.ntask <= 0) return;
    const long long nblk = 8 * ((L.ntask + 7) / 8);
    if (L.queue) launch_zero_int(st, (int *)L.queue, 8);
    if (L.prec >= 64) {
        const int nr = L.prec - 64;   // 64: no Newton step, 65: one step
        if (L.group == 2) {
            if (L.m_flag) nr ? launch_pass1_f64<true, 2, 1>(st, L, nblk)
                             : launch_pass1_f64<true, 2, 0>(st, L, nblk);
            else nr ? launch_pass1_f64<false, 2, 1>(st, L, nblk)
                    : launch_pass1_f64<false, 2, 0>(st, L, nblk);
        } else {
            if (L.m_flag) nr ? launch_pass1_f64<true, 4, 1>(st, L, nblk)
                             : launch_pass1_f64<true, 4, 0>(st, L, nblk);
            else nr ? launch_pass1_f64<false, 4, 1>(st, L, nblk)
                    : launch_pass1_f64<false, 4, 0>(st, L, nblk);
        }
        return;
    }
    if (L.group == 2) {
        if (L.m_flag) launch_pass1<true, 2>(st, L, nblk);
        else launch_pass1<false, 2>(st, L, nblk);
    } else {
        if (L.m_flag) launch_pass1<true, 4>(st, L, nblk);
        else launch_pass1<false, 4>(st, L, nblk);
    }
}

void launch_regf_pass2(hipStream_t st, const RegfLaunch &L) {
    if (L.nblk2 <= 0) return;
    const int sortcap = pass2_sortcap(L.nnbmax);
    hipLaunchKernelGGL(k_regf_pass2, dim3((unsigned)L.nblk2), dim3(256),
                       4 * sortcap * sizeof(int), st, L.jps, L.jidx, L.desc,
                       L.ncl, L.xi, L.part, L.nbc, L.nbl, L.acc, L.jrk, L.pot,
                       L.nnb, L.nnid, L.list, L.off, L.total, L.list_cap,
                       L.nnbmax, sortcap, L.perm, L.err);
}

int pass2_sortcap(int nnbmax) {
    return std::max(64, (nnbmax + 64) / 64 * 64);
}

int pass2_ipb(int ni) {
    // About 2048 workgroups for large calls (one list-space atomic per 16 i),
    // one i per wavefront for small ones.
    return std::max(4, std::min(kP2I, ni / 2048));
}

void launch_pack_p(hipStream_t st, const double *m, const double *x, int n,
                   int n_pad, PRec *p, unsigned *err) {
    if (n_pad <= 0) return;
    hipLaunchKernelGGL(k_pack_p, dim3(cdiv(n_pad, 256)), dim3(256), 0, st, m, x,
                       n, n_pad, p, err);
}

void launch_pot(hipStream_t st, const PRec *p, int n, int ntile, int i0,
                int ni, int nslice, double *part, double *pot) {
    (void)n;
    if (ni <= 0) return;
    const int n_iw = cdiv(ni, 64);
    const int ntask = n_iw * nslice;
    const int nblk = 8 * ((ntask + 7) / 8);
    hipLaunchKernelGGL(k_pot_pass1, dim3(nblk), dim3(64), 0, st, p, ntile, i0,
                       ni, n_iw, nslice, part);
    hipLaunchKernelGGL(k_pot_pass2, dim3(cdiv(ni, 256)), dim3(256), 0, st, part,
                       ni, nslice, pot);
}

// Resident wavefronts of a kernel launched in blocks of `block` threads.
static int resident_waves_of(const void *fn, int block = 64) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0) !=
        hipSuccess)
        return 0;
    return nb * (block / 64) * prop.multiProcessorCount;
}

int regf_resident_waves(int m_flag, int group, int prec) {
    if (prec >= 64) {
        if (group == 2)
            return m_flag ? resident_waves_of((const void *)k_regf_pass1_f64<true, 2, 1>, 64 * kP1W)
                          : resident_waves_of((const void *)k_regf_pass1_f64<false, 2, 1>, 64 * kP1W);
        return m_flag ? resident_waves_of((const void *)k_regf_pass1_f64<true, 4, 1>, 64 * kP1W)
                      : resident_waves_of((const void *)k_regf_pass1_f64<false, 4, 1>, 64 * kP1W);
    }
    if (group == 2)
        return m_flag ? resident_waves_of((const void *)k_regf_pass1<true, 2>, 64 * kP1W)
                      : resident_waves_of((const void *)k_regf_pass1<false, 2>, 64 * kP1W);
    return m_flag ? resident_waves_of((const void *)k_regf_pass1<true, 4>, 64 * kP1W)
                  : resident_waves_of((const void *)k_regf_pass1<false, 4>, 64 * kP1W);
}

int pot_resident_waves() { return resident_waves_of((const void *)k_pot_pass1); }

}  // namespace gpunb
