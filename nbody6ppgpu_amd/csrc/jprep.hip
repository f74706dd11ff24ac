// jprep.hip -- the sorted j view of the regular-force engine.
//
// After every send / predict the j set is reordered along a Hilbert curve, so
// that 64 consecutive sorted j form a spatially compact tile, and each j is
// stored as an FP32 offset from its tile's FP64 centre (JRec32).  The pair
// loop then forms x_j - x_i as (C_t - x_i) + o_j: the first term once per
// (i, tile) in FP64 and rounded, the second one FP32 add per pair, instead of
// an FP64 subtraction and a conversion per pair component.
//
// Steps (all on the library stream, deterministic):
//   k_jstat1   per block of kStatChunk j: bounding box and mass sum, reduced
//              in a fixed order
//   k_jstat2   per cluster: the blocks' partials in block order -> JClus
//              (bounding box, mean mass)
//   k_jkey     64-bit keys (cluster << 32 | heavy << 30 | 30-bit Hilbert index
//              of the position in the cluster's box; padding records sort
//              last)
//   rocprim radix sort of (key, original index): stable, so equal codes keep
//              their index order
//   k_jtile    one wavefront per tile: sorted FP64 records, sorted -> original
//              index map, tile centres and offset bounds, FP32 records
#include "gpunb_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <cfloat>
#include <cmath>

namespace gpunb {
namespace {

__device__ inline double wmin(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wmax(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
// The same within each kVTile-lane segment of the wavefront.
__device__ inline double smin(double v) {
    for (int o = kVTile / 2; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double smax(double v) {
    for (int o = kVTile / 2; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// FP32 value >= v (v finite, >= 0).
__device__ inline float up32(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

// Cluster owning position g of the concatenated ranges (last c with
// pc[c].j0 <= g).
__device__ inline int clus_of(const PrepClus *__restrict__ pc, int ncl, int g) {
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pc[mid].j0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__device__ inline int clus_of_blk(const PrepClus *__restrict__ pc, int ncl,
                                  int b) {
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pc[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// 10 bits spread to every third position.
__device__ inline unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ inline unsigned cell10(double x, double lo, double scale) {
    const double q = (x - lo) * scale;
    if (!(q > 0.0)) return 0u;           // also NaN
    if (q >= 1023.0) return 1023u;
    return (unsigned)q;
}

// Index of cell (x, y, z) (10 bits each) along a 3-D Hilbert curve of the
// 1024^3 grid, 30 bits: J. Skilling, "Programming the Hilbert curve", AIP
// Conf. Proc. 707 (2004) 381 -- the axes are turned into the curve's
// "transpose" form (undo the excess work of each level, then Gray-encode),
// whose bits interleaved from the top are the index.  Consecutive indices are
// adjacent cells, so 64 consecutive bodies form a more compact tile than a
// Morton run, which jumps at every power-of-two cell boundary (King N = 131072:
// median tile extent 1.71 against 2.00 neighbour radii).
__device__ inline unsigned hilbert30(unsigned x, unsigned y, unsigned z) {
    unsigned X[3] = {x, y, z};
    for (unsigned Q = 1u << 9; Q > 1u; Q >>= 1) {
        const unsigned P = Q - 1u;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    unsigned t = 0u;
    for (unsigned Q = 1u << 9; Q > 1u; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1u;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    return (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
}

// Curve key of a position in the cube of a cluster's bounding box.
__device__ inline unsigned cell_key(double x, double y, double z,
                                    const JClus &J) {
    const double ext = fmax(fmax(J.hi[0] - J.lo[0], J.hi[1] - J.lo[1]),
                            J.hi[2] - J.lo[2]);
    const double scale = ext > 0.0 ? 1024.0 / ext : 0.0;
    return hilbert30(cell10(x, J.lo[0], scale), cell10(y, J.lo[1], scale),
                     cell10(z, J.lo[2], scale));
}

// Fixed-order tree reduction of 8 doubles over a 256-thread workgroup
// (components 0-2 min, 3-5 max, 6-7 sum).
__device__ inline void block_reduce8(double v[8], double (*sh)[8]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[t][k] = v[k];
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < 3; ++k) sh[t][k] = fmin(sh[t][k], sh[t + s][k]);
#pragma unroll
            for (int k = 3; k < 6; ++k) sh[t][k] = fmax(sh[t][k], sh[t + s][k]);
            sh[t][6] += sh[t + s][6];
            sh[t][7] += sh[t + s][7];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = sh[0][k];
}

}  // namespace

__global__ __launch_bounds__(256) void k_jstat1(const JRec *__restrict__ jp,
                                                const PrepClus *__restrict__ pc,
                                                int ncl,
                                                double *__restrict__ part) {
    __shared__ double sh[256][8];
    const int b = blockIdx.x;
    const PrepClus P = pc[clus_of_blk(pc, ncl, b)];
    const int beg = P.j0 + (b - P.blk0) * kStatChunk;
    const int end = min(beg + kStatChunk, P.j0 + P.nj);
    double v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY,
                   -INFINITY, 0.0, 0.0};
    for (int j = beg + (int)threadIdx.x; j < end; j += 256) {
        const JRec &r = jp[j];
        v[0] = fmin(v[0], r.x);
        v[1] = fmin(v[1], r.y);
        v[2] = fmin(v[2], r.z);
        v[3] = fmax(v[3], r.x);
        v[4] = fmax(v[4], r.y);
        v[5] = fmax(v[5], r.z);
        v[6] += r.m;
    }
    block_reduce8(v, sh);
    if (threadIdx.x < 8) part[8 * (size_t)b + threadIdx.x] = v[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_jstat2(const PrepClus *__restrict__ pc,
                                                int ncl, int nblk,
                                                const double *__restrict__ part,
                                                JClus *__restrict__ jcl) {
    __shared__ double sh[256][8];
    const int c = blockIdx.x;
    const int b0 = pc[c].blk0;
    const int b1 = (c + 1 < ncl) ? pc[c + 1].blk0 : nblk;
    double v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY,
                   -INFINITY, 0.0, 0.0};
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 256) {
        const double *p = part + 8 * (size_t)b;
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = fmin(v[k], p[k]);
#pragma unroll
        for (int k = 3; k < 6; ++k) v[k] = fmax(v[k], p[k]);
        v[6] += p[6];
    }
    block_reduce8(v, sh);
    if (threadIdx.x == 0) {
        JClus J;
        const int nj = pc[c].nj;
        for (int k = 0; k < 3; ++k) {
            J.lo[k] = nj > 0 ? v[k] : 0.0;
            J.hi[k] = nj > 0 ? v[3 + k] : 0.0;
        }
        J.mbar = nj > 0 ? v[6] / nj : 0.0;
        J.pad = 0.0;
        jcl[c] = J;
    }
}

__global__ __launch_bounds__(256) void k_jkey(
    const JRec *__restrict__ jp, const PrepClus *__restrict__ pc, int ncl,
    const JClus *__restrict__ jcl, int n_all, unsigned long long *__restrict__ key,
    int *__restrict__ val) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_all) return;
    const int c = clus_of(pc, ncl, g);
    const PrepClus P = pc[c];
    const JClus &J = jcl[c];
    unsigned code = 0xffffffffu;       // padding records sort last
    if (g - P.j0 < P.nj) {
        const JRec &r = jp[g];
        code = cell_key(r.x, r.y, r.z, J);
        // Heavy bodies (m > kHeavy <m>) sort after the light ones, in tiles
        // of their own: their band weight w_j would otherwise widen the
        // near region of every tile that holds one.
        if (r.m > (double)kHeavy * J.mbar) code |= 1u << 30;
    }
    key[g] = ((unsigned long long)c << 32) | code;
    val[g] = g;
}

__global__ __launch_bounds__(64) void k_jtile(
    const JRec *__restrict__ jp, const PrepClus *__restrict__ pc, int ncl,
    const JClus *__restrict__ jcl, const int *__restrict__ sorted,
    JRec *__restrict__ jps, JRec32 *__restrict__ j32,
    JTile *__restrict__ tiles, int *__restrict__ jidx) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const int p = t * kTile + lane;
    const int c = clus_of(pc, ncl, t * kTile);
    const PrepClus P = pc[c];
    const int g = sorted[p];
    const int local = g - P.j0;
    const bool real = local < P.nj;
    const JRec r = jp[g];
    jps[p] = r;
    jidx[p] = local;

    // The wavefront's lanes hold kTile / kVTile view tiles, one per
    // kVTile-lane segment: every reduction below stays within a segment.
    double v[6] = {r.x, r.y, r.z, r.vx, r.vy, r.vz};
    double cen[6];
    const unsigned long long seg_mask =
        kVTile >= 64 ? ~0ull : ((1ull << kVTile) - 1ull);
    const bool any =
        ((__ballot(real) >> (lane & ~(kVTile - 1))) & seg_mask) != 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double lo = smin(real ? v[k] : INFINITY);
        const double hi = smax(real ? v[k] : -INFINITY);
        cen[k] = any ? 0.5 * lo + 0.5 * hi : 0.0;
    }
    JRec32 q;
    double e = 0.0, ev = 0.0, w = 1.0;
    if (real) {
        q.ox = (float)(r.x - cen[0]);
        q.oy = (float)(r.y - cen[1]);
        q.oz = (float)(r.z - cen[2]);
        q.ovx = (float)(r.vx - cen[3]);
        q.ovy = (float)(r.vy - cen[4]);
        q.ovz = (float)(r.vz - cen[5]);
        q.m = r.m32;
        const double ox = q.ox, oy = q.oy, oz = q.oz;
        const double ux = q.ovx, uy = q.ovy, uz = q.ovz;
        e = sqrt(ox * ox + oy * oy + oz * oz) * (1.0 + 1e-12);
        ev = sqrt(ux * ux + uy * uy + uz * uz) * (1.0 + 1e-12);
        const double mbar = jcl[c].mbar;
        if (mbar > 0.0 && r.m > mbar) {
            const double x = r.m / mbar;
            w = fmax(fmax(1.0, cbrt(x * x)), 1.4 * x / (double)kBand);
            w *= 1.0 + 1e-12;
        }
        q.w = up32(w);
    } else {
        // Padding: far away and massless, never screened.
        q.ox = q.oy = q.oz = 1.0e18f;
        q.ovx = q.ovy = q.ovz = 0.0f;
        q.m = 0.0f;
        q.w = 1.0f;
    }
    j32[p] = q;
    e = smax(e);
    ev = smax(ev);
    w = smax(real ? (double)q.w : 1.0);
    if ((lane & (kVTile - 1)) == 0) {
        JTile T;
        T.cx = cen[0];
        T.cy = cen[1];
        T.cz = cen[2];
        T.cu = cen[3];
        T.cv = cen[4];
        T.cw = cen[5];
        T.e = up32(e);
        T.ev = up32(ev);
        T.wmax = (float)w;
        T.eb = up32((double)kTileBand * (double)kTileBand * (double)T.e *
                    (double)T.e * (1.0 + 1e-12));
        tiles[t * (kTile / kVTile) + lane / kVTile] = T;
    }
}

__global__ void k_set_prepclus(PrepClus *pc, PrepClus v) { *pc = v; }

// Padding FP32 records and empty tiles over a whole allocation.
__global__ __launch_bounds__(256) void k_park_view(JRec32 *__restrict__ j32,
                                                   int n, JTile *__restrict__ tiles,
                                                   int ntile) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        JRec32 q;
        q.ox = q.oy = q.oz = 1.0e18f;
        q.ovx = q.ovy = q.ovz = 0.0f;
        q.m = 0.0f;
        q.w = 1.0f;
        j32[k] = q;
    }
    if (k < ntile) {
        JTile T;
        T.cx = T.cy = T.cz = T.cu = T.cv = T.cw = 0.0;
        T.e = T.ev = 0.0f;
        T.wmax = 1.0f;
        T.eb = 0.0f;
        tiles[k] = T;
    }
}

static unsigned key_bits(int ncl) {
    unsigned cbits = 0;
    while ((1 << cbits) < ncl) ++cbits;
    return 32u + cbits;
}

// Hilbert order of the i of a call in the j set's box (cluster 0), so that
// the 64 i of a pass-1 wavefront are spatially close: far tiles then skip
// the screen for the whole wave, and the near band's FP64 pairs gather in
// fewer tiles.
__global__ __launch_bounds__(256) void k_ikey(const double *__restrict__ xi,
                                              int ni, const JClus *__restrict__ jcl,
                                              unsigned *__restrict__ key,
                                              int *__restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    key[i] = cell_key(xi[3 * (size_t)i + 0], xi[3 * (size_t)i + 1],
                      xi[3 * (size_t)i + 2], jcl[0]);
    val[i] = i;
}

// The i data in sorted order: xs/vs [ni][3], hs/ds [ni].
__global__ __launch_bounds__(256) void k_igather(
    const int *__restrict__ perm, int ni, const double *__restrict__ xi,
    const double *__restrict__ vi, const double *__restrict__ h2,
    const double *__restrict__ dtr, double *__restrict__ xs,
    double *__restrict__ vs, double *__restrict__ hs, double *__restrict__ ds) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ni) return;
    const int i = perm[p];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        xs[3 * (size_t)p + c] = xi[3 * (size_t)i + c];
        vs[3 * (size_t)p + c] = vi[3 * (size_t)i + c];
    }
    hs[p] = h2[i];
    ds[p] = dtr[i];
}

// Ensemble i keys: (cluster << 32) | Hilbert index of x_i in the box of the
// cluster owning global i (the last c with cl_i0[c] <= i; clusters hold
// consecutive i ranges in cluster order, empty ones repeat the next start).
__global__ __launch_bounds__(256) void k_ikey_ens(
    const double *__restrict__ xi, int ni, const int *__restrict__ cl_i0,
    int ncl, const JClus *__restrict__ jcl, unsigned long long *__restrict__ key,
    int *__restrict__ val) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ni) return;
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cl_i0[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const unsigned code = cell_key(xi[3 * (size_t)g + 0], xi[3 * (size_t)g + 1],
                                   xi[3 * (size_t)g + 2], jcl[lo]);
    key[g] = ((unsigned long long)lo << 32) | code;
    val[g] = g;
}

size_t isort_ens_tmp_bytes(int n, int ncl) {
    size_t bytes = 0;
    const hipError_t rc = rocprim::radix_sort_pairs(
        nullptr, bytes, (const unsigned long long *)nullptr,
        (unsigned long long *)nullptr, (const int *)nullptr, (int *)nullptr,
        (size_t)std::max(n, 1), 0u, key_bits(ncl), (hipStream_t)0);
    return rc == hipSuccess ? bytes : 0;
}

hipError_t launch_isort_ens(hipStream_t st, const JClus *jcl, const int *cl_i0,
                            int ncl, const double *xi, const double *vi,
                            const double *h2, const double *dtr, int ni,
                            const ISortBufs &B) {
    if (ni <= 0) return hipSuccess;
    const dim3 g((ni + 255) / 256);
    hipLaunchKernelGGL(k_ikey_ens, g, dim3(256), 0, st, xi, ni, cl_i0, ncl, jcl,
                       B.key64_in, B.val_in);
    size_t bytes = B.tmp_bytes;
    const hipError_t rc = rocprim::radix_sort_pairs(
        B.tmp, bytes, B.key64_in, B.key64_out, B.val_in, B.perm, (size_t)ni, 0u,
        key_bits(ncl), st);
    if (rc != hipSuccess) return rc;
    hipLaunchKernelGGL(k_igather, g, dim3(256), 0, st, B.perm, ni, xi, vi, h2,
                       dtr, B.xs, B.vs, B.hs, B.ds);
    return hipGetLastError();
}

size_t isort_tmp_bytes(int n) {
    size_t bytes = 0;
    const hipError_t rc = rocprim::radix_sort_pairs(
        nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
        (const int *)nullptr, (int *)nullptr, (size_t)std::max(n, 1), 0u, 30u,
        (hipStream_t)0);
    return rc == hipSuccess ? bytes : 0;
}

hipError_t launch_isort(hipStream_t st, const JClus *jcl, const double *xi,
                        const double *vi, const double *h2, const double *dtr,
                        int ni, const ISortBufs &B) {
    if (ni <= 0) return hipSuccess;
    const dim3 g((ni + 255) / 256);
    hipLaunchKernelGGL(k_ikey, g, dim3(256), 0, st, xi, ni, jcl, B.key_in,
                       B.val_in);
    size_t bytes = B.tmp_bytes;
    const hipError_t rc = rocprim::radix_sort_pairs(
        B.tmp, bytes, B.key_in, B.key_out, B.val_in, B.perm, (size_t)ni, 0u, 30u,
        st);
    if (rc != hipSuccess) return rc;
    hipLaunchKernelGGL(k_igather, g, dim3(256), 0, st, B.perm, ni, xi, vi, h2,
                       dtr, B.xs, B.vs, B.hs, B.ds);
    return hipGetLastError();
}

size_t jprep_tmp_bytes(int n, int ncl) {
    size_t bytes = 0;
    const hipError_t rc = rocprim::radix_sort_pairs(
        nullptr, bytes, (const unsigned long long *)nullptr,
        (unsigned long long *)nullptr, (const int *)nullptr, (int *)nullptr,
        (size_t)std::max(n, 1), 0u, key_bits(ncl), (hipStream_t)0);
    return rc == hipSuccess ? bytes : 0;
}

void launch_set_prepclus(hipStream_t st, PrepClus *pc, const PrepClus &v) {
    hipLaunchKernelGGL(k_set_prepclus, dim3(1), dim3(1), 0, st, pc, v);
}

void launch_park_view(hipStream_t st, JRec32 *j32, int n, JTile *tiles,
                      int ntile) {
    const int m = std::max(n, ntile);
    if (m <= 0) return;
    hipLaunchKernelGGL(k_park_view, dim3((m + 255) / 256), dim3(256), 0, st,
                       j32, n, tiles, ntile);
}

hipError_t launch_jprep(hipStream_t st, const JView &V, const PrepClus *pc,
                        int ncl, int nblk, int ntile_all, const PrepBufs &B) {
    if (ntile_all <= 0) return hipSuccess;
    const int n_all = ntile_all * kTile;
    hipLaunchKernelGGL(k_jstat1, dim3(nblk), dim3(256), 0, st, V.jp, pc, ncl,
                       B.part);
    hipLaunchKernelGGL(k_jstat2, dim3(ncl), dim3(256), 0, st, pc, ncl, nblk,
                       B.part, V.jcl);
    hipLaunchKernelGGL(k_jkey, dim3((n_all + 255) / 256), dim3(256), 0, st, V.jp,
                       pc, ncl, V.jcl, n_all, B.key_in, B.val_in);
    size_t bytes = B.tmp_bytes;
    const hipError_t rc = rocprim::radix_sort_pairs(
        B.tmp, bytes, B.key_in, B.key_out, B.val_in, B.val_out, (size_t)n_all,
        0u, key_bits(ncl), st);
    if (rc != hipSuccess) return rc;
    hipLaunchKernelGGL(k_jtile, dim3(ntile_all), dim3(64), 0, st, V.jp, pc, ncl,
                       V.jcl, B.val_out, V.jps, V.j32, V.tiles, V.jidx);
    return hipGetLastError();
}

}  // namespace gpunb
