// irr.hip -- irregular (neighbour) force on the device: a drop-in for the
// reference's irr_simd library (irr.avx.cpp:565-603, the Fortran interface
// irr_simd_open_/close_/profile_/set_jp_/set_list_/firr_vec_), plus batched
// entry points for callers that hold many particles at once.
//
// State in HBM: one 128-byte predictor record per particle (x0, v0, f = a/2,
// fd = jerk/6, m, t0 in FP64) and a neighbour table of nmax rows of lmax ints
// (count, then 0-based indices).  set_jp / set_list stage their data on the
// host; the next firr_vec uploads the staged records and rows in one copy each
// and scatters them with a kernel, so per-particle Fortran calls cost no
// device round trip.
//
// k_irr_force: one wavefront per active particle.  Every lane predicts its
// list members with the FP64 polynomial of xbpredall.f:18-26 (the same
// association as the oracle, no contraction), evaluates the pair in FP64 and
// accumulates in FP64; the wave then reduces in a fixed order.  When a call
// has more than three pairs per slot, k_irr_predict first predicts every
// slot once with the same function, and the members are read predicted.  The nearest
// member is the smallest r^2, ties to the smaller index (the reference's
// packed (r^2, index) minimum, irr.avx.cpp:317-320), returned 1-based, 0 for
// an empty list.  The reference evaluates the pairs in FP32 on double-single
// positions; FP64 costs little here (a few tens of pairs per particle).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpunb.h"
#include "gpunb_internal.h"

namespace gpunb_irr {

struct alignas(128) IRec {
    double x[3], v[3], f[3], fd[3];
    double m, t0;
    double pad[2];
};
static_assert(sizeof(IRec) == 128, "IRec must be 128 bytes");

#define ICHK(e)                                                               \
    do {                                                                      \
        hipError_t r_ = (e);                                                  \
        if (r_ != hipSuccess)                                                 \
            throw std::runtime_error(std::string(#e) + ": " +                 \
                                     hipGetErrorString(r_));                  \
    } while (0)

#define IREQ(c, msg)                                                          \
    do {                                                                      \
        if (!(c)) throw std::runtime_error(msg);                              \
    } while (0)

// Predicted position and velocity at t (xbpredall.f:18-26 association).
__device__ inline void predict(const IRec &r, double t, double *x, double *v) {
    const double s = t - r.t0;
    const double s1 = 1.5 * s;
    const double s2 = 2.0 * s;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = ((r.fd[k] * s + r.f[k]) * s + r.v[k]) * s + r.x[k];
        v[k] = (r.fd[k] * s1 + r.f[k]) * s2 + r.v[k];
    }
}

constexpr int kWavesPerBlock = 4;
// Results of calls of at least kChunkMin bodies go to the host in kOutChunks
// parts, so the host moves one part into the caller's arrays while the next
// one crosses the host link.
constexpr int kOutChunks = 4;
constexpr int kChunkMin = 4 * 8192;

// A particle predicted to the call's time: position, velocity, mass.
struct alignas(64) PRec {
    double x[3], v[3], m, pad;
};
static_assert(sizeof(PRec) == 64, "PRec must be 64 bytes");

// Every slot predicted once (the same function and operation order as the
// per-member prediction of k_irr_force<false>, so the values are the same).
__global__ __launch_bounds__(256) void k_irr_predict(int n, double t,
                                                     const IRec *__restrict__ jp,
                                                     PRec *__restrict__ pj) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const IRec &r = jp[k];
    PRec o;
    predict(r, t, o.x, o.v);
    o.m = r.m;
    o.pad = 0.0;
    pj[k] = o;
}

// PRE: members read from the predicted records pj (64 B each) instead of
// being predicted from jp (128 B each) member by member.
template <bool PRE>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_irr_force(
    int ni, const int *__restrict__ idx, double t, const IRec *__restrict__ jp,
    const PRec *__restrict__ pj, const int *__restrict__ list, int lstride,
    double *__restrict__ acc, double *__restrict__ jrk,
    int *__restrict__ nnbid) {
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (q >= ni) return;                  // uniform per wavefront
    const int i = idx[q];
    double xi[3], vi[3];
    if (PRE) {
        const PRec &p = pj[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xi[c] = p.x[c];
            vi[c] = p.v[c];
        }
    } else {
        predict(jp[i], t, xi, vi);
    }
    const int *row = list + (size_t)i * lstride;
    const int nnb = row[0];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
    double best = INFINITY;
    int bestj = INT_MAX;
    for (int k = lane; k < nnb; k += 64) {
        const int j = row[1 + k];
        double xj[3], vj[3], mj;
        if (PRE) {
            const PRec &p = pj[j];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                xj[c] = p.x[c];
                vj[c] = p.v[c];
            }
            mj = p.m;
        } else {
            predict(jp[j], t, xj, vj);
            mj = jp[j].m;
        }
        const double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
        const double dvx = vj[0] - vi[0], dvy = vj[1] - vi[1], dvz = vj[2] - vi[2];
        double r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
        if (r2 < best || (r2 == best && j < bestj)) {
            best = r2;
            bestj = j;
        }
        if (r2 == 0.0) continue;
        const double rinv = 1.0 / sqrt(r2);
        const double rinv2 = rinv * rinv;
        const double mr3 = mj * rinv2 * rinv;
        double rv = dx * dvx;
        rv = rv + dy * dvy;
        rv = rv + dz * dvz;
        const double alpha = -3.0 * rv * rinv2;
        a0 += mr3 * dx;
        a1 += mr3 * dy;
        a2 += mr3 * dz;
        j0 += mr3 * (dvx + alpha * dx);
        j1 += mr3 * (dvy + alpha * dy);
        j2 += mr3 * (dvz + alpha * dz);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
        j0 += __shfl_xor(j0, o);
        j1 += __shfl_xor(j1, o);
        j2 += __shfl_xor(j2, o);
        const double ob = __shfl_xor(best, o);
        const int oj = __shfl_xor(bestj, o);
        if (ob < best || (ob == best && oj < bestj)) {
            best = ob;
            bestj = oj;
        }
    }
    if (lane == 0) {
        acc[3 * (size_t)q + 0] = a0;
        acc[3 * (size_t)q + 1] = a1;
        acc[3 * (size_t)q + 2] = a2;
        jrk[3 * (size_t)q + 0] = j0;
        jrk[3 * (size_t)q + 1] = j1;
        jrk[3 * (size_t)q + 2] = j2;
        nnbid[q] = bestj == INT_MAX ? 0 : bestj + 1;
    }
}

// Staged predictor records -> their slots.
__global__ void k_irr_scatter_jp(int n, const int *__restrict__ addr,
                                 const IRec *__restrict__ src,
                                 IRec *__restrict__ jp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) jp[addr[k]] = src[k];
}

// Staged list rows: hdr[3r] = slot, hdr[3r+1] = offset of the row in flat,
// hdr[3r+2] = count; one wavefront per row (entries already 0-based).
__global__ void k_irr_scatter_lists(int nrows, const int *__restrict__ hdr,
                                    const int *__restrict__ flat,
                                    int *__restrict__ list, int lstride) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= nrows) return;
    const int slot = hdr[3 * r], off = hdr[3 * r + 1], cnt = hdr[3 * r + 2];
    int *row = list + (size_t)slot * lstride;
    if (lane == 0) row[0] = cnt;
    for (int k = lane; k < cnt; k += 64) row[1 + k] = flat[off + k];
}

// Device -> mapped pinned host copy of three ranges of n16[b] 16-byte units,
// blockIdx.y selecting the range: one part of a call's acceleration, jerk and
// nearest-neighbour results.
struct Ranges {
    const uint4 *src[3];
    uint4 *dst[3];
    size_t n16[3];
};

__global__ __launch_bounds__(256) void k_irr_to_host(Ranges r) {
    const int b = blockIdx.y;
    const uint4 *__restrict__ src = r.src[b];
    uint4 *__restrict__ dst = r.dst[b];
    const size_t n16 = r.n16[b];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16;
         k += stride)
        dst[k] = src[k];
}

// Grow-only device / pinned host buffers.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    void ensure(size_t n) {
        if (n <= cap) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        ICHK(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
        cap = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

template <class T>
struct HBuf {
    T *p = nullptr;
    size_t cap = 0;
    void ensure(size_t n) {
        if (n <= cap) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        ICHK(hipHostMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
        cap = n;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct State {
    bool open = false;
    int nmax = 0, lstride = 0;
    hipStream_t st = nullptr;
    DBuf<IRec> jp;
    DBuf<PRec> pj;                            // slots predicted to a call's t
    DBuf<int> list;
    // staging (host), latest data per slot wins
    std::vector<IRec> jp_rec;
    std::vector<int> jp_addr, jp_slot;        // jp_slot[a] = staged index or -1
    std::vector<int> ls_hdr, ls_flat, ls_slot;
    std::vector<int> nnb;                     // host mirror of the counts
    HBuf<IRec> h_rec;
    HBuf<int> h_int;
    DBuf<IRec> d_rec;
    DBuf<int> d_int;
    // results of a call: acc[ni_pad][3], jerk[ni_pad][3] (d_out, h_out) and
    // the nearest neighbours[ni_pad] (d_nn, h_nn), ni_pad = ni rounded up to
    // 64, so every part of them starts on a 16-byte boundary
    DBuf<double> d_out;
    DBuf<int> d_nn;
    HBuf<double> h_out;
    HBuf<int> h_nn;
    // profile
    long long ncall = 0, nsteps = 0, ninter = 0;
    double t_force = 0.0;
    // k_irr_force of the last call, timed on the stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms = 0.0;
    hipEvent_t part[kOutChunks] = {};        // after the copy of part c
};
State I;

// Staging.  The reference's host calls set_jp / set_list for every particle
// inside an OpenMP parallel loop (intgrt.F:200-205), so they must be safe to
// call concurrently: each thread appends to its own buffer, and the next
// firr_vec (called from one thread, with no staging running) merges all
// buffers, keeping the last value staged for each particle.
struct Stage {
    std::vector<IRec> rec;
    std::vector<int> rec_addr;   // 0-based slot per record
    std::vector<int> hdr;        // (slot, offset in flat, count) per row
    std::vector<int> flat;       // members, 0-based
    void clear() {
        rec.clear();
        rec_addr.clear();
        hdr.clear();
        flat.clear();
    }
};
std::mutex g_stage_mu;
std::vector<Stage *> g_stages;   // kept for the life of the process
thread_local Stage *t_stage = nullptr;

Stage &my_stage() {
    if (!t_stage) {
        Stage *st = new Stage;
        std::lock_guard<std::mutex> g(g_stage_mu);
        g_stages.push_back(st);
        t_stage = st;
    }
    return *t_stage;
}

void stage_jp(int a, const double *x, const double *v, const double *f,
              const double *fd, double m, double t0) {
    IREQ(a >= 1 && a <= I.nmax, "address out of range [1, nmax]");
    IRec r{};
    for (int k = 0; k < 3; ++k) {
        r.x[k] = x[k];
        r.v[k] = v[k];
        r.f[k] = f[k];
        r.fd[k] = fd[k];
    }
    r.m = m;
    r.t0 = t0;
    Stage &st = my_stage();
    st.rec.push_back(r);
    st.rec_addr.push_back(a - 1);
}

// nb: 1-based member indices.
void stage_list(int a, int cnt, const int *nb) {
    IREQ(a >= 1 && a <= I.nmax, "address out of range [1, nmax]");
    IREQ(cnt >= 0 && cnt <= I.lstride - 1, "list longer than lmax - 1");
    for (int k = 0; k < cnt; ++k)
        IREQ(nb[k] >= 1 && nb[k] <= I.nmax, "list member out of range [1, nmax]");
    Stage &st = my_stage();
    st.hdr.push_back(a - 1);
    st.hdr.push_back((int)st.flat.size());
    st.hdr.push_back(cnt);
    for (int k = 0; k < cnt; ++k) st.flat.push_back(nb[k] - 1);
}

// All threads' staged data into I.jp_rec / jp_addr / ls_hdr / ls_flat, the
// last value per particle winning (a superseded row keeps slot -1).
void merge_stages() {
    std::lock_guard<std::mutex> g(g_stage_mu);
    for (Stage *st : g_stages) {
        for (size_t k = 0; k < st->rec.size(); ++k) {
            const int s = st->rec_addr[k];
            if (I.jp_slot[s] >= 0) {
                I.jp_rec[I.jp_slot[s]] = st->rec[k];
            } else {
                I.jp_slot[s] = (int)I.jp_rec.size();
                I.jp_rec.push_back(st->rec[k]);
                I.jp_addr.push_back(s);
            }
        }
        for (size_t r = 0; r < st->hdr.size(); r += 3) {
            const int s = st->hdr[r], off = st->hdr[r + 1], cnt = st->hdr[r + 2];
            if (I.ls_slot[s] >= 0) I.ls_hdr[3 * I.ls_slot[s]] = -1;
            I.ls_slot[s] = (int)(I.ls_hdr.size() / 3);
            I.ls_hdr.push_back(s);
            I.ls_hdr.push_back((int)I.ls_flat.size());
            I.ls_hdr.push_back(cnt);
            I.ls_flat.insert(I.ls_flat.end(), st->flat.begin() + off,
                             st->flat.begin() + off + cnt);
            I.nnb[s] = cnt;
        }
        st->clear();
    }
}

void clear_stages() {
    std::lock_guard<std::mutex> g(g_stage_mu);
    for (Stage *st : g_stages) st->clear();
}


double wtime() {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

bool quiet() {
    const char *q = getenv("GPUNB_QUIET");
    return q && *q && *q != '0';
}

void do_open(int nmax, int lmax, int rank) {
    if (I.open) {
        fprintf(stderr, "irr_simd: it is already open\n");
        return;
    }
    IREQ(nmax > 0, "irr_simd_open: nmax must be > 0");
    IREQ(lmax >= 1, "irr_simd_open: lmax must be >= 1");
    ICHK(hipStreamCreateWithFlags(&I.st, hipStreamNonBlocking));
    ICHK(hipEventCreate(&I.ev0));
    ICHK(hipEventCreate(&I.ev1));
    for (int c = 0; c < kOutChunks; ++c)
        ICHK(hipEventCreateWithFlags(&I.part[c], hipEventDisableTiming));
    I.last_kernel_ms = 0.0;
    I.nmax = nmax;
    I.lstride = lmax;
    I.jp.ensure((size_t)nmax);
    I.list.ensure((size_t)nmax * lmax);
    // Empty lists and a parked predictor state until the caller sets them.
    ICHK(hipMemsetAsync(I.list.p, 0, (size_t)nmax * lmax * sizeof(int), I.st));
    ICHK(hipMemsetAsync(I.jp.p, 0, (size_t)nmax * sizeof(IRec), I.st));
    ICHK(hipStreamSynchronize(I.st));
    I.jp_slot.assign(nmax, -1);
    I.ls_slot.assign(nmax, -1);
    I.nnb.assign(nmax, 0);
    I.jp_rec.clear();
    I.jp_addr.clear();
    I.ls_hdr.clear();
    I.ls_flat.clear();
    clear_stages();
    I.ncall = I.nsteps = I.ninter = 0;
    I.t_force = 0.0;
    I.open = true;
    if (!quiet())
        fprintf(stderr, "# Opening irr_simd lib. MI355X ver. - rank: %d; nmax: %d, lmax: %d\n",
                rank, nmax, lmax);
}

void do_close(int rank) {
    if (!I.open) {
        fprintf(stderr, "irr_simd: it is already close\n");
        return;
    }
    ICHK(hipStreamSynchronize(I.st));
    I.jp.release();
    I.pj.release();
    I.list.release();
    I.d_rec.release();
    I.d_int.release();
    I.d_out.release();
    I.d_nn.release();
    I.h_rec.release();
    I.h_int.release();
    I.h_out.release();
    I.h_nn.release();
    (void)hipEventDestroy(I.ev0);
    (void)hipEventDestroy(I.ev1);
    I.ev0 = I.ev1 = nullptr;
    for (int c = 0; c < kOutChunks; ++c) {
        (void)hipEventDestroy(I.part[c]);
        I.part[c] = nullptr;
    }
    (void)hipStreamDestroy(I.st);
    I.st = nullptr;
    I.open = false;
    if (!quiet()) fprintf(stderr, "Closing irr_simd lib. MI355X ver. - rank: %d\n", rank);
}

// Upload the staged records and list rows, scatter them, then (ni > 0) the
// force of addr[] at t; the host waits only for each part of the results
// (one part below kChunkMin bodies).  Layout of the pinned int buffer:
// [jp slots (nj)] [row headers (3 nr)] [members] [i (ni)].
void flush_and_force(double t, int ni, const int *addr, double *acc,
                     double *jrk, int *nnbid) {
    merge_stages();
    const size_t nj = I.jp_rec.size();
    std::vector<int> hdr;
    hdr.reserve(I.ls_hdr.size());
    for (size_t r = 0; r < I.ls_hdr.size(); r += 3)
        if (I.ls_hdr[r] >= 0)
            hdr.insert(hdr.end(), I.ls_hdr.begin() + r, I.ls_hdr.begin() + r + 3);
    const size_t nr = hdr.size() / 3, nf = I.ls_flat.size();
    const size_t o_hdr = nj, o_flat = o_hdr + 3 * nr, o_idx = o_flat + nf;
    const size_t nint = o_idx + (size_t)ni;
    long long pairs = 0;
    if (nint) {
        I.h_int.ensure(nint);
        I.d_int.ensure(nint);
        int *h = I.h_int.p;
        memcpy(h, I.jp_addr.data(), nj * sizeof(int));
        memcpy(h + o_hdr, hdr.data(), 3 * nr * sizeof(int));
        memcpy(h + o_flat, I.ls_flat.data(), nf * sizeof(int));
        for (int q = 0; q < ni; ++q) {
            const int a = addr[q];
            IREQ(a >= 1 && a <= I.nmax, "address out of range [1, nmax]");
            h[o_idx + q] = a - 1;
        }
        ICHK(hipMemcpyAsync(I.d_int.p, h, nint * sizeof(int),
                            hipMemcpyHostToDevice, I.st));
    }
    if (nj) {
        I.h_rec.ensure(nj);
        I.d_rec.ensure(nj);
        memcpy(I.h_rec.p, I.jp_rec.data(), nj * sizeof(IRec));
        ICHK(hipMemcpyAsync(I.d_rec.p, I.h_rec.p, nj * sizeof(IRec),
                            hipMemcpyHostToDevice, I.st));
        hipLaunchKernelGGL(k_irr_scatter_jp, dim3((unsigned)((nj + 255) / 256)),
                           dim3(256), 0, I.st, (int)nj, I.d_int.p, I.d_rec.p,
                           I.jp.p);
    }
    if (nr)
        hipLaunchKernelGGL(k_irr_scatter_lists, dim3((unsigned)((nr + 3) / 4)),
                           dim3(256), 0, I.st, (int)nr, I.d_int.p + o_hdr,
                           I.d_int.p + o_flat, I.list.p, I.lstride);
    for (int s : I.jp_addr) I.jp_slot[s] = -1;
    for (size_t r = 0; r < I.ls_hdr.size(); r += 3)
        if (I.ls_hdr[r] >= 0) I.ls_slot[I.ls_hdr[r]] = -1;
    // Counts of the members sent now are already in the host mirror.
    for (int q = 0; q < ni; ++q) pairs += I.nnb[addr[q] - 1];
    I.jp_rec.clear();
    I.jp_addr.clear();
    I.ls_hdr.clear();
    I.ls_flat.clear();
    if (ni > 0) {
        const size_t np = ((size_t)ni + 63) & ~(size_t)63;
        I.d_out.ensure(np * 6);
        I.d_nn.ensure(np);
        I.h_out.ensure(np * 6);
        I.h_nn.ensure(np);
        ICHK(hipEventRecord(I.ev0, I.st));
        // Predict every slot once when the call's pairs outnumber the slots
        // threefold: each member then costs a 64-byte read instead of a
        // 128-byte read and its prediction.
        const bool pre = pairs > 3LL * I.nmax;
        if (pre) {
            I.pj.ensure((size_t)I.nmax);
            hipLaunchKernelGGL(k_irr_predict, dim3((unsigned)((I.nmax + 255) / 256)),
                               dim3(256), 0, I.st, I.nmax, t, I.jp.p, I.pj.p);
        }
        const dim3 grid((unsigned)((ni + kWavesPerBlock - 1) / kWavesPerBlock));
        if (pre)
            hipLaunchKernelGGL(k_irr_force<true>, grid, dim3(64 * kWavesPerBlock),
                               0, I.st, ni, I.d_int.p + o_idx, t, I.jp.p,
                               I.pj.p, I.list.p, I.lstride, I.d_out.p,
                               I.d_out.p + 3 * np, I.d_nn.p);
        else
            hipLaunchKernelGGL(k_irr_force<false>, grid, dim3(64 * kWavesPerBlock),
                               0, I.st, ni, I.d_int.p + o_idx, t, I.jp.p,
                               (const PRec *)nullptr, I.list.p, I.lstride,
                               I.d_out.p, I.d_out.p + 3 * np, I.d_nn.p);
        ICHK(hipEventRecord(I.ev1, I.st));
        // Part c: bodies q0 .. q0 + nq (q0 a multiple of 64), copied by one
        // launch to the mapped h_out / h_nn; a byte count is rounded up to
        // 16, which stays inside the padded arrays.
        const int nchunk = ni >= kChunkMin ? kOutChunks : 1;
        const int per = (int)((((size_t)ni + nchunk - 1) / nchunk + 63) & ~(size_t)63);
        void *ho = nullptr, *hn = nullptr;
        ICHK(hipHostGetDevicePointer(&ho, I.h_out.p, 0));
        ICHK(hipHostGetDevicePointer(&hn, I.h_nn.p, 0));
        for (int c = 0; c < nchunk; ++c) {
            const size_t q0 = (size_t)c * per;
            if (q0 >= (size_t)ni) break;
            const size_t nq = std::min<size_t>(ni - q0, per);
            Ranges r;
            r.src[0] = (const uint4 *)(I.d_out.p + 3 * q0);
            r.dst[0] = (uint4 *)((double *)ho + 3 * q0);
            r.src[1] = (const uint4 *)(I.d_out.p + 3 * (np + q0));
            r.dst[1] = (uint4 *)((double *)ho + 3 * (np + q0));
            r.src[2] = (const uint4 *)(I.d_nn.p + q0);
            r.dst[2] = (uint4 *)((int *)hn + q0);
            r.n16[0] = r.n16[1] = (24 * nq + 15) / 16;
            r.n16[2] = (4 * nq + 15) / 16;
            const unsigned gx = (unsigned)std::min<size_t>(512, (r.n16[0] + 255) / 256);
            hipLaunchKernelGGL(k_irr_to_host, dim3(gx, 3), dim3(256), 0, I.st, r);
            ICHK(hipEventRecord(I.part[c], I.st));
        }
        ICHK(hipGetLastError());
        // Each part into the caller's arrays once its copy is done (the
        // uploads, the force and every earlier copy were before it on I.st).
        for (int c = 0; c < nchunk; ++c) {
            const size_t q0 = (size_t)c * per;
            if (q0 >= (size_t)ni) break;
            const size_t nq = std::min<size_t>(ni - q0, per);
            ICHK(hipEventSynchronize(I.part[c]));
            memcpy(acc + 3 * q0, I.h_out.p + 3 * q0, 3 * nq * sizeof(double));
            memcpy(jrk + 3 * q0, I.h_out.p + 3 * (np + q0), 3 * nq * sizeof(double));
            if (nnbid) memcpy(nnbid + q0, I.h_nn.p + q0, nq * sizeof(int));
        }
        float ms = 0.0f;
        ICHK(hipEventElapsedTime(&ms, I.ev0, I.ev1));
        I.last_kernel_ms = ms;
        I.ncall++;
        I.nsteps += ni;
        I.ninter += pairs;
    } else {
        ICHK(hipGetLastError());
        // The pinned staging buffers are reused by the next call.
        if (nint || nj) ICHK(hipStreamSynchronize(I.st));
    }
}

void firr(double t, int ni, const int *addr, double *acc, double *jrk,
          int *nnbid) {
    IREQ(I.open, "irr_simd is not open");
    IREQ(ni >= 0, "ni must be >= 0");
    const double w0 = wtime();
    flush_and_force(t, ni, addr, acc, jrk, nnbid);
    I.t_force += wtime() - w0;
}

void profile(int rank) {
    if (!I.open || quiet()) return;
    const double gf = I.t_force > 0 ? 60.0 * (double)I.ninter * 1e-9 / I.t_force : 0.0;
    fprintf(stderr,
            "[R.%d MI355X Irr.F ] Ncall: %lld <NI>: %lld <NB>: %f grav: %f s, %f Gflops, %f usec\n",
            rank, I.ncall, I.ncall ? I.nsteps / I.ncall : 0,
            I.nsteps ? (double)I.ninter / I.nsteps : 0.0, I.t_force, gf,
            I.ncall ? 1e6 * I.t_force / I.ncall : 0.0);
    I.ncall = I.nsteps = I.ninter = 0;
    I.t_force = 0.0;
}

}  // namespace gpunb_irr

using namespace gpunb_irr;

#define IRR_EXPORT __attribute__((visibility("default")))

extern "C" {

IRR_EXPORT void irr_simd_open_(int *nmax, int *lmax, int *rank) {
    gpunb::run_guarded("irr_simd_open_", [&] { do_open(*nmax, *lmax, *rank); });
}

IRR_EXPORT void irr_simd_close_(int *rank) {
    gpunb::run_guarded("irr_simd_close_", [&] { do_close(*rank); });
}

IRR_EXPORT void irr_simd_profile_(int *rank) {
    gpunb::run_guarded("irr_simd_profile_", [&] { profile(*rank); });
}

IRR_EXPORT void irr_simd_set_jp_(int *addr, double pos[3], double vel[3],
                                 double acc2[3], double jrk6[3], double *mass,
                                 double *time) {
    gpunb::run_guarded("irr_simd_set_jp_", [&] {
        IREQ(I.open, "irr_simd is not open");
        stage_jp(*addr, pos, vel, acc2, jrk6, *mass, *time);
    });
}

IRR_EXPORT void irr_simd_set_list_(int *addr, int *nblist) {
    gpunb::run_guarded("irr_simd_set_list_", [&] {
        IREQ(I.open, "irr_simd is not open");
        stage_list(*addr, nblist[0], nblist + 1);
    });
}

IRR_EXPORT void irr_simd_firr_vec_(double *ti, int *ni, int addr[],
                                   double acc[][3], double jrk[][3],
                                   int nnbid[]) {
    gpunb::run_guarded("irr_simd_firr_vec_", [&] {
        firr(*ti, *ni, addr, &acc[0][0], &jrk[0][0], nnbid);
    });
}

IRR_EXPORT int gpunb_irr_set_jp(int n, const int *addr, const double *x,
                                const double *v, const double *f,
                                const double *fd, const double *m,
                                const double *t0) {
    return gpunb::run_guarded("gpunb_irr_set_jp", [&] {
        IREQ(I.open, "irr_simd is not open");
        for (int q = 0; q < n; ++q)
            stage_jp(addr[q], x + 3 * (size_t)q, v + 3 * (size_t)q,
                     f + 3 * (size_t)q, fd + 3 * (size_t)q, m[q], t0[q]);
    });
}

IRR_EXPORT double gpunb_irr_kernel_ms(void) { return I.last_kernel_ms; }

IRR_EXPORT int gpunb_irr_set_lists(int n, const int *addr, const int *nnb,
                                   const int *members) {
    return gpunb::run_guarded("gpunb_irr_set_lists", [&] {
        IREQ(I.open, "irr_simd is not open");
        size_t off = 0;
        for (int q = 0; q < n; ++q) {
            stage_list(addr[q], nnb[q], members + off);
            off += (size_t)nnb[q];
        }
    });
}

}  // extern "C"
