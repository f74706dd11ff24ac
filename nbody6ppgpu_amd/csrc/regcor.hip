// regcor.hip -- the regular corrector of a block of bodies on the device:
// the work NBODY6++GPU's host does per body after gpunb_regf
// (src/Main/regcor_gpu.F:1-805, called from util_gpu.F:152-180), for an
// isolated cluster (no cloud or external force, no CHECKL / FPCORR):
//   * the Ahmad-Cohen neighbour-radius update (regcor_gpu.F:146-253),
//   * the old/new list comparison with the small-step retention
//     (:266-370) and the irregular-force change of the lost and gained
//     neighbours (:400-438),
//   * the Hermite corrector of the regular force and the new D0..D3R
//     (:440-551), the list copy (:605-613) and the new regular / irregular
//     step (:615-805).
//
// One thread per body: the work per body is a few hundred FP64 operations
// with data-dependent branches and list walks, far too little for a
// wavefront.  Neighbour positions, velocities and masses come from the
// library's predicted j records (the X, XDOT, BODY the host would read).
//
// The arithmetic is the reference's, operation for operation and in its
// evaluation order (the library is built with -ffp-contract=off), so the
// results match the CPU restatement bit for bit where the math library's
// sqrt / pow agree; REAL*4 literals and FLOAT() of the Fortran are float
// values here as there.  STEP(J) of other bodies is read as it was before
// the block: the new STEP / TIMENW are written by a second launch.

#include "gpunb_internal.h"

namespace gpunb {

namespace {

__device__ inline double sqd(double a) { return a * a; }

constexpr double kOne3 = 1.0 / 3.0;    // ONE3 (zero.f:113)
constexpr double kOne12 = 1.0 / 12.0;  // ONE12 (zero.f:116)
constexpr int kMaxHalve = 2200;        // > 2098 halvings: DBL_MAX -> 0
constexpr int kChunk = 8;              // list entries loaded per chunk
constexpr int kWin = 4;                // look-ahead of the list merge in NLIST
constexpr int kNWin = 16;              // LIST entries per round of the merge
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt not waited on): every vector
// memory operation of the wavefront complete.
constexpr int kVmcnt0 = 0x0F70;
// LDS rows of one workgroup (one wavefront, 64 bodies, an NLIST row each):
// 20 KB, so that the eight workgroups a CU holds at N = 131072 stay
// resident together.
constexpr int kStage = 5120;
constexpr int kStageU = 16;            // rows whose loads are in flight together
constexpr int kPairs = 4;              // list-change members loaded together

// Exclusive prefix sum of v over the 64 lanes; *total gets the full sum.
__device__ inline int wave_excl_scan(int v, int lane, int *total) {
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

__device__ inline int *lane_ptr(int *p, int r) {
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)a, r);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(a >> 32), r);
    return (int *)(((unsigned long long)hi << 32) | lo);
}

// Row copies between HBM and LDS for the 64 lanes of a wavefront: lane r's
// row (n_r ints at g_r in HBM, at s + b_r in LDS) is moved by the whole
// wavefront, 64 consecutive entries per instruction, so one row costs a few
// cache lines instead of one line per lane per entry.  kStageU rows are in
// flight at a time.  Every lane must be active.
template <bool kToLds>
__device__ void move_rows(int *s, int b, int *g, int n, int lane) {
    for (int r0 = 0; r0 < 64; r0 += kStageU) {
        int v[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
            const int nr = __builtin_amdgcn_readlane(n, r0 + u);
            const int br = __builtin_amdgcn_readlane(b, r0 + u);
            const int *gr = lane_ptr(g, r0 + u);
            v[u] = lane < nr ? (kToLds ? gr[lane] : s[br + lane]) : 0;
        }
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
            const int nr = __builtin_amdgcn_readlane(n, r0 + u);
            const int br = __builtin_amdgcn_readlane(b, r0 + u);
            int *gr = lane_ptr(g, r0 + u);
            if (lane < nr) {
                if (kToLds)
                    s[br + lane] = v[u];
                else
                    gr[lane] = v[u];
            }
        }
    }
    // Rows longer than 64 entries: the rest, row by row.
    unsigned long long rest = __ballot(n > 64);
    while (rest) {
        const int r = __builtin_ctzll(rest);
        rest &= rest - 1;
        const int nr = __builtin_amdgcn_readlane(n, r);
        const int br = __builtin_amdgcn_readlane(b, r);
        int *gr = lane_ptr(g, r);
        for (int e = 64 + lane; e < nr; e += 64) {
            if (kToLds)
                s[br + e] = gr[e];
            else
                gr[e] = s[br + e];
        }
    }
}

// Pair force / derivative of neighbour record R on (xi, xidot), added
// (sign > 0) to or subtracted from (f, fd) as regcor_gpu.F:349-357 and
// :409-416 form them.
__device__ inline void pair_add(const JRec &R, const double *xi,
                                const double *xidot, bool add, double *f,
                                double *fd) {
    const double a1 = R.x - xi[0];
    const double a2 = R.y - xi[1];
    const double a3 = R.z - xi[2];
    const double dv0 = R.vx - xidot[0];
    const double dv1 = R.vy - xidot[1];
    const double dv2 = R.vz - xidot[2];
    const double rij2 = a1 * a1 + a2 * a2 + a3 * a3;
    const double dr2i = 1.0 / rij2;
    const double dr3i = R.m * dr2i * sqrt(dr2i);
    const double drdv = a1 * dv0 + a2 * dv1 + a3 * dv2;
    const double drdp = 3.0 * drdv * dr2i;
    const double a[3] = {a1, a2, a3};
    const double dv[3] = {dv0, dv1, dv2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (add) {
            f[k] = f[k] + a[k] * dr3i;
            fd[k] = fd[k] + (dv[k] - a[k] * drdp) * dr3i;
        } else {
            f[k] = f[k] - a[k] * dr3i;
            fd[k] = fd[k] - (dv[k] - a[k] * drdp) * dr3i;
        }
    }
}

// Pre-pass: every body index in range and listed once.
__global__ void k_regcor_mark(RegcorLaunch L) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L.ni) return;
    const int I = L.idx[q];
    if (I < 1 || I > L.par.ntot) {
        atomicOr(L.err, ERR_REGCOR);
        return;
    }
    atomicAdd(&L.mark[I - 1], 1);
}

// NLIST entry k at nl[(k - 1) nls]: the body's LDS row, or the HBM scratch
// NLG.  NLG and JJLIST are interleaved by body (entry k of body q at
// (k - 1) ni + q): the lanes of a wavefront, which walk their lists nearly
// in step, then share cache lines.  LIST entries through lrow (HBM).
#define NL(k) nl[(size_t)((k) - 1) * nls]
#define NLG(k) nlg[(size_t)((k) - 1) * ni_s]
#define LI(k) lrow[(k) - 1]
#define JJ(k) jj[(size_t)((k) - 1) * ni_s]

// Everything after the list checks for one body (index I, lane q): radius,
// list merge, irregular-force change, corrector, new list and steps.  With
// kLds, NLIST is the body's LDS row; the new list is left there for the
// wavefront to write to LIST, and the count of its entries is returned (0:
// LIST written here, or unchanged).
template <bool kLds>
__device__ __forceinline__ int regcor_body(const RegcorLaunch &L, int q, int I,
                                           int nnb0, int nnb, int *nl,
                                           size_t nls, int *lrow) {
    const gpunb_regcor_par &P = L.par;
    const gpunb_regcor_state &S = L.st;
    const int p = I - 1;
    const size_t ni_s = (size_t)L.ni;
    int *nlg = L.nlist + q;
    int *jj = L.jj + q;
    auto rec = [&](int J) -> const JRec & { return L.jp[J - P.ifirst]; };
    auto in_j = [&](int J) {
        return J >= P.ifirst && J - P.ifirst < L.nj && J <= P.ntot;
    };
    int ncopy = 0;
    bool nl_hbm = !kLds;

    double xi[3], xidot[3], freg[3], fdr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xi[k] = L.xi[3 * (size_t)q + k];
        xidot[k] = L.vi[3 * (size_t)q + k];
        freg[k] = L.freg[3 * (size_t)q + k];
        fdr[k] = L.fdr[3 * (size_t)q + k];
    }
    // The body's own state, read before anything is stored: the state
    // arrays may alias as far as the compiler knows, so a load placed after
    // a store would wait for it.  No other lane writes this body's entries.
    double fi0[3], fid0[3], fr0[3], frd0[3], x00[3], x0d0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fi0[k] = S.fi[3 * (size_t)p + k];
        fid0[k] = S.fidot[3 * (size_t)p + k];
        fr0[k] = S.fr[3 * (size_t)p + k];
        frd0[k] = S.frdot[3 * (size_t)p + k];
        x00[k] = S.x0[3 * (size_t)p + k];
        x0d0[k] = S.x0dot[3 * (size_t)p + k];
    }
    const double step0 = S.step[p], stepr0 = S.stepr[p];
    const double tnw0 = S.timenw[p], t00 = S.t0[p];
    double dfirr[3] = {0.0, 0.0, 0.0}, dfd[3] = {0.0, 0.0, 0.0};
    const double dtr = P.time - S.t0r[p];
    double ri2, rh2;
    if (P.nb_flag == 1) {
        ri2 = sqd(xi[0] - P.rdens[0]) + sqd(xi[1] - P.rdens[1]) +
              sqd(xi[2] - P.rdens[2]);
        rh2 = sqd(P.rscale);
    } else {
        ri2 = sqd(xi[0]) + sqd(xi[1]) + sqd(xi[2]);
        rh2 = 9.0 * sqd(P.rscale);
    }
    double rs = S.rs[p];
    const double rs2 = sqd(rs);
    int c_void = 0, c_smin = 0, c_flux = 0, c_conv = 0;

    if (nnb == 0) {
        // No neighbours: a small mass at the centre (regcor_gpu.F:119-143);
        // the radius update is skipped.
        const double r2 = sqd(xi[0]) + sqd(xi[1]) + sqd(xi[2]);
        const double fij = (double)0.01f * P.bodym / (r2 * sqrt(r2));
        const double rdot =
            3.0 * (xi[0] * xidot[0] + xi[1] * xidot[1] + xi[2] * xidot[2]) /
            r2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dfirr[k] = -(fij * xi[k]);
            dfd[k] = -((xidot[k] - rdot * xi[k]) * fij);
        }
        if (rdot > 0.0) {
            const double a = 0.75 * rs, b = (double)0.1f * P.rscale;
            rs = a > b ? a : b;
        } else {
            rs = (double)1.1f * rs;
        }
        c_void = 1;
    } else {
        // Neighbour radius (regcor_gpu.F:146-253).
        float ft = 0.1f * ((float)P.nnbopt - (float)nnb);
        ft = ft / (float)nnb;
        const double fac = (double)(1.0f + ft);
        double a3 = (double)(float)nnb * fac;
        int nbave;
        {
            int mx = P.nnbmax - 50;
            const int zi = (int)P.znbmin;
            if (zi > mx) mx = zi;
            nbave = P.nnbopt + 150 < mx ? P.nnbopt + 150 : mx;
        }
        {
            const double cap = (double)(0.9f * (float)nbave);
            if (cap < a3) a3 = cap;
        }
        const int nbp = (int)a3;
        if (ri2 > rh2 && P.nb_flag == 1) {
            double g = P.rscale / sqrt(ri2);
            if (g < (double)0.1f) g = (double)0.1f;
            a3 = a3 * g;
        } else if (a3 < P.znbmin) {
            a3 = P.znbmin;
        }
        double a4 = a3 / (double)(float)nnb;
        if ((a3 - (double)(float)nnb0) * (a3 - (double)(float)nnb) < 0.0)
            a4 = sqrt(a4);
        const double ridot = (xi[0] - P.rdens[0]) * xidot[0] +
                             (xi[1] - P.rdens[1]) * xidot[1] +
                             (xi[2] - P.rdens[2]) * xidot[2];
        if (ri2 > P.rc2 && P.nb_flag == 1 && ri2 < 100.0 * rh2)
            a4 = a4 * (1.0 + ridot * dtr / ri2);
        if (I > P.n && ri2 <= rh2) {
            const double a2 = 100.0 * S.body[p] / fabs(S.h[I - P.n - 1]);
            if (a2 > rs) {
                const float qf = 1.0f - (float)nnb / (float)nbave;
                a3 = (double)qf > 0.0 ? (double)qf : 0.0;
                a4 = 1.0 + 0.5 * a3;
            }
        }
        double a5 = P.rsfac * dtr;
        if (0.25 < a5) a5 = 0.25;
        a4 = a4 - 1.0;
        if (a4 > a5)
            a4 = a5;
        else if (a4 < -a5)
            a4 = -a5;
        a3 = kOne3 * a4;
        double a1 = 1.0 + a3 - a3 * a3;
        if (nnb < P.znbmin && nbp > P.znbmin && a1 < 1.0) a1 = (double)1.05f;
        if (fabs(a1 - 1.0) > (double)0.003f) {
            if (sqd(rs) > ri2 && sqd(rs) < 2.0 * ri2 && nnb > 5) a1 = sqrt(a1);
            rs = a1 * rs;
        }
        if (nnb > nbave && ridot < 0.0) {
            const float drs =
                ((float)nnb - (float)nbave) / (3.0f * (float)nnb);
            rs = rs * (1.0 - (double)drs);
        }
        if (P.m_flag == 0 && nnb <= 3 && ri2 < 100.0 * rh2) {
            double b1 = 2.0 * rs;
            for (int l = 1; l <= nnb; ++l) {
                const JRec &R = rec(NL(l + 1));
                const double rij = sqrt(sqd(xi[0] - R.x) + sqd(xi[1] - R.y) +
                                        sqd(xi[2] - R.z));
                const double rsdot = ((xi[0] - R.x) * (xidot[0] - R.vx) +
                                      (xi[1] - R.y) * (xidot[1] - R.vy) +
                                      (xi[2] - R.z) * (xidot[2] - R.vz)) /
                                     rij;
                const double c = rij + rsdot * dtr;
                if (c < b1) b1 = c;
            }
            const double c = (double)1.1f * rs;
            rs = b1 > c ? b1 : c;
        }
    }

    // Losses and gains (regcor_gpu.F:266-370): one pass over both sorted
    // lists with the sentinel NTOT + 1 after each.
    int nbloss = 0, nbgain = 0;
    if (nnb0 == 0) {
        nbgain = nnb;
        for (int l = 1; l <= nnb; ++l) JJ(l) = NL(l + 1);
    } else {
        int jmin = 0;
        int l = 2, lg = 2;
        NL(nnb + 2) = P.ntot + 1;
        NL(1) = LI(nnb0 + 1);
        // wb[d] = NL(lg + d): each step advances l and lg by 0 or 1, so the
        // NLIST entries kWin ahead are read kWin steps before they are
        // compared.  LIST is taken in rounds: kNWin entries from LI(l) for
        // every lane at once, one wait, then the walk runs on that window
        // (wa[0] = LI(l)) until it is used up, with no LIST loads inside.
        // The loop writes LIST only at LI(l) (the sentinel, with l = NNB0 +
        // 1 from then on) and never writes NLIST.
        int wb[kWin];
#pragma unroll
        for (int d = 0; d < kWin; ++d) wb[d] = lg + d <= nnb + 2 ? NL(lg + d) : 0;
        // Every old member passes through a window once, in order, and is
        // checked there (in the j set, strictly ascending): the walk ends
        // only after the last member, and LIST is written only after that.
        // Rows 16-byte aligned (LMAX a multiple of 4): five 16-byte loads
        // per window instead of sixteen single ones.
        const bool row16 = (((unsigned long long)L.st.list & 15) == 0) &&
                           (P.lmax & 3) == 0;
        bool done = false, bad = false;
        int vprev = 0;
        while (!done) {
            int wa[kNWin];
            if (row16) {
                const int e0 = l - 1, o = e0 & 3;
                const int4 *bp = (const int4 *)lrow + (e0 >> 2);
                int e[kNWin + 4];
#pragma unroll
                for (int i = 0; i < kNWin / 4 + 1; ++i) {
                    const int4 v = e0 - o + 4 * i <= nnb0 ? bp[i]
                                                           : make_int4(0, 0, 0, 0);
                    e[4 * i] = v.x;
                    e[4 * i + 1] = v.y;
                    e[4 * i + 2] = v.z;
                    e[4 * i + 3] = v.w;
                }
#pragma unroll
                for (int d = 0; d < kNWin; ++d) {
                    const int x = o == 0 ? e[d] : o == 1 ? e[d + 1]
                                : o == 2 ? e[d + 2] : e[d + 3];
                    wa[d] = l + d <= nnb0 + 1 ? x : 0;
                }
            } else {
#pragma unroll
                for (int d = 0; d < kNWin; ++d)
                    wa[d] = l + d <= nnb0 + 1 ? LI(l + d) : 0;
            }
            __builtin_amdgcn_s_waitcnt(kVmcnt0);
#pragma unroll
            for (int d = 0; d < kNWin; ++d) {
                if (l + d <= nnb0 + 1) {
                    if (!in_j(wa[d]) || (l + d > 2 && wa[d] <= vprev)) bad = true;
                    vprev = wa[d];
                }
            }
            if (bad) break;
            int avail = kNWin;
            while (avail > 0) {
                const int a = wa[0], b = wb[0];
                int l2 = l, lg2 = lg;
                if (a != b) {
                    if (a >= b) {
                        ++nbgain;
                        JJ(nnb0 + nbgain) = b;
                        --l2;
                    } else {
                        ++nbloss;
                        JJ(nbloss) = a;
                        --lg2;
                    }
                }
                bool sent = false;
                if (l2 <= nnb0) {
                    ++l2;
                    ++lg2;
                } else if (lg2 <= nnb) {
                    ++lg2;
                    LI(l2) = P.ntot + 1;       // l2 == l here
                    sent = true;
                } else {
                    done = true;
                    break;
                }
                if (l2 != l) {
#pragma unroll
                    for (int d = 0; d + 1 < kNWin; ++d) wa[d] = wa[d + 1];
                    --avail;
                } else if (sent) {
                    wa[0] = P.ntot + 1;
                }
                if (lg2 != lg) {
#pragma unroll
                    for (int d = 0; d + 1 < kWin; ++d) wb[d] = wb[d + 1];
                    wb[kWin - 1] = lg + kWin <= nnb + 2 ? NL(lg + kWin) : 0;
                }
                l = l2;
                lg = lg2;
            }
        }
        if (bad) {
            atomicOr(L.err, ERR_REGCOR);
            return 0;
        }
        for (int k0 = 1; k0 <= nbloss; k0 += kChunk) {
            int v[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                v[u] = k0 + u <= nbloss ? JJ(k0 + u) : 1;
            double sv[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) sv[u] = S.step[v[u] - 1];
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                if (k0 + u <= nbloss && sv[u] < P.smin) jmin = 1;
        }
        if (jmin != 0) {
            // Keep a lost neighbour with a small step inside 2 RS
            // (regcor_gpu.F:328-370), on NLIST in HBM (the LDS row has no
            // room to grow).
            if (kLds) {
                for (int k = 1; k <= nnb + 1; ++k) NLG(k) = NL(k);
                nl_hbm = true;
            }
            int k = 1;
            for (;;) {
                if (nnb > P.nnbmax || I > P.n) break;
                const int J = JJ(k);
                bool skip = S.step[J - 1] > P.smin || J < P.ifirst || J > P.n;
                const JRec &R = rec(J);
                if (!skip) {
                    const double rij2 = sqd(xi[0] - R.x) + sqd(xi[1] - R.y) +
                                        sqd(xi[2] - R.z);
                    if (rij2 > 4.0 * rs2) skip = true;
                }
                if (!skip) {
                    int ll = nnb + 1;
                    for (;;) {
                        if (NLG(ll) < J) break;
                        NLG(ll + 1) = NLG(ll);
                        --ll;
                        if (!(ll > 1)) break;
                    }
                    NLG(ll + 1) = J;
                    ++nnb;
                    --nbloss;
                    LI(nnb0 + 1) = NLG(1);
                    ++c_smin;
                    pair_add(R, xi, xidot, false, freg, fdr);
                    if (k > nbloss) break;
                    for (int m = k; m <= nbloss; ++m) JJ(m) = JJ(m + 1);
                    --k;
                }
                ++k;
                if (!(k <= nbloss)) break;
            }
        }
    }

    // Irregular force change of the list change (regcor_gpu.F:400-438),
    // in list order, kPairs members at a time: their indices, then their
    // records, are loaded together (past the end, the first member again).
    auto pair_run = [&](int first, int count, bool add) {
        if (count <= 0) return;
        const int j0 = JJ(first);
        for (int l0 = 0; l0 < count; l0 += kPairs) {
            int jv[kPairs];
#pragma unroll
            for (int u = 0; u < kPairs; ++u)
                jv[u] = l0 + u < count ? JJ(first + l0 + u) : j0;
            JRec rv[kPairs];
#pragma unroll
            for (int u = 0; u < kPairs; ++u) rv[u] = rec(jv[u]);
#pragma unroll
            for (int u = 0; u < kPairs; ++u)
                if (l0 + u < count) pair_add(rv[u], xi, xidot, add, dfirr, dfd);
        }
    };
    pair_run(1, nbloss, false);
    pair_run(nnb0 + 1, nbgain, true);

    // Corrector (regcor_gpu.F:440-551).
    const double dtsq = dtr * dtr;
    const double dt6 = 6.0 / (dtr * dtsq);
    const double dt2 = 2.0 / dtsq;
    double dtsq12 = kOne12 * dtsq;
    double dtr13 = kOne3 * dtr;
    S.t0r[p] = P.time;
    double stepi = step0;
    if (stepi < 5.0 * P.dtmin && dtr > 50.0 * stepi) {
        dtr13 = 0.0;
        dtsq12 = 0.0;
    }
    double fi[3], fid[3], d2r[3], d3r[3], xnd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fi[k] = fi0[k];
        fid[k] = fid0[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double dfr = fr0[k] - freg[k] - dfirr[k];
        const double fdr0 = fdr[k] + dfd[k];
        const double frd = frd0[k];
        const double sum = frd + fdr0;
        const double at3 = 2.0 * dfr + dtr * sum;
        const double bt2 = -3.0 * dfr - dtr * (sum + frd);
        S.xn[3 * (size_t)p + k] =
            x00[k] + ((double)0.6f * at3 + bt2) * dtsq12;
        xnd[k] = x0d0[k] + (0.75 * at3 + bt2) * dtr13;
        fi[k] = fi[k] + dfirr[k];
        fid[k] = fid[k] + dfd[k];
        if (k == 1) fid[0] = fid[0] - P.tidal4 * dfr;
        if (k == 0) fid[1] = fid[1] + P.tidal4 * dfr;
        d3r[k] = at3 * dt6;
        d2r[k] = (3.0 * at3 + bt2) * dt2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S.xndot[3 * (size_t)p + k] = xnd[k];
        S.fi[3 * (size_t)p + k] = fi[k];
        S.fr[3 * (size_t)p + k] = freg[k];
        S.fidot[3 * (size_t)p + k] = fid[k];
        S.frdot[3 * (size_t)p + k] = fdr[k];
        S.d0[3 * (size_t)p + k] = fi[k];
        S.d0r[3 * (size_t)p + k] = freg[k];
        S.d1r[3 * (size_t)p + k] = fdr[k];
        S.d2r[3 * (size_t)p + k] = d2r[k];
        S.d3r[3 * (size_t)p + k] = d3r[k];
        L.freg[3 * (size_t)q + k] = freg[k];
        L.fdr[3 * (size_t)q + k] = fdr[k];
    }
    S.rs[p] = rs;

    // New list (regcor_gpu.F:605-613): LIST(1) = NNB, LIST(2..NNB+1) =
    // NLIST(2..NNB+1).  An NLIST row in LDS becomes that list by its first
    // entry alone and is written to LIST by the wavefront; one in HBM is
    // copied here.
    if (nbloss + nbgain > 0) {
        if (kLds && !nl_hbm) {
            NL(1) = nnb;
            ncopy = nnb + 1;
        } else {
            LI(1) = nnb;
            for (int l0 = 2; l0 <= nnb + 1; l0 += kNWin) {
                int v[kNWin];
#pragma unroll
                for (int u = 0; u < kNWin; ++u)
                    v[u] = l0 + u <= nnb + 1 ? NLG(l0 + u) : 0;
#pragma unroll
                for (int u = 0; u < kNWin; ++u)
                    if (l0 + u <= nnb + 1) LI(l0 + u) = v[u];
            }
        }
        c_flux = nbloss + nbgain;
    }

    // New regular step (regcor_gpu.F:615-805).
    const double fr2 = sqd(freg[0]) + sqd(freg[1]) + sqd(freg[2]);
    const double w1 = sqd(fdr[0]) + sqd(fdr[1]) + sqd(fdr[2]);
    const double w2 = sqd(d2r[0]) + sqd(d2r[1]) + sqd(d2r[2]);
    const double w3 = sqd(d3r[0]) + sqd(d3r[1]) + sqd(d3r[2]);
    double w0 = (sqrt(fr2 * w2) + w1) / (sqrt(w1 * w3) + w2);
    w0 = P.etar * w0;
    double ttmp = sqrt(w0);
    const double dt0 = ttmp;
    const double steprI = stepr0;
    if (ttmp > 2.0 * steprI) {
        if (fmod(P.time, 2.0 * dtr) == 0.0) {
            ttmp = 2.0 * dtr < P.smax ? 2.0 * dtr : P.smax;
            for (int it = 0; it < kMaxHalve && dt0 > 2.0 * ttmp &&
                             2.0 * ttmp < P.smax &&
                             fmod(P.time, 2.0 * ttmp) == 0.0;
                 ++it)
                ttmp = 2.0 * ttmp;
        } else {
            ttmp = steprI;
        }
    } else if (ttmp < dtr) {
        ttmp = 0.5 * dtr;
        if (ttmp > dt0) ttmp = 0.5 * ttmp;
    } else {
        ttmp = steprI;
    }
    if (ttmp > (double)0.1f && ttmp > dtr) {
        double dv2 = 0.0;
        for (int k = 0; k < 3; ++k) dv2 = dv2 + sqd(xidot[k] - xnd[k]);
        const double dtj =
            dtr * pow(1.0e-04 * sqd(dtr) * fr2 / dv2, (double)0.1f);
        if (dtj < ttmp) ttmp = dtr;
    }
    if (P.smax < ttmp) ttmp = P.smax;
    S.stepr[p] = ttmp;
    double tnw = tnw0;
    if (nnb0 == 0 && nnb >= 0) {
        stepi = 0.125 * stepi;
        tnw = t00 + stepi;
    }
    // Step halving (regcor_gpu.F:798-801), capped at kMaxHalve passes:
    // enough to take any finite double to zero, so the cap only ends the
    // loop where the reference's would not end (ttmp <= 0).
    for (int it = 0; it < kMaxHalve && ttmp < stepi; ++it) {
        stepi = 0.5 * stepi;
        tnw = t00 + stepi;
        ++c_conv;
    }
    L.step_new[q] = stepi;
    L.tnw_new[q] = tnw;
    if (c_void) atomicAdd(&L.cnt[0], c_void);
    if (c_smin) atomicAdd(&L.cnt[1], c_smin);
    if (c_flux) atomicAdd(&L.cnt[2], c_flux);
    if (c_conv) atomicAdd(&L.cnt[3], c_conv);
    return ncopy;
}

// One wavefront of 64 bodies (lanes).  With kLds, each lane's new list is
// copied into its LDS row (cg + 2 entries at base, the list two entries in)
// and NLIST built over it in place; the merge reads it there and the old
// LIST row from HBM in rounds, and the new lists go back to LIST in
// coalesced sweeps.  Without kLds (the workgroup's rows exceed kStage),
// NLIST lives in the HBM scratch.
template <bool kLds>
__device__ __forceinline__ void regcor_wave(const RegcorLaunch &L, int *s_reg,
                                            int lane, int q, bool lv, int I,
                                            bool ok, int cg, int og,
                                            int nnb0, int base) {
    const gpunb_regcor_par &P = L.par;
    const gpunb_regcor_state &S = L.st;
    const int lmax = P.lmax;
    const size_t ni_s = (size_t)L.ni;
    auto in_j = [&](int J) {
        return J >= P.ifirst && J - P.ifirst < L.nj && J <= P.ntot;
    };
    int *grow = ok ? S.list + (size_t)(I - 1) * lmax : S.list;
    int *lrow = grow;
    const int *nsrc = ok ? L.lst + og : L.lst;
    int *nl = L.nlist + q;
    size_t nls = ni_s;
    if (kLds) {
        move_rows<true>(s_reg, base + 2, const_cast<int *>(nsrc),
                        ok ? cg : 0, lane);
        __syncthreads();
        nsrc = s_reg + base + 2;
        nl = s_reg + base;
        nls = 1;
    }
    int nnb = 0;
    // The new list kChunk entries at a time.  In LDS, NLIST(k) goes to
    // nl[k - 1] while raw entry e sits at nl[e + 2]: a chunk writes below
    // the next chunk's entries.
    if (ok) {
        int prev = 0;
        for (int k0 = 0; ok && k0 < cg; k0 += kChunk) {
            int v[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u)
                v[u] = k0 + u < cg ? nsrc[k0 + u] : 0;
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                // Predicated rather than branched: after the first bad entry
                // nothing more is written, and the body is rejected below.
                const int J = v[u] + P.ifirst;
                const bool live = ok && k0 + u < cg && J != I;
                const bool good = in_j(J) && !(nnb > 0 && J <= prev) &&
                                  nnb + 3 <= lmax;
                if (live && !good) ok = false;
                if (live && good) {
                    NL(nnb + 2) = J;
                    prev = J;
                    ++nnb;
                }
            }
        }
        NL(1) = nnb;
    }
    if (lv && !ok) atomicOr(L.err, ERR_REGCOR);
    int ncopy = 0;
    if (ok) ncopy = regcor_body<kLds>(L, q, I, nnb0, nnb, nl, nls, lrow);
    if (kLds) {
        __syncthreads();
        move_rows<false>(s_reg, base, grow, ncopy, lane);
    }
}

__global__ __launch_bounds__(64) void k_regcor(RegcorLaunch L) {
    __shared__ int s_reg[kStage];
    const int lane = threadIdx.x;
    const int q = blockIdx.x * 64 + lane;
    const gpunb_regcor_par &P = L.par;
    const gpunb_regcor_state &S = L.st;
    // lv: index in range (others are flagged by the pre-pass) and listed
    // once (a body listed twice is skipped in both places).
    bool lv = false;
    int I = 0;
    if (q < L.ni) {
        I = L.idx[q];
        if (I >= 1 && I <= P.ntot) {
            L.step_new[q] = S.step[I - 1];
            L.tnw_new[q] = S.timenw[I - 1];
            if (L.mark[I - 1] != 1)
                atomicOr(L.err, ERR_REGCOR);
            else
                lv = true;
        }
    }
    // New list in util_gpu.F:103-110's form (1-based, the body dropped);
    // both lists checked (members in the j set, strictly ascending) before
    // anything is modified.  With sorted in-range lists every index of the
    // list walks stays inside its row.
    const int lmax = P.lmax;
    int cg = 0, og = 0, nnb0 = 0;
    bool ok = false;
    if (lv) {
        cg = L.nnb[q];
        og = L.off[q];
        nnb0 = S.list[(size_t)(I - 1) * lmax];
        ok = cg >= 0 && og >= 0 && cg + 2 <= lmax &&
             (long long)og + cg <= L.list_len && nnb0 >= 0 &&
             nnb0 + 1 <= lmax && (I <= P.n || S.h != nullptr);
    }
    // LDS row of a lane: NLIST, built over its new list (cg + 2 entries).
    const int rsz = ok ? cg + 2 : 0;
    int total;
    const int base = wave_excl_scan(rsz, lane, &total);
    if (__builtin_amdgcn_readfirstlane(total) <= kStage)
        regcor_wave<true>(L, s_reg, lane, q, lv, I, ok, cg, og, nnb0, base);
    else
        regcor_wave<false>(L, s_reg, lane, q, lv, I, ok, cg, og, nnb0, base);
}
#undef NL
#undef NLG
#undef LI
#undef JJ

// New STEP / TIMENW after every body is corrected, and the marks cleared.
// A body listed twice reads its mark as 2, or as 0 once the other thread
// has cleared it: skipped either way.
__global__ void k_regcor_commit(RegcorLaunch L) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L.ni) return;
    const int I = L.idx[q];
    if (I < 1 || I > L.par.ntot) return;
    const int p = I - 1;
    if (L.mark[p] == 1) {
        L.st.step[p] = L.step_new[q];
        L.st.timenw[p] = L.tnw_new[q];
    }
    L.mark[p] = 0;
}

}  // namespace

void launch_regcor(hipStream_t st, const RegcorLaunch &L) {
    if (L.ni <= 0) return;
    const dim3 grid((unsigned)((L.ni + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_regcor_mark, grid, block, 0, st, L);
    hipLaunchKernelGGL(k_regcor, dim3((unsigned)((L.ni + 63) / 64)), dim3(64),
                       0, st, L);
    hipLaunchKernelGGL(k_regcor_commit, grid, block, 0, st, L);
}

}  // namespace gpunb
