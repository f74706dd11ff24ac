// gpunb_internal.h -- shared declarations between the HIP kernels
// (kernels.hip) and the host side of the C-ABI (gpunb.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <functional>

#include "gpunb.h"

namespace gpunb {

// One j-particle in HBM: 64 bytes, so a wavefront fetches a whole record with
// one scalar load (s_load_dwordx16) and feeds it to VALU ops as SGPR operands.
// FP64 position, velocity and mass are kept because the pair differences and
// the neighbour predicate are evaluated from FP64 values; m32 is the FP32 mass
// used by the force sums.
struct alignas(64) JRec {
    double x, y, z;
    double vx, vy, vz;
    double m;
    float m32;
    float pad;
};
static_assert(sizeof(JRec) == 64, "JRec must be 64 bytes");

// One particle for the potential kernel: 32 bytes (s_load_dwordx8).
struct alignas(32) PRec {
    double x, y, z;
    float m32;
    float pad;
};
static_assert(sizeof(PRec) == 32, "PRec must be 32 bytes");

// The j set in Hilbert order, as the FP32 pair loop reads it.  The sorted
// positions are grouped in tiles of kTile spatially close j; every j is held
// as an FP32 offset from its tile's FP64 centre, so a pair difference is
// (C_t - x_i) + o_j: the first term once per (i, tile) in FP64 rounded to
// FP32, the second a plain FP32 add per pair.  32 bytes: one s_load_dwordx8.
struct alignas(32) JRec32 {
    float ox, oy, oz;    // position - tile centre (FP32 rounded)
    float m;             // mass (FP32)
    float ovx, ovy, ovz; // velocity - tile velocity centre
    float w;             // near-band weight max(1, x^(2/3), 1.4 x / K), x = m / <m>
};
static_assert(sizeof(JRec32) == 32, "JRec32 must be 32 bytes");

// One tile of the sorted j set: FP64 centres (midpoints of the bounding
// boxes of the tile's positions and velocities), upper bounds on |o| and
// |ov| of its FP32 records, and the largest weight w.  64 bytes.
struct alignas(64) JTile {
    double cx, cy, cz;
    double cu, cv, cw;
    float e;     // >= max |(ox, oy, oz)| over the tile's real j
    float ev;    // >= max |(ovx, ovy, ovz)|
    float wmax;  // >= max w
    float eb;    // (kTileBand e)^2: pairs closer than kTileBand e to the
                 // tile centre's offset scale are evaluated in FP64
};
static_assert(sizeof(JTile) == 64, "JTile must be 64 bytes");

// Per-cluster statistics of the j set, reduced in a fixed order (so every
// run derives the same thresholds from them): bounding box of the positions
// (the curve quantisation box) and the mean mass.
struct JClus {
    double lo[3];
    double hi[3];
    double mbar;
    double pad;
};

// Wavefronts per pass-1 workgroup: they split one j slice of one
// i-wavefront and reduce their sums in LDS, so a slice partial covers
// kP1W times the tiles one wavefront runs.
constexpr int kP1W = 4;

// Near band of the FP32 pair loop: a tile holding a pair with min(r2, q) <
// K_BAND * h2 * w_j (times <m> with mass-weighted lists) for some lane of a
// wave is evaluated in FP64 by that wave.
constexpr float kBand = 12.0f;
// j heavier than kHeavy <m> go to tiles of their own (after the light j of
// their cluster): light tiles then carry w_j <= max(1, kHeavy^(2/3)) ~ 2.
constexpr float kHeavy = 2.8f;
// Tile floor of the near band: the tile-relative FP32 difference of a pair
// carries up to 2u (|dx| + e_t) per component, so pairs with |dx| below
// kTileBand e_t are also evaluated in FP64 (this matters only where the
// neighbour sphere is small against the tile, e.g. h2 = 0).
constexpr float kTileBand = 0.5f;

// Cluster of a j preparation: j range [j0, j0 + ntile * kTile) of the
// original and sorted arrays (tile aligned), nj real j at its start, and its
// first statistics block.
struct PrepClus {
    int j0;
    int nj;
    int ntile;
    int blk0;
};

// Buffers of the sorted j view (one per engine; the ensemble has its own).
struct JView {
    const JRec *jp;     // original order (input of the preparation)
    JRec *jps;          // FP64 records in Hilbert order
    JRec32 *j32;        // FP32 tile-relative records in Hilbert order
    JTile *tiles;
    int *jidx;          // sorted position -> original j (local to the cluster)
    JClus *jcl;         // per cluster
};

// Statistics of the current j set, computed on the device at send/predict:
// max |v_j| and min positive m_j, stored as the bit patterns of the FP64
// values (non-negative doubles order like their bit patterns, so integer
// atomicMax/atomicMin reduce them).
struct JStats {
    unsigned long long vmax_bits;
    unsigned long long mmin_bits;
};

// Device error bits (OR-ed by kernels, checked by the host at sync points).
enum : unsigned {
    ERR_NAN_J = 1u,      // NaN in j data (send / jpset)
    ERR_NAN_I = 2u,      // NaN in i data (regf)
    ERR_LIST_CAP = 4u,   // compact list buffer too small (device entry point)
    ERR_REGCOR = 8u,     // regular corrector: invalid body index or list
    ERR_ENERGY = 16u,    // energy sum: invalid list row
};

// Far-away parking place of the padding j records (mass 0).
constexpr double kPadPos = 1.0e18;

// Tile of j-particles between FP32 -> FP64 folds of the partial sums.
constexpr int kTile = 64;
// j per tile of the sorted view (FP32 offsets from a tile centre, the far
// test of the mixed kernel): kTile / kVTile view tiles per tile.  32-j view
// tiles leave fewer tiles near some lane (King N = 131072: 22.6% against
// 26.9% of (wavefront, tile) pairs) but pay the per-tile set-up twice as
// often: whole step 15.40-15.46 against 15.28-15.34 ms on one GPU, so view
// tiles are whole tiles.
constexpr int kVTile = 64;
static_assert(kTile % kVTile == 0 && kVTile <= 64, "view tiles split a tile");

// One independent cluster of a regular-force launch.  A launch covers ncl
// clusters (the drop-in path uses one per i chunk, the ensemble path one per
// star cluster); pass-1 tasks and pass-2 workgroups of all clusters form one
// grid, each block finding its cluster by binary search over task0 / blk0.
struct ClusterDesc {
    long long task0;   // first pass-1 task (wavefront) of the cluster
    long long pbase;   // first (i, slice) pair of the cluster in part
    long long sbase;   // first element of the cluster's list entries in nbl
    long long blk0;    // first pass-2 workgroup of the cluster
    int j0;            // first JRec of the cluster (multiple of kTile)
    int nj;            // real j count
    int ntile;         // j tiles (kTile j each, padded)
    int i0;            // first i of the cluster in the i / output arrays
    int ni;
    int n_iw;          // i wavefronts: ceil(ni / 64)
    int nslice;        // j slices per i wavefront
    int cap;           // list entries stored per i (max(nnbmax, 1))
    int ipb;           // pass-2 i per workgroup
    int jc;            // index of the cluster's JClus
    int ibase;         // first i of the cluster in the list counts nbc
};

struct RegfLaunch {
    // j side: the sorted view (pass 1 reads the sorted records, pass 2 maps
    // sorted positions back to original indices)
    const JRec *jps;
    const JRec32 *j32;
    const JTile *tiles;
    const int *jidx;     // NULL: records in the caller's order
    const JClus *jcl;
    int nj;          // real j count
    int ntile;       // padded j count / kTile
    const JStats *jst;
    // i side (device, AoS like the ABI)
    const double *xi, *vi, *h2, *dtr;
    int ni;
    // decomposition
    int nslice;      // j slices per i-wave
    // pass-1 scratch
    double *part;    // [nslice][ni_pad][8] slice partials (slice-major: a
                     // pass-1 workgroup writes 64 consecutive records)
    int *nbc;        // [ni] list counts: zero before pass 1, reset by pass 2
    int *nbl;        // [ni][cap] list entries (sorted positions)
    // i in sorted order: output row of the i at (call-wide) position p is
    // perm[p]; NULL when the i arrays are in the caller's order
    const int *perm;
    // outputs
    double *acc, *jrk, *pot;
    int *nnb, *nnid, *list, *off, *total;
    long long list_cap;
    int nnbmax;
    int m_flag;
    int group;       // pairs per branch-free group of pass 1 (2 or 4)
    int prec;        // 32: FP32 pair arithmetic; 64/65: FP64 (+ Newton steps)
    unsigned *err;
    // cluster table (device) and grid sizes
    const ClusterDesc *desc;
    int ncl;
    long long ntask;   // pass-1 tasks over all clusters
    long long nblk2;   // pass-2 workgroups over all clusters
    // Persistent pass 1 (queue != NULL): pgrid workgroups take tasks from
    // 8 per-XCD counters (cleared by the launch); XCD x owns the tasks
    // [xrange[x], xrange[x + 1]).
    unsigned *queue;
    const long long *xrange;
    int pgrid;
};

// Kernel launchers (kernels.hip).  All asynchronous on `st`.
void launch_pack_j(hipStream_t st, const double *m, const double *x,
                   const double *v, int nj, int nj_pad, JRec *jp, JStats *jst,
                   unsigned *err);
void launch_stats_reset(hipStream_t st, JStats *jst, int *total);
void launch_predict(hipStream_t st, double t, int nj, int nj_pad,
                    const double *x0, const double *v0, const double *f,
                    const double *fd, const double *t0, const double *m,
                    JRec *jp, JStats *jst);
void launch_jpset(hipStream_t st, int n, const int *addr, const double *src,
                  double *x0, double *v0, double *f, double *fd, double *mass,
                  double *t0, unsigned *err);
void launch_jget(hipStream_t st, int n, const JRec *jp, double *x, double *v);

// j preparation (jprep.hip): per-cluster statistics, Hilbert sort and the
// records of the sorted view, for ncl clusters whose ranges (device table pc)
// tile [0, ntile_all * kTile).  nblk = sum over clusters of
// prep_stat_blocks(nj).
constexpr int kStatChunk = 2048;   // j per statistics block
inline int prep_stat_blocks(int nj) {
    return nj > 0 ? (nj + kStatChunk - 1) / kStatChunk : 1;
}
struct PrepBufs {
    unsigned long long *key_in, *key_out;
    int *val_in, *val_out;
    void *tmp;
    size_t tmp_bytes;
    double *part;      // 8 doubles per statistics block
};
// rocprim temporary storage for sorting n keys of ncl clusters.
size_t jprep_tmp_bytes(int n, int ncl);
hipError_t launch_jprep(hipStream_t st, const JView &V, const PrepClus *pc,
                        int ncl, int nblk, int ntile_all, const PrepBufs &B);
// Hilbert order of the i of one call (jprep.hip): perm[p] = the original i
// at sorted position p, and the i data gathered into that order.
struct ISortBufs {
    unsigned *key_in, *key_out;                 // one cluster (32-bit keys)
    unsigned long long *key64_in, *key64_out;   // ensemble (cluster << 32 | code)
    int *val_in, *perm;
    void *tmp;
    size_t tmp_bytes;
    double *xs, *vs, *hs, *ds;   // [ni][3], [ni][3], [ni], [ni]
};
size_t isort_tmp_bytes(int n);
hipError_t launch_isort(hipStream_t st, const JClus *jcl, const double *xi,
                        const double *vi, const double *h2, const double *dtr,
                        int ni, const ISortBufs &B);
// Ensemble: the i of every cluster in Hilbert order within the cluster (its
// range [cl_i0[c], cl_i0[c + 1]) of the concatenated i kept), keyed in the
// box of the cluster's j set.
size_t isort_ens_tmp_bytes(int n, int ncl);
hipError_t launch_isort_ens(hipStream_t st, const JClus *jcl, const int *cl_i0,
                            int ncl, const double *xi, const double *vi,
                            const double *h2, const double *dtr, int ni,
                            const ISortBufs &B);
// *pc = v, in stream order.
void launch_set_prepclus(hipStream_t st, PrepClus *pc, const PrepClus &v);
// Padding records in j32[0, n) and empty tiles in tiles[0, ntile).
void launch_park_view(hipStream_t st, JRec32 *j32, int n, JTile *tiles,
                      int ntile);
void launch_regf_pass1(hipStream_t st, const RegfLaunch &L);
void launch_regf_pass2(hipStream_t st, const RegfLaunch &L);
int pass2_ipb(int ni);   // pass-2 i per workgroup for a cluster of ni
// Entries of a pass-2 wavefront's LDS list row (>= nnbmax + 1); the row of
// the four wavefronts of a workgroup must fit the 64 KiB of LDS a
// workgroup can address, hence the largest supported nnbmax.
int pass2_sortcap(int nnbmax);
constexpr int kMaxNnbmax = 4000;
void launch_zero_int(hipStream_t st, int *p, int n);
// *d = v, in stream order (a cluster table without a host buffer).
void launch_set_desc(hipStream_t st, ClusterDesc *d, const ClusterDesc &v);
// bytes (a multiple of 16, both pointers 16-byte aligned) from device memory
// to a mapped pinned host buffer, and *err to *herr (host-mapped).
void launch_copy_to_host(hipStream_t st, const void *src, void *dst,
                         size_t bytes, const unsigned *err, unsigned *herr);
// bytes (a multiple of 16, both pointers 16-byte aligned) from a mapped
// pinned host buffer to device memory, and *zero = 0.
void launch_copy_from_host(hipStream_t st, const void *src, void *dst,
                           size_t bytes, int *zero);
void launch_pot(hipStream_t st, const PRec *p, int n, int ntile, int i0,
                int ni, int nslice, double *part, double *pot);
void launch_pack_p(hipStream_t st, const double *m, const double *x, int n,
                   int n_pad, PRec *p, unsigned *err);

// One regular-corrector call (regcor.hip).
struct RegcorLaunch {
    gpunb_regcor_par par;
    gpunb_regcor_state st;
    int ni;
    const int *idx;
    const double *xi, *vi;
    double *freg, *fdr;
    const int *nnb, *off, *lst;   // compact lists of the regular force
    long long list_len;           // readable entries of lst
    const JRec *jp;               // predicted j records
    int nj;
    int *nlist;                   // scratch [lmax][ni]: NLIST
    int *jj;                      // scratch [2 lmax][ni]: JJLIST
    double *step_new, *tnw_new;   // scratch [ni]
    int *mark;                    // [ntot], zero between calls
    int *cnt;                     // [4] counters
    unsigned *err;
};
void launch_regcor(hipStream_t st, const RegcorLaunch &L);

// One potential-energy sum (energy.hip).
struct EnergyLaunch {
    int n, ntot, ifirst, lmax;
    double cmsep2_rmin2;
    const double *body, *x, *xdot;
    const int *list;
    double *phidbl;
    double *cpot, *ckin;          // scratch [ntot - ifirst + 1]
    double *ppart, *kpart;        // scratch [ceil((ntot - ifirst + 1) / 2048)]
    double *out;                  // [2]: POT, ZKIN
    unsigned *err;
};
void launch_energy(hipStream_t st, const EnergyLaunch &L);

// Occupancy-derived resident-wave capacity of the regular-force kernel.
int regf_resident_waves(int m_flag, int group, int prec);
int pot_resident_waves();

// Runs f under the library's error convention (the message goes to
// gpunb_last_error(), the process aborts unless aborting is switched off):
// 0 on success, -1 after an exception.  For entry points defined outside
// gpunb.hip.
int run_guarded(const char *fn, const std::function<void()> &f);

}  // namespace gpunb
