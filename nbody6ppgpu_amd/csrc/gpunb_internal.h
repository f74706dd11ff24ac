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

// One independent cluster of a regular-force launch.  A launch covers ncl
// clusters (the drop-in path uses one per i chunk, the ensemble path one per
// star cluster); pass-1 tasks and pass-2 workgroups of all clusters form one
// grid, each block finding its cluster by binary search over task0 / blk0.
struct ClusterDesc {
    long long task0;   // first pass-1 task (wavefront) of the cluster
    long long pbase;   // first (i, slice) pair of the cluster in part / cnt
    long long sbase;   // first element of the cluster's segments in seg
    long long blk0;    // first pass-2 workgroup of the cluster
    int j0;            // first JRec of the cluster (multiple of kTile)
    int nj;            // real j count
    int ntile;         // j tiles (kTile j each, padded)
    int i0;            // first i of the cluster in the i / output arrays
    int ni;
    int n_iw;          // i wavefronts: ceil(ni / 64)
    int nslice;        // j slices per i wavefront
    int cap;           // neighbour segment capacity per (i, slice)
    int ipb;           // pass-2 i per workgroup
    int pad;
};

struct RegfLaunch {
    // j side
    const JRec *jp;
    int nj;          // real j count
    int ntile;       // padded j count / kTile
    const JStats *jst;
    // i side (device, AoS like the ABI)
    const double *xi, *vi, *h2, *dtr;
    int ni;
    // decomposition
    int nslice;      // j slices per i-wave
    int cap;         // neighbour segment capacity per (i, slice)
    // pass-1 scratch
    double *part;    // [ni_pad][nslice][8]
    int *cnt;        // [ni_pad][nslice]
    uint16_t *seg;   // [ni_pad][nslice][cap]
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
void launch_regf_pass1(hipStream_t st, const RegfLaunch &L);
void launch_regf_pass2(hipStream_t st, const RegfLaunch &L);
int pass2_ipb(int ni);   // pass-2 i per workgroup for a cluster of ni
void launch_zero_int(hipStream_t st, int *p, int n);
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
