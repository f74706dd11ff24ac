/*
 * gpunb.h -- C-ABI of the MI355X regular-force engine for NBODY6++GPU.
 *
 * The first block is the drop-in boundary: the exact symbols, argument order
 * and by-reference convention of the reference's regular-force libraries, so
 * the unchanged Fortran host (implicit interfaces, trailing underscore, every
 * argument by reference) links against this library in place of the CUDA
 * `gpunb` + `gpupot` objects.  Every entry point cites the reference function
 * it replaces (paths relative to the reference tree).
 *
 * Array conventions (Fortran column-major == C row-major of the transposed
 * shape):
 *   Fortran X(3,N)        == C double x[N][3]
 *   Fortran LISTGP(LMAX,NI)== C int    list[NI][LMAX]
 * Neighbour and nearest-neighbour indices are 0-based, relative to the first
 * j-particle passed to gpunb_send_ (the caller adds IFIRST, util_gpu.F:105).
 *
 * Error behaviour: like the reference (assert/abort on precondition or device
 * errors, gpunb.gpu.cu:703-709, cuda_pointer.h:20-42) the library prints a
 * line "gpunb: ..." on stderr and aborts.  A caller that prefers to handle
 * errors itself calls gpunb_set_abort_(&zero) once; failing calls then return
 * with their outputs untouched and gpunb_last_error() reports the message.
 * Neighbour-list overflow is not an error: it is reported in-band as a
 * negative count, as in the reference.
 *
 * Threading: calls come from one host thread (util_gpu.F:60 sits between
 * OpenMP loops); the library is not re-entrant, as the reference is not.
 */
#ifndef GPUNB_MI355X_H
#define GPUNB_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in boundary (reference ABI)                                          */
/* ------------------------------------------------------------------------ */

/* Select the device: first entry of the GPU_LIST environment variable if set,
 * otherwise irank % device_count; prints the device name once.  Idempotent.
 * Replaces GPUNB_devinit / gpunb_devinit_ (src/Main/gpunb.gpu.cu:557-590,822-824). */
void gpunb_devinit_(int *irank);

/* Open the library for up to *nbmax j-particles.  A second open without close
 * prints "gpunb: it is already open" and returns.
 * Replaces GPUNB_open / gpunb_open_ (src/Main/gpunb.gpu.cu:592-634,825-827). */
void gpunb_open_(int *nbmax, int *irank);

/* Release device and pinned buffers.  A second close prints a warning.
 * Replaces GPUNB_close / gpunb_close_ (src/Main/gpunb.gpu.cu:636-668,828-830). */
void gpunb_close_(void);

/* Upload *nj predicted j-particles (mass, position, velocity; FP64).  They stay
 * resident in device memory for any number of following gpunb_regf_ calls.
 * Requires *nj <= nbmax of the open.  NaN inputs are an error.
 * Replaces GPUNB_send / gpunb_send_ (src/Main/gpunb.gpu.cu:670-688,831-833). */
void gpunb_send_(int *nj, double *mj, double (*xj)[3], double (*vj)[3]);

/* Regular force of *ni i-particles against all j-particles last sent.
 *   h2[i]  squared neighbour radius RS^2 (RS^2/<m> when *m_flag != 0)
 *   dtr[i] regular step, used by the predicted-distance criterion
 * Outputs, for every i (FP64): far-field acceleration acc[i][3], jerk
 * jrk[i][3] and potential pot[i] = sum over non-neighbours of m_j/r_ij
 * (positive), and the neighbour list in list[i*lmax + 0 .. nnb]:
 *   list[i*lmax]      = nnb             if 0 <= nnb <= *nnbmax
 *                     = -nnb            if nnb > *nnbmax (overflow; rest undefined)
 *   list[i*lmax+1..]  = j indices, strictly ascending, self included
 * j is a neighbour of i iff min(r2, r2p) < h2[i]   (m_flag == 0)
 *                      or   min(r2, r2p) < m_j*h2[i] (m_flag != 0),
 * with dx = x_j - x_i, dv = v_j - v_i, r2 = |dx|^2 and r2p = |dx + dtr*dv|^2,
 * all evaluated in FP64 exactly as written (no fused multiply-add); pairs
 * with r2 == 0 never contribute to acc/jrk/pot.
 * Any *ni >= 0 is accepted (the reference asserts 0 < ni <= 1024 and the
 * Fortran host chunks at 1024, util_gpu.F:33-37); *lmax must be > *nnbmax.
 * Replaces GPUNB_regf / gpunb_regf_ (src/Main/gpunb.gpu.cu:690-796,834-838;
 * CPU twin src/Main/reg.avx.cpp:185-331). */
void gpunb_regf_(int *ni, double *h2, double *dtr, double (*xi)[3],
                 double (*vi)[3], double (*acc)[3], double (*jrk)[3],
                 double *pot, int *lmax, int *nnbmax, int *list, int *m_flag);

/* Print the per-interval profile line (send/grav/nb/out seconds and Gflops
 * at the reference's 60 flops per pair interaction) and reset the counters.
 * Replaces GPUNB_profile / gpunb_profile_ (src/Main/gpunb.gpu.cu:799-819,839-841). */
void gpunb_profile_(int *irank);

/* Potential phi_k = sum over j in [0,*n) with r2 > 0 of m_j / r_jk for the
 * particles k = *istart-1 .. *istart-2+*ni (1-based *istart, as energy.F:38
 * calls it); pot[0..*ni-1] receives them (positive values).  Independent of
 * the open/send state (it calls gpunb_devinit_ itself).
 * Replaces gpupot / gpupot_ (src/Main/gpupot.gpu.cu:61-127; CPU twin
 * src/Main/pot.avx.cpp:72-159). */
void gpupot_(int *irank, int *istart, int *ni, int *n, double *m,
             double (*x)[3], double *pot);

/* ------------------------------------------------------------------------ */
/* Additive entry points (no reference symbol is changed)                    */
/* ------------------------------------------------------------------------ */

/* gpunb_regf_ plus the nearest-neighbour index: nnid[i] = the listed j with
 * r2 > 0 of smallest FP64 r2 (ties: smaller j), -1 if there is none or the
 * list overflowed.  Semantics of IMINR in src/Main/nbint.F:29-47 (strict <,
 * first in ascending list order); the reference computes it on the host or in
 * the irregular library (src/Main/irr.avx.cpp:317-320), never in regf. */
void gpunb_regfnn_(int *ni, double *h2, double *dtr, double (*xi)[3],
                   double (*vi)[3], double (*acc)[3], double (*jrk)[3],
                   double *pot, int *lmax, int *nnbmax, int *list,
                   int *m_flag, int *nnid);

/* Device-resident j predictor, part 1: store the Hermite state of *n
 * j-particles at the 1-based j slots addr[k] (slot 1 = first j):
 * x0, v0 (X0, X0DOT), f = a/2 (F), fdot = jerk/6 (FDOT), mass and t0.
 * Mirrors the per-particle irr_simd_set_jp_ (src/Main/irr.avx.cpp:582-592)
 * as one bulk call.  Slots above the current j count extend it. */
void gpunb_jpset_(int *n, int *addr, double (*x0)[3], double (*v0)[3],
                  double (*f)[3], double (*fdot)[3], double *mass,
                  double *t0);

/* Device-resident j predictor, part 2: predict every stored j-particle to
 * *time and make the result the j set used by gpunb_regf_ (FP64):
 *   s = time - t0;  x = ((fdot*s + f)*s + v0)*s + x0;
 *   v = (fdot*1.5s + f)*2s + v0
 * Replaces the host XBPREDALL + GPUNB_SEND pair (src/Main/xbpredall.f:15-27,
 * src/Main/intgrt.F:515-521,912-918) for a host that keeps the j state on
 * the device; the KS pair resolution of xbpredall.f:29-35 stays on the host. */
void gpunb_predict_(double *time);

/* Copy the current (predicted or sent) j positions/velocities back to the
 * host for the first *n j-particles. */
void gpunb_jpget_(int *n, double (*x)[3], double (*v)[3]);

/* Error handling switch: *on != 0 (default) aborts on error like the
 * reference; *on == 0 returns instead and records the message. */
void gpunb_set_abort_(int *on);

/* Last error message ("" if none) and clear it.  Returns its length. */
int gpunb_last_error(char *buf, int len);

/* ------------------------------------------------------------------------ */
/* Device-resident entry points (plain device pointers, asynchronous on the  */
/* library stream).  For callers that keep i data and results in HBM.        */
/* ------------------------------------------------------------------------ */

/* Regular force with every input and output in device memory, enqueued on
 * the library stream and not synchronised.
 *   d_xi, d_vi : double[ni][3]     d_h2, d_dtr : double[ni]
 *   d_acc, d_jrk : double[ni][3]   d_pot : double[ni]
 *   d_nnb : int[ni]  (count, or -count on overflow, as list[i*lmax])
 *   d_nnid: int[ni]  (may be NULL: then the nearest neighbour is not searched)
 *   d_list: int[list_cap] compact neighbour lists; d_off[i] = start of i's
 *           list in d_list (-1 on overflow); d_total: int[1], set to the
 *           number of list entries written (the order of the per-i blocks in
 *           d_list is not fixed, their content is).
 * Returns 0, or -1 with gpunb_last_error() set (including list_cap too small,
 * which is reported after the fact by gpunb_sync(): the lists are then
 * incomplete). */
int gpunb_regf_dev(int ni, const double *d_h2, const double *d_dtr,
                   const double *d_xi, const double *d_vi, double *d_acc,
                   double *d_jrk, double *d_pot, int nnbmax, int m_flag,
                   int *d_nnb, int *d_nnid, int *d_list, int *d_off,
                   int *d_total, long long list_cap);

/* gpunb_regf_dev that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and asynchrony as gpunb_regf_dev; its scratch (cluster
 * table, list counts and entries, slice partials) lives in a 48 MiB device
 * array of the library, one per GPU, held until the process ends.  A call
 * needs (ni rounded up to 64) * (4 + 4 max(nnbmax, 1) + 64) bytes of it
 * (fewer j slices are used when the partials of the default decomposition
 * do not fit); a larger call is refused (-1, gpunb_last_error() says so).
 * The i run in the caller's order (no Hilbert sort).  Calls share the array:
 * two calls that overlap on one GPU race (calls on the library stream run
 * in order). */
int gpunb_regf_dev_noalloc(int ni, const double *d_h2, const double *d_dtr,
                           const double *d_xi, const double *d_vi,
                           double *d_acc, double *d_jrk, double *d_pot,
                           int nnbmax, int m_flag, int *d_nnb, int *d_nnid,
                           int *d_list, int *d_off, int *d_total,
                           long long list_cap);

/* ------------------------------------------------------------------------ */
/* Ensemble of independent clusters (no reference counterpart: the reference */
/* runs one cluster per process, -D ENSEMBLE, nbody6.F:35).  All clusters'   */
/* j sets stay in HBM and one launch computes every cluster's regular force; */
/* indices in lists are local to the cluster.  Asynchronous, device arrays.  */
/* ------------------------------------------------------------------------ */

/* Open for ncl clusters of at most nbmax[c] j-particles each. */
int gpunb_ens_open(int ncl, const int *nbmax);
int gpunb_ens_close(void);

/* Send all clusters' j-particles: host arrays concatenated in cluster order
 * (nj[c] bodies of cluster c; m[], x[][3], v[][3]). */
int gpunb_ens_send(int ncl, const int *nj, const double *m, const double *x,
                   const double *v);

/* Regular force of ni[c] i-particles of every cluster c against that cluster's
 * j set, one launch pair for all.  i data and outputs are device arrays
 * concatenated in cluster order (layouts as gpunb_regf_dev; d_off indexes
 * the shared compact list, d_total receives its length). */
int gpunb_ens_regf_dev(int ncl, const int *ni, const double *d_h2,
                       const double *d_dtr, const double *d_xi,
                       const double *d_vi, double *d_acc, double *d_jrk,
                       double *d_pot, int nnbmax, int m_flag, int *d_nnb,
                       int *d_nnid, int *d_list, int *d_off, int *d_total,
                       long long list_cap);

/* enable != 0: capture the ensemble launch sequence into a hipGraph and
 * replay it while sizes and pointers stay the same (re-captured otherwise). */
int gpunb_ens_graph(int enable);

/* enable != 0 (the default): pass 1 of the ensemble launch runs as a
 * persistent grid of resident workgroups that take (cluster, i-wavefront,
 * j slice) tasks from per-XCD atomic queues, largest tasks first, with the
 * task list cut into XCD ranges of equal work (a workgroup whose range is
 * used up takes tasks from the next ranges); 0: one workgroup per task.
 * Results are identical either way.  Returns 0 or -1. */
int gpunb_ens_persistent(int enable);

/* Autotuner output: j tiles (64 j) per slice for clusters whose j count lies
 * in [2^b, 2^(b+1)), b < nbucket; 0 = automatic.  Returns 0 or -1. */
int gpunb_ens_set_bucket_tiles(int nbucket, const int *tiles);

/* Pair interactions of the last ensemble launch (sum over clusters ni*nj). */
long long gpunb_ens_pairs(void);

/* ------------------------------------------------------------------------ */
/* Irregular (neighbour) force: drop-in for the reference's irr_simd library */
/* (irr.avx.cpp:565-603; callers intgrt.F:175 open, :202-204 and :1270      */
/* set_jp, :204/:971/:1194 set_list, :545/:680 firr_vec, :1502 close).      */
/* Addresses and list members are 1-based, as the Fortran host passes them.  */
/* Pair arithmetic is FP64 (the reference: FP32 on double-single positions). */
/* ------------------------------------------------------------------------ */

/* Replaces irr_simd_open_ (irr.avx.cpp:567-569): state for nmax particles,
 * list rows of lmax ints (count + members).  A second open warns. */
void irr_simd_open_(int *nmax, int *lmax, int *rank);
/* Replaces irr_simd_close_ (irr.avx.cpp:570-572). */
void irr_simd_close_(int *rank);
/* Replaces irr_simd_profile_ (irr.avx.cpp:573-575): calls, <NI>, <NB>, time,
 * Gflops at 60 flops per pair; counters reset. */
void irr_simd_profile_(int *rank);
/* Replaces irr_simd_set_jp_ (irr.avx.cpp:577-587): predictor state of one
 * particle: position, velocity, acc2 = a/2, jrk6 = jerk/6, mass, time.
 * Staged on the host, uploaded by the next firr_vec. */
void irr_simd_set_jp_(int *addr, double pos[3], double vel[3], double acc2[3],
                      double jrk6[3], double *mass, double *time);
/* Replaces irr_simd_set_list_ (irr.avx.cpp:588-593): nblist[0] = count,
 * nblist[1..count] = members (1-based).  Staged like set_jp. */
void irr_simd_set_list_(int *addr, int *nblist);
/* Replaces irr_simd_firr_vec_ (irr.avx.cpp:594-602): irregular acceleration
 * and jerk at time ti of the ni particles addr[] over their lists, every
 * particle predicted to ti from its own state; nnbid[q] = 1-based nearest
 * member (smallest r^2, ties to the smaller index), 0 for an empty list. */
void irr_simd_firr_vec_(double *ti, int *ni, int addr[], double acc[][3],
                        double jrk[][3], int nnbid[]);

/* Additive, batched forms of set_jp / set_list (no reference counterpart;
 * same staging).  x, v, f, fd: [n][3]; lists: nnb[q] members of addr[q]
 * concatenated in `members` (1-based).  Return 0 or -1 (error set). */
int gpunb_irr_set_jp(int n, const int *addr, const double *x, const double *v,
                     const double *f, const double *fd, const double *m,
                     const double *t0);
int gpunb_irr_set_lists(int n, const int *addr, const int *nnb,
                        const int *members);
/* Diagnostic, no reference counterpart (the reference's irr_simd_profile_
 * reports host wall time only): duration in ms of the force kernel of the
 * last irr_simd_firr_vec_ with ni > 0, measured with HIP events on the
 * library's stream; 0 before the first such call after irr_simd_open_. */
double gpunb_irr_kernel_ms(void);

/* ------------------------------------------------------------------------ */
/* Learned neighbour radius (no reference counterpart: the reference sets RS */
/* only by the Ahmad-Cohen volume rule, regcor_gpu.F:146-253, which stays    */
/* the caller's fallback).  An MLP 8 -> 64 -> 64 -> 1 with SiLU activations, */
/* BF16 weights and activations, FP32 accumulation, predicts log RS from 8   */
/* per-particle features.  Parameter vector (FP32, gpunb_mlp_nparams()       */
/* floats): W1[64][8] b1[64] W2[64][64] b2[64] W3[64] b3.  Inputs are        */
/* normalised as (x - xmean) / xstd before the first layer; the output is    */
/* yhat * ystd + ymean.  Host arrays; synchronous; 0 or -1 (error set).      */
/* ------------------------------------------------------------------------ */

int gpunb_mlp_nparams(void);

/* Glorot-uniform weights from seed, zero biases, Adam state reset. */
int gpunb_mlp_init(unsigned seed);

int gpunb_mlp_set_norm(const float *xmean, const float *xstd, float ymean,
                       float ystd);
int gpunb_mlp_get_params(float *p);
/* Sets the FP32 master parameters (and their BF16 copy); resets Adam. */
int gpunb_mlp_set_params(const float *p);

/* y[s] = network(x[s][0..7]) for n samples. */
int gpunb_mlp_predict(int n, const float *x, float *y);

/* Gradient of L = mean_s (yhat_s - (t_s - ymean)/ystd)^2 with respect to the
 * parameters (BF16 rounding differentiated as identity), and L itself. */
int gpunb_mlp_grad(int n, const float *x, const float *t, float *grad,
                   double *loss);

/* Adam (lr, beta 0.9/0.999, eps 1e-8) on minibatches of `batch` samples,
 * reshuffled (seed) every epoch; *loss_out = mean loss of the last epoch. */
int gpunb_mlp_train(int n, const float *x, const float *t, int epochs,
                    int batch, float lr, unsigned seed, double *loss_out);

/* Wait for all work on the library stream; returns 0 or -1 (error set). */
int gpunb_sync(void);

/* The library's HIP stream (as void*), for callers that time or order work
 * against it. */
void *gpunb_stream(void);

/* Library build/version string. */
const char *gpunb_version(void);

/* Work decomposition of the regular-force kernel (autotuner hook):
 * group = pairs per branch-free group of the inner loop (2 or 4; 0 = default),
 * rounds = resident-wave capacity multiples per launch (0 = default).
 * Returns 0, or -1 with gpunb_last_error() set. */
int gpunb_set_tuning(int group, double rounds);

/* Pair arithmetic of the regular-force kernel: 32 = FP32 pair arithmetic on
 * tile-relative differences, with the tiles near an i evaluated in FP64;
 * 64 = FP64 pair arithmetic with the hardware 1/sqrt; 65 = FP64 with one
 * Newton step on 1/sqrt; 0 = default: 32 for calls whose i are run in
 * Hilbert order (gpunb_regf_ / gpunb_regf_dev with ni >= 8192), 65 for the
 * others (scattered i leave few tiles far from every i of a wavefront).
 * Neighbour lists are identical in every mode.  Returns 0 or -1. */
int gpunb_set_precision(int prec);

/* Kernel timing: with enable != 0 the library records HIP events on its own
 * stream around every launch of the regular-force pair kernel; a read
 * synchronises and returns the summed kernel time (ms), the number of
 * launches and the pair interactions they computed, then resets. */
int gpunb_timing(int enable);
int gpunb_timing_read(double *ms_total, long long *launches,
                      long long *pairs);

/* ------------------------------------------------------------------------ */
/* Regular corrector on the device.  Replaces the host loop                  */
/*   DO II = 1, NI: CALL REGCOR_GPU(I, x_i, v_i, GPU_A, GPU_JERK, LISTGP)     */
/* of util_gpu.F:152-180 (REGCOR_GPU: regcor_gpu.F:1-805) for an isolated    */
/* cluster: KZ(13) = KZ(14) = KZ(37) = KZ(38) = 0 (no cloud or external      */
/* force, no CHECKL, no FPCORR; calls with these options are refused).       */
/* ------------------------------------------------------------------------ */

/* Scalars of the reference's common6.h that REGCOR_GPU reads. */
typedef struct {
    double time;       /* TIME */
    double rscale;     /* RSCALE */
    double rdens[3];   /* RDENS: density centre */
    double rc2;        /* RC2: core radius squared */
    double rsfac;      /* RSFAC (adjust.F:220) */
    double znbmin;     /* ZNBMIN (input.F:122) */
    double smax;       /* SMAX */
    double dtmin;      /* DTMIN */
    double etar;       /* ETAR */
    double smin;       /* SMIN */
    double bodym;      /* BODYM: mean mass */
    double tidal4;     /* TIDAL(4); 0 for an isolated cluster */
    int nnbopt;        /* NNBOPT */
    int nnbmax;        /* NNBMAX */
    int n;             /* N: single bodies are 1..N, c.m. bodies N+1..NTOT */
    int ntot;          /* NTOT */
    int ifirst;        /* IFIRST: j-particle 0 of the library is body IFIRST */
    int lmax;          /* LMAX: row length of LIST (>= nnbmax + 3) */
    int nb_flag;       /* NB_FLAG (KZ(39) = 0 or 2): radii about RDENS */
    int m_flag;        /* M_FLAG (KZ(39) = 2 or 3): mass-weighted lists */
    /* Options of branches the device corrector does not implement; a call
     * with any of them nonzero is refused (error, nothing changed):
     * KZ(13) cloud forces (FCLOUD, regcor_gpu.F:93-96), KZ(14) external
     * force (XTRNLF, :98-118), KZ(37) neighbour additions (CHECKL,
     * :255-260), KZ(38) force-polynomial corrections (FPCORR, :584-595). */
    int kz13, kz14, kz37, kz38;
} gpunb_regcor_par;

/* Device arrays of the body state, indexed by body I - 1 (vectors as three
 * consecutive doubles, the Fortran (3,NMAX) layout; LIST rows of lmax ints:
 * count, then the 1-based members in ascending order). */
typedef struct {
    const double *x0, *x0dot;            /* X0, X0DOT (read) */
    double *xn, *xndot;                  /* XN, XNDOT (written) */
    double *fi, *fidot, *fr, *frdot;     /* FI, FIDOT, FR, FRDOT */
    double *d0, *d0r, *d1r, *d2r, *d3r;  /* D0, D0R, D1R, D2R, D3R (written) */
    double *t0r, *stepr, *step;          /* T0R, STEPR, STEP */
    const double *t0;                    /* T0 (read) */
    double *timenw, *rs;                 /* TIMENW, RS */
    const double *body;                  /* BODY (read) */
    const double *h;                     /* H(NPAIRS) for I > N; NULL if none */
    int *list;                           /* LIST(LMAX,NMAX) */
} gpunb_regcor_state;

/* Correct the ni bodies d_idx[] (1-based I, distinct) at par->time, after
 * gpunb_regf_dev at the same time, on the library stream (not synchronised):
 *   d_xi, d_vi       : double[ni][3] predicted x_i, v_i (the regf inputs)
 *   d_freg, d_fdr    : double[ni][3] regular force / derivative from regf
 *                      (updated when a small-step neighbour is retained)
 *   d_nnb, d_off, d_list : the compact lists of gpunb_regf_dev (0-based j,
 *                      body itself included; overflows must be resolved)
 *   list_len         : entries of d_list that may be read (a body whose
 *                      list reaches past them is reported as invalid)
 *   d_counters       : int[4] accumulating NBVOID, NBSMIN, NBFLUX, NICONV
 * Neighbour positions, velocities and masses are the library's current
 * (predicted) j records: body J is j-particle J - IFIRST.  STEP(J) of
 * another body is read as it was before the call.  Returns 0, or -1 with
 * gpunb_last_error() set; invalid bodies (index, duplicate, list range or
 * order) are left unchanged and reported by gpunb_sync(). */
int gpunb_regcor_dev(const gpunb_regcor_par *par,
                     const gpunb_regcor_state *st, int ni, const int *d_idx,
                     const double *d_xi, const double *d_vi, double *d_freg,
                     double *d_fdr, const int *d_nnb, const int *d_off,
                     const int *d_list, long long list_len, int *d_counters);

/* ------------------------------------------------------------------------ */
/* Potential energy of the energy check on the device: the GPU branch of     */
/* ENERGY (energy.F:37-79: GPUPOT over IFIRST..NTOT, the c.m. potential      */
/* copied to the KS components, PHICOR's differential corrections, phicor.f: */
/* 1-112) and the kinetic energy (energy.F:141-146).                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;               /* N: singles and KS components are 1..N */
    int ntot;            /* NTOT = N + NPAIRS (c.m. bodies N+1..NTOT) */
    int ifirst;          /* IFIRST = 2*NPAIRS + 1 */
    int lmax;            /* LMAX: row length of LIST */
    double cmsep2_rmin2; /* CMSEP2*RMIN2: radius^2 of the c.m. resolution */
} gpunb_energy_par;

/* Device arrays indexed by body I - 1: d_body[NTOT], d_x / d_xdot
 * [NTOT][3] (components at their resolved positions), d_list LIST(LMAX,NTOT)
 * (singles: neighbour lists with c.m. members last; KS pair IPAIR: its
 * perturber list in row 2*IPAIR-1), d_phidbl[NTOT] PHIDBL.  run_pot != 0:
 * PHIDBL(IFIRST..NTOT) is computed first by the gpupot_ kernels; 0: the
 * caller's values are used.  On return (after gpunb_sync) PHIDBL holds the
 * corrected potentials of singles and components, d_out[0] = POT and
 * d_out[1] = ZKIN (both halved as energy.F does).  Asynchronous on the
 * library stream; invalid list rows are reported by gpunb_sync().
 * Returns 0, or -1 with gpunb_last_error() set. */
int gpunb_energy_dev(const gpunb_energy_par *par, const double *d_body,
                     const double *d_x, const double *d_xdot,
                     const int *d_list, int run_pot, double *d_phidbl,
                     double *d_out);

#ifdef __cplusplus
}
#endif

#endif /* GPUNB_MI355X_H */
