// gpunb.hip -- host side of the MI355X regular-force engine: the C-ABI of
// include/gpunb.h, device state, staging buffers and the work decomposition.
//
// One library stream (default priority) carries every copy and kernel.  The
// ABI entry points stage host arrays through pinned buffers, enqueue the
// kernels and synchronise once per call; the *_dev entry points work on
// caller-owned device memory and do not synchronise.
#include "gpunb_internal.h"
#include "gpunb.h"

#include <hip/hip_runtime.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <mutex>
#include <string>
#include <vector>

namespace gpunb {
namespace {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                         \
    do {                                                                     \
        hipError_t rc_ = (expr);                                             \
        if (rc_ != hipSuccess)                                               \
            throw Error(std::string(#expr) + " failed: " +                   \
                        hipGetErrorString(rc_));                             \
    } while (0)

#define REQUIRE(cond, msg)                     \
    do {                                       \
        if (!(cond)) throw Error(msg);         \
    } while (0)

double wtime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1.e-6 * tv.tv_usec;
}

bool env_flag(const char *name, bool dflt) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    return !(v[0] == '0' || v[0] == 'n' || v[0] == 'N' || v[0] == 'f' ||
             v[0] == 'F');
}

// Device buffer that only grows.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t need) {
        if (need <= n) return;
        release();
        size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
        HIPCHK(hipMalloc((void **)&p, bytes));
        n = need;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Pinned host buffer that only grows.
template <class T>
struct HostBuf {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t need) {
        if (need <= n) return;
        release();
        size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
        HIPCHK(hipHostMalloc((void **)&p, bytes, hipHostMallocDefault));
        n = need;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// The Hilbert-sorted view of a j set (see jprep.hip) and the scratch of its
// preparation.
struct SortedView {
    DevBuf<JRec> jps;
    DevBuf<JRec32> j32;
    DevBuf<JTile> tiles;
    DevBuf<int> jidx;
    DevBuf<JClus> jcl;
    DevBuf<PrepClus> pc;
    DevBuf<unsigned long long> kin, kout;
    DevBuf<int> vin, vout;
    DevBuf<char> tmp;
    size_t tmp_bytes = 0;
    DevBuf<double> part;

    // cap_j records (a whole number of tiles including one spare tile past
    // the last one the preparation writes) for ncl clusters.
    void alloc(hipStream_t st, size_t cap_j, int ncl) {
        jps.ensure(cap_j);
        j32.ensure(cap_j);
        jidx.ensure(cap_j);
        // View tiles (kVTile j each), plus those of one spare tile.
        const size_t nvt = cap_j / kVTile + kTile / kVTile;
        tiles.ensure(nvt);
        jcl.ensure((size_t)std::max(ncl, 1));
        pc.ensure((size_t)std::max(ncl, 1));
        launch_park_view(st, j32.p, (int)cap_j, tiles.p, (int)nvt);
    }
    void release() {
        jps.release();
        j32.release();
        tiles.release();
        jidx.release();
        jcl.release();
        pc.release();
        kin.release();
        kout.release();
        vin.release();
        vout.release();
        tmp.release();
        tmp_bytes = 0;
        part.release();
    }
    // Enqueue the preparation of ncl clusters whose table (pc, device) is in
    // place: they tile [0, ntile_all * kTile) and hold nblk statistics blocks.
    void prepare(hipStream_t st, const JRec *jp, int ncl, int nblk,
                 int ntile_all) {
        if (ntile_all <= 0) return;
        const size_t n = (size_t)ntile_all * kTile;
        kin.ensure(n);
        kout.ensure(n);
        vin.ensure(n);
        vout.ensure(n);
        part.ensure(8 * (size_t)nblk);
        const size_t need = jprep_tmp_bytes((int)n, ncl);
        if (need == 0) throw Error("j preparation: sort storage query failed");
        if (need > tmp_bytes) {
            tmp.ensure(need);
            tmp_bytes = need;
        }
        JView V;
        V.jp = jp;
        V.jps = jps.p;
        V.j32 = j32.p;
        V.tiles = tiles.p;
        V.jidx = jidx.p;
        V.jcl = jcl.p;
        PrepBufs B;
        B.key_in = kin.p;
        B.key_out = kout.p;
        B.val_in = vin.p;
        B.val_out = vout.p;
        B.tmp = tmp.p;
        B.tmp_bytes = tmp_bytes;
        B.part = part.p;
        HIPCHK(launch_jprep(st, V, pc.p, ncl, nblk, ntile_all, B));
    }
    // One cluster covering [0, round_up(nj, kTile)).
    void prepare_one(hipStream_t st, const JRec *jp, int nj) {
        const int ntile = (nj + kTile - 1) / kTile;
        if (ntile == 0) return;
        PrepClus v;
        v.j0 = 0;
        v.nj = nj;
        v.ntile = ntile;
        v.blk0 = 0;
        launch_set_prepclus(st, pc.p, v);
        prepare(st, jp, 1, prep_stat_blocks(nj), ntile);
    }
};

struct State {
    bool devinit = false;
    bool open = false;
    int dev = 0;
    int ndev = 0;
    int rank = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev_send = nullptr, ev_jp = nullptr, ev_desc = nullptr;
    bool ev_send_live = false, ev_jp_live = false, ev_desc_live = false;
    HostBuf<ClusterDesc> h_desc;
    DevBuf<ClusterDesc> d_desc;
    // What d_desc holds (uploaded by enqueue_regf); an unchanged table is not
    // uploaded again.
    std::vector<ClusterDesc> desc_up;

    int nbmax = 0;
    int nj = 0;       // current j count used by regf
    int njp = 0;      // j slots covered by the predictor state
    // The sorted view of the j set is rebuilt lazily: send and predict mark
    // it stale, and only calls that run the mixed kernel or sort their i
    // rebuild it (the FP64 kernel reads the caller-order records).
    bool sv_stale = false;
    DevBuf<JRec> jp;
    SortedView sv;
    DevBuf<JStats> jst;
    DevBuf<unsigned> err;
    HostBuf<unsigned> h_err;

    HostBuf<double> h_send;
    DevBuf<double> d_send;

    DevBuf<double> x0, v0, f, fd, t0, mass;
    HostBuf<double> h_jp;
    DevBuf<double> d_jp;
    HostBuf<int> h_addr;
    DevBuf<int> d_addr;

    HostBuf<double> h_in;
    DevBuf<double> d_in;
    // Regular-force scratch: slice partials, per-i list counts and entries.
    // A call of several chunks runs pass 2 of chunk c on st2 while pass 1 of
    // chunk c + 1 runs on st into the other scratch set (index 1 for odd
    // chunks).  The counts are zero between launches (zeroed when allocated,
    // reset by pass 2); nbc_dirty marks a set whose pass 1 was enqueued
    // without its pass 2.
    DevBuf<double> part[2];
    DevBuf<int> nbc[2];
    DevBuf<int> nbl[2];
    bool nbc_dirty[2] = {false, false};
    // Hilbert order of the i of a call (large calls): keys, permutation, sort
    // storage and the i data gathered into that order.
    DevBuf<unsigned> ikey_in, ikey_out;
    DevBuf<int> ival_in, iperm;
    DevBuf<char> itmp;
    size_t itmp_bytes = 0;
    DevBuf<double> isorted;   // xs [ni][3], vs [ni][3], hs [ni], ds [ni]
    hipStream_t st2 = nullptr;
    hipEvent_t ev_p1[2] = {nullptr, nullptr}, ev_p2[2] = {nullptr, nullptr};
    hipEvent_t ev_p2end = nullptr;
    bool p2_live[2] = {false, false};
    // gpunb_regf_ results in one buffer, so that one copy brings them back:
    // acc, jrk, pot (7 ni doubles), then nnb, off, nnid, total (3 ni + 1
    // ints), then the compact list from a 256-byte boundary.
    DevBuf<char> d_res;
    HostBuf<char> h_res;
    long long list_est = 0;

    DevBuf<PRec> pot_rec;
    DevBuf<double> d_pin, d_ppart, d_pout;
    HostBuf<double> h_pin, h_pout;

    // Regular corrector scratch (gpunb_regcor_dev); rc_mark stays zero
    // between calls.
    DevBuf<int> rc_nlist, rc_jj, rc_mark;
    DevBuf<double> rc_step, rc_tnw;

    // Energy-sum scratch (gpunb_energy_dev).
    DevBuf<double> en_cpot, en_ckin, en_part;

    double time_send = 0, time_grav = 0, time_nb = 0, time_out = 0;
    long long numInter = 0;
    int icall = 0, ini = 0, isend = 0;

    // Kernel timing (gpunb_timing): events around every pass-1 launch.
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    std::vector<std::pair<int, int>> ev_pairs;   // indices into ev_pool
    std::vector<long long> ev_pairs_work;        // pair interactions per launch

    int resident[2][2][3] = {};
    int pot_resident = 0;
    int group = 0;          // 0: choose automatically
    double rounds = 0.0;    // 0: default
    int prec = 0;           // 0: default pair arithmetic
};

State S;
bool g_abort = true;

// Scratch of gpunb_regf_dev_noalloc: a file-scope device array (each GPU
// holds its own copy until the process ends).  Host code never passes the
// array's name to a kernel: scratch_dev[d] is its device address on GPU d,
// taken once per device by hipGetSymbolAddress while that device is current.
constexpr size_t kNoAllocBytes = 48ull << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];
unsigned char *scratch_dev[kMaxDevices] = {};

bool g_abort_set = false;
std::string g_last_error;
std::mutex g_error_mu;   // entry points may fail on several threads at once

bool quiet() { return env_flag("GPUNB_QUIET", false); }

void report(const char *fn, const char *msg) {
    {
        std::lock_guard<std::mutex> g(g_error_mu);
        g_last_error = std::string(fn) + ": " + msg;
    }
    fprintf(stderr, "gpunb: %s: %s\n", fn, msg);
    fflush(stderr);
    bool ab = g_abort_set ? g_abort : env_flag("GPUNB_ABORT", true);
    if (ab) abort();
}

template <class F>
int guarded(const char *fn, F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        report(fn, e.what());
        return -1;
    }
}

int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }
int round_up(int a, int b) { return cdiv(a, b) * b; }

void devinit(int irank) {
    if (S.devinit) return;
    HIPCHK(hipGetDeviceCount(&S.ndev));
    REQUIRE(S.ndev > 0, "no HIP device visible");
    int dev = irank % S.ndev;
    const char *gl = getenv("GPU_LIST");
    if (gl && *gl) {
        // First entry of the space-separated list (gpunb.gpu.cu:562-571);
        // parsed from a copy so the environment is left untouched.
        std::string first(gl);
        size_t a = first.find_first_not_of(' ');
        REQUIRE(a != std::string::npos, "GPU_LIST is empty");
        dev = atoi(first.c_str() + a);
        REQUIRE(dev >= 0 && dev < S.ndev, "GPU_LIST names a missing device");
    }
    S.dev = dev;
    S.rank = irank;
    HIPCHK(hipSetDevice(dev));
    REQUIRE(dev < kMaxDevices, "device index beyond the library's table");
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = (unsigned char *)p;
    }
    HIPCHK(hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&S.st2, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipEventCreateWithFlags(&S.ev_p1[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&S.ev_p2[k], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&S.ev_p2end, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&S.ev_send, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&S.ev_jp, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&S.ev_desc, hipEventDisableTiming));
    S.jst.ensure(1);
    S.err.ensure(1);
    S.h_err.ensure(1);
    HIPCHK(hipMemsetAsync(S.err.p, 0, sizeof(unsigned), S.st));
    launch_stats_reset(S.st, S.jst.p, nullptr);
    HIPCHK(hipStreamSynchronize(S.st));
    if (!quiet()) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        char host[150];
        memset(host, 0, sizeof(host));
        gethostname(host, sizeof(host) - 1);
        fprintf(stderr,
                "# gpunb (MI355X) init - rank: %d; host %s; devices %d; "
                "device %d %s (%s, %d CUs)\n",
                irank, host, S.ndev, dev, prop.name, prop.gcnArchName,
                prop.multiProcessorCount);
    }
    S.devinit = true;
}

void set_device() { HIPCHK(hipSetDevice(S.dev)); }

// The device error word accumulates until the next synchronisation point
// (an error raised by an asynchronous send is reported by the following
// regf); err_check clears it when it reports.
size_t round_up16(size_t b) { return (b + 15) & ~(size_t)15; }

// Device address of a pinned host buffer (hipHostMalloc memory is mapped).
void *host_dev_ptr(void *h) {
    void *d = nullptr;
    HIPCHK(hipHostGetDevicePointer(&d, h, 0));
    return d;
}

void err_fetch() {
    HIPCHK(hipMemcpyAsync(S.h_err.p, S.err.p, sizeof(unsigned),
                          hipMemcpyDeviceToHost, S.st));
}
void err_check() {
    unsigned e = *S.h_err.p;
    if (!e) return;
    *S.h_err.p = 0;
    HIPCHK(hipMemsetAsync(S.err.p, 0, sizeof(unsigned), S.st));
    HIPCHK(hipStreamSynchronize(S.st));
    std::string m;
    if (e & ERR_NAN_J) m += "NaN in j-particle data; ";
    if (e & ERR_NAN_I) m += "NaN in i-particle data; ";
    if (e & ERR_LIST_CAP) m += "neighbour list buffer too small; ";
    if (e & ERR_ENERGY)
        m += "energy sum: a LIST row with a count out of range or a member "
             "outside 1..NTOT; ";
    if (e & ERR_REGCOR)
        m += "regular corrector: body index out of range or listed twice, or "
             "a list too long, unsorted or outside the j set (those bodies "
             "are left unchanged); ";
    throw Error(m);
}

int pick_group() {
    if (S.group) return S.group;
    const char *g = getenv("GPUNB_REGF_GROUP");
    if (g && atoi(g) == 2) return 2;
    if (g && atoi(g) == 4) return 4;
    // Two pairs per group: the mixed kernel fits 4 resident waves per SIMD
    // and the FP64 kernel 5 (four pairs: more than 128 VGPRs, 3 waves),
    // which hides the latency of the dependent chains better.
    return 2;
}

double pick_rounds(int ni) {
    if (S.rounds > 0) return S.rounds;
    const char *r = getenv("GPUNB_WAVE_ROUNDS");
    if (r && atof(r) > 0) return atof(r);
    // Resident-wave capacity multiples per launch, by block size: large
    // blocks take up to 12 (King N = 131072 in one launch, whole step:
    // 16.96 ms at 6, 16.32 at 12, 16.74 at 24), small ones fewer, since
    // every extra slice is one more FP64 partial per i to write and reduce
    // (whole step, 1024 i: 0.187 ms at 1, 0.206 at 4; 128 i: 0.108 ms at
    // 0.5, 0.121 at 4).  1024 scattered i through gpunb_regf_: 5.05e11
    // pairs/s at 0.75, 5.3e11 at 1, 5.2e11 at 1.5, 5.4e11 at 2.
    return std::min(12.0, std::max(0.5, ni / 512.0));
}

// Pair arithmetic of a call.  Default: the mixed kernel (32) when the i run
// in Hilbert order, since its FP32 far tiles need waves of nearby i; the
// FP64 kernel with one Newton step (65) otherwise: with i scattered over the
// cluster nearly every tile is near some lane, and the mixed kernel would
// run its FP64 loop almost everywhere with fewer resident waves (King
// N = 131072, 1024 scattered i: pass 1 220 us mixed).
int pick_prec(bool i_sorted) {
    if (S.prec) return S.prec;
    const char *p = getenv("GPUNB_PREC");
    if (p && (atoi(p) == 32 || atoi(p) == 64 || atoi(p) == 65)) return atoi(p);
    return i_sorted ? 32 : 65;
}

int resident_waves(int m_flag, int group, int prec) {
    int &r = S.resident[m_flag ? 1 : 0][group == 2 ? 0 : 1][prec >= 64 ? 1 : 0];
    if (!r) r = regf_resident_waves(m_flag, group, prec);
    if (r <= 0) r = 256 * 8;
    return r;
}

// j slices per i-wavefront: enough wavefronts (kP1W per slice task) to
// fill the resident capacity `rounds` times; a slice holds at least one
// tile per wavefront where the j set allows.
int decompose(int ni_c, int ntile, int m_flag, int group, int prec) {
    const int n_iw = cdiv(ni_c, 64);
    const double target = resident_waves(m_flag, group, prec) * pick_rounds(ni_c);
    int ns = (int)std::lround(target / ((double)n_iw * kP1W));
    ns = std::max(ns, 1);
    ns = std::min(ns, std::max(ntile / kP1W, 1));
    return ns;
}

// A count buffer of at least n ints, zero where newly allocated.
void ensure_counts(DevBuf<int> &b, size_t n) {
    if (n <= b.n) return;
    b.ensure(n);
    launch_zero_int(S.st, b.p, (int)b.n);
}

// Calls of at least this many i are run in Hilbert order of the i
// (GPUNB_ISORT=0 switches it off, GPUNB_ISORT=1 forces it on).
bool pick_isort(int ni) {
    const char *e = getenv("GPUNB_ISORT");
    if (e && *e) return atoi(e) != 0;
    return ni >= 8192;
}

// i-particles per pass-1/pass-2 launch pair (GPUNB_CHUNK overrides).  Long
// wavefronts pay their set-up and the end-of-task reduction once for more
// tiles: King N = 131072 at 12 rounds, whole step 17.92 ms in chunks of
// 32768, 16.99 of 65536, 16.32 in one launch of 131072.  The chunk bounds
// the list scratch: 4 max(nnbmax, 1) bytes per i, two sets.
int pick_chunk() {
    const char *c = getenv("GPUNB_CHUNK");
    if (c && atoi(c) >= 64) return atoi(c);
    return 131072;
}

// Next free timing event (created on first use, reused after a read).
int timing_event() {
    if (S.ev_used == (int)S.ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        S.ev_pool.push_back(e);
    }
    return S.ev_used++;
}

void ensure_regf_scratch(int ni_c, int nslice, int cap, int set) {
    const size_t nip = (size_t)round_up(ni_c, 64);
    S.part[set].ensure(nip * nslice * 8);
    ensure_counts(S.nbc[set], nip);
    S.nbl[set].ensure(nip * cap);
    if (S.nbc_dirty[set]) {
        launch_zero_int(S.st, S.nbc[set].p, (int)S.nbc[set].n);
        S.nbc_dirty[set] = false;
    }
}

// Waits for a recorded event (defined below).
void wait_event(hipEvent_t ev, bool &live);

// The sorted view of the current j set, rebuilt if a send or predict made it
// stale.
void ensure_view() {
    if (!S.sv_stale) return;
    S.sv.prepare_one(S.st, S.jp.p, S.nj);
    S.sv_stale = false;
}

// j side of a launch: the sorted view for the mixed kernel, the
// caller-order records (identity position map) for the FP64 kernel.
void set_j_side(RegfLaunch &L, int prec) {
    if (prec < 64) {
        L.jps = S.sv.jps.p;
        L.j32 = S.sv.j32.p;
        L.tiles = S.sv.tiles.p;
        L.jidx = S.sv.jidx.p;
        L.jcl = S.sv.jcl.p;
    } else {
        L.jps = S.jp.p;
        L.j32 = nullptr;
        L.tiles = nullptr;
        L.jidx = nullptr;
        L.jcl = nullptr;
    }
}

// Enqueue the regular-force kernels for device-resident i data: one launch
// pair per chunk of pick_chunk() i, each a one-cluster table.
void enqueue_regf(int ni, const double *xi, const double *vi, const double *h2,
                  const double *dtr, double *acc, double *jrk, double *pot,
                  int nnbmax, int m_flag, int *nnb, int *nnid, int *list,
                  int *off, int *total, long long list_cap) {
    const int group = pick_group();
    const bool isort = S.nj > 0 && pick_isort(ni);
    const int prec = pick_prec(isort);
    if (isort || prec < 64) ensure_view();
    const int ntile = cdiv(S.nj, kTile);
    const int chunk = pick_chunk();
    const int nch = cdiv(ni, chunk);
    const int *perm = nullptr;
    if (isort) {
        const size_t n = (size_t)ni;
        S.ikey_in.ensure(n);
        S.ikey_out.ensure(n);
        S.ival_in.ensure(n);
        S.iperm.ensure(n);
        S.isorted.ensure(8 * n);
        const size_t need = isort_tmp_bytes(ni);
        if (need == 0) throw Error("i sort: storage query failed");
        if (need > S.itmp_bytes) {
            S.itmp.ensure(need);
            S.itmp_bytes = need;
        }
        ISortBufs B = {};
        B.key_in = S.ikey_in.p;
        B.key_out = S.ikey_out.p;
        B.val_in = S.ival_in.p;
        B.perm = S.iperm.p;
        B.tmp = S.itmp.p;
        B.tmp_bytes = S.itmp_bytes;
        B.xs = S.isorted.p;
        B.vs = B.xs + 3 * n;
        B.hs = B.vs + 3 * n;
        B.ds = B.hs + n;
        HIPCHK(launch_isort(S.st, S.sv.jcl.p, xi, vi, h2, dtr, ni, B));
        xi = B.xs;
        vi = B.vs;
        h2 = B.hs;
        dtr = B.ds;
        perm = B.perm;
    }
    std::vector<ClusterDesc> nd((size_t)nch);
    for (int c = 0; c < nch; ++c) {
        const int c0 = c * chunk;
        const int nc = std::min(chunk, ni - c0);
        const int ns = ntile == 0 ? 1 : decompose(nc, ntile, m_flag, group, prec);
        const int cap = std::max(nnbmax, 1);
        ensure_regf_scratch(nc, ns, cap, c & 1);
        ClusterDesc &D = nd[c];
        memset(&D, 0, sizeof D);
        D.nj = S.nj;
        D.ntile = ntile;
        D.i0 = c0;
        D.ni = nc;
        D.n_iw = cdiv(nc, 64);
        D.nslice = ns;
        D.cap = cap;
        D.ipb = pass2_ipb(nc);
    }
    const size_t had = S.d_desc.n;
    S.d_desc.ensure((size_t)nch);
    if (S.d_desc.n != had) S.desc_up.clear();   // new buffer: nothing held
    if (S.desc_up.size() != nd.size() ||
        memcmp(S.desc_up.data(), nd.data(), nd.size() * sizeof(ClusterDesc))) {
        // h_desc may still be read by the previous upload.
        wait_event(S.ev_desc, S.ev_desc_live);
        S.h_desc.ensure((size_t)nch);
        memcpy(S.h_desc.p, nd.data(), nd.size() * sizeof(ClusterDesc));
        HIPCHK(hipMemcpyAsync(S.d_desc.p, S.h_desc.p, nch * sizeof(ClusterDesc),
                              hipMemcpyHostToDevice, S.st));
        HIPCHK(hipEventRecord(S.ev_desc, S.st));
        S.ev_desc_live = true;
        S.desc_up = nd;
    }
    for (int c = 0; c < nch; ++c) {
        const ClusterDesc &D = nd[c];
        RegfLaunch L;
        memset(&L, 0, sizeof L);
        set_j_side(L, prec);
        L.jst = S.jst.p;
        L.xi = xi;
        L.vi = vi;
        L.h2 = h2;
        L.dtr = dtr;
        const int b = c & 1;
        L.part = S.part[b].p;
        L.nbc = S.nbc[b].p;
        L.nbl = S.nbl[b].p;
        L.perm = perm;
        L.acc = acc;
        L.jrk = jrk;
        L.pot = pot;
        L.nnb = nnb;
        L.nnid = nnid;
        L.list = list;
        L.off = off;
        L.total = total;
        L.list_cap = list_cap;
        L.nnbmax = nnbmax;
        L.m_flag = m_flag;
        L.group = group;
        L.prec = prec;
        L.err = S.err.p;
        L.desc = S.d_desc.p + c;
        L.ncl = 1;
        L.ntask = (long long)D.n_iw * D.nslice;
        L.nblk2 = cdiv(D.ni, D.ipb);
        // Scratch set b was last read by the pass 2 of chunk c - 2.
        if (nch > 1 && S.p2_live[b])
            HIPCHK(hipStreamWaitEvent(S.st, S.ev_p2[b], 0));
        int e0 = -1, e1 = -1;
        if (S.timing) {
            e0 = timing_event();
            HIPCHK(hipEventRecord(S.ev_pool[e0], S.st));
        }
        S.nbc_dirty[b] = true;
        launch_regf_pass1(S.st, L);
        if (S.timing) {
            e1 = timing_event();
            HIPCHK(hipEventRecord(S.ev_pool[e1], S.st));
            S.ev_pairs.emplace_back(e0, e1);
            S.ev_pairs_work.push_back((long long)D.ni * S.nj);
        }
        if (nch > 1) {
            HIPCHK(hipEventRecord(S.ev_p1[b], S.st));
            HIPCHK(hipStreamWaitEvent(S.st2, S.ev_p1[b], 0));
            launch_regf_pass2(S.st2, L);
            HIPCHK(hipEventRecord(S.ev_p2[b], S.st2));
            S.p2_live[b] = true;
        } else {
            launch_regf_pass2(S.st, L);
        }
        HIPCHK(hipGetLastError());
        S.nbc_dirty[b] = false;
    }
    if (nch > 1) {
        // Everything after this call on st (copies, the next predictor,
        // synchronisations) follows the last pass 2.
        HIPCHK(hipEventRecord(S.ev_p2end, S.st2));
        HIPCHK(hipStreamWaitEvent(S.st, S.ev_p2end, 0));
    }
}

// gpunb_regf_dev without any allocation: one launch pair whose scratch (the
// cluster table, list counts and entries, slice partials) is carved out of
// the device array `scratch`.  The decomposition is the default one, with
// fewer slices when the partials would not fit; a call whose lists alone
// (or one slice of partials) do not fit is refused.  Calls share the array:
// two calls that overlap on one GPU race.
void enqueue_regf_noalloc(int ni, const double *xi, const double *vi,
                          const double *h2, const double *dtr, double *acc,
                          double *jrk, double *pot, int nnbmax, int m_flag,
                          int *nnb, int *nnid, int *list, int *off, int *total,
                          long long list_cap) {
    unsigned char *base = scratch_dev[S.dev];
    REQUIRE(base, "gpunb_regf_dev_noalloc: scratch address not initialised");
    const int group = pick_group();
    const int prec = pick_prec(false);
    const int ntile = cdiv(S.nj, kTile);
    const int cap = std::max(nnbmax, 1);
    const size_t nip = (size_t)round_up(ni, 64);
    auto a256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_nbc = a256(sizeof(ClusterDesc));
    const size_t o_nbl = o_nbc + a256(nip * sizeof(int));
    const size_t o_part = o_nbl + a256(nip * (size_t)cap * sizeof(int));
    const int ns_min = 1;
    REQUIRE(o_part + nip * ns_min * 8 * sizeof(double) <= kNoAllocBytes,
            "gpunb_regf_dev_noalloc: ni * nnbmax too large for the 48 MiB "
            "scratch (use gpunb_regf_dev)");
    int ns = ntile == 0 ? 1 : decompose(ni, ntile, m_flag, group, prec);
    const size_t fit = (kNoAllocBytes - o_part) / (nip * 8 * sizeof(double));
    ns = (int)std::max<size_t>(ns_min, std::min<size_t>(ns, fit));
    ClusterDesc D;
    memset(&D, 0, sizeof D);
    D.nj = S.nj;
    D.ntile = ntile;
    D.ni = ni;
    D.n_iw = cdiv(ni, 64);
    D.nslice = ns;
    D.cap = cap;
    D.ipb = pass2_ipb(ni);
    ClusterDesc *d_desc = (ClusterDesc *)base;
    int *nbc = (int *)(base + o_nbc);
    launch_set_desc(S.st, d_desc, D);
    launch_zero_int(S.st, nbc, (int)nip);
    RegfLaunch L;
    memset(&L, 0, sizeof L);
    if (prec < 64) ensure_view();
    set_j_side(L, prec);
    L.jst = S.jst.p;
    L.xi = xi;
    L.vi = vi;
    L.h2 = h2;
    L.dtr = dtr;
    L.part = (double *)(base + o_part);
    L.nbc = nbc;
    L.nbl = (int *)(base + o_nbl);
    L.acc = acc;
    L.jrk = jrk;
    L.pot = pot;
    L.nnb = nnb;
    L.nnid = nnid;
    L.list = list;
    L.off = off;
    L.total = total;
    L.list_cap = list_cap;
    L.nnbmax = nnbmax;
    L.m_flag = m_flag;
    L.group = group;
    L.prec = prec;
    L.err = S.err.p;
    L.desc = d_desc;
    L.ncl = 1;
    L.ntask = (long long)D.n_iw * D.nslice;
    L.nblk2 = cdiv(D.ni, D.ipb);
    launch_regf_pass1(S.st, L);
    launch_regf_pass2(S.st, L);
    HIPCHK(hipGetLastError());
}

void require_open(const char *what) {
    if (!S.open) throw Error(std::string(what) + ": library is not open");
}

void wait_event(hipEvent_t ev, bool &live) {
    if (live) {
        HIPCHK(hipEventSynchronize(ev));
        live = false;
    }
}

// ---------------------------------------------------------------------------
// Entry point bodies
// ---------------------------------------------------------------------------

void do_open(int nbmax, int irank) {
    S.time_send = S.time_grav = S.time_nb = S.time_out = 0.0;
    S.numInter = 0;
    S.icall = S.ini = S.isend = 0;
    devinit(irank);
    if (S.open) {
        fprintf(stderr, "gpunb: it is already open\n");
        return;
    }
    REQUIRE(nbmax >= 0, "nbmax must be >= 0");
    set_device();
    const size_t cap = (size_t)round_up(std::max(nbmax, 1), kTile) + kTile;
    S.jp.ensure(cap);
    S.x0.ensure(3 * (size_t)std::max(nbmax, 1));
    S.v0.ensure(3 * (size_t)std::max(nbmax, 1));
    S.f.ensure(3 * (size_t)std::max(nbmax, 1));
    S.fd.ensure(3 * (size_t)std::max(nbmax, 1));
    S.t0.ensure((size_t)std::max(nbmax, 1));
    S.mass.ensure((size_t)std::max(nbmax, 1));
    HIPCHK(hipMemsetAsync(S.x0.p, 0, S.x0.n * sizeof(double), S.st));
    HIPCHK(hipMemsetAsync(S.v0.p, 0, S.v0.n * sizeof(double), S.st));
    HIPCHK(hipMemsetAsync(S.f.p, 0, S.f.n * sizeof(double), S.st));
    HIPCHK(hipMemsetAsync(S.fd.p, 0, S.fd.n * sizeof(double), S.st));
    HIPCHK(hipMemsetAsync(S.t0.p, 0, S.t0.n * sizeof(double), S.st));
    HIPCHK(hipMemsetAsync(S.mass.p, 0, S.mass.n * sizeof(double), S.st));
    // Park every record (mass 0, far away) so spare records are harmless.
    launch_stats_reset(S.st, S.jst.p, nullptr);
    launch_pack_j(S.st, nullptr, nullptr, nullptr, 0, (int)cap, S.jp.p,
                  S.jst.p, S.err.p);
    S.sv.alloc(S.st, cap, 1);
    HIPCHK(hipStreamSynchronize(S.st));
    S.nbmax = nbmax;
    S.nj = 0;
    S.njp = 0;
    S.open = true;
    if (!quiet())
        fprintf(stderr, "# Open MI355X regular force - rank: %d; nbmax: %d\n",
                irank, nbmax);
}

void do_close() {
    if (!S.open) {
        fprintf(stderr, "gpunb: it is already close\n");
        return;
    }
    set_device();
    HIPCHK(hipStreamSynchronize(S.st));
    S.ev_send_live = S.ev_jp_live = S.ev_desc_live = false;
    S.jp.release();
    S.sv.release();
    S.d_send.release();
    S.h_send.release();
    S.x0.release();
    S.v0.release();
    S.f.release();
    S.fd.release();
    S.t0.release();
    S.mass.release();
    S.h_jp.release();
    S.d_jp.release();
    S.h_addr.release();
    S.d_addr.release();
    S.h_in.release();
    S.d_in.release();
    for (int b = 0; b < 2; ++b) {
        S.part[b].release();
        S.nbc[b].release();
        S.nbl[b].release();
        S.nbc_dirty[b] = false;
    }
    S.ikey_in.release();
    S.ikey_out.release();
    S.ival_in.release();
    S.iperm.release();
    S.itmp.release();
    S.itmp_bytes = 0;
    S.isorted.release();
    S.d_res.release();
    S.h_res.release();
    S.rc_nlist.release();
    S.rc_jj.release();
    S.rc_mark.release();
    S.rc_step.release();
    S.rc_tnw.release();
    S.en_cpot.release();
    S.en_ckin.release();
    S.en_part.release();
    S.nbmax = 0;
    S.nj = 0;
    S.njp = 0;
    S.open = false;
}

void do_send(int nj, const double *m, const double (*x)[3],
             const double (*v)[3]) {
    require_open("gpunb_send_");
    REQUIRE(nj >= 0, "gpunb_send_: nj must be >= 0");
    REQUIRE(nj <= S.nbmax, "gpunb_send_: nj exceeds nbmax of gpunb_open_");
    double t0 = wtime();
    set_device();
    S.isend++;
    wait_event(S.ev_send, S.ev_send_live);
    const size_t n7 = 7 * (size_t)nj;
    S.h_send.ensure(std::max<size_t>(n7, 7));
    S.d_send.ensure(std::max<size_t>(n7, 7));
    if (nj > 0) {
        memcpy(S.h_send.p, m, nj * sizeof(double));
        memcpy(S.h_send.p + nj, x, 3 * nj * sizeof(double));
        memcpy(S.h_send.p + 4 * (size_t)nj, v, 3 * nj * sizeof(double));
        HIPCHK(hipMemcpyAsync(S.d_send.p, S.h_send.p, n7 * sizeof(double),
                              hipMemcpyHostToDevice, S.st));
        HIPCHK(hipEventRecord(S.ev_send, S.st));
        S.ev_send_live = true;
    }
    const int old_pad = round_up(S.nj, kTile);
    const int pad = std::max(round_up(nj, kTile), old_pad);
    launch_stats_reset(S.st, S.jst.p, nullptr);
    launch_pack_j(S.st, S.d_send.p, S.d_send.p + nj, S.d_send.p + 4 * (size_t)nj,
                  nj, pad, S.jp.p, S.jst.p, S.err.p);
    S.sv_stale = true;
    HIPCHK(hipGetLastError());
    S.nj = nj;
    S.time_send += wtime() - t0;
}

void do_regf(int ni, const double *h2, const double *dtr, const double (*xi)[3],
             const double (*vi)[3], double (*acc)[3], double (*jrk)[3],
             double *pot, int lmax, int nnbmax, int *list, int m_flag,
             int *nnid) {
    require_open("gpunb_regf_");
    REQUIRE(ni >= 0, "gpunb_regf_: ni must be >= 0");
    REQUIRE(nnbmax >= 0, "gpunb_regf_: nnbmax must be >= 0");
    REQUIRE(nnbmax <= kMaxNnbmax, "gpunb_regf_: nnbmax above 4000 is not supported (a list row is sorted in LDS; the reference LMAX is at most 600)");
    REQUIRE(lmax > nnbmax, "gpunb_regf_: lmax must exceed nnbmax");
    if (ni == 0) return;
    double t0 = wtime();
    set_device();
    S.numInter += (long long)ni * S.nj;
    S.ini += ni;
    S.icall++;

    const size_t n8 = 8 * (size_t)ni;
    S.h_in.ensure(n8);
    S.d_in.ensure(n8);
    memcpy(S.h_in.p, xi, 3 * (size_t)ni * sizeof(double));
    memcpy(S.h_in.p + 3 * (size_t)ni, vi, 3 * (size_t)ni * sizeof(double));
    memcpy(S.h_in.p + 6 * (size_t)ni, h2, (size_t)ni * sizeof(double));
    memcpy(S.h_in.p + 7 * (size_t)ni, dtr, (size_t)ni * sizeof(double));

    const long long list_cap = std::max<long long>(1, (long long)ni * nnbmax);
    const size_t off_i = 7 * (size_t)ni * sizeof(double);
    const size_t off_l =
        (off_i + (3 * (size_t)ni + 1) * sizeof(int) + 255) & ~(size_t)255;
    S.d_res.ensure(round_up16(off_l + (size_t)list_cap * sizeof(int)));
    double *d_acc = (double *)S.d_res.p, *d_jrk = d_acc + 3 * (size_t)ni,
           *d_pot = d_acc + 6 * (size_t)ni;
    int *d_nnb = (int *)(S.d_res.p + off_i), *d_off = d_nnb + ni,
        *d_nnid = d_off + ni, *d_total = d_nnid + ni;
    int *d_lst = (int *)(S.d_res.p + off_l);
    // i block up (64 ni bytes) and the list cursor cleared, in stream order.
    launch_copy_from_host(S.st, host_dev_ptr(S.h_in.p), S.d_in.p,
                          n8 * sizeof(double), d_total);
    const double *dx = S.d_in.p, *dv = dx + 3 * (size_t)ni,
                 *dh = dx + 6 * (size_t)ni, *dt = dx + 7 * (size_t)ni;
    // The nearest neighbour is searched only for gpunb_regfnn_.
    enqueue_regf(ni, dx, dv, dh, dt, d_acc, d_jrk, d_pot, nnbmax, m_flag,
                 d_nnb, nnid ? d_nnid : nullptr, d_lst, d_off, d_total,
                 list_cap);

    // Results back in one copy: the fixed-size outputs plus a speculative
    // prefix of the compact list (sized from the previous call); the rest,
    // if any, after the synchronisation.
    long long est = std::max<long long>((long long)ni * 64, S.list_est);
    est = std::min(est, list_cap);
    const size_t first = round_up16(off_l + (size_t)est * sizeof(int));
    S.h_res.ensure(first);
    launch_copy_to_host(S.st, S.d_res.p, host_dev_ptr(S.h_res.p), first, S.err.p,
                        (unsigned *)host_dev_ptr(S.h_err.p));
    HIPCHK(hipStreamSynchronize(S.st));
    err_check();
    double t1 = wtime();
    S.time_grav += t1 - t0;

    const long long total =
        ((const int *)(S.h_res.p + off_i))[3 * (size_t)ni];
    S.list_est = total + total / 4 + 64;
    if (total > est) {
        // Grow the pinned buffer, keeping what was already copied.
        const size_t have = off_l + (size_t)est * sizeof(int);
        HostBuf<char> nb;
        nb.ensure(off_l + (size_t)total * sizeof(int));
        memcpy(nb.p, S.h_res.p, have);
        S.h_res.release();
        S.h_res = nb;
        nb.p = nullptr;
        HIPCHK(hipMemcpyAsync(S.h_res.p + have, S.d_res.p + have,
                              (size_t)(total - est) * sizeof(int),
                              hipMemcpyDeviceToHost, S.st));
        HIPCHK(hipStreamSynchronize(S.st));
    }
    double t2 = wtime();
    S.time_nb += t2 - t1;

    const int *h_nnb = (const int *)(S.h_res.p + off_i), *h_off = h_nnb + ni,
              *h_nnid = h_off + ni;
    const int *h_lst = (const int *)(S.h_res.p + off_l);
    const double *h_acc = (const double *)S.h_res.p,
                 *h_jrk = h_acc + 3 * (size_t)ni, *h_pot = h_acc + 6 * (size_t)ni;
    memcpy(acc, h_acc, 3 * (size_t)ni * sizeof(double));
    memcpy(jrk, h_jrk, 3 * (size_t)ni * sizeof(double));
    memcpy(pot, h_pot, (size_t)ni * sizeof(double));
    for (int i = 0; i < ni; ++i) {
        int *row = list + (size_t)i * lmax;
        const int nnb = h_nnb[i];
        row[0] = nnb;
        if (nnb > 0)
            memcpy(row + 1, h_lst + h_off[i], (size_t)nnb * sizeof(int));
    }
    if (nnid) memcpy(nnid, h_nnid, (size_t)ni * sizeof(int));
    S.time_out += wtime() - t2;
}

void do_profile(int irank) {
    if (S.icall && !quiet()) {
        fprintf(stderr,
                "[R.%d-D.%d MI355X Reg.F ] Nsend %d  Ngrav %d  <Ni> %d   "
                "send(s) %f grav(s) %f  nb(s) %f  out(s) %f  Perf.(Gflops) %f\n",
                irank, S.dev, S.isend, S.icall, S.isend ? S.ini / S.isend : S.ini,
                S.time_send, S.time_grav, S.time_nb, S.time_out,
                S.time_grav > 0 ? 60.e-9 * S.numInter / S.time_grav : 0.0);
    }
    S.time_send = S.time_grav = S.time_nb = S.time_out = 0.0;
    S.numInter = 0;
    S.icall = S.ini = S.isend = 0;
}

void do_pot(int irank, int istart, int ni, int n, const double *m,
            const double (*x)[3], double *pot) {
    devinit(irank);
    REQUIRE(n >= 0 && ni >= 0, "gpupot_: n and ni must be >= 0");
    REQUIRE(istart >= 1 || ni == 0, "gpupot_: istart is 1-based");
    REQUIRE((long long)istart - 1 + ni <= n, "gpupot_: istart-1+ni exceeds n");
    if (ni == 0) return;
    double t0 = wtime();
    set_device();
    const int ntile = cdiv(n, kTile);
    const int n_pad = ntile * kTile;
    S.h_pin.ensure(4 * (size_t)n);
    S.d_pin.ensure(4 * (size_t)n);
    S.pot_rec.ensure((size_t)n_pad + kTile);
    memcpy(S.h_pin.p, m, (size_t)n * sizeof(double));
    memcpy(S.h_pin.p + n, x, 3 * (size_t)n * sizeof(double));
    HIPCHK(hipMemcpyAsync(S.d_pin.p, S.h_pin.p, 4 * (size_t)n * sizeof(double),
                          hipMemcpyHostToDevice, S.st));
    launch_pack_p(S.st, S.d_pin.p, S.d_pin.p + n, n, n_pad + kTile, S.pot_rec.p,
                  S.err.p);
    if (!S.pot_resident) S.pot_resident = pot_resident_waves();
    const int target = std::max(S.pot_resident, 1024) * 2;
    const int n_iw = cdiv(ni, 64);
    int ns = std::max(1, (int)std::lround((double)target / n_iw));
    ns = std::min(ns, std::max(ntile, 1));
    S.d_ppart.ensure((size_t)round_up(ni, 64) * ns);
    S.d_pout.ensure((size_t)ni);
    S.h_pout.ensure((size_t)ni);
    launch_pot(S.st, S.pot_rec.p, n, ntile, istart - 1, ni, ns, S.d_ppart.p,
               S.d_pout.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(S.h_pout.p, S.d_pout.p, (size_t)ni * sizeof(double),
                          hipMemcpyDeviceToHost, S.st));
    err_fetch();
    HIPCHK(hipStreamSynchronize(S.st));
    err_check();
    memcpy(pot, S.h_pout.p, (size_t)ni * sizeof(double));
    if (!quiet())
        fprintf(stderr, "[R.%d MI355X Pot.A] Ni %d  NTOT %d  pot(s) %f\n",
                irank, ni, n, wtime() - t0);
}

void do_jpset(int n, const int *addr, const double (*x0)[3],
              const double (*v0)[3], const double (*f)[3],
              const double (*fd)[3], const double *mass, const double *t0) {
    require_open("gpunb_jpset_");
    REQUIRE(n >= 0, "gpunb_jpset_: n must be >= 0");
    if (n == 0) return;
    int amax = 0;
    for (int k = 0; k < n; ++k) {
        REQUIRE(addr[k] >= 1 && addr[k] <= S.nbmax,
                "gpunb_jpset_: address outside 1..nbmax");
        amax = std::max(amax, addr[k]);
    }
    set_device();
    wait_event(S.ev_jp, S.ev_jp_live);
    S.h_jp.ensure(14 * (size_t)n);
    S.d_jp.ensure(14 * (size_t)n);
    S.h_addr.ensure((size_t)n);
    S.d_addr.ensure((size_t)n);
    for (int k = 0; k < n; ++k) {
        double *r = S.h_jp.p + 14 * (size_t)k;
        for (int c = 0; c < 3; ++c) {
            r[c] = x0[k][c];
            r[3 + c] = v0[k][c];
            r[6 + c] = f[k][c];
            r[9 + c] = fd[k][c];
        }
        r[12] = mass[k];
        r[13] = t0[k];
    }
    memcpy(S.h_addr.p, addr, (size_t)n * sizeof(int));
    HIPCHK(hipMemcpyAsync(S.d_jp.p, S.h_jp.p, 14 * (size_t)n * sizeof(double),
                          hipMemcpyHostToDevice, S.st));
    HIPCHK(hipMemcpyAsync(S.d_addr.p, S.h_addr.p, (size_t)n * sizeof(int),
                          hipMemcpyHostToDevice, S.st));
    HIPCHK(hipEventRecord(S.ev_jp, S.st));
    S.ev_jp_live = true;
    launch_jpset(S.st, n, S.d_addr.p, S.d_jp.p, S.x0.p, S.v0.p, S.f.p, S.fd.p,
                 S.mass.p, S.t0.p, S.err.p);
    HIPCHK(hipGetLastError());
    S.njp = std::max(S.njp, amax);
}

void do_predict(double t) {
    require_open("gpunb_predict_");
    set_device();
    const int pad = std::max(round_up(S.njp, kTile), round_up(S.nj, kTile));
    launch_stats_reset(S.st, S.jst.p, nullptr);
    launch_predict(S.st, t, S.njp, pad, S.x0.p, S.v0.p, S.f.p, S.fd.p, S.t0.p,
                   S.mass.p, S.jp.p, S.jst.p);
    S.sv_stale = true;
    HIPCHK(hipGetLastError());
    S.nj = S.njp;
}

void do_jpget(int n, double (*x)[3], double (*v)[3]) {
    require_open("gpunb_jpget_");
    REQUIRE(n >= 0 && n <= S.nj, "gpunb_jpget_: n outside 0..nj");
    if (n == 0) return;
    set_device();
    S.d_pout.ensure(6 * (size_t)n);
    S.h_pout.ensure(6 * (size_t)n);
    launch_jget(S.st, n, S.jp.p, S.d_pout.p, S.d_pout.p + 3 * (size_t)n);
    HIPCHK(hipMemcpyAsync(S.h_pout.p, S.d_pout.p, 6 * (size_t)n * sizeof(double),
                          hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipStreamSynchronize(S.st));
    memcpy(x, S.h_pout.p, 3 * (size_t)n * sizeof(double));
    memcpy(v, S.h_pout.p + 3 * (size_t)n, 3 * (size_t)n * sizeof(double));
}

// ---------------------------------------------------------------------------
// Ensemble of independent clusters: every cluster's j set lives in one JRec
// array (each padded to whole tiles), and one launch pair computes the
// regular force of all clusters, each i against its own cluster only.  The
// launch sequence can be captured into a hipGraph and replayed while the
// sizes stay the same.
// ---------------------------------------------------------------------------

struct Ens {
    bool open = false;
    int ncl = 0;
    std::vector<int> capj, j0, nj;
    int jcap = 0;
    DevBuf<JRec> jp;
    SortedView sv;
    DevBuf<JStats> jst;
    HostBuf<double> h_send;
    DevBuf<double> d_send;
    HostBuf<ClusterDesc> h_desc;
    DevBuf<ClusterDesc> d_desc;
    DevBuf<double> part;
    DevBuf<int> nbc;   // zero between launches, as the engine's
    DevBuf<int> nbl;
    bool nbc_dirty = false;
    std::vector<int> bucket_tiles = std::vector<int>(24, 0);
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<long long> gkey;
    long long last_pairs = 0;
    // The i of every cluster in Hilbert order (for the mixed kernel): keys,
    // permutation, gathered i data, and the clusters' i starts.
    DevBuf<unsigned long long> ikey_in, ikey_out;
    DevBuf<int> ival_in, iperm, d_i0;
    DevBuf<char> itmp;
    size_t itmp_bytes = 0;
    DevBuf<double> isorted;
    HostBuf<int> h_i0;
    std::vector<int> i0_up;   // the starts last uploaded to d_i0
    // Persistent pass 1: task queues per XCD and their task ranges.
    bool persistent = true;
    DevBuf<unsigned> queue;
    DevBuf<long long> d_xr;
    HostBuf<long long> h_xr;
};
Ens E;

void ens_graph_drop() {
    if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    E.gexec = nullptr;
    E.graph = nullptr;
    E.gkey.clear();
}

void ens_close() {
    if (!E.open) return;
    set_device();
    HIPCHK(hipStreamSynchronize(S.st));
    ens_graph_drop();
    E.jp.release();
    E.sv.release();
    E.jst.release();
    E.h_send.release();
    E.d_send.release();
    E.h_desc.release();
    E.d_desc.release();
    E.part.release();
    E.nbc.release();
    E.nbl.release();
    E.ikey_in.release();
    E.ikey_out.release();
    E.ival_in.release();
    E.iperm.release();
    E.d_i0.release();
    E.itmp.release();
    E.itmp_bytes = 0;
    E.isorted.release();
    E.h_i0.release();
    E.i0_up.clear();
    E.queue.release();
    E.d_xr.release();
    E.h_xr.release();
    E.nbc_dirty = false;
    E.open = false;
    E.ncl = 0;
}

void ens_open(int ncl, const int *nbmax) {
    devinit(0);
    REQUIRE(ncl > 0, "gpunb_ens_open: ncl must be > 0");
    if (E.open) ens_close();
    set_device();
    E.ncl = ncl;
    E.capj.assign(nbmax, nbmax + ncl);
    E.j0.resize(ncl);
    E.nj.assign(ncl, 0);
    long long off = 0;
    for (int c = 0; c < ncl; ++c) {
        REQUIRE(nbmax[c] >= 0, "gpunb_ens_open: nbmax must be >= 0");
        E.j0[c] = (int)off;
        off += round_up(std::max(nbmax[c], 1), kTile);
    }
    REQUIRE(off + kTile < (1LL << 31), "gpunb_ens_open: ensemble too large");
    E.jcap = (int)off + kTile;
    E.jp.ensure((size_t)E.jcap);
    E.jst.ensure(1);
    launch_stats_reset(S.st, E.jst.p, nullptr);
    launch_pack_j(S.st, nullptr, nullptr, nullptr, 0, E.jcap, E.jp.p, E.jst.p,
                  S.err.p);
    E.sv.alloc(S.st, (size_t)E.jcap, ncl);
    HIPCHK(hipStreamSynchronize(S.st));
    E.open = true;
}

void ens_send(int ncl, const int *nj, const double *m, const double (*x)[3],
              const double (*v)[3]) {
    REQUIRE(E.open, "gpunb_ens_send: ensemble is not open");
    REQUIRE(ncl == E.ncl, "gpunb_ens_send: ncl differs from gpunb_ens_open");
    set_device();
    long long tot = 0;
    for (int c = 0; c < ncl; ++c) {
        REQUIRE(nj[c] >= 0 && nj[c] <= E.capj[c],
                "gpunb_ens_send: nj outside 0..nbmax of the cluster");
        tot += nj[c];
    }
    HIPCHK(hipStreamSynchronize(S.st));   // staging and graphs may be in use
    E.h_send.ensure(std::max<size_t>(7 * (size_t)tot, 7));
    E.d_send.ensure(std::max<size_t>(7 * (size_t)tot, 7));
    memcpy(E.h_send.p, m, tot * sizeof(double));
    memcpy(E.h_send.p + tot, x, 3 * tot * sizeof(double));
    memcpy(E.h_send.p + 4 * tot, v, 3 * tot * sizeof(double));
    HIPCHK(hipMemcpyAsync(E.d_send.p, E.h_send.p, 7 * tot * sizeof(double),
                          hipMemcpyHostToDevice, S.st));
    launch_stats_reset(S.st, E.jst.p, nullptr);
    long long o = 0;
    for (int c = 0; c < ncl; ++c) {
        const int pad = std::max(round_up(nj[c], kTile), round_up(E.nj[c], kTile));
        launch_pack_j(S.st, E.d_send.p + o, E.d_send.p + tot + 3 * o,
                      E.d_send.p + 4 * tot + 3 * o, nj[c], pad,
                      E.jp.p + E.j0[c], E.jst.p, S.err.p);
        E.nj[c] = nj[c];
        o += nj[c];
    }
    // Sorted view of every cluster: each one's whole capacity range (its
    // spare records are padding and sort last), so the ranges tile the
    // array.
    std::vector<PrepClus> pcs((size_t)ncl);
    int blk = 0;
    for (int c = 0; c < ncl; ++c) {
        pcs[c].j0 = E.j0[c];
        pcs[c].nj = nj[c];
        pcs[c].ntile = (c + 1 < ncl ? E.j0[c + 1] : E.jcap - kTile) / kTile -
                       E.j0[c] / kTile;
        pcs[c].blk0 = blk;
        blk += prep_stat_blocks(nj[c]);
    }
    HIPCHK(hipMemcpyAsync(E.sv.pc.p, pcs.data(), ncl * sizeof(PrepClus),
                          hipMemcpyHostToDevice, S.st));
    E.sv.prepare(S.st, E.jp.p, ncl, blk, (E.jcap - kTile) / kTile);
    HIPCHK(hipStreamSynchronize(S.st));   // pcs leaves scope
    HIPCHK(hipGetLastError());
}

// Tiles per j slice shared by every cluster so that the launch holds about
// `rounds` times the resident-wave capacity; a per-size-bucket table
// (gpunb_ens_set_bucket_tiles, the autotuner's output) overrides it.
int ens_slice_tiles(const int *ni, int group, int prec, int m_flag) {
    long long ni_all = 0;
    for (int c = 0; c < E.ncl; ++c) ni_all += ni[c];
    const double target =
        resident_waves(m_flag, group, prec) *
        pick_rounds((int)std::min<long long>(ni_all, 1 << 30));
    int lo = 1, hi = 1024;
    auto waves = [&](int L) {
        double w = 0;
        for (int c = 0; c < E.ncl; ++c)
            w += (double)cdiv(ni[c], 64) * kP1W *
                 std::max(1, cdiv(cdiv(E.nj[c], kTile), L));
        return w;
    };
    while (lo < hi) {               // smallest L with waves(L) <= target
        const int mid = (lo + hi) / 2;
        if (waves(mid) <= target) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

void ens_regf_dev(int ncl, const int *ni, const double *h2, const double *dtr,
                  const double *xi, const double *vi, double *acc, double *jrk,
                  double *pot, int nnbmax, int m_flag, int *nnb, int *nnid,
                  int *list, int *off, int *total, long long list_cap) {
    REQUIRE(E.open, "gpunb_ens_regf_dev: ensemble is not open");
    REQUIRE(ncl == E.ncl, "gpunb_ens_regf_dev: ncl differs from gpunb_ens_open");
    REQUIRE(nnbmax >= 0, "gpunb_ens_regf_dev: nnbmax must be >= 0");
    REQUIRE(nnbmax <= kMaxNnbmax, "gpunb_ens_regf_dev: nnbmax above 4000 is not supported (a list row is sorted in LDS; the reference LMAX is at most 600)");
    set_device();
    const int group = pick_group();
    // Every cluster's i run in Hilbert order (sorted here, before the graph
    // replay below, so a replay reads this call's order): the mixed kernel
    // by default.
    const int prec = pick_prec(true);
    const int *perm = nullptr;
    long long ni_all = 0;
    for (int c = 0; c < ncl; ++c) {
        REQUIRE(ni[c] >= 0, "gpunb_ens_regf_dev: ni must be >= 0");
        ni_all += ni[c];
    }
    REQUIRE(ni_all < (1LL << 31), "gpunb_ens_regf_dev: too many i");
    if (prec < 64 && ni_all > 0) {
        std::vector<int> i0v((size_t)ncl + 1, 0);
        for (int c = 0; c < ncl; ++c) i0v[c + 1] = i0v[c] + ni[c];
        if (i0v != E.i0_up) {
            HIPCHK(hipStreamSynchronize(S.st));   // h_i0 may be in flight
            E.h_i0.ensure((size_t)ncl + 1);
            E.d_i0.ensure((size_t)ncl + 1);
            memcpy(E.h_i0.p, i0v.data(), ((size_t)ncl + 1) * sizeof(int));
            HIPCHK(hipMemcpyAsync(E.d_i0.p, E.h_i0.p, ((size_t)ncl + 1) * sizeof(int),
                                  hipMemcpyHostToDevice, S.st));
            E.i0_up = i0v;
        }
        const size_t n = (size_t)ni_all;
        E.ikey_in.ensure(n);
        E.ikey_out.ensure(n);
        E.ival_in.ensure(n);
        E.iperm.ensure(n);
        E.isorted.ensure(8 * n);
        const size_t need = isort_ens_tmp_bytes((int)n, ncl);
        if (need == 0) throw Error("ensemble i sort: storage query failed");
        if (need > E.itmp_bytes) {
            E.itmp.ensure(need);
            E.itmp_bytes = need;
        }
        ISortBufs B = {};
        B.key64_in = E.ikey_in.p;
        B.key64_out = E.ikey_out.p;
        B.val_in = E.ival_in.p;
        B.perm = E.iperm.p;
        B.tmp = E.itmp.p;
        B.tmp_bytes = E.itmp_bytes;
        B.xs = E.isorted.p;
        B.vs = B.xs + 3 * n;
        B.hs = B.vs + 3 * n;
        B.ds = B.hs + n;
        HIPCHK(launch_isort_ens(S.st, E.sv.jcl.p, E.d_i0.p, ncl, xi, vi, h2, dtr,
                                (int)n, B));
        xi = B.xs;
        vi = B.vs;
        h2 = B.hs;
        dtr = B.ds;
        perm = B.perm;
    }
    const int L = ens_slice_tiles(ni, group, prec, m_flag);
    // The launch key: a graph is replayed only for identical launches.
    std::vector<long long> key = {ncl, nnbmax, m_flag, group, prec, L,
                                  (long long)h2, (long long)dtr,
                                  (long long)xi, (long long)vi,
                                  (long long)acc, (long long)jrk,
                                  (long long)pot, (long long)nnb,
                                  (long long)nnid, (long long)list,
                                  (long long)off, (long long)total, list_cap,
                                  (long long)perm, (long long)E.persistent};
    for (int c = 0; c < ncl; ++c) {
        key.push_back(ni[c]);
        key.push_back(E.nj[c]);
        key.push_back(E.bucket_tiles[std::min(23, std::max(0, (int)std::log2(std::max(E.nj[c], 1))))]);
    }
    if (E.graph_on && E.gexec && key == E.gkey) {
        HIPCHK(hipGraphLaunch(E.gexec, S.st));
        return;
    }
    // Cluster table, one entry per cluster; scratch offsets in cluster order.
    HIPCHK(hipStreamSynchronize(S.st));   // the table may be in use
    E.h_desc.ensure((size_t)ncl);
    E.d_desc.ensure((size_t)ncl);
    std::vector<ClusterDesc> cd((size_t)ncl);
    long long pairs = 0, i0 = 0, work = 0;
    for (int c = 0; c < ncl; ++c) {
        ClusterDesc &D = cd[c];
        memset(&D, 0, sizeof D);
        const int ntile = cdiv(E.nj[c], kTile);
        const int b = std::min(23, std::max(0, (int)std::log2(std::max(E.nj[c], 1))));
        const int Lc = E.bucket_tiles[b] > 0 ? E.bucket_tiles[b] : L;
        D.j0 = E.j0[c];
        D.nj = E.nj[c];
        D.ntile = ntile;
        D.i0 = (int)i0;
        D.ni = ni[c];
        D.n_iw = cdiv(ni[c], 64);
        D.nslice = std::max(1, cdiv(ntile, Lc));
        D.nslice = std::min(D.nslice, std::max(ntile, 1));
        D.cap = std::max(nnbmax, 1);
        D.ipb = pass2_ipb(ni[c]);
        D.jc = c;
        D.ibase = (int)i0;
        D.pbase = pairs;
        D.sbase = i0 * D.cap;
        pairs += (long long)round_up(std::max(ni[c], 1), 64) * D.nslice;
        i0 += ni[c];
        work += (long long)ni[c] * E.nj[c];
    }
    // Table order = task order.  Persistent launch: clusters by decreasing
    // tiles per task (the largest tasks are taken first, small ones even out
    // the end), and the task list cut into 8 ranges of equal tile count, one
    // per XCD queue.  Plain launch: cluster order.
    std::vector<int> order((size_t)ncl);
    for (int c = 0; c < ncl; ++c) order[c] = c;
    auto tiles_per_task = [&](int c) {
        return (double)cd[c].ntile / std::max(cd[c].nslice, 1);
    };
    const bool persist = E.persistent;
    if (persist)
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return tiles_per_task(a) > tiles_per_task(b);
        });
    long long task = 0, blk = 0;
    double cost_all = 0.0;
    for (int k = 0; k < ncl; ++k) {
        ClusterDesc &D = cd[order[k]];
        D.task0 = task;
        D.blk0 = blk;
        task += (long long)D.n_iw * D.nslice;
        blk += cdiv(D.ni, D.ipb);
        cost_all += (double)D.n_iw * D.nslice * tiles_per_task(order[k]);
        E.h_desc.p[k] = D;
    }
    long long xr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    {
        // xr[x] = first task whose preceding tasks hold x/8 of the tiles.
        double cum = 0.0;
        int x = 1;
        for (int k = 0; k < ncl && x < 8; ++k) {
            const ClusterDesc &D = cd[order[k]];
            const long long nt = (long long)D.n_iw * D.nslice;
            const double w = tiles_per_task(order[k]);
            while (x < 8 && nt > 0 && cum + nt * w >= cost_all * x / 8.0) {
                const long long m = std::min<long long>(
                    nt, std::max<long long>(
                            0, (long long)std::ceil((cost_all * x / 8.0 - cum) / w)));
                xr[x++] = D.task0 + m;
            }
            cum += nt * w;
        }
        for (; x < 8; ++x) xr[x] = task;
        xr[8] = task;
    }
    E.last_pairs = work;
    E.part.ensure((size_t)pairs * 8);
    ensure_counts(E.nbc, (size_t)std::max<long long>(i0, 1));
    E.nbl.ensure((size_t)std::max<long long>(i0 * std::max(nnbmax, 1), 1));
    if (E.nbc_dirty) {
        launch_zero_int(S.st, E.nbc.p, (int)E.nbc.n);
        E.nbc_dirty = false;
    }
    HIPCHK(hipMemcpyAsync(E.d_desc.p, E.h_desc.p, ncl * sizeof(ClusterDesc),
                          hipMemcpyHostToDevice, S.st));
    if (persist) {
        E.queue.ensure(8);
        E.d_xr.ensure(9);
        E.h_xr.ensure(9);
        memcpy(E.h_xr.p, xr, sizeof xr);
        HIPCHK(hipMemcpyAsync(E.d_xr.p, E.h_xr.p, sizeof xr,
                              hipMemcpyHostToDevice, S.st));
    }
    HIPCHK(hipStreamSynchronize(S.st));
    RegfLaunch R;
    memset(&R, 0, sizeof R);
    R.jps = E.sv.jps.p;
    R.j32 = E.sv.j32.p;
    R.tiles = E.sv.tiles.p;
    R.jidx = E.sv.jidx.p;
    R.jcl = E.sv.jcl.p;
    R.jst = E.jst.p;
    R.xi = xi;
    R.vi = vi;
    R.h2 = h2;
    R.dtr = dtr;
    R.part = E.part.p;
    R.nbc = E.nbc.p;
    R.nbl = E.nbl.p;
    R.acc = acc;
    R.jrk = jrk;
    R.pot = pot;
    R.nnb = nnb;
    R.nnid = nnid;
    R.list = list;
    R.off = off;
    R.total = total;
    R.list_cap = list_cap;
    R.nnbmax = nnbmax;
    R.m_flag = m_flag;
    R.group = group;
    R.prec = prec;
    R.err = S.err.p;
    R.desc = E.d_desc.p;
    R.ncl = ncl;
    R.ntask = task;
    R.nblk2 = blk;
    R.perm = perm;
    if (persist) {
        // One workgroup per resident slot, a multiple of 8 (one share per
        // XCD), no more than the tasks need.
        const long long slots = resident_waves(m_flag, group, prec) / kP1W;
        long long g = std::min<long long>(slots, 8 * ((task + 7) / 8));
        g = std::max<long long>(8, g / 8 * 8);
        R.queue = E.queue.p;
        R.xrange = E.d_xr.p;
        R.pgrid = (int)g;
    }
    auto enqueue = [&]() {
        E.nbc_dirty = true;
        launch_zero_int(S.st, total, 1);
        launch_regf_pass1(S.st, R);
        launch_regf_pass2(S.st, R);
        E.nbc_dirty = false;
    };
    if (!E.graph_on) {
        enqueue();
        HIPCHK(hipGetLastError());
        return;
    }
    ens_graph_drop();
    HIPCHK(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal));
    enqueue();
    HIPCHK(hipStreamEndCapture(S.st, &E.graph));
    HIPCHK(hipGraphInstantiate(&E.gexec, E.graph, nullptr, nullptr, 0));
    E.gkey = key;
    HIPCHK(hipGraphLaunch(E.gexec, S.st));
}

}  // namespace

int run_guarded(const char *fn, const std::function<void()> &f) {
    return guarded(fn, f);
}

}  // namespace gpunb

using namespace gpunb;

// Exported symbols (the library is built with -fvisibility=hidden).
#define GPUNB_EXPORT __attribute__((visibility("default")))

extern "C" {

GPUNB_EXPORT void gpunb_devinit_(int *irank) {
    guarded("gpunb_devinit_", [&] { devinit(*irank); });
}

GPUNB_EXPORT void gpunb_open_(int *nbmax, int *irank) {
    guarded("gpunb_open_", [&] { do_open(*nbmax, *irank); });
}

GPUNB_EXPORT void gpunb_close_(void) {
    guarded("gpunb_close_", [&] { do_close(); });
}

GPUNB_EXPORT void gpunb_send_(int *nj, double *mj, double (*xj)[3], double (*vj)[3]) {
    guarded("gpunb_send_", [&] { do_send(*nj, mj, xj, vj); });
}

GPUNB_EXPORT void gpunb_regf_(int *ni, double *h2, double *dtr, double (*xi)[3],
                 double (*vi)[3], double (*acc)[3], double (*jrk)[3],
                 double *pot, int *lmax, int *nnbmax, int *list, int *m_flag) {
    guarded("gpunb_regf_", [&] {
        do_regf(*ni, h2, dtr, xi, vi, acc, jrk, pot, *lmax, *nnbmax, list,
                *m_flag, nullptr);
    });
}

GPUNB_EXPORT void gpunb_regfnn_(int *ni, double *h2, double *dtr, double (*xi)[3],
                   double (*vi)[3], double (*acc)[3], double (*jrk)[3],
                   double *pot, int *lmax, int *nnbmax, int *list,
                   int *m_flag, int *nnid) {
    guarded("gpunb_regfnn_", [&] {
        do_regf(*ni, h2, dtr, xi, vi, acc, jrk, pot, *lmax, *nnbmax, list,
                *m_flag, nnid);
    });
}

GPUNB_EXPORT void gpunb_profile_(int *irank) {
    guarded("gpunb_profile_", [&] { do_profile(*irank); });
}

GPUNB_EXPORT void gpupot_(int *irank, int *istart, int *ni, int *n, double *m,
             double (*x)[3], double *pot) {
    guarded("gpupot_", [&] { do_pot(*irank, *istart, *ni, *n, m, x, pot); });
}

GPUNB_EXPORT void gpunb_jpset_(int *n, int *addr, double (*x0)[3], double (*v0)[3],
                  double (*f)[3], double (*fdot)[3], double *mass,
                  double *t0) {
    guarded("gpunb_jpset_",
            [&] { do_jpset(*n, addr, x0, v0, f, fdot, mass, t0); });
}

GPUNB_EXPORT void gpunb_predict_(double *time) {
    guarded("gpunb_predict_", [&] { do_predict(*time); });
}

GPUNB_EXPORT void gpunb_jpget_(int *n, double (*x)[3], double (*v)[3]) {
    guarded("gpunb_jpget_", [&] { do_jpget(*n, x, v); });
}

GPUNB_EXPORT void gpunb_set_abort_(int *on) {
    g_abort = (*on != 0);
    g_abort_set = true;
}

GPUNB_EXPORT int gpunb_last_error(char *buf, int len) {
    std::lock_guard<std::mutex> g(g_error_mu);
    const int n = (int)g_last_error.size();
    if (buf && len > 0) {
        const int k = std::min(n, len - 1);
        memcpy(buf, g_last_error.data(), (size_t)k);
        buf[k] = '\0';
    }
    g_last_error.clear();
    return n;
}

GPUNB_EXPORT int gpunb_regf_dev(int ni, const double *d_h2, const double *d_dtr,
                   const double *d_xi, const double *d_vi, double *d_acc,
                   double *d_jrk, double *d_pot, int nnbmax, int m_flag,
                   int *d_nnb, int *d_nnid, int *d_list, int *d_off,
                   int *d_total, long long list_cap) {
    return guarded("gpunb_regf_dev", [&] {
        require_open("gpunb_regf_dev");
        REQUIRE(ni >= 0 && nnbmax >= 0, "gpunb_regf_dev: bad sizes");
        REQUIRE(nnbmax <= kMaxNnbmax, "gpunb_regf_dev: nnbmax above 4000 is not supported (a list row is sorted in LDS; the reference LMAX is at most 600)");
        set_device();
        launch_zero_int(S.st, d_total, 1);
        if (ni == 0) return;
        S.numInter += (long long)ni * S.nj;
        S.ini += ni;
        S.icall++;
        enqueue_regf(ni, d_xi, d_vi, d_h2, d_dtr, d_acc, d_jrk, d_pot, nnbmax,
                     m_flag, d_nnb, d_nnid, d_list, d_off, d_total, list_cap);
    });
}

GPUNB_EXPORT int gpunb_regf_dev_noalloc(
    int ni, const double *d_h2, const double *d_dtr, const double *d_xi,
    const double *d_vi, double *d_acc, double *d_jrk, double *d_pot,
    int nnbmax, int m_flag, int *d_nnb, int *d_nnid, int *d_list, int *d_off,
    int *d_total, long long list_cap) {
    return guarded("gpunb_regf_dev_noalloc", [&] {
        require_open("gpunb_regf_dev_noalloc");
        REQUIRE(ni >= 0 && nnbmax >= 0, "gpunb_regf_dev_noalloc: bad sizes");
        REQUIRE(nnbmax <= kMaxNnbmax,
                "gpunb_regf_dev_noalloc: nnbmax above 4000 is not supported");
        set_device();
        launch_zero_int(S.st, d_total, 1);
        if (ni == 0) return;
        S.numInter += (long long)ni * S.nj;
        S.ini += ni;
        S.icall++;
        enqueue_regf_noalloc(ni, d_xi, d_vi, d_h2, d_dtr, d_acc, d_jrk, d_pot,
                             nnbmax, m_flag, d_nnb, d_nnid, d_list, d_off,
                             d_total, list_cap);
    });
}

GPUNB_EXPORT int gpunb_energy_dev(const gpunb_energy_par *par,
                                  const double *d_body, const double *d_x,
                                  const double *d_xdot, const int *d_list,
                                  int run_pot, double *d_phidbl,
                                  double *d_out) {
    return guarded("gpunb_energy_dev", [&] {
        REQUIRE(par, "gpunb_energy_dev: par must not be NULL");
        const gpunb_energy_par &P = *par;
        REQUIRE(P.n >= 0 && P.ntot >= P.n && P.ntot >= 1,
                "gpunb_energy_dev: need 0 <= n <= ntot, ntot >= 1");
        REQUIRE(P.ifirst == 2 * (P.ntot - P.n) + 1 && P.ifirst <= P.n + 1,
                "gpunb_energy_dev: need ifirst = 2*npairs + 1 with npairs = "
                "ntot - n");
        REQUIRE(P.lmax >= 1, "gpunb_energy_dev: lmax must be >= 1");
        REQUIRE(d_body && d_x && d_xdot && d_list && d_phidbl && d_out,
                "gpunb_energy_dev: a device array is NULL");
        devinit(0);
        set_device();
        const int nb = P.ntot - P.ifirst + 1;
        if (run_pot && nb > 0) {
            const int ntile = cdiv(nb, kTile);
            const int n_pad = ntile * kTile;
            S.pot_rec.ensure((size_t)n_pad + kTile);
            launch_pack_p(S.st, d_body + (P.ifirst - 1),
                          d_x + 3 * (size_t)(P.ifirst - 1), nb, n_pad + kTile,
                          S.pot_rec.p, S.err.p);
            if (!S.pot_resident) S.pot_resident = pot_resident_waves();
            const int target = std::max(S.pot_resident, 1024) * 2;
            const int n_iw = cdiv(nb, 64);
            int ns = std::max(1, (int)std::lround((double)target / n_iw));
            ns = std::min(ns, std::max(ntile, 1));
            S.d_ppart.ensure((size_t)round_up(nb, 64) * ns);
            launch_pot(S.st, S.pot_rec.p, nb, ntile, 0, nb, ns, S.d_ppart.p,
                       d_phidbl + (P.ifirst - 1));
        }
        S.en_cpot.ensure((size_t)std::max(nb, 1));
        S.en_ckin.ensure((size_t)std::max(nb, 1));
        S.en_part.ensure(2 * (size_t)std::max((nb + 2047) / 2048, 1));
        EnergyLaunch L;
        L.n = P.n;
        L.ntot = P.ntot;
        L.ifirst = P.ifirst;
        L.lmax = P.lmax;
        L.cmsep2_rmin2 = P.cmsep2_rmin2;
        L.body = d_body;
        L.x = d_x;
        L.xdot = d_xdot;
        L.list = d_list;
        L.phidbl = d_phidbl;
        L.cpot = S.en_cpot.p;
        L.ckin = S.en_ckin.p;
        L.ppart = S.en_part.p;
        L.kpart = S.en_part.p + std::max((nb + 2047) / 2048, 1);
        L.out = d_out;
        L.err = S.err.p;
        launch_energy(S.st, L);
    });
}

static_assert(sizeof(gpunb_energy_par) == 4 * 4 + 8, "gpunb_energy_par layout");

// Layouts shared with ctypes (gpunb.RegcorPar / RegcorState) and the
// Fortran bind(C) types of gpunb_iface.f90.
static_assert(sizeof(gpunb_regcor_par) == 14 * 8 + 12 * 4,
              "gpunb_regcor_par layout");
static_assert(sizeof(gpunb_regcor_state) == 22 * sizeof(void *),
              "gpunb_regcor_state layout");

GPUNB_EXPORT int gpunb_regcor_dev(const gpunb_regcor_par *par,
                                  const gpunb_regcor_state *st, int ni,
                                  const int *d_idx, const double *d_xi,
                                  const double *d_vi, double *d_freg,
                                  double *d_fdr, const int *d_nnb,
                                  const int *d_off, const int *d_list,
                                  long long list_len, int *d_counters) {
    return guarded("gpunb_regcor_dev", [&] {
        require_open("gpunb_regcor_dev");
        REQUIRE(par && st, "gpunb_regcor_dev: par and st must not be NULL");
        REQUIRE(ni >= 0, "gpunb_regcor_dev: ni must be >= 0");
        if (ni == 0) return;
        const gpunb_regcor_par &P = *par;
        REQUIRE(P.ntot >= 1 && P.n >= 0 && P.n <= P.ntot,
                "gpunb_regcor_dev: need 0 <= n <= ntot, ntot >= 1");
        REQUIRE(P.ifirst >= 1 && P.ifirst <= P.ntot,
                "gpunb_regcor_dev: need 1 <= ifirst <= ntot");
        REQUIRE(P.ntot - P.ifirst + 1 <= S.nj,
                "gpunb_regcor_dev: bodies ifirst..ntot must all be j-particles "
                "(send or jpset and predict them first)");
        REQUIRE(P.nnbmax >= 0 && P.lmax >= P.nnbmax + 3,
                "gpunb_regcor_dev: need lmax >= nnbmax + 3");
        REQUIRE(P.kz13 == 0 && P.kz14 == 0 && P.kz37 == 0 && P.kz38 == 0,
                "gpunb_regcor_dev: KZ(13), KZ(14), KZ(37) and KZ(38) must be 0 "
                "(cloud forces, external force, CHECKL and FPCORR of "
                "REGCOR_GPU are not on the device)");
        REQUIRE((P.nb_flag == 0 || P.nb_flag == 1) &&
                    (P.m_flag == 0 || P.m_flag == 1),
                "gpunb_regcor_dev: nb_flag and m_flag must be 0 or 1");
        const gpunb_regcor_state &T = *st;
        REQUIRE(T.x0 && T.x0dot && T.xn && T.xndot && T.fi && T.fidot &&
                    T.fr && T.frdot && T.d0 && T.d0r && T.d1r && T.d2r &&
                    T.d3r && T.t0r && T.stepr && T.step && T.t0 &&
                    T.timenw && T.rs && T.body && T.list,
                "gpunb_regcor_dev: a state array is NULL");
        REQUIRE(d_idx && d_xi && d_vi && d_freg && d_fdr && d_nnb && d_off &&
                    d_list && d_counters,
                "gpunb_regcor_dev: a block array is NULL");
        REQUIRE(list_len >= 0, "gpunb_regcor_dev: list_len must be >= 0");
        set_device();
        S.rc_nlist.ensure((size_t)ni * P.lmax);
        S.rc_jj.ensure((size_t)ni * 2 * P.lmax);
        S.rc_step.ensure((size_t)ni);
        S.rc_tnw.ensure((size_t)ni);
        if (S.rc_mark.n < (size_t)P.ntot) {
            S.rc_mark.ensure((size_t)P.ntot);
            HIPCHK(hipMemsetAsync(S.rc_mark.p, 0, S.rc_mark.n * sizeof(int),
                                  S.st));
        }
        RegcorLaunch L;
        L.par = P;
        L.st = T;
        L.ni = ni;
        L.idx = d_idx;
        L.xi = d_xi;
        L.vi = d_vi;
        L.freg = d_freg;
        L.fdr = d_fdr;
        L.nnb = d_nnb;
        L.off = d_off;
        L.lst = d_list;
        L.list_len = list_len;
        L.jp = S.jp.p;
        L.nj = S.nj;
        L.nlist = S.rc_nlist.p;
        L.jj = S.rc_jj.p;
        L.step_new = S.rc_step.p;
        L.tnw_new = S.rc_tnw.p;
        L.mark = S.rc_mark.p;
        L.cnt = d_counters;
        L.err = S.err.p;
        launch_regcor(S.st, L);
    });
}

GPUNB_EXPORT int gpunb_sync(void) {
    return guarded("gpunb_sync", [&] {
        if (!S.devinit) return;
        set_device();
        err_fetch();
        HIPCHK(hipStreamSynchronize(S.st));
        err_check();
    });
}

GPUNB_EXPORT void *gpunb_stream(void) { return (void *)S.st; }

GPUNB_EXPORT const char *gpunb_version(void) {
    return "gpunb-mi355x 0.1 (gfx950, FP64-difference regular force)";
}

GPUNB_EXPORT int gpunb_timing(int enable) {
    return guarded("gpunb_timing", [&] {
        if (S.devinit) {
            set_device();
            HIPCHK(hipStreamSynchronize(S.st));
        }
        S.timing = enable != 0;
        S.ev_pairs.clear();
        S.ev_pairs_work.clear();
        S.ev_used = 0;
    });
}

GPUNB_EXPORT int gpunb_timing_read(double *ms_total, long long *launches,
                                   long long *pairs) {
    return guarded("gpunb_timing_read", [&] {
        double t = 0.0;
        long long w = 0;
        if (S.devinit && !S.ev_pairs.empty()) {
            set_device();
            HIPCHK(hipStreamSynchronize(S.st));
            for (size_t k = 0; k < S.ev_pairs.size(); ++k) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, S.ev_pool[S.ev_pairs[k].first],
                                           S.ev_pool[S.ev_pairs[k].second]));
                t += ms;
                w += S.ev_pairs_work[k];
            }
        }
        if (ms_total) *ms_total = t;
        if (launches) *launches = (long long)S.ev_pairs.size();
        if (pairs) *pairs = w;
        S.ev_pairs.clear();
        S.ev_pairs_work.clear();
        S.ev_used = 0;
    });
}

GPUNB_EXPORT int gpunb_ens_open(int ncl, const int *nbmax) {
    return guarded("gpunb_ens_open", [&] { ens_open(ncl, nbmax); });
}

GPUNB_EXPORT int gpunb_ens_close(void) {
    return guarded("gpunb_ens_close", [&] { ens_close(); });
}

GPUNB_EXPORT int gpunb_ens_send(int ncl, const int *nj, const double *m,
                                const double *x, const double *v) {
    return guarded("gpunb_ens_send", [&] {
        ens_send(ncl, nj, m, (const double(*)[3])x, (const double(*)[3])v);
    });
}

GPUNB_EXPORT int gpunb_ens_regf_dev(int ncl, const int *ni, const double *d_h2,
                                    const double *d_dtr, const double *d_xi,
                                    const double *d_vi, double *d_acc,
                                    double *d_jrk, double *d_pot, int nnbmax,
                                    int m_flag, int *d_nnb, int *d_nnid,
                                    int *d_list, int *d_off, int *d_total,
                                    long long list_cap) {
    return guarded("gpunb_ens_regf_dev", [&] {
        ens_regf_dev(ncl, ni, d_h2, d_dtr, d_xi, d_vi, d_acc, d_jrk, d_pot,
                     nnbmax, m_flag, d_nnb, d_nnid, d_list, d_off, d_total,
                     list_cap);
    });
}

GPUNB_EXPORT int gpunb_ens_graph(int enable) {
    return guarded("gpunb_ens_graph", [&] {
        if (S.devinit) {
            set_device();
            HIPCHK(hipStreamSynchronize(S.st));
        }
        ens_graph_drop();
        E.graph_on = enable != 0;
    });
}

GPUNB_EXPORT int gpunb_ens_persistent(int enable) {
    return guarded("gpunb_ens_persistent", [&] {
        if (S.devinit) {
            set_device();
            HIPCHK(hipStreamSynchronize(S.st));
        }
        ens_graph_drop();
        E.persistent = enable != 0;
    });
}

GPUNB_EXPORT int gpunb_ens_set_bucket_tiles(int nbucket, const int *tiles) {
    return guarded("gpunb_ens_set_bucket_tiles", [&] {
        REQUIRE(nbucket >= 0 && nbucket <= 24,
                "gpunb_ens_set_bucket_tiles: at most 24 buckets");
        for (int b = 0; b < 24; ++b)
            E.bucket_tiles[b] = (b < nbucket && tiles) ? std::max(0, tiles[b]) : 0;
        ens_graph_drop();
    });
}

GPUNB_EXPORT long long gpunb_ens_pairs(void) { return E.last_pairs; }

GPUNB_EXPORT int gpunb_set_tuning(int group, double rounds) {
    return guarded("gpunb_set_tuning", [&] {
        REQUIRE(group == 0 || group == 2 || group == 4,
                "gpunb_set_tuning: group must be 0, 2 or 4");
        REQUIRE(rounds >= 0.0, "gpunb_set_tuning: rounds must be >= 0");
        S.group = group;
        S.rounds = rounds;
    });
}

GPUNB_EXPORT int gpunb_set_precision(int prec) {
    return guarded("gpunb_set_precision", [&] {
        REQUIRE(prec == 0 || prec == 32 || prec == 64 || prec == 65,
                "gpunb_set_precision: prec must be 0, 32, 64 or 65");
        S.prec = prec;
    });
}

}  // extern "C"
