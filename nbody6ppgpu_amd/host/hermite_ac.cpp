// hermite_ac.cpp -- a host Hermite / Ahmad-Cohen block-step integrator that
// drives the regular-force engine through its C-ABI the way the reference's
// Fortran host does (intgrt.F: predict all j, send, regular force for the
// particles whose regular step is due, irregular force from the returned
// neighbour lists, correct, new steps; energy.F: gpupot at check times).
//
// It exists to measure what the engine must sustain in a real run: pair
// interactions per second over a whole integration and the relative energy
// error |dE/E| over N-body time units.  It is a compact restatement of the
// fourth-order Hermite Ahmad-Cohen scheme (Makino & Aarseth 1992, PASJ 44,
// 141), not a port of NBODY6++: no KS or chain regularisation, no stellar
// evolution, no tidal field.
//
// Per particle i the force is split in an irregular part (neighbours in the
// list L_i, summed in FP64 every irregular step dt_i, on the device through
// the irr_simd_* entry points or, with --irr host, on the host) and a regular part
// (all other particles, from gpunb_regf_ every regular step dtr_i, extended
// between regular steps by its Taylor polynomial).  Hermite interpolation of
// two force/jerk pairs gives the second and third derivatives for the
// corrector; when a regular step changes the list, the regular force seen
// with the old list is recovered as F_reg(L') + F_irr(L') - F_irr(L), so the
// regular polynomial stays that of one consistent split (regcor_gpu.F's
// gain/loss bookkeeping in its simplest form).  Steps follow Aarseth's
// criterion with ETAI / ETAR, quantised to powers of two and kept
// commensurate (block steps); an overflowing list is retried with a smaller
// radius (util_gpu.F:70-101).
//
// Neighbour radius RS after each regular step (--rs):
//   ac     the Ahmad-Cohen volume rule of regcor_gpu.F:146-253 (default);
//   simple RS *= clamp((NNBOPT/NNB)^(1/3), 0.8, 1.25);
//   mlp    the learned predictor (gpunb_mlp_predict, weights from --mlp),
//          with the ac rule as the fallback for a body whose learned radius
//          gave a list outside [NNBMIN, NNBMAX] or an overflow.
// --log-train writes up to --log-max (feature, target) pairs drawn uniformly
// (reservoir sampling) from all regular steps: the features of a body at one
// regular step and the radius that would have given exactly NNBOPT
// neighbours at its next one.
//
// Usage:  hermite_ac <in.bin> <t_end> [options]   (see parse_args)
// in.bin: int32 n, then FP64 m[n], x[n][3], v[n][3].
// Output: one JSON object per energy check on stdout; with --conf3 DIR also
// a conf.3 snapshot per check (the reference's binary format, output.F).
#include "gpunb.h"

#include <omp.h>
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <random>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

double wtime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct Opts {
    double t_end = 1.0;
    double etai = 0.05;
    double etar = 0.1;
    double dtadj = 0.125;      // energy check interval (power of two)
    double dtmin = 1.0 / (1 << 30);
    int nnbopt = 48;
    int lmax = 601;            // LMAX (params.h), list row stride
    int nnbmax = 550;          // NNBMAX = min(N/2, LMAX-50) (input.F:121)
    int use_predict = 1;       // 1: device predictor; 0: host predict + send
    int prec = 0;              // gpunb_set_precision
    double rs0 = 0.0;          // initial RS (0: from nearest neighbours)
    int rs_rule = 0;           // 0 ac, 1 simple, 2 mlp
    double nnbmin = 0.0;       // 0: ZNBMIN = max(0.01 NNBMAX, 1) (input.F:122)
    std::string mlp_path;      // weights file for --rs mlp
    std::string log_path;      // training pairs out (--log-train)
    std::string conf3_dir;     // conf.3 snapshots at every energy check
    std::string dump_path;     // checkpoint written at every energy check
    std::string restart_path;  // checkpoint to continue from
    long long log_max = 2000000;
    unsigned log_seed = 12345;
    double progress = 0.0;     // > 0: a progress line on stderr every so many s
    int irr_dev = 1;           // irregular force: 1 device (irr_simd_*), 0 host
    int irr_min = 256;         // blocks with fewer bodies: host irregular force
};

// Features of the learned radius predictor (8 floats per body), all taken
// after the body's regular step:
//   0 log10 local mass density within RS, (m_i + sum_L m_j) / (4/3 pi RS^3)
//   1 NNB / NNBOPT of the list just built
//   2 log10 |a|        3 log10 |jerk|      (total force and jerk)
//   4 log10 |x - density centre|           5 log10 (m / mean mass)
//   6 log2 (1 / new regular step)          7 log10 RS used for the list
// Target: log10 of the RS that gives NNBOPT neighbours at the next regular
// step.  A logged record is 10 floats: features, target, censored flag (1
// when the next list held fewer than NNBOPT, the target then being the
// volume estimate RS (NNBOPT / NNB)^(1/3), at most 2 RS).
constexpr int kNF = 8;
constexpr int kRec = kNF + 2;

struct Body {
    double m;
    double x0[3], v0[3];       // state at t0
    double a[3], j[3];         // total force and jerk at t0
    double fi[3], fid[3];      // irregular force, jerk at t0 (list L)
    double fi2[3], fi3[3];     // irregular 2nd/3rd derivatives at t0
    double fr[3], frd[3], fr2[3], fr3[3];  // regular polynomial at t0r
    double t0, dt, t0r, dtr;
    double rs;
    std::vector<int> list;     // neighbours (0-based, self excluded)
    float feat[kNF];           // features at the last regular step
    bool has_feat = false;
    bool learned = false;      // rs came from the learned predictor
};

inline double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Largest power of two <= x (x > 0).
double pow2_floor(double x) {
    int e;
    double f = std::frexp(x, &e);  // x = f 2^e, f in [0.5, 1)
    (void)f;
    return std::ldexp(1.0, e - 1);
}

// Aarseth's step criterion.
double aarseth(const double *f, const double *f1, const double *f2,
               const double *f3, double eta) {
    const double a0 = norm3(f), a1 = norm3(f1), a2 = norm3(f2), a3 = norm3(f3);
    const double num = a0 * a2 + a1 * a1;
    const double den = a1 * a3 + a2 * a2;
    if (den <= 0.0 || num <= 0.0) return 1e30;
    return std::sqrt(eta * num / den);
}

// Predict body b to time t (cubic, the order the engine's predictor uses).
inline void predict_cubic(const Body &b, double t, double *x, double *v) {
    const double s = t - b.t0;
    for (int k = 0; k < 3; ++k) {
        x[k] = ((b.j[k] * (s / 6.0) + 0.5 * b.a[k]) * s + b.v0[k]) * s + b.x0[k];
        v[k] = (b.j[k] * (0.5 * s) + b.a[k]) * s + b.v0[k];
    }
}

// Force and jerk on (xi, vi) from the listed bodies predicted to t (FP64).
void irr_force(const std::vector<Body> &B, const std::vector<int> &list,
               double t, const double *xi, const double *vi, double *f,
               double *fd) {
    f[0] = f[1] = f[2] = fd[0] = fd[1] = fd[2] = 0.0;
    for (int jj : list) {
        double xj[3], vj[3];
        predict_cubic(B[jj], t, xj, vj);
        const double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
        const double ux = vj[0] - vi[0], uy = vj[1] - vi[1], uz = vj[2] - vi[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0.0) continue;
        const double rinv2 = 1.0 / r2;
        const double mr3 = B[jj].m * rinv2 * std::sqrt(rinv2);
        const double al = 3.0 * (dx * ux + dy * uy + dz * uz) * rinv2;
        f[0] += mr3 * dx;
        f[1] += mr3 * dy;
        f[2] += mr3 * dz;
        fd[0] += mr3 * (ux - al * dx);
        fd[1] += mr3 * (uy - al * dy);
        fd[2] += mr3 * (uz - al * dz);
    }
}

// Second and third time derivatives of the force m r/|r|^3 that body j
// (relative position r, velocity v, acceleration a, jerk jk) exerts
// (Nitadori & Makino 2008, eqs. for the sixth-order Hermite scheme).
inline void pair_d23(double m, const double *r, const double *v,
                     const double *a, const double *jk, double *f2,
                     double *f3) {
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double ri2 = 1.0 / r2;
    const double rv = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
    const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double ra = r[0] * a[0] + r[1] * a[1] + r[2] * a[2];
    const double va = v[0] * a[0] + v[1] * a[1] + v[2] * a[2];
    const double rj = r[0] * jk[0] + r[1] * jk[1] + r[2] * jk[2];
    const double al = rv * ri2;
    const double be = (vv + ra) * ri2 + al * al;
    const double ga = (3.0 * va + rj) * ri2 + al * (3.0 * be - 4.0 * al * al);
    const double mr3 = m * ri2 * std::sqrt(ri2);
    for (int k = 0; k < 3; ++k) {
        const double F0 = mr3 * r[k];
        const double F1 = mr3 * v[k] - 3.0 * al * F0;
        const double F2 = mr3 * a[k] - 6.0 * al * F1 - 3.0 * be * F0;
        f2[k] = F2;
        f3[k] = mr3 * jk[k] - 9.0 * al * F2 - 9.0 * be * F1 - 3.0 * ga * F0;
    }
}

// Hermite interpolation: derivatives 2 and 3 at the start of an interval of
// length h from (f0, f0') and (f1, f1').
inline void hermite23(const double *f0, const double *d0, const double *f1,
                      const double *d1, double h, double *f2, double *f3) {
    const double h2 = h * h;
    for (int k = 0; k < 3; ++k) {
        f2[k] = (-6.0 * (f0[k] - f1[k]) - h * (4.0 * d0[k] + 2.0 * d1[k])) / h2;
        f3[k] = (12.0 * (f0[k] - f1[k]) + 6.0 * h * (d0[k] + d1[k])) / (h2 * h);
    }
}

struct Run {
    Opts o;
    int n = 0;
    std::vector<Body> B;
    double t = 0.0;
    // work counters
    long long reg_pairs = 0, irr_pairs = 0, reg_calls = 0, irr_steps = 0,
              reg_steps = 0, blocks = 0, overflows = 0;
    double t_reg = 0.0, t_irr = 0.0, t_pred = 0.0, t_pot = 0.0, t_mlp = 0.0,
           t_jp = 0.0;
    // Block-step schedule: next time -> bodies due then (times are dyadic).
    std::map<double, std::vector<int>> due;
    // staging for the ABI
    std::vector<double> h2, dtrv, xi, vi, acc, jrk, pot, jm, jx, jv;
    std::vector<int> lst;
    // body -> position in the current block's active / regular set (-1)
    std::vector<int> rpos, rk;

    // Cluster quantities of the Ahmad-Cohen radius rule, refreshed at every
    // energy check as adjust.F does: density centre and core radius
    // (Casertano & Hut 1985, 6th-neighbour densities from the lists),
    // half-mass radius RSCALE about the density centre, and
    // RSFAC = max(25/TCR, 1.5 VC/RSMIN) (adjust.F:220).
    double rdens[3] = {0.0, 0.0, 0.0};
    double rscale = 1.0, rc2 = 1.0, rsfac = 10.0, znbmin = 1.0, nnbmin = 1.0;
    double tcr = 2.0 * std::sqrt(2.0), vc = std::sqrt(0.5), mmean = 1.0;
    // Five-neighbour density and number density per body (core.f:90-98),
    // from the last update_centre; potentials of the last energy check.
    std::vector<double> rho_b, xn_b, phidbl;
    // Radius statistics.
    long long learned_steps = 0, fallbacks = 0, nnb_dev_n = 0;
    double nnb_dev_sum = 0.0;
    // Training pairs (reservoir sample of all regular steps).
    std::vector<float> logbuf;
    long long log_seen = 0;
    std::mt19937_64 log_rng;

    void abi_check() {
        char buf[512];
        if (gpunb_last_error(buf, sizeof buf) > 0) {
            fprintf(stderr, "hermite_ac: library error: %s\n", buf);
            exit(2);
        }
    }

    // Positions/velocities of every body at time t on the device: either
    // the device predictor (state sent with gpunb_jpset_ after each
    // correction) or a host prediction sent with gpunb_send_.
    void send_j(double tt) {
        double w0 = wtime();
        if (o.use_predict) {
            gpunb_predict_(&tt);
        } else {
            jm.resize(n);
            jx.resize(3 * (size_t)n);
            jv.resize(3 * (size_t)n);
#pragma omp parallel for schedule(static)
            for (int q = 0; q < n; ++q) {
                jm[q] = B[q].m;
                predict_cubic(B[q], tt, &jx[3 * (size_t)q], &jv[3 * (size_t)q]);
            }
            gpunb_send_(&n, jm.data(), (double(*)[3])jx.data(),
                        (double(*)[3])jv.data());
        }
        abi_check();
        t_pred += wtime() - w0;
    }

    void jpset(const std::vector<int> &idx) {
        if ((!o.use_predict && !o.irr_dev) || idx.empty()) return;
        const double w0 = wtime();
        const int k = (int)idx.size();
        std::vector<int> addr(k);
        std::vector<double> X(3 * k), V(3 * k), F(3 * k), FD(3 * k), M(k), T(k);
#pragma omp parallel for schedule(static) if (k >= 4096)
        for (int q = 0; q < k; ++q) {
            const Body &b = B[idx[q]];
            addr[q] = idx[q] + 1;
            for (int c = 0; c < 3; ++c) {
                X[3 * q + c] = b.x0[c];
                V[3 * q + c] = b.v0[c];
                F[3 * q + c] = 0.5 * b.a[c];
                FD[3 * q + c] = b.j[c] / 6.0;
            }
            M[q] = b.m;
            T[q] = b.t0;
        }
        int kk = k;
        if (o.use_predict)
        gpunb_jpset_(&kk, addr.data(), (double(*)[3])X.data(),
                     (double(*)[3])V.data(), (double(*)[3])F.data(),
                     (double(*)[3])FD.data(), M.data(), T.data());
        abi_check();
        if (o.irr_dev) {
            gpunb_irr_set_jp(k, addr.data(), X.data(), V.data(), F.data(),
                             FD.data(), M.data(), T.data());
            abi_check();
        }
        t_jp += wtime() - w0;
    }

    // Lists of bodies idx (0-based members) to the irregular library.
    void irr_lists(const std::vector<int> &idx,
                   const std::vector<std::vector<int>> &L) {
        if (!o.irr_dev || idx.empty()) return;
        std::vector<int> addr(idx.size()), cnt(idx.size()), flat;
        for (size_t q = 0; q < idx.size(); ++q) {
            addr[q] = idx[q] + 1;
            cnt[q] = (int)L[q].size();
            for (int j : L[q]) flat.push_back(j + 1);
        }
        if (flat.empty()) flat.push_back(0);
        gpunb_irr_set_lists((int)idx.size(), addr.data(), cnt.data(), flat.data());
        abi_check();
    }

    // Irregular force at t of bodies idx on the device (their lists as last
    // sent): F, FD [idx.size()][3].
    void irr_dev(double t, const std::vector<int> &idx, double *F, double *FD) {
        std::vector<int> addr(idx.size()), nn(idx.size());
        for (size_t q = 0; q < idx.size(); ++q) addr[q] = idx[q] + 1;
        int ni = (int)idx.size();
        double tt = t;
        irr_simd_firr_vec_(&tt, &ni, addr.data(), (double(*)[3])F,
                           (double(*)[3])FD, nn.data());
        abi_check();
    }

    // Regular force + new lists for bodies R at their positions (x, v).
    // Retries overflowing lists with a smaller radius (util_gpu.F:70-101).
    void regular(const std::vector<int> &R, const std::vector<double> &X,
                 const std::vector<double> &V, std::vector<double> &FR,
                 std::vector<double> &FRD, std::vector<std::vector<int>> &L,
                 std::vector<char> *ovf = nullptr) {
        const int nr = (int)R.size();
        FR.assign(3 * (size_t)nr, 0.0);
        FRD.assign(3 * (size_t)nr, 0.0);
        L.assign(nr, {});
        if (ovf) ovf->assign(nr, 0);
        std::vector<int> todo(nr);
        for (int q = 0; q < nr; ++q) todo[q] = q;
        while (!todo.empty()) {
            int ni = (int)todo.size();
            h2.resize(ni);
            dtrv.resize(ni);
            xi.resize(3 * (size_t)ni);
            vi.resize(3 * (size_t)ni);
            acc.resize(3 * (size_t)ni);
            jrk.resize(3 * (size_t)ni);
            pot.resize(ni);
            lst.resize((size_t)ni * o.lmax);
            for (int q = 0; q < ni; ++q) {
                const int r = todo[q];
                const Body &b = B[R[r]];
                h2[q] = b.rs * b.rs;
                dtrv[q] = b.dtr;
                for (int c = 0; c < 3; ++c) {
                    xi[3 * q + c] = X[3 * (size_t)r + c];
                    vi[3 * q + c] = V[3 * (size_t)r + c];
                }
            }
            double w0 = wtime();
            int mflag = 0;
            gpunb_regf_(&ni, h2.data(), dtrv.data(), (double(*)[3])xi.data(),
                        (double(*)[3])vi.data(), (double(*)[3])acc.data(),
                        (double(*)[3])jrk.data(), pot.data(), &o.lmax,
                        &o.nnbmax, lst.data(), &mflag);
            abi_check();
            t_reg += wtime() - w0;
            reg_calls++;
            reg_pairs += (long long)ni * n;
            std::vector<int> again;
            for (int q = 0; q < ni; ++q) {
                const int r = todo[q];
                Body &b = B[R[r]];
                const int *row = &lst[(size_t)q * o.lmax];
                const int nnb = row[0];
                if (nnb < 0) {
                    // Overflow: shrink RS (NNBOPT/NNB)^(1/3) and redo.
                    overflows++;
                    if (ovf) (*ovf)[r] = 1;
                    b.rs *= std::pow((double)o.nnbopt / (double)(-nnb), 1.0 / 3.0);
                    again.push_back(r);
                    continue;
                }
                L[r].clear();
                for (int k = 1; k <= nnb; ++k)
                    if (row[k] != R[r]) L[r].push_back(row[k]);
                for (int c = 0; c < 3; ++c) {
                    FR[3 * (size_t)r + c] = acc[3 * q + c];
                    FRD[3 * (size_t)r + c] = jrk[3 * q + c];
                }
            }
            todo.swap(again);
        }
    }

    // New irregular step for b at time tt from its irregular derivatives.
    double new_dt(const Body &b, double tt, double dt_old, double cap) {
        double f2[3], f3[3];
        for (int k = 0; k < 3; ++k) {
            f2[k] = b.fi2[k];
            f3[k] = b.fi3[k];
        }
        double d = b.list.empty() ? cap : aarseth(b.fi, b.fid, f2, f3, o.etai);
        d = std::min(d, 2.0 * dt_old);
        d = std::min(d, cap);
        d = std::max(d, o.dtmin);
        double q = pow2_floor(d);
        while (std::fmod(tt, q) != 0.0 && q > o.dtmin) q *= 0.5;
        return q;
    }

    double new_dtr(const Body &b, double tt, double dtr_old, double dt) {
        double d = aarseth(b.fr, b.frd, b.fr2, b.fr3, o.etar);
        d = std::min(d, 2.0 * dtr_old);
        d = std::min(d, o.dtadj);
        double q = pow2_floor(std::max(d, dt));
        while (std::fmod(tt, q) != 0.0 && q > dt) q *= 0.5;
        return std::max(q, dt);
    }

    // Ahmad-Cohen neighbour-radius update after a regular step
    // (regcor_gpu.F:146-253 with NB_FLAG = 1; no KS c.m. branch, no CHECKL).
    // nnb: new list size, nnb0: previous list size, xi/vi at the step time,
    // dtr: the regular step just completed, m3: min over the members of
    // RIJ + RSDOT*DTR (used when nnb <= 3; +inf without members).
    double ac_rs(double rs, int nnb, int nnb0, const double *xi,
                 const double *vi, double dtr, double m3) const {
        double d[3];
        for (int c = 0; c < 3; ++c) d[c] = xi[c] - rdens[c];
        const double ri2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double rh2 = rscale * rscale;
        // NNB = 0 takes the reference's zero-neighbour path first; the
        // volume rule below then sees NNB = 1.
        const double fnb = std::max(nnb, 1);
        const double fac = 1.0 + 0.1 * (o.nnbopt - fnb) / fnb;
        double a3 = fnb * fac;
        const double nbave = std::min(o.nnbopt + 150,
                                      std::max(o.nnbmax - 50, (int)znbmin));
        a3 = std::min(a3, 0.9 * nbave);
        const int nbp = (int)a3;
        if (ri2 > rh2) a3 *= std::max(rscale / std::sqrt(ri2), 0.1);
        else a3 = std::max(a3, znbmin);
        double a4 = a3 / fnb;
        // Inertial factor against oscillations of RS.
        if ((a3 - nnb0) * (a3 - fnb) < 0.0) a4 = std::sqrt(a4);
        const double ridot = d[0] * vi[0] + d[1] * vi[1] + d[2] * vi[2];
        if (ri2 > rc2 && ri2 < 100.0 * rh2) a4 *= 1.0 + ridot * dtr / ri2;
        const double a5 = std::min(rsfac * dtr, 0.25);
        a4 = std::min(std::max(a4 - 1.0, -a5), a5);
        const double t3 = a4 / 3.0;
        double a1 = 1.0 + t3 - t3 * t3;   // second-order cube root
        if (nnb < znbmin && nbp > znbmin && a1 < 1.0) a1 = 1.05;
        if (std::fabs(a1 - 1.0) > 0.003) {
            if (rs * rs > ri2 && rs * rs < 2.0 * ri2 && nnb > 5) a1 = std::sqrt(a1);
            rs *= a1;
        }
        if (nnb > nbave && ridot < 0.0) rs *= 1.0 - (nnb - nbave) / (3.0 * nnb);
        if (nnb <= 3 && ri2 < 100.0 * rh2) rs = std::max(std::min(2.0 * rs, m3), 1.1 * rs);
        return rs;
    }

    // The simple volume rule (previous default of this driver).
    double simple_rs(double rs, int nnb) const {
        const double nb = std::max<double>(1.0, (double)nnb);
        double f = std::cbrt((double)o.nnbopt / nb);
        return rs * std::min(std::max(f, 0.8), 1.25);
    }

    // Density centre, core radius, half-mass radius and RSFAC from the
    // current state (every body at t0 == t) and lists.
    void update_centre() {
        std::vector<double> rho(n, 0.0), xn(n, 0.0);
#pragma omp parallel for schedule(dynamic, 256)
        for (int q = 0; q < n; ++q) {
            const Body &b = B[q];
            const size_t k = b.list.size();
            if (k < 6) continue;
            std::vector<std::pair<double, double>> dm(k);
            for (size_t u = 0; u < k; ++u) {
                const Body &bj = B[b.list[u]];
                double r2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double e = bj.x0[c] - b.x0[c];
                    r2 += e * e;
                }
                dm[u] = {r2, bj.m};
            }
            std::partial_sort(dm.begin(), dm.begin() + 6, dm.end());
            double ms = 0.0;
            for (int u = 0; u < 5; ++u) ms += dm[u].second;
            const double r6 = std::sqrt(dm[5].first);
            if (r6 > 0.0) {
                xn[q] = 1.0 / (r6 * r6 * r6);
                rho[q] = ms * xn[q];
            }
        }
        rho_b = rho;
        xn_b = xn;
        double sw = 0.0, sx[3] = {0, 0, 0};
        for (int q = 0; q < n; ++q) {
            sw += rho[q];
            for (int c = 0; c < 3; ++c) sx[c] += rho[q] * B[q].x0[c];
        }
        if (sw > 0.0)
            for (int c = 0; c < 3; ++c) rdens[c] = sx[c] / sw;
        double s2 = 0.0, sr = 0.0;
        std::vector<std::pair<double, double>> rm(n);
        double mt = 0.0;
        for (int q = 0; q < n; ++q) {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double e = B[q].x0[c] - rdens[c];
                r2 += e * e;
            }
            s2 += rho[q] * rho[q];
            sr += rho[q] * rho[q] * r2;
            rm[q] = {r2, B[q].m};
            mt += B[q].m;
        }
        if (s2 > 0.0) rc2 = sr / s2;
        std::sort(rm.begin(), rm.end());
        double cum = 0.0;
        for (int q = 0; q < n; ++q) {
            cum += rm[q].second;
            if (cum >= 0.5 * mt) {
                rscale = std::sqrt(rm[q].first);
                break;
            }
        }
        double rsmin = 1e30;
        for (int q = 0; q < n; ++q)
            if (!B[q].list.empty()) rsmin = std::min(rsmin, B[q].rs);
        if (rsmin >= 1e30) rsmin = rscale;
        rsfac = std::max(25.0 / tcr, 1.5 * vc / rsmin);
    }

    // Features of body b after its regular step (see kNF above); rsl = RS
    // of the list just built, dtrn = the new regular step.
    void features(const Body &b, double rsl, double dtrn, float *f) const {
        double ml = b.m;
        for (int j : b.list) ml += B[j].m;
        const double vol = 4.0 / 3.0 * M_PI * rsl * rsl * rsl;
        double d[3];
        for (int c = 0; c < 3; ++c) d[c] = b.x0[c] - rdens[c];
        auto lg = [](double v) { return (float)std::log10(std::max(v, 1e-30)); };
        f[0] = lg(ml / vol);
        f[1] = (float)((double)b.list.size() / o.nnbopt);
        f[2] = lg(norm3(b.a));
        f[3] = lg(norm3(b.j));
        f[4] = lg(norm3(d));
        f[5] = lg(b.m / mmean);
        f[6] = (float)std::log2(1.0 / dtrn);
        f[7] = lg(rsl);
    }

    // Target of the training pair closed by a new list L of body i at tn:
    // log10 of the RS that gives exactly NNBOPT members, from the members'
    // list-criterion distances min(r, r at t + dtr) (the engine's predicate);
    // returns the censored flag.
    int target(const std::vector<int> &L, double tn, const double *xi,
               const double *vi, double dtr, double rs, float *t) const {
        const size_t k = (size_t)o.nnbopt;
        if (L.size() < k) {
            const double f = std::min(std::cbrt((double)k / std::max<size_t>(L.size(), 1)), 2.0);
            *t = (float)std::log10(rs * f);
            return 1;
        }
        std::vector<double> c(L.size());
        for (size_t u = 0; u < L.size(); ++u) {
            double xj[3], vj[3], r2 = 0.0, p2 = 0.0;
            predict_cubic(B[L[u]], tn, xj, vj);
            for (int q = 0; q < 3; ++q) {
                const double dx = xj[q] - xi[q], dv = vj[q] - vi[q];
                r2 += dx * dx;
                p2 += (dx + dtr * dv) * (dx + dtr * dv);
            }
            c[u] = std::max(std::min(r2, p2), 1e-300);
        }
        std::nth_element(c.begin(), c.begin() + (k - 1), c.end());
        const double ck = c[k - 1];
        double ck1 = rs * rs;
        if (c.size() > k) ck1 = *std::min_element(c.begin() + k, c.end());
        *t = (float)(0.25 * std::log10(ck * std::max(ck1, ck)));
        return 0;
    }

    void log_add(const float *feat, float tgt, int cens) {
        if (o.log_path.empty() || o.log_max <= 0) return;
        long long slot = log_seen;
        if (log_seen >= o.log_max) {
            slot = (long long)(log_rng() % (unsigned long long)(log_seen + 1));
            if (slot >= o.log_max) slot = -1;
        } else {
            logbuf.resize((size_t)(log_seen + 1) * kRec);
        }
        log_seen++;
        if (slot < 0) return;
        float *r = &logbuf[(size_t)slot * kRec];
        for (int u = 0; u < kNF; ++u) r[u] = feat[u];
        r[kNF] = tgt;
        r[kNF + 1] = (float)cens;
    }

    void write_log() const {
        if (o.log_path.empty()) return;
        FILE *f = fopen(o.log_path.c_str(), "wb");
        if (!f) {
            perror(o.log_path.c_str());
            exit(1);
        }
        fwrite(logbuf.data(), sizeof(float), logbuf.size(), f);
        fclose(f);
    }

    // Weights file: gpunb_mlp_nparams() FP32 parameters, then xmean[8],
    // xstd[8], ymean, ystd.
    void load_mlp() {
        const int np = gpunb_mlp_nparams();
        std::vector<float> w((size_t)np + 2 * kNF + 2);
        FILE *f = fopen(o.mlp_path.c_str(), "rb");
        if (!f) {
            perror(o.mlp_path.c_str());
            exit(1);
        }
        const size_t got = fread(w.data(), sizeof(float), w.size(), f);
        fclose(f);
        if (got != w.size()) {
            fprintf(stderr, "hermite_ac: %s: expected %zu floats\n",
                    o.mlp_path.c_str(), w.size());
            exit(1);
        }
        gpunb_mlp_set_params(w.data());
        gpunb_mlp_set_norm(&w[np], &w[np + kNF], w[np + 2 * kNF],
                           w[np + 2 * kNF + 1]);
        abi_check();
    }

    void init(const std::vector<double> &m, const std::vector<double> &x,
              const std::vector<double> &v) {
        B.resize(n);
        rpos.assign(n, -1);
        rk.assign(n, -1);
        double mt = 0.0;
        for (int q = 0; q < n; ++q) mt += m[q];
        mmean = mt / n;
        if (!o.log_path.empty()) log_rng.seed(o.log_seed);
        for (int q = 0; q < n; ++q) {
            Body &b = B[q];
            b = Body{};
            b.m = m[q];
            for (int c = 0; c < 3; ++c) {
                b.x0[c] = x[3 * (size_t)q + c];
                b.v0[c] = v[3 * (size_t)q + c];
            }
            const double r = norm3(b.x0);
            // FPOLY0 start values (fpoly0.F:52-55): RS0 sqrt(1 + r^2),
            // STEPR = min(SMAX/8 sqrt(1 + r^2), SMAX).
            b.rs = o.rs0 * std::sqrt(1.0 + r * r);
            b.dtr = std::min(o.dtadj / 8.0 * std::sqrt(1.0 + r * r), o.dtadj);
        }
        // Everybody at t = 0: regular force and lists for all.
        std::vector<int> R(n);
        for (int q = 0; q < n; ++q) R[q] = q;
        std::vector<double> X(3 * (size_t)n), V(3 * (size_t)n), FR, FRD;
        for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) {
                X[3 * (size_t)q + c] = B[q].x0[c];
                V[3 * (size_t)q + c] = B[q].v0[c];
            }
        std::vector<int> all(n);
        for (int q = 0; q < n; ++q) all[q] = q;
        jpset(all);   // device predictor state (a = j = 0 until the first force)
        send_j(0.0);
        std::vector<std::vector<int>> L;
        regular(R, X, V, FR, FRD, L);
        double w0 = wtime();
        std::vector<double> FI0, FID0;
        if (o.irr_dev) {
            irr_lists(all, L);
            FI0.resize(3 * (size_t)n);
            FID0.resize(3 * (size_t)n);
            irr_dev(0.0, all, FI0.data(), FID0.data());
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int q = 0; q < n; ++q) {
            Body &b = B[q];
            b.list = L[q];
            if (o.irr_dev) {
                for (int c = 0; c < 3; ++c) {
                    b.fi[c] = FI0[3 * (size_t)q + c];
                    b.fid[c] = FID0[3 * (size_t)q + c];
                }
            } else {
                irr_force(B, b.list, 0.0, b.x0, b.v0, b.fi, b.fid);
            }
            for (int c = 0; c < 3; ++c) {
                b.fr[c] = FR[3 * (size_t)q + c];
                b.frd[c] = FRD[3 * (size_t)q + c];
                b.a[c] = b.fi[c] + b.fr[c];
                b.j[c] = b.fid[c] + b.frd[c];
            }
            // Start with a conservative second-order step.
            const double a0 = norm3(b.a), a1 = norm3(b.j);
            double d = (a1 > 0) ? 0.1 * o.etai * a0 / a1 : o.dtadj;
            const double ai = norm3(b.fi), ai1 = norm3(b.fid);
            if (!b.list.empty() && ai1 > 0) d = std::min(d, 0.1 * o.etai * ai / ai1);
            d = std::min(std::max(d, o.dtmin), o.dtadj / 8.0);
            b.dt = pow2_floor(d);
            double dr = (norm3(b.frd) > 0) ? 0.1 * o.etar * norm3(b.fr) / norm3(b.frd) : o.dtadj;
            dr = std::min(std::max(dr, b.dt), o.dtadj);
            b.dtr = std::max(pow2_floor(dr), b.dt);
            b.t0 = 0.0;
            b.t0r = 0.0;
        }
        t_irr += wtime() - w0;
        irr_pairs += 0;
        for (int q = 0; q < n; ++q) irr_pairs += (long long)B[q].list.size();
        for (int q = 0; q < n; ++q) due[B[q].t0 + B[q].dt].push_back(q);
        jpset(all);
    }

    double energy(double *ekin, double *epot) {
        // Called at a synchronisation time: every body has t0 == t.
        std::vector<double> m(n), x(3 * (size_t)n), ph(n);
        double ke = 0.0;
        for (int q = 0; q < n; ++q) {
            m[q] = B[q].m;
            for (int c = 0; c < 3; ++c) x[3 * (size_t)q + c] = B[q].x0[c];
            ke += 0.5 * B[q].m *
                  (B[q].v0[0] * B[q].v0[0] + B[q].v0[1] * B[q].v0[1] +
                   B[q].v0[2] * B[q].v0[2]);
        }
        double w0 = wtime();
        int irank = 0, istart = 1, nn = n;
        gpupot_(&irank, &istart, &nn, &nn, m.data(), (double(*)[3])x.data(),
                ph.data());
        abi_check();
        t_pot += wtime() - w0;
        double pe = 0.0;
        for (int q = 0; q < n; ++q) pe -= 0.5 * m[q] * ph[q];
        phidbl = ph;
        *ekin = ke;
        *epot = pe;
        return ke + pe;
    }

    // Checkpoint of the whole integration state at an energy check (all
    // bodies synchronised): the options that shape the run, the cluster
    // quantities, the counters, E0 and every body's state and list.  A run
    // restarted from it continues bit for bit as the uninterrupted run does
    // (the device state is rebuilt from the bodies; wall-clock timers are
    // not part of it).  Layout: "NB6AMDCK", int32 version, int32 n, then
    // native doubles / int64 / int32 as written below.
    static constexpr int kCkptVersion = 1;
    void dump(const std::string &path, double e0) const {
        const std::string tmp = path + ".tmp";
        FILE *f = fopen(tmp.c_str(), "wb");
        if (!f) {
            perror(tmp.c_str());
            exit(1);
        }
        bool ok = true;
        auto put = [&](const void *p, size_t bytes) {
            ok = ok && fwrite(p, 1, bytes, f) == bytes;
        };
        put("NB6AMDCK", 8);
        const int32_t hdr[2] = {kCkptVersion, n};
        put(hdr, sizeof hdr);
        const double od[4] = {o.etai, o.etar, o.dtadj, o.dtmin};
        const int32_t oi[8] = {o.nnbopt, o.lmax, o.nnbmax, o.rs_rule,
                               o.irr_dev, o.prec, o.use_predict ? 1 : 0,
                               o.irr_min};
        put(od, sizeof od);
        put(oi, sizeof oi);
        const double sc[14] = {t, e0, rdens[0], rdens[1], rdens[2], rscale,
                               rc2, rsfac, znbmin, nnbmin, tcr, vc, mmean,
                               nnb_dev_sum};
        put(sc, sizeof sc);
        const long long cnt[11] = {reg_pairs, irr_pairs, reg_calls, irr_steps,
                                   reg_steps, blocks, overflows, learned_steps,
                                   fallbacks, nnb_dev_n, log_seen};
        put(cnt, sizeof cnt);
        for (const Body &b : B) {
            put(&b.m, sizeof b.m);
            put(b.x0, sizeof b.x0);
            put(b.v0, sizeof b.v0);
            put(b.a, sizeof b.a);
            put(b.j, sizeof b.j);
            put(b.fi, sizeof b.fi);
            put(b.fid, sizeof b.fid);
            put(b.fi2, sizeof b.fi2);
            put(b.fi3, sizeof b.fi3);
            put(b.fr, sizeof b.fr);
            put(b.frd, sizeof b.frd);
            put(b.fr2, sizeof b.fr2);
            put(b.fr3, sizeof b.fr3);
            const double tt[5] = {b.t0, b.dt, b.t0r, b.dtr, b.rs};
            put(tt, sizeof tt);
            put(b.feat, sizeof b.feat);
            const int32_t fl[3] = {b.has_feat ? 1 : 0, b.learned ? 1 : 0,
                                   (int32_t)b.list.size()};
            put(fl, sizeof fl);
            put(b.list.data(), b.list.size() * sizeof(int));
        }
        ok = (fclose(f) == 0) && ok;
        if (!ok || rename(tmp.c_str(), path.c_str()) != 0) {
            fprintf(stderr, "checkpoint: write to %s failed\n", path.c_str());
            exit(1);
        }
    }

    // Restores a checkpoint written by dump() for the same n and options;
    // returns E0.  The device state (predictor records, irregular-force
    // records and lists) and the block schedule are rebuilt from the bodies.
    double restore(const std::string &path) {
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) {
            perror(path.c_str());
            exit(1);
        }
        auto get = [&](void *p, size_t bytes) {
            if (fread(p, 1, bytes, f) != bytes) {
                fprintf(stderr, "checkpoint %s: truncated\n", path.c_str());
                exit(1);
            }
        };
        auto fail = [&](const char *what) {
            fprintf(stderr, "checkpoint %s: %s\n", path.c_str(), what);
            exit(1);
        };
        char magic[8];
        get(magic, 8);
        if (memcmp(magic, "NB6AMDCK", 8) != 0) fail("not a checkpoint");
        int32_t hdr[2];
        get(hdr, sizeof hdr);
        if (hdr[0] != kCkptVersion) fail("unknown version");
        if (hdr[1] != n) fail("written for another N");
        double od[4];
        int32_t oi[8];
        get(od, sizeof od);
        get(oi, sizeof oi);
        const double odn[4] = {o.etai, o.etar, o.dtadj, o.dtmin};
        const int32_t oin[8] = {o.nnbopt, o.lmax, o.nnbmax, o.rs_rule,
                                o.irr_dev, o.prec, o.use_predict ? 1 : 0,
                                o.irr_min};
        if (memcmp(od, odn, sizeof od) != 0 || memcmp(oi, oin, sizeof oi) != 0)
            fail("written with other integration options");
        double sc[14];
        get(sc, sizeof sc);
        t = sc[0];
        const double e0 = sc[1];
        rdens[0] = sc[2];
        rdens[1] = sc[3];
        rdens[2] = sc[4];
        rscale = sc[5];
        rc2 = sc[6];
        rsfac = sc[7];
        znbmin = sc[8];
        nnbmin = sc[9];
        tcr = sc[10];
        vc = sc[11];
        mmean = sc[12];
        nnb_dev_sum = sc[13];
        long long cnt[11];
        get(cnt, sizeof cnt);
        reg_pairs = cnt[0];
        irr_pairs = cnt[1];
        reg_calls = cnt[2];
        irr_steps = cnt[3];
        reg_steps = cnt[4];
        blocks = cnt[5];
        overflows = cnt[6];
        learned_steps = cnt[7];
        fallbacks = cnt[8];
        nnb_dev_n = cnt[9];
        log_seen = cnt[10];
        B.assign(n, Body{});
        for (Body &b : B) {
            get(&b.m, sizeof b.m);
            get(b.x0, sizeof b.x0);
            get(b.v0, sizeof b.v0);
            get(b.a, sizeof b.a);
            get(b.j, sizeof b.j);
            get(b.fi, sizeof b.fi);
            get(b.fid, sizeof b.fid);
            get(b.fi2, sizeof b.fi2);
            get(b.fi3, sizeof b.fi3);
            get(b.fr, sizeof b.fr);
            get(b.frd, sizeof b.frd);
            get(b.fr2, sizeof b.fr2);
            get(b.fr3, sizeof b.fr3);
            double tt[5];
            get(tt, sizeof tt);
            b.t0 = tt[0];
            b.dt = tt[1];
            b.t0r = tt[2];
            b.dtr = tt[3];
            b.rs = tt[4];
            get(b.feat, sizeof b.feat);
            int32_t fl[3];
            get(fl, sizeof fl);
            b.has_feat = fl[0] != 0;
            b.learned = fl[1] != 0;
            if (fl[2] < 0 || fl[2] > o.lmax) fail("bad list length");
            b.list.resize(fl[2]);
            get(b.list.data(), b.list.size() * sizeof(int));
            for (int jb : b.list)
                if (jb < 0 || jb >= n) fail("bad list member");
        }
        fclose(f);
        rpos.assign(n, -1);
        rk.assign(n, -1);
        due.clear();
        for (int q = 0; q < n; ++q) due[B[q].t0 + B[q].dt].push_back(q);
        std::vector<int> all(n);
        for (int q = 0; q < n; ++q) all[q] = q;
        jpset(all);
        if (o.irr_dev) {
            std::vector<std::vector<int>> L(n);
            for (int q = 0; q < n; ++q) L[q] = B[q].list;
            irr_lists(all, L);
        }
        return e0;
    }

    // conf.3 snapshot of the synchronised state at t (output.F:488-535):
    // records (NTOT, MODEL, NRUN, NK) and (AS(1..20), BODYS, RHOS, XNS, XS,
    // VS, PHI = -PHIDBL, NAME), REAL*4 and INTEGER*4, each framed by 4-byte
    // byte counts as Fortran unformatted sequential files are; file name
    // conf.3_ + string_left(t, dtadj).  N-body units: RBAR = ZMBAR = TSCALE
    // = VSTAR = 1; RTIDE, TIDAL(4), CMAX and DMIN1 (no tidal field, no
    // close-encounter bookkeeping here) are 0.
    void write_conf3(double tt) const {
        int ip = 0;
        double dtmp = o.dtadj;
        while (dtmp - std::trunc(dtmp) != 0.0 && ip < 40) {
            ++ip;
            dtmp *= 10.0;
        }
        char stamp[64];
        snprintf(stamp, sizeof stamp, "%.*f", ip, tt);
        std::string st(stamp);
        if (st.size() > 20) st = std::string(20, '*');
        const std::string path = o.conf3_dir + "/conf.3_" + st;
        const long long rec2 = 4LL * (20 + 11LL * n);
        if (rec2 >= (1LL << 31)) {
            fprintf(stderr, "conf.3: NTOT too large for one record\n");
            exit(1);
        }
        const double rc = std::sqrt(rc2);
        int nc = 0;
        double rhom = 0.0;
        for (int q = 0; q < n; ++q) {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double e = B[q].x0[c] - rdens[c];
                r2 += e * e;
            }
            if (r2 < rc2) ++nc;
            if (q < (int)rho_b.size()) rhom = std::max(rhom, rho_b[q]);
        }
        double rsmin = 1e30;
        for (int q = 0; q < n; ++q) rsmin = std::min(rsmin, B[q].rs);
        const float as[20] = {(float)tt, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f,
                              (float)rdens[0], (float)rdens[1], (float)rdens[2],
                              (float)(tt / tcr), 1.0f, 1.0f, (float)rc,
                              (float)nc, (float)vc, (float)rhom, 0.0f,
                              (float)rscale, (float)rsmin, 0.0f};
        std::vector<float> fl;
        fl.reserve(20 + 10 * (size_t)n);
        fl.insert(fl.end(), as, as + 20);
        for (int q = 0; q < n; ++q) fl.push_back((float)B[q].m);
        for (int q = 0; q < n; ++q)
            fl.push_back(q < (int)rho_b.size() ? (float)rho_b[q] : 0.0f);
        for (int q = 0; q < n; ++q)
            fl.push_back(q < (int)xn_b.size() ? (float)xn_b[q] : 0.0f);
        for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) fl.push_back((float)B[q].x0[c]);
        for (int q = 0; q < n; ++q)
            for (int c = 0; c < 3; ++c) fl.push_back((float)B[q].v0[c]);
        for (int q = 0; q < n; ++q)
            fl.push_back(q < (int)phidbl.size() ? (float)(-phidbl[q]) : 0.0f);
        std::vector<int32_t> name(n);
        for (int q = 0; q < n; ++q) name[q] = q + 1;
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) {
            perror(path.c_str());
            exit(1);
        }
        const int32_t r1[4] = {n, 1, 1, 20};
        const int32_t l1 = 16, l2 = (int32_t)rec2;
        bool ok = fwrite(&l1, 4, 1, f) == 1 && fwrite(r1, 4, 4, f) == 4 &&
                  fwrite(&l1, 4, 1, f) == 1 && fwrite(&l2, 4, 1, f) == 1 &&
                  fwrite(fl.data(), 4, fl.size(), f) == fl.size() &&
                  fwrite(name.data(), 4, name.size(), f) == name.size() &&
                  fwrite(&l2, 4, 1, f) == 1;
        ok = (fclose(f) == 0) && ok;
        if (!ok) {
            fprintf(stderr, "conf.3: write to %s failed\n", path.c_str());
            exit(1);
        }
    }

    // One block step to the earliest due time.
    void block() {
        auto it = due.begin();
        const double tn = it->first;
        std::vector<int> A;
        A.swap(it->second);
        due.erase(it);
        std::sort(A.begin(), A.end());
        std::vector<int> R;
        for (int q : A)
            if (B[q].t0r + B[q].dtr == tn) R.push_back(q);
        blocks++;
        const int na = (int)A.size();
        // Predicted position of every active body at tn.
        std::vector<double> XP(3 * (size_t)na), VP(3 * (size_t)na);
#pragma omp parallel for schedule(static) if (na >= 4096)
        for (int k = 0; k < na; ++k)
            predict_cubic(B[A[k]], tn, &XP[3 * (size_t)k], &VP[3 * (size_t)k]);
        // Irregular force with the current lists.
        std::vector<double> FI(3 * (size_t)na), FID(3 * (size_t)na);
        double w0 = wtime();
        // Device for large blocks; small ones stay on the host, where a call
        // costs less than a device round trip (the device state is staged
        // either way).
        if (o.irr_dev && na >= o.irr_min) {
            irr_dev(tn, A, FI.data(), FID.data());
        } else {
#pragma omp parallel for schedule(dynamic, 16) if (na >= 64)
            for (int k = 0; k < na; ++k)
                irr_force(B, B[A[k]].list, tn, &XP[3 * (size_t)k],
                          &VP[3 * (size_t)k], &FI[3 * (size_t)k],
                          &FID[3 * (size_t)k]);
        }
        t_irr += wtime() - w0;
        for (int k = 0; k < na; ++k) irr_pairs += (long long)B[A[k]].list.size();
        irr_steps += na;

        // Regular step for the bodies in R (engine call).
        for (int k = 0; k < na; ++k) rpos[A[k]] = k;
        std::vector<double> FRN, FRDN;
        std::vector<std::vector<int>> LN;
        std::vector<char> OVF;
        if (!R.empty()) {
            send_j(tn);
            std::vector<double> X(3 * R.size()), V(3 * R.size());
            for (size_t k = 0; k < R.size(); ++k) {
                const int a = rpos[R[k]];
                for (int c = 0; c < 3; ++c) {
                    X[3 * k + c] = XP[3 * (size_t)a + c];
                    V[3 * k + c] = VP[3 * (size_t)a + c];
                }
            }
            regular(R, X, V, FRN, FRDN, LN, &OVF);
            reg_steps += (long long)R.size();
        }
        for (size_t k = 0; k < R.size(); ++k) rk[R[k]] = (int)k;
        // Irregular force of the new lists at tn (read-only pass: the
        // corrector below modifies bodies that are neighbours of others).
        // Also the gain/loss corrections of the 2nd/3rd force derivatives:
        // a body that leaves the list carries its pair derivatives from the
        // irregular to the regular polynomial, one that enters the list the
        // other way (regcor_gpu.F's neighbour-change corrections).
        std::vector<double> FIL(3 * R.size()), FIDL(3 * R.size());
        std::vector<double> DF2(3 * R.size(), 0.0), DF3(3 * R.size(), 0.0);
        // Radius-rule inputs: min over members of RIJ + RSDOT*DTR for short
        // lists (regcor_gpu.F:236-247) and the training target closed by
        // the new list.
        std::vector<double> M3(R.size(), 1e300);
        std::vector<float> TG(R.size(), 0.0f);
        std::vector<int> TC(R.size(), -1);
        const bool logging = !o.log_path.empty();
        w0 = wtime();
        const int nr = (int)R.size();
        const bool dev_r = o.irr_dev && nr >= o.irr_min;
        if (o.irr_dev && nr > 0) irr_lists(R, LN);
        if (dev_r) irr_dev(tn, R, FIL.data(), FIDL.data());
#pragma omp parallel for schedule(dynamic, 16) if (nr >= 64)
        for (int q = 0; q < nr; ++q) {
            const int a = rpos[R[q]];
            const double *xi = &XP[3 * (size_t)a], *vi = &VP[3 * (size_t)a];
            if (!dev_r)
                irr_force(B, LN[q], tn, xi, vi, &FIL[3 * (size_t)q],
                          &FIDL[3 * (size_t)q]);
            const Body &bi = B[R[q]];
            if (LN[q].size() <= 3) {
                for (int jb : LN[q]) {
                    double xj[3], vj[3], r2 = 0.0, rv = 0.0;
                    predict_cubic(B[jb], tn, xj, vj);
                    for (int c = 0; c < 3; ++c) {
                        r2 += (xi[c] - xj[c]) * (xi[c] - xj[c]);
                        rv += (xi[c] - xj[c]) * (vi[c] - vj[c]);
                    }
                    const double rij = std::sqrt(r2);
                    M3[q] = std::min(M3[q], rij + rv / rij * bi.dtr);
                }
            }
            if (logging && bi.has_feat)
                TC[q] = target(LN[q], tn, xi, vi, bi.dtr, bi.rs, &TG[q]);
            // Total acceleration and jerk of i at tn.
            double ai[3], ji[3];
            for (int c = 0; c < 3; ++c) {
                ai[c] = FIL[3 * (size_t)q + c] + FRN[3 * (size_t)q + c];
                ji[c] = FIDL[3 * (size_t)q + c] + FRDN[3 * (size_t)q + c];
            }
            const std::vector<int> &Lo = B[R[q]].list, &Ln = LN[q];
            size_t u = 0, w = 0;
            while (u < Lo.size() || w < Ln.size()) {
                int jb, sign;
                if (w == Ln.size() || (u < Lo.size() && Lo[u] < Ln[w])) {
                    jb = Lo[u++];
                    sign = +1;            // lost: now a regular pair
                } else if (u == Lo.size() || Ln[w] < Lo[u]) {
                    jb = Ln[w++];
                    sign = -1;            // gained: now an irregular pair
                } else {
                    ++u;
                    ++w;
                    continue;
                }
                const Body &bj = B[jb];
                const double sj = tn - bj.t0;
                double xj[3], vj[3], r_[3], v_[3], a_[3], j_[3], f2[3], f3[3];
                predict_cubic(bj, tn, xj, vj);
                for (int c = 0; c < 3; ++c) {
                    const double F2j = bj.fi2[c] + bj.fr2[c] + bj.fr3[c] * (bj.t0 - bj.t0r);
                    r_[c] = xj[c] - xi[c];
                    v_[c] = vj[c] - vi[c];
                    a_[c] = bj.a[c] + sj * (bj.j[c] + 0.5 * sj * F2j) - ai[c];
                    j_[c] = bj.j[c] + sj * F2j - ji[c];
                }
                pair_d23(bj.m, r_, v_, a_, j_, f2, f3);
                for (int c = 0; c < 3; ++c) {
                    DF2[3 * (size_t)q + c] += sign * f2[c];
                    DF3[3 * (size_t)q + c] += sign * f3[c];
                }
            }
        }
        for (size_t q = 0; q < R.size(); ++q) {
            irr_pairs += (long long)LN[q].size();
            nnb_dev_sum += std::fabs((double)LN[q].size() - o.nnbopt) / o.nnbopt;
            if (TC[q] >= 0) log_add(B[R[q]].feat, TG[q], TC[q]);
        }
        nnb_dev_n += (long long)R.size();
        std::vector<char> WANT(R.size(), 0), FELL(R.size(), 0);


#pragma omp parallel for schedule(dynamic, 16) if (na >= 64)
        for (int k = 0; k < na; ++k) {
            Body &b = B[A[k]];
            const double h = tn - b.t0;
            const double *fin = &FI[3 * (size_t)k];
            const double *fidn = &FID[3 * (size_t)k];
            double fi2[3], fi3[3];
            hermite23(b.fi, b.fid, fin, fidn, h, fi2, fi3);
            const int r = rk[A[k]];
            int nnb0 = 0;
            double newfi[3], newfid[3];
            for (int c = 0; c < 3; ++c) {
                newfi[c] = fin[c];
                newfid[c] = fidn[c];
            }
            if (r >= 0) {
                // Regular force of the old list at tn, regular Hermite over
                // the regular interval, then switch to the new list.
                const double H = tn - b.t0r;
                const double *fiL2 = &FIL[3 * (size_t)r];
                const double *fidL2 = &FIDL[3 * (size_t)r];
                double frL[3], frdL[3], fr2[3], fr3[3];
                for (int c = 0; c < 3; ++c) {
                    frL[c] = FRN[3 * (size_t)r + c] + fiL2[c] - fin[c];
                    frdL[c] = FRDN[3 * (size_t)r + c] + fidL2[c] - fidn[c];
                }
                hermite23(b.fr, b.frd, frL, frdL, H, fr2, fr3);
                for (int c = 0; c < 3; ++c) {
                    b.fr2[c] = fr2[c];
                    b.fr3[c] = fr3[c];
                }
                // Corrector below uses the regular derivatives at t0.
                const double s0 = b.t0 - b.t0r;
                double F2[3], F3[3], xp[3], vp[3];
                for (int c = 0; c < 3; ++c) {
                    F2[c] = fi2[c] + b.fr2[c] + b.fr3[c] * s0;
                    F3[c] = fi3[c] + b.fr3[c];
                    xp[c] = XP[3 * (size_t)k + c];
                    vp[c] = VP[3 * (size_t)k + c];
                }
                const double h3 = h * h * h, h4 = h3 * h, h5 = h4 * h;
                for (int c = 0; c < 3; ++c) {
                    b.x0[c] = xp[c] + F2[c] * h4 / 24.0 + F3[c] * h5 / 120.0;
                    b.v0[c] = vp[c] + F2[c] * h3 / 6.0 + F3[c] * h4 / 24.0;
                }
                for (int c = 0; c < 3; ++c) {
                    // Regular derivatives at tn, moved to the new list.
                    b.fr2[c] = fr2[c] + fr3[c] * H + DF2[3 * (size_t)r + c];
                    b.fr3[c] = fr3[c] + DF3[3 * (size_t)r + c];
                    b.fr[c] = FRN[3 * (size_t)r + c];
                    b.frd[c] = FRDN[3 * (size_t)r + c];
                    newfi[c] = fiL2[c];
                    newfid[c] = fidL2[c];
                }
                nnb0 = (int)b.list.size();
                b.list = LN[r];
                b.t0r = tn;
            } else {
                const double s0 = b.t0 - b.t0r;
                double F2[3], F3[3];
                for (int c = 0; c < 3; ++c) {
                    F2[c] = fi2[c] + b.fr2[c] + b.fr3[c] * s0;
                    F3[c] = fi3[c] + b.fr3[c];
                }
                const double h3 = h * h * h, h4 = h3 * h, h5 = h4 * h;
                for (int c = 0; c < 3; ++c) {
                    b.x0[c] = XP[3 * (size_t)k + c] + F2[c] * h4 / 24.0 + F3[c] * h5 / 120.0;
                    b.v0[c] = VP[3 * (size_t)k + c] + F2[c] * h3 / 6.0 + F3[c] * h4 / 24.0;
                }
            }
            // State at tn (irregular derivatives moved to the new list).
            const double sr = tn - b.t0r;
            for (int c = 0; c < 3; ++c) {
                const double d2 = (r >= 0) ? DF2[3 * (size_t)r + c] : 0.0;
                const double d3 = (r >= 0) ? DF3[3 * (size_t)r + c] : 0.0;
                b.fi[c] = newfi[c];
                b.fid[c] = newfid[c];
                b.fi2[c] = fi2[c] + fi3[c] * h - d2;
                b.fi3[c] = fi3[c] - d3;
                const double frt = b.fr[c] + sr * (b.frd[c] + sr * (0.5 * b.fr2[c] + sr * b.fr3[c] / 6.0));
                const double frdt = b.frd[c] + sr * (b.fr2[c] + 0.5 * sr * b.fr3[c]);
                b.a[c] = b.fi[c] + frt;
                b.j[c] = b.fid[c] + frdt;
            }
            b.t0 = tn;
            const double dtn = new_dt(b, tn, h, std::min(b.dtr, o.dtadj));
            if (r >= 0) {
                // New neighbour radius (rule above; the learned radii are
                // filled in after this loop), then the new regular step.
                const int nnb = (int)b.list.size();
                const double rsl = b.rs, dtr_done = b.dtr;
                if (o.rs_rule == 1) {
                    b.rs = simple_rs(rsl, nnb);
                } else {
                    const double *xi = b.x0, *vi = b.v0;
                    b.rs = ac_rs(rsl, nnb, nnb0, xi, vi, dtr_done, M3[r]);
                }
                b.dtr = new_dtr(b, tn, dtr_done, dtn);
                features(b, rsl, b.dtr, b.feat);
                b.has_feat = true;
                if (o.rs_rule == 2) {
                    const bool bad = OVF[r] || nnb < nnbmin || nnb > o.nnbmax;
                    FELL[r] = b.learned && bad;
                    WANT[r] = !FELL[r];
                }
                b.learned = false;
            }
            // Irregular step never beyond the next regular time.
            double d = dtn;
            while (tn + d > b.t0r + b.dtr && d > o.dtmin) d *= 0.5;
            b.dt = d;
        }
        t_irr += wtime() - w0;
        if (o.rs_rule == 2 && !R.empty()) learned_radii(R, WANT, FELL);
        for (int q : A) due[B[q].t0 + B[q].dt].push_back(q);
        for (int q : A) rpos[q] = -1;
        for (int q : R) rk[q] = -1;
        jpset(A);
        t = tn;
    }

    // Learned radii for the regular bodies flagged in want (one batched
    // inference); a non-finite prediction keeps the rule's radius.
    void learned_radii(const std::vector<int> &R, const std::vector<char> &want,
                       const std::vector<char> &fell) {
        std::vector<int> idx;
        for (size_t q = 0; q < R.size(); ++q) {
            if (fell[q]) fallbacks++;
            if (want[q]) idx.push_back(R[q]);
        }
        if (idx.empty()) return;
        const int k = (int)idx.size();
        std::vector<float> F((size_t)k * kNF), Y(k);
        for (int u = 0; u < k; ++u)
            std::copy(B[idx[u]].feat, B[idx[u]].feat + kNF, &F[(size_t)u * kNF]);
        double w0 = wtime();
        gpunb_mlp_predict(k, F.data(), Y.data());
        abi_check();
        t_mlp += wtime() - w0;
        for (int u = 0; u < k; ++u) {
            const double rs = std::pow(10.0, (double)Y[u]);
            if (!std::isfinite(rs) || rs <= 0.0) continue;
            B[idx[u]].rs = rs;
            B[idx[u]].learned = true;
            learned_steps++;
        }
    }
};

void parse_args(int argc, char **argv, Opts &o, std::string &in) {
    if (argc < 3) {
        fprintf(stderr,
                "usage: %s in.bin t_end [--etai x] [--etar x] [--dtadj x] "
                "[--nnbopt k] [--lmax k] [--nnbmax k] [--send] [--prec p] "
                "[--rs0 x] [--rs ac|simple|mlp] [--mlp weights] "
                "[--nnbmin x] [--log-train file] [--log-max n] "
                "[--log-seed s] [--progress seconds] [--irr device|host] "
                "[--irr-min k] [--conf3 dir] [--dump file] [--restart file]\n",
                argv[0]);
        exit(1);
    }
    in = argv[1];
    o.t_end = atof(argv[2]);
    for (int a = 3; a < argc; ++a) {
        std::string s = argv[a];
        auto nxt = [&]() { return a + 1 < argc ? argv[++a] : (char *)"0"; };
        if (s == "--etai") o.etai = atof(nxt());
        else if (s == "--etar") o.etar = atof(nxt());
        else if (s == "--dtadj") o.dtadj = atof(nxt());
        else if (s == "--nnbopt") o.nnbopt = atoi(nxt());
        else if (s == "--lmax") o.lmax = atoi(nxt());
        else if (s == "--nnbmax") o.nnbmax = atoi(nxt());
        else if (s == "--send") o.use_predict = 0;
        else if (s == "--prec") o.prec = atoi(nxt());
        else if (s == "--rs0") o.rs0 = atof(nxt());
        else if (s == "--rs") {
            const std::string r = nxt();
            if (r == "ac") o.rs_rule = 0;
            else if (r == "simple") o.rs_rule = 1;
            else if (r == "mlp") o.rs_rule = 2;
            else {
                fprintf(stderr, "unknown --rs %s\n", r.c_str());
                exit(1);
            }
        }
        else if (s == "--mlp") o.mlp_path = nxt();
        else if (s == "--nnbmin") o.nnbmin = atof(nxt());
        else if (s == "--log-train") o.log_path = nxt();
        else if (s == "--conf3") o.conf3_dir = nxt();
        else if (s == "--dump") o.dump_path = nxt();
        else if (s == "--restart") o.restart_path = nxt();
        else if (s == "--log-max") o.log_max = atoll(nxt());
        else if (s == "--log-seed") o.log_seed = (unsigned)atol(nxt());
        else if (s == "--progress") o.progress = atof(nxt());
        else if (s == "--irr-min") o.irr_min = atoi(nxt());
        else if (s == "--irr") {
            const std::string r = nxt();
            if (r == "device") o.irr_dev = 1;
            else if (r == "host") o.irr_dev = 0;
            else {
                fprintf(stderr, "unknown --irr %s\n", r.c_str());
                exit(1);
            }
        }
        else {
            fprintf(stderr, "unknown option %s\n", s.c_str());
            exit(1);
        }
    }
}

}  // namespace

int main(int argc, char **argv) {
    Opts o;
    std::string in;
    parse_args(argc, argv, o, in);
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) {
        perror(in.c_str());
        return 1;
    }
    int n = 0;
    if (fread(&n, sizeof(int), 1, f) != 1 || n <= 0) {
        fprintf(stderr, "bad input\n");
        return 1;
    }
    std::vector<double> m(n), x(3 * (size_t)n), v(3 * (size_t)n);
    if (fread(m.data(), sizeof(double), n, f) != (size_t)n ||
        fread(x.data(), sizeof(double), 3 * (size_t)n, f) != 3 * (size_t)n ||
        fread(v.data(), sizeof(double), 3 * (size_t)n, f) != 3 * (size_t)n) {
        fprintf(stderr, "short input\n");
        return 1;
    }
    fclose(f);
    o.nnbmax = std::min(o.nnbmax, std::min(n / 2, o.lmax - 50));
    if (o.rs0 <= 0.0) {
        // RS0 such that a uniform sphere of the half-mass density holds
        // NNBOPT bodies: rs0^3 = NNBOPT/N (in units of the half-mass radius
        // ~0.8 of a virialised model).
        o.rs0 = 0.8 * std::cbrt((double)o.nnbopt / (double)n);
    }

    int one = 0;
    gpunb_set_abort_(&one);
    if (o.prec) gpunb_set_precision(o.prec);
    Run R;
    R.o = o;
    R.n = n;
    int nbmax = n + 10, irank = 0;
    gpunb_open_(&nbmax, &irank);
    R.abi_check();
    if (o.irr_dev) {
        int nm = n, lm = o.lmax;
        irr_simd_open_(&nm, &lm, &irank);
        R.abi_check();
    }
    if (o.rs_rule == 2) {
        if (o.mlp_path.empty()) {
            fprintf(stderr, "--rs mlp needs --mlp weights\n");
            return 1;
        }
        R.load_mlp();
    }
    if (!o.log_path.empty() && (!o.dump_path.empty() || !o.restart_path.empty())) {
        fprintf(stderr, "--log-train cannot be combined with checkpoints\n");
        return 1;
    }
    double w0 = wtime();
    double ek, ep, e0;
    if (!o.restart_path.empty()) {
        // Continue from a checkpoint (written at an energy check).
        e0 = R.restore(o.restart_path);
        const double e = R.energy(&ek, &ep);
        printf("{\"t\": %.6f, \"E\": %.15e, \"EK\": %.15e, \"EP\": %.15e, "
               "\"dE_E\": %.6e, \"restart\": true}\n",
               R.t, e, ek, ep, (e - e0) / std::fabs(e0));
    } else {
        R.init(m, x, v);
        e0 = R.energy(&ek, &ep);
        // Crossing time and velocity scale of scale.F, then the cluster
        // quantities of the radius rule.
        double mt = 0.0;
        for (int q = 0; q < n; ++q) mt += m[q];
        const double e2 = 2.0 * std::fabs(e0);
        R.tcr = std::pow(mt, 2.5) / std::pow(e2, 1.5);
        R.vc = std::sqrt(e2 / mt);
        R.znbmin = std::max(0.01 * o.nnbmax, 1.0);
        R.nnbmin = o.nnbmin > 0.0 ? o.nnbmin : R.znbmin;
        R.update_centre();
        if (!o.conf3_dir.empty()) R.write_conf3(0.0);
        printf("{\"t\": 0.0, \"E\": %.15e, \"EK\": %.15e, \"EP\": %.15e, "
               "\"dE_E\": 0.0}\n", e0, ek, ep);
    }
    fflush(stdout);
    double next = R.t + o.dtadj;
    double last_progress = wtime();
    while (R.t < o.t_end) {
        R.block();
        if (o.progress > 0.0 && wtime() - last_progress >= o.progress) {
            last_progress = wtime();
            fprintf(stderr, "progress t=%.6f blocks=%lld reg_steps=%lld wall=%.1f\n",
                    R.t, R.blocks, R.reg_steps, last_progress - w0);
            fflush(stderr);
        }
        if (R.t >= next) {
            const double e = R.energy(&ek, &ep);
            const double wall = wtime() - w0;
            long long nnbs = 0;
            for (int q = 0; q < n; ++q) nnbs += (long long)R.B[q].list.size();
            printf("{\"t\": %.6f, \"E\": %.15e, \"EK\": %.15e, \"EP\": %.15e, "
                   "\"dE_E\": %.6e, \"blocks\": %lld, \"irr_steps\": %lld, "
                   "\"reg_steps\": %lld, \"reg_calls\": %lld, \"reg_pairs\": %lld, "
                   "\"irr_pairs\": %lld, \"overflows\": %lld, \"nnb_mean\": %.2f, "
                   "\"nnb_dev\": %.4f, \"rs_rule\": \"%s\", \"irr\": \"%s\", "
                   "\"learned_steps\": %lld, \"fallbacks\": %lld, "
                   "\"log_seen\": %lld, \"rscale\": %.4f, \"rcore\": %.4f, "
                   "\"t_reg\": %.3f, \"t_irr\": %.3f, \"t_pred\": %.3f, "
                   "\"t_pot\": %.3f, \"t_mlp\": %.3f, \"t_jp\": %.3f, \"wall\": %.3f}\n",
                   R.t, e, ek, ep, (e - e0) / std::fabs(e0), R.blocks,
                   R.irr_steps, R.reg_steps, R.reg_calls, R.reg_pairs,
                   R.irr_pairs, R.overflows, (double)nnbs / n,
                   R.nnb_dev_n ? R.nnb_dev_sum / R.nnb_dev_n : 0.0,
                   o.rs_rule == 0 ? "ac" : o.rs_rule == 1 ? "simple" : "mlp",
                   o.irr_dev ? "device" : "host",
                   R.learned_steps, R.fallbacks, R.log_seen, R.rscale,
                   std::sqrt(R.rc2), R.t_reg, R.t_irr, R.t_pred, R.t_pot,
                   R.t_mlp, R.t_jp, wall);
            fflush(stdout);
            next += o.dtadj;
            R.update_centre();
            if (!o.conf3_dir.empty()) R.write_conf3(R.t);
            if (!o.dump_path.empty()) R.dump(o.dump_path, e0);
        }
    }
    R.write_log();
    if (o.irr_dev) irr_simd_close_(&irank);
    gpunb_close_();
    return 0;
}
