/*
 * oracle.c -- CPU restatement of the NBODY6++GPU regular-force contract.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and never as the thing measured or shipped.  The product library does not
 * link it and has no CPU fallback.
 *
 * Every function is plain FP64 C compiled with -ffp-contract=off, so each
 * expression is evaluated exactly as written (IEEE round-to-nearest, no fused
 * multiply-add).  The neighbour predicate below is therefore bit-defined; the
 * engine must reproduce its decisions exactly.
 *
 * What each function follows (reference tree paths):
 *   orc_regf     src/Main/gpunb.gpu.cu:149-241 (pair body, criterion, self
 *                exclusion), :752-794 (list layout and overflow codes),
 *                src/Main/reg.avx.cpp:239-328 (CPU twin: list order,
 *                -nnb on overflow), src/Main/regint.f:39-79 (FP64 form).
 *   orc_nearest  src/Main/nbint.F:29-47 (IMINR: strict <, first in list order).
 *   orc_pot      src/Main/gpupot.gpu.cu:32-57, src/Main/pot.avx.cpp:87-137.
 *   orc_predict  src/Main/xbpredall.f:18-26.
 *   orc_irr      src/Main/irr.avx.cpp:406-438 (irr_simd_firr) with the
 *                predictor polynomial of xbpredall.f:18-26, in FP64.
 *
 * How each is pinned (tests/test_oracle_*.py): closed forms, Newton's third
 * law, long-double brute force on small inputs, cubic-trajectory exactness,
 * and the reference's own AVX/SSE code compiled from its sources under
 * oracle/_ref (golden vectors in tests/golden/).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef __FP_FAST_FMA
#error "build the oracle with -ffp-contract=off"
#endif

/* Squared distance, association ((a*a + b*b) + c*c). */
static double sq3(double a, double b, double c) {
    double s = a * a;
    s = s + b * b;
    s = s + c * c;
    return s;
}

/*
 * Neighbour predicate for one pair (gpunb.gpu.cu:162-173 / :220, restated in
 * FP64 with strict <).  Returns 1 iff j is a neighbour of i.
 * Also returns r2 through *r2_out.
 */
int orc_pair_is_nb(const double xi[3], const double vi[3], double h2,
                   double dtr, const double xj[3], const double vj[3],
                   double mj, int m_flag, double *r2_out) {
    double dx = xj[0] - xi[0];
    double dy = xj[1] - xi[1];
    double dz = xj[2] - xi[2];
    double dvx = vj[0] - vi[0];
    double dvy = vj[1] - vi[1];
    double dvz = vj[2] - vi[2];
    double r2 = sq3(dx, dy, dz);
    double px = dx + dtr * dvx;
    double py = dy + dtr * dvy;
    double pz = dz + dtr * dvz;
    double r2p = sq3(px, py, pz);
    double rmin = r2 < r2p ? r2 : r2p;
    double thr = m_flag ? mj * h2 : h2;
    if (r2_out) *r2_out = r2;
    return rmin < thr;
}

/*
 * Regular force, neighbour lists and nearest neighbour for ni i-particles
 * against nj j-particles.  Layout and overflow coding of the ABI:
 * list[i*lmax] = nnb (0..nnbmax) or -nnb (> nnbmax); entries ascending.
 * nnid may be NULL.  Far-field acc/jrk/pot exclude neighbours and r2 == 0.
 * Pair terms follow gpunb.gpu.cu:170-193:
 *   rinv = 1/sqrt(r2), mr3 = m*rinv^3, alpha = -3 (dx.dv) rinv^2,
 *   acc += mr3*dx, jrk += mr3*(dv + alpha*dx), pot += m*rinv.
 */
void orc_regf(int nj, const double *m, const double *xj, const double *vj,
              int ni, const double *h2, const double *dtr, const double *xi,
              const double *vi, double *acc, double *jrk, double *pot,
              int lmax, int nnbmax, int *list, int m_flag, int *nnid) {
    for (int i = 0; i < ni; i++) {
        const double *x_i = xi + 3 * (size_t)i;
        const double *v_i = vi + 3 * (size_t)i;
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, ph = 0;
        int nnb = 0;
        int *row = list ? list + (size_t)i * lmax : NULL;
        double best = INFINITY;
        int bestj = -1;
        for (int j = 0; j < nj; j++) {
            const double *x_j = xj + 3 * (size_t)j;
            const double *v_j = vj + 3 * (size_t)j;
            double r2;
            int nb = orc_pair_is_nb(x_i, v_i, h2[i], dtr[i], x_j, v_j, m[j],
                                    m_flag, &r2);
            if (nb) {
                if (row && nnb < nnbmax) row[1 + nnb] = j;
                nnb++;
                if (r2 > 0.0 && r2 < best) {
                    best = r2;
                    bestj = j;
                }
                continue;
            }
            if (r2 == 0.0) continue;
            double dx = x_j[0] - x_i[0];
            double dy = x_j[1] - x_i[1];
            double dz = x_j[2] - x_i[2];
            double dvx = v_j[0] - v_i[0];
            double dvy = v_j[1] - v_i[1];
            double dvz = v_j[2] - v_i[2];
            double rinv = 1.0 / sqrt(r2);
            double rinv2 = rinv * rinv;
            double mr = m[j] * rinv;
            double mr3 = mr * rinv2;
            double rv = dx * dvx + dy * dvy + dz * dvz;
            double alpha = -3.0 * rv * rinv2;
            ph += mr;
            ax += mr3 * dx;
            ay += mr3 * dy;
            az += mr3 * dz;
            jx += mr3 * (dvx + alpha * dx);
            jy += mr3 * (dvy + alpha * dy);
            jz += mr3 * (dvz + alpha * dz);
        }
        acc[3 * (size_t)i + 0] = ax;
        acc[3 * (size_t)i + 1] = ay;
        acc[3 * (size_t)i + 2] = az;
        jrk[3 * (size_t)i + 0] = jx;
        jrk[3 * (size_t)i + 1] = jy;
        jrk[3 * (size_t)i + 2] = jz;
        pot[i] = ph;
        int over = nnb > nnbmax;
        if (row) row[0] = over ? -nnb : nnb;
        if (nnid) nnid[i] = over ? -1 : bestj;
    }
}

/*
 * Nearest neighbour over a given list (nbint.F:29-47): the first entry of
 * smallest FP64 r2 among entries with r2 > 0; -1 if none.
 */
int orc_nearest(const double *xi, int nnb, const int *idx, const double *xj) {
    double best = INFINITY;
    int bestj = -1;
    for (int k = 0; k < nnb; k++) {
        const double *p = xj + 3 * (size_t)idx[k];
        double r2 = sq3(p[0] - xi[0], p[1] - xi[1], p[2] - xi[2]);
        if (r2 > 0.0 && r2 < best) {
            best = r2;
            bestj = idx[k];
        }
    }
    return bestj;
}

/*
 * Potential (gpupot.gpu.cu:32-57): for k in [istart-1, istart-1+ni),
 * pot[k-istart+1] = sum over j in [0,n) with r2 > 0 of m_j / r_jk.
 */
void orc_pot(int istart, int ni, int n, const double *m, const double *x,
             double *pot) {
    for (int q = 0; q < ni; q++) {
        const double *p = x + 3 * (size_t)(istart - 1 + q);
        double s = 0.0;
        for (int j = 0; j < n; j++) {
            const double *o = x + 3 * (size_t)j;
            double r2 = sq3(o[0] - p[0], o[1] - p[1], o[2] - p[2]);
            if (r2 > 0.0) s += m[j] / sqrt(r2);
        }
        pot[q] = s;
    }
}

/*
 * Hermite predictor (xbpredall.f:18-26), F = a/2 and FDOT = jerk/6:
 *   x = ((fd*s + f)*s + v0)*s + x0,  v = (fd*(1.5s) + f)*(2s) + v0.
 */
void orc_predict(int n, double t, const double *t0, const double *x0,
                 const double *v0, const double *f, const double *fd,
                 double *x, double *v) {
    for (int j = 0; j < n; j++) {
        double s = t - t0[j];
        double s1 = 1.5 * s;
        double s2 = 2.0 * s;
        for (int k = 0; k < 3; k++) {
            size_t q = 3 * (size_t)j + k;
            x[q] = ((fd[q] * s + f[q]) * s + v0[q]) * s + x0[q];
            v[q] = (fd[q] * s1 + f[q]) * s2 + v0[q];
        }
    }
}

/*
 * Long-double brute force of the far-field sums for a given neighbour mask
 * (used by the tests as an error budget for the FP64 oracle itself).
 * mask[i*nj + j] != 0 marks pairs excluded from the sums.
 */
void orc_regf_ld(int nj, const double *m, const double *xj, const double *vj,
                 int ni, const double *xi, const double *vi,
                 const unsigned char *mask, double *acc, double *jrk,
                 double *pot) {
    for (int i = 0; i < ni; i++) {
        long double a[3] = {0, 0, 0}, jk[3] = {0, 0, 0}, ph = 0;
        for (int j = 0; j < nj; j++) {
            if (mask[(size_t)i * nj + j]) continue;
            long double d[3], dv[3], r2 = 0, rv = 0;
            for (int k = 0; k < 3; k++) {
                d[k] = (long double)xj[3 * (size_t)j + k] - xi[3 * (size_t)i + k];
                dv[k] = (long double)vj[3 * (size_t)j + k] - vi[3 * (size_t)i + k];
                r2 += d[k] * d[k];
                rv += d[k] * dv[k];
            }
            if (r2 == 0) continue;
            long double rinv = 1.0L / sqrtl(r2);
            long double mr3 = m[j] * rinv * rinv * rinv;
            ph += m[j] * rinv;
            for (int k = 0; k < 3; k++) {
                a[k] += mr3 * d[k];
                jk[k] += mr3 * (dv[k] - 3.0L * rv * rinv * rinv * d[k]);
            }
        }
        for (int k = 0; k < 3; k++) {
            acc[3 * (size_t)i + k] = (double)a[k];
            jrk[3 * (size_t)i + k] = (double)jk[k];
        }
        pot[i] = (double)ph;
    }
}

/*
 * Irregular force of ni particles (0-based indices idx[]) at time t over
 * their neighbour lists (irr.avx.cpp:406-438, restated in FP64).  Every
 * particle is predicted from its own (x0, v0, f = a/2, fd = jerk/6, t0) with
 * the xbpredall.f polynomial (orc_predict's association).  Lists: row i of
 * `list` (stride lstride) holds nnb then nnb 0-based indices.
 * acc/jrk per pair: rinv = 1/sqrt(r2), mr3 = m rinv^3,
 * alpha = -3 (dx.dv) rinv^2, acc += mr3 dx, jrk += mr3 (dv + alpha dx),
 * summed in list order; a pair with r2 == 0 is skipped.
 * nnbid = 1 + the member with the smallest r2 (ties: the smaller index),
 * 0 for an empty list (the reference's packed (r2, index) minimum).
 */
void orc_irr(int ni, const int *idx, double t, const double *t0,
             const double *x0, const double *v0, const double *f,
             const double *fd, const double *m, const int *list, int lstride,
             double *acc, double *jrk, int *nnbid) {
    for (int q = 0; q < ni; q++) {
        const int i = idx[q];
        double xi[3], vi[3];
        orc_predict(1, t, t0 + i, x0 + 3 * (size_t)i, v0 + 3 * (size_t)i,
                    f + 3 * (size_t)i, fd + 3 * (size_t)i, xi, vi);
        const int *row = list + (size_t)i * lstride;
        const int nnb = row[0];
        double a[3] = {0.0, 0.0, 0.0}, jk[3] = {0.0, 0.0, 0.0};
        double best = INFINITY;
        int bestj = -1;
        for (int k = 0; k < nnb; k++) {
            const int j = row[1 + k];
            double xj[3], vj[3];
            orc_predict(1, t, t0 + j, x0 + 3 * (size_t)j, v0 + 3 * (size_t)j,
                        f + 3 * (size_t)j, fd + 3 * (size_t)j, xj, vj);
            const double dx = xj[0] - xi[0], dy = xj[1] - xi[1],
                         dz = xj[2] - xi[2];
            const double dvx = vj[0] - vi[0], dvy = vj[1] - vi[1],
                         dvz = vj[2] - vi[2];
            const double r2 = sq3(dx, dy, dz);
            if (r2 < best || (r2 == best && j < bestj)) {
                best = r2;
                bestj = j;
            }
            if (r2 == 0.0) continue;
            const double rinv = 1.0 / sqrt(r2);
            const double rinv2 = rinv * rinv;
            const double mr3 = m[j] * rinv2 * rinv;
            double rv = dx * dvx;
            rv = rv + dy * dvy;
            rv = rv + dz * dvz;
            const double alpha = -3.0 * rv * rinv2;
            a[0] += mr3 * dx;
            a[1] += mr3 * dy;
            a[2] += mr3 * dz;
            jk[0] += mr3 * (dvx + alpha * dx);
            jk[1] += mr3 * (dvy + alpha * dy);
            jk[2] += mr3 * (dvz + alpha * dz);
        }
        for (int c = 0; c < 3; c++) {
            acc[3 * (size_t)q + c] = a[c];
            jrk[3 * (size_t)q + c] = jk[c];
        }
        nnbid[q] = bestj + 1;
    }
}
