/*
 * fp32_model.c -- CPU model of the rounding of the regular-force pair
 * kernel's FP32 path, for the precision study in test_precision_model.py.
 *
 * TEST INFRASTRUCTURE ONLY.  It models the engine's arithmetic (not the
 * reference's) so that the error budget of each design choice can be shown
 * on the CPU against the FP64 oracle:
 *
 *   mode 0  pair differences x_j - x_i in FP64, rounded to FP32
 *   mode 1  tile-relative differences: j in the given order is cut into
 *           tiles of TILE, C_t = midpoint of the tile's bounding box (positions
 *           and velocities apart), dx = fl32(C_t - x_i) + fl32(x_j - C_t)
 *   mode 2  everything in FP64 except 1/r, rounded to FP32 (the floor any
 *           FP32 pair arithmetic has)
 *
 * Then, unless the pair is in the near band, FP32 r2 = |dx|^2 and
 * dx.dv with fused multiply-adds, 1/r = the correctly rounded FP32 value
 * moved by -1, 0 or +1 ulp (a deterministic hash of (i, j); the hardware
 * v_rsq_f32 is specified to 1 ulp), and the FP32 force terms accumulated in
 * FP32 and folded into FP64 every TILE j.  Near-band pairs -- min(r2, q) <
 * band * h2 * w_j with q = r2 + 2 dtr dx.dv, w_j = max(1, x^(2/3),
 * 1.4 x / K), x = m_j / mbar, band = K (times mbar for mass-weighted lists) --
 * and list neighbours are evaluated in FP64 as the oracle does; in mode 1
 * the threshold band * h2 is raised per tile to (e_t / 2)^2, e_t = the
 * largest |fl32(x_j - C_t)| of the tile, as the engine does.  band = 0
 * switches the band off (only neighbours and r2 == 0 leave the FP32 sums).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* j per tile of the engine's sorted view (kVTile). */
#define TILE 64

static uint32_t hash2(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

static float rsq_model(float r2, uint32_t i, uint32_t j) {
    float r = (float)(1.0 / sqrt((double)r2));
    switch (hash2(i, j) % 3u) {
    case 1: return nextafterf(r, INFINITY);
    case 2: return nextafterf(r, 0.0f);
    default: return r;
    }
}

static double sq3(double a, double b, double c) {
    double s = a * a;
    s = s + b * b;
    s = s + c * c;
    return s;
}

/* stats[0] = band pairs evaluated in FP64, stats[1] = neighbour pairs. */
void model_regf(int nj, const double *m, const double *xj, const double *vj,
                int ni, const double *h2, const double *dtr, const double *xi,
                const double *vi, int m_flag, int mode, double band,
                double K, double *acc, double *jrk, double *pot,
                long long *stats) {
    double mbar = 0.0;
    for (int j = 0; j < nj; j++) mbar += m[j];
    mbar = nj > 0 ? mbar / nj : 1.0;
    int ntile = (nj + TILE - 1) / TILE;
    double *cen = malloc(sizeof(double) * 7 * (size_t)(ntile > 0 ? ntile : 1));
    for (int t = 0; t < ntile; t++) {
        double lo[6], hi[6];
        for (int c = 0; c < 6; c++) {
            lo[c] = INFINITY;
            hi[c] = -INFINITY;
        }
        for (int j = TILE * t; j < TILE * t + TILE && j < nj; j++)
            for (int c = 0; c < 6; c++) {
                double v = c < 3 ? xj[3 * (size_t)j + c] : vj[3 * (size_t)j + c - 3];
                if (v < lo[c]) lo[c] = v;
                if (v > hi[c]) hi[c] = v;
            }
        double *C = cen + 7 * t;
        for (int c = 0; c < 6; c++) C[c] = 0.5 * lo[c] + 0.5 * hi[c];
        double e = 0.0;
        for (int j = TILE * t; j < TILE * t + TILE && j < nj; j++) {
            double o2 = 0.0;
            for (int c = 0; c < 3; c++) {
                const double o = (float)(xj[3 * (size_t)j + c] - C[c]);
                o2 += o * o;
            }
            if (sqrt(o2) > e) e = sqrt(o2);
        }
        C[6] = 0.25 * e * e;
    }
    long long s_band = 0, s_nb = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : s_band, s_nb)
    for (int i = 0; i < ni; i++) {
        const double *x0 = xi + 3 * (size_t)i, *v0 = vi + 3 * (size_t)i;
        float ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, ph = 0;
        double A[7] = {0, 0, 0, 0, 0, 0, 0};
        const double Bh = band * h2[i] * (m_flag ? mbar : 1.0);
        const float dtr2x = 2.0f * (float)dtr[i];
        for (int j = 0; j < nj; j++) {
            const double *X = xj + 3 * (size_t)j, *V = vj + 3 * (size_t)j;
            double d[6];
            for (int c = 0; c < 3; c++) {
                d[c] = X[c] - x0[c];
                d[3 + c] = V[c] - v0[c];
            }
            const double r2d = sq3(d[0], d[1], d[2]);
            const double r2pd = sq3(d[0] + dtr[i] * d[3], d[1] + dtr[i] * d[4],
                                    d[2] + dtr[i] * d[5]);
            const double mn = r2d < r2pd ? r2d : r2pd;
            const double H = m_flag ? m[j] * h2[i] : h2[i];
            int fp64 = 0;
            if (mn < H) {
                s_nb++;
                fp64 = -1;
            } else if (r2d == 0.0) {
                fp64 = -1;
            }
            float f[6];
            if (mode == 1) {
                const double *C = cen + 7 * (j / TILE);
                for (int c = 0; c < 3; c++) {
                    f[c] = (float)(C[c] - x0[c]) + (float)(X[c] - C[c]);
                    f[3 + c] = (float)(C[3 + c] - v0[c]) + (float)(V[c] - C[3 + c]);
                }
            } else {
                for (int c = 0; c < 6; c++) f[c] = (float)d[c];
            }
            const float r2 = fmaf(f[2], f[2], fmaf(f[1], f[1], f[0] * f[0]));
            const float rv = fmaf(f[2], f[5], fmaf(f[1], f[4], f[0] * f[3]));
            if (fp64 == 0 && band > 0.0) {
                const double x = m[j] / mbar;
                double w = 1.0;
                if (x > 1.0) {
                    w = cbrt(x * x);
                    if (1.4 * x / K > w) w = 1.4 * x / K;
                }
                const float q = fmaf(dtr2x, rv, r2);
                const float s = r2 < q ? r2 : q;
                double T = Bh;
                if (mode == 1 && cen[7 * (j / TILE) + 6] > T) T = cen[7 * (j / TILE) + 6];
                if ((double)s < T * w) fp64 = 1;
            }
            if (fp64 < 0) {
                /* list neighbour or r2 == 0: no force term */
            } else if (fp64 > 0 || mode == 2) {
                double rinv = 1.0 / sqrt(r2d);
                if (mode == 2 && fp64 == 0)
                    rinv = (double)rsq_model((float)r2d, (uint32_t)i, (uint32_t)j);
                else
                    s_band++;
                const double rinv2 = rinv * rinv;
                const double mr = m[j] * rinv, mr3 = mr * rinv2;
                const double rvd = d[0] * d[3] + d[1] * d[4] + d[2] * d[5];
                const double al = -3.0 * rvd * rinv2;
                A[6] += mr;
                for (int c = 0; c < 3; c++) {
                    A[c] += mr3 * d[c];
                    A[3 + c] += mr3 * (d[3 + c] + al * d[c]);
                }
            } else {
                const float rinv = rsq_model(r2, (uint32_t)i, (uint32_t)j);
                const float rinv2 = rinv * rinv;
                const float mr = (float)m[j] * rinv;
                const float mr3 = mr * rinv2;
                const float al = (-3.0f * rv) * rinv2;
                ph += mr;
                ax = fmaf(mr3, f[0], ax);
                ay = fmaf(mr3, f[1], ay);
                az = fmaf(mr3, f[2], az);
                jx = fmaf(mr3, fmaf(al, f[0], f[3]), jx);
                jy = fmaf(mr3, fmaf(al, f[1], f[4]), jy);
                jz = fmaf(mr3, fmaf(al, f[2], f[5]), jz);
            }
            if ((j % TILE) == TILE - 1 || j == nj - 1) {
                A[0] += ax; A[1] += ay; A[2] += az;
                A[3] += jx; A[4] += jy; A[5] += jz; A[6] += ph;
                ax = ay = az = jx = jy = jz = ph = 0;
            }
        }
        for (int c = 0; c < 3; c++) {
            acc[3 * (size_t)i + c] = A[c];
            jrk[3 * (size_t)i + c] = A[3 + c];
        }
        pot[i] = A[6];
    }
    free(cen);
    stats[0] = s_band;
    stats[1] = s_nb;
}

/*
 * model_regf_tiles -- the engine's mixed kernel as shipped: j in tile order
 * (the caller passes them light then heavy, each along the engine's curve,
 * cut into tiles of TILE), i in wavefronts of 64 (the caller passes them in
 * curve order; a partial last wavefront has inactive lanes).  Per (wavefront, tile) every
 * lane evaluates the far test in FP32 as the kernel does
 *   D = fl32(C_t - x_i), R = |D| (1 - 1e-5) - e_t,
 *   A = |dtr| (1 + 2u) (|V| + ev_t) (1 + 1e-5),
 *   far: R > sqrt(A^2 + max(B, eb_t) wmax_t (1 + 1e-5)) (1 + 1e-5) + A,
 *   B = fl32(K h2 (<m>) (1 + 8u)) (1 + 2u);
 * a tile that is near for any lane is evaluated in FP64 by every lane (the
 * oracle's arithmetic, neighbours and r2 == 0 excluded), a far tile in FP32
 * with tile-relative differences, its sums folded into FP64 at its end.
 * stats[0] = near (wavefront, tile) pairs, stats[1] = all such pairs.
 */
void model_regf_tiles(int nj, const double *m, const double *xj,
                      const double *vj, int ni, const double *h2,
                      const double *dtr, const double *xi, const double *vi,
                      int m_flag, double K, double *acc, double *jrk,
                      double *pot, long long *stats) {
    const float u32 = 5.9604645e-08f;
    double mbar = 0.0;
    for (int j = 0; j < nj; j++) mbar += m[j];
    mbar = nj > 0 ? mbar / nj : 1.0;
    const int ntile = (nj + TILE - 1) / TILE;
    /* per tile: C[6], e, ev, wmax, eb (floats kept as doubles) */
    double *T = malloc(sizeof(double) * 10 * (size_t)(ntile > 0 ? ntile : 1));
    float *o = malloc(sizeof(float) * 6 * (size_t)(nj > 0 ? nj : 1));
    for (int t = 0; t < ntile; t++) {
        double lo[6], hi[6];
        for (int c = 0; c < 6; c++) {
            lo[c] = INFINITY;
            hi[c] = -INFINITY;
        }
        for (int j = TILE * t; j < TILE * t + TILE && j < nj; j++)
            for (int c = 0; c < 6; c++) {
                double v = c < 3 ? xj[3 * (size_t)j + c] : vj[3 * (size_t)j + c - 3];
                if (v < lo[c]) lo[c] = v;
                if (v > hi[c]) hi[c] = v;
            }
        double *C = T + 10 * t;
        for (int c = 0; c < 6; c++) C[c] = 0.5 * lo[c] + 0.5 * hi[c];
        double e = 0.0, ev = 0.0, wm = 1.0;
        for (int j = TILE * t; j < TILE * t + TILE && j < nj; j++) {
            double a = 0.0, b = 0.0;
            for (int c = 0; c < 3; c++) {
                o[6 * (size_t)j + c] = (float)(xj[3 * (size_t)j + c] - C[c]);
                o[6 * (size_t)j + 3 + c] = (float)(vj[3 * (size_t)j + c] - C[3 + c]);
                a += (double)o[6 * (size_t)j + c] * o[6 * (size_t)j + c];
                b += (double)o[6 * (size_t)j + 3 + c] * o[6 * (size_t)j + 3 + c];
            }
            a = sqrt(a) * (1.0 + 1e-12);
            b = sqrt(b) * (1.0 + 1e-12);
            if (a > e) e = a;
            if (b > ev) ev = b;
            double w = 1.0;
            const double x = m[j] / mbar;
            if (mbar > 0.0 && m[j] > mbar) {
                w = fmax(fmax(1.0, cbrt(x * x)), 1.4 * x / K) * (1.0 + 1e-12);
                float wf = (float)w;
                if ((double)wf < w) wf = nextafterf(wf, INFINITY);
                w = wf;
            }
            if (w > wm) wm = w;
        }
        float ef = (float)e, evf = (float)ev;
        if ((double)ef < e) ef = nextafterf(ef, INFINITY);
        if ((double)evf < ev) evf = nextafterf(evf, INFINITY);
        C[6] = ef;
        C[7] = evf;
        C[8] = (float)wm;
        double eb = 0.25 * (double)ef * ef * (1.0 + 1e-12);
        float ebf = (float)eb;
        if ((double)ebf < eb) ebf = nextafterf(ebf, INFINITY);
        C[9] = ebf;
    }
    const int nw = (ni + 63) / 64;
    long long s_near = 0, s_all = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : s_near, s_all)
    for (int w = 0; w < nw; w++) {
        const int i0 = 64 * w, nl = ni - i0 < 64 ? ni - i0 : 64;
        double A[64][7];
        float Bl[64], adtr[64];
        for (int l = 0; l < nl; l++) {
            for (int c = 0; c < 7; c++) A[l][c] = 0.0;
            const int i = i0 + l;
            const float d32 = (float)dtr[i];
            adtr[l] = fabsf(d32) * (1.0f + 2.0f * u32);
            Bl[l] = 0.0f;
            if (h2[i] > 0.0) {
                const double bb = K * h2[i] * (m_flag ? mbar : 1.0) *
                                  (1.0 + 8.0 * (double)u32);
                Bl[l] = (float)bb * (1.0f + 2.0f * u32);
            }
        }
        for (int t = 0; t < ntile; t++) {
            const double *C = T + 10 * t;
            int near = 0;
            for (int l = 0; l < nl && !near; l++) {
                const int i = i0 + l;
                const float Dx = (float)(C[0] - xi[3 * (size_t)i + 0]);
                const float Dy = (float)(C[1] - xi[3 * (size_t)i + 1]);
                const float Dz = (float)(C[2] - xi[3 * (size_t)i + 2]);
                const float Vx = (float)(C[3] - vi[3 * (size_t)i + 0]);
                const float Vy = (float)(C[4] - vi[3 * (size_t)i + 1]);
                const float Vz = (float)(C[5] - vi[3 * (size_t)i + 2]);
                const float Bt = Bl[l] > (float)C[9] ? Bl[l] : (float)C[9];
                const float d = sqrtf(fmaf(Dz, Dz, fmaf(Dy, Dy, Dx * Dx)));
                const float dvn = sqrtf(fmaf(Vz, Vz, fmaf(Vy, Vy, Vx * Vx)));
                const float R = fmaf(d, 1.0f - 1e-5f, -(float)C[6]);
                const float Aa = adtr[l] * (dvn + (float)C[7]) * (1.0f + 1e-5f);
                const float Tm = Bt * (float)C[8] * (1.0f + 1e-5f);
                const int far =
                    R > fmaf(sqrtf(fmaf(Aa, Aa, Tm)), 1.0f + 1e-5f, Aa);
                if (!far) near = 1;
            }
            s_all++;
            s_near += near;
            const int j1 = TILE * t + TILE < nj ? TILE * t + TILE : nj;
            for (int l = 0; l < nl; l++) {
                const int i = i0 + l;
                const double *x0 = xi + 3 * (size_t)i, *v0 = vi + 3 * (size_t)i;
                float ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, ph = 0;
                float F[6];
                for (int c = 0; c < 3; c++) {
                    F[c] = (float)(C[c] - x0[c]);
                    F[3 + c] = (float)(C[3 + c] - v0[c]);
                }
                for (int j = TILE * t; j < j1; j++) {
                    if (near) {
                        const double *X = xj + 3 * (size_t)j, *V = vj + 3 * (size_t)j;
                        double dd[6];
                        for (int c = 0; c < 3; c++) {
                            dd[c] = X[c] - x0[c];
                            dd[3 + c] = V[c] - v0[c];
                        }
                        const double r2d = sq3(dd[0], dd[1], dd[2]);
                        const double r2pd = sq3(dd[0] + dtr[i] * dd[3],
                                                dd[1] + dtr[i] * dd[4],
                                                dd[2] + dtr[i] * dd[5]);
                        const double mn = r2d < r2pd ? r2d : r2pd;
                        const double H = m_flag ? m[j] * h2[i] : h2[i];
                        if (mn < H || r2d == 0.0) continue;
                        const double rinv = 1.0 / sqrt(r2d);
                        const double rinv2 = rinv * rinv;
                        const double mr = m[j] * rinv, mr3 = mr * rinv2;
                        const double rvd = dd[0] * dd[3] + dd[1] * dd[4] + dd[2] * dd[5];
                        const double al = -3.0 * rvd * rinv2;
                        A[l][6] += mr;
                        for (int c = 0; c < 3; c++) {
                            A[l][c] += mr3 * dd[c];
                            A[l][3 + c] += mr3 * (dd[3 + c] + al * dd[c]);
                        }
                    } else {
                        const float *oj = o + 6 * (size_t)j;
                        float f[6];
                        for (int c = 0; c < 6; c++) f[c] = F[c] + oj[c];
                        const float r2 = fmaf(f[2], f[2], fmaf(f[1], f[1], f[0] * f[0]));
                        const float rv = fmaf(f[2], f[5], fmaf(f[1], f[4], f[0] * f[3]));
                        const float rinv = rsq_model(r2, (uint32_t)i, (uint32_t)j);
                        const float rinv2 = rinv * rinv;
                        const float mr = (float)m[j] * rinv;
                        const float mr3 = mr * rinv2;
                        const float al = (-3.0f * rv) * rinv2;
                        ph += mr;
                        ax = fmaf(mr3, f[0], ax);
                        ay = fmaf(mr3, f[1], ay);
                        az = fmaf(mr3, f[2], az);
                        jx = fmaf(mr3, fmaf(al, f[0], f[3]), jx);
                        jy = fmaf(mr3, fmaf(al, f[1], f[4]), jy);
                        jz = fmaf(mr3, fmaf(al, f[2], f[5]), jz);
                    }
                }
                A[l][0] += ax; A[l][1] += ay; A[l][2] += az;
                A[l][3] += jx; A[l][4] += jy; A[l][5] += jz; A[l][6] += ph;
            }
        }
        for (int l = 0; l < nl; l++) {
            const int i = i0 + l;
            for (int c = 0; c < 3; c++) {
                acc[3 * (size_t)i + c] = A[l][c];
                jrk[3 * (size_t)i + c] = A[l][3 + c];
            }
            pot[i] = A[l][6];
        }
    }
    free(T);
    free(o);
    stats[0] = s_near;
    stats[1] = s_all;
}
