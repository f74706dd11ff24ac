/*
 * regcor.c -- CPU restatement of NBODY6++GPU's regular corrector.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the device
 * corrector (gpunb_regcor_dev): only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and never as the thing measured
 * or shipped.  The product library does not link it.
 *
 * orc_regcor follows src/Main/regcor_gpu.F:1-805 (REGCOR_GPU) for a block of
 * bodies, as util_gpu.F:152-180 calls it after gpunb_regf, in the branch set
 * of an isolated cluster: KZ(13) = KZ(14) = 0 (no cloud or external force;
 * the KZ(14) <= 2 tidal FIDOT term is kept with the caller's TIDAL(4)),
 * KZ(37) = 0 (no CHECKL), KZ(38) = 0 (no FPCORR).  NB_FLAG and M_FLAG are
 * inputs (regcor_gpu.F:19-20 derives them from KZ(39)).
 *
 * Arithmetic is FP64 evaluated as written (-ffp-contract=off), left to
 * right as gfortran evaluates it.  The reference is compiled without
 * -fdefault-real-8 (configure.ac:116-138), so FLOAT() and the literals 0.1,
 * 0.9, 0.01, 0.003, 1.05, 1.1 and 0.6 are REAL*4: they are restated below as
 * float values (and float arithmetic where both operands are REAL*4),
 * promoted to double where they meet a REAL*8 operand.
 *
 * One deliberate difference from the reference's serial loop: STEP(J) of
 * another body is read as it was before the block (the reference's loop may
 * already have halved STEP(J) for a J corrected earlier in the same block).
 * The new STEP / TIMENW values are written after all bodies are corrected,
 * which is what a parallel corrector needs to be deterministic.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef __FP_FAST_FMA
#error "build the oracle with -ffp-contract=off"
#endif

/* Scalars of common6.h that REGCOR_GPU reads.  Same layout as the device
 * entry point's gpunb_regcor_par. */
typedef struct {
    double time, rscale, rdens[3], rc2, rsfac, znbmin, smax, dtmin, etar,
        smin, bodym, tidal4;
    int nnbopt, nnbmax, n, ntot, ifirst, lmax, nb_flag, m_flag;
    int kz13, kz14, kz37, kz38;   /* must be 0 (not restated) */
} orc_regcor_par;

/* Per-body arrays (0-based body p = I - 1; vectors as 3 consecutive
 * doubles, LIST rows of lmax ints: LIST(1,I) = count, then members).
 * x, xdot: predicted positions / velocities of all bodies at TIME (the
 * neighbour side, X and XDOT of common6.h). */
typedef struct {
    double *x0, *x0dot, *xn, *xndot, *fi, *fidot, *fr, *frdot, *d0, *d0r,
        *d1r, *d2r, *d3r, *t0r, *stepr, *step, *t0, *timenw, *rs, *body, *h;
    int *list;
    const double *x, *xdot;
} orc_regcor_state;

static const double ONE3 = 1.0 / 3.0;   /* zero.f:113 */
static const double ONE12 = 1.0 / 12.0; /* zero.f:116 */
#define MAX_HALVE 2200                  /* > 2098 halvings: DBL_MAX -> 0 */

static double sq(double a) { return a * a; }

/* Pair force / jerk of body j on (xi, xidot) as regcor_gpu.F:349-357 /
 * :409-416 form them; sign +1 adds, -1 subtracts (FREG/FDR or DFIRR/DFD). */
static void pair_add(const orc_regcor_state *S, int j, const double *xi,
                     const double *xidot, double sign, double *f,
                     double *fd) {
    const int p = j - 1;
    const double a1 = S->x[3 * (size_t)p + 0] - xi[0];
    const double a2 = S->x[3 * (size_t)p + 1] - xi[1];
    const double a3 = S->x[3 * (size_t)p + 2] - xi[2];
    double dv[3];
    dv[0] = S->xdot[3 * (size_t)p + 0] - xidot[0];
    dv[1] = S->xdot[3 * (size_t)p + 1] - xidot[1];
    dv[2] = S->xdot[3 * (size_t)p + 2] - xidot[2];
    const double rij2 = a1 * a1 + a2 * a2 + a3 * a3;
    const double dr2i = 1.0 / rij2;
    const double dr3i = S->body[p] * dr2i * sqrt(dr2i);
    const double drdv = a1 * dv[0] + a2 * dv[1] + a3 * dv[2];
    const double drdp = 3.0 * drdv * dr2i;
    const double a[3] = {a1, a2, a3};
    for (int k = 0; k < 3; k++) {
        if (sign > 0) {
            f[k] = f[k] + a[k] * dr3i;
            fd[k] = fd[k] + (dv[k] - a[k] * drdp) * dr3i;
        } else {
            f[k] = f[k] - a[k] * dr3i;
            fd[k] = fd[k] - (dv[k] - a[k] * drdp) * dr3i;
        }
    }
}

/*
 * One body.  nlist: NLIST(1..) of regcor_gpu.F (count, members, room for two
 * more entries); freg/fdr modified in place (JMIN retention); step_old: STEP
 * before the block; *step_new / *timenw_new: this body's new STEP / TIMENW.
 * cnt: NBVOID, NBSMIN, NBFLUX, NICONV increments.
 */
static void regcor_one(const orc_regcor_par *P, orc_regcor_state *S,
                       const double *step_old, int I, const double *xi,
                       const double *xidot, double *freg, double *fdr,
                       int *nlist, int *jjlist, double *step_new,
                       double *timenw_new, int *cnt) {
    const int p = I - 1;
    int *LIST = S->list + (size_t)p * P->lmax;   /* LIST(L,I) = LIST[L-1] */
#define NL(k) nlist[(k) - 1]
#define LI(k) LIST[(k) - 1]
#define JJ(k) jjlist[(k) - 1]
    double dfirr[3], dfd[3];
    int nnb = NL(1);
    const int nnb0 = LI(1);
    const double dtr = P->time - S->t0r[p];
    double ri2, rh2;
    if (P->nb_flag == 1) {
        ri2 = sq(xi[0] - P->rdens[0]) + sq(xi[1] - P->rdens[1]) +
              sq(xi[2] - P->rdens[2]);
        rh2 = sq(P->rscale);
    } else {
        ri2 = sq(xi[0]) + sq(xi[1]) + sq(xi[2]);
        rh2 = 9.0 * sq(P->rscale);
    }
    const double rs2 = sq(S->rs[p]);
    double ridot = 0.0;
    int nbave = 0;

    if (nnb == 0) {
        /* regcor_gpu.F:119-143: small mass at the centre. */
        const double r2 = sq(xi[0]) + sq(xi[1]) + sq(xi[2]);
        const double fij = (double)0.01f * P->bodym / (r2 * sqrt(r2));
        const double rdot =
            3.0 * (xi[0] * xidot[0] + xi[1] * xidot[1] + xi[2] * xidot[2]) /
            r2;
        for (int k = 0; k < 3; k++) {
            dfirr[k] = -(fij * xi[k]);
            dfd[k] = -((xidot[k] - rdot * xi[k]) * fij);
        }
        if (rdot > 0.0) {
            const double a = 0.75 * S->rs[p];
            const double b = (double)0.1f * P->rscale;
            S->rs[p] = a > b ? a : b;
        } else {
            S->rs[p] = (double)1.1f * S->rs[p];
        }
        cnt[0] += 1;
        goto label50;   /* IRSKIP = 1: the radius update below is skipped */
    }

    {
        /* regcor_gpu.F:146-253: Ahmad-Cohen neighbour-radius update. */
        float ft = 0.1f * ((float)P->nnbopt - (float)nnb);
        ft = ft / (float)nnb;
        const double fac = (double)(1.0f + ft);
        double a3 = (double)(float)nnb * fac;
        {
            int mx = P->nnbmax - 50;
            const int zi = (int)P->znbmin;
            if (zi > mx) mx = zi;
            nbave = P->nnbopt + 150 < mx ? P->nnbopt + 150 : mx;
        }
        {
            const double c = (double)(0.9f * (float)nbave);
            if (c < a3) a3 = c;
        }
        const int nbp = (int)a3;
        if (ri2 > rh2 && P->nb_flag == 1) {
            double g = P->rscale / sqrt(ri2);
            if (g < (double)0.1f) g = (double)0.1f;
            a3 = a3 * g;
        } else {
            if (a3 < P->znbmin) a3 = P->znbmin;
        }
        double a4 = a3 / (double)(float)nnb;
        if ((a3 - (double)(float)nnb0) * (a3 - (double)(float)nnb) < 0.0)
            a4 = sqrt(a4);
        ridot = (xi[0] - P->rdens[0]) * xidot[0] +
                (xi[1] - P->rdens[1]) * xidot[1] +
                (xi[2] - P->rdens[2]) * xidot[2];
        if (ri2 > P->rc2 && P->nb_flag == 1 && ri2 < 100.0 * rh2)
            a4 = a4 * (1.0 + ridot * dtr / ri2);
        if (I > P->n && ri2 <= rh2) {
            const double a2 = 100.0 * S->body[p] / fabs(S->h[I - P->n - 1]);
            if (a2 > S->rs[p]) {
                const float q = 1.0f - (float)nnb / (float)nbave;
                a3 = (double)q > 0.0 ? (double)q : 0.0;
                a4 = 1.0 + 0.5 * a3;
            }
        }
        double a5 = P->rsfac * dtr;
        if (0.25 < a5) a5 = 0.25;
        a4 = a4 - 1.0;
        if (a4 > a5) {
            a4 = a5;
        } else {
            if (a4 < -a5) a4 = -a5;
        }
        a3 = ONE3 * a4;
        double a1 = 1.0 + a3 - a3 * a3;
        if (nnb < P->znbmin && nbp > P->znbmin) {
            if (a1 < 1.0) a1 = (double)1.05f;
        }
        if (fabs(a1 - 1.0) > (double)0.003f) {
            if (sq(S->rs[p]) > ri2 && sq(S->rs[p]) < 2.0 * ri2 && nnb > 5)
                a1 = sqrt(a1);
            S->rs[p] = a1 * S->rs[p];
        }
        if (nnb > nbave) {
            if (ridot < 0.0) {
                const float drs = ((float)nnb - (float)nbave) / (3.0f * (float)nnb);
                S->rs[p] = S->rs[p] * (1.0 - (double)drs);
            }
        }
        if (P->m_flag == 0) {
            if (nnb <= 3 && ri2 < 100.0 * rh2) {
                double b1 = 2.0 * S->rs[p];
                for (int l = 1; l <= nnb; l++) {
                    const int j = NL(l + 1);
                    const double *xj = S->x + 3 * (size_t)(j - 1);
                    const double *vj = S->xdot + 3 * (size_t)(j - 1);
                    const double rij = sqrt(sq(xi[0] - xj[0]) + sq(xi[1] - xj[1]) +
                                            sq(xi[2] - xj[2]));
                    const double rsdot = ((xi[0] - xj[0]) * (xidot[0] - vj[0]) +
                                          (xi[1] - xj[1]) * (xidot[1] - vj[1]) +
                                          (xi[2] - xj[2]) * (xidot[2] - vj[2])) /
                                         rij;
                    const double c = rij + rsdot * dtr;
                    if (c < b1) b1 = c;
                }
                const double c = (double)1.1f * S->rs[p];
                S->rs[p] = b1 > c ? b1 : c;
            }
        }
        for (int k = 0; k < 3; k++) {
            dfirr[k] = 0.0;
            dfd[k] = 0.0;
        }
    }

label50:;
    /* regcor_gpu.F:266-402: neighbour losses and gains. */
    int nbloss = 0, nbgain = 0;
    if (nnb0 == 0) {
        nbgain = nnb;
        for (int l = 1; l <= nnb; l++) JJ(l) = NL(l + 1);
        goto label70;
    }
    {
        int jmin = 0;
        int l = 2, lg = 2;
        NL(nnb + 2) = P->ntot + 1;
        NL(1) = LI(nnb0 + 1);
        for (;;) {
            /* label 56 */
            if (LI(l) != NL(lg)) {
                if (LI(l) >= NL(lg)) {
                    nbgain += 1;
                    JJ(nnb0 + nbgain) = NL(lg);
                    l = l - 1;
                } else {
                    nbloss += 1;
                    const int j = LI(l);
                    JJ(nbloss) = j;
                    if (step_old[j - 1] < P->smin) jmin = j;
                    lg = lg - 1;
                }
            }
            /* label 58 */
            if (l <= nnb0) {
                l = l + 1;
                lg = lg + 1;
            } else if (lg <= nnb) {
                lg = lg + 1;
                LI(l) = P->ntot + 1;
            } else {
                break;
            }
        }
        if (jmin == 0) goto label70;
        /* regcor_gpu.F:328-370: keep an old neighbour with a small step. */
        int k = 1;
        for (;;) {
            /* label 60 */
            if (nnb > P->nnbmax || I > P->n) goto label70;
            const int j = JJ(k);
            int skip = step_old[j - 1] > P->smin || j < P->ifirst || j > P->n;
            if (!skip) {
                const double *xj = S->x + 3 * (size_t)(j - 1);
                const double rij2 = sq(xi[0] - xj[0]) + sq(xi[1] - xj[1]) +
                                    sq(xi[2] - xj[2]);
                if (rij2 > 4.0 * rs2) skip = 1;
            }
            if (!skip) {
                int ll = nnb + 1;
                for (;;) {
                    /* label 62 */
                    if (NL(ll) < j) break;
                    NL(ll + 1) = NL(ll);
                    ll = ll - 1;
                    if (!(ll > 1)) break;
                }
                /* label 64 */
                NL(ll + 1) = j;
                nnb += 1;
                nbloss -= 1;
                LI(nnb0 + 1) = NL(1);
                cnt[1] += 1;
                pair_add(S, j, xi, xidot, -1.0, freg, fdr);
                if (k > nbloss) goto label70;
                for (int m = k; m <= nbloss; m++) JJ(m) = JJ(m + 1);
                k = k - 1;
            }
            /* label 68 */
            k = k + 1;
            if (!(k <= nbloss)) break;
        }
    }

label70:;
    /* regcor_gpu.F:400-438: irregular force change from the list change. */
    for (int l = 1; l <= nbloss; l++)
        pair_add(S, JJ(l), xi, xidot, -1.0, dfirr, dfd);
    for (int l = 1; l <= nbgain; l++)
        pair_add(S, JJ(nnb0 + l), xi, xidot, +1.0, dfirr, dfd);

    /* regcor_gpu.F:440-551: corrector. */
    const double dtsq = dtr * dtr;
    const double dt6 = 6.0 / (dtr * dtsq);
    const double dt2 = 2.0 / dtsq;
    double dtsq12 = ONE12 * dtsq;
    double dtr13 = ONE3 * dtr;
    S->t0r[p] = P->time;
    double stepi = step_old[p];
    if (stepi < 5.0 * P->dtmin && dtr > 50.0 * stepi) {
        dtr13 = 0.0;
        dtsq12 = 0.0;
    }
    for (int k = 0; k < 3; k++) {
        const double dfr = S->fr[3 * (size_t)p + k] - freg[k] - dfirr[k];
        const double fdr0 = fdr[k] + dfd[k];
        const double frd = S->frdot[3 * (size_t)p + k];
        const double sum = frd + fdr0;
        const double at3 = 2.0 * dfr + dtr * sum;
        const double bt2 = -3.0 * dfr - dtr * (sum + frd);
        S->xn[3 * (size_t)p + k] =
            S->x0[3 * (size_t)p + k] + ((double)0.6f * at3 + bt2) * dtsq12;
        S->xndot[3 * (size_t)p + k] =
            S->x0dot[3 * (size_t)p + k] + (0.75 * at3 + bt2) * dtr13;
        S->fi[3 * (size_t)p + k] = S->fi[3 * (size_t)p + k] + dfirr[k];
        S->fr[3 * (size_t)p + k] = freg[k];
        S->fidot[3 * (size_t)p + k] = S->fidot[3 * (size_t)p + k] + dfd[k];
        S->frdot[3 * (size_t)p + k] = fdr[k];
        if (k == 1)
            S->fidot[3 * (size_t)p + 0] =
                S->fidot[3 * (size_t)p + 0] - P->tidal4 * dfr;
        if (k == 0)
            S->fidot[3 * (size_t)p + 1] =
                S->fidot[3 * (size_t)p + 1] + P->tidal4 * dfr;
        S->d0[3 * (size_t)p + k] = S->fi[3 * (size_t)p + k];
        S->d0r[3 * (size_t)p + k] = freg[k];
        S->d1r[3 * (size_t)p + k] = fdr[k];
        S->d3r[3 * (size_t)p + k] = at3 * dt6;
        S->d2r[3 * (size_t)p + k] = (3.0 * at3 + bt2) * dt2;
    }

    /* regcor_gpu.F:605-613: new list. */
    if (nbloss + nbgain > 0) {
        LI(1) = nnb;
        for (int l = 2; l <= nnb + 1; l++) LI(l) = NL(l);
        cnt[2] += nbloss + nbgain;
    }

    /* regcor_gpu.F:615-805: new regular step. */
    const double fr2 = sq(freg[0]) + sq(freg[1]) + sq(freg[2]);
    const double w1 = sq(fdr[0]) + sq(fdr[1]) + sq(fdr[2]);
    const double *d2r = S->d2r + 3 * (size_t)p, *d3r = S->d3r + 3 * (size_t)p;
    const double w2 = sq(d2r[0]) + sq(d2r[1]) + sq(d2r[2]);
    const double w3 = sq(d3r[0]) + sq(d3r[1]) + sq(d3r[2]);
    double w0 = (sqrt(fr2 * w2) + w1) / (sqrt(w1 * w3) + w2);
    w0 = P->etar * w0;
    double ttmp = sqrt(w0);
    const double dt0 = ttmp;
    const double steprI = S->stepr[p];
    if (ttmp > 2.0 * steprI) {
        if (fmod(P->time, 2.0 * dtr) == 0.0) {
            ttmp = 2.0 * dtr < P->smax ? 2.0 * dtr : P->smax;
            for (int it = 0; it < MAX_HALVE && dt0 > 2.0 * ttmp &&
                             2.0 * ttmp < P->smax &&
                             fmod(P->time, 2.0 * ttmp) == 0.0;
                 it++)
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
        for (int k = 0; k < 3; k++)
            dv2 = dv2 + sq(xidot[k] - S->xndot[3 * (size_t)p + k]);
        const double dtj =
            dtr * pow(1.0e-04 * sq(dtr) * fr2 / dv2, (double)0.1f);
        if (dtj < ttmp) ttmp = dtr;
    }
    if (P->smax < ttmp) ttmp = P->smax;
    S->stepr[p] = ttmp;
    double tnw = S->timenw[p];
    if (nnb0 == 0 && nnb >= 0) {
        stepi = 0.125 * stepi;
        tnw = S->t0[p] + stepi;
    }
    /* Capped at MAX_HALVE passes (any finite step reaches zero before):
     * the cap only ends loops the reference would not end (ttmp <= 0). */
    for (int it = 0; it < MAX_HALVE && ttmp < stepi; it++) {
        stepi = 0.5 * stepi;
        tnw = S->t0[p] + stepi;
        cnt[3] += 1;
    }
    *step_new = stepi;
    *timenw_new = tnw;
#undef NL
#undef LI
#undef JJ
}

/*
 * A block of ni bodies idx[q] (1-based I), corrected one after the other as
 * util_gpu.F:152-180 calls REGCOR_GPU, with the STEP snapshot described in
 * the file header.  xi/vi: predicted position / velocity of each (x_i,
 * v_i); freg/fdr: regular force / derivative from gpunb_regf (modified by
 * the small-step retention); nlist: ni rows of lmax ints in the form
 * util_gpu.F:103-110 builds (count, members 1-based without the body
 * itself); rows must have room for count + 3 entries.  cnt: NBVOID, NBSMIN,
 * NBFLUX, NICONV, accumulated.  Returns 0, or -1 on an invalid argument.
 */
int orc_regcor(const orc_regcor_par *P, orc_regcor_state *S, int ni,
               const int *idx, const double *xi, const double *vi,
               double *freg, double *fdr, int *nlist, int *cnt) {
    if (ni < 0 || P->lmax < 4) return -1;
    double *step_old = (double *)malloc(sizeof(double) * (size_t)(P->ntot > 0 ? P->ntot : 1));
    double *step_new = (double *)malloc(sizeof(double) * (size_t)(ni > 0 ? ni : 1));
    double *tnw_new = (double *)malloc(sizeof(double) * (size_t)(ni > 0 ? ni : 1));
    int *jj = (int *)malloc(sizeof(int) * 2 * (size_t)P->lmax);
    if (!step_old || !step_new || !tnw_new || !jj) {
        free(step_old); free(step_new); free(tnw_new); free(jj);
        return -1;
    }
    memcpy(step_old, S->step, sizeof(double) * (size_t)P->ntot);
    for (int q = 0; q < ni; q++)
        regcor_one(P, S, step_old, idx[q], xi + 3 * (size_t)q,
                   vi + 3 * (size_t)q, freg + 3 * (size_t)q,
                   fdr + 3 * (size_t)q, nlist + (size_t)q * P->lmax, jj,
                   &step_new[q], &tnw_new[q], cnt);
    for (int q = 0; q < ni; q++) {
        S->step[idx[q] - 1] = step_new[q];
        S->timenw[idx[q] - 1] = tnw_new[q];
    }
    free(step_old);
    free(step_new);
    free(tnw_new);
    free(jj);
    return 0;
}
