/*
 * energy.c -- CPU restatement of NBODY6++GPU's potential-energy sum.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for gpunb_energy_dev:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * orc_energy follows the GPU branch of ENERGY (src/Main/energy.F:37-79)
 * after its GPUPOT call: the c.m. potential copied to the KS components,
 * PHICOR (src/Main/phicor.f:1-112) per body I = IFIRST..NTOT, the serial sum
 * POT = 1/2 sum BODY*PHIDBL, and ZKIN (energy.F:141-146).  FP64 evaluated as
 * written (-ffp-contract=off), left to right.
 */
#include <math.h>
#include <stddef.h>

#ifdef __FP_FAST_FMA
#error "build the oracle with -ffp-contract=off"
#endif

typedef struct {
    int n, ntot, ifirst, lmax;
    double cmsep2_rmin2;
} orc_energy_par;

/* RIJ2 = 0; RIJ2 = RIJ2 + (X(K,a) - X(K,b))**2 (1-based bodies). */
static double d2(const double *x, int a, int b) {
    double s = 0.0;
    for (int k = 0; k < 3; k++) {
        const double d = x[3 * (size_t)(a - 1) + k] - x[3 * (size_t)(b - 1) + k];
        s = s + d * d;
    }
    return s;
}

/* PHICOR(I, DPHI1, DPHI2), phicor.f:1-112.  Returns -1 on an invalid row. */
int orc_phicor(const orc_energy_par *P, const double *body, const double *x,
               const int *list, int I, double *dphi1_out,
               double *dphi2_out) {
    const int N = P->n;
    double dphi1 = 0.0, dphi2 = 0.0;
    if (I <= N) {
        const int *row = list + (size_t)(I - 1) * P->lmax;
        const int nnb = row[0];
        if (nnb < 0 || nnb + 1 > P->lmax) return -1;
        for (int l = nnb + 1; l >= 2; l--) {
            const int J = row[l - 1];
            if (J < 1 || J > P->ntot) return -1;
            if (J <= N) break;
            const double rij2 = d2(x, I, J);
            if (rij2 < P->cmsep2_rmin2) {
                const int j1 = 2 * (J - N) - 1, j2 = j1 + 1;
                const double rj1 = d2(x, I, j1);
                const double rj2 = d2(x, I, j2);
                dphi1 = dphi1 - body[J - 1] / sqrt(rij2) +
                        body[j1 - 1] / sqrt(rj1) + body[j2 - 1] / sqrt(rj2);
            }
        }
    } else {
        const int i1 = 2 * (I - N) - 1, i2 = i1 + 1;
        const int *row = list + (size_t)(i1 - 1) * P->lmax;
        const int np = row[0];
        if (np < 0 || np + 1 > P->lmax) return -1;
        for (int l = 2; l <= np + 1; l++) {
            const int J = row[l - 1];
            if (J < 1 || J > P->ntot) return -1;
            const double rij2 = d2(x, I, J);
            if (J <= N) {
                const double ri1 = d2(x, i1, J);
                const double ri2 = d2(x, i2, J);
                dphi1 = dphi1 - body[J - 1] / sqrt(rij2) + body[J - 1] / sqrt(ri1);
                dphi2 = dphi2 - body[J - 1] / sqrt(rij2) + body[J - 1] / sqrt(ri2);
            } else {
                dphi1 = dphi1 - body[J - 1] / sqrt(rij2);
                dphi2 = dphi2 - body[J - 1] / sqrt(rij2);
                int j1 = 2 * (J - N) - 1;
                const int j2 = j1 + 1;
                for (;;) {   /* label 25: J1 then J2 */
                    const double rj1 = d2(x, i1, j1);
                    const double rj2 = d2(x, i2, j1);
                    dphi1 = dphi1 + body[j1 - 1] / sqrt(rj1);
                    dphi2 = dphi2 + body[j1 - 1] / sqrt(rj2);
                    j1 = j1 + 1;
                    if (j1 != j2) break;
                }
            }
        }
    }
    *dphi1_out = dphi1;
    *dphi2_out = dphi2;
    return 0;
}

/*
 * phidbl: PHIDBL[NTOT], on entry with GPUPOT's values at IFIRST..NTOT; on
 * return corrected for singles and KS components.  out[0] = POT, out[1] =
 * ZKIN.  Returns 0, or -1 if a LIST row is invalid.
 */
int orc_energy(const orc_energy_par *P, const double *body, const double *x,
               const double *xdot, const int *list, double *phidbl,
               double *out) {
    const int N = P->n, npairs = P->ntot - P->n;
    for (int ipair = 1; ipair <= npairs; ipair++) {
        const int i1 = 2 * ipair - 1;
        phidbl[i1 - 1] = phidbl[N + ipair - 1];
        phidbl[i1] = phidbl[N + ipair - 1];
    }
    double pot = 0.0;
    int rc = 0;
    for (int I = P->ifirst; I <= P->ntot; I++) {
        double dphi1, dphi2;
        if (orc_phicor(P, body, x, list, I, &dphi1, &dphi2) != 0) {
            rc = -1;
            continue;
        }
        if (I <= N) {
            phidbl[I - 1] = phidbl[I - 1] + dphi1;
            pot = pot + body[I - 1] * phidbl[I - 1];
        } else {
            const int i1 = 2 * (I - N) - 1;
            phidbl[i1 - 1] = phidbl[i1 - 1] + dphi1;
            phidbl[i1] = phidbl[i1] + dphi2;
            pot = pot + body[i1 - 1] * phidbl[i1 - 1] + body[i1] * phidbl[i1];
        }
    }
    pot = 0.5 * pot;
    double zkin = 0.0;
    for (int I = P->ifirst; I <= P->ntot; I++) {
        const double *v = xdot + 3 * (size_t)(I - 1);
        zkin = zkin + body[I - 1] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    zkin = 0.5 * zkin;
    out[0] = pot;
    out[1] = zkin;
    return rc;
}
