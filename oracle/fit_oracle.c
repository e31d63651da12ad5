/*
 * fit_oracle.c -- FP64 CPU restatement of MAOS's DM fitting solve.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle of the GPU fitting solve
 * (gpu_fit_*).  Only tests/ and bench.py's baseline leg load it.
 *
 * Restates, in plain C:
 *   mkh (linear influence)  lib/mkh.c:47-135: HA_f (DM grid -> floc) and
 *                           HX_f (tomography layer -> floc) as sparse rows,
 *                           built here from the grid parameters with the
 *                           reference's own formula for the interpolation
 *                           point (scaled coordinate plus the displacement
 *                           in grid units), taps of weight <= EPS dropped,
 *                           no renormalisation
 *   applyW                  maos/recon_utils.c:155-180
 *   FitR / FitL             maos/recon_utils.c:483-560, plus the assembled
 *                           FL.M's Tikhonov diagonal
 *                           (maos/setup_recon_fit.c:270-283)
 *   pcg, Mmul = NULL        lib/pcg.c:28-113
 * W0/W1, NW and the slaving regulariser S come from the setup's
 * mt_fit_recon_t (they are pinned by their own tests: W0/W1 against the
 * defining overlap integrals by brute-force quadrature, NW by its closed
 * form, S by its construction rules).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/maos_tomo.h"

#define ORF_EPS 1e-15 /* sys/common.h:88 */

/* one sparse row set: rows = floc points, CSR */
typedef struct {
    long nrow;
    long *p, *i;
    double *x;
} orsp_t;

typedef struct {
    int nfit, ndm, npsr, nnw, nthread;
    long fnx, fny, nfloc, nact, nxtot;
    long *aoff, *xoff;
    double *wt;
    orsp_t *HA; /* [f*ndm + d]: columns are the DM's actuators */
    orsp_t *HX; /* [f*npsr + l]: columns are the layer's nodes */
    double *W0, *W1, *NW, *slx, tikdiag;
    long *slp, *sli;
} orf_t;

/* mkhb's rows for a square source grid (all map points valid): the
 * interpolation point of output point (px, py) is
 *   fx = px * scale / dx + (displace_x - ox) / dx    (lib/mkh.c:67-97)
 * split into node and fraction, taps of weight <= EPS or off the grid
 * dropped. */
static void mkh_rows(orsp_t *sp, long snx, long sny, double sox, double soy, double sdx, long fnx, long fny,
                     double fox, double foy, double fdx, double dispx, double dispy, double scale) {
    const long n = fnx * fny;
    sp->nrow = n;
    sp->p = (long *)malloc(sizeof(long) * (n + 1));
    sp->i = (long *)malloc(sizeof(long) * 4 * n);
    sp->x = (double *)malloc(sizeof(double) * 4 * n);
    const double dx_in1 = 1. / sdx, dx_in2 = scale * dx_in1;
    const double dsx = (dispx - sox) * dx_in1, dsy = (dispy - soy) * dx_in1;
    long cnt = 0;
    for (long iy = 0; iy < fny; iy++) {
        for (long ix = 0; ix < fnx; ix++) {
            const long row = ix + iy * fnx;
            sp->p[row] = cnt;
            const double px = fox + ix * fdx, py = foy + iy * fdx;
            double fx1 = px * dx_in2 + dsx, fy1 = py * dx_in2 + dsy;
            const long nx0 = (long)floor(fx1), ny0 = (long)floor(fy1);
            fx1 -= nx0;
            fy1 -= ny0;
            const double fx[2] = {1. - fx1, fx1}, fy[2] = {1. - fy1, fy1};
            for (int b = 0; b < 2; b++) {
                for (int a = 0; a < 2; a++) {
                    const double w = fx[a] * fy[b];
                    const long jx = nx0 + a, jy = ny0 + b;
                    if (w > ORF_EPS && jx >= 0 && jx < snx && jy >= 0 && jy < sny) {
                        sp->i[cnt] = jx + jy * snx;
                        sp->x[cnt] = w;
                        cnt++;
                    }
                }
            }
        }
    }
    sp->p[n] = cnt;
}

static void orsp_free(orsp_t *sp) {
    free(sp->p);
    free(sp->i);
    free(sp->x);
}

orf_t *orf_create(const mt_recon_t *R, const mt_fit_recon_t *F, int nthread) {
    orf_t *o = (orf_t *)calloc(1, sizeof(orf_t));
    o->nfit = F->nfit;
    o->ndm = F->ndm;
    o->npsr = R->npsr;
    o->nnw = F->nnw;
    o->nthread = nthread > 0 ? nthread : 1;
    o->fnx = F->fnx;
    o->fny = F->fny;
    o->nfloc = F->fnx * F->fny;
    o->aoff = (long *)malloc(sizeof(long) * (F->ndm + 1));
    o->aoff[0] = 0;
    for (int d = 0; d < F->ndm; d++) o->aoff[d + 1] = o->aoff[d] + F->anx[d] * F->any[d];
    o->nact = o->aoff[F->ndm];
    o->xoff = (long *)malloc(sizeof(long) * (R->npsr + 1));
    o->xoff[0] = 0;
    for (int l = 0; l < R->npsr; l++) o->xoff[l + 1] = o->xoff[l] + R->xnx[l] * R->xny[l];
    o->nxtot = o->xoff[R->npsr];
    o->wt = (double *)malloc(sizeof(double) * F->nfit);
    memcpy(o->wt, F->wt, sizeof(double) * F->nfit);
    o->HA = (orsp_t *)calloc((size_t)F->nfit * F->ndm, sizeof(orsp_t));
    o->HX = (orsp_t *)calloc((size_t)F->nfit * R->npsr, sizeof(orsp_t));
    for (int f = 0; f < F->nfit; f++) {
        for (int d = 0; d < F->ndm; d++) {
            /* setup_fit_HA (maos/setup_recon_fit.c:49-66): floc at height 0 */
            const double ht = F->aht[d], scale = 1. - ht / F->hs[f];
            mkh_rows(&o->HA[f * F->ndm + d], F->anx[d], F->any[d], F->aox[d], F->aoy[d], F->adx[d], F->fnx, F->fny,
                     F->fox, F->foy, F->fdx, F->thetax[f] * ht, F->thetay[f] * ht, scale);
        }
        for (int l = 0; l < R->npsr; l++) {
            /* setup_fit_HXF (maos/setup_recon_fit.c:25-43) */
            const double ht = R->ht[l], scale = 1. - ht / F->hs[f];
            mkh_rows(&o->HX[f * R->npsr + l], R->xnx[l], R->xny[l], R->xox[l], R->xoy[l], R->xdx[l], F->fnx, F->fny,
                     F->fox, F->foy, F->fdx, F->thetax[f] * ht, F->thetay[f] * ht, scale);
        }
    }
    o->W0 = (double *)malloc(sizeof(double) * 9 * o->nfloc);
    memcpy(o->W0, F->W0, sizeof(double) * 9 * o->nfloc);
    o->W1 = (double *)malloc(sizeof(double) * o->nfloc);
    memcpy(o->W1, F->W1, sizeof(double) * o->nfloc);
    if (F->nnw) {
        o->NW = (double *)malloc(sizeof(double) * F->nnw * o->nact);
        memcpy(o->NW, F->NW, sizeof(double) * F->nnw * o->nact);
    }
    if (F->sl_p) {
        const long nz = F->sl_p[o->nact];
        o->slp = (long *)malloc(sizeof(long) * (o->nact + 1));
        memcpy(o->slp, F->sl_p, sizeof(long) * (o->nact + 1));
        o->sli = (long *)malloc(sizeof(long) * (nz ? nz : 1));
        memcpy(o->sli, F->sl_i, sizeof(long) * nz);
        o->slx = (double *)malloc(sizeof(double) * (nz ? nz : 1));
        memcpy(o->slx, F->sl_x, sizeof(double) * nz);
    }
    o->tikdiag = F->tikdiag;
    return o;
}

void orf_free(orf_t *o) {
    if (!o) return;
    for (int k = 0; k < o->nfit * o->ndm; k++) orsp_free(&o->HA[k]);
    for (int k = 0; k < o->nfit * o->npsr; k++) orsp_free(&o->HX[k]);
    free(o->HA);
    free(o->HX);
    free(o->aoff);
    free(o->xoff);
    free(o->wt);
    free(o->W0);
    free(o->W1);
    free(o->NW);
    free(o->slp);
    free(o->sli);
    free(o->slx);
    free(o);
}

long orf_nact(const orf_t *o) { return o->nact; }
long orf_nfloc(const orf_t *o) { return o->nfloc; }

/* y += H x for one sparse row set */
static void sp_mv(const orsp_t *sp, const double *x, double *y) {
    for (long r = 0; r < sp->nrow; r++) {
        double s = 0;
        for (long k = sp->p[r]; k < sp->p[r + 1]; k++) s += sp->x[k] * x[sp->i[k]];
        y[r] += s;
    }
}
/* y += alpha H' x */
static void sp_mtv(const orsp_t *sp, const double *x, double *y, double alpha) {
    for (long r = 0; r < sp->nrow; r++) {
        const double v = alpha * x[r];
        for (long k = sp->p[r]; k < sp->p[r + 1]; k++) y[sp->i[k]] += sp->x[k] * v;
    }
}

/* applyWeach: xp <- wt (W0 xp - W1 (W1' xp)) */
static void apply_w(const orf_t *o, double *xp, double wt) {
    const long n = o->nfloc, nx = o->fnx, ny = o->fny;
    double *out = (double *)malloc(sizeof(double) * n);
    double dot = 0;
    for (long i = 0; i < n; i++) dot += o->W1[i] * xp[i];
    for (long iy = 0; iy < ny; iy++)
        for (long ix = 0; ix < nx; ix++) {
            const long i = ix + iy * nx;
            double s = 0;
            for (int k = 0; k < 9; k++) {
                const long jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
                const double c = o->W0[9 * i + k];
                if (c != 0 && jx >= 0 && jx < nx && jy >= 0 && jy < ny) s += c * xp[jx + jy * nx];
            }
            out[i] = wt * s - wt * o->W1[i] * dot;
        }
    memcpy(xp, out, sizeof(double) * n);
    free(out);
}

/* xout += alpha FitR x (x: tomography layers) */
void orf_fit_r(const orf_t *o, const double *x, double *xout, double alpha) {
    double *xp = (double *)calloc((size_t)o->nfit * o->nfloc, sizeof(double));
#pragma omp parallel for num_threads(o->nthread)
    for (int f = 0; f < o->nfit; f++) {
        double *xf = xp + (long)f * o->nfloc;
        for (int l = 0; l < o->npsr; l++) sp_mv(&o->HX[f * o->npsr + l], x + o->xoff[l], xf);
        apply_w(o, xf, o->wt[f]);
    }
    for (int f = 0; f < o->nfit; f++)
        for (int d = 0; d < o->ndm; d++) sp_mtv(&o->HA[f * o->ndm + d], xp + (long)f * o->nfloc, xout + o->aoff[d], alpha);
    free(xp);
}

/* xout += alpha FitL a */
void orf_fit_l(const orf_t *o, const double *a, double *xout, double alpha) {
    double *xp = (double *)calloc((size_t)o->nfit * o->nfloc, sizeof(double));
#pragma omp parallel for num_threads(o->nthread)
    for (int f = 0; f < o->nfit; f++) {
        double *xf = xp + (long)f * o->nfloc;
        for (int d = 0; d < o->ndm; d++) sp_mv(&o->HA[f * o->ndm + d], a + o->aoff[d], xf);
        apply_w(o, xf, o->wt[f]);
    }
    for (int f = 0; f < o->nfit; f++)
        for (int d = 0; d < o->ndm; d++) sp_mtv(&o->HA[f * o->ndm + d], xp + (long)f * o->nfloc, xout + o->aoff[d], alpha);
    free(xp);
    /* NW (NW' a) */
    for (int k = 0; k < o->nnw; k++) {
        const double *c = o->NW + (long)k * o->nact;
        double t = 0;
        for (long i = 0; i < o->nact; i++) t += c[i] * a[i];
        for (long i = 0; i < o->nact; i++) xout[i] += alpha * c[i] * t;
    }
    /* slaving regulariser */
    if (o->slp)
        for (long r = 0; r < o->nact; r++) {
            double s = 0;
            for (long k = o->slp[r]; k < o->slp[r + 1]; k++) s += o->slx[k] * a[o->sli[k]];
            xout[r] += alpha * s;
        }
    if (o->tikdiag != 0)
        for (long i = 0; i < o->nact; i++) xout[i] += alpha * o->tikdiag * a[i];
}

static double orf_inn(const double *a, const double *b, long n) {
    double s = 0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* pcg with Mmul = NULL (lib/pcg.c:28-113): x in (warm) / out; returns
 * sqrt(r'r/b'b) after the last iteration, NAN if alpha became non-finite. */
double orf_pcg(const orf_t *o, const double *b, double *x, int warm, int maxit) {
    const long n = o->nact;
    double *r = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *Ap = (double *)malloc(sizeof(double) * n);
    if (!warm) memset(x, 0, sizeof(double) * n);
    const double bb = orf_inn(b, b, n);
    double rr = 0, res = NAN;
    for (int k = 0; k < maxit; k++) {
        if (k % 100 == 0) {
            memcpy(r, b, sizeof(double) * n);
            orf_fit_l(o, x, r, -1);
            memcpy(p, r, sizeof(double) * n);
            rr = orf_inn(r, r, n);
        }
        memset(Ap, 0, sizeof(double) * n);
        orf_fit_l(o, p, Ap, 1);
        const double ak = rr / orf_inn(p, Ap, n);
        if (!isfinite(ak)) {
            res = NAN;
            break;
        }
        for (long i = 0; i < n; i++) x[i] += ak * p[i];
        for (long i = 0; i < n; i++) r[i] -= ak * Ap[i];
        const double rr2 = orf_inn(r, r, n);
        if (k + 1 == maxit) {
            res = sqrt(rr2 / bb);
            break;
        }
        const double bk = rr2 / rr;
        for (long i = 0; i < n; i++) p[i] = bk * p[i] + r[i];
        rr = rr2;
    }
    free(r);
    free(p);
    free(Ap);
    return res;
}

/* The fit as curecon_t::fit runs it: rhs = FitR(opdr); a = pcg(FitL, rhs). */
double orf_solve(const orf_t *o, const double *opdr, double *a, int warm, int maxit) {
    double *rhs = (double *)calloc(o->nact, sizeof(double));
    orf_fit_r(o, opdr, rhs, 1);
    const double res = orf_pcg(o, rhs, a, warm, maxit);
    free(rhs);
    return res;
}
