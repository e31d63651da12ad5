/*
 * tomo_oracle.c -- FP64 CPU restatement of MAOS's tomography solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU
 * baseline ("port") of the GPU solver.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It consumes the same reconstruction operators the GPU path consumes (an
 * mt_recon_t: the RECON_T fields of this path) and restates, in plain C with
 * OpenMP threading where the reference threads:
 *
 *   prop_grid_stat / _transpose   lib/prop_grid.c:47-343 (FUN_NAME_BLOCK, one
 *                                 output column at a time) + :476-506
 *   TTFR                          maos/recon_utils.c:129-137
 *   Tomo_prop / Tomo_nea_gpt /
 *   Tomo_iprop                    maos/recon_utils.c:224-380
 *   apply_L2 (L2'L2)              maos/recon_utils.c:42-48, 360-366 with L2
 *                                 from lib/laplacian.c:70-97
 *   TomoR / TomoL                 maos/recon_utils.c:397-412, 447-477
 *   fdpcg_precond                 maos/fdpcg.c:593-726 (C2C FFT per layer,
 *                                 cvecperm / block multiply / cvecpermi,
 *                                 math/mathmisc.c:264-288)
 *   pcg                           lib/pcg.c:28-113
 *
 * FFT: the reference calls FFTW3 (math/fft.c:278-303, FFTW_FORWARD = -1,
 * unnormalized; version not pinned by the reference's build).  Restated here
 * as a recursive mixed-radix Cooley-Tukey DFT (same definition and sign
 * convention), checked against a brute-force DFT by the tests.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/maos_tomo.h"

typedef double complex dcomplex;

/* ------------------------------------------------------------------ */
/* FFT                                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    int nfac;
    int fac[64]; /* (p, m) pairs */
    dcomplex *tw[2]; /* tw[0]: exp(-2 pi i k/n), tw[1]: exp(+2 pi i k/n) */
} or_fft_t;

static void fft_plan(or_fft_t *p, int n) {
    p->n = n;
    int nf = 0, m = n;
    const int pref[4] = {4, 2, 3, 5};
    while (m > 1) {
        int r = 0;
        for (int i = 0; i < 4; i++) {
            if (m % pref[i] == 0) {
                r = pref[i];
                break;
            }
        }
        if (!r) {
            for (r = 7; r <= m; r += 2)
                if (m % r == 0) break;
        }
        m /= r;
        p->fac[2 * nf] = r;
        p->fac[2 * nf + 1] = m;
        nf++;
    }
    p->nfac = nf;
    for (int s = 0; s < 2; s++) {
        p->tw[s] = (dcomplex *)malloc(sizeof(dcomplex) * (n > 0 ? n : 1));
        for (int k = 0; k < n; k++) {
            /* exact phase reduction keeps the table accurate */
            double ph = 2.0 * M_PI * (double)k / (double)n;
            p->tw[s][k] = cos(ph) + (s ? 1.0 : -1.0) * I * sin(ph);
        }
    }
}
static void fft_free(or_fft_t *p) {
    free(p->tw[0]);
    free(p->tw[1]);
}
static void fft_work(dcomplex *out, const dcomplex *f, long fstride, long istride,
                     const int *fac, const or_fft_t *p, const dcomplex *tw) {
    const int r = fac[0], m = fac[1];
    const dcomplex *fo = f;
    if (m == 1) {
        for (int k = 0; k < r; k++) out[k] = f[(long)k * fstride * istride];
    } else {
        for (int q = 0; q < r; q++) {
            fft_work(out + q * m, fo, fstride * r, istride, fac + 2, p, tw);
            fo += fstride * istride;
        }
    }
    /* generic radix-r butterfly (twiddle index = fstride*k*q mod n) */
    dcomplex scratch[64];
    const int n = p->n;
    for (int u = 0; u < m; u++) {
        for (int q1 = 0, k = u; q1 < r; q1++, k += m) scratch[q1] = out[k];
        for (int q1 = 0, k = u; q1 < r; q1++, k += m) {
            long twidx = 0;
            dcomplex acc = scratch[0];
            for (int q = 1; q < r; q++) {
                twidx += fstride * k;
                twidx %= n;
                acc += scratch[q] * tw[twidx];
            }
            out[k] = acc;
        }
    }
}
/* 1-D FFT of n points read with stride istride into contiguous out. */
static void fft1(const or_fft_t *p, const dcomplex *in, long istride, dcomplex *out, int sign) {
    if (p->n == 1) {
        out[0] = in[0];
        return;
    }
    fft_work(out, in, 1, istride, p->fac, p, p->tw[sign > 0 ? 1 : 0]);
}

/* In-place 2-D FFT of an nx*ny column-major array (x fastest), FFTW sign. */
static void fft2(const or_fft_t *px, const or_fft_t *py, dcomplex *a, long nx, long ny,
                 int sign, dcomplex *buf) {
    for (long iy = 0; iy < ny; iy++) {
        fft1(px, a + iy * nx, 1, buf, sign);
        memcpy(a + iy * nx, buf, sizeof(dcomplex) * nx);
    }
    for (long ix = 0; ix < nx; ix++) {
        fft1(py, a + ix, nx, buf, sign);
        for (long iy = 0; iy < ny; iy++) a[ix + iy * nx] = buf[iy];
    }
}

/* Exported for the tests: out = FFT(in), n points. */
void or_fft1d(const double *in_reim, double *out_reim, int n, int sign) {
    or_fft_t p;
    fft_plan(&p, n);
    fft1(&p, (const dcomplex *)in_reim, 1, (dcomplex *)out_reim, sign);
    fft_free(&p);
}

/* ------------------------------------------------------------------ */
/* ray tracing: prop_grid_stat on the square ploc                      */
/* ------------------------------------------------------------------ */
#define PROP_EPS 1.e-12 /* lib/accphi.c:26 */
#define ifloor(A) ((int)floor(A))
#define iceil(A) ((int)ceil(A))
#define SPLIT(A, B, C) \
    {                  \
        C = ifloor(A); \
        B = (A) - (C); \
    }

/*
 * One output column (nyout = 1) of FUN_NAME_BLOCK, lib/prop_grid.c:47-343,
 * wrap = 0.  trans = 0: phiout += alpha*H*phiin; trans = 1: phiin += alpha*H'*phiout.
 */
static void prop_block_col(double *phiin, long nxin, long nyin, double *phiout, long nxout,
                           double dxout, double dyout, double oxout, double oyout,
                           double alpha, int trans) {
    const long wrapx = nxin - 1, wrapy = nyin - 1;
    long nyout = 1;
    if (fabs(dxout - 1.) < PROP_EPS && fabs(dyout - 1.) < PROP_EPS) {
        /* matched sampling, lib/prop_grid.c:59-143 */
        int irows = 0;
        double dplocx = oxout, dplocy = oyout;
        int nplocx, nplocy0;
        if (dplocx < 0) irows = iceil(-dplocx);
        SPLIT(dplocx, dplocx, nplocx);
        const int do_edge = dplocx < PROP_EPS;
        long rowdiv0 = wrapx - nplocx;
        if (rowdiv0 > nxout) rowdiv0 = nxout;
        if (rowdiv0 < 0) rowdiv0 = 0;
        int icols = 0;
        if (dplocy < 0) icols = iceil(-dplocy);
        int coldiv = ifloor(wrapy - oyout) + 1;
        if (coldiv < nyout) nyout = coldiv;
        SPLIT(dplocy, dplocy, nplocy0);
        const double bl = dplocx * dplocy;
        const double br = (1. - dplocx) * dplocy;
        const double tl = dplocx * (1 - dplocy);
        const double tr = (1. - dplocx) * (1 - dplocy);
        for (long icol = icols; icol < nyout; icol++) {
            long nplocy = nplocy0 + icol;
            while (nplocy >= nyin) nplocy -= nyin;
            /* rows addressed relative to column nplocx of the input */
            const long c1 = nplocy * nxin + nplocx;
            const long c2 = (nplocy == wrapy ? 0 : (nplocy + 1)) * nxin + nplocx;
            long irow = irows;
            for (; irow < rowdiv0; irow++) {
                if (!trans) {
                    phiout[irow] += alpha * (bl * phiin[c2 + irow + 1] + br * phiin[c2 + irow] +
                                             tl * phiin[c1 + irow + 1] + tr * phiin[c1 + irow]);
                } else {
                    double tmp = alpha * phiout[irow];
                    phiin[c2 + irow + 1] += tmp * bl;
                    phiin[c2 + irow] += tmp * br;
                    phiin[c1 + irow + 1] += tmp * tl;
                    phiin[c1 + irow] += tmp * tr;
                }
            }
            /* point on the last input column: right neighbour wraps to 0.
               (The reference would read past the row when the whole column
               lies right of the grid; such rays are treated as missing.) */
            if (irow < nxout && do_edge && nplocx + irow == wrapx) {
                if (!trans) {
                    phiout[irow] += alpha * (bl * phiin[c2 + irow - wrapx] + br * phiin[c2 + irow] +
                                             tl * phiin[c1 + irow - wrapx] + tr * phiin[c1 + irow]);
                } else {
                    double tmp = alpha * phiout[irow];
                    phiin[c2 + irow - wrapx] += tmp * bl;
                    phiin[c2 + irow] += tmp * br;
                    phiin[c1 + irow - wrapx] += tmp * tl;
                    phiin[c1 + irow] += tmp * tr;
                }
            }
        }
    } else {
        /* generic sampling, lib/prop_grid.c:144-315 (no cache: nyout = 1) */
        const double dxout1 = 1. / dxout, dyout1 = 1. / dyout;
        long icols = 0;
        if (oyout < 0) icols = iceil(-oyout * dyout1);
        int coldiv = ifloor((wrapy - oyout) * dyout1) + 1;
        if (coldiv < nyout) nyout = coldiv;
        int irows = 0;
        if (oxout < 0) irows = iceil(-oxout * dxout1);
        int rowdiv0 = iceil((wrapx - oxout) * dxout1);
        if (rowdiv0 <= 0) return;
        if (rowdiv0 > nxout) rowdiv0 = nxout;
        const double dplocx00 = oxout + irows * dxout;
        const double xover = PROP_EPS;
        for (long icol = icols; icol < nyout; icol++) {
            double dplocy = icol * dyout + oyout;
            double dplocx0, dplocy0;
            int nplocx0, nplocy;
            SPLIT(dplocy, dplocy0, nplocy);
            const long c1 = nplocy * nxin;
            const long c2 = (nplocy == wrapy ? 0 : (nplocy + 1)) * nxin;
            const double dplocy1 = 1. - dplocy0;
            double dplocx = dplocx00;
            int irow = irows;
            /* First the points whose right neighbour exists (offset 1), then
               the points right on the last column (offset -wrapx: the right
               neighbour wraps to column 0 with a vanishing weight).  Written
               with an explicit right-neighbour index n1; the bounds guard only
               matters where accumulated rounding would make the reference read
               outside the row. */
            for (int pass = 0; pass < 2; pass++) {
                for (; pass == 0 ? irow < rowdiv0 : (irow < nxout && dplocx < wrapx + xover);
                     irow++, dplocx += dxout) {
                    SPLIT(dplocx, dplocx0, nplocx0);
                    if (nplocx0 < 0 || nplocx0 > wrapx) continue;
                    const long n1 = nplocx0 == wrapx ? 0 : nplocx0 + 1;
                    if (!trans) {
                        phiout[irow] += alpha * (dplocy1 * ((1 - dplocx0) * phiin[c1 + nplocx0] +
                                                            dplocx0 * phiin[c1 + n1]) +
                                                 dplocy0 * ((1 - dplocx0) * phiin[c2 + nplocx0] +
                                                            dplocx0 * phiin[c2 + n1]));
                    } else {
                        double tmp = phiout[irow] * alpha;
                        phiin[c1 + nplocx0] += tmp * dplocy1 * (1 - dplocx0);
                        phiin[c1 + n1] += tmp * dplocy1 * dplocx0;
                        phiin[c2 + nplocx0] += tmp * dplocy0 * (1 - dplocx0);
                        phiin[c2 + n1] += tmp * dplocy0 * dplocx0;
                    }
                }
            }
        }
    }
}

/*
 * prop_grid_stat (trans = 0) / prop_grid_stat_transpose (trans = 1),
 * lib/prop_grid.c:476-506, for a square output grid (one locstat column per
 * grid row).  mapin: nxin x nyin at (ox, oy, dx).
 */
static void prop_grid_stat(double *mapin, long nxin, long nyin, double mox, double moy,
                           double mdx, double *phiout, long pnx, long pny, double pox,
                           double poy, double pdx, double alpha, double displacex,
                           double displacey, double scale, int trans) {
    const double dx_in1 = 1. / mdx;
    const double dxout = pdx * dx_in1 * scale;
    const double dyout = pdx * dx_in1 * scale;
    displacex = (displacex - mox) * dx_in1;
    displacey = (displacey - moy) * dx_in1;
    for (long icol = 0; icol < pny; icol++) {
        const double oxout = pox * dx_in1 * scale + displacex;
        const double oyout = (poy + icol * pdx) * dx_in1 * scale + displacey;
        prop_block_col(mapin, nxin, nyin, phiout + icol * pnx, pnx, dxout, dyout, oxout,
                       oyout, alpha, trans);
    }
}

/* ------------------------------------------------------------------ */
/* oracle state                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int npsr, nwfs;
    long *xnx, *xny, *xoff;
    double *xox, *xoy, *xdx, *ht;
    long pnx, pny;
    double pox, poy, pdx;
    double *thetax, *thetay, *hs;
    int *skip, *ttr;
    long *nsa, *goff;
    long **gp_p, **gp_i;
    double **gp_x, **neai, **ptt;
    int ptt_lhs;
    double *l2cf;
    int precond, bs, scale;
    long nb;
    dcomplex *Minv;
    long *perm;
    long nxtot, ngtot;
    or_fft_t *fftx, *ffty;
    int nthread;
} or_t;

static void *dupmem(const void *p, size_t n) {
    if (!p || !n) return NULL;
    void *q = malloc(n);
    memcpy(q, p, n);
    return q;
}

or_t *or_create(const mt_recon_t *r, int nthread) {
    or_t *o = (or_t *)calloc(1, sizeof(or_t));
    o->npsr = r->npsr;
    o->nwfs = r->nwfs;
    o->xnx = (long *)dupmem(r->xnx, sizeof(long) * r->npsr);
    o->xny = (long *)dupmem(r->xny, sizeof(long) * r->npsr);
    o->xox = (double *)dupmem(r->xox, sizeof(double) * r->npsr);
    o->xoy = (double *)dupmem(r->xoy, sizeof(double) * r->npsr);
    o->xdx = (double *)dupmem(r->xdx, sizeof(double) * r->npsr);
    o->ht = (double *)dupmem(r->ht, sizeof(double) * r->npsr);
    o->xoff = (long *)calloc(r->npsr + 1, sizeof(long));
    for (int l = 0; l < r->npsr; l++) o->xoff[l + 1] = o->xoff[l] + r->xnx[l] * r->xny[l];
    o->nxtot = o->xoff[r->npsr];
    o->pnx = r->pnx;
    o->pny = r->pny;
    o->pox = r->pox;
    o->poy = r->poy;
    o->pdx = r->pdx;
    const int nw = r->nwfs;
    o->thetax = (double *)dupmem(r->thetax, sizeof(double) * nw);
    o->thetay = (double *)dupmem(r->thetay, sizeof(double) * nw);
    o->hs = (double *)dupmem(r->hs, sizeof(double) * nw);
    o->skip = (int *)dupmem(r->skip, sizeof(int) * nw);
    o->ttr = (int *)dupmem(r->trs, sizeof(int) * nw);
    o->nsa = (long *)dupmem(r->nsa, sizeof(long) * nw);
    o->goff = (long *)calloc(nw + 1, sizeof(long));
    o->gp_p = (long **)calloc(nw, sizeof(long *));
    o->gp_i = (long **)calloc(nw, sizeof(long *));
    o->gp_x = (double **)calloc(nw, sizeof(double *));
    o->neai = (double **)calloc(nw, sizeof(double *));
    o->ptt = (double **)calloc(nw, sizeof(double *));
    const long nploc = r->pnx * r->pny;
    for (int w = 0; w < nw; w++) {
        o->goff[w + 1] = o->goff[w] + 2 * r->nsa[w];
        if (r->skip[w]) continue;
        const long nnz = r->gp_p[w][nploc];
        o->gp_p[w] = (long *)dupmem(r->gp_p[w], sizeof(long) * (nploc + 1));
        o->gp_i[w] = (long *)dupmem(r->gp_i[w], sizeof(long) * nnz);
        o->gp_x[w] = (double *)dupmem(r->gp_x[w], sizeof(double) * nnz);
        o->neai[w] = (double *)dupmem(r->neai[w], sizeof(double) * 3 * r->nsa[w]);
        if (r->trs[w] && r->ptt[w]) o->ptt[w] = (double *)dupmem(r->ptt[w], sizeof(double) * 4 * r->nsa[w]);
    }
    o->ngtot = o->goff[nw];
    o->ptt_lhs = r->ptt_lhs;
    o->l2cf = (double *)calloc(r->npsr, sizeof(double));
    for (int l = 0; l < r->npsr; l++) o->l2cf[l] = 4.0 * sqrt(r->l2c[l]);
    o->precond = r->precond;
    if (r->precond) {
        o->bs = r->fd_bs;
        o->nb = r->fd_nb;
        o->scale = r->fd_scale;
        o->Minv = (dcomplex *)dupmem(r->fd_Minv, sizeof(dcomplex) * r->fd_nb * r->fd_bs * r->fd_bs);
        o->perm = (long *)dupmem(r->fd_perm, sizeof(long) * r->fd_nb * r->fd_bs);
        o->fftx = (or_fft_t *)calloc(r->npsr, sizeof(or_fft_t));
        o->ffty = (or_fft_t *)calloc(r->npsr, sizeof(or_fft_t));
        for (int l = 0; l < r->npsr; l++) {
            fft_plan(&o->fftx[l], (int)r->xnx[l]);
            fft_plan(&o->ffty[l], (int)r->xny[l]);
        }
    }
    o->nthread = nthread > 0 ? nthread : 1;
    return o;
}

void or_free(or_t *o) {
    if (!o) return;
    for (int w = 0; w < o->nwfs; w++) {
        free(o->gp_p[w]);
        free(o->gp_i[w]);
        free(o->gp_x[w]);
        free(o->neai[w]);
        free(o->ptt[w]);
    }
    if (o->fftx) {
        for (int l = 0; l < o->npsr; l++) {
            fft_free(&o->fftx[l]);
            fft_free(&o->ffty[l]);
        }
    }
    free(o->fftx);
    free(o->ffty);
    free(o->gp_p);
    free(o->gp_i);
    free(o->gp_x);
    free(o->neai);
    free(o->ptt);
    free(o->xnx);
    free(o->xny);
    free(o->xox);
    free(o->xoy);
    free(o->xdx);
    free(o->ht);
    free(o->xoff);
    free(o->thetax);
    free(o->thetay);
    free(o->hs);
    free(o->skip);
    free(o->ttr);
    free(o->nsa);
    free(o->goff);
    free(o->l2cf);
    free(o->Minv);
    free(o->perm);
    free(o);
}

long or_nx(const or_t *o) { return o->nxtot; }
long or_ng(const or_t *o) { return o->ngtot; }

/* HX for one WFS: opd (pnx*pny) += sum_l prop(x_l) (maos/recon_utils.c:224-265) */
static void hx_wfs(const or_t *o, int w, const double *x, double *opd) {
    for (int l = 0; l < o->npsr; l++) {
        const double ht = o->ht[l];
        const double dispx = o->thetax[w] * ht, dispy = o->thetay[w] * ht;
        const double scale = 1. - ht / o->hs[w];
        prop_grid_stat((double *)(x + o->xoff[l]), o->xnx[l], o->xny[l], o->xox[l], o->xoy[l],
                       o->xdx[l], opd, o->pnx, o->pny, o->pox, o->poy, o->pdx, 1, dispx,
                       dispy, scale, 0);
    }
}
/* Exported helpers for the operator-level tests. */
void or_hx(const or_t *o, int w, const double *x, double *opd) { hx_wfs(o, w, x, opd); }
void or_hxt(const or_t *o, int w, const double *opd, double *x) {
    for (int l = 0; l < o->npsr; l++) {
        const double ht = o->ht[l];
        prop_grid_stat(x + o->xoff[l], o->xnx[l], o->xny[l], o->xox[l], o->xoy[l], o->xdx[l],
                       (double *)opd, o->pnx, o->pny, o->pox, o->poy, o->pdx, 1,
                       o->thetax[w] * ht, o->thetay[w] * ht, 1. - ht / o->hs[w], 1);
    }
}

/* gg = GP * opd (dspmm "nn", CSC 2nsa x nploc) */
static void gp_apply(const or_t *o, int w, const double *opd, double *gg) {
    const long nploc = o->pnx * o->pny;
    const long *p = o->gp_p[w], *ii = o->gp_i[w];
    const double *xx = o->gp_x[w];
    memset(gg, 0, sizeof(double) * 2 * o->nsa[w]);
    for (long c = 0; c < nploc; c++) {
        const double v = opd[c];
        for (long k = p[c]; k < p[c + 1]; k++) gg[ii[k]] += xx[k] * v;
    }
}
/* out += alpha * GP' * gg (dspmm "tn") */
static void gpt_apply(const or_t *o, int w, const double *gg, double *out, double alpha) {
    const long nploc = o->pnx * o->pny;
    const long *p = o->gp_p[w], *ii = o->gp_i[w];
    const double *xx = o->gp_x[w];
    for (long c = 0; c < nploc; c++) {
        double s = 0;
        for (long k = p[c]; k < p[c + 1]; k++) s += xx[k] * gg[ii[k]];
        out[c] += alpha * s;
    }
}
/* gg = NEAI * gg (saneai: diagonal + x/y cross terms) */
static void nea_apply(const or_t *o, int w, double *gg) {
    const long nsa = o->nsa[w];
    const double *ne = o->neai[w];
    for (long i = 0; i < nsa; i++) {
        const double gx = gg[i], gy = gg[i + nsa];
        gg[i] = ne[3 * i] * gx + ne[3 * i + 2] * gy;
        gg[i + nsa] = ne[3 * i + 2] * gx + ne[3 * i + 1] * gy;
    }
}
/* TTFR: g -= TT*(PTT*g) (maos/recon_utils.c:129-137) */
static void ttfr(const or_t *o, int w, double *gg) {
    if (!o->ttr[w] || !o->ptt[w]) return;
    const long nsa = o->nsa[w];
    const double *P = o->ptt[w]; /* 2 x 2nsa column-major */
    double j0 = 0, j1 = 0;
    for (long i = 0; i < 2 * nsa; i++) {
        j0 += P[2 * i] * gg[i];
        j1 += P[2 * i + 1] * gg[i];
    }
    for (long i = 0; i < nsa; i++) {
        gg[i] -= j0;
        gg[i + nsa] -= j1;
    }
}
/* xout += alpha * L2' L2 xin for layer l (periodic 5-point, lib/laplacian.c:70-97) */
static void l2_apply(const or_t *o, int l, const double *xin, double *xout, double alpha,
                     double *tmp) {
    const long nx = o->xnx[l], ny = o->xny[l];
    const double cf = o->l2cf[l];
    for (long iy = 0; iy < ny; iy++) {
        const long iym = iy == 0 ? ny - 1 : iy - 1, iyp = iy == ny - 1 ? 0 : iy + 1;
        for (long ix = 0; ix < nx; ix++) {
            const long ixm = ix == 0 ? nx - 1 : ix - 1, ixp = ix == nx - 1 ? 0 : ix + 1;
            tmp[ix + iy * nx] = 0.25 * cf * xin[ix + iym * nx] + 0.25 * cf * xin[ixm + iy * nx] -
                                cf * xin[ix + iy * nx] + 0.25 * cf * xin[ixp + iy * nx] +
                                0.25 * cf * xin[ix + iyp * nx];
        }
    }
    for (long iy = 0; iy < ny; iy++) {
        const long iym = iy == 0 ? ny - 1 : iy - 1, iyp = iy == ny - 1 ? 0 : iy + 1;
        for (long ix = 0; ix < nx; ix++) {
            const long ixm = ix == 0 ? nx - 1 : ix - 1, ixp = ix == nx - 1 ? 0 : ix + 1;
            xout[ix + iy * nx] += alpha * (0.25 * cf * tmp[ix + iym * nx] + 0.25 * cf * tmp[ixm + iy * nx] -
                                           cf * tmp[ix + iy * nx] + 0.25 * cf * tmp[ixp + iy * nx] +
                                           0.25 * cf * tmp[ix + iyp * nx]);
        }
    }
}

/* Exported for the tests: xout += alpha * L2'L2 xin on layer l. */
void or_l2(const or_t *o, int l, const double *xin, double *xout, double alpha) {
    double *tmp = (double *)malloc(sizeof(double) * o->xnx[l] * o->xny[l]);
    l2_apply(o, l, xin, xout, alpha, tmp);
    free(tmp);
}

/* iprop for layer l: xout_l += sum_w HX_wl' gg_w (+ alpha L2'L2 xin_l) */
static void iprop_layer(const or_t *o, int l, double *const *gg, const double *xin, double *xout,
                        double alpha) {
    double *xo = xout + o->xoff[l];
    const double ht = o->ht[l];
    for (int w = 0; w < o->nwfs; w++) {
        if (o->skip[w] || !gg[w]) continue;
        prop_grid_stat(xo, o->xnx[l], o->xny[l], o->xox[l], o->xoy[l], o->xdx[l], gg[w], o->pnx,
                       o->pny, o->pox, o->poy, o->pdx, 1, o->thetax[w] * ht, o->thetay[w] * ht,
                       1. - ht / o->hs[w], 1);
    }
    if (xin) {
        double *tmp = (double *)malloc(sizeof(double) * o->xnx[l] * o->xny[l]);
        l2_apply(o, l, xin + o->xoff[l], xo, alpha, tmp);
        free(tmp);
    }
}

/* TomoR: xout = alpha * HX' GP' W (I - TT PTT) g   (xout overwritten) */
void or_tomo_r(const or_t *o, const double *g, double *xout, double alpha) {
    const long nploc = o->pnx * o->pny;
    double **opd = (double **)calloc(o->nwfs, sizeof(double *));
    memset(xout, 0, sizeof(double) * o->nxtot);
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int w = 0; w < o->nwfs; w++) {
        if (o->skip[w]) continue;
        const long ng = 2 * o->nsa[w];
        double *gg = (double *)malloc(sizeof(double) * ng);
        memcpy(gg, g + o->goff[w], sizeof(double) * ng);
        ttfr(o, w, gg);
        nea_apply(o, w, gg);
        opd[w] = (double *)calloc(nploc, sizeof(double));
        gpt_apply(o, w, gg, opd[w], alpha);
        free(gg);
    }
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int l = 0; l < o->npsr; l++) iprop_layer(o, l, opd, NULL, xout, alpha);
    for (int w = 0; w < o->nwfs; w++) free(opd[w]);
    free(opd);
}

/* TomoL: xout += alpha * (HX' GP' W GP HX + L2'L2) xin */
void or_tomo_l(const or_t *o, const double *xin, double *xout, double alpha) {
    const long nploc = o->pnx * o->pny;
    double **opd = (double **)calloc(o->nwfs, sizeof(double *));
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int w = 0; w < o->nwfs; w++) {
        if (o->skip[w]) continue;
        const long ng = 2 * o->nsa[w];
        double *xx = (double *)calloc(nploc, sizeof(double));
        hx_wfs(o, w, xin, xx);
        double *gg = (double *)malloc(sizeof(double) * ng);
        gp_apply(o, w, xx, gg);
        if (o->ptt_lhs) ttfr(o, w, gg);
        nea_apply(o, w, gg);
        memset(xx, 0, sizeof(double) * nploc);
        gpt_apply(o, w, gg, xx, alpha);
        free(gg);
        opd[w] = xx;
    }
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int l = 0; l < o->npsr; l++) iprop_layer(o, l, opd, xin, xout, alpha);
    for (int w = 0; w < o->nwfs; w++) free(opd[w]);
    free(opd);
}

/* fdpcg_precond: xout = IFFT( perm^-1 ( Minv * perm( FFT(xin) ) ) ) */
void or_fdpcg(const or_t *o, const double *xin, double *xout) {
    const long nxtot = o->nxtot, bs = o->bs, nb = o->nb;
    dcomplex *xhat = (dcomplex *)malloc(sizeof(dcomplex) * nxtot);
    dcomplex *xhat2 = (dcomplex *)malloc(sizeof(dcomplex) * nxtot);
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int l = 0; l < o->npsr; l++) {
        const long n = o->xnx[l] * o->xny[l];
        dcomplex *a = xhat + o->xoff[l];
        for (long i = 0; i < n; i++) a[i] = xin[o->xoff[l] + i];
        long nbuf = o->xnx[l] > o->xny[l] ? o->xnx[l] : o->xny[l];
        dcomplex *buf = (dcomplex *)malloc(sizeof(dcomplex) * nbuf);
        fft2(&o->fftx[l], &o->ffty[l], a, o->xnx[l], o->xny[l], -1, buf);
        free(buf);
        if (o->scale) {
            const double s = 1. / sqrt((double)n);
            for (long i = 0; i < n; i++) a[i] *= s;
        }
    }
    /* cvecperm (math/mathmisc.c:264-275) */
    for (long i = 0; i < nxtot; i++) {
        const long pm = o->perm[i];
        xhat2[i] = pm > 0 ? xhat[pm] : conj(xhat[-pm]);
    }
#pragma omp parallel for num_threads(o->nthread) schedule(static)
    for (long ib = 0; ib < nb; ib++) {
        const dcomplex *M = o->Minv + ib * bs * bs;
        const dcomplex *v = xhat2 + ib * bs;
        dcomplex *y = xhat + ib * bs;
        for (long a = 0; a < bs; a++) y[a] = 0;
        for (long b = 0; b < bs; b++)
            for (long a = 0; a < bs; a++) y[a] += M[a + b * bs] * v[b];
    }
    /* cvecpermi (math/mathmisc.c:277-288) */
    for (long i = 0; i < nxtot; i++) {
        const long pm = o->perm[i];
        if (pm > 0) xhat2[pm] = xhat[i];
        else xhat2[-pm] = conj(xhat[i]);
    }
#pragma omp parallel for num_threads(o->nthread) schedule(dynamic, 1)
    for (int l = 0; l < o->npsr; l++) {
        const long n = o->xnx[l] * o->xny[l];
        dcomplex *a = xhat2 + o->xoff[l];
        long nbuf = o->xnx[l] > o->xny[l] ? o->xnx[l] : o->xny[l];
        dcomplex *buf = (dcomplex *)malloc(sizeof(dcomplex) * nbuf);
        fft2(&o->fftx[l], &o->ffty[l], a, o->xnx[l], o->xny[l], 1, buf);
        free(buf);
        const double s = o->scale ? 1. / sqrt((double)n) : 1.;
        for (long i = 0; i < n; i++) xout[o->xoff[l] + i] = creal(a[i]) * s;
    }
    free(xhat);
    free(xhat2);
}

static double dinn(const double *a, const double *b, long n) {
    double s = 0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/*
 * pcg (lib/pcg.c:28-113).  x: in (warm start when warm) / out.  Returns
 * sqrt(r'r/b'b) after the last iteration; NAN if alpha became non-finite
 * (the reference then returns an unset value).  res_hist (maxit+1 entries,
 * may be NULL) receives the residual history the reference prints with
 * PRINT_RES: res[k] = sqrt(r'r/b'b) at the start of iteration k.
 */
double or_pcg(const or_t *o, const double *b, double *x, int warm, int maxit, int use_precond,
              double *res_hist) {
    const long n = o->nxtot;
    double *r = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *Ap = (double *)malloc(sizeof(double) * n);
    const int M = use_precond && o->precond;
    if (!warm) memset(x, 0, sizeof(double) * n);
    double r0z1 = 0, r0z2;
    const double r0z0 = dinn(b, b, n);
    double res = NAN;
    for (int k = 0; k < maxit; k++) {
        if (k % 100 == 0) {
            memcpy(r, b, sizeof(double) * n);
            or_tomo_l(o, x, r, -1);
            if (M) or_fdpcg(o, r, z);
            else memcpy(z, r, sizeof(double) * n);
            memcpy(p, z, sizeof(double) * n);
            r0z1 = dinn(r, z, n);
        }
        if (res_hist) res_hist[k] = sqrt(dinn(r, r, n) / r0z0);
        memset(Ap, 0, sizeof(double) * n);
        or_tomo_l(o, p, Ap, 1);
        double ak = r0z1 / dinn(p, Ap, n);
        if (!isfinite(ak)) {
            res = NAN;
            goto end;
        }
        for (long i = 0; i < n; i++) x[i] += ak * p[i];
        for (long i = 0; i < n; i++) r[i] += -ak * Ap[i];
        if (k + 1 == maxit) {
            res = sqrt(dinn(r, r, n) / r0z0);
            if (res_hist) res_hist[k + 1] = res;
            break;
        }
        if (M) or_fdpcg(o, r, z);
        else memcpy(z, r, sizeof(double) * n);
        r0z2 = dinn(r, z, n);
        const double bk = r0z2 / r0z1;
        for (long i = 0; i < n; i++) p[i] = bk * p[i] + z[i];
        r0z1 = r0z2;
    }
end:
    free(r);
    free(z);
    free(p);
    free(Ap);
    return res;
}

/* Full solve as tomofit does it: rhs = TomoR(g); x = pcg(...) (lib/muv.c:417-447). */
double or_solve(const or_t *o, const double *g, double *x, int warm, int maxit) {
    double *rhs = (double *)malloc(sizeof(double) * o->nxtot);
    or_tomo_r(o, g, rhs, 1);
    double res = or_pcg(o, rhs, x, warm, maxit, 1, NULL);
    free(rhs);
    return res;
}
