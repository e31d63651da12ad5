// DM fitting on the GPU (gpu_fit_*): the CG solve after tomography,
//
//   (sum_f wt_f HA_f' W HA_f + NW NW' + S + tikdiag I) a = sum_f wt_f HA_f' W HX_f x,
//   W = W0 - W1 W1',
//
// replacing curecon_t::fit with cufit_grid (cuda/recon.cu:277-288,
// cuda/fit.cu:47-230, cuda/recon_base.cu:75-139) and the CPU FitR / FitL
// (maos/recon_utils.c:483-560) with fit.precond=0 (plain CG, lib/pcg.c:28-113
// with Mmul = NULL).
//
// Design.  A fit is small (the 30 m system with two DMs: 8.5k actuators,
// 4k fitting-grid points per direction) and its CG is a chain of dependent
// passes, so the whole solve of one system runs in ONE workgroup of 1024
// threads (16 waves) that loops over the CG iterations with workgroup
// barriers between the passes: one launch per batch of systems instead of
// ~6 launches per iteration, each system on its own CU (a batch of n systems
// fills n CUs; MVM assembly solves thousands).  Every operator is a gather:
//   HX_f, HA_f  bilinear interpolation of the layers / DM grids at the
//               fitting-grid points along direction f, from per-axis tap
//               tables (the grids are square and axis-aligned, so a 2-D tap
//               is the product of an x and a y tap, as mkh builds it);
//   W           9-point W0 stencil per point (rows of the reference's sparse
//               W0) minus the W1 piston projection, per direction;
//   HA_f'       per actuator, the fitting-grid columns and rows whose taps
//               land on it (transposed per-axis lists), no atomics;
//   NW, S       low-rank dots (block reductions) and a CSR row gather.
// Arithmetic is FP64 throughout: the fitting matrix carries a Tikhonov term
// of 4e-7/nact relative to its largest eigenvalue, too ill-conditioned for
// FP32 vectors at 1e-4 parity, and the solve is latency-bound, not
// FLOP-bound.  Reductions are fixed-order (a system's bits do not depend on
// its batch neighbours).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fit_internal.hpp"
#include "maos_tomo.h"

namespace mt {
void set_error(const std::string &msg);
}

namespace {

struct FitErr {
    int code;
    std::string msg;
};
#define FCHK(call)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw FitErr{MT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
[[noreturn]] void ffail(int code, const std::string &m) { throw FitErr{code, m}; }

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        free();
        n = count;
        if (count) FCHK(hipMalloc((void **)&p, sizeof(T) * count));
    }
    void upload(const std::vector<T> &h) {
        alloc(h.size());
        if (n) FCHK(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { free(); }
};

constexpr int FMAXDM = 8, FMAXFIT = 32, FMAXL = 16, FMAXNW = 32;
constexpr int FTHREADS = 1024;
constexpr double FEPS = 1e-15;  // tap weights not above this are dropped (lib/mkh.c:109, sys/common.h:88)

// One axis tap pair of a forward interpolation: nodes n and n+1 with
// weights w0, w1 (0 where the node is off the grid).
struct Tap2 {
    int n;
    int pad_;
    double w0, w1;
};
// One entry of a transposed axis list: output-grid index and weight.
struct TapT {
    int i;
    int pad_;
    double w;
};

// Everything the kernel reads of one plan (device resident, read through
// scalar loads as kernel arguments are).
struct FitDev {
    int nfit, ndm, npsr, nnw;
    int fnx, fny;
    long nfloc, nact, nxtot;
    int anx[FMAXDM], any[FMAXDM];
    long aoff[FMAXDM + 1];
    int xnx[FMAXL], xny[FMAXL];
    long xoff[FMAXL + 1];
    double wt[FMAXFIT];
    // forward taps: HA (f, dm) and HX (f, layer), per fitting-grid column / row
    const Tap2 *hax, *hay;  // [(f*ndm + d)*fnx + px], [(f*ndm + d)*fny + py]
    const Tap2 *hxx, *hxy;  // [(f*npsr + l)*fnx + px], ...
    // transposed HA lists: per (f, d) and actuator column ia (row ja) the
    // entries [tp[b + ia], tp[b + ia + 1]) of te, b = txb[f*ndm + d]
    const int *txb, *tyb;
    const int *txp, *typ;
    const TapT *txe, *tye;
    const double *W0, *W1;  // 9 per point (row form), per point
    const double *NW;       // nact x nnw
    const long *slp, *sli;  // S in CSR (slp null: none)
    const double *slx;
    double tikdiag;
    // per-system workspace, stride given by the fields below
    double *opd, *opd2;     // [s][nfit][nfloc]
    double *r, *p, *Ap, *b; // [s][nact]
    double *res;            // [s]
};

enum { FIT_SOLVE = 0, FIT_R = 1, FIT_L = 2 };

__device__ double fit_bsum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sh[i];
        sh[FTHREADS / 64] = s;
    }
    __syncthreads();
    return sh[FTHREADS / 64];
}

// opd_f = sum over sources (DMs or layers) of the bilinear interpolation
// at the fitting-grid points; v: stacked source vector (double or float).
template <typename V>
__device__ void fwd_interp(const FitDev &D, const Tap2 *tx, const Tap2 *ty, int nsrc, const int *snx, const long *soff,
                           const V *v, double *opd) {
    const int fnx = D.fnx;
    const long nf = D.nfloc;
    for (long t = threadIdx.x; t < (long)D.nfit * nf; t += blockDim.x) {
        const int f = (int)(t / nf);
        const long pt = t - (long)f * nf;
        const int py = (int)(pt / fnx), px = (int)(pt - (long)py * fnx);
        double acc = 0;
        for (int d = 0; d < nsrc; d++) {
            const Tap2 a = tx[((long)f * nsrc + d) * fnx + px];
            const Tap2 b = ty[((long)f * nsrc + d) * D.fny + py];
            const V *src = v + soff[d];
            const int nx = snx[d];
            const double wx[2] = {a.w0, a.w1}, wy[2] = {b.w0, b.w1};
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const double w = wx[i] * wy[j];
                    if (w > FEPS) acc += w * (double)src[(long)(b.n + j) * nx + a.n + i];
                }
        }
        opd[t] = acc;
    }
}

// opd2_f = wt_f (W0 opd_f - W1 (W1' opd_f)) (applyW, maos/recon_utils.c:155-180)
__device__ void apply_w(const FitDev &D, const double *opd, double *opd2, double *sh) {
    const int fnx = D.fnx, fny = D.fny;
    const long nf = D.nfloc;
    for (int f = 0; f < D.nfit; f++) {
        const double *o = opd + (long)f * nf;
        double dot = 0;
        for (long i = threadIdx.x; i < nf; i += blockDim.x) dot += D.W1[i] * o[i];
        dot = fit_bsum(dot, sh);
        const double wf = D.wt[f];
        for (long i = threadIdx.x; i < nf; i += blockDim.x) {
            const int iy = (int)(i / fnx), ix = (int)(i - (long)iy * fnx);
            const double *w0 = D.W0 + 9 * i;
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
                const double c = w0[k];
                if (c != 0.0 && jx >= 0 && jx < fnx && jy >= 0 && jy < fny) acc += c * o[(long)jy * fnx + jx];
            }
            opd2[(long)f * nf + i] = wf * (acc - D.W1[i] * dot);
        }
    }
}

// out[a] = alpha * sum_f HA_f' opd2_f at actuator a (+ beta * out[a])
__device__ void hat_gather(const FitDev &D, const double *opd2, double *out, double alpha) {
    const long nf = D.nfloc;
    const int fnx = D.fnx;
    for (long a = threadIdx.x; a < D.nact; a += blockDim.x) {
        int d = 0;
        while (a >= D.aoff[d + 1]) d++;
        const long k = a - D.aoff[d];
        const int ja = (int)(k / D.anx[d]), ia = (int)(k - (long)ja * D.anx[d]);
        double acc = 0;
        for (int f = 0; f < D.nfit; f++) {
            const int fd = f * D.ndm + d;
            const int bx = D.txb[fd] + ia, by = D.tyb[fd] + ja;
            const int x0 = D.txp[bx], x1 = D.txp[bx + 1];
            const int y0 = D.typ[by], y1 = D.typ[by + 1];
            const double *o = opd2 + (long)f * nf;
            for (int ey = y0; ey < y1; ey++) {
                const TapT ty = D.tye[ey];
                for (int ex = x0; ex < x1; ex++) {
                    const TapT tx = D.txe[ex];
                    const double w = tx.w * ty.w;
                    if (w > FEPS) acc += w * o[(long)ty.i * fnx + tx.i];
                }
            }
        }
        out[a] = alpha * acc;
    }
}

// Ap = FitL p: HA, W, HA', then the low-rank, slaving and Tikhonov terms.
// Returns p.Ap.
__device__ double apply_L(const FitDev &D, const double *p, double *Ap, double *opd, double *opd2, double *sh,
                          double *tnw) {
    fwd_interp(D, D.hax, D.hay, D.ndm, D.anx, D.aoff, p, opd);
    __syncthreads();
    apply_w(D, opd, opd2, sh);
    __syncthreads();
    hat_gather(D, opd2, Ap, 1.0);
    // NW t, t = NW' p (dots over all actuators, in a fixed order; t in LDS)
    for (int k = 0; k < D.nnw; k++) {
        double s = 0;
        const double *c = D.NW + (long)k * D.nact;
        for (long a = threadIdx.x; a < D.nact; a += blockDim.x) s += c[a] * p[a];
        s = fit_bsum(s, sh);
        if (threadIdx.x == 0) tnw[k] = s;
    }
    __syncthreads();
    double pap = 0;
    for (long a = threadIdx.x; a < D.nact; a += blockDim.x) {
        double v = Ap[a];
        for (int k = 0; k < D.nnw; k++) v += D.NW[(long)k * D.nact + a] * tnw[k];
        if (D.slp) {
            double s = 0;
            for (long e = D.slp[a]; e < D.slp[a + 1]; e++) s += D.slx[e] * p[D.sli[e]];
            v += s;
        }
        v += D.tikdiag * p[a];
        Ap[a] = v;
        pap += p[a] * v;
    }
    return fit_bsum(pap, sh);
}

// One workgroup per system.  MODE FIT_R: b = FitR xin (into b, and out if
// given); FIT_L: out = FitL ain; FIT_SOLVE: the whole CG into x.
template <int MODE>
__global__ __launch_bounds__(FTHREADS) void k_fit(const FitDev *__restrict__ Dp, const float *__restrict__ xin,
                                                  long xstride, double *__restrict__ xout,
                                                  const double *__restrict__ ain, int s0, int warm, int maxit) {
    const FitDev &D = *Dp;
    __shared__ double sh[FTHREADS / 64 + 1];
    __shared__ double tnw[FMAXNW];
    const int s = s0 + blockIdx.x;
    const long nf = D.nfit * D.nfloc, na = D.nact;
    double *opd = D.opd + (long)s * nf, *opd2 = D.opd2 + (long)s * nf;
    double *r = D.r + (long)s * na, *p = D.p + (long)s * na, *Ap = D.Ap + (long)s * na, *b = D.b + (long)s * na;
    if (MODE == FIT_L) {
        (void)apply_L(D, ain + (long)blockIdx.x * na, xout + (long)blockIdx.x * na, opd, opd2, sh, tnw);
        return;
    }
    // RHS: b = sum_f HA_f' W HX_f x
    fwd_interp(D, D.hxx, D.hxy, D.npsr, D.xnx, D.xoff, xin + (long)blockIdx.x * xstride, opd);
    __syncthreads();
    apply_w(D, opd, opd2, sh);
    __syncthreads();
    hat_gather(D, opd2, b, 1.0);
    __syncthreads();
    if (MODE == FIT_R) {
        for (long a = threadIdx.x; a < na; a += blockDim.x) xout[(long)blockIdx.x * na + a] = b[a];
        return;
    }
    // CG (lib/pcg.c:28-113, Mmul = NULL: z = r), restart every 100 iterations
    double *x = xout + (long)blockIdx.x * na;
    if (!warm)
        for (long a = threadIdx.x; a < na; a += blockDim.x) x[a] = 0;
    double bb = 0;
    for (long a = threadIdx.x; a < na; a += blockDim.x) bb += b[a] * b[a];
    bb = fit_bsum(bb, sh);
    double rr = 0, res = NAN;
    for (int k = 0; k < maxit; k++) {
        if (k % 100 == 0) {
            // r = b - A x; p = r; rr = r.r
            __syncthreads();
            (void)apply_L(D, x, Ap, opd, opd2, sh, tnw);
            double t = 0;
            for (long a = threadIdx.x; a < na; a += blockDim.x) {
                const double v = b[a] - Ap[a];
                r[a] = v;
                p[a] = v;
                t += v * v;
            }
            rr = fit_bsum(t, sh);
        }
        const double pap = apply_L(D, p, Ap, opd, opd2, sh, tnw);
        const double ak = rr / pap;
        if (!isfinite(ak)) {
            res = NAN;
            break;
        }
        double t = 0;
        for (long a = threadIdx.x; a < na; a += blockDim.x) {
            x[a] += ak * p[a];
            const double v = r[a] - ak * Ap[a];
            r[a] = v;
            t += v * v;
        }
        const double rr2 = fit_bsum(t, sh);
        if (k + 1 == maxit) {
            res = sqrt(rr2 / bb);
            break;
        }
        const double bk = rr2 / rr;
        for (long a = threadIdx.x; a < na; a += blockDim.x) p[a] = bk * p[a] + r[a];
        rr = rr2;
        __syncthreads();
    }
    if (threadIdx.x == 0) D.res[s] = res;
}

// Forward taps of one axis: output coordinates u_k (in source grid units).
std::vector<Tap2> axis_taps(const std::vector<double> &u, long nsrc) {
    std::vector<Tap2> t(u.size());
    for (size_t k = 0; k < u.size(); k++) {
        const double fl = std::floor(u[k]);
        const long n0 = (long)fl;
        const double f = u[k] - fl;
        Tap2 &a = t[k];
        a.n = (int)n0;
        a.pad_ = 0;
        a.w0 = (n0 >= 0 && n0 < nsrc) ? 1 - f : 0;
        a.w1 = (n0 + 1 >= 0 && n0 + 1 < nsrc) ? f : 0;
        // weights of off-grid nodes are zero; keep the node index in range
        // so the (never used) address stays inside the source
        if (a.w0 == 0 && a.w1 == 0) a.n = 0;
        else if (a.n < 0) a.n = -1;
    }
    return t;
}

}  // namespace

struct gpu_fit {
    int device = 0, nbatch = 1, maxit = 10, warm = 0;
    hipStream_t stream = nullptr;
    FitDev H{};
    DBuf<FitDev> dev;
    DBuf<Tap2> hax, hay, hxx, hxy;
    DBuf<int> txb, tyb, txp, typ;
    DBuf<TapT> txe, tye;
    DBuf<double> W0, W1, NW, slx;
    DBuf<long> slp, sli;
    DBuf<double> opd, opd2, r, p, Ap, b, res, x;
    DBuf<float> xin;
    std::vector<float> hstage;
};

namespace {

template <typename F>
int fguarded(F &&f) {
    try {
        f();
        return MT_OK;
    } catch (const FitErr &e) {
        mt::set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        mt::set_error("out of host memory");
        return MT_ENOMEM;
    } catch (const std::exception &e) {
        mt::set_error(e.what());
        return MT_EINVAL;
    }
}

void build_fit_plan(gpu_fit *g, const mt_recon_t *R, const mt_fit_recon_t *F) {
    FitDev &D = g->H;
    if (F->ndm <= 0 || F->ndm > FMAXDM) ffail(MT_EINVAL, "gpu_fit_setup: 1 to 8 DMs");
    if (F->nfit <= 0 || F->nfit > FMAXFIT) ffail(MT_EINVAL, "gpu_fit_setup: 1 to 32 fitting directions");
    if (R->npsr <= 0 || R->npsr > FMAXL) ffail(MT_EINVAL, "gpu_fit_setup: 1 to 16 layers");
    if (F->nnw < 0 || F->nnw > FMAXNW) ffail(MT_EINVAL, "gpu_fit_setup: at most 32 low-rank terms");
    D.nfit = F->nfit;
    D.ndm = F->ndm;
    D.npsr = R->npsr;
    D.nnw = F->nnw;
    D.fnx = (int)F->fnx;
    D.fny = (int)F->fny;
    D.nfloc = F->fnx * F->fny;
    D.aoff[0] = 0;
    for (int d = 0; d < D.ndm; d++) {
        D.anx[d] = (int)F->anx[d];
        D.any[d] = (int)F->any[d];
        D.aoff[d + 1] = D.aoff[d] + F->anx[d] * F->any[d];
    }
    D.nact = D.aoff[D.ndm];
    D.xoff[0] = 0;
    for (int l = 0; l < D.npsr; l++) {
        D.xnx[l] = (int)R->xnx[l];
        D.xny[l] = (int)R->xny[l];
        D.xoff[l + 1] = D.xoff[l] + R->xnx[l] * R->xny[l];
    }
    D.nxtot = D.xoff[D.npsr];
    for (int f = 0; f < D.nfit; f++) D.wt[f] = F->wt[f];
    // forward taps, as mkh computes the interpolation point: source-grid
    // coordinate = (x * scale + theta * ht - origin) / dx
    const long fnx = F->fnx, fny = F->fny;
    std::vector<Tap2> hax, hay, hxx, hxy;
    std::vector<int> txb, tyb, txp, typ;
    std::vector<TapT> txe, tye;
    for (int f = 0; f < D.nfit; f++) {
        for (int d = 0; d < D.ndm; d++) {
            const double ht = F->aht[d], sc = 1. - ht / F->hs[f];
            std::vector<double> ux(fnx), uy(fny);
            for (long i = 0; i < fnx; i++) ux[i] = ((F->fox + i * F->fdx) * sc + F->thetax[f] * ht - F->aox[d]) / F->adx[d];
            for (long i = 0; i < fny; i++) uy[i] = ((F->foy + i * F->fdx) * sc + F->thetay[f] * ht - F->aoy[d]) / F->adx[d];
            const std::vector<Tap2> tx = axis_taps(ux, F->anx[d]), ty = axis_taps(uy, F->any[d]);
            hax.insert(hax.end(), tx.begin(), tx.end());
            hay.insert(hay.end(), ty.begin(), ty.end());
            // transposed lists per actuator column / row
            auto transpose = [](const std::vector<Tap2> &t, long nsrc, std::vector<int> &bases, std::vector<int> &ptr,
                                std::vector<TapT> &ent) {
                std::vector<std::vector<TapT>> lists(nsrc);
                for (size_t k = 0; k < t.size(); k++) {
                    if (t[k].w0 != 0) lists[t[k].n].push_back(TapT{(int)k, 0, t[k].w0});
                    if (t[k].w1 != 0) lists[t[k].n + 1].push_back(TapT{(int)k, 0, t[k].w1});
                }
                bases.push_back((int)ptr.size());
                for (long a = 0; a < nsrc; a++) {
                    ptr.push_back((int)ent.size());
                    ent.insert(ent.end(), lists[a].begin(), lists[a].end());
                }
                ptr.push_back((int)ent.size());
            };
            transpose(tx, F->anx[d], txb, txp, txe);
            transpose(ty, F->any[d], tyb, typ, tye);
        }
        for (int l = 0; l < D.npsr; l++) {
            const double ht = R->ht[l], sc = 1. - ht / F->hs[f];
            std::vector<double> ux(fnx), uy(fny);
            for (long i = 0; i < fnx; i++) ux[i] = ((F->fox + i * F->fdx) * sc + F->thetax[f] * ht - R->xox[l]) / R->xdx[l];
            for (long i = 0; i < fny; i++) uy[i] = ((F->foy + i * F->fdx) * sc + F->thetay[f] * ht - R->xoy[l]) / R->xdx[l];
            const std::vector<Tap2> tx = axis_taps(ux, R->xnx[l]), ty = axis_taps(uy, R->xny[l]);
            hxx.insert(hxx.end(), tx.begin(), tx.end());
            hxy.insert(hxy.end(), ty.begin(), ty.end());
        }
    }
    if (txe.size() > (size_t)INT32_MAX || tye.size() > (size_t)INT32_MAX) ffail(MT_EINVAL, "fitting tables too large");
    g->hax.upload(hax);
    g->hay.upload(hay);
    g->hxx.upload(hxx);
    g->hxy.upload(hxy);
    g->txb.upload(txb);
    g->tyb.upload(tyb);
    g->txp.upload(txp);
    g->typ.upload(typ);
    g->txe.upload(txe);
    g->tye.upload(tye);
    g->W0.upload(std::vector<double>(F->W0, F->W0 + 9 * D.nfloc));
    g->W1.upload(std::vector<double>(F->W1, F->W1 + D.nfloc));
    if (D.nnw) g->NW.upload(std::vector<double>(F->NW, F->NW + (size_t)D.nnw * D.nact));
    if (F->sl_p) {
        std::vector<long> p(F->sl_p, F->sl_p + D.nact + 1);
        const long nz = p[D.nact];
        g->slp.upload(p);
        g->sli.upload(std::vector<long>(F->sl_i, F->sl_i + nz));
        g->slx.upload(std::vector<double>(F->sl_x, F->sl_x + nz));
    }
    D.hax = g->hax.p;
    D.hay = g->hay.p;
    D.hxx = g->hxx.p;
    D.hxy = g->hxy.p;
    D.txb = g->txb.p;
    D.tyb = g->tyb.p;
    D.txp = g->txp.p;
    D.typ = g->typ.p;
    D.txe = g->txe.p;
    D.tye = g->tye.p;
    D.W0 = g->W0.p;
    D.W1 = g->W1.p;
    D.NW = g->NW.p;
    D.slp = g->slp.p;
    D.sli = g->sli.p;
    D.slx = g->slx.p;
    D.tikdiag = F->tikdiag;
    // workspace
    const long ns = g->nbatch;
    g->opd.alloc(ns * D.nfit * D.nfloc);
    g->opd2.alloc(ns * D.nfit * D.nfloc);
    g->r.alloc(ns * D.nact);
    g->p.alloc(ns * D.nact);
    g->Ap.alloc(ns * D.nact);
    g->b.alloc(ns * D.nact);
    g->res.alloc(ns);
    g->x.alloc(ns * D.nact);
    g->xin.alloc(ns * D.nxtot);
    FCHK(hipMemset(g->x.p, 0, sizeof(double) * ns * D.nact));
    D.opd = g->opd.p;
    D.opd2 = g->opd2.p;
    D.r = g->r.p;
    D.p = g->p.p;
    D.Ap = g->Ap.p;
    D.b = g->b.p;
    D.res = g->res.p;
    g->dev.alloc(1);
    FCHK(hipMemcpy(g->dev.p, &D, sizeof(FitDev), hipMemcpyHostToDevice));
}

void stage_x(gpu_fit *g, int isys, const double *opdr) {
    const long n = g->H.nxtot;
    g->hstage.resize(n);
    for (long i = 0; i < n; i++) {
        if (!std::isfinite(opdr[i])) ffail(MT_ENOTFINITE, "non-finite tomography input");
        g->hstage[i] = (float)opdr[i];
    }
    FCHK(hipMemcpyAsync(g->xin.p + (long)isys * n, g->hstage.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                        g->stream));
}

}  // namespace

extern "C" {

int gpu_fit_setup(const mt_recon_t *recon, const mt_fit_recon_t *fit, int device, int nbatch, gpu_fit_t **out) {
    if (!out) return MT_EINVAL;
    *out = nullptr;
    if (!recon || !fit || nbatch <= 0 || fit->maxit <= 0 || fit->maxit > 100000) {
        mt::set_error("gpu_fit_setup: invalid arguments");
        return MT_EINVAL;
    }
    std::unique_ptr<gpu_fit> g(new (std::nothrow) gpu_fit());
    if (!g) return MT_ENOMEM;
    int rc = fguarded([&] {
        int ndev = 0;
        FCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) ffail(MT_EINVAL, "gpu_fit_setup: no such device");
        g->device = device;
        FCHK(hipSetDevice(device));
        g->nbatch = nbatch;
        g->maxit = fit->maxit;
        g->warm = fit->warm ? 1 : 0;
        build_fit_plan(g.get(), recon, fit);
        FCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        FCHK(hipDeviceSynchronize());
    });
    if (rc == MT_OK) *out = g.release();
    return rc;
}

int gpu_fit_solve(gpu_fit_t *g, int isys, const double *opdr, double *dmfit, double *cgres) {
    if (!g || !opdr || !dmfit || isys < 0 || isys >= g->nbatch) return MT_EINVAL;
    double res = NAN;
    int rc = fguarded([&] {
        FCHK(hipSetDevice(g->device));
        stage_x(g, isys, opdr);
        const long na = g->H.nact;
        hipLaunchKernelGGL(k_fit<FIT_SOLVE>, dim3(1), dim3(FTHREADS), 0, g->stream, g->dev.p,
                           g->xin.p + (long)isys * g->H.nxtot, g->H.nxtot, g->x.p + (long)isys * na,
                           (const double *)nullptr, isys, g->warm, g->maxit);
        FCHK(hipGetLastError());
        FCHK(hipMemcpyAsync(dmfit, g->x.p + (long)isys * na, sizeof(double) * na, hipMemcpyDeviceToHost, g->stream));
        FCHK(hipMemcpyAsync(&res, g->res.p + isys, sizeof(double), hipMemcpyDeviceToHost, g->stream));
        FCHK(hipStreamSynchronize(g->stream));
    });
    if (cgres) *cgres = res;
    if (rc == MT_OK && !(res <= 1.0)) rc = MT_NOTCONV;
    return rc;
}

int gpu_fit_solve_dev(gpu_fit_t *g, const float *d_opdr, double *d_dmfit, int nsys, void *stream) {
    if (!g || !d_opdr || !d_dmfit || nsys <= 0 || nsys > g->nbatch) return MT_EINVAL;
    return fguarded([&] {
        FCHK(hipSetDevice(g->device));
        hipStream_t st = stream ? (hipStream_t)stream : g->stream;
        hipLaunchKernelGGL(k_fit<FIT_SOLVE>, dim3(nsys), dim3(FTHREADS), 0, st, g->dev.p, d_opdr, g->H.nxtot,
                           d_dmfit, (const double *)nullptr, 0, g->warm, g->maxit);
        FCHK(hipGetLastError());
    });
}

int gpu_fit_R(gpu_fit_t *g, const double *opdr, double *out, double alpha) {
    if (!g || !opdr || !out) return MT_EINVAL;
    return fguarded([&] {
        FCHK(hipSetDevice(g->device));
        stage_x(g, 0, opdr);
        const long na = g->H.nact;
        hipLaunchKernelGGL(k_fit<FIT_R>, dim3(1), dim3(FTHREADS), 0, g->stream, g->dev.p, g->xin.p, g->H.nxtot,
                           g->Ap.p, (const double *)nullptr, 0, 0, 0);
        FCHK(hipGetLastError());
        std::vector<double> h(na);
        FCHK(hipMemcpyAsync(h.data(), g->Ap.p, sizeof(double) * na, hipMemcpyDeviceToHost, g->stream));
        FCHK(hipStreamSynchronize(g->stream));
        for (long i = 0; i < na; i++) out[i] = alpha * h[i];
    });
}

int gpu_fit_L(gpu_fit_t *g, const double *a, double *out, double alpha, double beta) {
    if (!g || !a || !out) return MT_EINVAL;
    return fguarded([&] {
        FCHK(hipSetDevice(g->device));
        const long na = g->H.nact;
        for (long i = 0; i < na; i++)
            if (!std::isfinite(a[i])) ffail(MT_ENOTFINITE, "non-finite input vector");
        // the input goes to the p slot of system 0, the output to its r slot
        FCHK(hipMemcpyAsync(g->p.p, a, sizeof(double) * na, hipMemcpyHostToDevice, g->stream));
        hipLaunchKernelGGL(k_fit<FIT_L>, dim3(1), dim3(FTHREADS), 0, g->stream, g->dev.p, (const float *)nullptr,
                           0L, g->r.p, (const double *)g->p.p, 0, 0, 0);
        FCHK(hipGetLastError());
        std::vector<double> h(na);
        FCHK(hipMemcpyAsync(h.data(), g->r.p, sizeof(double) * na, hipMemcpyDeviceToHost, g->stream));
        FCHK(hipStreamSynchronize(g->stream));
        for (long i = 0; i < na; i++) out[i] = beta * out[i] + alpha * h[i];
    });
}

int gpu_fit_residuals(gpu_fit_t *g, double *cgres, int nsys) {
    if (!g || !cgres || nsys <= 0 || nsys > g->nbatch) return MT_EINVAL;
    return fguarded([&] {
        FCHK(hipSetDevice(g->device));
        FCHK(hipStreamSynchronize(g->stream));
        FCHK(hipMemcpy(cgres, g->res.p, sizeof(double) * nsys, hipMemcpyDeviceToHost));
    });
}

void gpu_fit_free(gpu_fit_t *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    (void)hipDeviceSynchronize();
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

}  // extern "C"

int fit_enqueue_cold(gpu_fit_t *g, const float *d_x, long xstride, double *d_dm, int nsys, hipStream_t st) {
    if (!g || !d_x || !d_dm || nsys <= 0) return MT_EINVAL;
    return fguarded([&] {
        const long na = g->H.nact;
        for (int s0 = 0; s0 < nsys; s0 += g->nbatch) {
            const int n = std::min(g->nbatch, nsys - s0);
            hipLaunchKernelGGL(k_fit<FIT_SOLVE>, dim3(n), dim3(FTHREADS), 0, st, g->dev.p, d_x + (long)s0 * xstride,
                               xstride, d_dm + (long)s0 * na, (const double *)nullptr, 0, 0, g->maxit);
            FCHK(hipGetLastError());
        }
    });
}
long fit_nact(const gpu_fit_t *g) { return g->H.nact; }
long fit_nxtot(const gpu_fit_t *g) { return g->H.nxtot; }
int fit_device(const gpu_fit_t *g) { return g->device; }

