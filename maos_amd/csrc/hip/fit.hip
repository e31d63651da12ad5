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
// passes.  Each pass is one launch over the whole GPU, with work items
// (point, direction, system) or (actuator, direction, system), so that a
// thread does one short chain of dependent loads per pass; an iteration is
// five launches (HA, W, HA' + low-rank terms, x/r update, p update) and a
// batch of systems shares every launch.  Scalars (alpha, beta, the W1 and NW
// projections) are fixed-order sums of per-workgroup partials, summed again
// by every workgroup that needs them, so no launch waits on the host and a
// system's bits do not depend on its batch neighbours.  (Running the whole
// solve of a system inside one 1024-thread workgroup, with workgroup
// barriers between the passes, was 20x slower for one system: every thread
// then walks ~35 dependent load chains per pass.)  Every operator is a
// gather:
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
    // entries [tp[b + ia], tp[b + ia + 1]) of te, b = f tpx + tdx[d] (the
    // pointer arrays of the (f, d) pairs in order, anx[d] + 1 entries each)
    int tpx, tpy, tdx[FMAXDM], tdy[FMAXDM];
    const int *txp, *typ;
    const TapT *txe, *tye;
    const double *W0, *W1;  // 9 per point (row form), per point
    const double *NW;       // nact x nnw
    const long *slp, *sli;  // S in CSR (slp null: none)
    const double *slx;
    double tikdiag;
    // launch shape: point chunks of 256 per direction (nc1), actuator chunks
    // of hab actuators x nfit directions (nc3), actuator chunks of 256 (nc4)
    int nc1, nc3, nc4, hab;
    // per-system partials, pstride doubles each: W1 . opd_f [nfit][nc1] at
    // o1, p . Ap [nc3] at o2, NW' p [nnw][nc4] at o3, r . r in two slots
    // [2][nc4] at o4 (iteration parity), b . b [nc4] at o5
    int o1, o2, o3, o4, o5;
    long pstride;
    // per-system workspace
    double *opd, *opd2;     // [s][nfit][nfloc]
    double *tnw;            // [s][FMAXNW]: NW' p of the current p
    double *r, *p, *Ap, *b; // [s][nact]
    double *part;           // [s][pstride]
    int *stop;              // [s]: 1 once a step size was not finite
    double *res;            // [s]
};

constexpr int FB = 256;  // threads per workgroup of every fit kernel

// Sum of v over the 256 threads of a workgroup in a fixed order (lane tree,
// then the four waves in order); the result is valid in thread 0.  Every
// thread of the workgroup must call it.
__device__ double fit_bred(double v, double *wsum) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) s = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    return s;
}

// Sum of the n partials at q in a fixed order (thread-strided sums, then the
// lane tree and the four waves in order), returned to every thread.  Every
// thread of the workgroup must call it; psh is its own LDS (not fit_bred's).
__device__ double psum(const double *q, int n, double *psh) {
    double v = 0;
    for (int i = threadIdx.x; i < n; i += FB) v += q[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();  // the previous call's readers are done with psh
    if ((threadIdx.x & 63) == 0) psh[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((psh[0] + psh[1]) + psh[2]) + psh[3];
}

__device__ __forceinline__ double *fpart(const FitDev &D, int s) { return D.part + (long)s * D.pstride; }

// opd_f = sum over sources (DMs or layers) of the bilinear interpolation at
// the fitting-grid point, and the workgroup's partial of W1 . opd_f.
// Grid (point chunks, directions, systems).  SRC 0: actuator vectors
// (double), 1: layer vectors (float).  Sources are taken GS at a time with
// clamped indexes so that the tap loads of a group issue together.
template <int SRC, int GS>
__global__ __launch_bounds__(FB) void kf_fwd(const FitDev *__restrict__ Dp, const void *__restrict__ srcv,
                                              long sstride, int ws0) {
    const FitDev &D = *Dp;
    __shared__ double wsum[4];
    const int f = blockIdx.y, s = ws0 + blockIdx.z;
    const long nf = D.nfloc;
    const long pt = (long)blockIdx.x * FB + threadIdx.x;
    const bool on = pt < nf;
    const long ptc = on ? pt : nf - 1;
    // 32-bit index arithmetic (a 64-bit division is a long instruction
    // sequence; every grid here is far below 2^31 points)
    const int py = (int)ptc / D.fnx, px = (int)ptc - py * D.fnx;
    const int nsrc = SRC ? D.npsr : D.ndm;
    const Tap2 *tx = SRC ? D.hxx : D.hax, *ty = SRC ? D.hxy : D.hay;
    const int *snx = SRC ? D.xnx : D.anx;
    const long *soff = SRC ? D.xoff : D.aoff;
    double acc = 0;
    for (int d0 = 0; d0 < nsrc; d0 += GS) {
        Tap2 a[GS], b[GS];
#pragma unroll
        for (int j = 0; j < GS; j++) {
            const int d = min(d0 + j, nsrc - 1);
            a[j] = tx[((long)f * nsrc + d) * D.fnx + px];
            b[j] = ty[((long)f * nsrc + d) * D.fny + py];
        }
        double v[GS][4];
#pragma unroll
        for (int j = 0; j < GS; j++) {
            const int d = min(d0 + j, nsrc - 1);
            const long base = soff[d] + (long)blockIdx.z * sstride;
            const int nx = snx[d];
            const long i00 = base + (long)b[j].n * nx + a[j].n;
            // taps of zero weight point at node 0 or -1 of an axis; clamp the
            // address into the source so the load stays in bounds
            const long lo = base, hi = base + (long)snx[d] * (SRC ? D.xny[d] : D.any[d]) - 1;
            const long k[4] = {i00, i00 + 1, i00 + nx, i00 + nx + 1};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long kk = k[q] < lo ? lo : (k[q] > hi ? hi : k[q]);
                v[j][q] = SRC ? (double)((const float *)srcv)[kk] : ((const double *)srcv)[kk];
            }
        }
#pragma unroll
        for (int j = 0; j < GS; j++) {
            if (d0 + j >= nsrc) break;
            const double wx[2] = {a[j].w0, a[j].w1}, wy[2] = {b[j].w0, b[j].w1};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double w = wx[q & 1] * wy[q >> 1];
                if (w > FEPS) acc += w * v[j][q];
            }
        }
    }
    if (on) D.opd[((long)s * D.nfit + f) * nf + pt] = acc;
    const double pw = fit_bred(on ? D.W1[pt] * acc : 0.0, wsum);
    if (threadIdx.x == 0) fpart(D, s)[D.o1 + f * D.nc1 + blockIdx.x] = pw;
}

// opd2_f = wt_f (W0 opd_f - W1 (W1' opd_f)) (applyW, maos/recon_utils.c:155-180).
// With nwdots, the first workgroup of each system also sums the partials of
// NW' p once for the kf_hat that follows (instead of every kf_hat workgroup:
// those sums were a quarter of its time).
__global__ __launch_bounds__(FB) void kf_w(const FitDev *__restrict__ Dp, int ws0, int nwdots) {
    const FitDev &D = *Dp;
    __shared__ double psh[4];
    const int f = blockIdx.y, s = ws0 + blockIdx.z;
    const long nf = D.nfloc;
    const long i = (long)blockIdx.x * FB + threadIdx.x;
    if (nwdots && blockIdx.x == 0 && blockIdx.y == 0)
        for (int k = 0; k < D.nnw; k++) {
            const double v = psum(fpart(D, s) + D.o3 + k * D.nc4, D.nc4, psh);
            if (threadIdx.x == 0) D.tnw[(long)s * FMAXNW + k] = v;
        }
    const double dot = psum(fpart(D, s) + D.o1 + f * D.nc1, D.nc1, psh);
    if (i >= nf) return;
    const int fnx = D.fnx, fny = D.fny;
    const int iy = (int)i / fnx, ix = (int)i - iy * fnx;
    const double *o = D.opd + ((long)s * D.nfit + f) * nf;
    const double *w0 = D.W0 + 9 * i;
    double acc = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
        const double c = w0[k];
        if (c != 0.0 && jx >= 0 && jx < fnx && jy >= 0 && jy < fny) acc += c * o[(long)jy * fnx + jx];
    }
    D.opd2[((long)s * D.nfit + f) * nf + i] = D.wt[f] * (acc - D.W1[i] * dot);
}

enum { HAT_R = 0, HAT_L = 1 };
constexpr int HAH = 2;  // actuator groups per kf_hat workgroup (each thread gathers for HAH actuators)

// out[a] = sum_f HA_f' opd2_f at actuator a; HAT_L adds NW (NW' p) + S p +
// tikdiag p and writes the workgroup's partial of p . out.  A workgroup holds
// HAH groups of hab actuators x nfit directions, one thread per (actuator of
// each group, direction) (two actuators per thread halve the workgroups and
// their fixed reduction and barrier cost),
// gathering through the transposed per-axis lists (no atomics); the
// directions of an actuator are summed in index order.  Actuators are the
// fast index of the threads, so a wave's gathers of one direction read
// neighbouring fitting-grid points (with directions fastest, every lane read
// another direction's grid: one cache line per lane and load).  Grid
// (actuator chunks, systems); p and out are per-system vectors of stride nact.
template <int MODE>
__global__ __launch_bounds__(FB) void kf_hat(const FitDev *__restrict__ Dp, const double *__restrict__ p,
                                              double *__restrict__ out, int ws0) {
    const FitDev &D = *Dp;
    __shared__ double red[HAH * FB];
    __shared__ double tnw[FMAXNW];
    __shared__ double wsum[4];
    const int s = ws0 + blockIdx.y;
    const int nfit = D.nfit, t = threadIdx.x, hab = D.hab;
    const int f = t / hab, al = t - f * hab;
    const long na = D.nact, nf = D.nfloc;
    const long a0 = (long)blockIdx.x * HAH * hab;  // first actuator of the workgroup
    if (MODE == HAT_L && t < D.nnw) tnw[t] = D.tnw[(long)s * FMAXNW + t];
    const double *o = D.opd2 + ((long)s * nfit + f) * nf;
    const int fnx = D.fnx;
#pragma unroll
    for (int j = 0; j < HAH; j++) {
        const long a = a0 + j * hab + al;
        double acc = 0;
        if (f < nfit && a < na) {
            int d = 0;
            while (a >= D.aoff[d + 1]) d++;
            const int k = (int)(a - D.aoff[d]);
            const int ja = k / D.anx[d], ia = k - ja * D.anx[d];
            const int bx = f * D.tpx + D.tdx[d] + ia, by = f * D.tpy + D.tdy[d] + ja;
            const int x0 = D.txp[bx], x1 = D.txp[bx + 1];
            const int y0 = D.typ[by], y1 = D.typ[by + 1];
            for (int ey = y0; ey < y1; ey++) {
                const TapT ty = D.tye[ey];
                for (int ex = x0; ex < x1; ex++) {
                    const TapT tx = D.txe[ex];
                    const double w = tx.w * ty.w;
                    if (w > FEPS) acc += w * o[(long)ty.i * fnx + tx.i];
                }
            }
        }
        red[j * FB + t] = acc;
    }
    __syncthreads();
    double pap = 0;
    for (int e = t; e < HAH * hab; e += FB) {
        const int j = e / hab, el = e - j * hab;
        const long a = a0 + e;
        if (a >= na) break;
        double v = 0;
        for (int q = 0; q < nfit; q++) v += red[j * FB + q * hab + el];
        if (MODE == HAT_L) {
            const double *ps = p + (long)blockIdx.y * na;
            for (int q = 0; q < D.nnw; q++) v += D.NW[(long)q * na + a] * tnw[q];
            if (D.slp) {
                double sl = 0;
                for (long e2 = D.slp[a]; e2 < D.slp[a + 1]; e2++) sl += D.slx[e2] * ps[D.sli[e2]];
                v += sl;
            }
            v += D.tikdiag * ps[a];
            pap += ps[a] * v;
        }
        out[(long)blockIdx.y * na + a] = v;
    }
    if (MODE == HAT_L) {
        pap = fit_bred(pap, wsum);
        if (t == 0) fpart(D, s)[D.o2 + blockIdx.x] = pap;
    }
}

// Partials of NW' v (into o3) for a per-system actuator vector v; with
// INIT also zeroes v first when the start is cold, writes the partials of
// b . b (o5) and clears the system's stop flag.  Grid (actuator chunks of
// 256, systems).
template <bool INIT>
__global__ __launch_bounds__(FB) void kf_init(const FitDev *__restrict__ Dp, double *__restrict__ v, int ws0,
                                               int warm) {
    const FitDev &D = *Dp;
    __shared__ double wsum[4];
    const int s = ws0 + blockIdx.y;
    const long na = D.nact;
    const long a = (long)blockIdx.x * FB + threadIdx.x;
    const bool on = a < na;
    double *vs = v + (long)blockIdx.y * na;
    double val = 0;
    if (on) {
        if (INIT && !warm) vs[a] = 0;
        else val = vs[a];
    }
    double *q = fpart(D, s);
    for (int k = 0; k < D.nnw; k++) {
        const double c = on ? D.NW[(long)k * na + a] * val : 0.0;
        const double t = fit_bred(c, wsum);
        if (threadIdx.x == 0) q[D.o3 + k * D.nc4 + blockIdx.x] = t;
        __syncthreads();
    }
    if (INIT) {
        const double bv = on ? D.b[(long)s * na + a] : 0.0;
        const double t = fit_bred(bv * bv, wsum);
        if (threadIdx.x == 0) {
            q[D.o5 + blockIdx.x] = t;
            if (blockIdx.x == 0) D.stop[s] = 0;
        }
    }
}

// CG (re)start at iteration k: r = b - Ap (COLD: r = b, x being zero),
// p = r; partials of r . r into slot k & 1 and of NW' p.
template <bool COLD>
__global__ __launch_bounds__(FB) void kf_restart(const FitDev *__restrict__ Dp, int ws0, int k) {
    const FitDev &D = *Dp;
    __shared__ double wsum[4];
    const int s = ws0 + blockIdx.y;
    const long na = D.nact;
    const long a = (long)blockIdx.x * FB + threadIdx.x;
    const bool on = a < na;
    double v = 0;
    if (on) {
        const long i = (long)s * na + a;
        v = COLD ? D.b[i] : D.b[i] - D.Ap[i];
        D.r[i] = v;
        D.p[i] = v;
    }
    double *q = fpart(D, s);
    const double t = fit_bred(v * v, wsum);
    if (threadIdx.x == 0) q[D.o4 + (k & 1) * D.nc4 + blockIdx.x] = t;
    for (int j = 0; j < D.nnw; j++) {
        __syncthreads();
        const double c = on ? D.NW[(long)j * na + a] * v : 0.0;
        const double u = fit_bred(c, wsum);
        if (threadIdx.x == 0) q[D.o3 + j * D.nc4 + blockIdx.x] = u;
    }
}

// Iteration k, first half: alpha = rr_k / p.Ap; x += alpha p; r -= alpha Ap;
// partials of r . r into slot (k + 1) & 1.  A non-finite alpha stops the
// system (x keeps its value, the residual is NaN).
__global__ __launch_bounds__(FB) void kf_upd(const FitDev *__restrict__ Dp, double *__restrict__ x, int ws0,
                                             int k) {
    const FitDev &D = *Dp;
    __shared__ double wsum[4], psh[4];
    const int s = ws0 + blockIdx.y;
    if (D.stop[s]) return;
    const double *q = fpart(D, s);
    const double rr = psum(q + D.o4 + (k & 1) * D.nc4, D.nc4, psh);
    const double pap = psum(q + D.o2, D.nc3, psh);
    const double ak = rr / pap;
    if (!isfinite(ak)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            D.stop[s] = 1;
            D.res[s] = NAN;
        }
        return;
    }
    const long na = D.nact;
    const long a = (long)blockIdx.x * FB + threadIdx.x;
    double v = 0;
    if (a < na) {
        const long i = (long)s * na + a;
        x[(long)blockIdx.y * na + a] += ak * D.p[i];
        v = D.r[i] - ak * D.Ap[i];
        D.r[i] = v;
    }
    const double t = fit_bred(v * v, wsum);
    if (threadIdx.x == 0) fpart(D, s)[D.o4 + ((k + 1) & 1) * D.nc4 + blockIdx.x] = t;
}

// Iteration k, second half: beta = rr_(k+1) / rr_k; p = beta p + r and the
// partials of NW' p; after the last iteration the residual sqrt(rr / b.b).
__global__ __launch_bounds__(FB) void kf_p(const FitDev *__restrict__ Dp, int ws0, int k, int last) {
    const FitDev &D = *Dp;
    __shared__ double wsum[4], psh[4];
    const int s = ws0 + blockIdx.y;
    if (D.stop[s]) return;
    const double *qs = fpart(D, s);
    const double rr2 = psum(qs + D.o4 + ((k + 1) & 1) * D.nc4, D.nc4, psh);
    if (last) {
        if (blockIdx.x == 0) {
            const double bb = psum(qs + D.o5, D.nc4, psh);
            if (threadIdx.x == 0) D.res[s] = sqrt(rr2 / bb);
        }
        return;
    }
    const double rr = psum(qs + D.o4 + (k & 1) * D.nc4, D.nc4, psh);
    const double bk = rr2 / rr;
    const long na = D.nact;
    const long a = (long)blockIdx.x * FB + threadIdx.x;
    const bool on = a < na;
    double v = 0;
    if (on) {
        const long i = (long)s * na + a;
        v = bk * D.p[i] + D.r[i];
        D.p[i] = v;
    }
    double *q = fpart(D, s);
    for (int j = 0; j < D.nnw; j++) {
        const double c = on ? D.NW[(long)j * na + a] * v : 0.0;
        const double u = fit_bred(c, wsum);
        if (threadIdx.x == 0) q[D.o3 + j * D.nc4 + blockIdx.x] = u;
        __syncthreads();
    }
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
    DBuf<int> txp, typ;
    DBuf<TapT> txe, tye;
    DBuf<double> W0, W1, NW, slx;
    DBuf<long> slp, sli;
    DBuf<double> opd, opd2, tnw, r, p, Ap, b, part, res, x;
    DBuf<int> stop;
    DBuf<float> xin;
    std::vector<float> hstage;
    // captured solves (the ~500 launches of a 100-iteration solve replay with
    // graph-node gaps instead of launch gaps), keyed by their arguments
    struct Graph {
        const float *xin;
        long xstride;
        double *x;
        int ws0, n, warm, maxit;
        hipGraph_t graph;
        hipGraphExec_t exec;
        hipEvent_t done;  // recorded after each launch
    };
    std::vector<Graph> graphs;  // most recently used last
    hipStream_t cap = nullptr;  // capture stream
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
    if (txe.size() > (size_t)INT32_MAX || tye.size() > (size_t)INT32_MAX || D.nfloc * D.nfit >= INT32_MAX ||
        D.nact >= INT32_MAX)
        ffail(MT_EINVAL, "fitting tables too large");
    g->hax.upload(hax);
    g->hay.upload(hay);
    g->hxx.upload(hxx);
    g->hxy.upload(hxy);
    // the pointer arrays' bases follow from the grid sizes
    D.tpx = D.tpy = 0;
    for (int d = 0; d < D.ndm; d++) {
        D.tdx[d] = D.tpx;
        D.tdy[d] = D.tpy;
        D.tpx += D.anx[d] + 1;
        D.tpy += D.any[d] + 1;
    }
    for (int f = 0; f < D.nfit; f++)
        for (int d = 0; d < D.ndm; d++)
            if (txb[f * D.ndm + d] != f * D.tpx + D.tdx[d] || tyb[f * D.ndm + d] != f * D.tpy + D.tdy[d])
                ffail(MT_EINVAL, "fitting tables: unexpected layout");
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
    g->tnw.alloc(ns * FMAXNW);
    D.tnw = g->tnw.p;
    D.nc1 = (int)((D.nfloc + FB - 1) / FB);
    D.hab = FB / D.nfit;
    D.nc3 = (int)((D.nact + (long)HAH * D.hab - 1) / ((long)HAH * D.hab));
    D.nc4 = (int)((D.nact + FB - 1) / FB);
    D.o1 = 0;
    D.o2 = D.o1 + D.nfit * D.nc1;
    D.o3 = D.o2 + D.nc3;
    D.o4 = D.o3 + D.nnw * D.nc4;
    D.o5 = D.o4 + 2 * D.nc4;
    D.pstride = D.o5 + D.nc4;
    g->part.alloc(ns * D.pstride);
    g->stop.alloc(ns);
    FCHK(hipMemset(g->stop.p, 0, sizeof(int) * ns));
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
    D.part = g->part.p;
    D.stop = g->stop.p;
    D.res = g->res.p;
    g->dev.alloc(1);
    FCHK(hipMemcpy(g->dev.p, &D, sizeof(FitDev), hipMemcpyHostToDevice));
}

// Launch sequences.  ws0: first workspace slot; n: systems; external
// vectors are given for the first system of the launch, stride nact (actuator
// vectors) or xstride (layer vectors).
dim3 grid_pts(const FitDev &D, int n) { return dim3(D.nc1, D.nfit, n); }
dim3 grid_hat(const FitDev &D, int n) { return dim3(D.nc3, n); }
dim3 grid_act(const FitDev &D, int n) { return dim3(D.nc4, n); }

// Interpolation from nsrc sources, GS of them per group of loads: 2 for up
// to two sources, 3 for three, else 4 (a partial last group reloads the
// last source).
template <int SRC>
void launch_fwd(const FitDev &D, int nsrc, dim3 grid, hipStream_t st, const FitDev *dev, const void *src, long sstride,
                int ws0) {
    if (nsrc <= 2) hipLaunchKernelGGL((kf_fwd<SRC, 2>), grid, dim3(FB), 0, st, dev, src, sstride, ws0);
    else if (nsrc == 3) hipLaunchKernelGGL((kf_fwd<SRC, 3>), grid, dim3(FB), 0, st, dev, src, sstride, ws0);
    else hipLaunchKernelGGL((kf_fwd<SRC, 4>), grid, dim3(FB), 0, st, dev, src, sstride, ws0);
}

// out = sum_f wt_f HA_f' W HX_f x (FitR)
void enq_rhs(gpu_fit *g, hipStream_t st, const float *xin, long xstride, double *out, int ws0, int n) {
    const FitDev &D = g->H;
    launch_fwd<1>(D, D.npsr, grid_pts(D, n), st, g->dev.p, (const void *)xin, xstride, ws0);
    hipLaunchKernelGGL(kf_w, grid_pts(D, n), dim3(FB), 0, st, g->dev.p, ws0, 0);
    hipLaunchKernelGGL(kf_hat<HAT_R>, grid_hat(D, n), dim3(FB), 0, st, g->dev.p, (const double *)nullptr, out, ws0);
}

// out = FitL p; the partials of NW' p must be in place (kf_init, kf_restart
// or kf_p wrote them).
void enq_L(gpu_fit *g, hipStream_t st, const double *p, double *out, int ws0, int n) {
    const FitDev &D = g->H;
    launch_fwd<0>(D, D.ndm, grid_pts(D, n), st, g->dev.p, (const void *)p, D.nact, ws0);
    hipLaunchKernelGGL(kf_w, grid_pts(D, n), dim3(FB), 0, st, g->dev.p, ws0, 1);
    hipLaunchKernelGGL(kf_hat<HAT_L>, grid_hat(D, n), dim3(FB), 0, st, g->dev.p, p, out, ws0);
}

// The whole solve of n systems (lib/pcg.c:28-113 with Mmul = NULL: z = r,
// restart every 100 iterations): 5 launches per iteration.
void enq_solve(gpu_fit *g, hipStream_t st, const float *xin, long xstride, double *x, int ws0, int n, int warm,
               int maxit) {
    const FitDev &D = g->H;
    const long na = D.nact;
    double *b = g->b.p + (long)ws0 * na, *p = g->p.p + (long)ws0 * na, *Ap = g->Ap.p + (long)ws0 * na;
    enq_rhs(g, st, xin, xstride, b, ws0, n);
    hipLaunchKernelGGL(kf_init<true>, grid_act(D, n), dim3(FB), 0, st, g->dev.p, x, ws0, warm);
    for (int k = 0; k < maxit; k++) {
        if (k % 100 == 0) {
            if (k == 0 && !warm) {
                hipLaunchKernelGGL(kf_restart<true>, grid_act(D, n), dim3(FB), 0, st, g->dev.p, ws0, k);
            } else {
                if (k > 0) hipLaunchKernelGGL(kf_init<false>, grid_act(D, n), dim3(FB), 0, st, g->dev.p, x, ws0, 1);
                enq_L(g, st, x, Ap, ws0, n);
                hipLaunchKernelGGL(kf_restart<false>, grid_act(D, n), dim3(FB), 0, st, g->dev.p, ws0, k);
            }
        }
        enq_L(g, st, p, Ap, ws0, n);
        hipLaunchKernelGGL(kf_upd, grid_act(D, n), dim3(FB), 0, st, g->dev.p, x, ws0, k);
        hipLaunchKernelGGL(kf_p, grid_act(D, n), dim3(FB), 0, st, g->dev.p, ws0, k, (int)(k + 1 == maxit));
    }
    FCHK(hipGetLastError());
}

constexpr size_t FIT_GRAPHS = 8;  // captured solves kept per plan

// Destroy a captured solve once its last launch (which an asynchronous
// caller may not have waited for) has completed.
void drop_graph(gpu_fit::Graph &e) {
    if (e.done) {
        (void)hipEventSynchronize(e.done);
        (void)hipEventDestroy(e.done);
    }
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
}
void drop_graphs(gpu_fit *g) {
    for (auto &e : g->graphs) drop_graph(e);
    g->graphs.clear();
}

// The solve as a replay of a captured graph on st (captured on first use of
// these arguments); direct launches when st is itself being captured.
void run_solve(gpu_fit *g, hipStream_t st, const float *xin, long xstride, double *x, int ws0, int n, int warm,
               int maxit) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FCHK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        enq_solve(g, st, xin, xstride, x, ws0, n, warm, maxit);
        return;
    }
    size_t k = 0;
    for (; k < g->graphs.size(); k++) {
        const auto &e = g->graphs[k];
        if (e.xin == xin && e.xstride == xstride && e.x == x && e.ws0 == ws0 && e.n == n && e.warm == warm &&
            e.maxit == maxit)
            break;
    }
    if (k == g->graphs.size()) {
        if (!g->cap) FCHK(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
        gpu_fit::Graph e{xin, xstride, x, ws0, n, warm, maxit, nullptr, nullptr, nullptr};
        FCHK(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
        try {
            enq_solve(g, g->cap, xin, xstride, x, ws0, n, warm, maxit);
        } catch (...) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(g->cap, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            throw;
        }
        FCHK(hipStreamEndCapture(g->cap, &e.graph));
        hipError_t ie = hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0);
        if (ie == hipSuccess) ie = hipEventCreateWithFlags(&e.done, hipEventDisableTiming);
        if (ie != hipSuccess) {
            drop_graph(e);
            FCHK(ie);
        }
        if (g->graphs.size() >= FIT_GRAPHS) {
            drop_graph(g->graphs.front());
            g->graphs.erase(g->graphs.begin());
        }
        g->graphs.push_back(e);
        k = g->graphs.size() - 1;
    } else if (k + 1 != g->graphs.size()) {
        std::rotate(g->graphs.begin() + k, g->graphs.begin() + k + 1, g->graphs.end());
        k = g->graphs.size() - 1;
    }
    FCHK(hipGraphLaunch(g->graphs[k].exec, st));
    FCHK(hipEventRecord(g->graphs[k].done, st));
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
    if (!recon || !fit || nbatch <= 0 || nbatch > 65535 || fit->maxit <= 0 || fit->maxit > 100000) {
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
        run_solve(g, g->stream, g->xin.p + (long)isys * g->H.nxtot, g->H.nxtot, g->x.p + (long)isys * na, isys, 1,
                  g->warm, g->maxit);
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
        run_solve(g, st, d_opdr, g->H.nxtot, d_dmfit, 0, nsys, g->warm, g->maxit);
    });
}

int gpu_fit_R(gpu_fit_t *g, const double *opdr, double *out, double alpha) {
    if (!g || !opdr || !out) return MT_EINVAL;
    return fguarded([&] {
        FCHK(hipSetDevice(g->device));
        stage_x(g, 0, opdr);
        const long na = g->H.nact;
        enq_rhs(g, g->stream, g->xin.p, g->H.nxtot, g->Ap.p, 0, 1);
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
        hipLaunchKernelGGL(kf_init<false>, grid_act(g->H, 1), dim3(FB), 0, g->stream, g->dev.p, g->p.p, 0, 1);
        enq_L(g, g->stream, g->p.p, g->r.p, 0, 1);
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
    drop_graphs(g);
    if (g->cap) (void)hipStreamDestroy(g->cap);
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
            run_solve(g, st, d_x + (long)s0 * xstride, xstride, d_dm + (long)s0 * na, 0, n, 0, g->maxit);
        }
    });
}
long fit_nact(const gpu_fit_t *g) { return g->H.nact; }
long fit_nxtot(const gpu_fit_t *g) { return g->H.nxtot; }
int fit_device(const gpu_fit_t *g) { return g->device; }

