// All-FP64 tomography solve (precision mode 4): the FDPCG-preconditioned CG
// of (HX' GP' W GP HX + L2' L2) x = HX' GP' W (I - TT PTT) g with every
// vector, operator coefficient, FFT and reduction in FP64 -- the GPU
// counterpart of the reference's CUDA_DOUBLE build (cuda/common.h:39-45),
// whose tomography is cuda/tomo.cu:174-312 and whose PCG is
// cuda/pcg.cu:135-226 / lib/pcg.c:28-113.
//
// It serves the precision study (what FP64 everywhere buys against the FP32
// operator modes) and callers that need FP64 end to end; the FP32-operator
// modes remain the fast path.  Design:
//  - one launch per pass over (work item, system) grids, a batch of systems
//    sharing every launch; gathers only (no atomics): HX from per-axis ray
//    tables, GP and GP' from the CSR / CSC forms of GP, HX' from transposed
//    per-axis ray lists, L2'L2 as two periodic 5-point passes composed at
//    each node;
//  - FDPCG: the full-spectrum 2-D FFT of every layer (rows, then columns)
//    in LDS by a two-factor Cooley-Tukey split n = n1 n2 with direct DFTs
//    of lengths n1 and n2 and one exactly reduced twiddle table per axis,
//    the spectrum permutation and bs x bs block multiply of fdpcg_precond
//    (maos/fdpcg.c:593-726), the inverse permutation as a gather, and the
//    inverse FFT;
//  - scalars as fixed-order sums of per-workgroup partials, summed again by
//    every workgroup that needs them: no host round trip inside the solve,
//    and a system's bits do not depend on its batch neighbours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "maos_tomo.h"
#include "rays.hpp"
#include "solve64.hpp"

namespace mt {
void set_error(const std::string &msg);
}

namespace {

struct Err64 {
    int code;
    std::string msg;
};
#define C64(call)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) throw Err64{MT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
[[noreturn]] void fail64(int code, const std::string &m) { throw Err64{code, m}; }

template <typename T>
struct Buf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        free();
        n = count;
        if (count) C64(hipMalloc((void **)&p, sizeof(T) * count));
    }
    void upload(const std::vector<T> &h) {
        if (h.size() != n) alloc(h.size());
        if (n) C64(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
    void zero() {
        if (n) C64(hipMemset(p, 0, sizeof(T) * n));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~Buf() { free(); }
};

constexpr int B64 = 256;  // threads per workgroup
constexpr int X64W = 16, X64L = 16;
constexpr long IPCONJ = 1L << 62;  // inverse-permutation entry: conjugate the source

// Device record of a plan (read through scalar loads, like kernel arguments).
struct D64 {
    int npsr, nw, pnx, pny;
    long nploc;
    int lnx[X64L], lny[X64L];  // layer grids
    long loff[X64L + 1];       // layer offsets in a system's vector
    int nx, ny;     // FFT size: the layer grid (every layer the same size with FDPCG)
    long nxl;       // nx * ny
    long nxtot, ngtot;
    long goff[X64W + 1];
    int nsa[X64W];
    long saoff[X64W + 1];
    int ttr[X64W];
    long pttoff[X64W];
    int ptt_lhs, precond, fd_scale, bs;
    long nb;
    double l2cf[X64L];
    double fscale;  // 1/sqrt(nx ny) when fd_scale, else 1
    // rays: per (k, l) and ploc index, the two nodes and the fraction of the
    // second (n0 < 0: the ray misses the layer)
    const int2 *rnx, *rny;
    const double *rfx, *rfy;
    // transposed rays: per (k, l) and layer node, the ploc indices and
    // weights that land on it (pointers at txo[k npsr + l] + node)
    const int *txp, *typ, *txi, *tyi;
    const double *txw, *tyw;
    int txo[X64W * X64L], tyo[X64W * X64L];
    // GP of WFS k: rows (CSR, pointer array at rpo[k]) and columns (CSC,
    // pointer array at cpo[k]); entry indices absolute
    const long *grp, *gri, *gcp, *gci;
    const double *grx, *gcx;
    long rpo[X64W], cpo[X64W];
    const double *neai;  // 3 per subaperture (xx, yy, xy), at 3 saoff[k]
    const double *ptt;   // 2 x 2nsa per TT WFS, at pttoff[k]
    // FDPCG
    const double2 *twx, *twy;
    int fx1, fx2, fy1, fy2;  // nx = fx1 fx2, ny = fy1 fy2
    const double2 *Minv;     // nb blocks, bs x bs column-major
    const long *perm;        // nxtot: > 0 source index, <= 0 conjugate of -v
    const long *iperm;       // nxtot: source of each spectral index (| IPCONJ)
    int rrows, rcols;        // rows (columns) per row (column) FFT workgroup
    int nrw, ncw;            // row / column FFT workgroups per system
    int nch, nchg, nchs;     // chunks of 256: nodes, max 2nsa, max nsa
    // partials per system (pstride doubles): TT projections [nw][2][nchg],
    // p.Ap [nch], r.r [nch], r.z [2][nrz] by iteration parity, b.b [nch]
    int o_tt, o_pap, o_rr, o_rz, o_bb, nrz;
    long pstride;
    // state per system
    double *x, *r, *z, *p, *Ap, *b;  // [s][nxtot]
    double *opd, *q;                 // [s][nw][nploc]
    double *gin, *g, *wg;            // [s][ngtot]
    double2 *sa, *sb;                // [s][nxtot]
    double *part;
    int *stop;
    double *res;
};

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ double bred(double v, double *wsum) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) s = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    return s;
}
// Sum of the n partials at q in a fixed order (thread-strided sums, then the
// lane tree and the four waves in order), returned to every thread.  Every
// thread of the workgroup must call it; psh is its own LDS (not bred's).
__device__ double psum(const double *q, int n, double *psh) {
    double v = 0;
    for (int i = threadIdx.x; i < n; i += B64) v += q[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();  // the previous call's readers are done with psh
    if ((threadIdx.x & 63) == 0) psh[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((psh[0] + psh[1]) + psh[2]) + psh[3];
}
__device__ __forceinline__ double *prt(const D64 &D, int s) { return D.part + (long)s * D.pstride; }

// ---------------------------------------------------------------------------
// Operators
// ---------------------------------------------------------------------------

// opd_k = HX_k v (prop_grid_stat, lib/prop_grid.c:47-343).  Grid (ploc
// chunks, WFS, systems); v: per-system layer vectors, stride nxtot.
__global__ __launch_bounds__(B64) void k64_hx(const D64 *__restrict__ Dp, const double *__restrict__ v, int ws0) {
    const D64 &D = *Dp;
    const int k = blockIdx.y, z = blockIdx.z, s = ws0 + z;
    const long pt = (long)blockIdx.x * B64 + threadIdx.x;
    if (pt >= D.nploc) return;
    const int iy = (int)(pt / D.pnx), ix = (int)(pt - (long)iy * D.pnx);
    const double *src = v + (long)z * D.nxtot;
    double acc = 0;
    for (int l = 0; l < D.npsr; l++) {
        const long kl = (long)k * D.npsr + l;
        const int2 a = D.rnx[kl * D.pnx + ix], b = D.rny[kl * D.pny + iy];
        if (a.x < 0 || b.x < 0) continue;
        const double fx = D.rfx[kl * D.pnx + ix], fy = D.rfy[kl * D.pny + iy];
        const double *m = src + D.loff[l];
        const int nx = D.lnx[l];
        const double v00 = m[(long)b.x * nx + a.x], v01 = m[(long)b.x * nx + a.y];
        const double v10 = m[(long)b.y * nx + a.x], v11 = m[(long)b.y * nx + a.y];
        acc += (1 - fy) * ((1 - fx) * v00 + fx * v01) + fy * ((1 - fx) * v10 + fx * v11);
    }
    D.opd[((long)s * D.nw + k) * D.nploc + pt] = acc;
}

// Gradients of WFS k: FROM_OPD g = GP opd_k (into g), else the input
// gradients gin; when TT is removed here (the RHS of a TT WFS, or the LHS
// with ptt_lhs) the workgroup's partials of PTT g (maos/recon_utils.c:
// 129-137).  Grid (chunks of 2nsa, WFS, systems).
template <bool FROM_OPD>
__global__ __launch_bounds__(B64) void k64_grad(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    __shared__ double wsum[4];
    const int k = blockIdx.y, s = ws0 + blockIdx.z;
    const long ng = 2L * D.nsa[k];
    const long i = (long)blockIdx.x * B64 + threadIdx.x;
    const bool on = i < ng;
    double gv = 0;
    if (on) {
        const long o = (long)s * D.ngtot + D.goff[k] + i;
        if (FROM_OPD) {
            const double *opd = D.opd + ((long)s * D.nw + k) * D.nploc;
            const long *rp = D.grp + D.rpo[k];
            for (long e = rp[i]; e < rp[i + 1]; e++) gv += D.grx[e] * opd[D.gri[e]];
            D.g[o] = gv;
        } else {
            gv = D.gin[o];
        }
    }
    const bool tt = D.ttr[k] && (!FROM_OPD || D.ptt_lhs);
    if (tt) {
        const double *P = D.ptt + D.pttoff[k];
        double *q = prt(D, s) + D.o_tt + (long)k * 2 * D.nchg;
        for (int c = 0; c < 2; c++) {
            const double t = bred(on ? P[2 * i + c] * gv : 0.0, wsum);
            if (threadIdx.x == 0) q[c * D.nchg + blockIdx.x] = t;
            __syncthreads();
        }
    }
}

// wg = NEA^-1 (g - TT (PTT g)) per subaperture (nea_apply with its x/y
// cross term).  Grid (chunks of nsa, WFS, systems).
template <bool FROM_OPD>
__global__ __launch_bounds__(B64) void k64_wg(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    __shared__ double psh[4];
    const int k = blockIdx.y, s = ws0 + blockIdx.z;
    const long nsa = D.nsa[k];
    const long i = (long)blockIdx.x * B64 + threadIdx.x;
    double j0 = 0, j1 = 0;
    if (D.ttr[k] && (!FROM_OPD || D.ptt_lhs)) {
        const double *q = prt(D, s) + D.o_tt + (long)k * 2 * D.nchg;
        j0 = psum(q, D.nchg, psh);
        j1 = psum(q + D.nchg, D.nchg, psh);
    }
    if (i >= nsa) return;
    const long o = (long)s * D.ngtot + D.goff[k];
    const double *gsrc = FROM_OPD ? D.g : D.gin;
    const double gx = gsrc[o + i] - j0, gy = gsrc[o + nsa + i] - j1;
    const double *ne = D.neai + 3 * (D.saoff[k] + i);
    D.wg[o + i] = ne[0] * gx + ne[2] * gy;
    D.wg[o + nsa + i] = ne[2] * gx + ne[1] * gy;
}

// q_k = GP_k' wg_k (column gather of the CSC form).  Grid (ploc chunks, WFS,
// systems).
__global__ __launch_bounds__(B64) void k64_gpt(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    const int k = blockIdx.y, s = ws0 + blockIdx.z;
    const long pt = (long)blockIdx.x * B64 + threadIdx.x;
    if (pt >= D.nploc) return;
    const long *cp = D.gcp + D.cpo[k];
    const double *wg = D.wg + (long)s * D.ngtot + D.goff[k];
    double acc = 0;
    for (long e = cp[pt]; e < cp[pt + 1]; e++) acc += D.gcx[e] * wg[D.gci[e]];
    D.q[((long)s * D.nw + k) * D.nploc + pt] = acc;
}

// out = sum_k HX_k' q_k; LHS adds L2' L2 v (two periodic 5-point passes,
// lib/laplacian.c:70-97, composed at the node) and writes the workgroup's
// partial of v . out.  Grid (node chunks, systems); v, out: per-system
// vectors of stride nxtot.
template <bool LHS>
__global__ __launch_bounds__(B64) void k64_hxt(const D64 *__restrict__ Dp, const double *__restrict__ v,
                                               double *__restrict__ out, int ws0) {
    const D64 &D = *Dp;
    __shared__ double wsum[4];
    const int z = blockIdx.y, s = ws0 + z;
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    const bool on = n < D.nxtot;
    double acc = 0, pv = 0;
    if (on) {
        int l = 0;
        while (n >= D.loff[l + 1]) l++;
        const int nx = D.lnx[l], ny = D.lny[l];
        const long rem = n - D.loff[l];
        const int iy = (int)(rem / nx), ix = (int)(rem - (long)iy * nx);
        for (int k = 0; k < D.nw; k++) {
            const int kl = k * D.npsr + l;
            const int x0 = D.txp[D.txo[kl] + ix], x1 = D.txp[D.txo[kl] + ix + 1];
            const int y0 = D.typ[D.tyo[kl] + iy], y1 = D.typ[D.tyo[kl] + iy + 1];
            const double *q = D.q + ((long)s * D.nw + k) * D.nploc;
            for (int ey = y0; ey < y1; ey++) {
                const double wy = D.tyw[ey];
                const long row = (long)D.tyi[ey] * D.pnx;
                for (int ex = x0; ex < x1; ex++) acc += D.txw[ex] * wy * q[row + D.txi[ex]];
            }
        }
        if (LHS) {
            const double cf = D.l2cf[l];
            const double *m = v + (long)z * D.nxtot + D.loff[l];
            auto at = [&](int jx, int jy) { return m[(long)jy * nx + jx]; };
            // t(jx, jy) = cf (0.25 (N + W + E + S) - C) at a node, periodic
            auto t = [&](int jx, int jy) {
                const int xm = jx == 0 ? nx - 1 : jx - 1, xp = jx == nx - 1 ? 0 : jx + 1;
                const int ym = jy == 0 ? ny - 1 : jy - 1, yp = jy == ny - 1 ? 0 : jy + 1;
                return 0.25 * cf * at(jx, ym) + 0.25 * cf * at(xm, jy) - cf * at(jx, jy) + 0.25 * cf * at(xp, jy) +
                       0.25 * cf * at(jx, yp);
            };
            const int xm = ix == 0 ? nx - 1 : ix - 1, xp = ix == nx - 1 ? 0 : ix + 1;
            const int ym = iy == 0 ? ny - 1 : iy - 1, yp = iy == ny - 1 ? 0 : iy + 1;
            acc += 0.25 * cf * t(ix, ym) + 0.25 * cf * t(xm, iy) - cf * t(ix, iy) + 0.25 * cf * t(xp, iy) +
                   0.25 * cf * t(ix, yp);
            pv = at(ix, iy);
        }
        out[(long)z * D.nxtot + n] = acc;
    }
    if (LHS) {
        const double t = bred(pv * acc, wsum);
        if (threadIdx.x == 0) prt(D, s)[D.o_pap + blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// FDPCG
// ---------------------------------------------------------------------------

// R transforms of length n = n1 n2 held in LDS (rows of a, stride n), in
// place, through tmp: X[k1 + n1 k2] = sum_j2 W_n^(j2 k1) W_n2^(j2 k2)
// sum_j1 x[j1 n2 + j2] W_n1^(j1 k1), with W_n^m = tw[m mod n] (conjugated
// for the inverse, unscaled).
__device__ void lds_dft(double2 *a, double2 *tmp, int R, int n, int n1, int n2, const double2 *__restrict__ tw,
                        bool inv) {
    for (int e = threadIdx.x; e < R * n; e += blockDim.x) {
        const int rl = e / n, t = e - rl * n;
        const int k1 = t / n2, j2 = t - k1 * n2;
        const double2 *x = a + rl * n;
        double2 acc = make_double2(0, 0);
        for (int j1 = 0; j1 < n1; j1++) {
            double2 w = tw[((j1 * k1) % n1) * n2];
            if (inv) w.y = -w.y;
            const double2 v = x[j1 * n2 + j2];
            acc.x += v.x * w.x - v.y * w.y;
            acc.y += v.x * w.y + v.y * w.x;
        }
        double2 w2 = tw[j2 * k1];
        if (inv) w2.y = -w2.y;
        tmp[rl * n + k1 * n2 + j2] = cm(acc, w2);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < R * n; e += blockDim.x) {
        const int rl = e / n, t = e - rl * n;
        const int k1 = t % n1, k2 = t / n1;
        const double2 *y = tmp + rl * n + k1 * n2;
        double2 acc = make_double2(0, 0);
        for (int j2 = 0; j2 < n2; j2++) {
            double2 w = tw[((j2 * k2) % n2) * n1];
            if (inv) w.y = -w.y;
            const double2 v = y[j2];
            acc.x += v.x * w.x - v.y * w.y;
            acc.y += v.x * w.y + v.y * w.x;
        }
        a[rl * n + t] = acc;
    }
    __syncthreads();
}

// Row transforms (x axis) of every layer.  FWD: real r into sa; else the
// block-multiplied spectrum sb gathered through the inverse permutation
// (cvecpermi) into sa.  Grid (row groups, systems); LDS 2 rrows nx double2.
template <bool FWD>
__global__ __launch_bounds__(B64) void k64_frow(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    extern __shared__ double2 lds64[];
    const int R = D.rrows, n = D.nx;
    double2 *a = lds64, *tmp = lds64 + R * n;
    const int s = ws0 + blockIdx.y;
    const int nrows = D.npsr * D.ny;
    const int row0 = blockIdx.x * R;
    for (int e = threadIdx.x; e < R * n; e += blockDim.x) {
        const int rl = e / n, ix = e - rl * n;
        const int row = min(row0 + rl, nrows - 1);
        const long j = (long)row * n + ix;  // row = l ny + iy: node index l nxl + iy nx + ix
        double2 v;
        if (FWD) {
            v = make_double2(D.r[(long)s * D.nxtot + j], 0.0);
        } else {
            const long ip = D.iperm[j];
            v = D.sb[(long)s * D.nxtot + (ip & (IPCONJ - 1))];
            if (ip & IPCONJ) v.y = -v.y;
        }
        a[e] = v;
    }
    __syncthreads();
    lds_dft(a, tmp, R, n, D.fx1, D.fx2, D.twx, !FWD);
    for (int e = threadIdx.x; e < R * n; e += blockDim.x) {
        const int rl = e / n, ix = e - rl * n;
        const int row = row0 + rl;
        if (row < nrows) D.sa[(long)s * D.nxtot + (long)row * n + ix] = a[e];
    }
}

// Column transforms (y axis) of sa.  FWD: scaled, in place; else the real
// part scaled into z, and the workgroup's partial of r . z into slot `slot`.
// Grid (column groups, systems); LDS 2 rcols ny double2.
template <bool FWD>
__global__ __launch_bounds__(B64) void k64_fcol(const D64 *__restrict__ Dp, int ws0, int slot) {
    const D64 &D = *Dp;
    extern __shared__ double2 lds64[];
    __shared__ double wsum[4];
    const int C = D.rcols, n = D.ny, nx = D.nx;
    double2 *a = lds64, *tmp = lds64 + C * n;
    const int s = ws0 + blockIdx.y;
    const int ncols = D.npsr * nx;
    const int col0 = blockIdx.x * C;
    double2 *sa = D.sa + (long)s * D.nxtot;
    // element e: column c = e % C of the group, row iy = e / C (columns of a
    // group are adjacent in memory, so each row's loads are contiguous)
    for (int e = threadIdx.x; e < C * n; e += blockDim.x) {
        const int iy = e / C, cl = e - iy * C;
        const int col = min(col0 + cl, ncols - 1);
        const int l = col / nx, ix = col - l * nx;
        a[cl * n + iy] = sa[(long)l * D.nxl + (long)iy * nx + ix];
    }
    __syncthreads();
    lds_dft(a, tmp, C, n, D.fy1, D.fy2, D.twy, !FWD);
    double rz = 0;
    for (int e = threadIdx.x; e < C * n; e += blockDim.x) {
        const int iy = e / C, cl = e - iy * C;
        const int col = col0 + cl;
        if (col >= ncols) continue;
        const int l = col / nx, ix = col - l * nx;
        const long j = (long)l * D.nxl + (long)iy * nx + ix;
        const double2 v = a[cl * n + iy];
        if (FWD) {
            sa[j] = make_double2(v.x * D.fscale, v.y * D.fscale);
        } else {
            const double zv = v.x * D.fscale;
            D.z[(long)s * D.nxtot + j] = zv;
            rz += D.r[(long)s * D.nxtot + j] * zv;
        }
    }
    if (!FWD) {
        const double t = bred(rz, wsum);
        if (threadIdx.x == 0) prt(D, s)[D.o_rz + slot * D.nrz + blockIdx.x] = t;
    }
}

// sb = Minv (perm sa) per bs x bs block (cvecperm, then the block multiply).
// Grid (chunks of nb bs, systems).
__global__ __launch_bounds__(B64) void k64_blk(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    const int s = ws0 + blockIdx.y;
    const long t = (long)blockIdx.x * B64 + threadIdx.x;
    const int bs = D.bs;
    if (t >= D.nb * bs) return;
    const long ib = t / bs;
    const int a = (int)(t - ib * bs);
    const double2 *sa = D.sa + (long)s * D.nxtot;
    const double2 *M = D.Minv + ib * bs * bs;
    double2 y = make_double2(0, 0);
    for (int b = 0; b < bs; b++) {
        const long pm = D.perm[ib * bs + b];
        double2 v = sa[pm > 0 ? pm : -pm];
        if (pm <= 0) v.y = -v.y;
        const double2 m = M[a + (long)b * bs];
        y.x += m.x * v.x - m.y * v.y;
        y.y += m.x * v.y + m.y * v.x;
    }
    D.sb[(long)s * D.nxtot + t] = y;
}

// ---------------------------------------------------------------------------
// CG (lib/pcg.c:28-113): vectors over node chunks, grid (chunks, systems)
// ---------------------------------------------------------------------------

// Start of a solve: x = 0 when cold, partials of b . b, clear the stop flag.
__global__ __launch_bounds__(B64) void k64_init(const D64 *__restrict__ Dp, int ws0, int warm) {
    const D64 &D = *Dp;
    __shared__ double wsum[4];
    const int s = ws0 + blockIdx.y;
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    double bv = 0;
    if (n < D.nxtot) {
        const long i = (long)s * D.nxtot + n;
        if (!warm) D.x[i] = 0;
        bv = D.b[i];
    }
    const double t = bred(bv * bv, wsum);
    if (threadIdx.x == 0) {
        prt(D, s)[D.o_bb + blockIdx.x] = t;
        if (blockIdx.x == 0) D.stop[s] = 0;
    }
}

// (Re)start at iteration k: r = b - Ap (COLD: r = b, x being zero); without
// a preconditioner also p = r and the partials of r . r into slot k & 1.
template <bool COLD>
__global__ __launch_bounds__(B64) void k64_restart(const D64 *__restrict__ Dp, int ws0, int k) {
    const D64 &D = *Dp;
    __shared__ double wsum[4];
    const int s = ws0 + blockIdx.y;
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    double v = 0;
    if (n < D.nxtot) {
        const long i = (long)s * D.nxtot + n;
        v = COLD ? D.b[i] : D.b[i] - D.Ap[i];
        D.r[i] = v;
        if (!D.precond) D.p[i] = v;
    }
    if (!D.precond) {
        const double t = bred(v * v, wsum);
        if (threadIdx.x == 0) prt(D, s)[D.o_rz + (k & 1) * D.nrz + blockIdx.x] = t;
    }
}

// p = z after a (re)start with the preconditioner.
__global__ __launch_bounds__(B64) void k64_pfirst(const D64 *__restrict__ Dp, int ws0) {
    const D64 &D = *Dp;
    const int s = ws0 + blockIdx.y;
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    if (n < D.nxtot) D.p[(long)s * D.nxtot + n] = D.z[(long)s * D.nxtot + n];
}

// alpha = r.z / p.Ap; x += alpha p; r -= alpha Ap; partials of r . r (and,
// without a preconditioner, of r . z = r . r into slot (k + 1) & 1).
__global__ __launch_bounds__(B64) void k64_upd(const D64 *__restrict__ Dp, int ws0, int k) {
    const D64 &D = *Dp;
    __shared__ double wsum[4], psh[4];
    const int s = ws0 + blockIdx.y;
    if (D.stop[s]) return;
    const double *q = prt(D, s);
    const int nrzk = D.precond ? D.ncw : D.nch;
    const double rz = psum(q + D.o_rz + (k & 1) * D.nrz, nrzk, psh);
    const double ak = rz / psum(q + D.o_pap, D.nch, psh);
    if (!isfinite(ak)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            D.stop[s] = 1;
            D.res[s] = NAN;
        }
        return;
    }
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    double v = 0;
    if (n < D.nxtot) {
        const long i = (long)s * D.nxtot + n;
        D.x[i] += ak * D.p[i];
        v = D.r[i] - ak * D.Ap[i];
        D.r[i] = v;
    }
    const double t = bred(v * v, wsum);
    if (threadIdx.x == 0) {
        double *qw = prt(D, s);
        qw[D.o_rr + blockIdx.x] = t;
        if (!D.precond) qw[D.o_rz + ((k + 1) & 1) * D.nrz + blockIdx.x] = t;
    }
}

// beta = r.z_(k+1) / r.z_k; p = beta p + z (z = r without a
// preconditioner); after the last iteration the residual sqrt(r.r / b.b).
__global__ __launch_bounds__(B64) void k64_p(const D64 *__restrict__ Dp, int ws0, int k, int last) {
    const D64 &D = *Dp;
    __shared__ double psh[4];
    const int s = ws0 + blockIdx.y;
    if (D.stop[s]) return;
    const double *q = prt(D, s);
    if (last) {
        if (blockIdx.x == 0) {
            const double rr = psum(q + D.o_rr, D.nch, psh);
            const double bb = psum(q + D.o_bb, D.nch, psh);
            if (threadIdx.x == 0) D.res[s] = sqrt(rr / bb);
        }
        return;
    }
    const int nrzk = D.precond ? D.ncw : D.nch;
    const double rz2 = psum(q + D.o_rz + ((k + 1) & 1) * D.nrz, nrzk, psh);
    const double bk = rz2 / psum(q + D.o_rz + (k & 1) * D.nrz, nrzk, psh);
    const long n = (long)blockIdx.x * B64 + threadIdx.x;
    if (n < D.nxtot) {
        const long i = (long)s * D.nxtot + n;
        D.p[i] = bk * D.p[i] + (D.precond ? D.z[i] : D.r[i]);
    }
}

__global__ void k64_cvt_in(double *dst, const float *src, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}
__global__ void k64_cvt_out(float *dst, const double *src, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

}  // namespace

struct tomo64 {
    int nbatch = 1;
    D64 H{};
    Buf<D64> dev;
    Buf<int2> rnx, rny;
    Buf<double> rfx, rfy;
    Buf<int> txp, typ, txi, tyi;
    Buf<double> txw, tyw;
    Buf<long> grp, gri, gcp, gci;
    Buf<double> grx, gcx, neai, ptt;
    Buf<double2> twx, twy, Minv;
    Buf<long> perm, iperm;
    Buf<double> x, r, z, p, Ap, b, opd, q, gin, g, wg, part, res;
    Buf<double2> sa, sb;
    Buf<int> stop;
    std::vector<long> goff_all;  // caller gradient layout (all WFS)
    std::vector<int> active;     // active WFS -> caller WFS index
    long ng_all = 0;
};

namespace {

template <typename F>
int g64(F &&f) {
    try {
        f();
        return MT_OK;
    } catch (const Err64 &e) {
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

// n = n1 n2 with n1 the largest divisor not above sqrt(n).
void split2(int n, int &n1, int &n2) {
    n1 = 1;
    for (int d = 1; (long)d * d <= n; d++)
        if (n % d == 0) n1 = d;
    n2 = n / n1;
}

std::vector<double2> tw_table(int n) {
    std::vector<double2> t(n);
    for (int k = 0; k < n; k++) {
        const double ph = 2.0 * M_PI * (double)k / (double)n;
        t[k] = make_double2(std::cos(ph), -std::sin(ph));
    }
    return t;
}

void upload_blocks(tomo64 *t, const mt_recon_t *R) {
    D64 &D = t->H;
    for (int l = 0; l < D.npsr; l++) D.l2cf[l] = 4.0 * std::sqrt(R->l2c[l]);
    if (D.precond) {
        const long nbs = D.nb * D.bs * D.bs;
        std::vector<double2> M(nbs);
        for (long i = 0; i < nbs; i++) M[i] = make_double2(R->fd_Minv[2 * i], R->fd_Minv[2 * i + 1]);
        t->Minv.upload(M);
        D.Minv = t->Minv.p;
    }
}

void build64(tomo64 *t, const mt_recon_t *R, int nbatch) {
    D64 &D = t->H;
    if (R->npsr <= 0 || R->npsr > X64L) fail64(MT_EINVAL, "FP64 solve: 1 to 16 layers");
    D.npsr = R->npsr;
    D.precond = R->precond ? 1 : 0;
    D.nx = (int)R->xnx[0];
    D.ny = (int)R->xny[0];
    D.loff[0] = 0;
    for (int l = 0; l < R->npsr; l++) {
        if (R->xnx[l] < 2 || R->xny[l] < 2) fail64(MT_EINVAL, "FP64 solve: layer grid below 2 x 2");
        if (D.precond && (R->xnx[l] != D.nx || R->xny[l] != D.ny))
            fail64(MT_EINVAL, "FP64 solve: FDPCG needs layers of one size");
        D.lnx[l] = (int)R->xnx[l];
        D.lny[l] = (int)R->xny[l];
        D.loff[l + 1] = D.loff[l] + R->xnx[l] * R->xny[l];
    }
    // a row or column of the FFT kernels and its work copy (32 bytes per
    // point) stay within 32 KB of LDS
    if (D.precond && (D.nx > 1024 || D.ny > 1024)) fail64(MT_EINVAL, "FP64 solve: layer grid above 1024");
    D.nxl = (long)D.nx * D.ny;
    D.nxtot = D.loff[D.npsr];
    D.pnx = (int)R->pnx;
    D.pny = (int)R->pny;
    D.nploc = (long)D.pnx * D.pny;
    // WFS
    t->goff_all.assign(R->nwfs + 1, 0);
    for (int w = 0; w < R->nwfs; w++) {
        const long nsa = R->skip[w] ? 0 : R->nsa[w];
        t->goff_all[w + 1] = t->goff_all[w] + 2 * nsa;
        if (!R->skip[w] && R->nsa[w] > 0) t->active.push_back(w);
    }
    t->ng_all = t->goff_all[R->nwfs];
    const int nw = (int)t->active.size();
    if (nw == 0 || nw > X64W) fail64(MT_EINVAL, "FP64 solve: 1 to 16 active WFS");
    D.nw = nw;
    long go = 0, so = 0, po = 0, maxng = 0, maxsa = 0;
    std::vector<double> hptt, hnea;
    for (int k = 0; k < nw; k++) {
        const int w = t->active[k];
        D.nsa[k] = (int)R->nsa[w];
        D.goff[k] = go;
        D.saoff[k] = so;
        go += 2 * R->nsa[w];
        so += R->nsa[w];
        maxng = std::max(maxng, 2 * R->nsa[w]);
        maxsa = std::max(maxsa, R->nsa[w]);
        hnea.insert(hnea.end(), R->neai[w], R->neai[w] + 3 * R->nsa[w]);
        D.ttr[k] = R->trs[w] && R->ptt[w] ? 1 : 0;
        D.pttoff[k] = po;
        if (D.ttr[k]) {
            hptt.insert(hptt.end(), R->ptt[w], R->ptt[w] + 4 * R->nsa[w]);
            po += 4 * R->nsa[w];
        }
    }
    D.goff[nw] = go;
    D.saoff[nw] = so;
    D.ngtot = go;
    D.ptt_lhs = R->ptt_lhs ? 1 : 0;
    D.fd_scale = R->fd_scale ? 1 : 0;
    D.fscale = D.fd_scale ? 1.0 / std::sqrt((double)D.nxl) : 1.0;
    D.bs = R->fd_bs;
    D.nb = R->fd_nb;
    // rays and their transposes
    const long pnx = D.pnx, pny = D.pny;
    std::vector<int2> hnx((size_t)nw * D.npsr * pnx), hny((size_t)nw * D.npsr * pny);
    std::vector<double> hfx(hnx.size()), hfy(hny.size());
    std::vector<int> htxp, htyp, htxi, htyi;
    std::vector<double> htxw, htyw;
    for (int k = 0; k < nw; k++) {
        const int w = t->active[k];
        for (int l = 0; l < D.npsr; l++) {
            const double ht = R->ht[l];
            mtk::Axis ax, ay;
            mtk::ray_axes(R->xdx[l], R->xox[l], R->xoy[l], R->xnx[l], R->xny[l], pnx, pny, R->pox, R->poy, R->pdx,
                          R->thetax[w] * ht, R->thetay[w] * ht, 1. - ht / R->hs[w], ax, ay);
            const size_t kl = (size_t)k * D.npsr + l;
            D.txo[kl] = (int)htxp.size();
            D.tyo[kl] = (int)htyp.size();
            for (long i = 0; i < pnx; i++) {
                hnx[kl * pnx + i] = ax.nn[i];
                hfx[kl * pnx + i] = ax.f[i];
            }
            for (long i = 0; i < pny; i++) {
                hny[kl * pny + i] = ay.nn[i];
                hfy[kl * pny + i] = ay.f[i];
            }
            auto transpose = [](const mtk::Axis &a, int nnode, std::vector<int> &ptr, std::vector<int> &idx,
                                std::vector<double> &wt) {
                std::vector<std::vector<std::pair<int, double>>> lists(nnode);
                for (size_t i = 0; i < a.nn.size(); i++) {
                    if (a.nn[i].x < 0) continue;
                    lists[a.nn[i].x].push_back({(int)i, 1 - a.f[i]});
                    lists[a.nn[i].y].push_back({(int)i, a.f[i]});
                }
                for (int n = 0; n < nnode; n++) {
                    ptr.push_back((int)idx.size());
                    for (auto &e : lists[n]) {
                        idx.push_back(e.first);
                        wt.push_back(e.second);
                    }
                }
                ptr.push_back((int)idx.size());
            };
            transpose(ax, D.lnx[l], htxp, htxi, htxw);
            transpose(ay, D.lny[l], htyp, htyi, htyw);
        }
    }
    t->rnx.upload(hnx);
    t->rny.upload(hny);
    t->rfx.upload(hfx);
    t->rfy.upload(hfy);
    t->txp.upload(htxp);
    t->typ.upload(htyp);
    t->txi.upload(htxi);
    t->tyi.upload(htyi);
    t->txw.upload(htxw);
    t->tyw.upload(htyw);
    // GP: CSC as given, CSR by transposition
    std::vector<long> hcp, hci, hrp, hri;
    std::vector<double> hcx, hrx;
    for (int k = 0; k < nw; k++) {
        const int w = t->active[k];
        const long ng = 2 * R->nsa[w];
        const long *p = R->gp_p[w], *ii = R->gp_i[w];
        const double *xx = R->gp_x[w];
        const long base_c = (long)hci.size();
        D.cpo[k] = (long)hcp.size();
        for (long c = 0; c <= D.nploc; c++) hcp.push_back(base_c + p[c]);
        hci.insert(hci.end(), ii, ii + p[D.nploc]);
        hcx.insert(hcx.end(), xx, xx + p[D.nploc]);
        std::vector<long> cnt(ng + 1, 0);
        for (long e = 0; e < p[D.nploc]; e++) {
            if (ii[e] < 0 || ii[e] >= ng) fail64(MT_EINVAL, "FP64 solve: GP row index out of range");
            cnt[ii[e] + 1]++;
        }
        for (long i = 0; i < ng; i++) cnt[i + 1] += cnt[i];
        const long base_r = (long)hri.size();
        D.rpo[k] = (long)hrp.size();
        for (long i = 0; i <= ng; i++) hrp.push_back(base_r + cnt[i]);
        hri.resize(base_r + p[D.nploc]);
        hrx.resize(base_r + p[D.nploc]);
        std::vector<long> fill(cnt.begin(), cnt.end() - 1);
        for (long c = 0; c < D.nploc; c++)
            for (long e = p[c]; e < p[c + 1]; e++) {
                const long d = base_r + fill[ii[e]]++;
                hri[d] = c;
                hrx[d] = xx[e];
            }
    }
    t->gcp.upload(hcp);
    t->gci.upload(hci);
    t->gcx.upload(hcx);
    t->grp.upload(hrp);
    t->gri.upload(hri);
    t->grx.upload(hrx);
    t->neai.upload(hnea);
    if (!hptt.empty()) t->ptt.upload(hptt);
    // FDPCG
    D.rrows = std::max(1, B64 / D.nx);
    D.rcols = std::max(1, B64 / D.ny);
    D.nrw = (D.npsr * D.ny + D.rrows - 1) / D.rrows;
    D.ncw = (D.npsr * D.nx + D.rcols - 1) / D.rcols;
    split2(D.nx, D.fx1, D.fx2);
    split2(D.ny, D.fy1, D.fy2);
    if (D.precond) {
        if (D.nb * D.bs != D.nxtot) fail64(MT_EINVAL, "FP64 solve: FDPCG blocks do not cover the layers");
        std::vector<long> hp(R->fd_perm, R->fd_perm + D.nxtot), hip(D.nxtot, -1);
        for (long i = 0; i < D.nxtot; i++) {
            const long pm = hp[i];
            const long j = pm > 0 ? pm : -pm;
            if (j >= D.nxtot) fail64(MT_EINVAL, "FP64 solve: FDPCG permutation out of range");
            hip[j] = i | (pm > 0 ? 0 : IPCONJ);
        }
        for (long j = 0; j < D.nxtot; j++)
            if (hip[j] < 0) fail64(MT_EINVAL, "FP64 solve: FDPCG permutation is not one to one");
        t->perm.upload(hp);
        t->iperm.upload(hip);
        t->twx.upload(tw_table(D.nx));
        t->twy.upload(tw_table(D.ny));
    }
    upload_blocks(t, R);
    // chunks and partials
    D.nch = (int)((D.nxtot + B64 - 1) / B64);
    D.nchg = (int)std::max<long>(1, (maxng + B64 - 1) / B64);
    D.nchs = (int)std::max<long>(1, (maxsa + B64 - 1) / B64);
    D.nrz = std::max(D.ncw, D.nch);
    D.o_tt = 0;
    D.o_pap = D.o_tt + nw * 2 * D.nchg;
    D.o_rr = D.o_pap + D.nch;
    D.o_rz = D.o_rr + D.nch;
    D.o_bb = D.o_rz + 2 * D.nrz;
    D.pstride = D.o_bb + D.nch;
    // state
    t->nbatch = nbatch;
    const long ns = nbatch;
    for (Buf<double> *v : {&t->x, &t->r, &t->z, &t->p, &t->Ap, &t->b}) v->alloc(ns * D.nxtot);
    t->x.zero();
    t->opd.alloc(ns * nw * D.nploc);
    t->q.alloc(ns * nw * D.nploc);
    t->gin.alloc(ns * D.ngtot);
    t->g.alloc(ns * D.ngtot);
    t->wg.alloc(ns * D.ngtot);
    t->sa.alloc(ns * D.nxtot);
    t->sb.alloc(ns * D.nxtot);
    t->part.alloc(ns * D.pstride);
    t->part.zero();
    t->stop.alloc(ns);
    t->stop.zero();
    t->res.alloc(ns);
    D.rnx = t->rnx.p;
    D.rny = t->rny.p;
    D.rfx = t->rfx.p;
    D.rfy = t->rfy.p;
    D.txp = t->txp.p;
    D.typ = t->typ.p;
    D.txi = t->txi.p;
    D.tyi = t->tyi.p;
    D.txw = t->txw.p;
    D.tyw = t->tyw.p;
    D.grp = t->grp.p;
    D.gri = t->gri.p;
    D.grx = t->grx.p;
    D.gcp = t->gcp.p;
    D.gci = t->gci.p;
    D.gcx = t->gcx.p;
    D.neai = t->neai.p;
    D.ptt = t->ptt.p;
    D.twx = t->twx.p;
    D.twy = t->twy.p;
    D.perm = t->perm.p;
    D.iperm = t->iperm.p;
    D.x = t->x.p;
    D.r = t->r.p;
    D.z = t->z.p;
    D.p = t->p.p;
    D.Ap = t->Ap.p;
    D.b = t->b.p;
    D.opd = t->opd.p;
    D.q = t->q.p;
    D.gin = t->gin.p;
    D.g = t->g.p;
    D.wg = t->wg.p;
    D.sa = t->sa.p;
    D.sb = t->sb.p;
    D.part = t->part.p;
    D.stop = t->stop.p;
    D.res = t->res.p;
    t->dev.alloc(1);
    C64(hipMemcpy(t->dev.p, &D, sizeof(D64), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// Launch sequences (ws0: first system slot, n: systems)
// ---------------------------------------------------------------------------
dim3 g_ploc(const D64 &D, int n) { return dim3((unsigned)((D.nploc + B64 - 1) / B64), D.nw, n); }
dim3 g_node(const D64 &D, int n) { return dim3(D.nch, n); }

// out = HX' GP' W (I - TT PTT) gin (TomoR, without the alpha scale)
void enq_R(tomo64 *t, hipStream_t st, double *out, int ws0, int n) {
    const D64 &D = t->H;
    hipLaunchKernelGGL(k64_grad<false>, dim3(D.nchg, D.nw, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_wg<false>, dim3(D.nchs, D.nw, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_gpt, g_ploc(D, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_hxt<false>, g_node(D, n), dim3(B64), 0, st, t->dev.p, (const double *)nullptr, out, ws0);
}
// out = (HX' GP' W GP HX + L2'L2) v and the partials of v . out
void enq_L(tomo64 *t, hipStream_t st, const double *v, double *out, int ws0, int n) {
    const D64 &D = t->H;
    hipLaunchKernelGGL(k64_hx, g_ploc(D, n), dim3(B64), 0, st, t->dev.p, v, ws0);
    hipLaunchKernelGGL(k64_grad<true>, dim3(D.nchg, D.nw, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_wg<true>, dim3(D.nchs, D.nw, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_gpt, g_ploc(D, n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_hxt<true>, g_node(D, n), dim3(B64), 0, st, t->dev.p, v, out, ws0);
}
// z = FDPCG^-1 r and the partials of r . z into slot `slot`
void enq_P(tomo64 *t, hipStream_t st, int ws0, int n, int slot) {
    const D64 &D = t->H;
    const size_t lr = sizeof(double2) * 2 * D.rrows * D.nx, lc = sizeof(double2) * 2 * D.rcols * D.ny;
    hipLaunchKernelGGL(k64_frow<true>, dim3(D.nrw, n), dim3(B64), lr, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_fcol<true>, dim3(D.ncw, n), dim3(B64), lc, st, t->dev.p, ws0, slot);
    hipLaunchKernelGGL(k64_blk, dim3((unsigned)((D.nb * D.bs + B64 - 1) / B64), n), dim3(B64), 0, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_frow<false>, dim3(D.nrw, n), dim3(B64), lr, st, t->dev.p, ws0);
    hipLaunchKernelGGL(k64_fcol<false>, dim3(D.ncw, n), dim3(B64), lc, st, t->dev.p, ws0, slot);
}

void enq_solve(tomo64 *t, hipStream_t st, int ws0, int n, int warm, int maxit) {
    const D64 &D = t->H;
    const long nx = D.nxtot;
    double *b = t->b.p + (long)ws0 * nx, *x = t->x.p + (long)ws0 * nx, *p = t->p.p + (long)ws0 * nx,
           *Ap = t->Ap.p + (long)ws0 * nx;
    enq_R(t, st, b, ws0, n);
    hipLaunchKernelGGL(k64_init, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0, warm);
    for (int k = 0; k < maxit; k++) {
        if (k % 100 == 0) {
            if (k == 0 && !warm) {
                hipLaunchKernelGGL(k64_restart<true>, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0, k);
            } else {
                enq_L(t, st, x, Ap, ws0, n);
                hipLaunchKernelGGL(k64_restart<false>, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0, k);
            }
            if (D.precond) {
                enq_P(t, st, ws0, n, k & 1);
                hipLaunchKernelGGL(k64_pfirst, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0);
            }
        }
        enq_L(t, st, p, Ap, ws0, n);
        hipLaunchKernelGGL(k64_upd, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0, k);
        const int last = k + 1 == maxit;
        if (D.precond && !last) enq_P(t, st, ws0, n, (k + 1) & 1);
        hipLaunchKernelGGL(k64_p, g_node(D, n), dim3(B64), 0, st, t->dev.p, ws0, k, last);
    }
    C64(hipGetLastError());
}

void grad_in(tomo64 *t, int isys, const double *grad, hipStream_t st) {
    const D64 &D = t->H;
    std::vector<double> h(D.ngtot);
    for (int k = 0; k < D.nw; k++) {
        const long o = t->goff_all[t->active[k]];
        for (long i = 0; i < 2L * D.nsa[k]; i++) {
            const double v = grad[o + i];
            if (!std::isfinite(v)) fail64(MT_ENOTFINITE, "non-finite gradient input");
            h[D.goff[k] + i] = v;
        }
    }
    C64(hipMemcpyAsync(t->gin.p + (long)isys * D.ngtot, h.data(), sizeof(double) * D.ngtot, hipMemcpyHostToDevice,
                       st));
    C64(hipStreamSynchronize(st));
}
void vec_in(tomo64 *t, double *dst, const double *src, hipStream_t st) {
    const long n = t->H.nxtot;
    for (long i = 0; i < n; i++)
        if (!std::isfinite(src[i])) fail64(MT_ENOTFINITE, "non-finite input vector");
    C64(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, st));
    C64(hipStreamSynchronize(st));
}
void vec_out(tomo64 *t, double *dst, const double *src, hipStream_t st) {
    C64(hipMemcpyAsync(dst, src, sizeof(double) * t->H.nxtot, hipMemcpyDeviceToHost, st));
    C64(hipStreamSynchronize(st));
}

}  // namespace

// Host copy of the recon fields the FP64 plan is built from.
struct recon64 {
    mt_recon_t v{};
    std::vector<long> xnx, xny, nsa;
    std::vector<double> xox, xoy, xdx, ht, thetax, thetay, hs, l2c, Minv;
    std::vector<int> skip, trs;
    std::vector<long> perm;
    std::vector<std::vector<long>> gp_p, gp_i;
    std::vector<std::vector<double>> gp_x, neai, ptt;
    std::vector<const long *> gp_pv, gp_iv;
    std::vector<const double *> gp_xv, neaiv, pttv;
};

int t64_keep(const mt_recon_t *R, recon64 **out) {
    if (!R || !out) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<recon64> c(new (std::nothrow) recon64());
    if (!c) return MT_ENOMEM;
    const int rc = g64([&] {
        const int L = R->npsr, W = R->nwfs;
        if (L <= 0 || W <= 0) fail64(MT_EINVAL, "empty reconstruction");
        mt_recon_t &v = c->v;
        v = *R;
        c->xnx.assign(R->xnx, R->xnx + L);
        c->xny.assign(R->xny, R->xny + L);
        c->xox.assign(R->xox, R->xox + L);
        c->xoy.assign(R->xoy, R->xoy + L);
        c->xdx.assign(R->xdx, R->xdx + L);
        c->ht.assign(R->ht, R->ht + L);
        c->l2c.assign(R->l2c, R->l2c + L);
        c->thetax.assign(R->thetax, R->thetax + W);
        c->thetay.assign(R->thetay, R->thetay + W);
        c->hs.assign(R->hs, R->hs + W);
        c->skip.assign(R->skip, R->skip + W);
        c->trs.assign(R->trs, R->trs + W);
        c->nsa.assign(R->nsa, R->nsa + W);
        const long nploc = R->pnx * R->pny;
        c->gp_p.resize(W);
        c->gp_i.resize(W);
        c->gp_x.resize(W);
        c->neai.resize(W);
        c->ptt.resize(W);
        for (int w = 0; w < W; w++) {
            if (R->skip[w] || R->nsa[w] <= 0) continue;
            const long nz = R->gp_p[w][nploc];
            c->gp_p[w].assign(R->gp_p[w], R->gp_p[w] + nploc + 1);
            c->gp_i[w].assign(R->gp_i[w], R->gp_i[w] + nz);
            c->gp_x[w].assign(R->gp_x[w], R->gp_x[w] + nz);
            c->neai[w].assign(R->neai[w], R->neai[w] + 3 * R->nsa[w]);
            if (R->trs[w] && R->ptt[w]) c->ptt[w].assign(R->ptt[w], R->ptt[w] + 4 * R->nsa[w]);
        }
        for (int w = 0; w < W; w++) {
            c->gp_pv.push_back(c->gp_p[w].empty() ? nullptr : c->gp_p[w].data());
            c->gp_iv.push_back(c->gp_i[w].empty() ? nullptr : c->gp_i[w].data());
            c->gp_xv.push_back(c->gp_x[w].empty() ? nullptr : c->gp_x[w].data());
            c->neaiv.push_back(c->neai[w].empty() ? nullptr : c->neai[w].data());
            c->pttv.push_back(c->ptt[w].empty() ? nullptr : c->ptt[w].data());
        }
        if (R->precond) {
            const long nbs = R->fd_nb * R->fd_bs * R->fd_bs;
            c->Minv.assign(R->fd_Minv, R->fd_Minv + 2 * nbs);
            c->perm.assign(R->fd_perm, R->fd_perm + R->fd_nb * R->fd_bs);
        }
        v.xnx = c->xnx.data();
        v.xny = c->xny.data();
        v.xox = c->xox.data();
        v.xoy = c->xoy.data();
        v.xdx = c->xdx.data();
        v.ht = c->ht.data();
        v.l2c = c->l2c.data();
        v.thetax = c->thetax.data();
        v.thetay = c->thetay.data();
        v.hs = c->hs.data();
        v.skip = c->skip.data();
        v.trs = c->trs.data();
        v.nsa = c->nsa.data();
        v.gp_p = c->gp_pv.data();
        v.gp_i = c->gp_iv.data();
        v.gp_x = c->gp_xv.data();
        v.neai = c->neaiv.data();
        v.ptt = c->pttv.data();
        v.fd_Minv = c->Minv.empty() ? nullptr : c->Minv.data();
        v.fd_perm = c->perm.empty() ? nullptr : c->perm.data();
        v.fd_permhf = nullptr;
        v.dsa = nullptr;
        v.saloc = nullptr;
        v.saa = nullptr;
    });
    if (rc == MT_OK) *out = c.release();
    return rc;
}

int t64_keep_cn2(recon64 *c, const mt_recon_t *R) {
    if (!c || !R) return MT_EINVAL;
    return g64([&] {
        if (R->npsr != c->v.npsr || R->precond != c->v.precond ||
            (R->precond && (R->fd_nb != c->v.fd_nb || R->fd_bs != c->v.fd_bs)))
            fail64(MT_EINVAL, "Cn2 update: the geometry changed");
        c->l2c.assign(R->l2c, R->l2c + R->npsr);
        if (R->precond) c->Minv.assign(R->fd_Minv, R->fd_Minv + 2 * R->fd_nb * R->fd_bs * R->fd_bs);
        c->v.l2c = c->l2c.data();
        c->v.fd_Minv = c->Minv.empty() ? nullptr : c->Minv.data();
    });
}

const mt_recon_t *t64_view(const recon64 *c) { return c ? &c->v : nullptr; }
void t64_keep_free(recon64 *c) { delete c; }

int t64_create(const mt_recon_t *recon, int nbatch, tomo64 **out) {
    if (!recon || !out || nbatch <= 0 || nbatch > 65535) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<tomo64> t(new (std::nothrow) tomo64());
    if (!t) return MT_ENOMEM;
    const int rc = g64([&] { build64(t.get(), recon, nbatch); });
    if (rc == MT_OK) *out = t.release();
    return rc;
}

void t64_free(tomo64 *t) {
    if (!t) return;
    (void)hipDeviceSynchronize();
    delete t;
}

int t64_update_cn2(tomo64 *t, const mt_recon_t *recon) {
    if (!t || !recon) return MT_EINVAL;
    return g64([&] {
        if (recon->npsr != t->H.npsr || (t->H.precond && (recon->fd_nb != t->H.nb || recon->fd_bs != t->H.bs)))
            fail64(MT_EINVAL, "Cn2 update: the geometry changed");
        upload_blocks(t, recon);
        C64(hipMemcpy(t->dev.p, &t->H, sizeof(D64), hipMemcpyHostToDevice));
    });
}

int t64_load_x(tomo64 *t, const float *d_x, hipStream_t st) {
    if (!t || !d_x) return MT_EINVAL;
    return g64([&] {
        hipLaunchKernelGGL(k64_cvt_in, dim3(1024), dim3(256), 0, st, t->x.p, d_x, (long)t->nbatch * t->H.nxtot);
        C64(hipGetLastError());
        C64(hipStreamSynchronize(st));
    });
}

int t64_store_x(tomo64 *t, float *d_x, double *d_res, hipStream_t st) {
    if (!t || !d_x) return MT_EINVAL;
    return g64([&] {
        hipLaunchKernelGGL(k64_cvt_out, dim3(1024), dim3(256), 0, st, d_x, t->x.p, (long)t->nbatch * t->H.nxtot);
        C64(hipGetLastError());
        if (d_res)
            C64(hipMemcpyAsync(d_res, t->res.p, sizeof(double) * t->nbatch, hipMemcpyDeviceToDevice, st));
        C64(hipStreamSynchronize(st));
    });
}

int t64_reset(tomo64 *t) {
    if (!t) return MT_EINVAL;
    return g64([&] { t->x.zero(); });
}

int t64_solve(tomo64 *t, int s0, int nsys, const double *grad, double *x, double *res, int maxit, int warm,
              hipStream_t st) {
    if (!t || !grad || !x || s0 < 0 || nsys <= 0 || s0 + nsys > t->nbatch || maxit <= 0) return MT_EINVAL;
    return g64([&] {
        for (int s = 0; s < nsys; s++) grad_in(t, s0 + s, grad + (long)s * t->ng_all, st);
        enq_solve(t, st, s0, nsys, warm, maxit);
        const long n = t->H.nxtot;
        C64(hipMemcpyAsync(x, t->x.p + (long)s0 * n, sizeof(double) * n * nsys, hipMemcpyDeviceToHost, st));
        if (res) C64(hipMemcpyAsync(res, t->res.p + s0, sizeof(double) * nsys, hipMemcpyDeviceToHost, st));
        C64(hipStreamSynchronize(st));
    });
}

int t64_L(tomo64 *t, const double *xin, double *xout, hipStream_t st) {
    if (!t || !xin || !xout) return MT_EINVAL;
    return g64([&] {
        vec_in(t, t->z.p, xin, st);
        enq_L(t, st, t->z.p, t->Ap.p, 0, 1);
        C64(hipGetLastError());
        vec_out(t, xout, t->Ap.p, st);
    });
}

int t64_R(tomo64 *t, const double *grad, double *xout, hipStream_t st) {
    if (!t || !grad || !xout) return MT_EINVAL;
    return g64([&] {
        grad_in(t, 0, grad, st);
        enq_R(t, st, t->b.p, 0, 1);
        C64(hipGetLastError());
        vec_out(t, xout, t->b.p, st);
    });
}

int t64_P(tomo64 *t, const double *xin, double *xout, hipStream_t st) {
    if (!t || !xin || !xout) return MT_EINVAL;
    return g64([&] {
        if (!t->H.precond) fail64(MT_EINVAL, "no FDPCG preconditioner in this plan");
        vec_in(t, t->r.p, xin, st);
        enq_P(t, st, 0, 1, 0);
        C64(hipGetLastError());
        vec_out(t, xout, t->z.p, st);
    });
}

int t64_get_x(tomo64 *t, int isys, double *x, double *res, hipStream_t st) {
    if (!t || !x || isys < 0 || isys >= t->nbatch) return MT_EINVAL;
    return g64([&] {
        vec_out(t, x, t->x.p + (long)isys * t->H.nxtot, st);
        if (res) C64(hipMemcpy(res, t->res.p + isys, sizeof(double), hipMemcpyDeviceToHost));
    });
}

int t64_residuals(tomo64 *t, double *res, int n, hipStream_t st) {
    if (!t || !res || n <= 0 || n > t->nbatch) return MT_EINVAL;
    return g64([&] {
        C64(hipStreamSynchronize(st));
        C64(hipMemcpy(res, t->res.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}
