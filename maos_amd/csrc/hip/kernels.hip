// Kernels of the tomography PCG/FDPCG solve for gfx950.
//
// Reference counterparts (each replaced by a gather-only kernel):
//   gpu_map2map_do 'n'/'t'   cuda/prop_wrap.cu:34-349   -> k_wfs (HX), k_layer (HX')
//   gpu_gp_do / gpu_gpt_do   cuda/tomo.cu:361-573        -> k_wfs, k_ttr
//   gpu_laplacian_do         cuda/tomo.cu:317-355        -> k_layer
//   cuFFT R2C/C2R +
//   fdpcg_mul_block_sync_half cuda/fdpcg.cu:112-226      -> k_row, k_col, k_irow
//   inn_do / add_do / div_do cuda/kernel.cu, cuda/pcg.cu -> FP64 partials in
//                                                          k_layer/k_irow/k_row,
//                                                          fixed-order reduction
#include "kernels.hpp"
#include "launch.hpp"

namespace mtk {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
template <typename C> struct CT;
template <> struct CT<float2> {
    typedef float R;
    static __device__ __forceinline__ float2 make(float a, float b) { return make_float2(a, b); }
};
template <> struct CT<double2> {
    typedef double R;
    static __device__ __forceinline__ double2 make(double a, double b) { return make_double2(a, b); }
};

// ---------------------------------------------------------------------------
// Deterministic FP64 reductions (fixed order: per-lane strided sums, then a
// 64-lane shuffle tree, then wavefront partials summed by lane 0 in order).
// ---------------------------------------------------------------------------
__device__ double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) s += sh[i];
        sh[0] = s;
    }
    __syncthreads();
    double r = sh[0];
    __syncthreads();
    return r;
}
__device__ double reduce_partials(const double *part, int n, double *sh) {
    double v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    return block_sum(v, sh);
}

// ---------------------------------------------------------------------------
// Stockham mixed-radix FFT on LDS: nf transforms of length n, stored
// contiguously; radices 2,3,4,5,7 (sizes that nextfftsize produces).
// tw[k] = exp(-2 pi i k/n).  SIGN=-1 forward, +1 inverse, unnormalized.
// ---------------------------------------------------------------------------
template <int R, int SIGN, typename C>
__device__ __forceinline__ void small_dft(C *v, const C *tw, int n) {
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // forward: a3 = -i*d ; inverse: +i*d
        C a3 = SIGN < 0 ? CT<C>::make(d.y, -d.x) : CT<C>::make(-d.y, d.x);
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    } else {
        C o[R];
        const int step = n / R;
#pragma unroll
        for (int k = 0; k < R; k++) {
            C acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                C w = tw[((k * r) % R) * step];
                if (SIGN > 0) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            o[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = o[k];
    }
}

template <int R, int SIGN, typename C>
__device__ void fft_stage(const C *in, C *out, int n, int nf, int Ns, const C *tw) {
    const int nb = n / R;
    const int tot = nf * nb;
    const int tstep = n / (Ns * R);
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        const int f = t / nb, j = t - f * nb;
        const C *src = in + f * n;
        C *dst = out + f * n;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
        const int jm = j % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C w = tw[jm * r * tstep];
                if (SIGN > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        small_dft<R, SIGN, C>(v, tw, n);
        const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; r++) dst[idxD + r * Ns] = v[r];
    }
}

// Returns the buffer holding the result (a or b).
template <int SIGN, typename C>
__device__ C *fft_run(C *a, C *b, int n, int nf, int nst, const int *rad, const C *tw) {
    int Ns = 1;
    C *in = a, *out = b;
    for (int s = 0; s < nst; s++) {
        const int R = rad[s];
        switch (R) {
            case 2: fft_stage<2, SIGN, C>(in, out, n, nf, Ns, tw); break;
            case 3: fft_stage<3, SIGN, C>(in, out, n, nf, Ns, tw); break;
            case 4: fft_stage<4, SIGN, C>(in, out, n, nf, Ns, tw); break;
            case 5: fft_stage<5, SIGN, C>(in, out, n, nf, Ns, tw); break;
            case 7: fft_stage<7, SIGN, C>(in, out, n, nf, Ns, tw); break;
            default: break;
        }
        __syncthreads();
        C *t = in;
        in = out;
        out = t;
        Ns *= R;
    }
    return in;
}

template <typename C> __device__ __forceinline__ const C *tw_of(const Geom &G, bool y);
template <> __device__ __forceinline__ const float2 *tw_of<float2>(const Geom &G, bool y) { return y ? G.twy : G.twx; }
template <> __device__ __forceinline__ const double2 *tw_of<double2>(const Geom &G, bool y) { return y ? G.twy64 : G.twx64; }

// ---------------------------------------------------------------------------
// TT projection of the RHS gradients: tt = PTT * g per (system, WFS).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ttr(Geom G, const float *grad, float *tt) {
    __shared__ double sh[8];
    const int w = blockIdx.x, s = blockIdx.z;
    float *out = tt + ((long)s * G.nwfs + w) * 2;
    if (!G.ttr[w]) {
        if (threadIdx.x == 0) out[0] = out[1] = 0.f;
        return;
    }
    const int nsa = G.nsa[w];
    const float *g = grad + (long)s * G.ngtot + G.goff[w];
    const float *P = G.ptt + 4 * G.saoff[w];
    double a0 = 0, a1 = 0;
    for (int i = threadIdx.x; i < 2 * nsa; i += blockDim.x) {
        a0 += (double)P[2 * i] * g[i];
        a1 += (double)P[2 * i + 1] * g[i];
    }
    a0 = block_sum(a0, sh);
    a1 = block_sum(a1, sh);
    if (threadIdx.x == 0) {
        out[0] = (float)a0;
        out[1] = (float)a1;
    }
}

// ---------------------------------------------------------------------------
// WFS-side pass.  One workgroup = one TILE x TILE tile of ploc points of one
// WFS.  LHS: opd = HX p on the tile + halo (LDS), g = GP opd, W g, then
// q = GP' (W g) gathered from the <= (pos+1)^2 subapertures touching each
// point.  RHS: W (g - TT tt) then the same GP' gather.
// ---------------------------------------------------------------------------
template <int POS, bool RHS, typename RT>
__global__ __launch_bounds__(256) void k_wfs(Geom G, const float *x, const float *grad,
                                             const float *tt, RT *q) {
    constexpr int H = TILE + 2 * POS;
    constexpr int S = TILE + POS;
    constexpr int NP = POS + 1;
    __shared__ RT opd[RHS ? 1 : H * H];
    __shared__ RT wgx[S * S], wgy[S * S];
    __shared__ int sidx[S * S];
    const int w = blockIdx.y, s = blockIdx.z;
    const int pnx = G.pnx, pny = G.pny;
    const int ntx = (pnx + TILE - 1) / TILE;
    const int tx0 = (blockIdx.x % ntx) * TILE, ty0 = (blockIdx.x / ntx) * TILE;
    if (!RHS) {
        const float *xs = x + (long)s * G.nxtot;
        for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
            const int px = tx0 - POS + i % H, py = ty0 - POS + i / H;
            RT v = 0;
            if (px >= 0 && px < pnx && py >= 0 && py < pny) {
                for (int l = 0; l < G.npsr; l++) {
                    const int wl = w * MAXL + l;
                    const int2 nx_ = G.hx_nx[(long)wl * pnx + px];
                    const int2 ny_ = G.hx_ny[(long)wl * pny + py];
                    if (nx_.x >= 0 && ny_.x >= 0) {
                        const RT fx = G.hx_fx[(long)wl * pnx + px];
                        const RT fy = G.hx_fy[(long)wl * pny + py];
                        const float *xl = xs + G.xoff[l];
                        const int nxl = G.xnx[l];
                        const RT a = xl[ny_.x * nxl + nx_.x], b = xl[ny_.x * nxl + nx_.y];
                        const RT c = xl[ny_.y * nxl + nx_.x], d = xl[ny_.y * nxl + nx_.y];
                        v += (RT(1) - fy) * ((RT(1) - fx) * a + fx * b) + fy * ((RT(1) - fx) * c + fx * d);
                    }
                }
            }
            opd[i] = v;
        }
        __syncthreads();
    }
    const int nsa = G.nsa[w];
    const float *gw = grad + (long)s * G.ngtot + G.goff[w];
    RT ttx = 0, tty = 0;
    if (RHS && G.ttr[w]) {
        ttx = tt[((long)s * G.nwfs + w) * 2];
        tty = tt[((long)s * G.nwfs + w) * 2 + 1];
    }
    for (int i = threadIdx.x; i < S * S; i += blockDim.x) {
        const int a = i % S, b = i / S;
        const int cx = tx0 - POS + a, cy = ty0 - POS + b;
        int sa = -1;
        if (cx >= 0 && cx < pnx && cy >= 0 && cy < pny) sa = G.saidx[(long)w * G.nploc + cy * pnx + cx];
        sidx[i] = sa;
        RT ox = 0, oy = 0;
        if (sa >= 0) {
            const long so = G.saoff[w] + sa;
            RT gx = 0, gy = 0;
            if (RHS) {
                gx = (RT)gw[sa] - ttx;
                gy = (RT)gw[nsa + sa] - tty;
            } else {
                const float *cxp = G.gpx + so * NP * NP, *cyp = G.gpy + so * NP * NP;
#pragma unroll
                for (int zy = 0; zy < NP; zy++)
#pragma unroll
                    for (int zx = 0; zx < NP; zx++) {
                        const RT o = opd[(b + zy) * H + a + zx];
                        gx += (RT)cxp[zy * NP + zx] * o;
                        gy += (RT)cyp[zy * NP + zx] * o;
                    }
            }
            const float4 ne = G.neai[so];
            ox = (RT)ne.x * gx + (RT)ne.z * gy;
            oy = (RT)ne.z * gx + (RT)ne.y * gy;
        }
        wgx[i] = ox;
        wgy[i] = oy;
    }
    __syncthreads();
    const int ox = threadIdx.x % TILE, oy = threadIdx.x / TILE;
    const int px = tx0 + ox, py = ty0 + oy;
    if (threadIdx.x < TILE * TILE && px < pnx && py < pny) {
        RT acc = 0;
#pragma unroll
        for (int zy = 0; zy < NP; zy++)
#pragma unroll
            for (int zx = 0; zx < NP; zx++) {
                const int si = (oy + POS - zy) * S + (ox + POS - zx);
                const int sa = sidx[si];
                if (sa >= 0) {
                    const long so = G.saoff[w] + sa;
                    acc += (RT)G.gpx[so * NP * NP + zy * NP + zx] * wgx[si] +
                           (RT)G.gpy[so * NP * NP + zy * NP + zx] * wgy[si];
                }
            }
        q[((long)s * G.nwfs + w) * G.nploc + py * pnx + px] = acc;
    }
}

// ---------------------------------------------------------------------------
// Layer-side pass.  out = sum_w HX_wl' q_w  (+ L2'L2 p for the LHS), gathered
// per layer node from the <= TAPS^2 ploc samples whose bilinear footprint
// covers it.  Also writes the FP64 partial of p.Ap (LHS) or out.out (RHS).
// ---------------------------------------------------------------------------
template <bool LHS, typename RT>
__global__ __launch_bounds__(256) void k_layer(Geom G, const RT *q, const float *p, RT *out,
                                               double *part, int partstride) {
    __shared__ double sh[8];
    const int l = blockIdx.y, s = blockIdx.z;
    const int nxl = G.xnx[l], nyl = G.xny[l];
    const int ntx = (nxl + TILE - 1) / TILE, nty = (nyl + TILE - 1) / TILE;
    double dot = 0;
    if ((int)blockIdx.x < ntx * nty) {
        const int jx = (blockIdx.x % ntx) * TILE + threadIdx.x % TILE;
        const int jy = (blockIdx.x / ntx) * TILE + threadIdx.x / TILE;
        if (jx < nxl && jy < nyl) {
            RT acc = 0;
            const long xt = G.hxt_xoff[G.npsr];  // per-WFS stride of transposed tables
            for (int w = 0; w < G.nwfs; w++) {
                const int wl = w * MAXL + l;
                const int4 cx = G.hxt_x[(long)w * xt + G.hxt_xoff[l] + jx];
                const int4 cy = G.hxt_y[(long)w * xt + G.hxt_xoff[l] + jy];
                const int ix[TAPS] = {cx.x, cx.y, cx.z, cx.w};
                const int iy[TAPS] = {cy.x, cy.y, cy.z, cy.w};
                float wx[TAPS], wy[TAPS];
#pragma unroll
                for (int k = 0; k < TAPS; k++) {
                    wx[k] = 0.f;
                    wy[k] = 0.f;
                    if (ix[k] >= 0) {
                        const int2 nn = G.hx_nx[(long)wl * G.pnx + ix[k]];
                        const float f = G.hx_fx[(long)wl * G.pnx + ix[k]];
                        wx[k] = (nn.x == jx ? 1.f - f : 0.f) + (nn.y == jx ? f : 0.f);
                    }
                    if (iy[k] >= 0) {
                        const int2 nn = G.hx_ny[(long)wl * G.pny + iy[k]];
                        const float f = G.hx_fy[(long)wl * G.pny + iy[k]];
                        wy[k] = (nn.x == jy ? 1.f - f : 0.f) + (nn.y == jy ? f : 0.f);
                    }
                }
                const RT *qw = q + ((long)s * G.nwfs + w) * G.nploc;
#pragma unroll
                for (int ky = 0; ky < TAPS; ky++) {
                    if (iy[ky] < 0) continue;
                    RT ra = 0;
#pragma unroll
                    for (int kx = 0; kx < TAPS; kx++)
                        if (ix[kx] >= 0) ra += (RT)wx[kx] * qw[iy[ky] * G.pnx + ix[kx]];
                    acc += (RT)wy[ky] * ra;
                }
            }
            const long o = (long)s * G.nxtot + G.xoff[l];
            if (LHS) {
                // periodic L2'L2 stencil (20, -8, 2, 1) * l2c, accumulated in RT:
                // its terms are large and cancel for smooth p
                const float *pl = p + o;
                auto P = [&](int dx, int dy) {
                    int xx = jx + dx, yy = jy + dy;
                    xx = xx < 0 ? xx + nxl : (xx >= nxl ? xx - nxl : xx);
                    yy = yy < 0 ? yy + nyl : (yy >= nyl ? yy - nyl : yy);
                    return (RT)pl[yy * nxl + xx];
                };
                const RT c = P(0, 0);
                const RT s4 = P(-1, 0) + P(1, 0) + P(0, -1) + P(0, 1);
                const RT sd = P(-1, -1) + P(1, -1) + P(-1, 1) + P(1, 1);
                const RT s2 = P(-2, 0) + P(2, 0) + P(0, -2) + P(0, 2);
                acc += (RT)G.l2c[l] * (RT(20) * c - RT(8) * s4 + RT(2) * sd + s2);
                dot = (double)c * (double)acc;
            } else {
                dot = (double)acc * (double)acc;
            }
            out[o + jy * nxl + jx] = acc;
        }
    }
    dot = block_sum(dot, sh);
    if (threadIdx.x == 0 && part) part[(long)s * partstride + blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// ---------------------------------------------------------------------------
// Matched-sampling fast path (ploc spacing * cone scale == layer spacing for
// every WFS and layer: LGS in cone coordinates, NGS on an NGS-only system).
// The ray of (w, l) is then a constant integer offset plus a constant
// bilinear fraction, so HX is a 2x2 stencil on a shifted window and HX' its
// mirror.  Layer windows (k_wfs_m) and WFS-field windows (k_layer_m) are
// staged in LDS with coalesced loads.
// ---------------------------------------------------------------------------
template <int POS, typename RT>
__global__ __launch_bounds__(256) void k_wfs_m(Geom G, const float *x, RT *q) {
    constexpr int H = TILE + 2 * POS;
    constexpr int H1 = H + 1;
    constexpr int S = TILE + POS;
    constexpr int NP = POS + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float *xs = reinterpret_cast<float *>(lds_raw);  // [npsr][H1*H1]
    __shared__ RT opd[H * H];
    __shared__ RT wgx[S * S], wgy[S * S];
    __shared__ int sidx[S * S];
    const int w = blockIdx.y, s = blockIdx.z;
    const int pnx = G.pnx, pny = G.pny;
    const int ntx = (pnx + TILE - 1) / TILE;
    const int tx0 = (blockIdx.x % ntx) * TILE, ty0 = (blockIdx.x / ntx) * TILE;
    const float *xsys = x + (long)s * G.nxtot;
    for (int l = 0; l < G.npsr; l++) {
        const MRay m = G.mray[w * MAXL + l];
        const int nxl = G.xnx[l], nyl = G.xny[l];
        const float *xl = xsys + G.xoff[l];
        const int gx0 = tx0 - POS + m.cx, gy0 = ty0 - POS + m.cy;
        for (int i = threadIdx.x; i < H1 * H1; i += blockDim.x) {
            int gx = gx0 + i % H1, gy = gy0 + i / H1;
            if (gx == nxl && m.wx) gx = 0;  // edge ray: second node wraps
            if (gy == nyl && m.wy) gy = 0;
            xs[l * H1 * H1 + i] = (gx >= 0 && gx < nxl && gy >= 0 && gy < nyl) ? xl[gy * nxl + gx] : 0.f;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        const int a = i % H, b = i / H;
        const int px = tx0 - POS + a, py = ty0 - POS + b;
        RT v = 0;
        if (px >= 0 && px < pnx && py >= 0 && py < pny) {
            for (int l = 0; l < G.npsr; l++) {
                const MRay m = G.mray[w * MAXL + l];
                if (px >= m.ax0 && px < m.ax1 && py >= m.ay0 && py < m.ay1) {
                    const float *t = xs + l * H1 * H1 + b * H1 + a;
                    const RT fx = m.fx, fy = m.fy;
                    v += (RT(1) - fy) * ((RT(1) - fx) * (RT)t[0] + fx * (RT)t[1]) +
                         fy * ((RT(1) - fx) * (RT)t[H1] + fx * (RT)t[H1 + 1]);
                }
            }
        }
        opd[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S * S; i += blockDim.x) {
        const int a = i % S, b = i / S;
        const int cx = tx0 - POS + a, cy = ty0 - POS + b;
        int sa = -1;
        if (cx >= 0 && cx < pnx && cy >= 0 && cy < pny) sa = G.saidx[(long)w * G.nploc + cy * pnx + cx];
        sidx[i] = sa;
        RT ox = 0, oy = 0;
        if (sa >= 0) {
            const long so = G.saoff[w] + sa;
            RT gx = 0, gy = 0;
            const float *cxp = G.gpx + so * NP * NP, *cyp = G.gpy + so * NP * NP;
#pragma unroll
            for (int zy = 0; zy < NP; zy++)
#pragma unroll
                for (int zx = 0; zx < NP; zx++) {
                    const RT o = opd[(b + zy) * H + a + zx];
                    gx += (RT)cxp[zy * NP + zx] * o;
                    gy += (RT)cyp[zy * NP + zx] * o;
                }
            const float4 ne = G.neai[so];
            ox = (RT)ne.x * gx + (RT)ne.z * gy;
            oy = (RT)ne.z * gx + (RT)ne.y * gy;
        }
        wgx[i] = ox;
        wgy[i] = oy;
    }
    __syncthreads();
    const int ox = threadIdx.x % TILE, oy = threadIdx.x / TILE;
    const int px = tx0 + ox, py = ty0 + oy;
    if (threadIdx.x < TILE * TILE && px < pnx && py < pny) {
        RT acc = 0;
#pragma unroll
        for (int zy = 0; zy < NP; zy++)
#pragma unroll
            for (int zx = 0; zx < NP; zx++) {
                const int si = (oy + POS - zy) * S + (ox + POS - zx);
                const int sa = sidx[si];
                if (sa >= 0) {
                    const long so = G.saoff[w] + sa;
                    acc += (RT)G.gpx[so * NP * NP + zy * NP + zx] * wgx[si] +
                           (RT)G.gpy[so * NP * NP + zy * NP + zx] * wgy[si];
                }
            }
        q[((long)s * G.nwfs + w) * G.nploc + py * pnx + px] = acc;
    }
}

template <bool LHS, typename RT>
__global__ __launch_bounds__(256) void k_layer_m(Geom G, const RT *q, const float *p, RT *out,
                                                 double *part, int partstride) {
    constexpr int Q = TILE + 1;      // WFS-field window edge
    constexpr int PH = TILE + 4;     // p window edge (L2'L2 reach is 2)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    RT *qs = reinterpret_cast<RT *>(lds_raw);                       // [nwfs][Q*Q]
    float *ps = reinterpret_cast<float *>(qs + G.nwfs * Q * Q);     // [PH*PH]
    __shared__ double sh[8];
    const int l = blockIdx.y, s = blockIdx.z;
    const int nxl = G.xnx[l], nyl = G.xny[l];
    const int ntx = (nxl + TILE - 1) / TILE, nty = (nyl + TILE - 1) / TILE;
    const bool active = (int)blockIdx.x < ntx * nty;
    const int jx0 = (blockIdx.x % ntx) * TILE, jy0 = (blockIdx.x / ntx) * TILE;
    if (active) {
        for (int w = 0; w < G.nwfs; w++) {
            const MRay m = G.mray[w * MAXL + l];
            const RT *qw = q + ((long)s * G.nwfs + w) * G.nploc;
            const int ix0 = jx0 - m.cx - 1, iy0 = jy0 - m.cy - 1;
            for (int i = threadIdx.x; i < Q * Q; i += blockDim.x) {
                int ix = ix0 + i % Q, iy = iy0 + i / Q;
                // node 0 is the wrapped second node of the edge ray a1 - 1;
                // slot i = -c - 1 (node 0's second-node ray) is otherwise empty
                if (m.wx && ix == -m.cx - 1) ix = m.ax1 - 1;
                if (m.wy && iy == -m.cy - 1) iy = m.ay1 - 1;
                RT v = 0;
                if (ix >= m.ax0 && ix < m.ax1 && iy >= m.ay0 && iy < m.ay1) v = qw[iy * G.pnx + ix];
                qs[w * Q * Q + i] = v;
            }
        }
        if (LHS) {
            const float *pl = p + (long)s * G.nxtot + G.xoff[l];
            for (int i = threadIdx.x; i < PH * PH; i += blockDim.x) {
                // periodic; the last tile's window may run past a small layer
                const int xx = (jx0 - 2 + i % PH + nxl) % nxl, yy = (jy0 - 2 + i / PH + nyl) % nyl;
                ps[i] = pl[yy * nxl + xx];
            }
        }
    }
    __syncthreads();
    double dot = 0;
    const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
    const int jx = jx0 + tx, jy = jy0 + ty;
    if (active && jx < nxl && jy < nyl) {
        RT acc = 0;
        for (int w = 0; w < G.nwfs; w++) {
            const MRay m = G.mray[w * MAXL + l];
            const RT fx = m.fx, fy = m.fy;
            const RT *t = qs + w * Q * Q + ty * Q + tx;
            acc += fy * ((RT(1) - fx) * t[1] + fx * t[0]) + (RT(1) - fy) * ((RT(1) - fx) * t[Q + 1] + fx * t[Q]);
        }
        const long o = (long)s * G.nxtot + G.xoff[l];
        if (LHS) {
            auto P = [&](int dx, int dy) { return (RT)ps[(ty + 2 + dy) * PH + tx + 2 + dx]; };
            const RT c = P(0, 0);
            const RT s4 = P(-1, 0) + P(1, 0) + P(0, -1) + P(0, 1);
            const RT sd = P(-1, -1) + P(1, -1) + P(-1, 1) + P(1, 1);
            const RT s2 = P(-2, 0) + P(2, 0) + P(0, -2) + P(0, 2);
            acc += (RT)G.l2c[l] * (RT(20) * c - RT(8) * s4 + RT(2) * sd + s2);
            dot = (double)c * (double)acc;
        } else {
            dot = (double)acc * (double)acc;
        }
        out[o + jy * nxl + jx] = acc;
    }
    dot = block_sum(dot, sh);
    if (threadIdx.x == 0 && part) part[(long)s * partstride + blockIdx.y * gridDim.x + blockIdx.x] = dot;
}

// ---------------------------------------------------------------------------
// Row kernel: PCG vector update fused with the forward real-to-complex row
// FFT of r.  Two real rows are packed into one complex FFT.
// ---------------------------------------------------------------------------
enum { ROW_INIT_COLD = 0, ROW_INIT_WARM = 1, ROW_UPDATE = 2, ROW_LAST = 3 };

template <int MODE, typename RT, typename FT>
__global__ __launch_bounds__(256) void k_row(Geom G, Scal SC, int iter, float *x, RT *r,
                                             const RT *b, const float *p, const RT *Ap,
                                             FT *spec, int *stop) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FT *lds = reinterpret_cast<FT *>(lds_raw);
    typedef typename CT<FT>::R FR;
    __shared__ double sh[8];
    const int s = blockIdx.z;
    const int2 rb = G.rowblk[blockIdx.x];
    const int l = rb.x, row0 = rb.y;
    const int n = G.fnx, nyl = G.xny[l];
    const int nrows = min(ROWS, nyl - row0);
    const int nf = nrows / 2;
    double alpha = 0.;
    if (MODE >= ROW_UPDATE) {
        const double pap = reduce_partials(SC.part_pap + (long)s * SC.maxpart, SC.npart_layer, sh);
        const double rz = SC.rz[(long)s * (SC.maxit + 2) + iter];
        double a = rz / pap;
        const int stopped = stop[s];
        if (!isfinite(a) || stopped) {
            // lib/pcg.c:74-77: a non-finite alpha ends the solve; every
            // workgroup computes the same alpha, so all of them freeze x.
            a = 0;
            if (!stopped && threadIdx.x == 0 && blockIdx.x == 0) stop[s] = 1;
        }
        alpha = a;
    }
    // x += alpha p of an UPDATE iteration is applied by the inverse row
    // kernel, which reads p anyway for p = z + beta p.
    if (MODE == ROW_UPDATE && blockIdx.x == 0 && threadIdx.x == 0) SC.alpha[s] = alpha;
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    FR *bufA = reinterpret_cast<FR *>(lds);
    double dot = 0;
    for (int i = threadIdx.x; i < nrows * n; i += blockDim.x) {
        const int row = i / n, col = i - row * n;
        const long idx = base + (long)row * n + col;
        RT rv;
        if (MODE == ROW_INIT_COLD) {
            rv = b[idx];
            r[idx] = rv;
            x[idx] = 0.f;
        } else if (MODE == ROW_INIT_WARM) {
            rv = b[idx] - Ap[idx];
            r[idx] = rv;
        } else {
            if (MODE == ROW_LAST) x[idx] += (float)alpha * p[idx];
            rv = r[idx] - (RT)alpha * Ap[idx];
            r[idx] = rv;
        }
        if (MODE == ROW_LAST) {
            dot += (double)rv * (double)rv;
        } else {
            bufA[2 * ((row >> 1) * n + col) + (row & 1)] = (FR)rv;
        }
    }
    if (MODE == ROW_LAST) {
        dot = block_sum(dot, sh);
        if (threadIdx.x == 0) SC.part_rr[(long)s * SC.maxpart + blockIdx.x] = dot;
        return;
    }
    __syncthreads();
    FT *A = lds, *B = lds + ROWS / 2 * n;
    FT *res = fft_run<-1, FT>(A, B, n, nf, G.nstx, G.radx, tw_of<FT>(G, false));
    const int nh = n / 2 + 1;
    FT *S = spec + (long)s * G.nspec + G.soff[l];
    for (int i = threadIdx.x; i < nf * nh; i += blockDim.x) {
        const int f = i % nf, k = i / nf;
        const FT Z = res[f * n + k];
        const FT Zc = cconj(res[f * n + (n - k) % n]);
        const FT sum = cadd(Z, Zc), dif = csub(Z, Zc);
        S[(long)k * nyl + row0 + 2 * f] = CT<FT>::make(FR(0.5) * sum.x, FR(0.5) * sum.y);
        S[(long)k * nyl + row0 + 2 * f + 1] = CT<FT>::make(FR(0.5) * dif.y, FR(-0.5) * dif.x);
    }
}

// ---------------------------------------------------------------------------
// Column kernel: forward column FFTs of one column group (all layers),
// per-frequency bs x bs block multiply (half spectrum, conjugate flags), then
// inverse column FFTs.  Everything in LDS.
// ---------------------------------------------------------------------------
template <typename FT>
__global__ __launch_bounds__(256) void k_col(Geom G, Scal SC, const FT *specf, float2 *speci) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sh[8];
    typedef typename CT<FT>::R FR;
    const int g = blockIdx.x, s = blockIdx.z;
    const int ny = G.fny;
    const int c0 = G.grp_col[g], c1 = G.grp_col[g + 1];
    const int ncg = c1 - c0;
    const int nf = ncg * G.npsr;
    const int tot = nf * ny;
    const int cap = G.maxcol * G.npsr * ny;
    FT *A = reinterpret_cast<FT *>(lds_raw), *B = A + cap;
    float2 *C = reinterpret_cast<float2 *>(B + cap), *D = C + cap;
    const FT *Sf = specf + (long)s * G.nspec;
    float2 *Si = speci + (long)s * G.nspec;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int f = i / ny, ky = i - f * ny;
        const int ci = f / G.npsr, l = f - ci * G.npsr;
        const int kx = G.colkx[c0 + ci];
        A[i] = Sf[G.soff[l] + (long)kx * ny + ky];
    }
    __syncthreads();
    FT *res = fft_run<-1, FT>(A, B, ny, nf, G.nsty, G.rady, tw_of<FT>(G, true));
    if (G.scale) {
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            res[i].x *= (FR)G.fscale;
            res[i].y *= (FR)G.fscale;
        }
        __syncthreads();
    }
    // The per-layer piston is in the null space of the tomography operator
    // (GP annihilates constants, the periodic L2'L2 stencil sums to zero) and
    // the right-hand side is orthogonal to it, so in exact arithmetic every
    // residual has a zero DC bin.  FDPCG's DC block is regularised only by
    // max(psd)*sqrt(eps) (maos/fdpcg.c:396-401), so rounding left in that bin
    // would be amplified ~1e8 and change alpha/beta.  Zeroing the DC input is
    // therefore exact and keeps the iteration on the FP64 one.
    if (G.zero_dc) {
        for (int ci = 0; ci < ncg; ci++) {
            if (G.colkx[c0 + ci] != 0) continue;
            for (int l = threadIdx.x; l < G.npsr; l += blockDim.x) res[(ci * G.npsr + l) * ny] = CT<FT>::make(0, 0);
        }
        __syncthreads();
    }
    // Block multiply, and the FP64 partial of r.z taken on the spectrum:
    // sum_xy r z = sc * sum_k conj(R_k) Z_k over the full spectrum (Parseval;
    // sc = fscale when scaled, and R here already carries that fscale), i.e.
    // the half-spectrum sum with weight 2 off the kx = 0 and kx = n/2 columns.
    const int bs = G.bs;
    const int b0 = G.grp_blk[g], b1 = G.grp_blk[g + 1];
    double dot = 0;
    for (int t = threadIdx.x; t < (b1 - b0) * bs; t += blockDim.x) {
        const int bi = G.blkid[b0 + t / bs];
        const int m = t % bs;
        const float2 *M = G.Minv + (long)bi * bs * bs;
        const int *bp = G.blkpos + (long)bi * bs;
        FT y = CT<FT>::make(0, 0);
        for (int k = 0; k < bs; k++) {
            const int pe = bp[k];
            FT v = res[pe & 0x1fffffff];
            if ((pe >> 30) & 1) v.y = -v.y;
            const float2 mk = M[m + k * bs];
            y = cadd(y, cmul(CT<FT>::make(mk.x, mk.y), v));
        }
        const int pm = bp[m];
        if ((pm >> 30) & 1) y.y = -y.y;
        if (!((pm >> 29) & 1)) {
            const int pos = pm & 0x1fffffff;
            C[pos] = make_float2((float)y.x, (float)y.y);
            const int kx = G.colkx[c0 + pos / ny / G.npsr];
            const double wgt = (kx == 0 || 2 * kx == G.fnx) ? 1.0 : 2.0;
            dot += wgt * ((double)res[pos].x * (double)y.x + (double)res[pos].y * (double)y.y);
        }
    }
    dot = block_sum(dot, sh);  // contains the barrier before the inverse FFT
    if (threadIdx.x == 0) SC.part_rz[(long)s * SC.maxpart + g] = dot;
    float2 *res2 = fft_run<1, float2>(C, D, ny, nf, G.nsty, G.rady, G.twy);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int f = i / ny, ky = i - f * ny;
        const int ci = f / G.npsr, l = f - ci * G.npsr;
        const int kx = G.colkx[c0 + ci];
        Si[G.soff[l] + (long)kx * ny + ky] = res2[i];
    }
}

// ---------------------------------------------------------------------------
// Inverse row kernel: complex-to-real row FFT -> z (the imaginary parts of
// the DC and Nyquist bins are ignored, as C2R does), then by mode
//   IROW_Z       out = z                                  (P alone)
//   IROW_FIRST   p = z; rz[slot] = r.z                    (start / restart)
//   IROW_UPDATE  beta = r.z / rz[slot]; rz[slot+1] = r.z;
//                x += alpha p; p = z + beta p             (lib/pcg.c:703,714)
// r.z is reduced from the column kernel's spectral partials (fixed order,
// the same value in every workgroup).
// ---------------------------------------------------------------------------
enum { IROW_Z = 0, IROW_FIRST = 1, IROW_UPDATE = 2 };

template <int MODE>
__global__ __launch_bounds__(256) void k_irow(Geom G, Scal SC, int slot, const float2 *spec, float *x,
                                              float *out) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float2 *lds = reinterpret_cast<float2 *>(lds_raw);
    __shared__ double sh[8];
    const int s = blockIdx.z;
    double beta = 0, alpha = 0;
    if (MODE != IROW_Z) {
        const double rz = reduce_partials(SC.part_rz + (long)s * SC.maxpart, SC.npart_col, sh);
        double *hist = SC.rz + (long)s * (SC.maxit + 2);
        if (MODE == IROW_UPDATE) {
            beta = rz / hist[slot];
            if (!isfinite(beta)) beta = 0;
            alpha = SC.alpha[s];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) hist[MODE == IROW_FIRST ? slot : slot + 1] = rz;
    }
    const int2 rb = G.rowblk[blockIdx.x];
    const int l = rb.x, row0 = rb.y;
    const int n = G.fnx, nyl = G.xny[l];
    const int nrows = min(ROWS, nyl - row0);
    const int nf = nrows / 2;
    const float2 *S = spec + (long)s * G.nspec + G.soff[l];
    float2 *A = lds, *B = lds + ROWS / 2 * n;
    for (int i = threadIdx.x; i < nf * n; i += blockDim.x) {
        const int f = i % nf, k = i / nf;
        const int ya = row0 + 2 * f;
        float2 a, b;
        if (k <= n / 2) {
            a = S[(long)k * nyl + ya];
            b = S[(long)k * nyl + ya + 1];
            if (k == 0 || k == n / 2) {
                a.y = 0.f;
                b.y = 0.f;
            }
        } else {
            a = cconj(S[(long)(n - k) * nyl + ya]);
            b = cconj(S[(long)(n - k) * nyl + ya + 1]);
        }
        A[f * n + k] = make_float2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    float2 *res = fft_run<1, float2>(A, B, n, nf, G.nstx, G.radx, G.twx);
    const float sc = G.scale ? G.fscale : 1.f;
    const float fa = (float)alpha, fb = (float)beta;
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    for (int i = threadIdx.x; i < nrows * n; i += blockDim.x) {
        const int row = i / n, col = i - row * n;
        const float2 v = res[(row >> 1) * n + col];
        const float zv = ((row & 1) ? v.y : v.x) * sc;
        const long idx = base + (long)row * n + col;
        if (MODE == IROW_UPDATE) {
            const float pv = out[idx];
            x[idx] += fa * pv;
            out[idx] = zv + fb * pv;
        } else {
            out[idx] = zv;
        }
    }
}

// Residual sqrt(r.r / b.b) per system (NaN when the solve stopped early).
__global__ __launch_bounds__(256) void k_final(Geom G, Scal SC, int npart_rr, const int *stop) {
    __shared__ double sh[8];
    const int s = blockIdx.z;
    const double rr = reduce_partials(SC.part_rr + (long)s * SC.maxpart, npart_rr, sh);
    const double bb = reduce_partials(SC.part_bb + (long)s * SC.maxpart, SC.npart_layer, sh);
    if (threadIdx.x == 0) SC.res[s] = stop[s] ? NAN : sqrt(rr / bb);
}

// Copy / convert helpers.
__global__ void k_copy(const float *a, float *b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}
__global__ void k_zero_i(int *a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void launch_ttr(const Geom &G, int nsys, const float *grad, float *tt, hipStream_t st) {
    hipLaunchKernelGGL(k_ttr, dim3(G.nwfs, 1, nsys), dim3(256), 0, st, G, grad, tt);
}
template <typename RT>
static void launch_wfs_t(const Geom &G, int nsys, bool rhs, const float *x, const float *grad,
                         const float *tt, RT *q, hipStream_t st) {
    const int nt = ((G.pnx + TILE - 1) / TILE) * ((G.pny + TILE - 1) / TILE);
    dim3 grid(nt, G.nwfs, nsys);
    if (G.pos == 1) {
        if (rhs) hipLaunchKernelGGL((k_wfs<1, true, RT>), grid, dim3(256), 0, st, G, x, grad, tt, q);
        else hipLaunchKernelGGL((k_wfs<1, false, RT>), grid, dim3(256), 0, st, G, x, grad, tt, q);
    } else {
        if (rhs) hipLaunchKernelGGL((k_wfs<2, true, RT>), grid, dim3(256), 0, st, G, x, grad, tt, q);
        else hipLaunchKernelGGL((k_wfs<2, false, RT>), grid, dim3(256), 0, st, G, x, grad, tt, q);
    }
}
template <typename RT>
static void launch_wfs_m(const Geom &G, int nsys, const float *x, RT *q, hipStream_t st) {
    const int nt = ((G.pnx + TILE - 1) / TILE) * ((G.pny + TILE - 1) / TILE);
    dim3 grid(nt, G.nwfs, nsys);
    const int H1 = TILE + 2 * G.pos + 1;
    const size_t lds = sizeof(float) * G.npsr * H1 * H1;
    if (G.pos == 1) hipLaunchKernelGGL((k_wfs_m<1, RT>), grid, dim3(256), lds, st, G, x, q);
    else hipLaunchKernelGGL((k_wfs_m<2, RT>), grid, dim3(256), lds, st, G, x, q);
}
void launch_wfs(const Geom &G, int nsys, bool rhs, int rt64, const float *x, const float *grad,
                const float *tt, void *q, hipStream_t st) {
    if (!rhs && G.matched) {
        if (rt64) launch_wfs_m<double>(G, nsys, x, (double *)q, st);
        else launch_wfs_m<float>(G, nsys, x, (float *)q, st);
        return;
    }
    if (rt64) launch_wfs_t<double>(G, nsys, rhs, x, grad, tt, (double *)q, st);
    else launch_wfs_t<float>(G, nsys, rhs, x, grad, tt, (float *)q, st);
}
int layer_tiles(const Geom &G) {
    int mx = 0;
    for (int l = 0; l < G.npsr; l++) {
        int t = ((G.xnx[l] + TILE - 1) / TILE) * ((G.xny[l] + TILE - 1) / TILE);
        mx = t > mx ? t : mx;
    }
    return mx;
}
void launch_layer(const Geom &G, int nsys, bool lhs, int rt64, const void *q, const float *p, void *out,
                  double *part, int partstride, hipStream_t st) {
    dim3 grid(layer_tiles(G), G.npsr, nsys);
    if (G.matched) {
        const size_t Q2 = (TILE + 1) * (TILE + 1), P2 = (TILE + 4) * (TILE + 4);
        if (rt64) {
            const size_t lds = sizeof(double) * G.nwfs * Q2 + sizeof(float) * P2;
            double *o = (double *)out;
            const double *qq = (const double *)q;
            if (lhs) hipLaunchKernelGGL((k_layer_m<true, double>), grid, dim3(256), lds, st, G, qq, p, o, part, partstride);
            else hipLaunchKernelGGL((k_layer_m<false, double>), grid, dim3(256), lds, st, G, qq, p, o, part, partstride);
        } else {
            const size_t lds = sizeof(float) * G.nwfs * Q2 + sizeof(float) * P2;
            float *o = (float *)out;
            const float *qq = (const float *)q;
            if (lhs) hipLaunchKernelGGL((k_layer_m<true, float>), grid, dim3(256), lds, st, G, qq, p, o, part, partstride);
            else hipLaunchKernelGGL((k_layer_m<false, float>), grid, dim3(256), lds, st, G, qq, p, o, part, partstride);
        }
        return;
    }
    if (rt64) {
        double *o = (double *)out;
        const double *qq = (const double *)q;
        if (lhs) hipLaunchKernelGGL((k_layer<true, double>), grid, dim3(256), 0, st, G, qq, p, o, part, partstride);
        else hipLaunchKernelGGL((k_layer<false, double>), grid, dim3(256), 0, st, G, qq, p, o, part, partstride);
    } else {
        float *o = (float *)out;
        const float *qq = (const float *)q;
        if (lhs) hipLaunchKernelGGL((k_layer<true, float>), grid, dim3(256), 0, st, G, qq, p, o, part, partstride);
        else hipLaunchKernelGGL((k_layer<false, float>), grid, dim3(256), 0, st, G, qq, p, o, part, partstride);
    }
}
size_t row_lds(const Geom &G, int f64) { return (f64 ? sizeof(double2) : sizeof(float2)) * ROWS * G.fnx; }
template <typename RT, typename FT>
static void launch_row_t(const Geom &G, const Scal &SC, int nsys, int mode, int iter, float *x, RT *r,
                         const RT *b, const float *p, const RT *Ap, FT *spec, int *stop, hipStream_t st) {
    dim3 grid(G.nrowblk, 1, nsys);
    const size_t lds = row_lds(G, sizeof(FT) == sizeof(double2));
    switch (mode) {
        case ROW_INIT_COLD:
            hipLaunchKernelGGL((k_row<ROW_INIT_COLD, RT, FT>), grid, dim3(256), lds, st, G, SC, iter, x, r, b, p, Ap, spec, stop);
            break;
        case ROW_INIT_WARM:
            hipLaunchKernelGGL((k_row<ROW_INIT_WARM, RT, FT>), grid, dim3(256), lds, st, G, SC, iter, x, r, b, p, Ap, spec, stop);
            break;
        case ROW_UPDATE:
            hipLaunchKernelGGL((k_row<ROW_UPDATE, RT, FT>), grid, dim3(256), lds, st, G, SC, iter, x, r, b, p, Ap, spec, stop);
            break;
        default:
            hipLaunchKernelGGL((k_row<ROW_LAST, RT, FT>), grid, dim3(256), lds, st, G, SC, iter, x, r, b, p, Ap, spec, stop);
            break;
    }
}
void launch_row(const Geom &G, const Scal &SC, int nsys, int mode, int rt64, int iter, float *x, void *r,
                const void *b, const float *p, const void *Ap, void *spec, int *stop, hipStream_t st) {
    if (rt64)
        launch_row_t<double, double2>(G, SC, nsys, mode, iter, x, (double *)r, (const double *)b, p,
                                      (const double *)Ap, (double2 *)spec, stop, st);
    else
        launch_row_t<float, float2>(G, SC, nsys, mode, iter, x, (float *)r, (const float *)b, p, (const float *)Ap,
                                    (float2 *)spec, stop, st);
}
size_t col_lds(const Geom &G, int f64) {
    return ((f64 ? sizeof(double2) : sizeof(float2)) * 2 + sizeof(float2) * 2) * (size_t)G.maxcol * G.npsr * G.fny;
}
void launch_col(const Geom &G, const Scal &SC, int nsys, int f64, const void *specf, float2 *speci,
                hipStream_t st) {
    if (f64)
        hipLaunchKernelGGL(k_col<double2>, dim3(G.ngroup, 1, nsys), dim3(256), col_lds(G, 1), st, G, SC,
                           (const double2 *)specf, speci);
    else
        hipLaunchKernelGGL(k_col<float2>, dim3(G.ngroup, 1, nsys), dim3(256), col_lds(G, 0), st, G, SC,
                           (const float2 *)specf, speci);
}
void launch_irow(const Geom &G, const Scal &SC, int nsys, int mode, int slot, const float2 *spec, float *x,
                 float *out, hipStream_t st) {
    const dim3 grid(G.nrowblk, 1, nsys);
    const size_t lds = row_lds(G, 0);
    switch (mode) {
        case IROW_Z: hipLaunchKernelGGL(k_irow<IROW_Z>, grid, dim3(256), lds, st, G, SC, slot, spec, x, out); break;
        case IROW_FIRST:
            hipLaunchKernelGGL(k_irow<IROW_FIRST>, grid, dim3(256), lds, st, G, SC, slot, spec, x, out);
            break;
        default: hipLaunchKernelGGL(k_irow<IROW_UPDATE>, grid, dim3(256), lds, st, G, SC, slot, spec, x, out); break;
    }
}
void launch_final(const Geom &G, const Scal &SC, int nsys, int npart_rr, const int *stop, hipStream_t st) {
    hipLaunchKernelGGL(k_final, dim3(1, 1, nsys), dim3(256), 0, st, G, SC, npart_rr, stop);
}
void launch_copy(const float *a, float *b, long n, hipStream_t st) {
    const long nb = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, st, a, b, n);
}
void launch_zero_i(int *a, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i, dim3((n + 255) / 256), dim3(256), 0, st, a, n);
}

}  // namespace mtk
