// GPU plan of the tomography solve and its C ABI (gpu_tomo_*).
//
// gpu_tomo_setup plays the role of gpu_setup_recon -> curecon_t ->
// cutomo_grid / cufdpcg_t (cuda/recon.cu:50-230, cuda/tomo.cu:141-312,
// cuda/fdpcg.cu:32-102): it converts the host FP64 operators into FP32 device
// tables (ray tables instead of PROP_WRAP_T, GP stencils like prep_GP but in
// FP32 and refusing GP that leaves the stencil, NEA as float4 per subaperture,
// FDPCG half-spectrum blocks grouped by FFT column) and allocates the state of
// `nbatch` systems.  gpu_tomo_solve replaces gpu_tomo + gpu_pcg
// (cuda/recon.cu:259-276, cuda/pcg.cu:87-242): same PCG recurrence, FP32
// vectors, FP64 dot products and alpha/beta, no host synchronisation inside
// the iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "launch.hpp"
#include "maos_tomo.h"

namespace mt {
void set_error(const std::string &msg);
}

using namespace mtk;

namespace {

struct HipError {
    int code;
    std::string msg;
};

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw HipError{MT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)};           \
    } while (0)

[[noreturn]] void fail(int code, const std::string &m) { throw HipError{code, m}; }

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        free();
        n = count;
        if (count) HIPCHK(hipMalloc((void **)&p, sizeof(T) * count));
    }
    void upload(const std::vector<T> &h) {
        alloc(h.size());
        if (n) HIPCHK(hipMemcpy(p, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
};

// Kernel classes timed by the profiler.
// Kernel classes timed by the profiler: the iteration kernels (k_wfs ..
// k_irow, the UPDATE-iteration forms) and the once-per-solve forms apart.
enum KClass {
    KC_TTR, KC_WFS, KC_LAYER, KC_ROW, KC_COL, KC_IROW, KC_FINAL,
    KC_WFS_RHS, KC_LAYER_RHS, KC_ROW_START, KC_ROW_LAST, KC_IROW_START, KC_N
};
const char *kc_names[KC_N] = {"k_ttr", "k_wfs", "k_layer", "k_row", "k_col", "k_irow", "k_final",
                              "k_wfs_rhs", "k_layer_rhs", "k_row_start", "k_row_last", "k_irow_start"};

}  // namespace

struct gpu_tomo {
    int device = 0;
    int nbatch = 1, maxit = 30, warm = 0;
    Geom G{};
    Scal SC{};
    hipStream_t stream = nullptr;
    // host copies needed for validation/sizes
    int npsr = 0, nwfs_all = 0;
    std::vector<int> wfs_active;  // index of active WFS -> original WFS
    std::vector<long> nsa_all, goff_all;  // per original WFS
    long ngtot_all = 0;                    // gradient length in the caller's layout (all WFS)
    long nxtot = 0;
    // geometry tables
    DevBuf<int2> hx_nx, hx_ny;
    DevBuf<float> hx_fx, hx_fy;
    DevBuf<int4> hxt_x, hxt_y;
    DevBuf<MRay> mray;
    DevBuf<int> saidx;
    DevBuf<float> gpx, gpy, ptt;
    DevBuf<float4> neai, nead;
    DevBuf<float2> gpd;
    DevBuf<float2> twx, twy, Minv;
    DevBuf<double2> twx64, twy64;
    DevBuf<int> grp_col, colkx, grp_blk, blkpos;
    std::vector<int> blk_order;  // half-spectrum block id of each block in group order
    DevBuf<int2> rowblk;
    // state
    DevBuf<float> x, z, p, g, tt;
    DevBuf<double> r, Ap, b, q;  // residual side: double, or float in the first half when !rt64
    int rt64 = 1;             // 1: r, Ap, b in FP64 (default); 0: FP32
    int spec64 = 1;           // 1: forward FDPCG spectrum in FP64; 0: FP32
    DevBuf<double2> specf;  // forward spectrum of r (double2, or float2 in the first half when !rt64)
    DevBuf<float2> speci;   // block-multiplied spectrum, inverse transform input
    DevBuf<double> part_pap, part_rz, part_rr, part_bb, rz, alpha, res;
    DevBuf<int> stop;
    // host staging
    std::vector<float> hstage;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_nsys = 0;
    // profiling
    int prof = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
    double prof_ms[KC_N] = {0};
    int prof_cnt[KC_N] = {0};
};

namespace {

// ---------------------------------------------------------------------------
// Ray tables: the same column-by-column decisions as prop_grid_stat
// (lib/prop_grid.c:47-343) on the square ploc, evaluated once in FP64.
// For each ploc index along x (resp. y): n0 = left (lower) node, n1 = the next
// node (wrapping to 0 at the last node, as the reference's edge branch does),
// f = fractional weight of n1; n0 = -1 when the ray misses the layer grid.
// ---------------------------------------------------------------------------
struct Axis {
    std::vector<int2> nn;
    std::vector<double> f;
};

void ray_axes(double mdx, double mox, double moy, long nxin, long nyin, long pnx, long pny,
              double pox, double poy, double pdx, double dispx, double dispy, double scale, Axis &ax,
              Axis &ay) {
    const double EPSP = 1.e-12;
    const double dx_in1 = 1. / mdx;
    const double dxout = pdx * dx_in1 * scale, dyout = pdx * dx_in1 * scale;
    const double displacex = (dispx - mox) * dx_in1, displacey = (dispy - moy) * dx_in1;
    const long wrapx = nxin - 1, wrapy = nyin - 1;
    ax.nn.assign(pnx, make_int2(-1, -1));
    ax.f.assign(pnx, 0.);
    ay.nn.assign(pny, make_int2(-1, -1));
    ay.f.assign(pny, 0.);
    auto next = [](long n0, long wrap) { return n0 == wrap ? 0L : n0 + 1; };
    const double oxout = pox * dx_in1 * scale + displacex;
    const bool matched = std::fabs(dxout - 1.) < EPSP && std::fabs(dyout - 1.) < EPSP;
    if (matched) {
        int irows = 0;
        if (oxout < 0) irows = (int)std::ceil(-oxout);
        const int nplocx = (int)std::floor(oxout);
        const double fx = oxout - nplocx;
        const int do_edge = fx < EPSP;
        long rowdiv0 = wrapx - nplocx;
        rowdiv0 = std::min<long>(std::max<long>(rowdiv0, 0), pnx);
        long irow = irows;
        for (; irow < rowdiv0; irow++) {
            const long n0 = nplocx + irow;
            ax.nn[irow] = make_int2((int)n0, (int)(n0 + 1));
            ax.f[irow] = fx;
        }
        if (irow < pnx && do_edge && nplocx + irow == wrapx) {
            ax.nn[irow] = make_int2((int)wrapx, 0);
            ax.f[irow] = fx;
        }
    } else {
        const double dxout1 = 1. / dxout;
        int irows = 0;
        if (oxout < 0) irows = (int)std::ceil(-oxout * dxout1);
        int rowdiv0 = (int)std::ceil((wrapx - oxout) * dxout1);
        if (rowdiv0 > 0) {
            if (rowdiv0 > pnx) rowdiv0 = (int)pnx;
            double dplocx = oxout + irows * dxout;
            int irow = irows;
            for (; irow < rowdiv0; irow++, dplocx += dxout) {
                const long n0 = (long)std::floor(dplocx);
                if (n0 < 0 || n0 > wrapx) continue;
                ax.nn[irow] = make_int2((int)n0, (int)next(n0, wrapx));
                ax.f[irow] = dplocx - n0;
            }
            for (; irow < pnx && dplocx < wrapx + EPSP; irow++, dplocx += dxout) {
                const long n0 = (long)std::floor(dplocx);
                if (n0 < 0 || n0 > wrapx) continue;
                ax.nn[irow] = make_int2((int)n0, (int)next(n0, wrapx));
                ax.f[irow] = dplocx - n0;
            }
        }
    }
    for (long icol = 0; icol < pny; icol++) {
        const double oyout = (poy + icol * pdx) * dx_in1 * scale + displacey;
        if (oyout < 0) continue;
        const double dyout1 = matched ? 1. : 1. / dyout;
        if ((long)std::floor((wrapy - oyout) * dyout1) + 1 < 1) continue;
        const long n0 = (long)std::floor(oyout);
        if (n0 > wrapy) continue;
        ay.nn[icol] = make_int2((int)n0, (int)next(n0, wrapy));
        ay.f[icol] = oyout - n0;
    }
}

FDiv make_fdiv(unsigned d) {
    FDiv f;
    f.d = d;
    f.m = d <= 1 ? 0u : (unsigned)((((unsigned long long)1 << 32) + d - 1) / d);
    return f;
}
// Per-stage constants of the Stockham FFT of length n with radices rad.
void fft_stages(const std::vector<int> &rad, int n, FStage *out) {
    int Ns = 1;
    for (size_t i = 0; i < rad.size(); i++) {
        FStage &st = out[i];
        st.R = rad[i];
        st.Ns = Ns;
        st.nb = n / st.R;
        st.tstep = n / (Ns * st.R);
        st.dnb = make_fdiv((unsigned)st.nb);
        st.dNs = make_fdiv((unsigned)Ns);
        Ns *= st.R;
    }
}

// Radix plan for the Stockham FFT.
std::vector<int> fft_radices(int n) {
    std::vector<int> r;
    int m = n;
    const int pref[5] = {4, 2, 3, 5, 7};
    while (m > 1) {
        int rr = 0;
        for (int i = 0; i < 5; i++)
            if (m % pref[i] == 0) {
                rr = pref[i];
                break;
            }
        if (!rr) fail(MT_EINVAL, "FFT size " + std::to_string(n) + " has a prime factor above 7");
        r.push_back(rr);
        m /= rr;
    }
    if ((int)r.size() > MAXST) fail(MT_EINVAL, "too many FFT stages");
    return r;
}
std::vector<double2> twiddles64(int n) {
    std::vector<double2> t(n);
    for (int k = 0; k < n; k++) {
        const double ph = -2.0 * M_PI * (double)k / (double)n;
        t[k] = make_double2(std::cos(ph), std::sin(ph));
    }
    return t;
}
std::vector<float2> twiddles(int n) {
    std::vector<float2> t(n);
    for (int k = 0; k < n; k++) {
        const double ph = -2.0 * M_PI * (double)k / (double)n;
        t[k] = make_float2((float)std::cos(ph), (float)std::sin(ph));
    }
    return t;
}

// Half-spectrum FDPCG blocks for the R2C/C2R path.  The reference's CPU apply
// (maos/fdpcg.c:661-726) multiplies every block of the full C2C spectrum and
// keeps the real part of the inverse FFT, i.e. it uses the Hermitian part
// (Y(k) + conj(Y(-k)))/2 of the block output.  The reference's GPU apply
// (cuda/fdpcg.cu:112-137) multiplies only the half-spectrum blocks, which
// equals the CPU result only where Minv(-f) = conj(Minv(f)); the FD model
// breaks that on the Nyquist row (the shift factors are evaluated at -N/2 for
// both f and -f).  Using 1/2 (M_a + conj(P M_b P')) for each half block a,
// with b the block holding the conjugate frequencies (P: member matching),
// reproduces the CPU apply exactly and makes duplicated half-spectrum writes
// agree.
std::vector<float2> half_blocks(const mt_recon_t *R, long nx0, long ny0);
// Blocks of M (bs*bs each) permuted into the given order.
std::vector<float2> group_order(const std::vector<float2> &M, const std::vector<int> &order, long bs) {
    std::vector<float2> out(M.size());
    for (size_t j = 0; j < order.size(); j++)
        std::copy(M.begin() + (long)order[j] * bs * bs, M.begin() + ((long)order[j] + 1) * bs * bs,
                  out.begin() + (long)j * bs * bs);
    return out;
}
std::vector<float2> half_blocks(const mt_recon_t *R, long nx0, long ny0) {
    const long bs = R->fd_bs, nb = R->fd_nb, nbh = R->fd_nbhalf;
    const long nfull = nx0 * ny0;
    if (!R->fd_perm) fail(MT_EINVAL, "FDPCG full permutation missing");
    std::vector<int> posblk(nfull * R->npsr, -1), posmem(nfull * R->npsr, -1);
    for (long b = 0; b < nb; b++)
        for (long m = 0; m < bs; m++) {
            const long P = R->fd_perm[b * bs + m];
            if (P < 0 || P >= nfull * R->npsr) fail(MT_EINVAL, "FDPCG permutation out of range");
            posblk[P] = (int)b;
            posmem[P] = (int)m;
        }
    const double *M = R->fd_Minv;
    std::vector<float2> out(nbh * bs * bs);
    std::vector<int> pi(bs);
    for (long a = 0; a < nbh; a++) {
        int b = -1;
        for (long m = 0; m < bs; m++) {
            const long P = R->fd_perm[a * bs + m];
            const long l = P / nfull, rem = P - l * nfull, kx = rem % nx0, ky = rem / nx0;
            const long Q = l * nfull + (nx0 - kx) % nx0 + ((ny0 - ky) % ny0) * nx0;
            if (posblk[Q] < 0 || (b >= 0 && posblk[Q] != b)) fail(MT_EINVAL, "FDPCG blocks not closed under conjugation");
            b = posblk[Q];
            pi[m] = posmem[Q];
        }
        for (long j = 0; j < bs; j++)
            for (long i = 0; i < bs; i++) {
                const long ia = (a * bs * bs + i + j * bs) * 2;
                const long ib = (b * bs * bs + pi[i] + pi[j] * bs) * 2;
                out[a * bs * bs + i + j * bs] =
                    make_float2((float)(0.5 * (M[ia] + M[ib])), (float)(0.5 * (M[ia + 1] - M[ib + 1])));
            }
    }
    return out;
}

void build_plan(gpu_tomo *g, const mt_recon_t *R) {
    if (!R || R->npsr <= 0 || R->npsr > MAXL || R->nwfs <= 0) fail(MT_EINVAL, "invalid recon (npsr/nwfs)");
    if (R->pos != 1 && R->pos != 2) fail(MT_EINVAL, "tomo.pos must be 1 or 2");
    if (R->ptt_lhs) fail(MT_EINVAL, "TT removal in the LHS (recon.split=0 or tomo.splitlrt=1) is not supported");
    Geom &G = g->G;
    G = Geom();
    g->npsr = R->npsr;
    G.npsr = R->npsr;
    G.pos = R->pos;
    G.np = R->pos + 1;
    G.pnx = (int)R->pnx;
    G.pny = (int)R->pny;
    G.nploc = (int)(R->pnx * R->pny);
    long off = 0;
    for (int l = 0; l < R->npsr; l++) {
        G.xnx[l] = (int)R->xnx[l];
        G.xny[l] = (int)R->xny[l];
        G.xoff[l] = off;
        off += R->xnx[l] * R->xny[l];
    }
    G.xoff[R->npsr] = off;
    G.nxtot = off;
    g->nxtot = off;
    // active WFS
    g->nwfs_all = R->nwfs;
    g->nsa_all.assign(R->nwfs, 0);
    g->goff_all.assign(R->nwfs + 1, 0);
    for (int w = 0; w < R->nwfs; w++) {
        g->nsa_all[w] = R->skip[w] ? 0 : R->nsa[w];
        g->goff_all[w + 1] = g->goff_all[w] + 2 * g->nsa_all[w];
        if (!R->skip[w] && R->nsa[w] > 0) g->wfs_active.push_back(w);
    }
    g->ngtot_all = g->goff_all[R->nwfs];
    const int nw = (int)g->wfs_active.size();
    if (nw == 0 || nw > MAXW) fail(MT_EINVAL, "number of active WFS must be in [1, 16]");
    G.nwfs = nw;
    long goff = 0, saoff = 0;
    for (int k = 0; k < nw; k++) {
        const int w = g->wfs_active[k];
        G.nsa[k] = (int)R->nsa[w];
        G.goff[k] = goff;
        G.saoff[k] = saoff;
        goff += 2 * R->nsa[w];
        saoff += R->nsa[w];
        G.ttr[k] = R->trs[w] && R->ptt[w] ? 1 : 0;
    }
    G.goff[nw] = goff;
    G.saoff[nw] = saoff;
    G.ngtot = goff;
    for (int l = 0; l < R->npsr; l++) G.l2c[l] = (float)R->l2c[l];

    // --- ray tables -------------------------------------------------------
    const long pnx = R->pnx, pny = R->pny;
    std::vector<int2> hnx((size_t)nw * MAXL * pnx, make_int2(-1, -1)), hny((size_t)nw * MAXL * pny, make_int2(-1, -1));
    std::vector<float> hfx((size_t)nw * MAXL * pnx, 0.f), hfy((size_t)nw * MAXL * pny, 0.f);
    long xt = 0;
    for (int l = 0; l < R->npsr; l++) {
        G.hxt_xoff[l] = xt;
        xt += std::max(R->xnx[l], R->xny[l]);
    }
    G.hxt_xoff[R->npsr] = xt;
    std::vector<int4> htx((size_t)nw * xt, make_int4(-1, -1, -1, -1)), hty((size_t)nw * xt, make_int4(-1, -1, -1, -1));
    for (int k = 0; k < nw; k++) {
        const int w = g->wfs_active[k];
        for (int l = 0; l < R->npsr; l++) {
            const double ht = R->ht[l];
            Axis ax, ay;
            ray_axes(R->xdx[l], R->xox[l], R->xoy[l], R->xnx[l], R->xny[l], pnx, pny, R->pox, R->poy, R->pdx,
                     R->thetax[w] * ht, R->thetay[w] * ht, 1. - ht / R->hs[w], ax, ay);
            const size_t wl = (size_t)k * MAXL + l;
            for (long i = 0; i < pnx; i++) {
                hnx[wl * pnx + i] = ax.nn[i];
                hfx[wl * pnx + i] = (float)ax.f[i];
            }
            for (long i = 0; i < pny; i++) {
                hny[wl * pny + i] = ay.nn[i];
                hfy[wl * pny + i] = (float)ay.f[i];
            }
            // transposed lists
            auto addt = [&](std::vector<int4> &T, long node, int idx) {
                int4 &e = T[(size_t)k * xt + G.hxt_xoff[l] + node];
                int *v = &e.x;
                for (int t = 0; t < TAPS; t++) {
                    if (v[t] == idx) return;
                    if (v[t] < 0) {
                        v[t] = idx;
                        return;
                    }
                }
                fail(MT_EINVAL, "more than 4 ploc samples map to one layer node (ray ratio < 0.5)");
            };
            for (long i = 0; i < pnx; i++) {
                if (ax.nn[i].x < 0) continue;
                addt(htx, ax.nn[i].x, (int)i);
                addt(htx, ax.nn[i].y, (int)i);
            }
            for (long i = 0; i < pny; i++) {
                if (ay.nn[i].x < 0) continue;
                addt(hty, ay.nn[i].x, (int)i);
                addt(hty, ay.nn[i].y, (int)i);
            }
        }
    }
    // matched-sampling detection: on each axis the valid ploc indices form
    // one interval [a0, a1) whose rays land on node i + c with the next node
    // i + c + 1 and one fraction (float copies of FP64 values that differ by
    // rounding only).  The last ray may instead be the reference's edge case
    // (node n-1, next node wrapped to 0): flagged as wrap.
    std::vector<MRay> hm((size_t)nw * MAXL);
    int matched = 1;
    auto simple_axis = [&](const std::vector<int2> &nn, const std::vector<float> &ff, size_t o, long n,
                           int nl, int &c, float &f, int &a0, int &a1, int &wrap) {
        a0 = -1;
        a1 = -1;
        wrap = 0;
        for (long i = 0; i < n; i++) {
            const int2 v = nn[o + i];
            if (v.x < 0) {
                if (a0 >= 0 && a1 < 0) a1 = (int)i;
                continue;
            }
            if (a1 >= 0 || wrap) return false;  // second interval, or a ray after the edge
            if (a0 < 0) {
                a0 = (int)i;
                c = v.x - (int)i;
                f = ff[o + i];
            }
            if (v.x - (int)i != c || std::fabs(ff[o + i] - f) > 1e-6f) return false;
            if (v.y != v.x + 1) {
                if (v.x == nl - 1 && v.y == 0) wrap = 1;
                else return false;
            }
        }
        if (a0 < 0) {  // ray misses the layer entirely
            a0 = a1 = 0;
            c = 0;
            f = 0.f;
        } else if (a1 < 0) {
            a1 = (int)n;
        }
        return true;
    };
    for (int k = 0; k < nw && matched; k++)
        for (int l = 0; l < R->npsr && matched; l++) {
            const size_t wl = (size_t)k * MAXL + l;
            MRay &m = hm[wl];
            if (!simple_axis(hnx, hfx, wl * pnx, pnx, R->xnx[l], m.cx, m.fx, m.ax0, m.ax1, m.wx) ||
                !simple_axis(hny, hfy, wl * pny, pny, R->xny[l], m.cy, m.fy, m.ay0, m.ay1, m.wy))
                matched = 0;
        }
    // the matched kernels wrap the L2 window once: layers of at least TILE + 2
    for (int l = 0; l < R->npsr; l++)
        if (R->xnx[l] < TILE + 2 || R->xny[l] < TILE + 2) matched = 0;
    G.matched = matched;
    g->mray.upload(hm);
    G.mray = g->mray.p;
    g->hx_nx.upload(hnx);
    g->hx_ny.upload(hny);
    g->hx_fx.upload(hfx);
    g->hx_fy.upload(hfy);
    g->hxt_x.upload(htx);
    g->hxt_y.upload(hty);
    G.hx_nx = g->hx_nx.p;
    G.hx_ny = g->hx_ny.p;
    G.hx_fx = g->hx_fx.p;
    G.hx_fy = g->hx_fy.p;
    G.hxt_x = g->hxt_x.p;
    G.hxt_y = g->hxt_y.p;

    // --- subapertures, GP stencils, NEA, PTT ---------------------------------
    const int NP = G.np;
    std::vector<int> hsaidx((size_t)nw * G.nploc, -1);
    std::vector<float> hgpx((size_t)saoff * NP * NP, 0.f), hgpy((size_t)saoff * NP * NP, 0.f);
    std::vector<float4> hneai(saoff);
    std::vector<float> hptt((size_t)saoff * 4, 0.f);
    for (int k = 0; k < nw; k++) {
        const int w = g->wfs_active[k];
        const long nsa = R->nsa[w];
        const double *sx = R->saloc[w], *sy = R->saloc[w] + nsa;
        std::vector<long> cx(nsa), cy(nsa);
        for (long i = 0; i < nsa; i++) {
            // saloc mapped onto pmap (cuda/tomo.cu:91-111, prep_saptr)
            cx[i] = std::lround((sx[i] - R->pox) / R->pdx);
            cy[i] = std::lround((sy[i] - R->poy) / R->pdx);
            if (cx[i] < 0 || cy[i] < 0 || cx[i] + R->pos >= pnx || cy[i] + R->pos >= pny)
                fail(MT_EINVAL, "subaperture stencil leaves the ploc grid");
            int &slot = hsaidx[(size_t)k * G.nploc + cy[i] * pnx + cx[i]];
            if (slot >= 0) fail(MT_EINVAL, "two subapertures share a ploc corner");
            slot = (int)i;
        }
        const long *gp = R->gp_p[w], *gi = R->gp_i[w];
        const double *gv = R->gp_x[w];
        for (long c = 0; c < G.nploc; c++) {
            for (long e = gp[c]; e < gp[c + 1]; e++) {
                const long row = gi[e];
                if (row < 0 || row >= 2 * nsa) fail(MT_EINVAL, "GP row index out of range");
                const long isa = row < nsa ? row : row - nsa;
                const long zx = c % pnx - cx[isa], zy = c / pnx - cy[isa];
                if (zx < 0 || zx > R->pos || zy < 0 || zy > R->pos) {
                    if (std::fabs(gv[e]) > 0)
                        fail(MT_EINVAL, "GP couples a subaperture to ploc points outside its (pos+1)^2 stencil");
                    continue;
                }
                const size_t o = (size_t)(G.saoff[k] + isa) * NP * NP + zy * NP + zx;
                if (row < nsa) hgpx[o] += (float)gv[e];
                else hgpy[o] += (float)gv[e];
            }
        }
        const double *ne = R->neai[w];
        for (long i = 0; i < nsa; i++)
            hneai[G.saoff[k] + i] = make_float4((float)ne[3 * i], (float)ne[3 * i + 1], (float)ne[3 * i + 2], 0.f);
        if (G.ttr[k]) {
            const double *P = R->ptt[w];
            for (long i = 0; i < 4 * nsa; i++) hptt[(size_t)G.saoff[k] * 4 + i] = (float)P[i];
        }
    }
    {
        // dense per-point copies, deduplicated over WFS
        const size_t nl = (size_t)G.nploc;
        std::vector<float2> dgp;
        std::vector<float4> dne;
        int ncls = 0;
        for (int k = 0; k < nw; k++) {
            std::vector<float2> cg(nl * NP * NP, make_float2(0.f, 0.f));
            std::vector<float4> cn(nl, make_float4(0.f, 0.f, 0.f, 0.f));
            for (size_t c = 0; c < nl; c++) {
                const int sa = hsaidx[(size_t)k * nl + c];
                if (sa < 0) continue;
                const size_t so = (size_t)G.saoff[k] + sa;
                for (int z = 0; z < NP * NP; z++) cg[c * NP * NP + z] = make_float2(hgpx[so * NP * NP + z], hgpy[so * NP * NP + z]);
                cn[c] = hneai[so];
            }
            int found = -1;
            for (int c = 0; c < ncls && found < 0; c++)
                if (!std::memcmp(cg.data(), dgp.data() + (size_t)c * cg.size(), sizeof(float2) * cg.size()) &&
                    !std::memcmp(cn.data(), dne.data() + (size_t)c * nl, sizeof(float4) * nl))
                    found = c;
            if (found < 0) {
                found = ncls++;
                dgp.insert(dgp.end(), cg.begin(), cg.end());
                dne.insert(dne.end(), cn.begin(), cn.end());
            }
            G.gcls[k] = found;
        }
        g->gpd.upload(dgp);
        g->nead.upload(dne);
        G.gpd = g->gpd.p;
        G.nead = g->nead.p;
    }
    g->saidx.upload(hsaidx);
    g->gpx.upload(hgpx);
    g->gpy.upload(hgpy);
    g->neai.upload(hneai);
    g->ptt.upload(hptt);
    G.saidx = g->saidx.p;
    G.gpx = g->gpx.p;
    G.gpy = g->gpy.p;
    G.neai = g->neai.p;
    G.ptt = g->ptt.p;

    // --- FDPCG ---------------------------------------------------------------
    G.precond = R->precond ? 1 : 0;
    // FFT geometry is needed for the row blocks even without preconditioner
    // (the PCG update kernel walks rows).
    const int nx0 = (int)R->xnx[0], ny0 = (int)R->xny[0];
    for (int l = 0; l < R->npsr; l++)
        if (R->xnx[l] != nx0 || R->xny[l] != ny0)
            fail(MT_EINVAL, "all layers must share one grid size (equal atmr.os)");
    if ((nx0 & 1) || (ny0 & 1)) fail(MT_EINVAL, "layer grids must have even size");
    G.fnx = nx0;
    G.fny = ny0;
    {
        std::vector<int2> rb;
        for (int l = 0; l < R->npsr; l++)
            for (int r0 = 0; r0 < ny0; r0 += ROWS) rb.push_back(make_int2(l, r0));
        G.nrowblk = (int)rb.size();
        g->rowblk.upload(rb);
        G.rowblk = g->rowblk.p;
    }
    long so = 0;
    for (int l = 0; l < R->npsr; l++) {
        G.soff[l] = so;
        so += (long)(nx0 / 2 + 1) * ny0;
    }
    G.soff[R->npsr] = so;
    G.nspec = so;
    if (G.precond) {
        if (!R->fd_Minv || !R->fd_permhf || R->fd_bs <= 0 || R->fd_nbhalf <= 0)
            fail(MT_EINVAL, "FDPCG data missing");
        auto rx = fft_radices(nx0), ry = fft_radices(ny0);
        G.nstx = (int)rx.size();
        G.nsty = (int)ry.size();
        fft_stages(rx, nx0, G.fsx);
        fft_stages(ry, ny0, G.fsy);
        G.dfnx = make_fdiv((unsigned)nx0);
        G.dfny = make_fdiv((unsigned)ny0);
        G.dnpsr = make_fdiv((unsigned)R->npsr);
        for (int d = 0; d <= ROWS / 2; d++) G.dsmall[d] = make_fdiv((unsigned)std::max(d, 1));
        G.rmax = 2;
        for (int r : rx) G.rmax = std::max(G.rmax, r);
        for (int r : ry) G.rmax = std::max(G.rmax, r);
        g->twx.upload(twiddles(nx0));
        g->twy.upload(twiddles(ny0));
        G.twx = g->twx.p;
        G.twy = g->twy.p;
        g->twx64.upload(twiddles64(nx0));
        g->twy64.upload(twiddles64(ny0));
        G.twx64 = g->twx64.p;
        G.twy64 = g->twy64.p;
        G.bs = R->fd_bs;
        G.zero_dc = 1;
        G.scale = R->fd_scale;
        G.fscale = (float)(1. / std::sqrt((double)nx0 * ny0));
        const long nbh = R->fd_nbhalf, bs = R->fd_bs;
        // decode permhf entries -> (layer, kx, ky, conj)
        const long nfull = (long)nx0 * ny0;
        struct Ent {
            int l, kx, ky, cj;
        };
        std::vector<Ent> ent(nbh * bs);
        for (long e = 0; e < nbh * bs; e++) {
            long v = R->fd_permhf[e];
            const int cj = v < 0;
            v = std::labs(v);
            const int l = (int)(v / nfull);
            const long rem = v - (long)l * nfull;
            if (l >= R->npsr) fail(MT_EINVAL, "permhf index out of range");
            ent[e] = {l, (int)(rem % nx0), (int)(rem / nx0), cj};
            if (ent[e].kx > nx0 / 2) fail(MT_EINVAL, "permhf entry outside the half spectrum");
        }
        // union of columns coupled by a block
        const int nh = nx0 / 2 + 1;
        std::vector<int> par(nh);
        std::iota(par.begin(), par.end(), 0);
        std::function<int(int)> find = [&](int a) { return par[a] == a ? a : par[a] = find(par[a]); };
        for (long b = 0; b < nbh; b++)
            for (long m = 1; m < bs; m++) {
                int a0 = find(ent[b * bs].kx), a1 = find(ent[b * bs + m].kx);
                if (a0 != a1) par[a1] = a0;
            }
        std::map<int, std::vector<int>> groups;
        for (int c = 0; c < nh; c++) groups[find(c)].push_back(c);
        std::vector<int> hgc, hkx, hgb, hbid, hbpos(nbh * bs);
        std::vector<int> colslot(nh), colgrp(nh);
        int gi = 0, maxcol = 0;
        for (auto &kv : groups) {
            hgc.push_back((int)hkx.size());
            for (size_t ci = 0; ci < kv.second.size(); ci++) {
                colslot[kv.second[ci]] = (int)ci;
                colgrp[kv.second[ci]] = gi;
                hkx.push_back(kv.second[ci]);
            }
            maxcol = std::max<int>(maxcol, (int)kv.second.size());
            gi++;
        }
        hgc.push_back((int)hkx.size());
        G.ngroup = gi;
        G.maxcol = maxcol;
        std::vector<std::vector<int>> blocks_of(gi);
        for (long b = 0; b < nbh; b++) blocks_of[colgrp[ent[b * bs].kx]].push_back((int)b);
        // With os=2 the column kx = nx/4 appears in two blocks (once directly,
        // once conjugated); the reference's kernel writes it from both
        // (cuda/fdpcg.cu:112-137), values equal by Hermitian symmetry.  Here
        // the first block in order writes it and later copies carry a
        // no-write flag (bit 29), so the result does not depend on scheduling.
        std::vector<int> written((size_t)gi * maxcol * R->npsr * ny0, 0);
        for (int gg = 0; gg < gi; gg++) {
            hgb.push_back((int)hbid.size());
            for (int b : blocks_of[gg]) {
                hbid.push_back(b);
                for (long m = 0; m < bs; m++) {
                    const Ent &e = ent[b * bs + m];
                    const int pos = (colslot[e.kx] * R->npsr + e.l) * ny0 + e.ky;
                    int &wr = written[(size_t)gg * maxcol * R->npsr * ny0 + pos];
                    hbpos[b * bs + m] = pos | (e.cj << 30) | (wr ? (1 << 29) : 0);
                    wr++;
                }
            }
        }
        hgb.push_back((int)hbid.size());
        // every half-spectrum entry must be produced
        for (int gg = 0; gg < gi; gg++) {
            const int ncol = hgc[gg + 1] - hgc[gg];
            for (int pidx = 0; pidx < ncol * R->npsr * ny0; pidx++)
                if (written[(size_t)gg * maxcol * R->npsr * ny0 + pidx] < 1)
                    fail(MT_EINVAL, "FDPCG half-spectrum blocks do not cover the spectrum");
        }
        g->grp_col.upload(hgc);
        g->colkx.upload(hkx);
        // blocks (positions and Minv) are stored in group order, so a column
        // workgroup reads one contiguous range
        g->blk_order = hbid;
        std::vector<int> hbpos_g(nbh * bs);
        for (long j = 0; j < nbh; j++)
            for (long m = 0; m < bs; m++) hbpos_g[j * bs + m] = hbpos[(long)hbid[j] * bs + m];
        g->grp_blk.upload(hgb);
        g->blkpos.upload(hbpos_g);
        G.grp_col = g->grp_col.p;
        G.colkx = g->colkx.p;
        G.grp_blk = g->grp_blk.p;
        G.blkpos = g->blkpos.p;
        g->Minv.upload(group_order(half_blocks(R, nx0, ny0), g->blk_order, bs));
        G.Minv = g->Minv.p;
        if (col_lds(G, 1, 1) > 64 * 1024) fail(MT_EINVAL, "FDPCG column group too large for LDS");
    }
    if (row_lds(G, 1) > 64 * 1024) fail(MT_EINVAL, "layer grid too large for the row FFT kernel");
}

void alloc_state(gpu_tomo *g) {
    const Geom &G = g->G;
    const long ns = g->nbatch;
    g->x.alloc(ns * G.nxtot);
    g->r.alloc(ns * G.nxtot);
    g->z.alloc(ns * G.nxtot);
    g->p.alloc(ns * G.nxtot);
    g->Ap.alloc(ns * G.nxtot);
    g->b.alloc(ns * G.nxtot);
    g->g.alloc(ns * G.ngtot);
    g->q.alloc(ns * G.nwfs * G.nploc);
    g->tt.alloc(ns * G.nwfs * 2);
    g->specf.alloc(ns * std::max<long>(G.nspec, 1));
    g->speci.alloc(ns * std::max<long>(G.nspec, 1));
    const int maxpart = std::max({layer_tiles(G) * G.npsr, G.nrowblk, G.ngroup, 1});
    g->part_pap.alloc(ns * maxpart);
    g->part_rz.alloc(ns * maxpart);
    g->part_rr.alloc(ns * maxpart);
    g->part_bb.alloc(ns * maxpart);
    g->rz.alloc(ns * (g->maxit + 2));
    g->alpha.alloc(ns);
    g->res.alloc(ns);
    g->stop.alloc(ns);
    HIPCHK(hipMemset(g->x.p, 0, sizeof(float) * ns * G.nxtot));
    HIPCHK(hipMemset(g->p.p, 0, sizeof(float) * ns * G.nxtot));
    HIPCHK(hipMemset(g->part_pap.p, 0, sizeof(double) * ns * maxpart));
    HIPCHK(hipMemset(g->part_rz.p, 0, sizeof(double) * ns * maxpart));
    HIPCHK(hipMemset(g->part_rr.p, 0, sizeof(double) * ns * maxpart));
    HIPCHK(hipMemset(g->part_bb.p, 0, sizeof(double) * ns * maxpart));
    Scal &SC = g->SC;
    SC.part_pap = g->part_pap.p;
    SC.part_rz = g->part_rz.p;
    SC.part_rr = g->part_rr.p;
    SC.part_bb = g->part_bb.p;
    SC.rz = g->rz.p;
    SC.alpha = g->alpha.p;
    SC.res = g->res.p;
    SC.tt = g->tt.p;
    SC.maxpart = maxpart;
    SC.maxit = g->maxit;
    SC.npart_layer = layer_tiles(G) * G.npsr;
    SC.npart_row = G.nrowblk;
    SC.npart_col = G.ngroup;
}

// ---------------------------------------------------------------------------
// Launch sequence.  Every launch goes through rec(), which brackets it with
// HIP events on the plan's stream when profiling is on.
// ---------------------------------------------------------------------------
struct Rec {
    gpu_tomo *g;
    hipStream_t st;
    bool on;
    template <typename F>
    void operator()(int kc, F &&f) {
        if (on) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            HIPCHK(hipEventRecord(a, st));
            f();
            HIPCHK(hipEventRecord(b, st));
            g->evs.push_back({kc, {a, b}});
        } else {
            f();
        }
    }
};

// Operator A p -> Ap for all systems (WFS pass + layer pass).
void enqueue_L(gpu_tomo *g, int nsys, const float *xin, void *xout, double *part, Rec &rec) {
    const Geom &G = g->G;
    rec(KC_WFS, [&] { launch_wfs(G, nsys, false, g->rt64, xin, nullptr, nullptr, g->q.p, rec.st); });
    rec(KC_LAYER, [&] { launch_layer(G, nsys, true, g->rt64, g->q.p, xin, xout, part, g->SC.maxpart, rec.st); });
}
// RHS b = HX' GP' W (I - TT PTT) g for all systems.
void enqueue_R(gpu_tomo *g, int nsys, const float *grad, void *bout, Rec &rec) {
    const Geom &G = g->G;
    bool anyttr = false;
    for (int k = 0; k < G.nwfs; k++) anyttr |= G.ttr[k] != 0;
    if (anyttr) rec(KC_TTR, [&] { launch_ttr(G, nsys, grad, g->tt.p, rec.st); });
    rec(KC_WFS_RHS, [&] { launch_wfs(G, nsys, true, g->rt64, nullptr, grad, g->tt.p, g->q.p, rec.st); });
    rec(KC_LAYER_RHS, [&] { launch_layer(G, nsys, false, g->rt64, g->q.p, nullptr, bout, g->SC.part_bb, g->SC.maxpart, rec.st); });
}
// Preconditioner on the row-transformed residual (specf): column pass with
// the spectral r.z partials, then the inverse row pass in the given mode.
void enqueue_P_tail(gpu_tomo *g, int nsys, int irow_mode, int slot, float *out, Rec &rec) {
    const Geom &G = g->G;
    rec(KC_COL, [&] { launch_col(G, g->SC, nsys, g->spec64, g->specf.p, g->speci.p, rec.st); });
    rec(irow_mode == IROW_MODE_UPDATE ? KC_IROW : KC_IROW_START,
        [&] { launch_irow(G, g->SC, nsys, irow_mode, slot, g->speci.p, g->x.p, out, rec.st); });
}

// The solve as lib/pcg.c:28-113 runs it (restart every 100 iterations):
//   start:  r = b - A x (r = b cold, x = 0); z = M r; p = z; rz[0] = r.z
//   iter k: Ap; alpha = rz[k] / p.Ap; r -= alpha Ap;
//           last: x += alpha p, |r| -> residual; else z = M r,
//           beta = r.z / rz[k], x += alpha p, p = z + beta p
// Kernels per iteration: wfs, layer (Ap, p.Ap partials), row (alpha, r,
// forward row FFT), col (column FFTs, blocks, r.z partials), irow (inverse
// row FFT, beta, x and p).
void enqueue_solve(gpu_tomo *g, int nsys, int warm, Rec &rec) {
    const Geom &G = g->G;
    Scal &SC = g->SC;
    const int maxit = g->maxit;
    if (!G.precond) fail(MT_EINVAL, "the GPU solve requires the FDPCG preconditioner (tomo.precond=1)");
    launch_zero_i(g->stop.p, nsys, rec.st);
    enqueue_R(g, nsys, g->g.p, g->b.p, rec);
    auto start = [&](int k, bool cold) {
        if (cold) {
            rec(KC_ROW_START, [&] {
                launch_row(G, SC, nsys, ROW_MODE_INIT_COLD, g->rt64, g->spec64, k, g->x.p, g->r.p, g->b.p, g->p.p, g->Ap.p,
                           g->specf.p, g->stop.p, rec.st);
            });
        } else {
            enqueue_L(g, nsys, g->x.p, g->Ap.p, nullptr, rec);
            rec(KC_ROW_START, [&] {
                launch_row(G, SC, nsys, ROW_MODE_INIT_WARM, g->rt64, g->spec64, k, g->x.p, g->r.p, g->b.p, g->p.p, g->Ap.p,
                           g->specf.p, g->stop.p, rec.st);
            });
        }
        enqueue_P_tail(g, nsys, IROW_MODE_FIRST, k, g->p.p, rec);
    };
    start(0, !warm);
    for (int k = 0; k < maxit; k++) {
        if (k > 0 && k % 100 == 0) start(k, false);
        enqueue_L(g, nsys, g->p.p, g->Ap.p, SC.part_pap, rec);
        const bool last = k + 1 == maxit;
        rec(last ? KC_ROW_LAST : KC_ROW, [&] {
            launch_row(G, SC, nsys, last ? ROW_MODE_LAST : ROW_MODE_UPDATE, g->rt64, g->spec64, k, g->x.p, g->r.p, g->b.p, g->p.p,
                       g->Ap.p, g->specf.p, g->stop.p, rec.st);
        });
        if (last) break;
        enqueue_P_tail(g, nsys, IROW_MODE_UPDATE, k, g->p.p, rec);
    }
    rec(KC_FINAL, [&] { launch_final(G, SC, nsys, G.nrowblk, g->stop.p, rec.st); });
}

void prof_collect(gpu_tomo *g) {
    HIPCHK(hipStreamSynchronize(g->stream));
    for (int i = 0; i < KC_N; i++) {
        g->prof_ms[i] = 0;
        g->prof_cnt[i] = 0;
    }
    for (auto &e : g->evs) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
        g->prof_ms[e.first] += ms;
        g->prof_cnt[e.first]++;
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    g->evs.clear();
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return MT_OK;
    } catch (const HipError &e) {
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

void to_dev_grad(gpu_tomo *g, int isys, const double *grad) {
    // caller layout (all WFS, skipped ones empty) -> active-WFS layout
    const Geom &G = g->G;
    g->hstage.assign(G.ngtot, 0.f);
    for (int k = 0; k < G.nwfs; k++) {
        const int w = g->wfs_active[k];
        const long o = g->goff_all[w];
        for (long i = 0; i < 2 * G.nsa[k]; i++) {
            const double v = grad[o + i];
            if (!std::isfinite(v)) fail(MT_ENOTFINITE, "non-finite gradient input");
            g->hstage[G.goff[k] + i] = (float)v;
        }
    }
    HIPCHK(hipMemcpyAsync(g->g.p + (long)isys * G.ngtot, g->hstage.data(), sizeof(float) * G.ngtot,
                          hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
}
void to_dev_x(gpu_tomo *g, float *dst, const double *src) {
    g->hstage.resize(g->nxtot);
    for (long i = 0; i < g->nxtot; i++) {
        if (!std::isfinite(src[i])) fail(MT_ENOTFINITE, "non-finite input vector");
        g->hstage[i] = (float)src[i];
    }
    HIPCHK(hipMemcpyAsync(dst, g->hstage.data(), sizeof(float) * g->nxtot, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
}
void to_dev_rt(gpu_tomo *g, double *dst, const double *src) {
    for (long i = 0; i < g->nxtot; i++)
        if (!std::isfinite(src[i])) fail(MT_ENOTFINITE, "non-finite input vector");
    if (g->rt64) {
        HIPCHK(hipMemcpy(dst, src, sizeof(double) * g->nxtot, hipMemcpyHostToDevice));
    } else {
        g->hstage.resize(g->nxtot);
        for (long i = 0; i < g->nxtot; i++) g->hstage[i] = (float)src[i];
        HIPCHK(hipMemcpy(dst, g->hstage.data(), sizeof(float) * g->nxtot, hipMemcpyHostToDevice));
    }
}
void from_dev_rt(gpu_tomo *g, double *dst, const double *src) {
    HIPCHK(hipStreamSynchronize(g->stream));
    if (g->rt64) {
        HIPCHK(hipMemcpy(dst, src, sizeof(double) * g->nxtot, hipMemcpyDeviceToHost));
    } else {
        g->hstage.resize(g->nxtot);
        HIPCHK(hipMemcpy(g->hstage.data(), src, sizeof(float) * g->nxtot, hipMemcpyDeviceToHost));
        for (long i = 0; i < g->nxtot; i++) dst[i] = g->hstage[i];
    }
}
void from_dev_x(gpu_tomo *g, double *dst, const float *src) {
    g->hstage.resize(g->nxtot);
    HIPCHK(hipMemcpyAsync(g->hstage.data(), src, sizeof(float) * g->nxtot, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    for (long i = 0; i < g->nxtot; i++) dst[i] = g->hstage[i];
}

}  // namespace

extern "C" {

int gpu_tomo_setup(const mt_recon_t *recon, int device, int nbatch, int maxit, int warm, gpu_tomo_t **out) {
    if (!out) return MT_EINVAL;
    *out = nullptr;
    if (nbatch <= 0 || maxit <= 0 || maxit > 10000) {
        mt::set_error("gpu_tomo_setup: nbatch and maxit must be positive");
        return MT_EINVAL;
    }
    std::unique_ptr<gpu_tomo> g(new (std::nothrow) gpu_tomo());
    if (!g) return MT_ENOMEM;
    int rc = guarded([&] {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(MT_EINVAL, "gpu_tomo_setup: no such device");
        g->device = device;
        HIPCHK(hipSetDevice(device));
        g->nbatch = nbatch;
        g->maxit = maxit;
        g->warm = warm ? 1 : 0;
        build_plan(g.get(), recon);
        alloc_state(g.get());
        HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        HIPCHK(hipDeviceSynchronize());
    });
    if (rc == MT_OK) *out = g.release();
    return rc;
}

int gpu_tomo_update_cn2(gpu_tomo_t *g, const mt_recon_t *recon) {
    if (!g || !recon) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipStreamSynchronize(g->stream));
        if (recon->npsr != g->npsr) fail(MT_EINVAL, "layer count changed");
        for (int l = 0; l < recon->npsr; l++) g->G.l2c[l] = (float)recon->l2c[l];
        if (g->G.precond) {
            const long nbh = recon->fd_nbhalf, bs = recon->fd_bs;
            if (bs != g->G.bs || (long)g->Minv.n != nbh * bs * bs) fail(MT_EINVAL, "FDPCG layout changed");
            std::vector<float2> hM = group_order(half_blocks(recon, g->G.fnx, g->G.fny), g->blk_order, bs);
            HIPCHK(hipMemcpy(g->Minv.p, hM.data(), sizeof(float2) * hM.size(), hipMemcpyHostToDevice));
        }
        // kernel arguments changed: drop the captured graph
        if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->gexec = nullptr;
        g->graph = nullptr;
    });
}

int gpu_tomo_reset(gpu_tomo_t *g) {
    if (!g) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipMemsetAsync(g->x.p, 0, sizeof(float) * g->x.n, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

void gpu_tomo_free(gpu_tomo_t *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto &e : g->evs) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int gpu_tomo_sizes(const gpu_tomo_t *g, long *nx, long *ng) {
    if (!g) return MT_EINVAL;
    if (nx) *nx = g->nxtot;
    if (ng) *ng = g->ngtot_all;
    return MT_OK;
}

int gpu_tomo_solve(gpu_tomo_t *g, int isys, const double *grad, double *x, double *cgres) {
    if (!g || !grad || !x || isys < 0 || isys >= g->nbatch) return MT_EINVAL;
    double res = NAN;
    int rc = guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (g->nbatch != 1) fail(MT_EINVAL, "gpu_tomo_solve is the one-system entry point (nbatch must be 1)");
        to_dev_grad(g, 0, grad);
        Rec rec{g, g->stream, false};
        enqueue_solve(g, 1, g->warm, rec);
        from_dev_x(g, x, g->x.p);
        HIPCHK(hipMemcpy(&res, g->res.p, sizeof(double), hipMemcpyDeviceToHost));
    });
    if (cgres) *cgres = res;
    if (rc == MT_OK && !(res <= 1.0)) rc = MT_NOTCONV;  // cuda/solve.cu:23-32 warning case
    return rc;
}

int gpu_tomo_solve_batch(gpu_tomo_t *g, int nsys, const double *grad, double *x, double *cgres) {
    if (!g || !grad || !x || nsys <= 0 || nsys > g->nbatch) return MT_EINVAL;
    std::vector<double> res(nsys, NAN);
    int rc = guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        for (int s = 0; s < nsys; s++) to_dev_grad(g, s, grad + (long)s * g->ngtot_all);
        Rec rec{g, g->stream, false};
        enqueue_solve(g, nsys, g->warm, rec);
        for (int s = 0; s < nsys; s++) from_dev_x(g, x + (long)s * g->nxtot, g->x.p + (long)s * g->G.nxtot);
        HIPCHK(hipMemcpy(res.data(), g->res.p, sizeof(double) * nsys, hipMemcpyDeviceToHost));
    });
    if (cgres)
        for (int s = 0; s < nsys; s++) cgres[s] = res[s];
    if (rc == MT_OK)
        for (int s = 0; s < nsys; s++)
            if (!(res[s] <= 1.0)) rc = MT_NOTCONV;
    return rc;
}

int gpu_tomo_solve_dev(gpu_tomo_t *g, const float *d_grad, float *d_x, int nbatch, int mode, void *stream) {
    if (!g || nbatch <= 0 || nbatch > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        hipStream_t st = stream ? (hipStream_t)stream : g->stream;
        const Geom &G = g->G;
        if (d_grad && d_grad != g->g.p) launch_copy(d_grad, g->g.p, (long)nbatch * G.ngtot, st);
        if (d_x && d_x != g->x.p && g->warm) launch_copy(d_x, g->x.p, (long)nbatch * G.nxtot, st);
        if (mode == 1 && !g->prof) {
            if (!g->gexec || g->graph_nsys != nbatch) {
                if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
                if (g->graph) (void)hipGraphDestroy(g->graph);
                g->gexec = nullptr;
                g->graph = nullptr;
                hipStream_t cs;
                HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                Rec rec{g, cs, false};
                HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_solve(g, nbatch, g->warm, rec);
                } catch (...) {
                    hipGraph_t junk;
                    (void)hipStreamEndCapture(cs, &junk);
                    (void)hipStreamDestroy(cs);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(cs, &g->graph));
                (void)hipStreamDestroy(cs);
                HIPCHK(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
                g->graph_nsys = nbatch;
            }
            HIPCHK(hipGraphLaunch(g->gexec, st));
        } else {
            Rec rec{g, st, g->prof != 0};
            enqueue_solve(g, nbatch, g->warm, rec);
            if (g->prof) {
                if (st != g->stream) HIPCHK(hipStreamSynchronize(st));
                prof_collect(g);
            }
        }
        if (d_x && d_x != g->x.p) launch_copy(g->x.p, d_x, (long)nbatch * G.nxtot, st);
    });
}

int gpu_tomo_residuals(gpu_tomo_t *g, double *cgres, int nbatch) {
    if (!g || !cgres || nbatch <= 0 || nbatch > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(cgres, g->res.p, sizeof(double) * nbatch, hipMemcpyDeviceToHost));
    });
}

int gpu_tomo_L(gpu_tomo_t *g, const double *xin, double *xout, double alpha, double beta) {
    if (!g || !xin || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        to_dev_x(g, g->z.p, xin);
        Rec rec{g, g->stream, false};
        enqueue_L(g, 1, g->z.p, g->Ap.p, nullptr, rec);
        std::vector<double> tmp(g->nxtot);
        from_dev_rt(g, tmp.data(), g->Ap.p);
        for (long i = 0; i < g->nxtot; i++) xout[i] = beta * xout[i] + alpha * tmp[i];
    });
}

int gpu_tomo_R(gpu_tomo_t *g, const double *grad, double *xout, double alpha) {
    if (!g || !grad || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        to_dev_grad(g, 0, grad);
        Rec rec{g, g->stream, false};
        enqueue_R(g, 1, g->g.p, g->b.p, rec);
        from_dev_rt(g, xout, g->b.p);
        for (long i = 0; i < g->nxtot; i++) xout[i] *= alpha;
    });
}

int gpu_tomo_P(gpu_tomo_t *g, const double *xin, double *xout) {
    if (!g || !xin || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (!g->G.precond) fail(MT_EINVAL, "no FDPCG preconditioner in this plan");
        // b <- xin, r <- b through the cold-start row kernel (which also
        // clears its x argument: the p buffer is used as that scratch), then P
        to_dev_rt(g, g->b.p, xin);
        Rec rec{g, g->stream, false};
        launch_zero_i(g->stop.p, 1, g->stream);
        launch_row(g->G, g->SC, 1, ROW_MODE_INIT_COLD, g->rt64, g->spec64, 0, g->p.p, g->r.p, g->b.p, g->p.p, g->Ap.p,
                   g->specf.p, g->stop.p, g->stream);
        enqueue_P_tail(g, 1, IROW_MODE_Z, 0, g->z.p, rec);
        from_dev_x(g, xout, g->z.p);
    });
}

int gpu_tomo_precision(gpu_tomo_t *g, int mode) {
    if (!g || mode < 0 || mode > 2) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipStreamSynchronize(g->stream));
        g->rt64 = mode >= 1;
        g->spec64 = mode == 1;
        if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->gexec = nullptr;
        g->graph = nullptr;
    });
}

int gpu_tomo_profile(gpu_tomo_t *g, int enable) {
    if (!g) return MT_EINVAL;
    g->prof = enable ? 1 : 0;
    return MT_OK;
}

int gpu_tomo_profile_read(gpu_tomo_t *g, const char **names, double *ms, double *calls, int *n) {
    if (!g || !n) return MT_EINVAL;
    const int cap = *n;
    int k = 0;
    for (int i = 0; i < KC_N && k < cap; i++) {
        if (!g->prof_cnt[i]) continue;
        if (names) names[k] = kc_names[i];
        if (ms) ms[k] = g->prof_ms[i] / g->prof_cnt[i];
        if (calls) calls[k] = (double)g->prof_cnt[i];
        k++;
    }
    *n = k;
    return MT_OK;
}

}  // extern "C"
