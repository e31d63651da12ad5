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
#include <set>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "fit_internal.hpp"
#include "rays.hpp"
#include "solve64.hpp"
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
// Status of an FP64-plan call (which has set the error message) as an exception.
void chk64(int rc) {
    if (rc != MT_OK) fail(rc, mt_last_error());
}

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

// Kernel classes timed by the profiler: the iteration kernels (k_wfs ..
// k_irow, the UPDATE-iteration forms) and the once-per-solve forms apart.
enum KClass {
    KC_TTR, KC_WFS, KC_LAYER, KC_ROW, KC_COL, KC_IROW, KC_FINAL,
    KC_WFS_RHS, KC_LAYER_RHS, KC_ROW_START, KC_ROW_LAST, KC_IROW_START, KC_N
};
const char *kc_names[KC_N] = {"k_ttr", "k_wfs", "k_layer", "k_row", "k_col", "k_irow", "k_final",
                              "k_wfs_rhs", "k_layer_rhs", "k_row_start", "k_row_last", "k_irow_start"};

}  // namespace

struct gpu_tomo;

// ---------------------------------------------------------------------------
// Workspace of gpu_tomo_solve_dev_noalloc: one file-scope device array per
// GPU (held until the process ends), so that entry point needs no device
// allocation.  Host code never uses the array's name as an address (on the
// host it is only a stand-in): at set-up, with the device current,
// hipGetSymbolAddress gives that device's copy, kept per device below.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
namespace {
constexpr int NOALLOC_MAXDEV = 64;
std::mutex scratch_mu;
void *scratch_ptr[NOALLOC_MAXDEV] = {nullptr};
}  // namespace

namespace {
// One launch of a stage: a template variant over a subset of the jobs.
struct Sub {
    int matched = 0, pos = 1, cs = 1, bs = 0, rm = 3;
    int reg = 0;  // column stage: P of the register kernel (0: the LDS kernel)
    int rp = 0;   // row stages: P of the register row FFT (rows of 32 P points), or 0
    int nthr = 0; // its threads per workgroup
    int regL = 0; // its lanes per column FFT (shL of every job)
    size_t lds = 0;
    DevBuf<WGI> wgi;
    int nwg = 0;
};
typedef std::vector<std::unique_ptr<Sub>> Stage;
// A list of jobs (system ranges of plans, possibly of different geometry)
// solved in lockstep, and the workgroup tables of every stage.
struct Batch {
    std::vector<gpu_tomo *> plan;   // per job
    std::vector<Job> hjob;
    DevBuf<Job> jobs;
    int rt64 = 1, spec64 = 1, maxit = 30, warm = 0, precond = 1, anyttr = 0;
    Stage zero, ttr, wfs_r, layer, wfs_l, row, col, irow, fin;
    Stage wfs_g, ttl;  // LHS with TT removal: gradient tiles, TT projection (ptt_lhs jobs)
    Stage cg;          // plain CG vector updates over the row blocks (no preconditioner)
    // a job table to launch with instead of `jobs` (same jobs, other device
    // records: gpu_tomo_solve_dev_noalloc's workspace); null normally
    const Job *jobs_over = nullptr;
};
}  // namespace

// One profiled interval: n launches of kernel class kc between events a, b.
struct ProfEv {
    int kc, n;
    hipEvent_t a, b;
};

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
    DevBuf<float2> twx, twy, Minv, Mreg;
    DevBuf<double2> twx64, twy64;
    DevBuf<int> grp_col, colkx, grp_blk, blkpos;
    std::vector<int> blk_order;  // half-spectrum block id of each block in group order
    std::vector<int> blkpos_host;  // blkpos, host copy
    std::vector<int> grp_blk_host; // grp_blk, host copy (register column layout)
    DevBuf<int2> rowblk;
    // state
    DevBuf<float> x, z, p, g, tt;
    DevBuf<double> r, Ap, b, q;  // residual side: double, or float in the first half when !rt64
    DevBuf<double> gl, ttl;      // LHS gradients and their TT projection (ptt_lhs plans only)
    int rt64 = 1;             // 1: r, Ap, b in FP64 (default); 0: FP32
    int spec64 = 1;           // 1: forward FDPCG spectrum in FP64; 0: FP32
    size_t lds_max = 64 * 1024;  // LDS one workgroup may use on this device
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
    hipEvent_t gdone = nullptr;  // recorded after each launch of gexec
    // flights: a graph solve of n systems as k graphs over consecutive
    // system ranges, each on its own stream, forked from and joined to the
    // caller's stream (the kernels of one flight fill the ramps and tails of
    // the others'); graphs per (n, k), streams and events kept
    struct Flight {
        hipStream_t st = nullptr;
        hipEvent_t done = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Flight> fl;
    int fl_nsys = 0, fl_k = 0;
    hipEvent_t fork = nullptr;
    int tune_flights = 0;  // 1, 2 or 4 (0: 1)
    // device copy of geometry + buffers, and the launch tables of this plan's
    // own system ranges (key: first system, count)
    DevBuf<PlanDev> pdev;
    std::map<std::pair<int, int>, std::unique_ptr<Batch>> own;
    // systems per workgroup chosen by the autotuner (0: heuristic)
    int tune_wfs = 0, tune_layer = 0, tune_col = 0;
    long gen = 0;         // bumped whenever the launch tables must be rebuilt (shape, precision)
    int tune_lty = 0;     // layer tile rows: 0 = heuristic (LTY where it fits), TILE or LTY
    int tune_colfft = 0;  // column kernel: 0 = registers where available, 1 = registers, 2 = LDS stages
    bool colr_ok = false; // the register column kernel applies (per-frequency blocks, ny = shL x shP)
    bool lty_ok = false;  // LTY-row tiles fit every layer of this (matched) plan
    bool lty2_ok = false; // LTY2-row tiles do
    // precision mode 4 (all FP64): a separate FP64 plan, made from a host
    // copy of the recon when the mode is selected, which holds the warm-start
    // state while the mode is active
    recon64 *rc64 = nullptr;
    tomo64 *t64 = nullptr;
    // profiling
    int prof = 0;
    std::vector<ProfEv> evs;  // profiling events
    double prof_ms[KC_N] = {0};
    int prof_cnt[KC_N] = {0};
};

namespace {

// Entry points of the FP32-operator path (device buffers, groups, the
// autotuner, MVM assembly) refuse a plan in precision mode 4.
void no_mode4(const gpu_tomo *g, const char *what) {
    if (g->t64)
        fail(MT_EINVAL, std::string(what) + ": not available in precision mode 4 (all FP64), which runs through "
                                            "gpu_tomo_solve, gpu_tomo_solve_batch and the operator applies");
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
// The blocks as the column kernel reads them: group order, and with
// zero_dc the columns of DC entries zeroed, which applies each block to the
// residual spectrum with its DC bins set to zero (the kernel leaves the
// input as it is and drops those bins from r.z).
std::vector<float2> col_blocks(const mt_recon_t *R, const gpu_tomo *g) {
    const long bs = R->fd_bs;
    std::vector<float2> M = group_order(half_blocks(R, g->G.fnx, g->G.fny), g->blk_order, bs);
    if (g->G.zero_dc)
        for (size_t j = 0; j < g->blk_order.size(); j++)
            for (long k = 0; k < bs; k++)
                if (g->blkpos_host[j * bs + k] & BP_DC)
                    for (long m = 0; m < bs; m++) M[j * bs * bs + m + k * bs] = make_float2(0.f, 0.f);
    return M;
}

// Blocks for the register column kernel, which applies where every block
// couples one half-spectrum bin (kx, ky) across the layers (member m on layer
// m, one conjugation flag) and every column is a group of its own: per
// (column c, k1, entry e) the L lanes, lane l holding the block of the bin
// ky = k1 + P bitrev_L(l), conjugated where the block's frequency is the
// conjugate of its bin (conj(M) x = conj(M conj(x))).  Otherwise the plan
// keeps the LDS column kernel only (colr_ok false, shL 0).
// Returns false (the layout does not apply) or fills Mr.  hgb: first block
// of each column group (group order).
bool mreg_layout(const gpu_tomo *g, const std::vector<float2> &hM, const std::vector<int> &hgb,
                 std::vector<float2> &Mr) {
    const Geom &G = g->G;
    const int L = G.shL, P = G.shP, ny = G.fny, bs = G.bs, ngr = G.ngroup;
    auto none = [&] { return false; };
    if (!L || G.maxcol != 1 || bs != G.npsr) return none();
    std::vector<int> jof((size_t)ngr * ny, -1);  // block of (column, ky)
    std::vector<char> cjof((size_t)ngr * ny, 0);
    for (int c = 0; c < ngr; c++) {
        for (int j = hgb[c]; j < hgb[c + 1]; j++) {
            const int p0 = g->blkpos_host[(size_t)j * bs];
            const int ky = (p0 & BP_POS) % ny;
            for (int m = 0; m < bs; m++) {
                const int pe = g->blkpos_host[(size_t)j * bs + m];
                if ((pe & BP_POS) != m * ny + ky || (pe & BP_NOWRITE) || ((pe ^ p0) & BP_CONJ)) return none();
            }
            if (jof[(size_t)c * ny + ky] >= 0) return none();
            jof[(size_t)c * ny + ky] = j;
            cjof[(size_t)c * ny + ky] = (p0 & BP_CONJ) != 0;
        }
        for (int ky = 0; ky < ny; ky++)
            if (jof[(size_t)c * ny + ky] < 0) return none();
    }
    int lb = 0;
    while ((1 << lb) < L) lb++;
    auto bitrev = [&](int l) {
        int r = 0;
        for (int b = 0; b < lb; b++) r |= ((l >> b) & 1) << (lb - 1 - b);
        return r;
    };
    const size_t nb2 = (size_t)bs * bs;
    Mr.assign((size_t)ngr * ny * nb2, make_float2(0.f, 0.f));
    for (int c = 0; c < ngr; c++)
        for (int k1 = 0; k1 < P; k1++)
            for (int l = 0; l < L; l++) {
                const size_t q = (size_t)c * ny + k1 + (size_t)P * bitrev(l);
                const float2 *blk = hM.data() + (size_t)jof[q] * nb2;
                for (size_t e = 0; e < nb2; e++) {
                    float2 v = blk[e];
                    if (cjof[q]) v.y = -v.y;
                    Mr[(((size_t)c * P + k1) * nb2 + e) * L + l] = v;
                }
            }
    return true;
}
void build_mreg(gpu_tomo *g, const std::vector<float2> &hM, const std::vector<int> &hgb) {
    Geom &G = g->G;
    g->colr_ok = false;
    G.Mreg = nullptr;
    std::vector<float2> Mr;
    if (!mreg_layout(g, hM, hgb, Mr)) {
        G.shL = G.shP = 0;
        return;
    }
    g->Mreg.upload(Mr);
    G.Mreg = g->Mreg.p;
    g->colr_ok = true;
    g->grp_blk_host = hgb;
}

void build_plan(gpu_tomo *g, const mt_recon_t *R) {
    if (!R || R->npsr <= 0 || R->npsr > MAXL || R->nwfs <= 0) fail(MT_EINVAL, "invalid recon (npsr/nwfs)");
    if (R->pos != 1 && R->pos != 2) fail(MT_EINVAL, "tomo.pos must be 1 or 2");
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
    // TT removal inside the LHS: the same TTFR as the RHS, on the operator's
    // own gradients (maos/recon_utils.c:461-465); the GPU reference removes
    // it there too when !split || splitlrt (cuda/tomo.cu:183)
    G.ptt_lhs = R->ptt_lhs ? 1 : 0;

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
            m.nrow = std::max(m.ay1 - m.ay0, 0);
            m.pad_ = 0;
            m.qoff = (long)k * pnx * pny + (long)m.ay0 * pnx;
        }
    // the matched kernels wrap the L2 window once: layers of at least TILE + 2
    // (16 x 16 tiles), LTY2 + 2 (16 x 32) or LTY + 2 rows (16 x 64 tiles)
    G.lty = TILE;
    for (int l = 0; l < R->npsr; l++)
        if (R->xnx[l] < TILE + 2 || R->xny[l] < TILE + 2) matched = 0;
    if (matched) {
        G.lty = LTY;
        for (int l = 0; l < R->npsr; l++)
            if (R->xny[l] < LTY + 2) G.lty = TILE;
    }
    g->lty_ok = G.lty == LTY;
    g->lty2_ok = matched != 0;
    for (int l = 0; l < R->npsr; l++)
        if (R->xny[l] < LTY2 + 2) g->lty2_ok = false;
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
    // Row blocks (ROWS rows of one layer) carry the PCG vector updates: the
    // row FFT kernels with FDPCG, the plain CG kernels without it (where
    // layers may differ in size).
    const int nx0 = (int)R->xnx[0], ny0 = (int)R->xny[0];
    if (G.precond) {
        for (int l = 0; l < R->npsr; l++)
            if (R->xnx[l] != nx0 || R->xny[l] != ny0)
                fail(MT_EINVAL, "all layers must share one grid size (equal atmr.os)");
        if ((nx0 & 1) || (ny0 & 1)) fail(MT_EINVAL, "layer grids must have even size");
        if (nx0 > MAXN || ny0 > MAXN) fail(MT_EINVAL, "layer grid larger than 1024");
    }
    G.fnx = nx0;
    G.fny = ny0;
    {
        std::vector<int2> rb;
        for (int l = 0; l < R->npsr; l++)
            for (int r0 = 0; r0 < R->xny[l]; r0 += ROWS) rb.push_back(make_int2(l, r0));
        G.nrowblk = (int)rb.size();
        g->rowblk.upload(rb);
        G.rowblk = g->rowblk.p;
    }
    long so = 0;
    for (int l = 0; l < R->npsr; l++) {
        G.soff[l] = so;
        if (G.precond) so += (long)(nx0 / 2 + 1) * ny0;
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
        // register layout of the column FFT: ny = L lanes x P points, L a
        // power of two dividing 64 (the widest first), P an in-lane radix
        G.shL = G.shP = 0;
        for (int L : {32, 64, 16, 8, 4}) {
            const int P = ny0 % L ? 0 : ny0 / L;
            if (colr_supported(P, R->fd_bs)) {
                G.shL = L;
                G.shP = P;
                break;
            }
        }
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
                    hbpos[b * bs + m] = pos | (e.cj ? BP_CONJ : 0) | (wr ? BP_NOWRITE : 0) |
                                        (e.kx == 0 && e.ky == 0 ? BP_DC : 0) |
                                        (e.kx == 0 || 2 * e.kx == nx0 ? BP_W1 : 0);
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
        g->blkpos_host = hbpos_g;
        G.grp_col = g->grp_col.p;
        G.colkx = g->colkx.p;
        G.grp_blk = g->grp_blk.p;
        G.blkpos = g->blkpos.p;
        if ((long)G.maxcol * R->npsr * ny0 > BP_POS) fail(MT_EINVAL, "FDPCG column group too large");
        const std::vector<float2> hM = col_blocks(R, g);
        g->Minv.upload(hM);
        G.Minv = g->Minv.p;
        build_mreg(g, hM, hgb);
        if (col_lds(G.maxcol, G.npsr, G.fny, 1, 1) > 64 * 1024)
            fail(MT_EINVAL, "FDPCG column group too large for LDS");
    }
    if (G.precond && row_lds(G.fnx, 1) > 64 * 1024) fail(MT_EINVAL, "layer grid too large for the row FFT kernel");
}

// Tiles of the layer passes (the largest layer decides; all layers are
// equal when FDPCG is used).
int layer_tiles(const Geom &G) {
    int t = 1;
    for (int l = 0; l < G.npsr; l++)
        t = std::max(t, ((G.xnx[l] + TILE - 1) / TILE) * ((G.xny[l] + G.lty - 1) / G.lty));
    return t;
}

void upload_plandev(gpu_tomo *g);
void upload_plandev(gpu_tomo *g, hipStream_t st);
void drop_batches(gpu_tomo *g);
// Apply the layer tile height choice (the plan's device record and
// partial counts follow; launch tables are rebuilt on next use).
void apply_lty(gpu_tomo *g) {
    if (!g->G.matched) return;
    const int lty = g->tune_lty == TILE   ? TILE
                    : g->tune_lty == LTY2 ? (g->lty2_ok ? LTY2 : TILE)
                    : g->lty_ok           ? LTY
                                          : TILE;
    if (lty == g->G.lty) return;
    g->G.lty = lty;
    g->SC.npart_layer = layer_tiles(g->G) * g->G.npsr;
    upload_plandev(g);
    drop_batches(g);
}

void upload_plandev(gpu_tomo *g, hipStream_t st) {
    PlanDev h{};
    h.G = g->G;
    h.x = g->x.p;
    h.z = g->z.p;
    h.p = g->p.p;
    h.g = g->g.p;
    h.tt = g->tt.p;
    h.r = g->r.p;
    h.Ap = g->Ap.p;
    h.b = g->b.p;
    h.q = g->q.p;
    h.specf = g->specf.p;
    h.speci = g->speci.p;
    h.SC = g->SC;
    h.stop = g->stop.p;
    h.ltiles = layer_tiles(g->G);
    h.gl = g->gl.p;
    h.ttl = g->ttl.p;
    if (!g->pdev.p) g->pdev.alloc(1);
    HIPCHK(hipMemcpyAsync(g->pdev.p, &h, sizeof(PlanDev), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
}
void upload_plandev(gpu_tomo *g) { upload_plandev(g, g->stream); }

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
    if (G.ptt_lhs) {
        g->gl.alloc(ns * G.ngtot);
        g->ttl.alloc(ns * G.nwfs * 2);
    }
    g->specf.alloc(ns * std::max<long>(G.nspec, 1));
    g->speci.alloc(ns * std::max<long>(G.nspec, 1));
    Geom g16 = G;  // partials sized for the smaller layer tiles (either height may be chosen later)
    g16.lty = TILE;
    const int maxpart = std::max({layer_tiles(g16) * G.npsr, G.nrowblk, G.ngroup, 1});
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
    upload_plandev(g);
}

// ---------------------------------------------------------------------------
// Launch sequence.  Every launch goes through rec(), which brackets it with
// HIP events on the plan's stream when profiling is on.
// ---------------------------------------------------------------------------
typedef std::vector<ProfEv> EvList;
// Kernel classes whose launches only read their inputs (a repeated launch
// writes the same outputs), so the profiler may launch them several times
// back to back inside one event pair.
bool kc_idempotent(int kc) {
    return kc == KC_TTR || kc == KC_WFS || kc == KC_LAYER || kc == KC_COL || kc == KC_WFS_RHS ||
           kc == KC_LAYER_RHS;
}
struct Rec {
    EvList *evs;
    hipStream_t st;
    bool on;
    int reps = 1;  // launches per event pair of an idempotent class
    template <typename F>
    void operator()(int kc, F &&f) {
        if (on) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            HIPCHK(hipEventRecord(a, st));
            const int n = kc_idempotent(kc) ? std::max(reps, 1) : 1;
            for (int i = 0; i < n; i++) f();
            HIPCHK(hipEventRecord(b, st));
            evs->push_back(ProfEv{kc, n, a, b});
        } else {
            f();
        }
    }
};

// ---------------------------------------------------------------------------
// Batches: the workgroup tables of every stage for a list of jobs.
// ---------------------------------------------------------------------------
struct JobGrid {
    int gx, gy, gz;
};
// Workgroup i of a launch runs on XCD i mod 8 (round-robin dispatch), and
// each XCD has its own L2.  Reorder a table so that all workgroups of one
// (job, system group) land on one XCD: the data they share (the layers read
// by every WFS tile, the WFS fields read by every layer tile) then stays in
// one L2 instead of being fetched by all eight.  Buckets go to the least
// loaded XCD in order; positions whose XCD has run out take the longest
// remaining queue, so a table with fewer buckets than XCDs keeps its order.
// by: the bucket's block index, XCD_BY_Z (system group) by default;
// XCD_BY_X / XCD_BY_Y for the column kernels, whose shared data (the FDPCG
// blocks of a column group) follows x (LDS kernel) or y (register kernel).
constexpr int NXCD = 8;
enum { XCD_BY_Z = 0, XCD_BY_X = 1, XCD_BY_Y = 2 };
std::vector<WGI> xcd_order(std::vector<WGI> tab, int by) {
    auto bkey = [&](const WGI &a) { return by == XCD_BY_X ? a.bx : by == XCD_BY_Y ? a.by : a.bz; };
    if (by != XCD_BY_Z)
        std::stable_sort(tab.begin(), tab.end(), [&](const WGI &a, const WGI &b) {
            return a.job != b.job ? a.job < b.job : bkey(a) < bkey(b);
        });
    auto same = [&](const WGI &a, const WGI &b) { return a.job == b.job && bkey(a) == bkey(b); };
    std::vector<std::vector<WGI>> q(NXCD);
    std::vector<size_t> load(NXCD, 0);
    size_t i = 0;
    while (i < tab.size()) {
        size_t e = i;
        while (e < tab.size() && same(tab[e], tab[i])) e++;
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        q[k].insert(q[k].end(), tab.begin() + i, tab.begin() + e);
        load[k] += e - i;
        i = e;
    }
    std::vector<WGI> out;
    out.reserve(tab.size());
    std::vector<size_t> pos(NXCD, 0);
    for (size_t n = 0; n < tab.size(); n++) {
        int k = (int)(n % NXCD);
        if (pos[k] == q[k].size()) {
            k = 0;
            for (int c = 1; c < NXCD; c++)
                if (q[c].size() - pos[c] > q[k].size() - pos[k]) k = c;
        }
        out.push_back(q[k][pos[k]++]);
    }
    return out;
}

// Partition the jobs by a variant key and emit one Sub per partition.
template <typename KeyF, typename GridF, typename SubF>
void build_stage(Batch &B, Stage &out, KeyF key, GridF grid, SubF fill,
                 std::function<int(const Sub &)> xcd_by = nullptr) {
    std::map<std::vector<int>, std::vector<int>> parts;
    for (size_t j = 0; j < B.hjob.size(); j++) {
        std::vector<int> k;
        if (!key((int)j, k)) continue;  // job takes no part in this stage
        parts[k].push_back((int)j);
    }
    for (auto &kv : parts) {
        std::unique_ptr<Sub> sub(new Sub());
        fill(*sub, kv.first, kv.second);
        std::vector<WGI> tab;
        for (int j : kv.second) {
            const JobGrid jg = grid(*sub, j);
            for (int z = 0; z < jg.gz; z++)
                for (int y = 0; y < jg.gy; y++)
                    for (int x = 0; x < jg.gx; x++)
                        tab.push_back(WGI{(unsigned short)j, (unsigned short)x, (unsigned short)y, (unsigned short)z});
        }
        if (tab.empty()) continue;
        tab = xcd_order(std::move(tab), xcd_by ? xcd_by(*sub) : XCD_BY_Z);
        sub->nwg = (int)tab.size();
        sub->wgi.upload(tab);
        out.push_back(std::move(sub));
    }
}
int ceil_div(int a, int b) { return (a + b - 1) / b; }
// Systems per workgroup: the most of cmax, cmax/2, .., 1 that keeps >= thresh
// workgroups in the launch; jw = (workgroups per system, systems) per job.
int pick_cs(const std::vector<std::pair<long, int>> &jw, int cmax, int thresh) {
    for (int cs = cmax; cs > 1; cs >>= 1) {
        long tot = 0;
        for (auto &p : jw) tot += p.first * ceil_div(p.second, cs);
        if (tot >= thresh) return cs;
    }
    return 1;
}

std::unique_ptr<Batch> make_batch(const std::vector<gpu_tomo *> &plans, const std::vector<int> &s0,
                                  const std::vector<int> &nsys) {
    std::unique_ptr<Batch> B(new Batch());
    const gpu_tomo *g0 = plans.at(0);
    B->rt64 = g0->rt64;
    B->spec64 = g0->spec64;
    B->maxit = g0->maxit;
    B->warm = g0->warm;
    B->precond = g0->G.precond;
    for (size_t j = 0; j < plans.size(); j++) {
        gpu_tomo *g = plans[j];
        if (g->device != g0->device || g->rt64 != B->rt64 || g->spec64 != B->spec64 || g->maxit != B->maxit ||
            g->warm != B->warm || g->G.precond != B->precond)
            fail(MT_EINVAL, "plans of one batch must share device, precision, maxit, warm start and preconditioner");
        if (s0[j] < 0 || nsys[j] <= 0 || s0[j] + nsys[j] > g->nbatch) fail(MT_EINVAL, "system range out of the plan");
        B->plan.push_back(g);
        B->hjob.push_back(Job{g->pdev.p, s0[j], nsys[j], 1, 0});
        for (int k = 0; k < g->G.nwfs; k++) B->anyttr |= g->G.ttr[k] != 0;
    }
    if (B->hjob.size() > 65535) fail(MT_EINVAL, "too many jobs in one batch");
    B->jobs.upload(B->hjob);
    Batch &b = *B;
    auto G = [&](int j) -> const Geom & { return b.plan[j]->G; };
    auto NS = [&](int j) { return b.hjob[j].nsys; };
    auto ptiles = [&](int j) { return ceil_div(G(j).pnx, TILE) * ceil_div(G(j).pny, TILE); };
    auto jw_of = [&](const std::vector<int> &js, std::function<long(int)> per) {
        std::vector<std::pair<long, int>> v;
        for (int j : js) v.push_back({per(j), NS(j)});
        return v;
    };
    // A batch over several plans (uneven groups, the request loop) is mostly
    // small systems, whose kernels are short enough that the number of
    // launches sets the time.  There the FFT and column stages share one
    // launch across plans, using the variant with the radix-5/7 butterflies
    // compiled in and the runtime block size when the plans differ (per
    // element the same arithmetic as the specialised variants).
    const bool merge = std::set<const gpu_tomo *>(b.plan.begin(), b.plan.end()).size() > 1;
    int rm_all = 3, bs_all = -1;
    for (size_t j = 0; j < b.hjob.size(); j++) {
        if (G((int)j).rmax > 4) rm_all = 7;
        const int bs = G((int)j).bs;
        const int bk = (bs == 2 || bs == 3 || bs == 4 || bs == 6 || bs == 8) ? bs : 0;
        bs_all = bs_all < 0 || bs_all == bk ? bk : 0;
    }
    auto rm_key = [&](int j) { return merge ? rm_all : G(j).rmax > 4 ? 7 : 3; };
    // row FFTs in registers for rows of 64 or 96 points (32 lanes x 2 or 3),
    // where every layer is fnx wide and a whole number of 16-row blocks tall
    // (each workgroup then has its 8 row pairs), unless the plan asks for
    // LDS stages.  A property of the plan, not of the batch, so that a system
    // computes the same bits alone and in any mix; a merged batch issues one
    // launch per register length beside its shared LDS-stage launch.
    auto rp_key = [&](int j) {
        const Geom &g = G(j);
        if (b.plan[j]->tune_colfft == 2 || g.rmax > 4 || (g.fnx != 96 && g.fnx != 64)) return 0;
        for (int l = 0; l < g.npsr; l++)
            if (g.xnx[l] != g.fnx || g.xny[l] % ROWS != 0) return 0;
        return g.fnx / 32;
    };
    auto row_key = [&](int j) { return rp_key(j) ? std::vector<int>{3, rp_key(j)} : std::vector<int>{rm_key(j), 0}; };
    // per-stage tables
    build_stage(b, b.zero, [&](int, std::vector<int> &) { return true; },
                [&](const Sub &, int) { return JobGrid{1, 1, 1}; }, [&](Sub &, const std::vector<int> &, const std::vector<int> &) {});
    build_stage(b, b.ttr,
                [&](int j, std::vector<int> &) {
                    for (int k = 0; k < G(j).nwfs; k++)
                        if (G(j).ttr[k]) return true;
                    return false;
                },
                [&](const Sub &, int j) { return JobGrid{G(j).nwfs, 1, NS(j)}; },
                [&](Sub &, const std::vector<int> &, const std::vector<int> &) {});
    build_stage(b, b.wfs_r, [&](int j, std::vector<int> &k) { k = {G(j).pos}; return true; },
                [&](const Sub &, int j) { return JobGrid{ptiles(j), G(j).nwfs, NS(j)}; },
                [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &) { sb.pos = k[0]; });
    // LHS with TT removal: three launches (gradients, TT projection, W GP')
    // on the generic tables; other jobs take the one-pass kernels below
    build_stage(b, b.wfs_g, [&](int j, std::vector<int> &k) { k = {G(j).pos}; return G(j).ptt_lhs != 0; },
                [&](const Sub &, int j) { return JobGrid{ptiles(j), G(j).nwfs, NS(j)}; },
                [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &) { sb.pos = k[0]; });
    build_stage(b, b.ttl, [&](int j, std::vector<int> &k) { k = {G(j).pos}; return G(j).ptt_lhs != 0; },
                [&](const Sub &, int j) { return JobGrid{G(j).nwfs, 1, NS(j)}; },
                [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &) { sb.pos = k[0]; });
    build_stage(b, b.wfs_l, [&](int j, std::vector<int> &k) { k = {G(j).matched, G(j).pos}; return !G(j).ptt_lhs; },
                [&](const Sub &sb, int j) {
                    return JobGrid{ptiles(j), G(j).nwfs, sb.matched ? ceil_div(NS(j), sb.cs) : NS(j)};
                },
                [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &js) {
                    sb.matched = k[0];
                    sb.pos = k[1];
                    const int tuned = b.plan[js[0]]->tune_wfs;
                    sb.cs = !sb.matched ? 1
                            : tuned     ? std::min(tuned, sb.pos == 1 ? 4 : 2)
                                        : pick_cs(jw_of(js, [&](int j) { return (long)ptiles(j) * G(j).nwfs; }),
                                                  sb.pos == 1 ? 4 : 2, 512);
                });
    build_stage(b, b.layer,
                [&](int j, std::vector<int> &k) {
                    k = {G(j).matched ? (G(j).lty == LTY ? 2 : G(j).lty == LTY2 ? 3 : 1) : 0};
                    return true;
                },
                [&](const Sub &sb, int j) {
                    return JobGrid{layer_tiles(G(j)), G(j).npsr, sb.matched ? ceil_div(NS(j), sb.cs) : NS(j)};
                },
                [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &js) {
                    sb.matched = k[0];
                    const int tuned = b.plan[js[0]]->tune_layer;
                    sb.cs = !sb.matched ? 1
                            : tuned     ? tuned
                                        : pick_cs(jw_of(js, [&](int j) { return (long)layer_tiles(G(j)) * G(j).npsr; }),
                                                  4, 512);
                });
    if (b.precond) {
        build_stage(b, b.row, [&](int j, std::vector<int> &k) { k = row_key(j); return true; },
                    [&](const Sub &, int j) { return JobGrid{G(j).nrowblk, 1, NS(j)}; },
                    [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &js) {
                        sb.rm = k[0];
                        sb.rp = k[1];
                        for (int j : js) sb.lds = std::max(sb.lds, row_lds(G(j).fnx, b.spec64));
                    });
        build_stage(b, b.irow, [&](int j, std::vector<int> &k) { k = row_key(j); return true; },
                    [&](const Sub &, int j) { return JobGrid{G(j).nrowblk, 1, NS(j)}; },
                    [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &js) {
                        sb.rm = k[0];
                        sb.rp = k[1];
                        for (int j : js) sb.lds = std::max(sb.lds, row_lds(G(j).fnx, 0));
                    });
        // register kernel where the plan allows it and the tuning does not
        // ask for the LDS kernel (a property of the plan, as for the rows).
        // One plan: a launch per (P, lanes, block size), specialised where an
        // instantiation exists.  Several plans: one launch per P, the
        // runtime lanes / block size variant (the same operations in the
        // same order, so the same bits) with the largest workgroup.
        auto colreg = [&](int j) { return b.plan[j]->colr_ok && b.plan[j]->tune_colfft != 2 ? G(j).shP : 0; };
        std::map<int, int> nthr_of_p;  // merged: largest workgroup per P
        for (size_t j = 0; j < b.hjob.size(); j++)
            if (colreg((int)j)) {
                int &v = nthr_of_p[colreg((int)j)];
                v = std::max(v, colr_shape(G((int)j).bs, G((int)j).shL, nullptr));
            }
        auto colr_tasks_of = [&](int j) {
            int T = 1;
            colr_shape(G(j).bs, G(j).shL, &T);
            return T;
        };
        build_stage(b, b.col,
                    [&](int j, std::vector<int> &k) {
                        if (colreg(j)) {
                            if (merge) k = {-1, colreg(j), nthr_of_p[colreg(j)], 0, 0};
                            else k = {-1, colreg(j), colr_shape(G(j).bs, G(j).shL, nullptr), G(j).shL, G(j).bs};
                            return true;
                        }
                        const int bs = G(j).bs;
                        const int bk = (bs == 2 || bs == 3 || bs == 4 || bs == 6 || bs == 8) ? bs : 0;
                        k = {merge ? bs_all : bk, rm_key(j)};
                        return true;
                    },
                    [&](const Sub &sb, int j) {
                        return sb.reg ? JobGrid{ceil_div(NS(j), colr_tasks_of(j)), G(j).ngroup, 1}
                                      : JobGrid{G(j).ngroup, 1, ceil_div(NS(j), sb.cs)};
                    },
                    [&](Sub &sb, const std::vector<int> &k, const std::vector<int> &js) {
                        if (k[0] == -1) {
                            sb.reg = k[1];
                            sb.nthr = k[2];
                            sb.regL = k[3];
                            sb.bs = k[4];
                            for (int j : js)
                                sb.lds = std::max(sb.lds, colr_lds(G(j).bs, G(j).shL, sb.reg, b.spec64));
                            return;
                        }
                        sb.bs = k[0];
                        sb.rm = k[1];
                        // up to 4 systems per workgroup as far as the device's LDS
                        // per workgroup allows (the heuristic stays at 2 with an
                        // FP64 spectrum; the autotuner may choose 4)
                        const size_t lmax = b.plan[js[0]]->lds_max;
                        const int tuned = b.plan[js[0]]->tune_col;
                        sb.cs = tuned ? std::min(tuned, 4)
                                      : pick_cs(jw_of(js, [&](int j) { return (long)G(j).ngroup; }),
                                                b.spec64 ? 2 : 4, 384);
                        auto fits = [&](int c) {
                            for (int j : js)
                                if (col_lds(G(j).maxcol, G(j).npsr, G(j).fny, b.spec64, c) > lmax) return false;
                            return true;
                        };
                        while (sb.cs > 1 && !fits(sb.cs)) sb.cs >>= 1;
                        for (int j : js)
                            sb.lds = std::max(sb.lds, col_lds(G(j).maxcol, G(j).npsr, G(j).fny, b.spec64, sb.cs));
                        if (sb.lds > 64 * 1024) col_allow_lds(b.spec64, sb.cs, sb.bs, sb.rm, sb.lds);
                    },
                    [](const Sub &sb) { return sb.reg ? XCD_BY_Y : XCD_BY_X; });
    }
    if (!b.precond)
        build_stage(b, b.cg, [&](int, std::vector<int> &) { return true; },
                    [&](const Sub &, int j) { return JobGrid{G(j).nrowblk, 1, NS(j)}; },
                    [&](Sub &, const std::vector<int> &, const std::vector<int> &) {});
    build_stage(b, b.fin, [&](int, std::vector<int> &) { return true; },
                [&](const Sub &, int j) { return JobGrid{1, 1, NS(j)}; },
                [&](Sub &, const std::vector<int> &, const std::vector<int> &) {});
    return B;
}

// The plan's own batch for systems [s0, s0 + n).
Batch &own_batch(gpu_tomo *g, int s0, int n) {
    auto key = std::make_pair(s0, n);
    auto it = g->own.find(key);
    if (it != g->own.end()) return *it->second;
    std::unique_ptr<Batch> B = make_batch({g}, {s0}, {n});
    Batch &ref = *B;
    g->own[key] = std::move(B);
    return ref;
}
// Destroy the plan's solve graph / flight graphs once their last launches
// (which an asynchronous caller may not have waited for) have completed.
void drop_graph(gpu_tomo *g) {
    if (g->gexec) {
        if (g->gdone) (void)hipEventSynchronize(g->gdone);
        (void)hipGraphExecDestroy(g->gexec);
    }
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->gexec = nullptr;
    g->graph = nullptr;
}
void drop_flights(gpu_tomo *g) {
    for (auto &f : g->fl) {
        if (f.exec) {
            if (f.done) (void)hipEventSynchronize(f.done);
            (void)hipGraphExecDestroy(f.exec);
        }
        if (f.graph) (void)hipGraphDestroy(f.graph);
        f.exec = nullptr;
        f.graph = nullptr;
    }
    g->fl_nsys = g->fl_k = 0;
}
void drop_batches(gpu_tomo *g) {
    g->gen++;
    drop_graph(g);
    drop_flights(g);
    g->own.clear();
}

LaunchTab tab(const Batch &B, const Sub &s) { return LaunchTab{B.jobs_over ? B.jobs_over : B.jobs.p, s.wgi.p, s.nwg}; }

// Operator A v -> Ap for all jobs (WFS pass + layer pass).
void enqueue_L(Batch &B, int vsel, int partsel, Rec &rec) {
    rec(KC_WFS, [&] {
        for (auto &s : B.wfs_l) launch_wfs(tab(B, *s), false, s->matched, s->pos, s->cs, B.rt64, vsel, rec.st);
        // the two stages hold the same partitions (keyed by tomo.pos) in order
        for (size_t i = 0; i < B.wfs_g.size(); i++)
            launch_wfs_lhs_tt(tab(B, *B.wfs_g[i]), tab(B, *B.ttl[i]), B.wfs_g[i]->pos, B.rt64, vsel, rec.st);
    });
    rec(KC_LAYER, [&] {
        for (auto &s : B.layer)
            launch_layer(tab(B, *s), true, s->matched, s->cs, B.rt64, vsel, BUF_AP, partsel, rec.st);
    });
}
// RHS b = HX' GP' W (I - TT PTT) g for all jobs.
void enqueue_R(Batch &B, Rec &rec) {
    if (!B.ttr.empty()) rec(KC_TTR, [&] {
            for (auto &s : B.ttr) launch_ttr(tab(B, *s), rec.st);
        });
    rec(KC_WFS_RHS, [&] {
        for (auto &s : B.wfs_r) launch_wfs(tab(B, *s), true, 0, s->pos, 1, B.rt64, BUF_X, rec.st);
    });
    rec(KC_LAYER_RHS, [&] {
        for (auto &s : B.layer) launch_layer(tab(B, *s), false, s->matched, s->cs, B.rt64, BUF_X, BUF_B, PART_BB, rec.st);
    });
}
void enqueue_row(Batch &B, int kc, int mode, int iter, int xsel, Rec &rec) {
    rec(kc, [&] {
        for (auto &s : B.row)
            launch_row(tab(B, *s), mode, B.rt64, B.spec64, s->rm, s->rp, iter, xsel, s->lds, rec.st);
    });
}
// Preconditioner on the row-transformed residual (specf): column pass with
// the spectral r.z partials, then the inverse row pass in the given mode.
void enqueue_P_tail(Batch &B, int irow_mode, int slot, int osel, Rec &rec) {
    rec(KC_COL, [&] {
        for (auto &s : B.col) {
            if (s->reg) launch_colr(tab(B, *s), B.spec64, s->reg, s->regL, s->bs, s->nthr, s->lds, rec.st);
            else launch_col(tab(B, *s), B.spec64, s->cs, s->bs, s->rm, s->lds, rec.st);
        }
    });
    rec(irow_mode == IROW_MODE_UPDATE ? KC_IROW : KC_IROW_START, [&] {
        for (auto &s : B.irow) launch_irow(tab(B, *s), irow_mode, s->rm, s->rp, slot, osel, s->lds, rec.st);
    });
}

// The solve as lib/pcg.c:28-113 runs it (restart every 100 iterations):
//   start:  r = b - A x (r = b cold, x = 0); z = M r; p = z; rz[0] = r.z
//   iter k: Ap; alpha = rz[k] / p.Ap; r -= alpha Ap;
//           last: x += alpha p, |r| -> residual; else z = M r,
//           beta = r.z / rz[k], x += alpha p, p = z + beta p
// Kernels per iteration: wfs, layer (Ap, p.Ap partials), row (alpha, r,
// forward row FFT), col (column FFTs, blocks, r.z partials), irow (inverse
// row FFT, beta, x and p).  All jobs of the batch advance in lockstep.
// Without a preconditioner (tomo.precond=0) z = r, as lib/pcg.c does with
// Mmul = NULL: per iteration wfs, layer, cgv (alpha, x, r, r.r partials) and
// cgp (beta, p); the start is cgv (r = b, or b - A x warm) and cgp (p = r).
void enqueue_solve_cg(Batch &B, Rec &rec) {
    const int maxit = B.maxit;
    auto cgv = [&](int kc, int mode, int iter) {
        rec(kc, [&] {
            for (auto &s : B.cg) launch_cgv(tab(B, *s), mode, B.rt64, iter, rec.st);
        });
    };
    auto cgp = [&](int kc, int mode, int slot) {
        rec(kc, [&] {
            for (auto &s : B.cg) launch_cgp(tab(B, *s), mode, B.rt64, slot, rec.st);
        });
    };
    auto start = [&](int k, bool cold) {
        if (cold) {
            cgv(KC_ROW_START, CGV_MODE_INIT_COLD, k);
        } else {
            enqueue_L(B, BUF_X, PART_NONE, rec);
            cgv(KC_ROW_START, CGV_MODE_INIT_WARM, k);
        }
        cgp(KC_IROW_START, CGP_MODE_FIRST, k);
    };
    start(0, !B.warm);
    for (int k = 0; k < maxit; k++) {
        if (k > 0 && k % 100 == 0) start(k, false);
        enqueue_L(B, BUF_P, PART_PAP, rec);
        const bool last = k + 1 == maxit;
        cgv(last ? KC_ROW_LAST : KC_ROW, last ? CGV_MODE_LAST : CGV_MODE_UPDATE, k);
        if (last) break;
        cgp(KC_IROW, CGP_MODE_UPDATE, k);
    }
}

void enqueue_solve(Batch &B, Rec &rec) {
    const int maxit = B.maxit;
    for (auto &s : B.zero) launch_zero_stop(tab(B, *s), rec.st);
    enqueue_R(B, rec);
    if (!B.precond) {
        enqueue_solve_cg(B, rec);
        rec(KC_FINAL, [&] {
            for (auto &s : B.fin) launch_final(tab(B, *s), rec.st);
        });
        return;
    }
    auto start = [&](int k, bool cold) {
        if (cold) {
            enqueue_row(B, KC_ROW_START, ROW_MODE_INIT_COLD, k, BUF_X, rec);
        } else {
            enqueue_L(B, BUF_X, PART_NONE, rec);
            enqueue_row(B, KC_ROW_START, ROW_MODE_INIT_WARM, k, BUF_X, rec);
        }
        enqueue_P_tail(B, IROW_MODE_FIRST, k, BUF_P, rec);
    };
    start(0, !B.warm);
    for (int k = 0; k < maxit; k++) {
        if (k > 0 && k % 100 == 0) start(k, false);
        enqueue_L(B, BUF_P, PART_PAP, rec);
        const bool last = k + 1 == maxit;
        enqueue_row(B, last ? KC_ROW_LAST : KC_ROW, last ? ROW_MODE_LAST : ROW_MODE_UPDATE, k, BUF_X, rec);
        if (last) break;
        enqueue_P_tail(B, IROW_MODE_UPDATE, k, BUF_P, rec);
    }
    rec(KC_FINAL, [&] {
        for (auto &s : B.fin) launch_final(tab(B, *s), rec.st);
    });
}

// Fold the recorded events into the per-class averages; accumulate: add to
// the current sums (a call made of several solves) instead of starting anew.
void prof_collect(gpu_tomo *g, bool accumulate = false) {
    HIPCHK(hipStreamSynchronize(g->stream));
    if (!accumulate)
        for (int i = 0; i < KC_N; i++) {
            g->prof_ms[i] = 0;
            g->prof_cnt[i] = 0;
        }
    for (auto &e : g->evs) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
        g->prof_ms[e.kc] += ms;
        g->prof_cnt[e.kc] += e.n;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
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

// Caller layout (all WFS, skipped ones empty, double) -> active-WFS layout
// (float); rejects non-finite input.
void pack_grad(const gpu_tomo *g, const double *grad, float *dst) {
    const Geom &G = g->G;
    for (int k = 0; k < G.nwfs; k++) {
        const int w = g->wfs_active[k];
        const long o = g->goff_all[w];
        for (long i = 0; i < 2 * G.nsa[k]; i++) {
            const double v = grad[o + i];
            if (!std::isfinite(v)) fail(MT_ENOTFINITE, "non-finite gradient input");
            dst[G.goff[k] + i] = (float)v;
        }
    }
}
void to_dev_grad(gpu_tomo *g, int isys, const double *grad) {
    const Geom &G = g->G;
    g->hstage.assign(G.ngtot, 0.f);
    pack_grad(g, grad, g->hstage.data());
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

// gpu_tomo_solve_dev_noalloc's set-up, done where allocating is allowed:
// this device's address of the workspace array (looked up once per device,
// with the device current) and the plan's one-system launch tables, which
// that entry point reuses.  Called by gpu_tomo_setup and again by every call
// that rebuilds the plan's launch tables.
void noalloc_init(gpu_tomo *g) {
    if (g->device < 0 || g->device >= NOALLOC_MAXDEV) fail(MT_EINVAL, "device index too large");
    {
        std::lock_guard<std::mutex> lk(scratch_mu);
        if (!scratch_ptr[g->device]) {
            void *ptr = nullptr;
            HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            if (!ptr) fail(MT_EHIP, "no device address for the workspace array");
            scratch_ptr[g->device] = ptr;
        }
    }
    (void)own_batch(g, 0, 1);
}

// Workspace layout of one system (byte offsets from the start of the
// array, 256-byte aligned): the device record and job entry, then the
// state buffers a solve writes besides x.
struct NoallocLayout {
    size_t pd, job, r, Ap, b, z, p, q, tt, gl, ttl, specf, speci, part[4], rz, alpha, res, stop, total;
};
NoallocLayout noalloc_layout(const gpu_tomo *g) {
    const Geom &G = g->G;
    const size_t rt = g->rt64 ? 8 : 4, sp = g->spec64 ? 16 : 8;
    Geom g16 = G;
    g16.lty = TILE;
    const size_t maxpart = (size_t)std::max({layer_tiles(g16) * G.npsr, G.nrowblk, G.ngroup, 1});
    NoallocLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += (bytes + 255) & ~(size_t)255;
        return at;
    };
    L.pd = take(sizeof(PlanDev));
    L.job = take(sizeof(Job));
    L.r = take(rt * G.nxtot);
    L.Ap = take(rt * G.nxtot);
    L.b = take(rt * G.nxtot);
    L.z = take(4 * G.nxtot);
    L.p = take(4 * G.nxtot);
    L.q = take(rt * G.nwfs * G.nploc);
    L.tt = take(4 * 2 * G.nwfs);
    L.gl = G.ptt_lhs ? take(rt * G.ngtot) : 0;
    L.ttl = G.ptt_lhs ? take(8 * 2 * G.nwfs) : 0;
    L.specf = take(sp * std::max<long>(G.nspec, 1));
    L.speci = take(8 * std::max<long>(G.nspec, 1));
    for (int i = 0; i < 4; i++) L.part[i] = take(8 * maxpart);
    L.rz = take(8 * (g->maxit + 2));
    L.alpha = take(8);
    L.res = take(8);
    L.stop = take(4);
    L.total = o;
    return L;
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
        {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess &&
                v > 64 * 1024)
                g->lds_max = (size_t)v;
        }
        g->nbatch = nbatch;
        g->maxit = maxit;
        g->warm = warm ? 1 : 0;
        build_plan(g.get(), recon);
        chk64(t64_keep(recon, &g->rc64));
        alloc_state(g.get());
        HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        noalloc_init(g.get());
        HIPCHK(hipDeviceSynchronize());
    });
    if (rc == MT_OK) *out = g.release();
    return rc;
}

// New regulariser and FDPCG blocks of a plan, written in place (the device
// record, Minv and the register kernel's copy Mreg keep their addresses, so
// launch tables and captured graphs that point at them stay valid).  Waits
// for the plan's own streams only: the caller must not have a solve of this
// plan in flight on another stream.
static void update_cn2_plan(gpu_tomo *g, const mt_recon_t *recon) {
    if (recon->npsr != g->npsr) fail(MT_EINVAL, "layer count changed");
    if ((recon->precond ? 1 : 0) != g->G.precond) fail(MT_EINVAL, "preconditioner changed");
    std::vector<float2> hM, Mr;
    if (g->G.precond) {
        const long nbh = recon->fd_nbhalf, bs = recon->fd_bs;
        if (bs != g->G.bs || (long)g->Minv.n != nbh * bs * bs) fail(MT_EINVAL, "FDPCG layout changed");
        hM = col_blocks(recon, g);
        if (g->colr_ok) {
            if (!mreg_layout(g, hM, g->grp_blk_host, Mr) || Mr.size() != g->Mreg.n)
                fail(MT_EINVAL, "FDPCG block layout changed");
        }
    }
    HIPCHK(hipStreamSynchronize(g->stream));
    for (auto &f : g->fl)
        if (f.st) HIPCHK(hipStreamSynchronize(f.st));
    for (int l = 0; l < recon->npsr; l++) g->G.l2c[l] = (float)recon->l2c[l];
    if (!hM.empty())
        HIPCHK(hipMemcpyAsync(g->Minv.p, hM.data(), sizeof(float2) * hM.size(), hipMemcpyHostToDevice, g->stream));
    if (!Mr.empty())
        HIPCHK(hipMemcpyAsync(g->Mreg.p, Mr.data(), sizeof(float2) * Mr.size(), hipMemcpyHostToDevice, g->stream));
    upload_plandev(g, g->stream);
    HIPCHK(hipStreamSynchronize(g->stream));
}

int gpu_tomo_update_cn2(gpu_tomo_t *g, const mt_recon_t *recon) {
    if (!g || !recon) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        update_cn2_plan(g, recon);
        chk64(t64_keep_cn2(g->rc64, recon));
        if (g->t64) chk64(t64_update_cn2(g->t64, recon));
    });
}

int gpu_tomo_reset(gpu_tomo_t *g) {
    if (!g) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipMemsetAsync(g->x.p, 0, sizeof(float) * g->x.n, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
        if (g->t64) chk64(t64_reset(g->t64));
    });
}

void gpu_tomo_free(gpu_tomo_t *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    (void)hipDeviceSynchronize();
    drop_batches(g);
    for (auto &e : g->evs) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto &f : g->fl) {
        if (f.st) (void)hipStreamDestroy(f.st);
        if (f.done) (void)hipEventDestroy(f.done);
    }
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->gdone) (void)hipEventDestroy(g->gdone);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    t64_free(g->t64);
    t64_keep_free(g->rc64);
    delete g;
}

int gpu_tomo_info(const gpu_tomo_t *g, int *matched, int *fnx, int *fny, int *bs, int *rmax) {
    if (!g) return MT_EINVAL;
    if (matched) *matched = g->G.matched;
    if (fnx) *fnx = g->G.fnx;
    if (fny) *fny = g->G.fny;
    if (bs) *bs = g->G.bs;
    if (rmax) *rmax = g->G.rmax;
    return MT_OK;
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
        if (g->t64) {
            chk64(t64_solve(g->t64, isys, 1, grad, x, &res, g->maxit, g->warm, g->stream));
            return;
        }
        to_dev_grad(g, isys, grad);
        Rec rec{&g->evs, g->stream, false};
        enqueue_solve(own_batch(g, isys, 1), rec);
        from_dev_x(g, x, g->x.p + (long)isys * g->G.nxtot);
        HIPCHK(hipMemcpy(&res, g->res.p + isys, sizeof(double), hipMemcpyDeviceToHost));
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
        if (g->t64) {
            chk64(t64_solve(g->t64, 0, nsys, grad, x, res.data(), g->maxit, g->warm, g->stream));
            return;
        }
        for (int s = 0; s < nsys; s++) to_dev_grad(g, s, grad + (long)s * g->ngtot_all);
        Rec rec{&g->evs, g->stream, false};
        enqueue_solve(own_batch(g, 0, nsys), rec);
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

// Capture the solve of batch B on a private stream into (graph, exec).
// Capture pre(stream), the solve and post(stream) into one graph.
extern "C++" {
template <typename Pre, typename Post>
void capture_solve_with(Batch &B, hipGraph_t *graph, hipGraphExec_t *exec, EvList *evs, Pre pre, Post post) {
    hipStream_t cs;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    Rec rec{evs, cs, false};
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
        pre(cs);
        enqueue_solve(B, rec);
        post(cs);
    } catch (...) {
        hipGraph_t junk;
        (void)hipStreamEndCapture(cs, &junk);
        (void)hipStreamDestroy(cs);
        throw;
    }
    HIPCHK(hipStreamEndCapture(cs, graph));
    (void)hipStreamDestroy(cs);
    HIPCHK(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
}
}  // extern "C++"
void capture_solve(Batch &B, hipGraph_t *graph, hipGraphExec_t *exec, EvList *evs) {
    capture_solve_with(B, graph, exec, evs, [](hipStream_t) {}, [](hipStream_t) {});
}

// Flights of a graph solve of n systems: the tuned count (default 1),
// never more than n.  Whether flights overlap depends on the hardware queues
// their streams land on: streams share a small pool of queues (least used
// first), so two flights can share one and run one after the other,
// depending on the other streams the process holds.  Hence no default above
// 1: the autotuner times 1, 2 and 4 and keeps more only where they gain.
int flights_for(const gpu_tomo *g, int n) { return std::min(std::max(g->tune_flights, 1), n); }
// The graph solve of systems [0, n) on stream st: one graph, or k flights
// over consecutive ranges (the first n % k ranges one system longer).
void graph_solve(gpu_tomo *g, int n, hipStream_t st) {
    const int k = flights_for(g, n);
    if (k == 1) {
        if (!g->gexec || g->graph_nsys != n) {
            drop_graph(g);
            capture_solve(own_batch(g, 0, n), &g->graph, &g->gexec, &g->evs);
            g->graph_nsys = n;
        }
        if (!g->gdone) HIPCHK(hipEventCreateWithFlags(&g->gdone, hipEventDisableTiming));
        HIPCHK(hipGraphLaunch(g->gexec, st));
        HIPCHK(hipEventRecord(g->gdone, st));
        return;
    }
    if (!g->fork) HIPCHK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    if ((int)g->fl.size() < k) g->fl.resize(k);
    if (g->fl_nsys != n || g->fl_k != k) {
        drop_flights(g);
        // the flights' streams first, back to back (no other stream created
        // or destroyed between them, as the captures below do), so that the
        // runtime's least-used choice gives them different hardware queues
        for (int i = 0; i < k; i++) {
            auto &f = g->fl[i];
            if (!f.st) HIPCHK(hipStreamCreateWithFlags(&f.st, hipStreamNonBlocking));
            if (!f.done) HIPCHK(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
        }
        for (int i = 0, s0 = 0; i < k; i++) {
            const int ni = n / k + (i < n % k);
            capture_solve(own_batch(g, s0, ni), &g->fl[i].graph, &g->fl[i].exec, &g->evs);
            s0 += ni;
        }
        g->fl_nsys = n;
        g->fl_k = k;
    }
    HIPCHK(hipEventRecord(g->fork, st));
    for (int i = 0; i < k; i++) {
        auto &f = g->fl[i];
        HIPCHK(hipStreamWaitEvent(f.st, g->fork, 0));
        HIPCHK(hipGraphLaunch(f.exec, f.st));
        HIPCHK(hipEventRecord(f.done, f.st));
    }
    for (int i = 0; i < k; i++) HIPCHK(hipStreamWaitEvent(st, g->fl[i].done, 0));
}

int gpu_tomo_solve_dev(gpu_tomo_t *g, const float *d_grad, float *d_x, int nbatch, int mode, void *stream) {
    if (!g || nbatch <= 0 || nbatch > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        no_mode4(g, "gpu_tomo_solve_dev");
        hipStream_t st = stream ? (hipStream_t)stream : g->stream;
        const Geom &G = g->G;
        if (d_grad && d_grad != g->g.p) launch_copy(d_grad, g->g.p, (long)nbatch * G.ngtot, st);
        if (d_x && d_x != g->x.p && g->warm) launch_copy(d_x, g->x.p, (long)nbatch * G.nxtot, st);
        if (mode == 1 && !g->prof) {
            graph_solve(g, nbatch, st);
        } else {
            Rec rec{&g->evs, st, g->prof != 0, g->prof};
            enqueue_solve(own_batch(g, 0, nbatch), rec);
            if (g->prof) {
                if (st != g->stream) HIPCHK(hipStreamSynchronize(st));
                prof_collect(g);
            }
        }
        if (d_x && d_x != g->x.p) launch_copy(g->x.p, d_x, (long)nbatch * G.nxtot, st);
    });
}

}  // extern "C"

// gpu_tomo_solve_dev_noalloc's device side: the workspace is cleared by a
// fill kernel (no memset node, so the call can be captured), and the device
// record and job entry of the system are written from kernel arguments.
__global__ __launch_bounds__(256) void k_noalloc_clear(uint4 *p, long n16) {
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long)gridDim.x * blockDim.x) p[i] = z;
}
__global__ void k_noalloc_head(PlanDev h, Job j, PlanDev *dpd, Job *djob) {
    if (threadIdx.x == 0) {
        *dpd = h;
        j.pd = dpd;
        *djob = j;
    }
}
static_assert(sizeof(PlanDev) + sizeof(Job) + 2 * sizeof(void *) <= 4096, "kernel argument size");

extern "C" {

int gpu_tomo_solve_dev_noalloc(gpu_tomo_t *g, const float *d_grad, float *d_x, double *d_res, int nsys,
                               void *stream) {
    if (!g || !d_grad || !d_x || nsys <= 0 || nsys > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (g->prof) fail(MT_ESTATE, "gpu_tomo_solve_dev_noalloc: profiling is on (it allocates events)");
        no_mode4(g, "gpu_tomo_solve_dev_noalloc");
        unsigned char *base = nullptr;
        {
            std::lock_guard<std::mutex> lk(scratch_mu);
            if (g->device < NOALLOC_MAXDEV) base = (unsigned char *)scratch_ptr[g->device];
        }
        if (!base) fail(MT_ESTATE, "gpu_tomo_solve_dev_noalloc: workspace not set up on this device");
        auto it = g->own.find(std::make_pair(0, 1));
        if (it == g->own.end()) fail(MT_ESTATE, "gpu_tomo_solve_dev_noalloc: launch tables not built");
        Batch &B = *it->second;
        const NoallocLayout L = noalloc_layout(g);
        if (L.total > NOALLOC_BYTES)
            fail(MT_EINVAL, "gpu_tomo_solve_dev_noalloc: one system needs " + std::to_string(L.total) +
                                " bytes of workspace, more than the " + std::to_string(NOALLOC_BYTES) +
                                " bytes available");
        hipStream_t st = stream ? (hipStream_t)stream : g->stream;
        const Geom &G = g->G;
        Geom g16 = G;
        g16.lty = TILE;
        const int maxpart = std::max({layer_tiles(g16) * G.npsr, G.nrowblk, G.ngroup, 1});
        for (int s = 0; s < nsys; s++) {
            // a system at a time through the one-system tables; the
            // workspace is reused, in stream order
            PlanDev h{};
            h.G = G;
            h.x = d_x + (long)s * G.nxtot;
            h.g = const_cast<float *>(d_grad) + (long)s * G.ngtot;
            h.z = (float *)(base + L.z);
            h.p = (float *)(base + L.p);
            h.tt = (float *)(base + L.tt);
            h.r = base + L.r;
            h.Ap = base + L.Ap;
            h.b = base + L.b;
            h.q = base + L.q;
            h.specf = base + L.specf;
            h.speci = (float2 *)(base + L.speci);
            h.SC = g->SC;
            h.SC.part_pap = (double *)(base + L.part[0]);
            h.SC.part_rz = (double *)(base + L.part[1]);
            h.SC.part_rr = (double *)(base + L.part[2]);
            h.SC.part_bb = (double *)(base + L.part[3]);
            h.SC.rz = (double *)(base + L.rz);
            h.SC.alpha = (double *)(base + L.alpha);
            h.SC.res = d_res ? d_res + s : (double *)(base + L.res);
            h.SC.tt = h.tt;
            h.SC.maxpart = maxpart;
            h.stop = (int *)(base + L.stop);
            h.ltiles = layer_tiles(G);
            h.gl = G.ptt_lhs ? base + L.gl : nullptr;
            h.ttl = G.ptt_lhs ? (double *)(base + L.ttl) : nullptr;
            const long n16 = (long)(L.total / 16);
            hipLaunchKernelGGL(k_noalloc_clear, dim3((unsigned)std::min<long>((n16 + 255) / 256, 2048)), dim3(256), 0,
                               st, (uint4 *)base, n16);
            hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(64), 0, st, h, Job{nullptr, 0, 1, 1, 0},
                               (PlanDev *)(base + L.pd), (Job *)(base + L.job));
            HIPCHK(hipGetLastError());
            B.jobs_over = (const Job *)(base + L.job);
            try {
                Rec rec{&g->evs, st, false};
                enqueue_solve(B, rec);
            } catch (...) {
                B.jobs_over = nullptr;
                throw;
            }
            B.jobs_over = nullptr;
        }
        HIPCHK(hipGetLastError());
    });
}

}  // extern "C"

// Columns [col0, col0 + ncol) of MVM assembly, each the cold solve of a unit
// gradient vector e_j: with f null the tomography solution x = T e_j (nx
// per column), else the DM commands a = F T e_j of the fit of that solution
// (nact per column), written to out column-major as doubles.  Chunks of
// nbatch columns: the unit vectors are written on the device from an index
// table, the tomography runs as a graph solve, the fit (if any) reads its x
// in place; two pinned staging buffers let the host copy chunk t-1 out
// while the GPU solves chunk t.  The slots' gradients and solutions (the
// warm-start state) are restored at the end.
static void mvm_columns(gpu_tomo *g, gpu_fit_t *f, long col0, long ncol, double *out) {
    HIPCHK(hipSetDevice(g->device));
    no_mode4(g, "MVM assembly");
    const Geom &G = g->G;
    const long nx = g->nxtot;
    const long nrow = f ? fit_nact(f) : nx;  // output rows per column
    hipStream_t st = g->stream;
    // device gradient index of each requested column; the columns of a
    // WFS the solve does not use are zero
    std::vector<int> kof(g->goff_all.size() - 1, -1);
    for (int k = 0; k < G.nwfs; k++) kof[g->wfs_active[k]] = k;
    std::vector<int> idx;
    std::vector<long> col;
    for (long j = col0; j < col0 + ncol; j++) {
        const int w = (int)(std::upper_bound(g->goff_all.begin(), g->goff_all.end(), j) - g->goff_all.begin()) - 1;
        const int k = kof[w];
        if (k < 0) {
            std::fill(out + (j - col0) * nrow, out + (j - col0 + 1) * nrow, 0.0);
        } else {
            idx.push_back((int)(G.goff[k] + (j - g->goff_all[w])));
            col.push_back(j - col0);
        }
    }
    if (idx.empty()) return;
    DevBuf<int> didx;
    didx.upload(idx);
    // the chunks overwrite the slots' gradients and solutions: keep them
    // (the warm-start state of a closed loop) and put them back at the end
    DevBuf<float> keep_x, keep_g;
    keep_x.alloc(g->x.n);
    keep_g.alloc(g->g.n);
    HIPCHK(hipMemcpyAsync(keep_x.p, g->x.p, sizeof(float) * g->x.n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(keep_g.p, g->g.p, sizeof(float) * g->g.n, hipMemcpyDeviceToDevice, st));
    const int B = g->nbatch;
    DevBuf<double> dm;  // fit output of one chunk
    if (f) dm.alloc((size_t)B * nrow);
    const long nch = ((long)idx.size() + B - 1) / B;
    void *h[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    const size_t esz = f ? sizeof(double) : sizeof(float);
    auto widen = [&](long t) {
        HIPCHK(hipEventSynchronize(ev[t % 2]));
        const long c = t * B;
        const int n = (int)std::min<long>(B, (long)idx.size() - c);
        // columns over up to 8 host threads (the output's first touch and
        // the widening are host-bound otherwise)
        auto cols = [&](int k0, int k1) {
            for (int k = k0; k < k1; k++) {
                double *o = out + col[c + k] * nrow;
                if (f) {
                    std::memcpy(o, (const double *)h[t % 2] + (long)k * nrow, sizeof(double) * nrow);
                } else {
                    const float *hk = (const float *)h[t % 2] + (long)k * nrow;
                    for (long i = 0; i < nrow; i++) o[i] = hk[i];
                }
            }
        };
        const int nt = std::min(n, 8);
        std::vector<std::thread> th;
        for (int q = 1; q < nt; q++) th.emplace_back(cols, q * n / nt, (q + 1) * n / nt);
        cols(0, n / nt);
        for (auto &x : th) x.join();
    };
    try {
        for (int b = 0; b < 2; b++) {
            HIPCHK(hipHostMalloc(&h[b], esz * nrow * B));
            HIPCHK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
        for (long t = 0; t < nch; t++) {
            const long c = t * B;
            const int n = (int)std::min<long>(B, (long)idx.size() - c);
            launch_unit(g->g.p, G.ngtot, didx.p + c, n, st);
            if (g->warm) launch_unit(g->x.p, G.nxtot, nullptr, n, st);  // cold start
            if (!g->prof) {
                graph_solve(g, n, st);
            } else {  // profiled plans launch kernel by kernel, as gpu_tomo_solve_dev does
                Rec rec{&g->evs, st, true, g->prof};
                enqueue_solve(own_batch(g, 0, n), rec);
                prof_collect(g, t > 0);  // the averages cover every chunk
            }
            if (f) {
                const int rc = fit_enqueue_cold(f, g->x.p, G.nxtot, dm.p, n, st);
                if (rc != MT_OK) fail(rc, mt_last_error());
                HIPCHK(hipMemcpyAsync(h[t % 2], dm.p, sizeof(double) * nrow * n, hipMemcpyDeviceToHost, st));
            } else {
                HIPCHK(hipMemcpy2DAsync(h[t % 2], sizeof(float) * nx, g->x.p, sizeof(float) * G.nxtot,
                                        sizeof(float) * nx, n, hipMemcpyDeviceToHost, st));
            }
            HIPCHK(hipEventRecord(ev[t % 2], st));
            if (t > 0) widen(t - 1);
        }
        widen(nch - 1);
    } catch (...) {
        (void)hipStreamSynchronize(st);
        (void)hipMemcpy(g->x.p, keep_x.p, sizeof(float) * g->x.n, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(g->g.p, keep_g.p, sizeof(float) * g->g.n, hipMemcpyDeviceToDevice);
        for (int b = 0; b < 2; b++) {
            if (h[b]) (void)hipHostFree(h[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
        }
        throw;
    }
    HIPCHK(hipMemcpyAsync(g->x.p, keep_x.p, sizeof(float) * g->x.n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(g->g.p, keep_g.p, sizeof(float) * g->g.n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int b = 0; b < 2; b++) {
        (void)hipHostFree(h[b]);
        (void)hipEventDestroy(ev[b]);
    }
}

extern "C" {

int gpu_tomo_mvm_tomo(gpu_tomo_t *g, long col0, long ncol, double *out) {
    if (!g || !out || col0 < 0 || ncol <= 0 || col0 + ncol > g->ngtot_all) return MT_EINVAL;
    return guarded([&] { mvm_columns(g, nullptr, col0, ncol, out); });
}

int gpu_mvm_assemble(gpu_tomo_t *g, gpu_fit_t *f, long col0, long ncol, double *out) {
    if (!g || !f || !out || col0 < 0 || ncol <= 0 || col0 + ncol > g->ngtot_all) return MT_EINVAL;
    return guarded([&] {
        if (fit_nxtot(f) != g->G.nxtot) fail(MT_EINVAL, "gpu_mvm_assemble: the fit plan's layers are not the tomography's");
        if (fit_device(f) != g->device) fail(MT_EINVAL, "gpu_mvm_assemble: plans on different devices");
        mvm_columns(g, f, col0, ncol, out);
    });
}

int gpu_tomo_residuals(gpu_tomo_t *g, double *cgres, int nbatch) {
    if (!g || !cgres || nbatch <= 0 || nbatch > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipDeviceSynchronize());
        if (g->t64) {
            chk64(t64_residuals(g->t64, cgres, nbatch, g->stream));
            return;
        }
        HIPCHK(hipMemcpy(cgres, g->res.p, sizeof(double) * nbatch, hipMemcpyDeviceToHost));
    });
}

int gpu_tomo_L(gpu_tomo_t *g, const double *xin, double *xout, double alpha, double beta) {
    if (!g || !xin || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (g->t64) {
            std::vector<double> tmp(g->nxtot);
            chk64(t64_L(g->t64, xin, tmp.data(), g->stream));
            for (long i = 0; i < g->nxtot; i++) xout[i] = beta * xout[i] + alpha * tmp[i];
            return;
        }
        to_dev_x(g, g->z.p, xin);
        Rec rec{&g->evs, g->stream, false};
        enqueue_L(own_batch(g, 0, 1), BUF_Z, PART_NONE, rec);
        std::vector<double> tmp(g->nxtot);
        from_dev_rt(g, tmp.data(), g->Ap.p);
        for (long i = 0; i < g->nxtot; i++) xout[i] = beta * xout[i] + alpha * tmp[i];
    });
}

int gpu_tomo_R(gpu_tomo_t *g, const double *grad, double *xout, double alpha) {
    if (!g || !grad || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (g->t64) {
            chk64(t64_R(g->t64, grad, xout, g->stream));
            for (long i = 0; i < g->nxtot; i++) xout[i] *= alpha;
            return;
        }
        to_dev_grad(g, 0, grad);
        Rec rec{&g->evs, g->stream, false};
        enqueue_R(own_batch(g, 0, 1), rec);
        from_dev_rt(g, xout, g->b.p);
        for (long i = 0; i < g->nxtot; i++) xout[i] *= alpha;
    });
}

int gpu_tomo_P(gpu_tomo_t *g, const double *xin, double *xout) {
    if (!g || !xin || !xout) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        if (!g->G.precond) fail(MT_EINVAL, "no FDPCG preconditioner in this plan");
        if (g->t64) {
            chk64(t64_P(g->t64, xin, xout, g->stream));
            return;
        }
        // b <- xin, r <- b through the cold-start row kernel (which also
        // clears its x argument: the p buffer is used as that scratch), then P
        to_dev_rt(g, g->b.p, xin);
        Batch &B = own_batch(g, 0, 1);
        Rec rec{&g->evs, g->stream, false};
        for (auto &s : B.zero) launch_zero_stop(tab(B, *s), g->stream);
        enqueue_row(B, KC_ROW_START, ROW_MODE_INIT_COLD, 0, BUF_P, rec);
        enqueue_P_tail(B, IROW_MODE_Z, 0, BUF_Z, rec);
        from_dev_x(g, xout, g->z.p);
    });
}

int gpu_tomo_set_grad(gpu_tomo_t *g, int isys, const double *grad) {
    if (!g || !grad || isys < 0 || isys >= g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        no_mode4(g, "gpu_tomo_set_grad");
        to_dev_grad(g, isys, grad);
    });
}

int gpu_tomo_get_x(gpu_tomo_t *g, int isys, double *x, double *cgres) {
    if (!g || !x || isys < 0 || isys >= g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipDeviceSynchronize());
        if (g->t64) {
            chk64(t64_get_x(g->t64, isys, x, cgres, g->stream));
            return;
        }
        from_dev_x(g, x, g->x.p + (long)isys * g->G.nxtot);
        if (cgres) HIPCHK(hipMemcpy(cgres, g->res.p + isys, sizeof(double), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

struct gpu_tomo_group {
    std::vector<gpu_tomo *> plans;
    std::vector<int> s0, nsys;
    std::vector<long> gens;           // each member's gen when B was built
    std::unique_ptr<Batch> B;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // recorded after each solve (of B, by graph or launches)
    int device = 0;
    EvList evs;
};
namespace {
// Replace the group's batch (and drop its graph) once its last solve, which
// an asynchronous caller may not have waited for, has completed.
void group_rebuild(gpu_tomo_group *grp) {
    if (grp->done) HIPCHK(hipEventSynchronize(grp->done));
    if (grp->gexec) (void)hipGraphExecDestroy(grp->gexec);
    if (grp->graph) (void)hipGraphDestroy(grp->graph);
    grp->gexec = nullptr;
    grp->graph = nullptr;
    grp->B = make_batch(grp->plans, grp->s0, grp->nsys);
    for (size_t j = 0; j < grp->plans.size(); j++) grp->gens[j] = grp->plans[j]->gen;
}
}  // namespace

extern "C" {

int gpu_tomo_group_create(int n, gpu_tomo_t *const *plans, const int *s0, const int *nsys,
                          gpu_tomo_group_t **out) {
    if (!out || n <= 0 || !plans || !s0 || !nsys) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<gpu_tomo_group> grp(new (std::nothrow) gpu_tomo_group());
    if (!grp) return MT_ENOMEM;
    int rc = guarded([&] {
        std::vector<gpu_tomo *> pl(plans, plans + n);
        for (auto *g : pl) {
            if (!g) fail(MT_EINVAL, "null plan");
            no_mode4(g, "gpu_tomo_group_create");
        }
        grp->device = pl[0]->device;
        HIPCHK(hipSetDevice(grp->device));
        grp->plans = pl;
        grp->s0.assign(s0, s0 + n);
        grp->nsys.assign(nsys, nsys + n);
        grp->B = make_batch(pl, grp->s0, grp->nsys);
        for (auto *g : pl) grp->gens.push_back(g->gen);
        HIPCHK(hipStreamCreateWithFlags(&grp->stream, hipStreamNonBlocking));
    });
    if (rc == MT_OK) *out = grp.release();
    return rc;
}

// The group's tables follow its members: a member whose launch shape or
// precision changed since (gpu_tomo_tune_set, gpu_tomo_precision,
// gpu_tomo_autotune) has a new gen, and the group rebuilds its batch and
// captures its graph again before solving.
static void group_refresh(gpu_tomo_group *grp) {
    for (auto *g : grp->plans) no_mode4(g, "gpu_tomo_group_solve");
    bool moved = false;
    for (size_t j = 0; j < grp->plans.size(); j++) moved |= grp->plans[j]->gen != grp->gens[j];
    if (!moved) return;
    group_rebuild(grp);
}

int gpu_tomo_group_solve(gpu_tomo_group_t *grp, int mode, void *stream) {
    if (!grp) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(grp->device));
        group_refresh(grp);
        hipStream_t st = stream ? (hipStream_t)stream : grp->stream;
        if (!grp->done) HIPCHK(hipEventCreateWithFlags(&grp->done, hipEventDisableTiming));
        if (mode == 1) {
            if (!grp->gexec) capture_solve(*grp->B, &grp->graph, &grp->gexec, &grp->evs);
            HIPCHK(hipGraphLaunch(grp->gexec, st));
        } else {
            Rec rec{&grp->evs, st, false};
            enqueue_solve(*grp->B, rec);
        }
        HIPCHK(hipEventRecord(grp->done, st));
        if (!stream) HIPCHK(hipStreamSynchronize(st));
    });
}

// Launch shape of a group: the same choice set on every member plan (a
// plan property, so each system's bits stay those of the plan alone), the
// tables rebuilt and the graph captured again on next use.
namespace {
void group_apply(gpu_tomo_group *grp, const int *c) {
    for (auto *g : grp->plans) {
        g->tune_wfs = c[0];
        g->tune_layer = c[1];
        g->tune_col = c[2];
        g->tune_lty = c[3];
        g->tune_colfft = c[4];
        drop_batches(g);
        apply_lty(g);
        noalloc_init(g);
    }
    group_rebuild(grp);
}
}  // namespace

int gpu_tomo_group_autotune(gpu_tomo_group_t *grp, int *choice) {
    if (!grp) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(grp->device));
        HIPCHK(hipDeviceSynchronize());
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        auto timed = [&](const int *c) {
            group_apply(grp, c);
            capture_solve(*grp->B, &grp->graph, &grp->gexec, &grp->evs);
            std::vector<float> ts;
            for (int rep = 0; rep < 4; rep++) {
                HIPCHK(hipEventRecord(e0, grp->stream));
                HIPCHK(hipGraphLaunch(grp->gexec, grp->stream));
                HIPCHK(hipEventRecord(e1, grp->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) ts.push_back(ms);
            }
            std::sort(ts.begin(), ts.end());
            return (double)ts[ts.size() / 2];
        };
        // coordinate descent from the heuristic: column kernel, layer tile
        // rows, then systems per workgroup of the WFS and layer passes; a
        // change is kept when it saves more than 2 %
        int best[6] = {0, 0, 0, 0, 0, 1};
        double tbest = timed(best);
        auto trial = [&](int k, std::initializer_list<int> vals) {
            for (int v : vals) {
                if (v == best[k]) continue;
                int c[6];
                std::copy(best, best + 6, c);
                c[k] = v;
                const double t = timed(c);
                if (t < tbest * 0.98) {
                    tbest = t;
                    best[k] = v;
                }
            }
        };
        trial(4, {2});
        trial(3, {LTY2, TILE});
        trial(0, {1, 2, 4});
        trial(1, {1, 2, 4});
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        group_apply(grp, best);
        if (choice) std::copy(best, best + 6, choice);
    });
}

void gpu_tomo_group_free(gpu_tomo_group_t *grp) {
    if (!grp) return;
    (void)hipSetDevice(grp->device);
    (void)hipDeviceSynchronize();
    if (grp->gexec) (void)hipGraphExecDestroy(grp->gexec);
    if (grp->graph) (void)hipGraphDestroy(grp->graph);
    if (grp->done) (void)hipEventDestroy(grp->done);
    if (grp->stream) (void)hipStreamDestroy(grp->stream);
    delete grp;
}

int gpu_tomo_tune_set(gpu_tomo_t *g, const int *choice) {
    if (!g || !choice) return MT_EINVAL;
    for (int i = 0; i < 3; i++)
        if (choice[i] != 0 && choice[i] != 1 && choice[i] != 2 && choice[i] != 4) return MT_EINVAL;
    if (choice[3] != 0 && choice[3] != TILE && choice[3] != LTY2 && choice[3] != LTY) return MT_EINVAL;
    if (choice[4] < 0 || choice[4] > 2) return MT_EINVAL;
    if (choice[5] != 0 && choice[5] != 1 && choice[5] != 2 && choice[5] != 4) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipDeviceSynchronize());
        g->tune_wfs = choice[0];
        g->tune_layer = choice[1];
        g->tune_col = choice[2];
        g->tune_lty = choice[3];
        g->tune_colfft = choice[4];
        g->tune_flights = choice[5];
        drop_batches(g);
        apply_lty(g);
        noalloc_init(g);
    });
}

int gpu_tomo_autotune(gpu_tomo_t *g, int nsys, int *choice) {
    if (!g || nsys <= 0 || nsys > g->nbatch) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipDeviceSynchronize());
        // representative input: a small constant slope on every subaperture
        // (the work of a solve does not depend on the values); the slots'
        // previous gradients and solutions are overwritten
        float v = 1e-7f;
        unsigned bits;
        std::memcpy(&bits, &v, 4);
        HIPCHK(hipMemsetD32((hipDeviceptr_t)g->g.p, (int)bits, (size_t)nsys * g->G.ngtot));
        const int save_prof = g->prof;
        g->prof = 1;
        // time one stage class under each candidate; other stages keep their choice
        auto stage_ms = [&](int kc) {
            drop_batches(g);
            Batch &B = own_batch(g, 0, nsys);
            double best = 1e30;
            for (int rep = 0; rep < 3; rep++) {
                // the timed classes only read their inputs: 4 launches per
                // event pair keep the event overhead out of the comparison
                Rec rec{&g->evs, g->stream, true, 4};
                enqueue_solve(B, rec);
                prof_collect(g);
                best = std::min(best, g->prof_ms[kc] / std::max(g->prof_cnt[kc], 1));
            }
            return best;
        };
        struct Knob {
            int *field, kc, cmax;
        };
        const int colmax = 4;  // clamped by the LDS per workgroup when the launch is built
        // the layer stage: tile shape (16 x 64, 16 x 32 or 16 x 16, where
        // they fit) and systems per workgroup together
        {
            double best = 1e30;
            int bestl = g->G.lty, bestc = 1;
            for (int lty : {LTY, LTY2, TILE}) {
                if (lty == LTY && !(g->G.matched && g->lty_ok)) continue;
                if (lty == LTY2 && !(g->G.matched && g->lty2_ok)) continue;
                g->tune_lty = lty;
                apply_lty(g);
                for (int c = 1; c <= 4; c <<= 1) {
                    g->tune_layer = c;
                    const double t = stage_ms(KC_LAYER);
                    if (t < best * 0.98) {  // prefer the earlier (larger tile, fewer systems) within 2%
                        best = t;
                        bestl = lty;
                        bestc = c;
                    }
                }
            }
            g->tune_lty = bestl;
            g->tune_layer = bestc;
            apply_lty(g);
        }
        // the column stage: register kernel (where it applies; one system
        // per lane group, no systems-per-workgroup choice) or the LDS kernel
        // with 1, 2 or 4 systems per workgroup
        {
            double best = 1e30;
            int bestf = 2, bestc = 1;
            for (int f : {1, 2}) {
                if (f == 1 && !g->colr_ok) continue;
                g->tune_colfft = f;
                for (int c = 1; c <= (f == 1 ? 1 : colmax); c <<= 1) {
                    g->tune_col = c;
                    const double t = stage_ms(KC_COL);
                    if (t < best * 0.98) {
                        best = t;
                        bestf = f;
                        bestc = c;
                    }
                }
            }
            g->tune_colfft = bestf;
            g->tune_col = bestf == 1 ? 0 : bestc;
        }
        Knob knobs[1] = {{&g->tune_wfs, KC_WFS, g->G.pos == 1 ? 4 : 2}};
        for (auto &k : knobs) {
            int bestc = 1;
            double bestt = 1e30;
            for (int c = 1; c <= k.cmax; c <<= 1) {
                *k.field = c;
                const double t = stage_ms(k.kc);
                if (t < bestt * 0.98) {  // prefer the smaller choice within 2%
                    bestt = t;
                    bestc = c;
                }
            }
            *k.field = bestc;
        }
        drop_batches(g);
        g->prof = save_prof;
        // flights: whole graph solves of the nsys systems, timed on the
        // plan's stream (median of 5 after 2 warm-up solves)
        {
            const int save_pr = g->prof;
            g->prof = 0;
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            double best = 1e30;
            int bestk = 1;
            for (int k : {1, 2, 4}) {
                if (k > nsys) break;
                g->tune_flights = k;
                std::vector<float> ts;
                for (int rep = 0; rep < 7; rep++) {
                    HIPCHK(hipEventRecord(e0, g->stream));
                    graph_solve(g, nsys, g->stream);
                    HIPCHK(hipEventRecord(e1, g->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (rep >= 2) ts.push_back(ms);
                }
                std::sort(ts.begin(), ts.end());
                const double t = ts[ts.size() / 2];
                if (t < best * 0.98) {  // prefer fewer flights within 2%
                    best = t;
                    bestk = k;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            g->tune_flights = bestk;
            g->prof = save_pr;
            drop_batches(g);
        }
        noalloc_init(g);
        if (choice) {
            choice[0] = g->tune_wfs;
            choice[1] = g->tune_layer;
            choice[2] = g->tune_col;
            choice[3] = g->G.lty;
            choice[4] = g->tune_colfft;
            choice[5] = g->tune_flights;
        }
    });
}

int gpu_tomo_precision(gpu_tomo_t *g, int mode) {
    if (!g || mode < 0 || mode > 4) return MT_EINVAL;
    return guarded([&] {
        HIPCHK(hipSetDevice(g->device));
        HIPCHK(hipStreamSynchronize(g->stream));
        HIPCHK(hipDeviceSynchronize());
        if (mode == 4) {
            // the FP64 plan takes over the warm-start state
            if (!g->t64) {
                tomo64 *t = nullptr;
                chk64(t64_create(t64_view(g->rc64), g->nbatch, &t));
                g->t64 = t;
                chk64(t64_load_x(g->t64, g->x.p, g->stream));
            }
            return;
        }
        if (g->t64) {
            // back to an FP32-operator mode: the FP64 warm-start state, rounded
            const int rc = t64_store_x(g->t64, g->x.p, g->res.p, g->stream);
            t64_free(g->t64);
            g->t64 = nullptr;
            chk64(rc);
        }
        g->rt64 = mode == 1 || mode == 2;
        g->spec64 = mode == 1;
        g->G.dot32 = mode == 3;
        upload_plandev(g);
        drop_batches(g);
        noalloc_init(g);
    });
}

int gpu_tomo_profile(gpu_tomo_t *g, int enable) {
    if (!g || enable < 0) return MT_EINVAL;
    g->prof = enable;
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

// ===========================================================================
// Request loop (continuous batching).  Each AO loop owns a plan with one
// warm-start slot.  Every stream ("flight") holds a batch over all admitted
// loops, captured once in a hipGraph; a launch switches on the jobs of the
// loops that have a request in it (the others' workgroups exit at entry).
// A worker thread collects the pending requests of loops that are not on the
// GPU, launches them on a free flight and retires flights as their events
// complete, so requests that arrive while others run are admitted as soon
// as a stream and their loop are free, and flights overlap on the device.
// Loops are admitted by HBM footprint against a budget.
// ===========================================================================
namespace {
typedef std::chrono::steady_clock Clock;

struct SrvLoop {
    gpu_tomo *plan = nullptr;
    size_t bytes = 0;
    bool busy = false;                 // a request of this loop is on the GPU
    bool hold = false;                 // its plan is being updated: admit none of its requests
    std::vector<double> xlast;         // last completed solution (host)
    double reslast = NAN;
};
struct SrvReq {
    int loop;
    long tag;
    std::vector<float> grad;
    Clock::time_point t_submit;
};
struct SrvDone {
    int loop;
    long tag;
    double res, lat_us;
};
struct Flight {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool active = false;
    std::vector<SrvReq> reqs;
    std::vector<SrvLoop *> lp;   // the requests' loops
    // batch over all admitted loops (rebuilt when loops change)
    long gen = -1;
    std::unique_ptr<Batch> B;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Job *hjobs = nullptr;        // pinned copy of the job table (on flags set per launch)
    std::vector<int> job_of;     // loop id -> job index (-1: not in the batch)
    // packed staging of all loops of the batch: one copy each way per launch
    std::vector<StageOff> hoff;  // per job
    DevBuf<StageOff> doff;
    float *hin = nullptr, *hout = nullptr;  // pinned: gradients, solutions
    double *hres = nullptr;                 // pinned: residuals
    DevBuf<float> din, dout;
    DevBuf<double> dres;
    long nin = 0, nout = 0, nres = 0;
};
size_t plan_bytes(const gpu_tomo *g) {
    size_t b = 0;
    b += g->hx_nx.n * sizeof(int2) + g->hx_ny.n * sizeof(int2) + (g->hx_fx.n + g->hx_fy.n) * sizeof(float);
    b += (g->hxt_x.n + g->hxt_y.n) * sizeof(int4) + g->mray.n * sizeof(MRay) + g->saidx.n * sizeof(int);
    b += (g->gpx.n + g->gpy.n + g->ptt.n) * sizeof(float) + (g->neai.n + g->nead.n) * sizeof(float4);
    b += (g->gpd.n + g->twx.n + g->twy.n + g->Minv.n) * sizeof(float2) + (g->twx64.n + g->twy64.n) * sizeof(double2);
    b += (g->grp_col.n + g->colkx.n + g->grp_blk.n + g->blkpos.n) * sizeof(int) + g->rowblk.n * sizeof(int2);
    b += (g->x.n + g->z.n + g->p.n + g->g.n + g->tt.n) * sizeof(float);
    b += (g->r.n + g->Ap.n + g->b.n + g->q.n) * sizeof(double) + g->specf.n * sizeof(double2) + g->speci.n * sizeof(float2);
    b += (g->part_pap.n + g->part_rz.n + g->part_rr.n + g->part_bb.n + g->rz.n + g->alpha.n + g->res.n) * sizeof(double);
    return b;
}
}  // namespace

// Ownership: `flights` (every field of every Flight) belongs to the worker
// thread alone; other threads reach it only by asking the worker (rebuild
// requests below).  Everything else is shared and guarded by `mu`.
struct gpu_tomo_server {
    int device = 0, max_inflight = 2, max_batch = 64, use_graph = 1;
    size_t budget = 0, used = 0;
    std::vector<std::unique_ptr<SrvLoop>> loops;  // index = loop id (plan null once removed)
    std::deque<SrvReq> pending;
    std::vector<SrvDone> done;
    std::vector<Flight> flights;
    long gen = 0;                      // bumped when loops are added or removed
    long rebuild_gen = -1;             // ask the worker to rebuild every flight for this gen
    long built_gen = -1;               // the gen it last rebuilt every flight for
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::thread worker;
    std::atomic<bool> stop{false};
    long inflight_reqs = 0;
    std::string err;
    EvList evs;
};

namespace {
void flight_drop(Flight &f) {
    if (f.exec) (void)hipGraphExecDestroy(f.exec);
    if (f.graph) (void)hipGraphDestroy(f.graph);
    f.exec = nullptr;
    f.graph = nullptr;
    if (f.hjobs) (void)hipHostFree(f.hjobs);
    f.hjobs = nullptr;
    if (f.hin) (void)hipHostFree(f.hin);
    if (f.hout) (void)hipHostFree(f.hout);
    if (f.hres) (void)hipHostFree(f.hres);
    f.hin = f.hout = nullptr;
    f.hres = nullptr;
    f.din.free();
    f.dout.free();
    f.dres.free();
    f.doff.free();
    f.hoff.clear();
    f.B.reset();
}
// (Re)build an idle flight's batch over all admitted loops (worker thread,
// lock held).  The flight counts as current for S->gen only once the whole
// build has succeeded: a build that throws leaves it dropped and stale, so
// it is built again before any launch (a flight marked current with a
// partial build would be launched with null host buffers).
void flight_build_body(gpu_tomo_server *S, Flight &f);
void flight_build(gpu_tomo_server *S, Flight &f) {
    flight_drop(f);
    f.gen = -1;
    try {
        flight_build_body(S, f);
    } catch (...) {
        flight_drop(f);
        f.gen = -1;
        throw;
    }
    f.gen = S->gen;
}
void flight_build_body(gpu_tomo_server *S, Flight &f) {
    std::vector<gpu_tomo *> pl;
    f.job_of.assign(S->loops.size(), -1);
    for (size_t i = 0; i < S->loops.size(); i++)
        if (S->loops[i]->plan) {
            f.job_of[i] = (int)pl.size();
            pl.push_back(S->loops[i]->plan);
        }
    if (pl.empty()) return;
    f.B = make_batch(pl, std::vector<int>(pl.size(), 0), std::vector<int>(pl.size(), 1));
    HIPCHK(hipHostMalloc((void **)&f.hjobs, sizeof(Job) * pl.size()));
    f.hoff.resize(pl.size());
    f.nin = f.nout = f.nres = 0;
    for (size_t j = 0; j < pl.size(); j++) {
        f.hoff[j] = StageOff{f.nin, f.nout, f.nres};
        f.nin += pl[j]->G.ngtot;
        f.nout += pl[j]->G.nxtot;
        f.nres += 1;
    }
    HIPCHK(hipHostMalloc((void **)&f.hin, sizeof(float) * f.nin));
    HIPCHK(hipHostMalloc((void **)&f.hout, sizeof(float) * f.nout));
    HIPCHK(hipHostMalloc((void **)&f.hres, sizeof(double) * f.nres));
    std::memset(f.hin, 0, sizeof(float) * f.nin);
    f.din.alloc(f.nin);
    f.dout.alloc(f.nout);
    f.dres.alloc(f.nres);
    f.doff.upload(f.hoff);
    if (S->use_graph) {
        Flight *fp = &f;
        const int nj = (int)pl.size();
        capture_solve_with(
            *f.B, &f.graph, &f.exec, &S->evs,
            [&](hipStream_t cs) { launch_stage_in(fp->B->jobs.p, nj, fp->doff.p, fp->din.p, cs); },
            [&](hipStream_t cs) { launch_stage_out(fp->B->jobs.p, nj, fp->doff.p, fp->dout.p, fp->dres.p, cs); });
    }
}

// Launch the requests in f.reqs on f.st (worker thread; f.lp resolved and
// the batch built under the lock).
void srv_launch(gpu_tomo_server *S, Flight &f) {
    Batch &B = *f.B;
    for (size_t j = 0; j < B.hjob.size(); j++) {
        f.hjobs[j] = B.hjob[j];
        f.hjobs[j].on = 0;
    }
    for (auto &r : f.reqs) {
        const int j = f.job_of[r.loop];
        f.hjobs[j].on = 1;
        std::memcpy(f.hin + f.hoff[j].g, r.grad.data(), sizeof(float) * r.grad.size());
    }
    const int nj = (int)B.hjob.size();
    HIPCHK(hipMemcpyAsync(B.jobs.p, f.hjobs, sizeof(Job) * nj, hipMemcpyHostToDevice, f.st));
    HIPCHK(hipMemcpyAsync(f.din.p, f.hin, sizeof(float) * f.nin, hipMemcpyHostToDevice, f.st));
    if (S->use_graph) {
        HIPCHK(hipGraphLaunch(f.exec, f.st));
    } else {
        launch_stage_in(B.jobs.p, nj, f.doff.p, f.din.p, f.st);
        Rec rec{&S->evs, f.st, false};
        enqueue_solve(B, rec);
        launch_stage_out(B.jobs.p, nj, f.doff.p, f.dout.p, f.dres.p, f.st);
    }
    HIPCHK(hipMemcpyAsync(f.hout, f.dout.p, sizeof(float) * f.nout, hipMemcpyDeviceToHost, f.st));
    HIPCHK(hipMemcpyAsync(f.hres, f.dres.p, sizeof(double) * f.nres, hipMemcpyDeviceToHost, f.st));
    HIPCHK(hipEventRecord(f.done, f.st));
    f.active = true;
}

constexpr double HOLD_US = 250.0;

void srv_worker(gpu_tomo_server *S) {
    (void)hipSetDevice(S->device);
    try {
        while (!S->stop) {
            bool did = false;
            // rebuild requests (server autotune): every flight is idle then
            {
                std::lock_guard<std::mutex> lk(S->mu);
                if (S->rebuild_gen > S->built_gen) {
                    const long want = S->rebuild_gen;
                    for (auto &f : S->flights)
                        if (!f.active) flight_build(S, f);
                    S->built_gen = want;
                    S->cv_done.notify_all();
                }
            }
            // retire completed flights
            for (auto &f : S->flights) {
                if (!f.active) continue;
                const hipError_t q = hipEventQuery(f.done);
                if (q == hipErrorNotReady) continue;
                HIPCHK(q);
                const auto now = Clock::now();
                std::lock_guard<std::mutex> lk(S->mu);
                for (size_t i = 0; i < f.reqs.size(); i++) {
                    SrvLoop &L = *f.lp[i];
                    const SrvReq &r = f.reqs[i];
                    const StageOff &o = f.hoff[f.job_of[r.loop]];
                    L.xlast.assign(f.hout + o.x, f.hout + o.x + L.plan->G.nxtot);
                    L.reslast = f.hres[o.r];
                    L.busy = false;
                    S->done.push_back(SrvDone{r.loop, r.tag, L.reslast,
                                              std::chrono::duration<double, std::micro>(now - r.t_submit).count()});
                }
                S->inflight_reqs -= (long)f.reqs.size();
                f.reqs.clear();
                f.lp.clear();
                f.active = false;
                did = true;
                S->cv_done.notify_all();
            }
            // admit pending requests of idle loops into a free flight.  A
            // launch costs about the same for few loops as for all (the
            // small systems' kernels are launch-bound), so while another
            // flight is on the GPU a flight that would carry fewer than half
            // of the loops waits for more requests, up to HOLD_US after the
            // oldest of them arrived.
            bool anyactive = false;
            for (auto &f : S->flights) anyactive |= f.active;
            for (auto &f : S->flights) {
                if (f.active) continue;
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    if (anyactive) {
                        int nl = 0, ready = 0;
                        Clock::time_point oldest = Clock::now();
                        for (auto &L : S->loops) nl += L->plan != nullptr;
                        std::vector<int> seen;
                        for (auto &r : S->pending) {
                            if (S->loops[r.loop]->busy || S->loops[r.loop]->hold ||
                                std::find(seen.begin(), seen.end(), r.loop) != seen.end())
                                continue;
                            seen.push_back(r.loop);
                            ready++;
                            oldest = std::min(oldest, r.t_submit);
                        }
                        const double waited = std::chrono::duration<double, std::micro>(Clock::now() - oldest).count();
                        if (2 * ready < nl && waited < HOLD_US) break;
                    }
                    std::vector<int> taken;
                    for (auto it = S->pending.begin(); it != S->pending.end() && (int)f.reqs.size() < S->max_batch;) {
                        SrvLoop &L = *S->loops[it->loop];
                        if (L.busy || L.hold || std::find(taken.begin(), taken.end(), it->loop) != taken.end()) {
                            ++it;
                            continue;
                        }
                        L.busy = true;
                        taken.push_back(it->loop);
                        f.lp.push_back(&L);
                        f.reqs.push_back(std::move(*it));
                        it = S->pending.erase(it);
                    }
                    S->inflight_reqs += (long)f.reqs.size();
                    if (!f.reqs.empty() && f.gen != S->gen) flight_build(S, f);
                }
                if (f.reqs.empty()) break;
                srv_launch(S, f);
                anyactive = true;
                did = true;
            }
            if (!did) {
                std::unique_lock<std::mutex> lk(S->mu);
                anyactive = false;
                for (auto &f : S->flights) anyactive |= f.active;
                if (anyactive) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(5));
                } else {
                    S->cv.wait_for(lk, std::chrono::milliseconds(1));
                }
            }
        }
    } catch (const HipError &e) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = e.msg;
        S->stop = true;
        S->cv_done.notify_all();
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = e.what();
        S->stop = true;
        S->cv_done.notify_all();
    }
}
}  // namespace

extern "C" {

int gpu_tomo_server_create(int device, double mem_budget_gb, int max_inflight, int max_batch, int use_graph,
                           gpu_tomo_server_t **out) {
    if (!out || max_inflight < 1 || max_inflight > 8 || max_batch < 1 || mem_budget_gb <= 0) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<gpu_tomo_server> S(new (std::nothrow) gpu_tomo_server());
    if (!S) return MT_ENOMEM;
    int rc = guarded([&] {
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(MT_EINVAL, "no such device");
        S->device = device;
        HIPCHK(hipSetDevice(device));
        S->budget = (size_t)(mem_budget_gb * 1e9);
        S->max_inflight = max_inflight;
        S->max_batch = max_batch;
        S->use_graph = use_graph ? 1 : 0;
        S->flights.resize(max_inflight);
        for (auto &f : S->flights) {
            HIPCHK(hipStreamCreateWithFlags(&f.st, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&f.done, hipEventDisableTiming));
        }
        S->worker = std::thread(srv_worker, S.get());
    });
    if (rc == MT_OK) *out = S.release();
    return rc;
}

int gpu_tomo_server_add_loop(gpu_tomo_server_t *S, const mt_recon_t *recon, int maxit, int *loop) {
    if (!S || !recon || !loop) return MT_EINVAL;
    {
        // loops of one server advance in lockstep batches: one iteration count
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto &L : S->loops)
            if (L->plan && L->plan->maxit != maxit) {
                mt::set_error("all loops of a server use the same maxit");
                return MT_EINVAL;
            }
    }
    gpu_tomo_t *g = nullptr;
    int rc = gpu_tomo_setup(recon, S->device, 1, maxit, 1, &g);
    if (rc != MT_OK) return rc;
    // the server always solves its loops in batches over several plans:
    // FFTs in LDS stages, so each FFT stage stays one launch for all loops
    // (a loop's path is the same alone and in any mix, so its bits are too)
    g->tune_colfft = 2;
    const size_t bytes = plan_bytes(g);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->used + bytes > S->budget) {
        gpu_tomo_free(g);
        mt::set_error("loop refused: HBM budget exceeded");
        return MT_EFULL;
    }
    rc = guarded([&] {
        std::unique_ptr<SrvLoop> L(new SrvLoop());
        L->plan = g;
        L->bytes = bytes;
        S->used += bytes;
        S->loops.push_back(std::move(L));
        *loop = (int)S->loops.size() - 1;
        S->gen++;
    });
    if (rc != MT_OK) gpu_tomo_free(g);
    return rc;
}

int gpu_tomo_server_remove_loop(gpu_tomo_server_t *S, int loop) {
    if (!S) return MT_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    if (loop < 0 || loop >= (int)S->loops.size() || !S->loops[loop]->plan) return MT_EINVAL;
    for (auto &r : S->pending)
        if (r.loop == loop) return MT_ESTATE;
    if (S->loops[loop]->busy || S->loops[loop]->hold) return MT_ESTATE;
    SrvLoop &L = *S->loops[loop];
    // flights still holding a batch with this loop switch it off (its job is
    // never on again) and rebuild before their next launch
    S->gen++;
    gpu_tomo_free(L.plan);
    S->used -= L.bytes;
    L = SrvLoop();
    return MT_OK;
}

int gpu_tomo_server_autotune(gpu_tomo_server_t *S, int *choice) {
    if (!S) return MT_EINVAL;
    // the lock is held (but for the wait on the worker's rebuild): the worker
    // admits nothing meanwhile, and with no request pending or in flight no
    // batch is on the GPU
    std::unique_lock<std::mutex> lk(S->mu);
    if (!S->pending.empty() || S->inflight_reqs) {
        mt::set_error("server autotune needs an idle server (drain first)");
        return MT_ESTATE;
    }
    std::vector<gpu_tomo_t *> pl;
    for (auto &L : S->loops)
        if (L->plan) pl.push_back(L->plan);
    if (pl.empty()) return MT_EINVAL;
    const int n = (int)pl.size();
    gpu_tomo_group_t *grp = nullptr;
    std::vector<DevBuf<float>> keep(n);
    int rc = guarded([&] {
        HIPCHK(hipSetDevice(S->device));
        // the timed solves overwrite each plan's warm-start solution
        for (int j = 0; j < n; j++) {
            keep[j].alloc(pl[j]->x.n);
            HIPCHK(hipMemcpy(keep[j].p, pl[j]->x.p, sizeof(float) * pl[j]->x.n, hipMemcpyDeviceToDevice));
        }
    });
    if (rc != MT_OK) return rc;
    const std::vector<int> s0(n, 0), ns(n, 1);
    rc = gpu_tomo_group_create(n, pl.data(), s0.data(), ns.data(), &grp);
    if (rc == MT_OK) rc = gpu_tomo_group_autotune(grp, choice);
    gpu_tomo_group_free(grp);
    // every flight's batch tables and graph predate the new shapes: the
    // worker (which alone touches the flights) rebuilds them now rather than
    // at their first launch, inside the request stream
    S->gen++;
    int rc2 = guarded([&] {
        for (int j = 0; j < n; j++)
            HIPCHK(hipMemcpy(pl[j]->x.p, keep[j].p, sizeof(float) * pl[j]->x.n, hipMemcpyDeviceToDevice));
        HIPCHK(hipDeviceSynchronize());
    });
    if (rc == MT_OK && rc2 == MT_OK) {
        const long want = S->gen;
        S->rebuild_gen = want;
        S->cv.notify_all();
        S->cv_done.wait(lk, [&] { return S->built_gen >= want || !S->err.empty(); });
        if (!S->err.empty()) {
            mt::set_error(S->err);
            rc2 = MT_EHIP;
        }
    }
    return rc != MT_OK ? rc : rc2;
}

int gpu_tomo_server_update_cn2(gpu_tomo_server_t *S, int loop, const mt_recon_t *recon) {
    if (!S || !recon) return MT_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    if (loop < 0 || loop >= (int)S->loops.size() || !S->loops[loop]->plan || S->loops[loop]->hold)
        return MT_EINVAL;
    SrvLoop &L = *S->loops[loop];
    // no new flight takes this loop; wait for the one that carries it, if any
    // (other loops keep being served meanwhile)
    L.hold = true;
    S->cv_done.wait(lk, [&] { return !L.busy || !S->err.empty(); });
    int rc = MT_OK;
    if (!S->err.empty()) {
        mt::set_error(S->err);
        rc = MT_EHIP;
    } else {
        gpu_tomo *g = L.plan;
        lk.unlock();
        rc = guarded([&] {
            HIPCHK(hipSetDevice(S->device));
            update_cn2_plan(g, recon);
        });
        lk.lock();
    }
    L.hold = false;
    lk.unlock();
    S->cv.notify_all();
    return rc;
}

int gpu_tomo_server_submit(gpu_tomo_server_t *S, int loop, const double *grad, long tag) {
    if (!S || !grad) return MT_EINVAL;
    SrvReq r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->err.empty()) {
            mt::set_error(S->err);
            return MT_EHIP;
        }
        if (loop < 0 || loop >= (int)S->loops.size() || !S->loops[loop]->plan) return MT_EINVAL;
        r.grad.assign(S->loops[loop]->plan->G.ngtot, 0.f);
        const gpu_tomo *g = S->loops[loop]->plan;
        int rc = guarded([&] { pack_grad(g, grad, r.grad.data()); });
        if (rc != MT_OK) return rc;
    }
    r.loop = loop;
    r.tag = tag;
    r.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->pending.push_back(std::move(r));
    }
    S->cv.notify_one();
    return MT_OK;
}

int gpu_tomo_server_poll(gpu_tomo_server_t *S, int maxn, int *loop, long *tag, double *cgres, double *latency_us,
                         int *n) {
    if (!S || !n) return MT_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    int k = 0;
    while (k < maxn && !S->done.empty()) {
        const SrvDone d = S->done.front();
        S->done.erase(S->done.begin());
        if (loop) loop[k] = d.loop;
        if (tag) tag[k] = d.tag;
        if (cgres) cgres[k] = d.res;
        if (latency_us) latency_us[k] = d.lat_us;
        k++;
    }
    *n = k;
    if (!S->err.empty()) {
        mt::set_error(S->err);
        return MT_EHIP;
    }
    return MT_OK;
}

int gpu_tomo_server_drain(gpu_tomo_server_t *S) {
    if (!S) return MT_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return (S->pending.empty() && S->inflight_reqs == 0) || !S->err.empty(); });
    if (!S->err.empty()) {
        mt::set_error(S->err);
        return MT_EHIP;
    }
    return MT_OK;
}

int gpu_tomo_server_get_x(gpu_tomo_server_t *S, int loop, double *x, double *cgres) {
    if (!S || !x) return MT_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    if (loop < 0 || loop >= (int)S->loops.size() || !S->loops[loop]->plan) return MT_EINVAL;
    const SrvLoop &L = *S->loops[loop];
    if (L.xlast.empty()) return MT_ESTATE;
    std::copy(L.xlast.begin(), L.xlast.end(), x);
    if (cgres) *cgres = L.reslast;
    return MT_OK;
}

int gpu_tomo_server_usage(gpu_tomo_server_t *S, double *used_gb, double *budget_gb) {
    if (!S) return MT_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    if (used_gb) *used_gb = S->used / 1e9;
    if (budget_gb) *budget_gb = S->budget / 1e9;
    return MT_OK;
}

void gpu_tomo_server_free(gpu_tomo_server_t *S) {
    if (!S) return;
    S->stop = true;
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->device);
    (void)hipDeviceSynchronize();
    for (auto &f : S->flights) flight_drop(f);
    for (auto &L : S->loops) {
        if (!L->plan) continue;
        gpu_tomo_free(L->plan);
    }
    for (auto &f : S->flights) {
        if (f.done) (void)hipEventDestroy(f.done);
        if (f.st) (void)hipStreamDestroy(f.st);
    }
    delete S;
}

}  // extern "C"
