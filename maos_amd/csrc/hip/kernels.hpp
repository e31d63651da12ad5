// CDNA4 (gfx950) kernels of the tomography PCG/FDPCG solve.
//
// Data layout in HBM, per system s of a batch (systems share one geometry):
//   layer vectors  x, r, z, p, Ap, b : s*nxtot + xoff[l] + ix + iy*xnx[l]
//   gradients      g                 : s*ngtot + goff[w] + {x slopes, y slopes}
//   WFS fields     q                 : (s*nwfs + w)*nploc + ix + iy*pnx
//   spectra        S (complex)       : s*nspec + soff[l] + kx*ny + ky,
//                                      kx in [0, nx/2] (half spectrum,
//                                      column-major so FDPCG columns are
//                                      contiguous); forward spectrum of r
//                                      in the residual precision, inverse
//                                      spectrum in FP32
// All gathers, no atomics: every output element is written by exactly one
// thread, so results do not depend on scheduling (a system computes the same
// bits alone or inside a batch).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mtk {

constexpr int MAXL = 16;      // layers
constexpr int MAXW = 16;      // wavefront sensors (active)
constexpr int MAXST = 8;      // FFT stages
constexpr int TILE = 16;      // output tile edge of the WFS / layer passes
constexpr int TAPS = 4;       // max ploc samples landing on one layer node per axis
constexpr int ROWS = 16;      // rows per row-FFT workgroup (8 complex FFTs)
constexpr int LTY = 64;       // tile height of the 16 x 64 layer kernel (k_layer_m8)
constexpr int LTY2 = 32;      // tile height of its 16 x 32 form
constexpr int MAXN = 1024;    // largest FFT length (layer grid edge)

// Matched-sampling ray of one (WFS, layer) pair: ploc index i maps to layer
// nodes i + c and i + c + 1 with a constant fraction f, for i in [a0, a1)
// (per axis).  wrap: the last ray's second node is node 0 (edge case).
// qoff = w * nploc + ay0 * pnx: where the rows [ay0, ay1) of this WFS's
// field start within a system's WFS fields; nrow = ay1 - ay0 (0 if none).
struct MRay {
    int cx, cy, ax0, ax1, ay0, ay1;
    int wx, wy;
    float fx, fy;
    int nrow, pad_;
    long qoff;
};

// Division by a runtime constant d >= 1 as a multiply-high: m = ceil(2^32/d)
// (0 for d = 1); exact for dividends below 2^16.
struct FDiv {
    unsigned d, m;
};
// One Stockham stage of a length-n FFT: radix R, Ns = product of the
// previous radices, nb = n/R butterflies per transform, tstep = n/(Ns*R).
struct FStage {
    int R, Ns, nb, tstep;
    FDiv dnb, dNs;
};

// Geometry shared by all systems of a plan (device copy).
struct Geom {
    int npsr, nwfs, pos, np;      // np = pos+1
    int pnx, pny, nploc;
    long nxtot, ngtot;
    int xnx[MAXL], xny[MAXL];
    long xoff[MAXL + 1];
    int nsa[MAXW];
    long goff[MAXW + 1];
    float l2c[MAXL];
    int ttr[MAXW];
    // ray tables, per (w,l): forward per ploc index, transposed per layer index
    // forward x: hx_nx[(w*MAXL+l)*pnx + i] = (n0, n1) (n0 < 0: no ray), hx_fx: fraction
    const int2 *hx_nx, *hx_ny;
    const float *hx_fx, *hx_fy;
    const int4 *hxt_x, *hxt_y;    // per (w,l) per layer index: up to 4 ploc indices (-1 pad)
    long hxt_xoff[MAXL + 1];      // offset of layer l in a (w,*) transposed x table
    int matched;                  // every (w,l) ray is matched sampling (fast path)
    int lty;                      // layer tile height: TILE, or LTY / LTY2 (k_layer_m8)
    const MRay *mray;             // [w*MAXL + l]
    // subapertures
    const int *saidx;             // per WFS per ploc point: subaperture with that lower-left corner, -1
    const float *gpx, *gpy;       // per WFS per subap np*np coefficients
    const float4 *neai;           // per WFS per subap (xx, yy, xy, 0)
    const float *ptt;             // per WFS 2 x 2nsa (column-major), ttr only
    long saoff[MAXW + 1];         // subap offset of WFS w in gpx/neai (units of subaps)
    // the same GP / NEA laid out per ploc point (lower-left subaperture corner;
    // zero where no subaperture), one copy per class of WFS with identical
    // subapertures, GP and NEA (the WFS of one powfs)
    int gcls[MAXW];               // class of WFS w
    const float2 *gpd;            // [class][nploc][np*np] (x coef, y coef)
    const float4 *nead;           // [class][nploc] (xx, yy, xy, 0)
    // FDPCG
    int ptt_lhs;                  // the LHS removes TT from the ttr WFS' gradients (!split || splitlrt)
    int precond, bs, scale;
    int zero_dc;                  // zero the DC bin of the residual spectrum (see k_col)
    int dot32;                    // precision mode 3: dot products and alpha/beta in FP32
    int fnx, fny;                 // FFT sizes (all layers equal)
    long nspec;                   // complex entries per system (sum over layers (nx/2+1)*ny)
    long soff[MAXL + 1];
    float fscale;                 // 1/sqrt(nx*ny) when scale
    int nstx, nsty;               // FFT stages
    FStage fsx[MAXST], fsy[MAXST];
    int rmax;                     // largest radix of either axis
    int shL, shP;                 // column FFT in registers: ny = shL lanes x shP points (0: none)
    FDiv dfnx, dfny, dnpsr;       // division by fnx, fny, npsr
    FDiv dsmall[ROWS / 2 + 1];    // division by 1 .. ROWS/2 (complex rows per block)
    const float2 *twx, *twy;      // exp(-2 pi i k/n)
    const double2 *twx64, *twy64; // same in FP64 (forward transform of the FP64 residual)
    const float2 *Minv;           // nbhalf blocks bs*bs, column-major, group order
    const float2 *Mreg;           // the same blocks for the register column kernel: per (column,
                                  // k1, entry) the shL lanes (bin k1 + shP bitrev(lane)), or null
    int ngroup;                   // column groups of the block apply
    const int *grp_col;           // per group: first column index into colkx
    const int *colkx;             // column kx list (per group, contiguous)
    const int *grp_blk;           // per group: first block (group order), ngroup+1 entries
    const int *blkpos;            // per block in group order (nbhalf*bs): LDS position | BP_* flags
    int maxcol;                   // max columns in one group
    // row blocks of the row FFT kernels
    int nrowblk;
    const int2 *rowblk;           // (layer, first row), ROWS rows each
};


// Flags of a blkpos entry (the LDS position is the low BP_POS bits).
constexpr int BP_CONJ = 1 << 30;      // the entry is the conjugate of the stored half-spectrum bin
constexpr int BP_NOWRITE = 1 << 29;   // another block writes this bin
constexpr int BP_DC = 1 << 28;        // the bin is a layer's DC bin
constexpr int BP_W1 = 1 << 27;        // the bin is on the kx = 0 or kx = nx/2 column (Parseval weight 1)
constexpr int BP_POS = (1 << 27) - 1;

// Per-plan scalar state (device, FP64).
struct Scal {
    double *part_pap;   // [nsys][maxpart]
    double *part_rz;    // [nsys][maxpart]
    double *part_rr;    // [nsys][maxpart]
    double *part_bb;    // [nsys][maxpart]
    double *rz;         // [nsys][maxit+2] r.z history: rz[k] is the numerator of alpha_k
    double *alpha;      // [nsys] step of the current iteration (0 once stopped)
    double *res;        // [nsys]
    float *tt;          // [nsys][nwfs][2]
    int maxpart, maxit;
    int npart_layer;    // partial count written by the layer kernel
    int npart_row;      // partial count written by the row kernel (r.r)
    int npart_col;      // partial count written by the column kernel (r.z)
};

// Everything a kernel needs of one plan: geometry, the state buffers of its
// nbatch systems (residual-side arrays are double, or float when the plan
// runs precision mode 0) and the scalar state.  One device copy per plan.
struct PlanDev {
    Geom G;
    float *x, *z, *p, *g, *tt;
    void *r, *Ap, *b, *q;
    void *specf;        // forward spectrum (double2, or float2)
    float2 *speci;      // inverse spectrum
    Scal SC;
    int *stop;
    int ltiles;         // layer tiles per layer (layer partial index stride)
    // TT removal inside the LHS (ptt_lhs): the operator's own gradients
    // G HX p (residual precision, ngtot per system) and their TT projection
    // PTT G HX p (FP64, 2 per system and WFS); null otherwise
    void *gl;
    double *ttl;
};

// A job is a range of systems of one plan; a launch covers a list of jobs,
// and every workgroup finds its job and its (x, y, z) block in a table, so
// one launch serves systems of different geometries (an uneven batch) as
// well as one plan alone.
struct Job {
    const PlanDev *pd;
    int s0, nsys;
    int on;             // 0: the job's workgroups exit at once (idle slot)
    int pad_;
};
struct WGI {
    unsigned short job, bx, by, bz;
};
static_assert(sizeof(WGI) == 8, "workgroup entries are read as one 8-byte word");

// Offsets of one job's entries in packed staging buffers (gradients,
// solutions, residuals), in elements.
struct StageOff {
    long g, x, r;
};

// Buffer selectors (which state array of the plan a kernel reads/writes).
enum { BUF_X = 0, BUF_P = 1, BUF_Z = 2, BUF_AP = 3, BUF_B = 4 };
enum { PART_NONE = 0, PART_PAP = 1, PART_BB = 2 };

}  // namespace mtk

// Loads through the constant address space are uniform and read-only: the
// compiler uses scalar loads for them, as it does for kernel arguments.
#define CONSTAS __attribute__((address_space(4)))
namespace mtk {
typedef const CONSTAS Geom CG;
typedef const CONSTAS PlanDev CPD;
}  // namespace mtk
