// Internal declarations of the FP64 host setup (geometry, operators, FDPCG).
// Every function restates the MAOS routine cited beside it.
#pragma once
#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "maos_tomo.h"

namespace mt {

typedef std::complex<double> cplx;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void set_error(const std::string &msg);
const char *last_error_cstr();

// sys/common.h:88 (EPS seen by math/ and maos/)
constexpr double EPS_SYS = 1e-15;
// lib/accphi.c:26 (EPS inside the ray-tracing code)
constexpr double EPS_PROP = 1e-12;

// A regular grid (MAOS map_t without data): origin, spacing, size.
struct Grid {
    long nx = 0, ny = 0;
    double ox = 0, oy = 0, dx = 0, dy = 0;
};

// One beam direction (a column of parms->dirs: thetax, thetay, hs).
struct Dir {
    double thetax, thetay, hs;
};

// math/mathmisc.c:327-360
long nextfftsize(long n);
// math/loc.c:1471-1560 (square and non-square size computation; the map
// content is all ones for square=1, which is the only case on this path)
Grid create_metapupil(const std::vector<Dir> &dirs, double D, double ht, double dx,
                      double dy, double offset, double guard, long ninx, long niny,
                      int pad, int square);
// math/loc.c:815-864 (gray-pixel circle / annulus on point coordinates)
void loccircle(double *phi, const double *x, const double *y, long n, double dx,
               double dy, double cx, double cy, double r, double val);
void locannular(double *phi, const double *x, const double *y, long n, double dx,
                double dy, double cx, double cy, double r, double rin, double val);

// Subaperture geometry of one powfs (maos/setup_powfs.c:255-420).
struct SaGeom {
    double dsa = 0;
    std::vector<double> locx, locy;  // lower-left corners
    std::vector<double> saa;         // normalized illuminated area
};
SaGeom setup_saloc(double D, double Din, const mt_powfs_t &pw);

// One GP operator (lib/mkg.c:113-470 called as in
// maos/setup_recon_prep.c:522): per subaperture, x and y rows as sparse
// (ploc index, weight) lists.
struct SparseRows {
    long nrow = 0, ncol = 0;
    std::vector<long> p;       // nrow+1
    std::vector<long> i;       // column indices
    std::vector<double> x;
};
SparseRows mkg_gp(const Grid &pgrid, const SaGeom &sa, double D, double Din,
                  double gdx);

// Laplacian coefficient (lib/laplacian.c:26-32).
double laplacian_coef(double r0, double weight, double dx);

// FDPCG preparation (maos/fdpcg.c:70-580).
struct FdpcgIn {
    int nps = 0;
    std::vector<long> nx, ny, os;
    int pos = 1;
    double dxp = 0;             // ploc spacing
    std::vector<double> ht;     // layer heights
    std::vector<double> l2cf;   // per layer: laplacian_coef * sqrt(cxxscale*TOMOSCALE)
    const SaGeom *sa = nullptr; // high-order powfs subapertures
    std::vector<double> thetax, thetay, neam;  // per high-order WFS
};
struct FdpcgOut {
    int bs = 0;
    long nb = 0, nbhalf = 0;
    int scale = 0;
    std::vector<double> Mhat;   // nb*bs*bs complex (re,im), column-major blocks
    std::vector<double> Minv;   // same layout, after the SVD pseudo inverse
    std::vector<long> perm;     // nb*bs
    std::vector<long> permhf;   // nbhalf*bs
    // The part that does not depend on the turbulence profile, kept for
    // profile updates: the WFS sum of every block (before the regulariser
    // diagonal), and per block member its index in the stacked fftshifted
    // layer spectra and the size of its layer.
    std::vector<double> MhatW;  // nb*bs*bs complex
    std::vector<long> diag_at;  // nb*bs
    std::vector<double> lsize;  // nb*bs
};
// fdpcg_prepare = fdpcg_geometry, then fdpcg_finish.  A profile update
// (layer weights, r0: only in.l2cf changes) reruns fdpcg_finish alone.
FdpcgOut fdpcg_prepare(const FdpcgIn &in);
FdpcgOut fdpcg_geometry(const FdpcgIn &in);
void fdpcg_finish(FdpcgOut &out, const FdpcgIn &in);

// Hermitian pseudo inverse of a bs x bs block with relative singular value
// threshold (math/blas.c:456-500 csvd_pow(A,-1,thres)).
void herm_pinv(const cplx *A, cplx *out, int n, double thres);
// In-place 2-D DFT of a column-major nx*ny complex array, sign -1 forward
// (FFTW_FORWARD, unnormalized).  O(n^3) separable DFT in long double: used
// only in setup.
void dft2(std::vector<cplx> &a, long nx, long ny, int sign);
// Quadrant swap (math/fft.c cfftshift for even sizes).
void fftshift2(std::vector<cplx> &a, long nx, long ny);

}  // namespace mt

// The host setup of one tomography geometry (the RECON_T fields of this path).
struct mt_setup {
    // parameters
    double D = 0, Din = 0;
    int npsr = 0;
    std::vector<double> ht, wt0;
    std::vector<int> os;
    double atmr_dx = 0, atmr_hs = INFINITY, r0 = 0, L0 = 0, za = 0;
    int pos = 1, guard = 1, precond = 1, ptt_lhs = 0, split = 1;
    double cxxscale = 1, tomoscale = 1e-12;
    std::vector<mt_powfs_t> powfs;
    std::vector<mt_wfs_t> wfs;
    std::vector<mt::Dir> dirs;
    // geometry
    std::vector<double> wt;  // clipped + normalised layer weights
    mt::Grid pg;
    std::vector<mt::Grid> xg;
    std::vector<mt::SaGeom> sa;        // per powfs
    std::vector<mt::SparseRows> gprow; // per powfs (rows = gradients)
    // per-powfs CSC export of GP (2nsa x nploc)
    std::vector<std::vector<long>> gp_p, gp_i;
    std::vector<std::vector<double>> gp_x;
    // per WFS
    std::vector<int> skip, ttr;
    std::vector<long> nsa;
    std::vector<double> dsa, thetax, thetay, hs, neam;
    std::vector<std::vector<double>> saloc, neai, ptt;
    // regularisation
    std::vector<double> l2c, l2cf;
    // FDPCG
    mt::FdpcgOut fd;
    // export views
    std::vector<long> xnx, xny;
    std::vector<double> xox, xoy, xdx;
    std::vector<const double *> v_saloc, v_saa, v_gpx, v_neai, v_ptt;
    std::vector<const long *> v_gpp, v_gpi;
};

