// DM fitting geometry and operators, FP64, on the fit.square=1 path that the
// GPU grid fitting requires (maos/setup_parms.c:3001-3004).
//
// Restates:
//   setup_recon_floc     maos/setup_recon_prep.c:112-165 (square floc, W0/W1)
//   mkw_circular/annular lib/mkw.c:363-481 (bilinear-influence aperture
//                        weighting; calcwt* quadratures :190-356)
//   setup_recon_aloc     maos/setup_recon_prep.c:271-360 (square actuator maps)
//   mkh (linear)         lib/mkh.c:47-135 (HA rows, for the coupling)
//   genactcpl            lib/slaving.c:24-55
//   slaving              lib/slaving.c:63-227 (no stuck/floating actuators)
//   setup_fit_lrt        maos/setup_recon_fit.c:104-171
//   tikcr term           maos/setup_recon_fit.c:270-283
//   fit.maxit default    maos/setup_parms.c:2381-2385
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>

#include "setup_internal.hpp"

using namespace mt;

struct mt_fit {
    Grid fg;                          // floc (square map)
    std::vector<Grid> ag;             // actuator grids per DM
    std::vector<double> aht;          // DM heights
    std::vector<double> thetax, thetay, wt, hs;
    std::vector<double> W0, W1;       // 9 per floc point (row form), per point
    int nnw = 0;
    std::vector<double> NW;           // nact x nnw column-major
    std::vector<long> sl_p, sl_i;     // slaving regulariser, CSR (empty: none)
    std::vector<double> sl_x;
    double tikdiag = 0;
    std::vector<double> actcpl;
    int maxit = 0, warm = 0;
    // views
    std::vector<long> anx, any;
    std::vector<double> aox, aoy, adx;
    long nact() const {
        long n = 0;
        for (const Grid &g : ag) n += g.nx * g.ny;
        return n;
    }
};

namespace {

constexpr int WRES = 50;  // lib/mkw.c:20 (quadrature points per cell side)

inline double sq(double a) { return a * a; }

// Overlap integrals of two bilinear hat functions centred on grid points i
// and j, over the part of a circle (centre icx, icy, radius icr, in grid
// units) inside their common support, by midpoint quadrature with WRES^2
// points per cell wherever the support is cut by the rim (lib/mkw.c:190-356).
// Full cells have the closed forms 4/9 (self), 1/9 (edge neighbour), 1/36
// (corner neighbour).
double w_corner(int ix, int iy, int jx, int jy, double icx, double icy, double icr) {
    const double r2 = icr * icr;
    const double c[4] = {sq(ix - icx) + sq(iy - icy), sq(ix - icx) + sq(jy - icy), sq(jx - icx) + sq(iy - icy),
                         sq(jx - icx) + sq(jy - icy)};
    if (c[0] < r2 && c[1] < r2 && c[2] < r2 && c[3] < r2) return 1. / 36.;
    if (c[0] > r2 && c[1] > r2 && c[2] > r2 && c[3] > r2) return 0;
    const double h = 1. / WRES, sx = (double)(jx - ix) / WRES, sy = (double)(jy - iy) / WRES;
    double acc = 0;
    for (int b = 0; b < WRES; b++) {
        const double v = (b + 0.5) * sy;
        for (int a = 0; a < WRES; a++) {
            const double u = (a + 0.5) * sx;
            if (sq(ix + u - icx) + sq(iy + v - icy) < r2) acc += (1 - std::fabs(u)) * (1 - std::fabs(v)) * std::fabs(u * v);
        }
    }
    return acc * h * h;
}
// neighbours along x (same row)
double w_lr(int ix, int iy, int jx, int /*jy*/, double icx, double icy, double icr) {
    const double r2 = icr * icr;
    const double c[4] = {sq(ix - icx) + sq(iy + 1 - icy), sq(ix - icx) + sq(iy - 1 - icy),
                         sq(jx - icx) + sq(iy + 1 - icy), sq(jx - icx) + sq(iy - 1 - icy)};
    if (c[0] < r2 && c[1] < r2 && c[2] < r2 && c[3] < r2) return 1. / 9.;
    if (c[0] > r2 && c[1] > r2 && c[2] > r2 && c[3] > r2) return 0;
    const double h = 1. / WRES, sx = (double)(jx - ix) / WRES, sy = 1. / WRES;
    double acc = 0;
    for (int b = 0; b < WRES; b++) {
        const double v = (b + 0.5) * sy;
        for (int a = 0; a < WRES; a++) {
            const double u = (a + 0.5) * sx;
            const double f = (1 - std::fabs(u)) * sq(1 - std::fabs(v)) * std::fabs(u);
            if (sq(ix + u - icx) + sq(iy + v - icy) < r2) acc += f;  // upper half
            if (sq(ix + u - icx) + sq(iy - v - icy) < r2) acc += f;  // lower half
        }
    }
    return acc * h * h;
}
// neighbours along y (same column)
double w_ud(int ix, int iy, int /*jx*/, int jy, double icx, double icy, double icr) {
    const double r2 = icr * icr;
    const double c[4] = {sq(ix - 1 - icx) + sq(iy - icy), sq(ix - 1 - icx) + sq(jy - icy),
                         sq(ix + 1 - icx) + sq(iy - icy), sq(ix + 1 - icx) + sq(jy - icy)};
    if (c[0] < r2 && c[1] < r2 && c[2] < r2 && c[3] < r2) return 1. / 9.;
    if (c[0] > r2 && c[1] > r2 && c[2] > r2 && c[3] > r2) return 0;
    const double h = 1. / WRES, sy = (double)(jy - iy) / WRES, sx = 1. / WRES;
    double acc = 0;
    for (int b = 0; b < WRES; b++) {
        const double v = (b + 0.5) * sy;
        for (int a = 0; a < WRES; a++) {
            const double u = (a + 0.5) * sx;
            const double f = sq(1 - std::fabs(u)) * (1 - std::fabs(v)) * std::fabs(v);
            if (sq(ix + u - icx) + sq(iy + v - icy) < r2) acc += f;  // right half
            if (sq(ix - u - icx) + sq(iy + v - icy) < r2) acc += f;  // left half
        }
    }
    return acc * h * h;
}
// the point with itself
double w_self(int ix, int iy, double icx, double icy, double icr) {
    const double r2 = icr * icr;
    const double c[4] = {sq(ix + 1 - icx) + sq(iy + 1 - icy), sq(ix + 1 - icx) + sq(iy - 1 - icy),
                         sq(ix - 1 - icx) + sq(iy + 1 - icy), sq(ix - 1 - icx) + sq(iy - 1 - icy)};
    if (c[0] < r2 && c[1] < r2 && c[2] < r2 && c[3] < r2) return 4. / 9.;
    if (c[0] > r2 && c[1] > r2 && c[2] > r2 && c[3] > r2) return 0;
    const double h = 1. / WRES;
    double acc = 0;
    for (int b = 0; b < WRES; b++) {
        const double v = (b + 0.5) * h;
        for (int a = 0; a < WRES; a++) {
            const double u = (a + 0.5) * h;
            const double f = sq((1 - u) * (1 - v));
            if (sq(ix - u - icx) + sq(iy - v - icy) < r2) acc += f;
            if (sq(ix + u - icx) + sq(iy - v - icy) < r2) acc += f;
            if (sq(ix - u - icx) + sq(iy + v - icy) < r2) acc += f;
            if (sq(ix + u - icx) + sq(iy + v - icy) < r2) acc += f;
        }
    }
    return acc * h * h;
}

// mkw_circular on a full square grid: the MAOS sparse W0 holds in column i
// the weights w(i -> j) of i's neighbours j (rows), scaled by 1/(pi icr^2);
// W1(i) is the column sum.  Both are renormalised so that sum W1 = 1.
// Returned in row form: W0r[9*j + k] = W0(j, i) for the neighbour i of j at
// offset k = (iy-jy+1)*3 + (ix-jx+1), i.e. out_j = sum_k W0r[9j+k] in_i.
void mkw_circular(const Grid &g, double cx, double cy, double cr, std::vector<double> &W0r,
                  std::vector<double> &W1) {
    const long nx = g.nx, ny = g.ny, n = nx * ny;
    const double idx = 1. / g.dx;
    const double icx = (cx - g.ox) * idx, icy = (cy - g.oy) * idx, icr = cr * idx;
    const double sc = 1. / (M_PI * icr * icr);
    W0r.assign(9 * n, 0.);
    W1.assign(n, 0.);
    for (long iy = 0; iy < ny; iy++)
        for (long ix = 0; ix < nx; ix++) {
            double wsum = 0;
            for (long jy = iy - 1; jy <= iy + 1; jy++)
                for (long jx = ix - 1; jx <= ix + 1; jx++) {
                    if (jx < 0 || jx >= nx || jy < 0 || jy >= ny) continue;
                    double w;
                    const int X = (int)ix, Y = (int)iy, JX = (int)jx, JY = (int)jy;
                    if (jx != ix && jy != iy) w = w_corner(X, Y, JX, JY, icx, icy, icr);
                    else if (jy != iy) w = w_ud(X, Y, JX, JY, icx, icy, icr);
                    else if (jx != ix) w = w_lr(X, Y, JX, JY, icx, icy, icr);
                    else w = w_self(X, Y, icx, icy, icr);
                    if (w > 0) {
                        // W0(j, i) = w: row j, neighbour i at offset i - j
                        const long j = jx + jy * nx;
                        const int k = (int)((iy - jy + 1) * 3 + (ix - jx + 1));
                        W0r[9 * j + k] = w * sc;
                    }
                    wsum += w;
                }
            W1[ix + iy * nx] = wsum * sc;
        }
    double s = 0;
    for (double v : W1) s += v;
    if (std::fabs(s - 1) > 1e-12) {
        const double f = 1. / s;
        for (double &v : W1) v *= f;
        for (double &v : W0r) v *= f;
    }
}

// mkw_annular (lib/mkw.c:443-481): an annulus as the difference of two
// circles, renormalised.
void mkw_annular(const Grid &g, double cri, double cro, std::vector<double> &W0r, std::vector<double> &W1) {
    if (cri < g.dx) {
        mkw_circular(g, 0, 0, cro, W0r, W1);
        return;
    }
    std::vector<double> W0i, W1i;
    mkw_circular(g, 0, 0, cri, W0i, W1i);
    mkw_circular(g, 0, 0, cro, W0r, W1);
    const double idx = 1. / g.dx;
    const double sco = 1. / (M_PI * sq(cro * idx)), sci = 1. / (M_PI * sq(cri * idx));
    // undo the renormalisation of each circle only through their own
    // scale factors, as the reference does (it assumes sum W1 = 1 held)
    double s = 0;
    for (size_t i = 0; i < W1.size(); i++) {
        W1[i] = W1[i] / sco - W1i[i] / sci;
        s += W1[i];
    }
    const double f = 1. / s;
    for (double &v : W1) v *= f;
    for (size_t i = 0; i < W0r.size(); i++) W0r[i] = W0r[i] * (f / sco) - W0i[i] * (f / sci);
}

// One bilinear tap list: for output coordinate u (in input grid units) the
// nodes n and n+1 with weights 1-f and f, dropped where the node is off the
// grid or the weight is not above EPS (mkh, lib/mkh.c:90-120; no
// renormalisation of the remaining weights).
struct Taps {
    long n[2];
    double w[2];
};
Taps taps(double u, long nin) {
    Taps t;
    const long n0 = (long)std::floor(u);
    const double f = u - n0;
    t.n[0] = n0;
    t.n[1] = n0 + 1;
    t.w[0] = 1 - f;
    t.w[1] = f;
    for (int k = 0; k < 2; k++)
        if (t.n[k] < 0 || t.n[k] >= nin) t.w[k] = 0;
    return t;
}

// HA' W1 for one (direction, DM): the coupling sum of genactcpl.
void add_hat_w1(const Grid &ag, const Grid &fg, double dispx, double dispy, double scale, const std::vector<double> &W1,
                double *cpl) {
    for (long py = 0; py < fg.ny; py++) {
        const Taps ty = taps(((fg.oy + py * fg.dy) * scale + dispy - ag.oy) / ag.dy, ag.ny);
        for (long px = 0; px < fg.nx; px++) {
            const Taps tx = taps(((fg.ox + px * fg.dx) * scale + dispx - ag.ox) / ag.dx, ag.nx);
            const double w1 = W1[px + py * fg.nx];
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    const double w = tx.w[a] * ty.w[b];
                    if (w > EPS_SYS) cpl[tx.n[a] + ty.n[b] * ag.nx] += w * w1;
                }
        }
    }
}

void build_fit(mt_fit *F, const mt_setup *s, const mt_fit_parms_t *p) {
    if (!p || p->ndm <= 0 || !p->dm || p->nfit <= 0 || !p->thetax || !p->thetay || !p->wt || !p->hs)
        throw Error(MT_EINVAL, "mt_fit_create: missing parameters");
    if (p->actinterp) throw Error(MT_EINVAL, "fit.actinterp=1 is not supported (HA*actinterp); set fit.actinterp=0");
    const int ndm = p->ndm, nfit = p->nfit;
    for (int i = 0; i < ndm; i++) {
        if (p->dm[i].iac != 0) throw Error(MT_EINVAL, "dm.iac != 0 (cubic influence functions) is not supported");
        if (!(p->dm[i].dx > 0) || p->dm[i].guard < 0) throw Error(MT_EINVAL, "invalid dm.dx or dm.guard");
    }
    // fit directions, weights normalised to sum |wt| = 1 (maos/setup_parms.c:1847)
    double ws = 0;
    for (int f = 0; f < nfit; f++) ws += std::fabs(p->wt[f]);
    if (!(ws > 0)) throw Error(MT_EINVAL, "fit.wt sums to zero");
    for (int f = 0; f < nfit; f++) {
        F->thetax.push_back(p->thetax[f]);
        F->thetay.push_back(p->thetay[f]);
        F->wt.push_back(p->wt[f] / ws);
        F->hs.push_back(p->hs[f]);
    }
    // the beam directions that size the grids: those of the setup, plus the
    // fitting directions if the caller did not list them (MAOS includes
    // them in parms->dirs, maos/setup_parms.c:1945-1984)
    std::vector<Dir> dirs = s->dirs;
    for (int f = 0; f < nfit; f++) {
        bool have = false;
        for (const Dir &d : dirs)
            have |= d.thetax == p->thetax[f] && d.thetay == p->thetay[f] && d.hs == p->hs[f];
        if (!have) dirs.push_back({p->thetax[f], p->thetay[f], p->hs[f]});
    }
    // floc: the square meta-pupil at fit.pos sampling
    const int fpos = p->pos > 0 ? p->pos : s->pos;
    const double dxr = s->atmr_dx / fpos;
    F->fg = create_metapupil(dirs, s->D, 0, dxr, dxr, 0, s->guard * dxr, 0, 0, 0, 1);
    // W0, W1 (setup_recon_floc)
    const double rin = (p->annular_W && s->Din > 0) ? s->Din / 2 : 0;
    mkw_annular(F->fg, rin, s->D / 2, F->W0, F->W1);
    // actuator grids (setup_recon_aloc, square)
    for (int i = 0; i < ndm; i++) {
        const mt_dm_t &d = p->dm[i];
        const long order = (long)std::ceil(s->D / d.dx);
        const double offset = d.offset + (order % 2) * 0.5;
        F->ag.push_back(create_metapupil(dirs, s->D, d.ht, d.dx, d.dx, offset, d.guard * d.dx, 0, 0, 0, 1));
        F->aht.push_back(d.ht);
    }
    const long nact = F->nact();
    // genactcpl (coupling: sum over directions of HA' W1, no direction
    // weight; above max/nfit counts as fully coupled)
    F->actcpl.assign(nact, 0.);
    long aoff = 0;
    std::vector<long> aoffs;
    for (int i = 0; i < ndm; i++) {
        const Grid &ag = F->ag[i];
        aoffs.push_back(aoff);
        double *cpl = F->actcpl.data() + aoff;
        for (int f = 0; f < nfit; f++) {
            const double ht = F->aht[i];
            add_hat_w1(ag, F->fg, F->thetax[f] * ht, F->thetay[f] * ht, 1. - ht / F->hs[f], F->W1, cpl);
        }
        const long na = ag.nx * ag.ny;
        double mx = 0;
        for (long k = 0; k < na; k++) mx = std::max(mx, cpl[k]);
        const double thres = mx / nfit;
        for (long k = 0; k < na; k++) {
            if (cpl[k] > thres || thres == 0) cpl[k] = 1;
            else cpl[k] /= thres;
        }
        aoff += na;
    }
    const double fitscl = 1. / (double)(F->fg.nx * F->fg.ny);
    // low-rank terms (setup_fit_lrt)
    int lrt_tt = p->lrt_tt;
    if (ndm < 2) lrt_tt = 0;  // maos/setup_parms.c:2387-2389
    F->nnw = (p->lrt_piston ? ndm : 0) + (lrt_tt ? 2 * (ndm - 1) : 0);
    F->NW.assign((size_t)nact * F->nnw, 0.);
    {
        int inw = 0;
        if (p->lrt_piston) {
            for (int i = 0; i < ndm; i++) {
                const long na = F->ag[i].nx * F->ag[i].ny;
                double *col = F->NW.data() + (size_t)(inw + i) * nact + aoffs[i];
                for (long k = 0; k < na; k++)
                    if (F->actcpl[aoffs[i] + k] > 0.1) col[k] = fitscl;
            }
            inw += ndm;
        }
        if (lrt_tt) {
            // loc_diam (math/loc.c:525-534) of the first DM's actuators
            const Grid &a0 = F->ag[0];
            double r2max = 0;
            for (long iy = 0; iy < a0.ny; iy++)
                for (long ix = 0; ix < a0.nx; ix++)
                    r2max = std::max(r2max, sq(a0.ox + ix * a0.dx) + sq(a0.oy + iy * a0.dy));
            const double factor = fitscl * 2. / (2 * std::sqrt(r2max));
            for (int i = 1; i < ndm; i++) {
                const Grid &ag = F->ag[i];
                double *cx = F->NW.data() + (size_t)(inw + (i - 1) * 2) * nact + aoffs[i];
                double *cy = cx + nact;
                for (long iy = 0; iy < ag.ny; iy++)
                    for (long ix = 0; ix < ag.nx; ix++) {
                        const long k = ix + iy * ag.nx;
                        if (F->actcpl[aoffs[i] + k] > 0.1) {
                            cx[k] = (ag.ox + ix * ag.dx) * factor;
                            cy[k] = (ag.oy + iy * ag.dy) * factor;
                        }
                    }
            }
            inw += 2 * (ndm - 1);
        }
    }
    // slaving regulariser S = T T' with T's column per slave actuator:
    // its 4-neighbours pulled towards it (more coupled neighbours only, or
    // all with a weak weight if none is better coupled), the actuator itself
    // with minus their sum (lib/slaving.c:103-190).  The reference builds it
    // inside setup_fit_lrt after that function's early return for no
    // low-rank terms (maos/setup_recon_fit.c:119), so without NW there is no
    // slaving either.
    if (p->actslave && F->nnw > 0) {
        const double scl = std::sqrt(fitscl);
        // columns of T as (row, value) lists, per slave actuator (global index)
        std::vector<std::vector<std::pair<long, double>>> cols;
        for (int i = 0; i < ndm; i++) {
            const Grid &ag = F->ag[i];
            const double *cpl = F->actcpl.data() + aoffs[i];
            for (long iy = 0; iy < ag.ny; iy++)
                for (long ix = 0; ix < ag.nx; ix++) {
                    const long k = ix + iy * ag.nx;
                    if (!(cpl[k] < p->actthres)) continue;
                    const long nb[4][2] = {{ix, iy - 1}, {ix - 1, iy}, {ix + 1, iy}, {ix, iy + 1}};
                    const double thres2 = std::max(0.1, cpl[k]);
                    int near_exist = 0, near_active = 0;
                    for (auto &q : nb) {
                        if (q[0] < 0 || q[0] >= ag.nx || q[1] < 0 || q[1] >= ag.ny) continue;
                        near_exist++;
                        if (cpl[q[0] + q[1] * ag.nx] > thres2) near_active++;
                    }
                    if (!near_exist) throw Error(MT_EINVAL, "isolated actuator in slaving");
                    const double value = near_active ? 0 : -scl * 0.1 / near_exist;
                    std::vector<std::pair<long, double>> col;
                    double vsum = 0;
                    for (auto &q : nb) {
                        if (q[0] < 0 || q[0] >= ag.nx || q[1] < 0 || q[1] >= ag.ny) continue;
                        const long kn = q[0] + q[1] * ag.nx;
                        if (!near_active) {
                            col.push_back({aoffs[i] + kn, value});
                            vsum += value;
                        } else if (cpl[kn] > cpl[k]) {
                            const double v = -scl * std::max(0.1, cpl[kn]);
                            col.push_back({aoffs[i] + kn, v});
                            vsum += v;
                        }
                    }
                    col.push_back({aoffs[i] + k, -vsum});
                    cols.push_back(std::move(col));
                }
        }
        if (!cols.empty()) {
            // S = T T' as CSR (rows sorted, duplicates summed)
            std::vector<std::vector<std::pair<long, double>>> rows(nact);
            for (const auto &col : cols)
                for (const auto &a : col)
                    for (const auto &b : col) rows[a.first].push_back({b.first, a.second * b.second});
            F->sl_p.assign(nact + 1, 0);
            for (long r = 0; r < nact; r++) {
                auto &row = rows[r];
                std::sort(row.begin(), row.end(),
                          [](const std::pair<long, double> &u, const std::pair<long, double> &v) { return u.first < v.first; });
                for (size_t k = 0; k < row.size(); k++) {
                    if (k > 0 && row[k].first == row[k - 1].first) {
                        F->sl_x.back() += row[k].second;
                    } else {
                        F->sl_i.push_back(row[k].first);
                        F->sl_x.push_back(row[k].second);
                    }
                }
                F->sl_p[r + 1] = (long)F->sl_i.size();
            }
        }
    }
    // Tikhonov term of the assembled FL.M: tikcr * 4/nact on the diagonal
    F->tikdiag = std::fabs(p->tikcr) > 1e-15 ? p->tikcr * 4. / (double)nact : 0.;
    F->warm = p->warm ? 1 : 0;
    F->maxit = p->maxit > 0 ? p->maxit : (F->warm ? 10 : 100);
    for (const Grid &g : F->ag) {
        F->anx.push_back(g.nx);
        F->any.push_back(g.ny);
        F->aox.push_back(g.ox);
        F->aoy.push_back(g.oy);
        F->adx.push_back(g.dx);
    }
}

template <typename Fn>
int guarded(Fn &&f) {
    try {
        f();
        return MT_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return MT_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return MT_EINVAL;
    }
}

}  // namespace

extern "C" {

int mt_fit_create(const mt_setup_t *s, const mt_fit_parms_t *p, mt_fit_t **out) {
    if (!out || !s) return MT_EINVAL;
    *out = nullptr;
    std::unique_ptr<mt_fit> F(new (std::nothrow) mt_fit());
    if (!F) return MT_ENOMEM;
    int rc = guarded([&] { build_fit(F.get(), s, p); });
    if (rc == MT_OK) *out = F.release();
    return rc;
}

int mt_fit_recon(const mt_fit_t *F, mt_fit_recon_t *r) {
    if (!F || !r) return MT_EINVAL;
    std::memset(r, 0, sizeof(*r));
    r->ndm = (int)F->ag.size();
    r->anx = F->anx.data();
    r->any = F->any.data();
    r->aox = F->aox.data();
    r->aoy = F->aoy.data();
    r->adx = F->adx.data();
    r->aht = F->aht.data();
    r->fnx = F->fg.nx;
    r->fny = F->fg.ny;
    r->fox = F->fg.ox;
    r->foy = F->fg.oy;
    r->fdx = F->fg.dx;
    r->nfit = (int)F->thetax.size();
    r->thetax = F->thetax.data();
    r->thetay = F->thetay.data();
    r->wt = F->wt.data();
    r->hs = F->hs.data();
    r->W0 = F->W0.data();
    r->W1 = F->W1.data();
    r->nnw = F->nnw;
    r->NW = F->nnw ? F->NW.data() : nullptr;
    r->sl_p = F->sl_p.empty() ? nullptr : F->sl_p.data();
    r->sl_i = F->sl_p.empty() ? nullptr : F->sl_i.data();
    r->sl_x = F->sl_p.empty() ? nullptr : F->sl_x.data();
    r->tikdiag = F->tikdiag;
    r->actcpl = F->actcpl.data();
    r->maxit = F->maxit;
    r->warm = F->warm;
    return MT_OK;
}

int mt_fit_sizes(const mt_fit_t *F, long *nact, long *nfloc) {
    if (!F) return MT_EINVAL;
    if (nact) *nact = F->nact();
    if (nfloc) *nfloc = F->fg.nx * F->fg.ny;
    return MT_OK;
}

void mt_fit_free(mt_fit_t *F) { delete F; }

}  // extern "C"
