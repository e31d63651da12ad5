// Reconstruction geometry and the GP average-gradient operator, FP64.
//
// Restates, on the tomo.square=1 path that the GPU reconstructor forces
// (maos/setup_parms.c:3024-3027):
//   nextfftsize         math/mathmisc.c:327-360
//   create_metapupil    math/loc.c:1471-1560
//   loccircle/annular   math/loc.c:815-864
//   saloc + saa         maos/setup_powfs.c:255-404 (setup_powfs_geom)
//   sa_reduce           maos/setup_powfs.c:126-235
//   mkg (mkgt)          lib/mkg.c:113-470, called from
//                       maos/setup_recon_prep.c:481-555 with gloc/gamp from
//                       make_gloc (maos/setup_recon_prep.c:84-105)
#include <algorithm>
#include <climits>
#include <cmath>

#include "setup_internal.hpp"

namespace mt {

long nextfftsize(long n) {
    // math/mathmisc.c:327-360: smallest n3 >= n that factors into at most two
    // distinct radices from {2,3,5,7} and is even.
    if (n == 0) n = 1;
    const int radixs[4] = {2, 3, 5, 7};
    long n2, n3;
    do {
        n2 = n;
        n3 = 1;
        int selected[4] = {0, 0, 0, 0};
        for (int irad = 0; irad < 4;) {
            int radix = radixs[irad];
            long ratio = n2 / radix;
            if (ratio * radix == n2) {
                selected[irad] = 1;
                n2 = ratio;
                n3 *= radix;
            } else {
                irad++;
            }
        }
        int count = selected[0] + selected[1] + selected[2] + selected[3];
        if (count > 2) n2 = n;
        n++;
    } while (n2 > 1 || (n3 & 1) == 1);
    return n3;
}

Grid create_metapupil(const std::vector<Dir> &dirs, double D, double ht, double dx,
                      double dy, double offset, double guard, long ninx, long niny,
                      int pad, int square) {
    // math/loc.c:1471-1537
    const double R = D / 2;
    double minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
    if (dirs.empty()) throw Error(MT_EINVAL, "create_metapupil: no directions");
    for (const Dir &d : dirs) {
        double RR = (1. - ht / d.hs) * R + guard;
        double sx1 = d.thetax * ht - RR, sx2 = d.thetax * ht + RR;
        double sy1 = d.thetay * ht - RR, sy2 = d.thetay * ht + RR;
        minx = std::min(minx, std::min(sx1, sx2));
        maxx = std::max(maxx, std::max(sx1, sx2));
        miny = std::min(miny, std::min(sy1, sy2));
        maxy = std::max(maxy, std::max(sy1, sy2));
    }
    if (square) {
        maxx = std::max(std::fabs(minx), std::fabs(maxx));
        minx = -maxx;
        maxy = std::max(std::fabs(miny), std::fabs(maxy));
        miny = -maxy;
    }
    offset = offset - std::floor(offset);
    {
        double mind = minx / dx;
        double adjust = mind - std::floor(mind) - offset;
        if (adjust < 0) adjust++;
        minx -= adjust * dx;
        mind = miny / dy;
        adjust = mind - std::floor(mind) - offset;
        if (adjust < 0) adjust++;
        miny -= adjust * dy;
    }
    Grid g;
    g.ox = minx;
    g.oy = miny;
    g.dx = dx;
    g.dy = dy;
    long nx = (long)std::ceil((maxx - g.ox) / dx) + 1;
    long ny = (long)std::ceil((maxy - g.oy) / dy) + 1;
    if (square) nx = ny = std::max(nx, ny);
    if (pad) {
        ninx = 1L << (int)std::ceil(std::log2((double)nx));
        niny = 1L << (int)std::ceil(std::log2((double)ny));
    }
    if (ninx > 1) {
        if (ninx < nx) throw Error(MT_EINVAL, "create_metapupil: ninx too small");
        g.ox = g.ox - ((ninx - nx) / 2) * dx;
        nx = ninx;
    }
    if (niny > 1) {
        if (niny < ny) throw Error(MT_EINVAL, "create_metapupil: niny too small");
        g.oy = g.oy - ((niny - ny) / 2) * dy;
        ny = niny;
    }
    g.nx = nx;
    g.ny = ny;
    return g;
}

void loccircle(double *phi, const double *x, const double *y, long n, double dx,
               double dy, double cx, double cy, double r, double val) {
    // math/loc.c:815-852: points well inside get val; points within 1.5 dx of
    // the rim get val times the covered fraction of a 10x10 sub-grid.
    const int nres = 10;
    const double cres = (nres - 1.) / 2.;
    const double res = 1. / nres;
    const double dxres = res * dx, dyres = res * dy, res2 = res * res;
    const double r2 = r * r;
    const double r2l = (r - dx * 1.5) * (r - dy * 1.5);
    const double r2u = (r + dx * 1.5) * (r + dy * 1.5);
    for (long i = 0; i < n; i++) {
        double r2r = (x[i] - cx) * (x[i] - cx) + (y[i] - cy) * (y[i] - cy);
        if (r2r < r2l) {
            phi[i] += val;
        } else if (r2r < r2u) {
            long tot = 0;
            for (int jres = 0; jres < nres; jres++) {
                double iiy = y[i] + (jres - cres) * dyres;
                double rr2y = (iiy - cy) * (iiy - cy);
                for (int ires = 0; ires < nres; ires++) {
                    double iix = x[i] + (ires - cres) * dxres;
                    if ((iix - cx) * (iix - cx) + rr2y <= r2) tot++;
                }
            }
            phi[i] += (double)tot * res2 * val;
        }
    }
}

void locannular(double *phi, const double *x, const double *y, long n, double dx,
                double dy, double cx, double cy, double r, double rin, double val) {
    // math/loc.c:857-864
    loccircle(phi, x, y, n, dx, dy, cx, cy, r, val);
    if (rin > EPS_SYS) loccircle(phi, x, y, n, dx, dy, cx, cy, rin, -val);
}

SaGeom setup_saloc(double D, double Din, const mt_powfs_t &pw) {
    // maos/setup_powfs.c:255-404 for a Shack-Hartmann powfs with no saloc file,
    // no amplitude map file (annular pupil), no misregistration, safill2d=1.
    SaGeom sa;
    const double dsa = pw.dsa;
    sa.dsa = dsa;
    const int order = (int)std::ceil(D / dsa);  // maos/setup_parms.c:1352
    const double areafulli = order > 2 ? 1. : 4. / M_PI;
    const int nx = 2 * (int)std::lround(0.5 * dsa / pw.dx);
    if (nx <= 0) throw Error(MT_EINVAL, "powfs.dx larger than dsa");
    const double dx = dsa / nx;
    const double dxoffset = dx * 0.5;
    const int nxsa = nx * nx;
    // Candidate subapertures: the reference's test ORs the outer and inner
    // radius conditions, which keeps every candidate of the order x order
    // square; sa_reduce below then trims by illuminated area.
    const double offset = (order & 1) ? -0.5 : 0.0;
    const double r2max = std::pow(order * 0.5, 2);
    const double r2min = dsa < Din ? std::pow(Din / dsa / 2, 2) : -1;
    std::vector<double> cx, cy;
    for (int j = -order / 2; j <= (order - 1) / 2; j++) {
        for (int i = -order / 2; i <= (order - 1) / 2; i++) {
            double xc = (double)i + offset, yc = (double)j + offset;
            double r1 = std::pow(xc + 1, 2) + std::pow(yc + 1, 2);
            double r2 = std::pow(xc + 1, 2) + std::pow(yc, 2);
            double r3 = std::pow(xc, 2) + std::pow(yc + 1, 2);
            double r4 = std::pow(xc, 2) + std::pow(yc, 2);
            if (r1 < r2max || r2 < r2max || r3 < r2max || r4 < r2max || r1 > r2min ||
                r2 > r2min || r3 > r2min || r4 > r2min) {
                cx.push_back(xc * dsa);
                cy.push_back(yc * dsa);
            }
        }
    }
    const long nsa0 = (long)cx.size();
    // Amplitude on the nx*nx OPD points of each subaperture (pts2loc order:
    // math/loc.c:1110-1132), annular pupil (maos/utils.c:446-454), then
    // saa = sum(amp^2)/nxsa (maos/setup_powfs.c:92-105).
    std::vector<double> saa(nsa0, 0.);
    {
        std::vector<double> px(nxsa), py(nxsa), amp(nxsa);
        for (long isa = 0; isa < nsa0; isa++) {
            const double ox = cx[isa] + dxoffset, oy = cy[isa] + dxoffset;
            for (int jj = 0; jj < nx; jj++) {
                for (int ii = 0; ii < nx; ii++) {
                    px[jj * nx + ii] = ox + (double)ii * dx;
                    py[jj * nx + ii] = oy + (double)jj * dx;
                }
            }
            std::fill(amp.begin(), amp.end(), 0.);
            locannular(amp.data(), px.data(), py.data(), nxsa, dx, dx, 0, 0, D * 0.5,
                       Din * 0.5, 1);
            double area = 0;
            for (int k = 0; k < nxsa; k++) area += amp[k] * amp[k];
            saa[isa] = area / (double)nxsa;
            if (std::fabs(areafulli - 1) > EPS_SYS) saa[isa] *= areafulli;
        }
    }
    const double thresarea = pw.saat;
    // sa_reduce (maos/setup_powfs.c:126-235): disable valid subapertures that
    // have two or fewer valid neighbours, repeated until stable (in-place
    // sweep in subaperture order), then keep saa >= saat.
    if (nsa0 > 4) {
        // ptsloc map: subaperture origins on a dsa grid.
        double xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY;
        for (long i = 0; i < nsa0; i++) {
            double x = cx[i] + dxoffset, y = cy[i] + dxoffset;
            xmin = std::min(xmin, x);
            xmax = std::max(xmax, x);
            ymin = std::min(ymin, y);
            ymax = std::max(ymax, y);
        }
        const double d1 = 1. / dsa;
        const long mnx = (long)std::lround((xmax - xmin) * d1) + 1;
        const long mny = (long)std::lround((ymax - ymin) * d1) + 1;
        std::vector<long> map(mnx * mny, 0);
        std::vector<long> ixs(nsa0), iys(nsa0);
        for (long i = 0; i < nsa0; i++) {
            long ix = std::lround((cx[i] + dxoffset - xmin) * d1);
            long iy = std::lround((cy[i] + dxoffset - ymin) * d1);
            ixs[i] = ix;
            iys[i] = iy;
            map[ix + iy * mnx] = i + 1;
        }
        int changed;
        do {
            changed = 0;
            for (long isa = 0; isa < nsa0; isa++) {
                const long ix = ixs[isa], iy = iys[isa];
                int nedge = 0, ncorner = 0;
                const int nself = saa[isa] >= thresarea;
                for (int jx = -1; jx < 2; jx++) {
                    for (int jy = -1; jy < 2; jy++) {
                        if (jx == 0 && jy == 0) continue;
                        long kx = ix + jx, ky = iy + jy;
                        long jsa = 0;
                        if (kx >= 0 && kx < mnx && ky >= 0 && ky < mny) jsa = map[kx + ky * mnx];
                        if (jsa && saa[jsa - 1] >= thresarea) {
                            if (std::abs(jx + jy) == 1) nedge++;
                            else ncorner++;
                        }
                    }
                }
                if (nself && ncorner + nedge <= 2) {
                    saa[isa] = 0;
                    changed++;
                }
            }
        } while (changed);
    }
    for (long isa = 0; isa < nsa0; isa++) {
        if (saa[isa] >= thresarea) {
            sa.locx.push_back(cx[isa]);
            sa.locy.push_back(cy[isa]);
            sa.saa.push_back(saa[isa]);
        }
    }
    if (sa.locx.empty()) throw Error(MT_EINVAL, "no subapertures above threshold");
    return sa;
}

namespace {
const double MKG_TOL = 5.e-16;  // lib/mkg.c:52
inline int mkg_ceil(double a) {  // lib/mkg.c:49
    return std::fabs(a - std::round(a)) < MKG_TOL ? (int)std::lround(a) : (int)std::ceil(a);
}
inline int mkg_floor(double a) {  // lib/mkg.c:50
    return std::fabs(a - std::round(a)) < MKG_TOL ? (int)std::lround(a) : (int)std::floor(a);
}
inline void split(double a, double &frac, long &ip) {  // math/numtype.h:60
    ip = (long)std::floor(a);
    frac = a - (double)ip;
}
}  // namespace

SparseRows mkg_gp(const Grid &pg, const SaGeom &sa, double D, double Din, double gdx) {
    // lib/mkg.c:113-470 (mkgt) with xloc = recon ploc (square grid pg),
    // ploc = gloc (annular loc at the WFS sampling gdx, amplitude gamp),
    // scale = 1, no displacement, do_partial = 1, then transposed (mkg).
    //
    // gloc = mkannloc(D+1, 0, gdx, 0) (math/loc.c:511-521) lies on the lattice
    // x = gox + i*gdx, gox = -(ngx/2)*gdx; its bounding box is the whole
    // ngx x ngx map (the circle of diameter D+1 touches the map edges).
    // Points of that lattice outside the circle carry zero amplitude here,
    // which has the same effect in mkg as absent points: every use of a
    // gloc point is gated by nonzero amplitude of the point itself.
    const long ngx = (long)((D + 1) / gdx + 1);
    const double gox = -(double)(ngx / 2) * gdx;
    const double dp1 = 1. / gdx;          // 1/ploc->dx in mkg naming
    const double dx1 = 1. / pg.dx;        // 1/xloc->dx
    const double dx2 = dx1;               // scale*dx1, scale = 1
    const double dispx = (0. - pg.ox) * dx1;
    const double dispy = (0. - pg.oy) * dx1;
    const double dp2 = dp1;               // same == 0 branch
    const double dsa = sa.dsa;
    const double dsa2 = dsa / 2. * dp2;
    const double amp_thres = dsa2 * dsa2 * 4 * 0.01;
    const double poff0 = -gox * dp1 + dsa2;
    const double poff1 = -gox * dp1 + dsa2;
    const long nsa = (long)sa.locx.size();

    // Amplitude on the gloc lattice region that the subapertures touch
    // (+-2 pixels).  gamp = mkamp(gloc, NULL, 0, 0, D, Din)
    // (maos/utils.c:446-454).
    long bx0 = LONG_MAX, by0 = LONG_MAX, bx1 = LONG_MIN, by1 = LONG_MIN;
    for (long isa = 0; isa < nsa; isa++) {
        double scx = sa.locx[isa] * dp2 + poff0, scy = sa.locy[isa] * dp2 + poff1;
        bx0 = std::min(bx0, (long)std::floor(scx - dsa2) - 3);
        by0 = std::min(by0, (long)std::floor(scy - dsa2) - 3);
        bx1 = std::max(bx1, (long)std::ceil(scx + dsa2) + 3);
        by1 = std::max(by1, (long)std::ceil(scy + dsa2) + 3);
    }
    bx0 = std::max(bx0, 0L);
    by0 = std::max(by0, 0L);
    bx1 = std::min(bx1, ngx - 1);
    by1 = std::min(by1, ngx - 1);
    const long anx = bx1 - bx0 + 1, any = by1 - by0 + 1;
    std::vector<double> amp(anx * any, 0.);
    {
        std::vector<double> px(anx), py(anx);
        for (long iy = 0; iy < any; iy++) {
            for (long ix = 0; ix < anx; ix++) {
                px[ix] = gox + (double)(ix + bx0) * gdx;
                py[ix] = gox + (double)(iy + by0) * gdx;
            }
            locannular(&amp[iy * anx], px.data(), py.data(), anx, gdx, gdx, 0, 0, D * 0.5,
                       Din * 0.5, 1);
        }
    }
    // ampcopy: scale each gloc point by the bilinear weight that lands on
    // existing recon-ploc points (lib/mkg.c:211-232).
    auto ploc_exists = [&](long ix, long iy) {
        return ix >= 0 && ix < pg.nx && iy >= 0 && iy < pg.ny;
    };
    for (long iy = 0; iy < any; iy++) {
        for (long ix = 0; ix < anx; ix++) {
            double &a = amp[iy * anx + ix];
            if (a == 0) continue;
            double plocx = (gox + (double)(ix + bx0) * gdx) * dx2 + dispx;
            double plocy = (gox + (double)(iy + by0) * gdx) * dx2 + dispy;
            double fx, fy;
            long nx_, ny_;
            split(plocx, fx, nx_);
            split(plocy, fy, ny_);
            double wtsum = 0;
            if (ploc_exists(nx_, ny_)) wtsum += (1 - fx) * (1 - fy);
            if (ploc_exists(nx_ + 1, ny_)) wtsum += fx * (1 - fy);
            if (ploc_exists(nx_, ny_ + 1)) wtsum += (1 - fx) * fy;
            if (ploc_exists(nx_ + 1, ny_ + 1)) wtsum += fx * fy;
            a *= wtsum;
        }
    }
    auto ampat = [&](long gx, long gy) -> double {
        // loc_map_get on the gloc map; 0 outside (absent point)
        if (gx < 0 || gx >= ngx || gy < 0 || gy >= ngx) return 0.;
        long ix = gx - bx0, iy = gy - by0;
        if (ix < 0 || ix >= anx || iy < 0 || iy >= any) return 0.;
        return amp[iy * anx + ix];
    };

    // Accumulate the transposed operator column by column (subaperture).
    std::vector<std::vector<long>> colI[2];
    std::vector<std::vector<double>> colX[2];
    for (int iw = 0; iw < 2; iw++) {
        colI[iw].resize(nsa);
        colX[iw].resize(nsa);
    }
    double amp2[3][3];
    double wtalpha[4], wtbeta[4];
    for (long isa = 0; isa < nsa; isa++) {
        const double scx = sa.locx[isa] * dp2 + poff0;
        const double scy = sa.locy[isa] * dp2 + poff1;
        double ampsum = 0;
        const double limx1 = scx - dsa2, limx2 = scx + dsa2;
        const double limy1 = scy - dsa2, limy2 = scy + dsa2;
        for (int limy = mkg_ceil(limy1) - 1; limy < mkg_floor(limy2) + 2; limy++) {
            for (int limx = mkg_ceil(limx1) - 1; limx < mkg_floor(limx2) + 2; limx++) {
                if (limx < 0 || limx >= ngx || limy < 0 || limy >= ngx) continue;
                double weight[2] = {0, 0};
                for (int jt = 0; jt < 3; jt++)
                    for (int it = 0; it < 3; it++) amp2[jt][it] = ampat(limx + it - 1, limy + jt - 1);
                double alpha1, alpha2, beta1, beta2, signx, signy;
                int indx[2], indy[2];
                // lib/mkg.c:58-92 (DO_CALC with do_partial)
                auto do_calc = [&]() {
                    alpha1 = alpha1 < 0. ? 0. : alpha1;
                    alpha2 = alpha2 > 1. ? 1. : alpha2;
                    beta1 = beta1 < 0. ? 0. : beta1;
                    beta2 = beta2 > 1. ? 1. : beta2;
                    const int valid = !(alpha1 > 1. - MKG_TOL || alpha2 < -MKG_TOL ||
                                        beta1 > 1. - MKG_TOL || beta2 < -MKG_TOL);
                    const double a1 = alpha1, a2 = alpha2, b1 = beta1, b2 = beta2;
                    wtalpha[0] = (a2 * a2 - a1 * a1) * 0.5 - (a2 * a2 * a2 - a1 * a1 * a1) / 3;
                    wtalpha[1] = ((1 - a1) * (1 - a1) * (1 - a1) - (1 - a2) * (1 - a2) * (1 - a2)) / 3;
                    wtalpha[2] = (a2 * a2 - a1 * a1) * 0.5;
                    wtalpha[3] = (a2 - a1) - (a2 * a2 - a1 * a1) * 0.5;
                    wtbeta[0] = (b2 * b2 - b1 * b1) * 0.5 - (b2 * b2 * b2 - b1 * b1 * b1) / 3;
                    wtbeta[1] = ((1 - b1) * (1 - b1) * (1 - b1) - (1 - b2) * (1 - b2) * (1 - b2)) / 3;
                    wtbeta[2] = (b2 * b2 - b1 * b1) * 0.5;
                    wtbeta[3] = (b2 - b1) - (b2 * b2 - b1 * b1) * 0.5;
                    const double a00 = amp2[indy[0]][indx[0]], a01 = amp2[indy[0]][indx[1]];
                    const double a10 = amp2[indy[1]][indx[0]], a11 = amp2[indy[1]][indx[1]];
                    if (valid && std::fabs(a00) > MKG_TOL && std::fabs(a01) > MKG_TOL &&
                        std::fabs(a10) > MKG_TOL && std::fabs(a11) > MKG_TOL) {
                        weight[0] += ((wtbeta[0] * a00 + wtbeta[1] * a10) * wtalpha[2] +
                                      (wtbeta[0] * a01 + wtbeta[1] * a11) * wtalpha[3]) *
                                     dp1 * signx;
                        weight[1] += ((wtbeta[2] * a00 + wtbeta[3] * a10) * wtalpha[0] +
                                      (wtbeta[2] * a01 + wtbeta[3] * a11) * wtalpha[1]) *
                                     dp1 * signy;
                        ampsum += wtalpha[3] * wtbeta[3] * amp2[1][1];
                    }
                };
                // left quadrants use 0 1
                alpha1 = limx - limx2;
                alpha2 = limx - limx1;
                indx[0] = 0; indx[1] = 1; signx = 1.;
                beta1 = limy - limy2; beta2 = limy - limy1;
                indy[0] = 0; indy[1] = 1; signy = 1.;
                do_calc();
                beta1 = limy1 - limy; beta2 = limy2 - limy;
                indy[0] = 2; indy[1] = 1; signy = -1.;
                do_calc();
                // right quadrants use 2 1
                alpha1 = limx1 - limx;
                alpha2 = limx2 - limx;
                indx[0] = 2; indx[1] = 1; signx = -1.;
                beta1 = limy - limy2; beta2 = limy - limy1;
                indy[0] = 0; indy[1] = 1; signy = 1.;
                do_calc();
                beta1 = limy1 - limy; beta2 = limy2 - limy;
                indy[0] = 2; indy[1] = 1; signy = -1.;
                do_calc();
                for (int iw = 0; iw < 2; iw++) {
                    if (!(std::fabs(weight[iw]) > MKG_TOL)) continue;
                    // map the gloc point onto recon ploc (lib/mkg.c:377-426)
                    double plocx = (gox + (double)limx * gdx) * dx2 + dispx;
                    double plocy = (gox + (double)limy * gdx) * dx2 + dispy;
                    double fx, fy;
                    long nx_, ny_;
                    split(plocx, fx, nx_);
                    split(plocy, fy, ny_);
                    double wtsum = 0;
                    if (ploc_exists(nx_, ny_)) wtsum += (1 - fx) * (1 - fy);
                    if (ploc_exists(nx_ + 1, ny_)) wtsum += fx * (1 - fy);
                    if (ploc_exists(nx_, ny_ + 1)) wtsum += (1 - fx) * fy;
                    if (ploc_exists(nx_ + 1, ny_ + 1)) wtsum += fx * fy;
                    wtsum = 1. / wtsum;  // scale == 1
                    auto interp = [&](long ax, double B, long cy_, double Dd) {
                        if (!ploc_exists(ax, cy_)) return;
                        long iphi = ax + cy_ * pg.nx;
                        double wt0 = weight[iw] * B * Dd * wtsum;
                        if (!(std::fabs(wt0) > MKG_TOL)) return;
                        std::vector<long> &ci = colI[iw][isa];
                        std::vector<double> &cxv = colX[iw][isa];
                        size_t k = 0;
                        for (; k < ci.size(); k++) {
                            if (ci[k] == iphi) {
                                cxv[k] += wt0;
                                break;
                            }
                        }
                        if (k == ci.size()) {
                            ci.push_back(iphi);
                            cxv.push_back(wt0);
                        }
                    };
                    interp(nx_, 1 - fx, ny_, 1 - fy);
                    interp(1 + nx_, fx, ny_, 1 - fy);
                    interp(nx_, 1 - fx, 1 + ny_, fy);
                    interp(1 + nx_, fx, 1 + ny_, fy);
                }
            }
        }
        ampsum = ampsum > amp_thres ? 1. / ampsum : 0.;
        for (int iw = 0; iw < 2; iw++)
            for (double &v : colX[iw][isa]) v *= ampsum;
    }
    // dspdroptol(GS0, 1e-14): keep |v| > max(v)*1e-14 (math/sp.c:836-843).
    double maxv = -INFINITY;
    for (int iw = 0; iw < 2; iw++)
        for (long isa = 0; isa < nsa; isa++)
            for (double v : colX[iw][isa]) maxv = std::max(maxv, v);
    const double tol = maxv * 1e-14;
    // GP = transpose: row r = iw*nsa+isa holds that subaperture's column,
    // sorted by ploc index.
    SparseRows gp;
    gp.nrow = 2 * nsa;
    gp.ncol = pg.nx * pg.ny;
    gp.p.push_back(0);
    for (int iw = 0; iw < 2; iw++) {
        for (long isa = 0; isa < nsa; isa++) {
            std::vector<std::pair<long, double>> ent;
            for (size_t k = 0; k < colI[iw][isa].size(); k++) {
                if (std::fabs(colX[iw][isa][k]) > tol)
                    ent.emplace_back(colI[iw][isa][k], colX[iw][isa][k]);
            }
            std::sort(ent.begin(), ent.end());
            for (auto &e : ent) {
                gp.i.push_back(e.first);
                gp.x.push_back(e.second);
            }
            gp.p.push_back((long)gp.i.size());
        }
    }
    return gp;
}

double laplacian_coef(double r0, double weight, double dx) {
    // lib/laplacian.c:26-32
    return (2. * M_PI / 0.5e-6) * std::sqrt(std::pow(r0 / dx, 5. / 3.) / 3.44 / weight) *
           std::pow(2. - std::pow(2, -1. / 3.) - std::pow(2, -1. / 6.), -1. / 2.);
}

}  // namespace mt
