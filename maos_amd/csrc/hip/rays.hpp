// Ray tables of HX on a square ploc (shared by the FP32 plan and the FP64
// solve path).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace mtk {

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

inline void ray_axes(double mdx, double mox, double moy, long nxin, long nyin, long pnx, long pny,
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

}  // namespace mtk
