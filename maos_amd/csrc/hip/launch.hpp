// Launch wrappers of the solve kernels (kernels.hip).  Every launch is a
// 1-D grid over a workgroup table (LaunchTab): one plan alone or the systems
// of several plans of different geometry in one launch.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mtk {
enum { ROW_MODE_INIT_COLD = 0, ROW_MODE_INIT_WARM = 1, ROW_MODE_UPDATE = 2, ROW_MODE_LAST = 3 };
enum { IROW_MODE_Z = 0, IROW_MODE_FIRST = 1, IROW_MODE_UPDATE = 2 };

struct LaunchTab {
    const Job *jobs;   // device
    const WGI *wgi;    // device, nwg entries
    int nwg;
};

void launch_ttr(const LaunchTab &t, hipStream_t st);
// rt64: the residual-side arrays (q, Ap, b, r) are double instead of float.
void launch_wfs(const LaunchTab &t, bool rhs, int matched, int pos, int cs, int rt64, int xsel, hipStream_t st);
void launch_layer(const LaunchTab &t, bool lhs, int matched, int cs, int rt64, int psel, int osel, int partsel,
                  hipStream_t st);
size_t row_lds(int fnx, int f64);
// spec64: the forward row/column spectrum is double2 (requires rt64).
// rp: 2 or 3 for the row FFT in registers (rows of 32 rp points, radices up
// to 4), else 0 (LDS stages).
void launch_row(const LaunchTab &t, int mode, int rt64, int spec64, int rm, int rp, int iter, int xsel, size_t lds,
                hipStream_t st);
size_t col_lds(int maxcol, int npsr, int fny, int f64, int cs);
void launch_col(const LaunchTab &t, int spec64, int cs, int bs, int rm, size_t lds, hipStream_t st);
// Raise the dynamic-LDS limit of the column kernel instantiation that a
// launch of more than 64 KB will use (before any launch or capture).
void col_allow_lds(int spec64, int cs, int bs, int rm, size_t lds);
// Column stage with the column FFTs in registers (blocks of one frequency
// across the layers): a task is one (system, column) on bs x L threads,
// P = ny / L points per lane.  colr_supported: an instantiation exists for P
// and the block size.  colr_shape: threads per workgroup (tasks per
// workgroup in *tasks); colr_lds: its dynamic LDS bytes.  Every job of a
// launch has the same P, L (shL), block size and threads per workgroup.
bool colr_supported(int P, int bs);
int colr_shape(int bs, int L, int *tasks);
size_t colr_lds(int bs, int L, int P, int f64);
void launch_colr(const LaunchTab &t, int spec64, int P, int L, int bs, int nthr, size_t lds, hipStream_t st);
void launch_irow(const LaunchTab &t, int mode, int rm, int rp, int slot, int osel, size_t lds, hipStream_t st);
// LHS WFS pass with TT removal (ptt_lhs): gradients (tg: the WFS tile
// grid), their TT projection (tt: one workgroup per WFS and system), then
// W (g - TT tt) and GP' into q (tg again).
void launch_wfs_lhs_tt(const LaunchTab &tg, const LaunchTab &tt, int pos, int rt64, int xsel, hipStream_t st);
// Plain CG vector kernels over the row blocks (tomo.precond=0).
enum { CGV_MODE_INIT_COLD = 0, CGV_MODE_INIT_WARM = 1, CGV_MODE_UPDATE = 2, CGV_MODE_LAST = 3 };
enum { CGP_MODE_FIRST = 0, CGP_MODE_UPDATE = 1 };
void launch_cgv(const LaunchTab &t, int mode, int rt64, int iter, hipStream_t st);
void launch_cgp(const LaunchTab &t, int mode, int rt64, int slot, hipStream_t st);
void launch_final(const LaunchTab &t, hipStream_t st);
void launch_zero_stop(const LaunchTab &t, hipStream_t st);
void launch_copy(const float *a, float *b, long n, hipStream_t st);
// n unit vectors of len values at v (stride len), 1 at idx[k]; zeros if idx is null
void launch_unit(float *v, long len, const int *idx, int n, hipStream_t st);
// Packed staging of the jobs that are on: gradients from src into each
// plan (in), solutions and residuals from each plan into dx, dres (out).
void launch_stage_in(const Job *jobs, int njobs, const StageOff *off, const float *src, hipStream_t st);
void launch_stage_out(const Job *jobs, int njobs, const StageOff *off, float *dx, double *dres, hipStream_t st);
}  // namespace mtk
