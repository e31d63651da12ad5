// Launch wrappers of the solve kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mtk {
enum { ROW_MODE_INIT_COLD = 0, ROW_MODE_INIT_WARM = 1, ROW_MODE_UPDATE = 2, ROW_MODE_LAST = 3 };
void launch_ttr(const Geom &G, int nsys, const float *grad, float *tt, hipStream_t st);
void launch_wfs(const Geom &G, int nsys, bool rhs, int rt64, const float *x, const float *grad,
                const float *tt, void *q, hipStream_t st);
int layer_tiles(const Geom &G);
// rt64: the residual-side vectors (Ap/b out, r) are double instead of float.
void launch_layer(const Geom &G, int nsys, bool lhs, int rt64, const void *q, const float *p, void *out,
                  double *part, int partstride, hipStream_t st);
size_t row_lds(const Geom &G, int f64);
// spec64: the forward row/column spectrum is double2 (requires rt64).
void launch_row(const Geom &G, const Scal &SC, int nsys, int mode, int rt64, int spec64, int iter, float *x,
                void *r, const void *b, const float *p, const void *Ap, void *spec, int *stop, hipStream_t st);
size_t col_lds(const Geom &G, int f64, int cs);
int col_systems(const Geom &G, int nsys, int f64);
void launch_col(const Geom &G, const Scal &SC, int nsys, int f64, const void *specf, float2 *speci,
                hipStream_t st);
enum { IROW_MODE_Z = 0, IROW_MODE_FIRST = 1, IROW_MODE_UPDATE = 2 };
void launch_irow(const Geom &G, const Scal &SC, int nsys, int mode, int slot, const float2 *spec, float *x,
                 float *out, hipStream_t st);
void launch_final(const Geom &G, const Scal &SC, int nsys, int npart_rr, const int *stop, hipStream_t st);
void launch_copy(const float *a, float *b, long n, hipStream_t st);
void launch_zero_i(int *a, int n, hipStream_t st);
}  // namespace mtk
