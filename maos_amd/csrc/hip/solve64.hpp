// Internal interface of the all-FP64 solve (precision mode 4 of a plan):
// every vector, operator, transform and reduction in FP64, the GPU
// counterpart of the reference's CUDA_DOUBLE build (cuda/common.h:39-45).
// Functions return an MT_* status and set the library's error message.
#pragma once
#include <hip/hip_runtime.h>

#include "maos_tomo.h"

struct tomo64;
struct recon64;

// Host copy of the fields of a recon the FP64 plan is built from (the plan
// is made when a caller asks for mode 4, after the caller's recon may be
// gone); the Cn2 update refreshes its L2 coefficients and FDPCG blocks.
int t64_keep(const mt_recon_t *recon, recon64 **out);
int t64_keep_cn2(recon64 *c, const mt_recon_t *recon);
const mt_recon_t *t64_view(const recon64 *c);
void t64_keep_free(recon64 *c);

// Plan of nbatch systems of recon's geometry on `device` (current).
int t64_create(const mt_recon_t *recon, int nbatch, tomo64 **out);
void t64_free(tomo64 *t);
// New L2 coefficients and FDPCG blocks (Cn2 update), same geometry.
int t64_update_cn2(tomo64 *t, const mt_recon_t *recon);
// Warm-start state: the nbatch FP32 solutions d_x (stride nxtot) into the
// FP64 state, or the FP64 state back into d_x (rounded) and the last
// residuals into d_res (nbatch doubles, may be null).
int t64_load_x(tomo64 *t, const float *d_x, hipStream_t st);
int t64_store_x(tomo64 *t, float *d_x, double *d_res, hipStream_t st);
// Zero the warm-start state.
int t64_reset(tomo64 *t);
// Solve systems [s0, s0 + nsys) from host gradients (caller layout: all WFS,
// skipped ones empty; nsys stacked, ng each) into host x (nsys stacked);
// res: nsys residuals sqrt(r'r / b'b) (NaN where the step size was not
// finite).  warm: start from the slot's resident solution.
int t64_solve(tomo64 *t, int s0, int nsys, const double *grad, double *x, double *res, int maxit, int warm,
              hipStream_t st);
// Operator applies on system slot 0 (host buffers): L, R, FDPCG^-1.
int t64_L(tomo64 *t, const double *xin, double *xout, hipStream_t st);
int t64_R(tomo64 *t, const double *grad, double *xout, hipStream_t st);
int t64_P(tomo64 *t, const double *xin, double *xout, hipStream_t st);
// Resident solution / last residual of slot isys.
int t64_get_x(tomo64 *t, int isys, double *x, double *res, hipStream_t st);
int t64_residuals(tomo64 *t, double *res, int n, hipStream_t st);
