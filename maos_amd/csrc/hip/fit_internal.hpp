// Internal entry of the DM fitting plan for other parts of the library
// (MVM assembly): enqueue the fits of nsys systems whose tomography layers
// are d_x (FP32, stride xstride per system) into d_dm (FP64, nact per
// system), cold start, on stream st.  Returns an MT_* status and sets the
// error message; nsys may exceed the plan's batch (sub-batches in order).
#pragma once
#include <hip/hip_runtime.h>

#include "maos_tomo.h"

int fit_enqueue_cold(gpu_fit_t *f, const float *d_x, long xstride, double *d_dm, int nsys, hipStream_t st);
long fit_nact(const gpu_fit_t *f);
long fit_nxtot(const gpu_fit_t *f);
int fit_device(const gpu_fit_t *f);
