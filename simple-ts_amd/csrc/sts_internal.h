// Library-internal interfaces between the solver context (sts_api.cpp) and
// the warm-start network (unet_api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "simple_ts.h"
#include "sts_unet.h"

// Device view of a solver context.
struct StsView {
  int Nx, Ny;
  const double *f[STS_NFIELDS];
  const uint8_t *mask;
  long long fluid_cells;
  int box_o[4];  // p/T update box: Nx_o0, Nx_o1, Ny_o0, Ny_o1
  int box_u[4], box_v[4];
  int px;
  const double *xe[4];  // the last step's predicted start u, v, p, T (null unless kept)
  hipStream_t stream;
};

int sts_internal_view(sts_ctx *c, StsView *v);
void sts_internal_set_error(const std::string &msg);
// Network prediction of the step's p, T into pw, Tw (device, FP64).
int sts_internal_unet_predict(sts_unet *u, const StsView &v, const sts_unet_norm *nrm, double *pw, double *Tw);
// Corrector: from x^n (the *_PR fields) and the predicted start v.xe, the
// corrected start written into out[4] (device u, v, p, T; may be the state
// arrays themselves).
int sts_internal_unet_correct(sts_unet *u, const StsView &v, const sts_unet_corr_norm *nrm, double *const out[4]);
