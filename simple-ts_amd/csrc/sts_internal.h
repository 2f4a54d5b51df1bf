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
  hipStream_t stream;
};

int sts_internal_view(sts_ctx *c, StsView *v);
void sts_internal_set_error(const std::string &msg);
// Network prediction of the step's p, T into pw, Tw (device, FP64).
int sts_internal_unet_predict(sts_unet *u, const StsView &v, const sts_unet_norm *nrm, double *pw, double *Tw);
