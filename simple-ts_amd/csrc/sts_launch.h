// Launch wrappers of the kernels in sts_kernels.hip (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "sts_device.h"

hipError_t launch_assemble(const KGrid &g, const KFields &f, hipStream_t s, Gate gate = Gate{nullptr, 0});
hipError_t launch_sweep(const KGrid &g, const KFields &f, const double *Po, const double *To, double *Pn, double *Tn,
                        unsigned long long *rp, unsigned long long *rT, hipStream_t s,
                        unsigned long long *rp2 = nullptr, unsigned long long *rT2 = nullptr,
                        Gate gate = Gate{nullptr, 0});
hipError_t launch_sweep2(const KGrid &g, const KFields &f, const double *Po, const double *To, double *Pn, double *Tn,
                         unsigned long long *r1p, unsigned long long *r1T, unsigned long long *r2p,
                         unsigned long long *r2T, hipStream_t s, Gate gate = Gate{nullptr, 0});
hipError_t launch_correct(const KGrid &g, const KFields &f, unsigned long long *ru, unsigned long long *rv,
                          int do_correct, hipStream_t s, Gate gate = Gate{nullptr, 0});
hipError_t launch_properties(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_rho_mid(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_wall(const KGrid &g, const KFields &f, int parts, hipStream_t s, Gate gate = Gate{nullptr, 0});
hipError_t launch_edges(const KGrid &g, const KFields &f, int parts, double *probe, hipStream_t s,
                        Gate gate = Gate{nullptr, 0}, LoopTest lt = LoopTest{nullptr, nullptr, nullptr, 0, 0, 0, 0});
hipError_t launch_flux_diff(const KGrid &g, const KFields &f, hipStream_t s, Gate gate = Gate{nullptr, 0});
hipError_t launch_loop_open(LoopCtrl *ctrl, unsigned long long *log, int n, unsigned long long *lred,
                            hipStream_t s);
hipError_t launch_fold(unsigned long long *slots, int n, unsigned long long *out, hipStream_t s);
hipError_t launch_apply_bodies(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_jump_overwrite(const KGrid &g, const KFields &f, hipStream_t s);
hipError_t launch_copy5(const double *const *src, double *const *dst, int n, hipStream_t s);
hipError_t launch_set_inflow(const KGrid &g, const KFields &f, double u_in, hipStream_t s);
int continuity_blocks(const KGrid &g);
hipError_t launch_continuity(const KGrid &g, const KFields &f, double *cs, double *cf, double *partial,
                             hipStream_t s);
