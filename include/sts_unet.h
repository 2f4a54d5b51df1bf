/* Warm-start U-Net of the SIMPLE-TS time step (no counterpart in the
   reference): a depth-4 U-Net, 3x3 convolutions, BF16 weights and
   activations with FP32 accumulation on the MI355X matrix cores, trained on
   the GPU by its own forward and backward kernels.  It predicts, from the
   state u, v, p, T of time step n and the body mask, the change of p and T
   over the step; sts_time_step_ws uses p + dp, T + dT as the starting guess
   of loop 3 and falls back to the previous step's p, T whenever the
   predicted guess raises the first sweep's residual.

   Tensors crossing this interface are FP32, NHWC (channels fastest), host
   memory.  Parameters are one flat FP32 array: for each convolution in the
   order e1a e1b e2a e2b e3a e3b b_a b_b d3a d3b d2a d2b d1a d1b head, the
   weights W[tap][ci][co] (tap = 3*(dy+1)+(dx+1), 1 tap for the head) then
   the bias b[co]. */
#ifndef STS_UNET_H
#define STS_UNET_H

#include "simple_ts.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sts_unet sts_unet;

/* Network for inputs of cin channels and outputs of cout channels, base
   width `base` (a multiple of 8), batches of up to max_n samples of H x W
   cells (H a multiple of 8, W a multiple of 64).  Parameters start at zero. */
int sts_unet_create(int cin, int cout, int base, int max_n, int H, int W, sts_unet **out);
int sts_unet_destroy(sts_unet *net);
long long sts_unet_param_count(const sts_unet *net);
int sts_unet_set_params(sts_unet *net, const float *host);
int sts_unet_get_params(sts_unet *net, float *host);
/* Gradients of the last backward pass, same layout as the parameters. */
int sts_unet_get_grads(sts_unet *net, float *host);

/* y[n][H][W][cout] = net(x[n][H][W][cin]). */
int sts_unet_forward(sts_unet *net, int n, const float *x, float *y);
/* The same outputs, bit for bit, without keeping what the backward pass
   needs (the decoder reads its concatenated inputs in place). */
int sts_unet_infer(sts_unet *net, int n, const float *x, float *y);
/* After sts_unet_forward of the same batch: gradients (overwritten) of
   L = sum over cells with mask != 0 of |y - t|^2 / (2 * count * cout),
   count = number of such cells; *loss gets L. */
int sts_unet_backward_mse(sts_unet *net, int n, const float *t, const unsigned char *mask, double *loss);
/* Gradients (overwritten) for a given dL/dy[n][H][W][cout]. */
int sts_unet_backward(sts_unet *net, int n, const float *dy);
/* One Adam step on the FP32 master parameters with the current gradients. */
int sts_unet_adam(sts_unet *net, float lr, float beta1, float beta2, float eps);

/* Normalisation between solver fields and network tensors:
   inputs  (u/u_ref, v/u_ref, p/p_ref - 1, T - 1, fluid mask),
   targets (dp/(p_ref dp_scale), dT/dT_scale). */
typedef struct sts_unet_norm {
  double u_ref, p_ref, dp_scale, dT_scale;
} sts_unet_norm;

/* Device-resident training set of `slots` samples of the network's H x W. */
int sts_unet_dataset(sts_unet *net, int slots);
/* Records sample `slot` from a solver context (Ny x Nx = H x W) right after
   a time step: input = the state the step started from (u_pr, v_pr, p_pr,
   T_pr), target = (p - p_pr, T - T_pr), normalised. */
int sts_unet_record(sts_unet *net, int slot, sts_ctx *ctx, const sts_unet_norm *nrm);
/* One training step on the recorded samples slots[0..n-1]: forward, masked
   MSE, backward, Adam.  *loss gets the batch loss before the update. */
int sts_unet_train_step(sts_unet *net, const int *slots, int n, float lr, double *loss);
/* Loss of the current parameters on recorded samples (no update). */
int sts_unet_eval(sts_unet *net, const int *slots, int n, double *loss);

/* The corrector: the same network with cin = 9, cout = 4, refining the
   start that the polynomial extrapolation in time predicts
   (sts_mixed_control.extrapolate).  Input: the state x^n the step starts
   from (u/u_ref, v/u_ref, p/p_ref - 1, T - 1), the predicted start's change
   x_e - x^n in the same units divided by `step`, and the fluid mask.
   Output: (x^(n+1) - x_e) in the same units divided by `corr` -- the part of
   the converged step that the extrapolation misses. */
typedef struct sts_unet_corr_norm {
  double u_ref, p_ref, step, corr;
} sts_unet_corr_norm;

/* The context keeps a copy of each step's start of loop 2 (the predicted
   start, or x^n when the step is not predicted), for recording; on != 0
   allocates it (4 fields), 0 frees it. */
int sts_keep_predicted_start(sts_ctx *ctx, int on);
/* Records corrector sample `slot` right after a time step of a context that
   keeps its predicted start: input from x^n and the kept start, target from
   the converged state. */
int sts_unet_record_correction(sts_unet *net, int slot, sts_ctx *ctx, const sts_unet_corr_norm *nrm);
/* Attaches a corrector (net == NULL detaches): every predicted step of
   sts_time_step_mixed then refines the extrapolated start with the network
   on the loop's update sets, and keeps the refinement only if it lowers the
   first sweep's residual max(max|dp|/tol_p, max|dT|/tol_T) (each start with
   its own assembly); otherwise the step runs from the extrapolated start.
   The network must be of the context's grid (H = Ny, W = Nx). */
int sts_set_corrector(sts_ctx *ctx, sts_unet *net, const sts_unet_corr_norm *nrm);
/* out[0]: predicted steps the corrector refined and kept, out[1]: refinements
   rejected by the residual test. */
int sts_corrector_stats(sts_ctx *ctx, long long out[2]);
/* Back-off of the corrector (no counterpart in the reference): with
   max_pause > 0, a refinement that does not pay makes the corrector sit out
   the next 1, 2, 4, ... predicted steps (at most max_pause; back to 1 after
   one that pays), so a network that does not help costs little; those steps
   run from the extrapolated start as without a corrector.  A refinement
   pays when it is kept and its step then needs at least one outer
   iteration fewer than the running mean (weights 1/2, 1/4, ...) of the
   predicted steps that ran without one; a rejected one, or one whose predicted attempt is abandoned,
   never pays.  0 (the default of sts_set_corrector): every predicted step
   is refined and tested. */
int sts_set_corrector_pause(sts_ctx *ctx, int max_pause);
/* Predicted steps the corrector sat out under sts_set_corrector_pause. */
int sts_corrector_skipped(sts_ctx *ctx, long long *skipped);

/* Time step with the network's warm start: prologue, prediction, choice of
   the starting p, T by the first sweep's residual (the prediction is used
   only if it lowers max(max|dp|/tol_p, max|dT|/tol_T) of that sweep), then
   loop 2 as sts_time_step (mc == NULL) or sts_time_step_mixed.  *used gets 1
   if the prediction was taken. */
int sts_time_step_ws(sts_ctx *ctx, sts_unet *net, const sts_unet_norm *nrm, const sts_loop_control *lc,
                     const sts_mixed_control *mc, int *used, int *iters32, int *iters64, double *res_log,
                     double *p_probe);

#ifdef __cplusplus
}
#endif
#endif
