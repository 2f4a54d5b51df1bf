/*
 * simple_ts.h -- C ABI of the MI355X SIMPLE-TS time-step library
 * (libsimple_ts_hip.so, HIP/CDNA4, gfx950).
 *
 * The reference (kshterev/SIMPLE-TS) has no operator or FFI API: its time
 * step is inline code in main() of SIMPLE_TS_MPI.cpp plus void(void) helpers
 * on file-scope globals.  Each entry point below replaces one of those call
 * sites; the citation "S:n" is the line in SIMPLE_TS_MPI.cpp.  Arguments are
 * plain pointers and sizes; device memory is owned by the context.
 *
 * Layout of every field: double[Nx*Ny], node = i + j*Nx (i fastest), the
 * reference's own layout (S:140-153, WriteReadData.h:107-134).
 *
 * Error convention: every function returns STS_OK (0) or a negative code and
 * records a message readable with sts_last_error().  The reference has no
 * error codes (I/O failures print and continue, WriteReadData.h:64-67) and
 * silently ends loop 2 on a NaN residual (S:4433-4437); this library fails
 * loudly instead: any non-finite residual returns STS_ERR_NONFINITE.
 *
 * Threading: one context per host thread; all work of a context is issued on
 * the context's own HIP stream, in order.
 */
#ifndef SIMPLE_TS_H
#define SIMPLE_TS_H

#ifdef __cplusplus
extern "C" {
#endif

#define STS_OK 0
#define STS_ERR_ARG (-1)
#define STS_ERR_HIP (-2)
#define STS_ERR_NONFINITE (-3)
#define STS_ERR_UNSUPPORTED (-4)
#define STS_ERR_IO (-5)

/* Body parameters (Circles_Objects.h:38-78, Polyhedrons_Objects.h:32-100).
   Body k (0-based) is the one whose cells hold vol_inf_fb == k+1: circles
   first, then polyhedrons (S:6306-6466). */
typedef struct sts_body {
  double u_body, v_body, p_body, T_body;
  int is_slip;            /* is_VelocitySlipBC_body */
  double F_slip;          /* F_VelocitySlip */
  int is_jump;            /* is_TemperatureJumpBC_body */
  double F_jump;          /* F_TemperatureJump */
  double dTdn_on_wall_0;
} sts_body;

/* Everything the time step consumes besides the fields: mesh, body mask,
   solve boxes (define_area_for_solving, S:6009-6267), scalars (S:2013-2071,
   S:4347-4361) and boundary-condition switches (EnteredData.txt:62-110). */
typedef struct sts_grid {
  int Nx, Ny;
  const double *hx, *hy;            /* [Nx], [Ny] */
  const unsigned int *vol_inf_fb;   /* [Nx*Ny]: 0 fluid, k>=1 body k-1 */
  int nbodies;
  const sts_body *bodies;
  int Nx_u[2], Ny_u[2], Nx_v[2], Ny_v[2], Nx_o[2], Ny_o[2];
  double ht, gamma1, Kn, cT, cT1, cT2, cPch, c_Fr_v;
  int given_xb, given_xe, kn_local;
  int dudx_0_xb, durhodx_0_xb, drhodt_xb, dudx_0_xe, durhodx_0_xe, drhodt_xe,
      u_Min_Mout_xe, dudy_0_yb, dudy_0_ye;
  int dvdx_0_xb, dvdx_0_xe, dvdy_0_yb, dvdy_0_ye;
  int dpdx_0_xb, dpdx_0_xe, dpdy_0_yb, dpdy_0_ye;
  int dTdx_0_xb, dTdx_0_xe, dTdy_0_yb, dTdy_0_ye;
  /* Periodic_boundary_conditions_about_OX: x neighbours wrap (S:201-232),
     the x solve boxes span 0..Nx (S:6055-6102) and the box-edge conditions
     are off (S:6795, S:7340, S:7436, S:7655). */
  int periodic_x;
} sts_grid;

/* Field identifiers, the reference's array names (S:245-1792). */
enum sts_field {
  STS_U = 0, STS_V, STS_P, STS_T, STS_RHO, STS_SQRT_T,
  STS_U_PR, STS_V_PR, STS_P_PR, STS_T_PR, STS_RHO_PR,
  STS_RHO_U, STS_RHO_V, STS_FX, STS_FY,
  STS_GAMMA_XF, STS_GAMMA_YF, STS_D_UX, STS_D_UY, STS_D_VX, STS_D_VY, STS_DTX, STS_DTY,
  STS_U_PSEUDO, STS_V_PSEUDO, STS_D_P_C_12, STS_D_P_C_34, STS_APX, STS_APY, STS_BPX, STS_BPY,
  STS_SCT1, STS_TW_U, STS_TW_V,
  STS_NFIELDS
};

/* Parts of the per-iteration boundary conditions, run in the reference's
   order S:4460-4527 when several are requested. */
#define STS_BC_SLIP_U 0x01        /* BoundaryConditions_Velocities_SlipBC_u, S:7098 */
#define STS_BC_SLIP_V 0x02        /* BoundaryConditions_Velocities_SlipBC_v, S:7215 */
#define STS_BC_VELOCITIES 0x04    /* BoundaryConditions_Velocities, S:6790 */
#define STS_BC_JUMP 0x08          /* BoundaryConditions_Temperature_TemperatureJumpBC, S:7744 */
#define STS_BC_TEMPERATURE 0x10   /* BoundaryConditions_Temperature, S:7650 */
#define STS_BC_PRESSURE 0x20      /* BoundaryConditions_Pressure, S:7335 */
#define STS_BC_ALL 0x3f

typedef struct sts_ctx sts_ctx;

/* Last error message of the calling thread ("" if none). */
const char *sts_last_error(void);

/* Allocates the device state for one case on the current HIP device.  All
   fields start at zero, as the reference zero-fills them before loop 1
   (S:2872-2930).  precision: 64 (FP64, reference arithmetic), or 32: the
   same FP64 state plus an FP32 copy for sts_time_step_mixed. */
int sts_create(const sts_grid *g, int precision, sts_ctx **out);
int sts_destroy(sts_ctx *ctx);

/* Host <-> device copies of one field (double[Nx*Ny]); synchronous. */
int sts_set_field(sts_ctx *ctx, int field, const double *host);
int sts_get_field(sts_ctx *ctx, int field, double *host);

/* Per-iteration scalar recomputation of cT and cPch, S:4347-4361. */
int sts_set_scalars(sts_ctx *ctx, double cT, double cPch);

/* Time-step prologue, S:3233-3239: p_pr, u_pr, v_pr, Temper_pr, rho_pr <- current. */
int sts_save_previous(sts_ctx *ctx);

/* Coefficient assembly: u pseudo-velocity S:3269-3344 + S:3478-3545, v
   pseudo-velocity S:3373-3450 + S:3561-3630, energy source S:3655-3756. */
int sts_assemble(sts_ctx *ctx);

/* Loop 3, S:3763-4131: n fused T/p Jacobi sweeps.  res (may be NULL) gets
   2*n doubles: max|p-p_old|, max|T-T_old| of each sweep. */
int sts_pt_jacobi(sts_ctx *ctx, int n, double *res);

/* Velocity correction S:4170-4305, then sqrt_T and rho S:4323-4330, then the
   face densities S:4368 (Solve_rho_in_middle_points).  res (may be NULL)
   gets max|du|, max|dv|. */
int sts_correct(sts_ctx *ctx, double *res);

/* sqrt_T and rho over all cells, S:4323-4330 (also S:2948-2953). */
int sts_properties(sts_ctx *ctx);

/* Solve_rho_in_middle_points, S:10382-10437. */
int sts_rho_in_middle_points(sts_ctx *ctx);

/* Boundary conditions of one outer iteration, S:4460-4527; parts: STS_BC_*.
   p_probe (may be NULL) gets p[Nx_o[0] + j*Nx], j = int((Ny_o[1]-1.0)-Ny_o[0]),
   the value BoundaryConditions_Pressure_begin_iteration_for_it reads into
   p_BC_xb when the inlet velocity is given (S:7509-7514). */
int sts_boundary_conditions(sts_ctx *ctx, int parts, double *p_probe);

/* SolveFluxes S:10440-10454 and SolveDiffisionTerms S:10457-10618. */
int sts_fluxes_diffusion(sts_ctx *ctx);

/* ApplyBodiesConditionsToVariables, S:6470-6584. */
int sts_apply_bodies(sts_ctx *ctx);

/* Body-cell temperature overwrite done before every save, S:5566-5567 ->
   S:8048-8099. */
int sts_jump_overwrite(sts_ctx *ctx);

/* Inflow velocity u[i < Nx_u[0]] = u_in on fluid cells
   (BoundaryConditions_function_of_time, S:8102-8144). */
int sts_set_inflow_u(sts_ctx *ctx, double u_in);

/* Continuity residual S:10621-10646 (SolveContinuityEquation): writes the
   stationary and full fields into cont_stat / cont (host, double[Nx*Ny], may
   be NULL) and norms[3] = {max|cont|, sqrt(sum cont^2), sum|cont|} over the
   cells the reference fills (i < Nx-1, j < Ny-1) that are fluid. */
int sts_continuity(sts_ctx *ctx, double *cont_stat, double *cont, double *norms);

/* One outer iteration of loop 2, S:3263-4608: assembly, n_pc sweeps,
   correction, boundary conditions, fluxes and diffusion.  res[4] = max|du|,
   max|dv| and the last sweep's max|dp|, max|dT| -- the four numbers the
   reference's loop-2 test reads (S:4433-4437).  res_p0T0 (may be NULL) gets
   the first sweep's max|dp|, max|dT| (max_residual_in_*_Iter_0).  p_probe as
   in sts_boundary_conditions. */
int sts_outer_iteration(sts_ctx *ctx, int n_pc, double *res, double *res_p0T0,
                        double *p_probe);

/* Loop-2 control of one time step (S:4406-4450). */
typedef struct sts_loop_control {
  int N_I;            /* cap on outer iterations per step */
  int N_I_p_c;        /* fused T/p sweeps per outer iteration (loop 3) */
  double tol_uv;      /* MaxError_Velocities */
  double tol_p;       /* MaxError_p_c */
  double tol_T;       /* MaxError_T */
  int batch;          /* outer iterations in the first batch enqueued (the host reads the loop control one
                         iteration before a batch ends); 0 = from the last step's count */
} sts_loop_control;

/* One time step of loop 1 (S:3230-4608): the prologue S:3233-3239, then
   outer iterations (as sts_outer_iteration) until the reference's loop-2 test
   S:4433-4437 stops them or N_I is reached.  The test is evaluated on the GPU
   after each outer iteration; iterations are enqueued in batches and those
   past the stopping one exit without work, so the host waits about once per
   step.  *iterations gets Iter; res_log (may be NULL) gets 6 doubles per outer
   iteration: max|du|, max|dv|, max|dp|, max|dT| (last sweep), max|dp|,
   max|dT| (first sweep).  p_probe as in sts_boundary_conditions, for the last
   outer iteration.  With res_log and p_probe both NULL the call returns as
   soon as the iteration count is known, without waiting for the step's last
   launches (the flux pass of its last iteration, iterations gated out past
   the stop); later calls on the context queue behind them, and every call
   that hands device data back waits for them.  Returns STS_ERR_NONFINITE if
   a residual is not finite. */
int sts_time_step(sts_ctx *ctx, const sts_loop_control *lc, int *iterations, double *res_log, double *p_probe);

/* No counterpart in the reference: sts_time_step for callers that cannot
   allocate memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid allocation).  Same arguments, arithmetic and results,
   and no call of any allocation routine: the loop control, the residual
   words, the probe and the residual log live in one fixed device array of the
   library per GPU (about 260 KB, held until the process ends), which every
   context on that GPU shares and which keeps the last call's contents (each
   call clears what it reads before writing it).  Limits: N_I <=
   STS_NOALLOC_MAX_N_I (else STS_ERR_ARG); the context's device must be the
   current device.  Calls that overlap on one GPU (from several threads, or on
   several contexts) race on that array.  sts_time_kernel does not time these
   calls. */
#define STS_NOALLOC_MAX_N_I 4096
int sts_time_step_noalloc(sts_ctx *ctx, const sts_loop_control *lc, int *iterations, double *res_log,
                          double *p_probe);

/* Mixed-precision time step (no counterpart in the reference; needs a context
   created with precision 32).  After the prologue, loop 2 first runs on an
   FP32 copy of the state -- assembly, sweeps, correction, boundary conditions,
   fluxes and diffusion coefficients all in FP32 -- until its successive
   differences fall below the FP32 tolerances, stagnate (the largest of
   max|d*|/tol shrinks by less than stall_ratio over stall_w iterations) or
   N_I32 iterations have run.  The FP32 state is then widened to FP64 and loop
   2 continues in FP64 exactly as in sts_time_step: the reference's test on
   FP64 successive differences alone decides convergence, so the step ends at
   the same tolerance as the FP64 path.  res_log (room for 6*(N_I32+N_I)
   doubles) gets 6 doubles per outer iteration, FP32 iterations first;
   res_log and p_probe may both be NULL, with the meaning they have in
   sts_time_step.
   N_I32 = 0 gives sts_time_step. */
typedef struct sts_mixed_control {
  int N_I32;                          /* cap on FP32 outer iterations */
  double tol32_uv, tol32_p, tol32_T;  /* FP32-phase tolerances */
  int stall_w;                        /* stagnation window (0: off) */
  double stall_ratio;                 /* required shrink over the window */
  int extrapolate;                    /* q = 1..5: start loop 2 from the degree-q polynomial in time
                                         through the last q+1 converged steps (q = 1: 2 x^n - x^(n-1);
                                         q = 2: 3 x^n - 3 x^(n-1) + x^(n-2); ...) for u, v, p, T on every
                                         cell, the derived fields rebuilt from them, instead of x^n; a
                                         cell whose guess is not finite (or p, T not positive) keeps x^n;
                                         fewer stored steps give a lower degree; 0: off */
  int fp32_above;                     /* > 0: run the FP32 phase only when the context's previous step
                                         took more than this many outer iterations (a well-started step
                                         goes straight to FP64); 0: always */
  int guard;                          /* > 0 (with extrapolate): a predicted step whose FP64 phase has not
                                         met the reference's test within as many outer iterations as
                                         steps from x^n recently took (their running mean) is restarted
                                         from x^n, FP64 only (accelerated if anderson asks for it, as
                                         every step from x^n), and the degree in use drops by one (it
                                         climbs back after 4 accepted predicted steps); an abandoned
                                         degree-1 attempt pauses prediction for 1, 2, 4, ... steps (at
                                         most guard).  0: every step predicted at the full degree */
  int anderson;                       /* m = 1..8: Anderson acceleration of the FP64 phase (Walker and Ni
                                         2011) over u, v, p, T of every cell: every anderson_every outer
                                         iterations the next start becomes the affine combination of the
                                         last m+1 stored iterates whose combined change is least in the
                                         2-norm, and the derived fields are rebuilt from it; the test
                                         still judges each outer iteration's own change, and the
                                         iteration that meets it is the state the step keeps.  0: off */
  int anderson_every;                 /* period of the Anderson step in outer iterations (0 or 1: every) */
  int anderson32;                     /* m = 1..8: the same acceleration of the FP32 phase (its iterates
                                         and mixes in FP32, the acceleration's sums in FP64); the FP64
                                         phase starts its own history.  0: off.  Batched steps run the
                                         FP32 phase without it. */
} sts_mixed_control;
int sts_time_step_mixed(sts_ctx *ctx, const sts_loop_control *lc, const sts_mixed_control *mc, int *iters32,
                        int *iters64, double *res_log, double *p_probe);
/* Counters of the guarded predicted start since the context was created:
   out[0] predicted steps, out[1] of them restarted from x^n (batched steps:
   that went over the budget), out[2] FP32 and out[3] FP64 outer iterations
   spent in the abandoned attempts (not included in the counts
   sts_time_step_mixed returns), out[4] steps run from x^n by the guard. */
int sts_predict_stats(sts_ctx *ctx, long long out[5]);

/* Batched time steps (no counterpart in the reference).  The cases of n
   contexts (any grid sizes) advance one time step together: every kernel of
   loop 2 is launched once for all of them over a list of tiles of every
   case, each workgroup following its own case's loop-2 control, so a case
   that has converged drops out while the others iterate.  Each case's
   arithmetic is that of sts_time_step / sts_time_step_mixed (mc may be
   NULL), bit for bit, except that the dot products of an Anderson step
   (mc->anderson > 0) are summed over the case's tiles in another order;
   a guarded predicted step that exceeds its budget pauses prediction for
   that case instead of restarting inside the shared launches.  All
   contexts need the same N_I_p_c; their own streams must be idle.
   iters32/iters64 (may be NULL) get n counts each. */
typedef struct sts_batch sts_batch;
int sts_batch_create(sts_ctx *const *ctxs, int n, sts_batch **out);
int sts_batch_destroy(sts_batch *b);
int sts_batch_time_step(sts_batch *b, const sts_loop_control *lc, const sts_mixed_control *mc, int *iters32,
                        int *iters64);
/* Tile shapes of the batched launches: variant v of the cell kernels is
   64x4, 64x8, 128x2, 32x8 cells, of the fused sweep pass 64x8, 64x4, 128x4,
   64x16 (v = 0..3).  sts_batch_autotune times each variant `reps` times on
   the current state, keeps the fastest and reports the mean milliseconds per
   variant in ms[8] (cell kernels, then sweeps; may be NULL). */
int sts_batch_set_tiles(sts_batch *b, int cell_variant, int sweep_variant);
int sts_batch_autotune(sts_batch *b, int reps, int *cell_variant, int *sweep_variant, double *ms);

/* Drag-coefficient sums of body `body` (0-based, circles first, the
   numbering of vol_inf_fb minus one), WriteDragCoefficient S:8203-8262: over
   the body's cells of the p/T box in the reference's order, for the faces
   bottom, front, top, behind (in that order) the friction integral, the
   pressure integral and the face length -- sums[12].  The caller applies the
   nondimensional factors of S:8265-8283 and writes CD.<k>.txt. */
int sts_drag_sums(sts_ctx *ctx, int body, double *sums);

/* The HIP stream all work of the context is issued on (diagnostics). */
int sts_stream_info(sts_ctx *ctx, void **hip_stream);

/* Kernel timing (diagnostics, no counterpart in the reference): every launch
   of a kernel in the timed set is bracketed by HIP events on the context's
   stream; launches that a device-side loop gates out (outer iterations
   enqueued past the iteration where the loop stopped, which return at once)
   are not counted.  sts_time_kernel adds `kernel` to the set and clears its
   totals (-1: empties the set); sts_kernel_time reports the counted launches
   of `kernel` and their summed duration.  kernel: STS_K_* for the FP64
   kernel, STS_K_* + STS_K_COUNT for its FP32 instance.  STS_K_ASSEMBLE is
   the assembly that reads Gamma_* and D_*, STS_K_ASSEMBLE_INL the one inside
   the loops that forms them per face; STS_K_AA_* are the Anderson kernels. */
enum sts_kernel { STS_K_ASSEMBLE = 0, STS_K_SWEEP, STS_K_CORRECT, STS_K_WALL, STS_K_EDGES, STS_K_FLUX_DIFF,
                  STS_K_ASSEMBLE_INL, STS_K_AA_STORE, STS_K_AA_MIX, STS_K_COUNT };
int sts_time_kernel(sts_ctx *ctx, int kernel);
int sts_kernel_time(sts_ctx *ctx, int kernel, long long *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
