/*
 * CPU FP64 restatement of the SIMPLE-TS time step (test infrastructure).
 *
 * This is the checker, not the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path (the HIP
 * library under simple-ts_amd/) never calls into it.
 *
 * Every function restates one call site of the reference's main() loop
 * (SIMPLE_TS_MPI.cpp) as a dense loop over the whole Nx*Ny grid whose update
 * predicate equals the union of the reference's "bc" and "nch" rule vectors
 * (S:8581-9061), evaluated with the reference's "bc" formula, which reduces
 * bit-exactly to its "nch" formula where every mask is 1.  Expression trees
 * follow the reference term by term; build with -ffp-contract=off.
 *
 * Layout: node = i + j*Nx, i fastest (S:140-153).
 */
#ifndef STS_ORACLE_H
#define STS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_body {
  double u_body, v_body, p_body, T_body;
  int is_slip;
  double F_slip;
  int is_jump;
  double F_jump;
  double dTdn_on_wall_0;
} ora_body;

typedef struct ora_grid {
  int Nx, Ny;
  const double *hx, *hy;            /* [Nx], [Ny] */
  const unsigned int *vol;          /* vol_inf_fb [Na]: 0 fluid, k>=1 body k-1 */
  const unsigned char *m;           /* vol_inf_fb_bool [Na] */
  int nbodies;
  const ora_body *bodies;           /* circles first, then polyhedrons */
  int Nx_u[2], Ny_u[2], Nx_v[2], Ny_v[2], Nx_o[2], Ny_o[2];
  double ht, gamma1, Kn, cT, cT1, cT2, cPch, c_Fr_v;
  int given_xb, given_xe, kn_local;
  /* edge boundary-condition switches, EnteredData.txt:81-109 */
  int dudx_0_xb, durhodx_0_xb, drhodt_xb, dudx_0_xe, durhodx_0_xe, drhodt_xe,
      u_Min_Mout_xe, dudy_0_yb, dudy_0_ye;
  int dvdx_0_xb, dvdx_0_xe, dvdy_0_yb, dvdy_0_ye;
  int dpdx_0_xb, dpdx_0_xe, dpdy_0_yb, dpdy_0_ye;
  int dTdx_0_xb, dTdx_0_xe, dTdy_0_yb, dTdy_0_ye;
  int periodic_x;                   /* Periodic_boundary_conditions_about_OX */
} ora_grid;

typedef struct ora_fields {
  double *u, *v, *p, *T, *rho, *sqrt_T;
  double *u_pr, *v_pr, *p_pr, *T_pr, *rho_pr;
  double *rho_u, *rho_v, *Fx, *Fy;
  double *Gamma_xf, *Gamma_yf, *D_ux, *D_uy, *D_vx, *D_vy, *DTx, *DTy;
  double *u_pseudo, *v_pseudo, *d_p_c_12, *d_p_c_34, *apx, *apy, *bpx, *bpy;
  double *ScT1, *Tw_u, *Tw_v;
  double *p_old, *T_old;            /* loop-3 work buffers */
} ora_fields;

/* A1, S:3233-3239: *_pr <- current */
void ora_save_previous(const ora_grid *g, ora_fields *f);
/* A2, S:3269-3344 + S:3478-3545 */
void ora_assemble_u(const ora_grid *g, ora_fields *f);
/* A3, S:3373-3450 + S:3561-3630 */
void ora_assemble_v(const ora_grid *g, ora_fields *f);
/* A4, S:3655-3756 */
void ora_energy_source(const ora_grid *g, ora_fields *f);
/* A5, S:3763-4131: n sweeps; res[2k] = max|dp|, res[2k+1] = max|dT| of sweep k */
void ora_pt_jacobi(const ora_grid *g, ora_fields *f, int n, double *res);
/* A6, S:4170-4305: res[0] = max|du|, res[1] = max|dv| */
void ora_correct_velocities(const ora_grid *g, ora_fields *f, double *res);
/* A7, S:4323-4330 */
void ora_properties(const ora_grid *g, ora_fields *f);
/* A9, S:10382-10437 */
void ora_rho_in_middle_points(const ora_grid *g, ora_fields *f);
/* A12 pieces, S:7098-7331, S:6790-7096, S:7744-8045, S:7650-7741, S:7335-7428 */
void ora_bc_slip_u(const ora_grid *g, ora_fields *f);
void ora_bc_slip_v(const ora_grid *g, ora_fields *f);
void ora_bc_velocities(const ora_grid *g, ora_fields *f);
void ora_bc_temperature_jump(const ora_grid *g, ora_fields *f);
void ora_bc_temperature(const ora_grid *g, ora_fields *f);
void ora_bc_pressure(const ora_grid *g, ora_fields *f);
/* S:8048-8099: body-cell Temper overwrite done before every save */
void ora_jump_overwrite(const ora_grid *g, ora_fields *f);
/* A10, S:10440-10454 */
void ora_fluxes(const ora_grid *g, ora_fields *f);
/* A11, S:10457-10618 */
void ora_diffusion(const ora_grid *g, ora_fields *f);
/* S:6470-6584 */
void ora_apply_bodies(const ora_grid *g, ora_fields *f);
/* S:10621-10646: continuity residual field (stationary part and full) */
void ora_continuity(const ora_grid *g, const ora_fields *f, double *cont_stat,
                    double *cont);

/* One outer iteration of loop 2 (S:3263-4608) with N_I_p_c sweeps; res[4]:
   max|du|, max|dv|, max|dp|, max|dT| (the last sweep's), as tested at
   S:4433-4437.  The scalar side effects of the pressure inflow bookkeeping
   (S:7430-7647) are not modelled here. */
void ora_outer_iteration(const ora_grid *g, ora_fields *f, int n_pc, double *res);

#ifdef __cplusplus
}
#endif
#endif
