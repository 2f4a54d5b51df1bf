/*
 * CPU FP64 restatement of the SIMPLE-TS time step -- test infrastructure only
 * (the checker; never linked into the product).  See sts_oracle.h.
 *
 * Citations "S:n" are lines of SIMPLE_TS_MPI.cpp in the reference.
 */
#include "sts_oracle.h"

#include <math.h>
#include <string.h>

#define MAXI(a, b) (((a) > (b)) ? (a) : (b)) /* S:121 maximum */
/* S:130 upwind_rho: the mean at exactly zero velocity */
#define UPWIND_RHO(r0, r1, v1) (((v1) == 0) ? (0.5 * ((r0) + (r1))) : (((v1) > 0) ? (r0) : (r1)))
/* S:133 Li, S:136 Hi */
#define LI(f1, f2, h1, h2) (((f1) * (h2) + (f2) * (h1)) / ((h1) + (h2)))
#define HI(f1, f2, h1, h2) (((h1) + (h2)) * (f1) * (f2) / ((h1) * (f2) + (h2) * (f1)))

/* Neighbour indices: clamped, S:140-153 (ij_function), or wrapping in x,
   S:201-232 (ij_function_PeriodicBC; j stays clamped). */
typedef struct {
  unsigned i, j, ij, i_1, i_2, i1, i_1j, i1j, ij_1, ij1, i_1j_1, i_1j1, i1j_1, i1j1, i_2j;
} nb_t;

static void nbr_px(const ora_grid *g, unsigned i, unsigned j, int px, nb_t *n) {
  const unsigned Nx = (unsigned)g->Nx, Ny = (unsigned)g->Ny;
  unsigned ij = i + j * Nx;
  n->i = i;
  n->j = j;
  n->ij = ij;
  if (px) {
    n->i_1 = (i == 0) ? Nx - 1 : i - 1;
    n->i_2 = (i == 1) ? Nx - 1 : (i == 0) ? Nx - 2 : i - 2;
    n->i1 = (i == Nx - 1) ? 0 : i + 1;
  } else {
    n->i_1 = i - 1 * (0 < i);
    n->i_2 = i - 1 * (0 < i) - 1 * (1 < i);
    n->i1 = i + 1 * (i < (Nx - 1));
  }
  const unsigned js = j - 1 * (0 < j), jn = j + 1 * (j < (Ny - 1));
  n->i_1j = n->i_1 + j * Nx;
  n->i1j = n->i1 + j * Nx;
  n->ij_1 = i + js * Nx;
  n->ij1 = i + jn * Nx;
  n->i_1j_1 = n->i_1 + js * Nx;
  n->i_1j1 = n->i_1 + jn * Nx;
  n->i1j_1 = n->i1 + js * Nx;
  n->i1j1 = n->i1 + jn * Nx;
  n->i_2j = n->i_2 + j * Nx;
}

/* The hot-path call sites take the wrapping form when the case is periodic
   about OX (S:3276, S:3381, S:3709, S:3813, S:4180, S:4224, S:7112, S:7230,
   S:7752, S:8058, S:10174-10354, S:10390, S:10465). */
static void nbr(const ora_grid *g, unsigned i, unsigned j, nb_t *n) { nbr_px(g, i, j, g->periodic_x, n); }

/* Body parameters of vol_inf_fb value k >= 1 (S:7114-7130). */
static const ora_body *body_of(const ora_grid *g, unsigned k) {
  if (k < 1 || (int)k > g->nbodies) return 0;
  return &g->bodies[k - 1];
}

void ora_save_previous(const ora_grid *g, ora_fields *f) {
  size_t n = (size_t)g->Nx * g->Ny * sizeof(double);
  memcpy(f->p_pr, f->p, n);
  memcpy(f->u_pr, f->u, n);
  memcpy(f->v_pr, f->v, n);
  memcpy(f->T_pr, f->T, n);
  memcpy(f->rho_pr, f->rho, n);
}

/* A2: u pseudo-velocity (bc formula S:3283-3343; update set = rule vectors
   *_u, S:8618-8715: the u box with fluid on both sides of the face). */
void ora_assemble_u(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const double *hx = g->hx, *hy = g->hy;
  const double cT = g->cT, ht = g->ht, cPch = g->cPch;
  const double *Fx = f->Fx, *Fy = f->Fy, *u = f->u, *v = f->v, *sqrt_T = f->sqrt_T;
  const double *Gamma_yf = f->Gamma_yf, *D_ux = f->D_ux, *D_uy = f->D_uy;
  const double *rho_pr = f->rho_pr, *u_pr = f->u_pr, *rho = f->rho, *rho_u = f->rho_u;
  for (unsigned j = (unsigned)g->Ny_u[0]; j < (unsigned)g->Ny_u[1]; j++)
    for (unsigned i = (unsigned)g->Nx_u[0]; i < (unsigned)g->Nx_u[1]; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, i1j = n.i1j, ij_1 = n.ij_1, ij1 = n.ij1;
      const unsigned i_1j_1 = n.i_1j_1, i_1j1 = n.i_1j1, i_2j = n.i_2j, i_1 = n.i_1;
      if (!(m[i_1j] && m[ij])) continue;
      double F_xu1 = 0.5 * (Fx[i_1j] + Fx[ij]);
      double F_xu2 = 0.5 * (Fx[ij] + Fx[i1j]);
      double Scu =
          cT * ((double)m[i1j] * u[i1j] * sqrt_T[ij] * hy[j] / (hx[i] * 3.0)
                + (double)m[i_2j] * u[i_1j] * sqrt_T[i_1j] * hy[j] / (hx[i_1] * 3.0)
                + Gamma_yf[ij1] * (v[ij1] * (double)m[ij1] - v[i_1j1] * (double)m[i_1j1])
                - Gamma_yf[ij] * (v[ij] * (double)m[ij_1] - v[i_1j] * (double)m[i_1j_1])
                - sqrt_T[ij] * (v[ij1] * (double)m[ij1] - v[ij] * (double)m[ij_1]) * 2.0 / 3.0
                + sqrt_T[i_1j] * (v[i_1j1] * (double)m[i_1j1] - v[i_1j] * (double)m[i_1j_1]) * 2.0 / 3.0);
      double Spu = -(D_ux[ij] + D_ux[i1j]) / 3.0;
      double au1 = MAXI(0.0, F_xu1) + D_ux[ij];
      double au2 = MAXI(0.0, -F_xu2) + D_ux[i1j];
      double au3 = 0.5 * (MAXI(0.0, Fy[i_1j]) + MAXI(0.0, Fy[ij])) + D_uy[ij];
      double au4 = 0.5 * (MAXI(0.0, -Fy[i_1j1]) + MAXI(0.0, -Fy[ij1])) + D_uy[ij1];
      double bu3 = (rho_pr[i_1j] * hx[i_1] + rho_pr[ij] * hx[i]) * u_pr[ij] * hy[j] * 0.5 / ht + Scu;
      double au0 = au1 + au2 + au3 + au4 + 0.5 * (-Fy[i_1j] - Fy[ij] + Fy[i_1j1] + Fy[ij1]) + F_xu2 - F_xu1
                   - Spu + (rho[i_1j] * hx[i_1] + rho[ij] * hx[i]) * hy[j] * 0.5 / ht;
      f->u_pseudo[ij] = (au1 * (double)m[i1j] * u[i_1j] + au2 * (double)m[i_2j] * u[i1j] + au3 * u[ij_1]
                         + au4 * u[ij1] + bu3) / au0;
      f->d_p_c_12[ij] = cPch * hy[j] / au0;
      double h = hy[j] * rho_u[ij];
      f->apx[ij] = f->d_p_c_12[ij] * h;
      f->bpx[ij] = (double)(m[i_1j] * m[ij]) * h * f->u_pseudo[ij];
    }
}

/* A3: v pseudo-velocity (bc formula S:3387-3449, gravity S:3404; update set
   S:8717-8814). */
void ora_assemble_v(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx;
  const double *hx = g->hx, *hy = g->hy;
  const double cT = g->cT, ht = g->ht, cPch = g->cPch, c_Fr_v = g->c_Fr_v;
  const double *Fx = f->Fx, *Fy = f->Fy, *u = f->u, *v = f->v, *sqrt_T = f->sqrt_T;
  const double *Gamma_xf = f->Gamma_xf, *D_vx = f->D_vx, *D_vy = f->D_vy;
  const double *rho_pr = f->rho_pr, *v_pr = f->v_pr, *rho = f->rho, *rho_v = f->rho_v;
  for (unsigned j = (unsigned)g->Ny_v[0]; j < (unsigned)g->Ny_v[1]; j++)
    for (unsigned i = (unsigned)g->Nx_v[0]; i < (unsigned)g->Nx_v[1]; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, i1j = n.i1j, ij_1 = n.ij_1, ij1 = n.ij1;
      const unsigned i_1j_1 = n.i_1j_1, i1j_1 = n.i1j_1;
      if (!(m[ij_1] && m[ij])) continue;
      /* S:3399: ij_1 - Nx*(0 <= (int)(j-2)); S:3400: hy[j - 1*(0 <= (int)(j-1))] */
      const unsigned jm2 = ij_1 - Nx * (0 <= (int)(j - 2));
      const double hyjm = hy[j - 1 * (0 <= (int)(j - 1))];
      double F_yv3 = 0.5 * (Fy[ij_1] + Fy[ij]);
      double F_yv4 = 0.5 * (Fy[ij] + Fy[ij1]);
      double Scv =
          cT * ((double)m[ij1] * v[ij1] * sqrt_T[ij] * hx[i] / (hy[j] * 3.0)
                + (double)m[jm2] * v[ij_1] * sqrt_T[ij_1] * hx[i] / (hyjm * 3.0)
                + Gamma_xf[i1j] * (u[i1j] * (double)m[i1j] - u[i1j_1] * (double)m[i1j_1])
                - Gamma_xf[ij] * (u[ij] * (double)m[i_1j] - u[ij_1] * (double)m[i_1j_1])
                - sqrt_T[ij] * (u[i1j] * (double)m[i1j] - u[ij] * (double)m[i_1j]) * 2.0 / 3.0
                + sqrt_T[ij_1] * (u[i1j_1] * (double)m[i1j_1] - u[ij_1] * (double)m[i_1j_1]) * 2.0 / 3.0)
          - c_Fr_v * (rho[ij_1] * hx[i] * 0.5 * hyjm + rho[ij] * hx[i] * 0.5 * hy[j]);
      double Spv = -(D_vy[ij] + D_vy[ij1]) / 3.0;
      double av1 = 0.5 * (MAXI(0.0, Fx[ij]) + MAXI(0.0, Fx[ij_1])) + D_vx[ij];
      double av2 = 0.5 * (MAXI(0.0, -Fx[i1j]) + MAXI(0.0, -Fx[i1j_1])) + D_vx[i1j];
      double av3 = MAXI(0.0, F_yv3) + D_vy[ij];
      double av4 = MAXI(0.0, -F_yv4) + D_vy[ij1];
      double bv3 = (rho_pr[ij_1] * hy[j - 1] + rho_pr[ij] * hy[j]) * v_pr[ij] * hx[i] * 0.5 / ht + Scv;
      double av0 = av1 + av2 + av3 + av4 + 0.5 * (Fx[i1j] - Fx[ij] + Fx[i1j_1] - Fx[ij_1]) + F_yv4 - F_yv3
                   - Spv + (rho[ij_1] * hy[j - 1] + rho[ij] * hy[j]) * hx[i] * 0.5 / ht;
      f->v_pseudo[ij] = (av1 * v[i_1j] + av2 * v[i1j] + av3 * (double)m[jm2] * v[ij_1]
                         + av4 * (double)m[ij1] * v[ij1] + bv3) / av0;
      f->d_p_c_34[ij] = cPch * hx[i] / av0;
      double h = hx[i] * rho_v[ij];
      f->apy[ij] = f->d_p_c_34[ij] * h;
      f->bpy[ij] = (double)(m[ij - Nx] * m[ij]) * h * f->v_pseudo[ij];
    }
}

/* A4: energy source ScT1 (bc formula S:3715-3755; sub-volume terms S:3729-3740
   with the j < Nx-1 guard of the reference). */
void ora_energy_source(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx;
  const double *hx = g->hx, *hy = g->hy;
  const double ht = g->ht, cT2 = g->cT2;
  const double *u = f->u, *v = f->v, *sqrt_T = f->sqrt_T, *p_pr = f->p_pr;
  for (unsigned j = (unsigned)g->Ny_o[0]; j < (unsigned)g->Ny_o[1]; j++)
    for (unsigned i = (unsigned)g->Nx_o[0]; i < (unsigned)g->Nx_o[1]; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, i1j = n.i1j, ij_1 = n.ij_1, ij1 = n.ij1;
      const unsigned i_1j_1 = n.i_1j_1, i_1j1 = n.i_1j1, i1j_1 = n.i1j_1, i1j1 = n.i1j1;
      const unsigned i_1 = n.i_1, i1 = n.i1;
      if (!m[ij]) continue;
      const unsigned jm = j - 1 * (0 < j), jp = j + 1 * (j < (Nx - 1));
      double dudx_T = ((double)m[i1j] * u[i1j] - (double)m[i_1j] * u[ij]) / hx[i];
      double dvdy_T = ((double)m[ij1] * v[ij1] - (double)m[ij_1] * v[ij]) / hy[j];
      double s0 = ((u[ij] - u[ij_1]) / ((double)(m[i_1j_1] && m[ij_1]) * hy[jm] + hy[j]))
                  + ((v[ij] - v[i_1j]) / ((double)(m[i_1j_1] && m[i_1j]) * hx[i_1] + hx[i]));
      double s1 = ((u[i1j] - u[i1j_1]) / ((double)(m[ij_1] && m[i1j_1]) * hy[jm] + hy[j]))
                  + ((v[i1j] - v[ij]) / (hx[i] + (double)(m[i1j_1] && m[i1j]) * hx[i1]));
      double s2 = ((u[ij1] - u[ij]) / (hy[j] + (double)(m[i_1j1] && m[ij1]) * hy[jp]))
                  + ((v[ij1] - v[i_1j1]) / ((double)(m[i_1j] && m[i_1j1]) * hx[i_1] + hx[i]));
      double s3 = ((u[i1j1] - u[i1j]) / (hy[j] + (double)(m[ij1] && m[i1j1]) * hy[jp]))
                  + ((v[i1j1] - v[ij1]) / (hx[i] + (double)(m[i1j] && m[i1j1]) * hx[i1]));
      double dsum = dudx_T + dvdy_T;
      f->ScT1[ij] = ht * sqrt_T[ij] * cT2 * hx[i] * hy[j]
                        * (2.0 * (dudx_T * dudx_T + dvdy_T * dvdy_T) + s0 * s0 + s1 * s1 + s2 * s2 + s3 * s3
                           - dsum * dsum * 2.0 / 3.0)
                    + p_pr[ij] * hx[i] * hy[j];
    }
}

/* A5: fused T/p Jacobi sweeps, S:3763-4131 (bc formula S:3837-3912).  The
   entry copy S:3763-3764 and the pointer swap S:3792-3795 are restated by
   ping-ponging between f->T/f->p and f->T_old/f->p_old; on return the latest
   sweep is in f->T/f->p and the one before in f->T_old/f->p_old. */
void ora_pt_jacobi(const ora_grid *g, ora_fields *f, int nsweep, double *res) {
  const unsigned char *m = g->m;
  const unsigned Na = (unsigned)(g->Nx * g->Ny);
  const double *hx = g->hx, *hy = g->hy;
  const double ht = g->ht, gamma1 = g->gamma1;
  double *const p0 = f->p, *const T0 = f->T, *const po0 = f->p_old, *const To0 = f->T_old;
  memcpy(f->p_old, f->p, Na * sizeof(double));
  memcpy(f->T_old, f->T, Na * sizeof(double));
  for (int s = 0; s < nsweep; s++) {
    /* swap so that the sweep reads *_old and writes the current arrays */
    double *t;
    t = f->p; f->p = f->p_old; f->p_old = t;
    t = f->T; f->T = f->T_old; f->T_old = t;
    /* the written buffer keeps the read buffer's values outside the update set */
    memcpy(f->p, f->p_old, Na * sizeof(double));
    memcpy(f->T, f->T_old, Na * sizeof(double));
    const double *P = f->p_old, *TO = f->T_old;
    const double *Fx = f->Fx, *Fy = f->Fy, *DTx = f->DTx, *DTy = f->DTy, *u = f->u, *v = f->v;
    const double *Twu = f->Tw_u, *Twv = f->Tw_v, *ScT1 = f->ScT1, *rho = f->rho;
    const double *apx = f->apx, *apy = f->apy, *bpx = f->bpx, *bpy = f->bpy;
    const double *p_pr = f->p_pr, *T_pr = f->T_pr;
    double rp = 0, rT = 0;
    for (unsigned j = (unsigned)g->Ny_o[0]; j < (unsigned)g->Ny_o[1]; j++)
      for (unsigned i = (unsigned)g->Nx_o[0]; i < (unsigned)g->Nx_o[1]; i++) {
        nb_t n;
        nbr(g, i, j, &n);
        const unsigned ij = n.ij, i_1j = n.i_1j, i1j = n.i1j, ij_1 = n.ij_1, ij1 = n.ij1;
        if (!m[ij]) continue;
        double aT1 = MAXI(0.0, Fx[ij]) + DTx[ij];
        double aT2 = MAXI(0.0, -Fx[i1j]) + DTx[i1j];
        double aT3 = MAXI(0.0, Fy[ij]) + DTy[ij];
        double aT4 = MAXI(0.0, -Fy[ij1]) + DTy[ij1];
        double Tn =
            (ht * (aT1 * ((double)m[i_1j] * TO[i_1j] + (double)(!m[i_1j]) * Twu[ij])
                   + aT2 * ((double)m[i1j] * TO[i1j] + (double)(!m[i1j]) * Twu[i1j])
                   + aT3 * ((double)m[ij_1] * TO[ij_1] + (double)(!m[ij_1]) * Twv[ij])
                   + aT4 * ((double)m[ij1] * TO[ij1] + (double)(!m[ij1]) * Twv[ij1])
                   + (1.0 - gamma1) * P[ij] * ((u[i1j] - u[ij]) * hy[j] + (v[ij1] - v[ij]) * hx[i]))
             + ScT1[ij])
            / (ht * (aT1 + aT2 + aT3 + aT4 + Fx[i1j] - Fx[ij] + Fy[ij1] - Fy[ij]) + hx[i] * hy[j] * rho[ij]);
        f->T[ij] = Tn;
        double gxb = (double)(g->given_xb * ((int)i == g->Nx_o[0]));
        double gxe = (double)(g->given_xe * ((int)i == g->Nx_o[1] - 1));
        double pn = (ht * (apx[ij] * P[i_1j] + apx[i1j] * P[i1j] + apy[ij] * P[ij_1] + apy[ij1] * P[ij1]
                           + bpx[ij] - bpx[i1j] + bpy[ij] - bpy[ij1] + gxb * Fx[ij] - gxe * Fx[i1j])
                     + p_pr[ij] * hx[i] * hy[j] / (T_pr[ij]))
                    / (ht * (apx[ij] + apx[i1j] + apy[ij] + apy[ij1]) + hx[i] * hy[j] / (Tn));
        f->p[ij] = pn;
        double r_T = Tn - TO[ij], r_p = pn - P[ij];
        if (rT < fabs(r_T)) rT = fabs(r_T);
        if (rp < fabs(r_p)) rp = fabs(r_p);
      }
    if (res) {
      res[2 * s] = rp;
      res[2 * s + 1] = rT;
    }
  }
  if (f->p != p0) { /* odd sweep count: swap contents back into the caller's buffers */
    for (unsigned c = 0; c < Na; c++) {
      double t;
      t = p0[c]; p0[c] = po0[c]; po0[c] = t;
      t = T0[c]; T0[c] = To0[c]; To0[c] = t;
    }
    f->p = p0; f->p_old = po0; f->T = T0; f->T_old = To0;
  }
}

/* A6: velocity correction, S:4170-4305. */
void ora_correct_velocities(const ora_grid *g, ora_fields *f, double *res) {
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx;
  double ru = 0, rv = 0;
  for (unsigned j = (unsigned)g->Ny_u[0]; j < (unsigned)g->Ny_u[1]; j++)
    for (unsigned i = (unsigned)g->Nx_u[0]; i < (unsigned)g->Nx_u[1]; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      if (!(m[n.i_1j] && m[n.ij])) continue;
      double old = f->u[n.ij];
      f->u[n.ij] = f->u_pseudo[n.ij] - (f->p[n.ij] - f->p[n.i_1j]) * f->d_p_c_12[n.ij];
      double r = f->u[n.ij] - old;
      if (ru < fabs(r)) ru = fabs(r);
    }
  for (unsigned j = (unsigned)g->Ny_v[0]; j < (unsigned)g->Ny_v[1]; j++)
    for (unsigned i = (unsigned)g->Nx_v[0]; i < (unsigned)g->Nx_v[1]; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      if (!(m[n.ij_1] && m[n.ij])) continue;
      double old = f->v[n.ij];
      f->v[n.ij] = f->v_pseudo[n.ij] - (f->p[n.ij] - f->p[n.ij - Nx]) * f->d_p_c_34[n.ij];
      double r = f->v[n.ij] - old;
      if (rv < fabs(r)) rv = fabs(r);
    }
  if (res) {
    res[0] = ru;
    res[1] = rv;
  }
}

/* A7: S:4323-4330 */
void ora_properties(const ora_grid *g, ora_fields *f) {
  const unsigned Na = (unsigned)(g->Nx * g->Ny);
  for (unsigned c = 0; c < Na; c++) {
    f->sqrt_T[c] = sqrt(f->T[c]);
    if (g->m[c]) f->rho[c] = f->p[c] / f->T[c];
    else f->rho[c] = 0;
  }
}

/* A9: S:10382-10437 */
void ora_rho_in_middle_points(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const double *rho = f->rho;
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, ij_1 = n.ij_1;
      if (m[i_1j] && m[ij]) f->rho_u[ij] = UPWIND_RHO(rho[i_1j], rho[ij], f->u[ij]);
      else if (m[i_1j] != m[ij]) f->rho_u[ij] = m[i_1j] ? rho[i_1j] : rho[ij];
      else f->rho_u[ij] = 0.0;
      if (m[ij_1] && m[ij]) f->rho_v[ij] = UPWIND_RHO(rho[ij_1], rho[ij], f->v[ij]);
      else if (m[ij_1] != m[ij]) f->rho_v[ij] = m[ij_1] ? rho[ij_1] : rho[ij];
      else f->rho_v[ij] = 0.0;
    }
}

/* S:7098-7213: velocity slip at body faces for u (list S:10157-10206). */
void ora_bc_slip_u(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const unsigned *vol = g->vol;
  const double *hy = g->hy;
  const double Kn = g->Kn;
  double *u = f->u;
  const double *rho_u = f->rho_u;
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, ij_1 = n.ij_1, ij1 = n.ij1, i1j = n.i1j;
      if (!((0 < vol[ij]) && ((0 == vol[ij_1]) || (0 == vol[ij1])))) continue;
      const ora_body *b = body_of(g, vol[ij]);
      if (!b || !b->is_slip) continue;
      double F = b->F_slip, ub = b->u_body, a_ij, a_i1j;
      if (0 == vol[ij_1]) {
        const double hh = hy[j - 1 * (0 < j)];
        if (g->kn_local) {
          a_ij = (F * Kn) / (rho_u[ij_1] * 0.5 * hh);
          a_i1j = (F * Kn) / (rho_u[n.i1j_1] * 0.5 * hh);
        } else {
          a_ij = (F * Kn) / (0.5 * hh);
          a_i1j = (F * Kn) / (0.5 * hh);
        }
        if (m[n.i_1j_1]) u[ij] = (u[ij_1] * a_ij + ub) / (1.0 + a_ij);
        if (m[n.i1j_1]) u[i1j] = (u[n.i1j_1] * a_i1j + ub) / (1.0 + a_i1j);
      } else {
        const double hh = hy[j + 1 * (j < (unsigned)(g->Ny - 1))];
        if (g->kn_local) {
          a_ij = (F * Kn) / (rho_u[ij1] * 0.5 * hh);
          a_i1j = (F * Kn) / (rho_u[n.i1j1] * 0.5 * hh);
        } else {
          a_ij = (F * Kn) / (0.5 * hh);
          a_i1j = (F * Kn) / (0.5 * hh);
        }
        if (m[n.i_1j1]) u[ij] = (u[ij1] * a_ij + ub) / (1.0 + a_ij);
        if (m[n.i1j1]) u[i1j] = (u[n.i1j1] * a_i1j + ub) / (1.0 + a_i1j);
      }
    }
}

/* S:7215-7331: velocity slip for v (list S:10208-10256). */
void ora_bc_slip_v(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const unsigned *vol = g->vol;
  const double *hx = g->hx;
  const double Kn = g->Kn;
  double *v = f->v;
  const double *rho_v = f->rho_v;
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, i1j = n.i1j, ij1 = n.ij1;
      if (!((0 < vol[ij]) && ((0 == vol[i_1j]) || (0 == vol[i1j])))) continue;
      const ora_body *b = body_of(g, vol[ij]);
      if (!b || !b->is_slip) continue;
      double F = b->F_slip, vb = b->v_body, a_ij, a_ij1;
      if (0 == vol[i_1j]) {
        const double hh = hx[n.i_1];
        if (g->kn_local) {
          a_ij = (F * Kn) / (rho_v[i_1j] * 0.5 * hh);
          a_ij1 = (F * Kn) / (rho_v[n.i_1j1] * 0.5 * hh);
        } else {
          a_ij = (F * Kn) / (0.5 * hh);
          a_ij1 = (F * Kn) / (0.5 * hh);
        }
        if (m[n.i_1j_1]) v[ij] = (v[i_1j] * a_ij + vb) / (1.0 + a_ij);
        if (m[n.i_1j1]) v[ij1] = (v[n.i_1j1] * a_ij1 + vb) / (1.0 + a_ij1);
      } else {
        const double hh = hx[n.i1];
        if (g->kn_local) {
          a_ij = (F * Kn) / (rho_v[i1j] * 0.5 * hh);
          a_ij1 = (F * Kn) / (rho_v[n.i1j1] * 0.5 * hh);
        } else {
          a_ij = (F * Kn) / (0.5 * hh);
          a_ij1 = (F * Kn) / (0.5 * hh);
        }
        if (m[n.i1j_1]) v[ij] = (v[i1j] * a_ij + vb) / (1.0 + a_ij);
        if (m[n.i1j1]) v[ij1] = (v[n.i1j1] * a_ij1 + vb) / (1.0 + a_ij1);
      }
    }
}

/* S:6790-7096: edge copies and inflow/outflow mass conditions for u and v. */
void ora_bc_velocities(const ora_grid *g, ora_fields *f) {
  if (g->periodic_x) return; /* S:6795 */
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx, Ny = (unsigned)g->Ny;
  const double *hx = g->hx, *hy = g->hy;
  const double ht = g->ht;
  double *u = f->u, *v = f->v;
  const double *rho = f->rho, *rho_pr = f->rho_pr, *rho_u = f->rho_u, *rho_v = f->rho_v;
  unsigned i, j, node;
  if (g->dvdx_0_xb)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_v[0] + j * Nx;
      for (i = 0; i < (unsigned)g->Nx_v[0]; i++)
        if (m[node]) v[i + j * Nx] = v[node];
    }
  if (g->dvdx_0_xe)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_v[1] - 1 + j * Nx;
      for (i = (unsigned)g->Nx_v[1]; i < Nx; i++)
        if (m[node]) v[i + j * Nx] = v[node];
    }
  if (g->dvdy_0_yb)
    for (i = 0; i < Nx; i++) {
      node = i + (unsigned)g->Ny_v[0] * Nx;
      for (j = 0; j < (unsigned)g->Ny_v[0]; j++)
        if (m[node]) v[i + j * Nx] = v[node];
    }
  if (g->dvdy_0_ye)
    for (i = 0; i < Nx; i++) {
      node = i + ((unsigned)g->Ny_v[1] - 1) * Nx;
      for (j = (unsigned)g->Ny_v[1]; j < Ny; j++)
        if (m[node]) v[i + j * Nx] = v[node];
    }
  if (g->dudx_0_xb)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_u[0] + j * Nx;
      for (i = 0; i < (unsigned)g->Nx_u[0]; i++)
        if (m[node]) u[i + j * Nx] = u[node];
    }
  if (g->durhodx_0_xb)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_u[0] - 1 + j * Nx;
      if (m[node]) u[node] = (rho_u[node + 1] * u[node + 1]) / rho_u[node];
    }
  if (g->drhodt_xb) {
    i = (unsigned)g->Nx_u[0] - 1;
    for (j = 0; j < Ny; j++) {
      node = i + j * Nx;
      if (m[node])
        u[node] = ((rho[node] - rho_pr[node]) * hx[i] * hy[j] / ht
                   + (rho_v[node + Nx] * v[node + Nx] - rho_v[node] * v[node]) * hx[i]
                   + rho_u[node + 1] * u[node + 1] * hy[j])
                  / (rho_u[node] * hy[j]);
    }
  }
  if (g->dudx_0_xe)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_u[1] - 1 + j * Nx;
      for (i = (unsigned)g->Nx_u[1]; i < Nx; i++)
        if (m[node]) u[i + j * Nx] = u[node];
    }
  if (g->durhodx_0_xe)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_u[1] - 1 + j * Nx;
      if (m[node]) u[node + 1] = (rho_u[node] * u[node]) / rho_u[node + 1];
    }
  if (g->drhodt_xe) {
    i = (unsigned)g->Nx_u[1] - 1;
    for (j = 0; j < Ny; j++) {
      node = i + j * Nx;
      if (m[node])
        u[node + 1] = (-(rho[node] - rho_pr[node]) * hx[i] * hy[j] / ht
                       - (rho_v[node + Nx] * v[node + Nx] - rho_v[node] * v[node]) * hx[i]
                       + rho_u[node] * u[node] * hy[j])
                      / (rho_u[node + 1] * hy[j]);
    }
  }
  if (g->u_Min_Mout_xe) { /* serial build only, S:7031-7059 */
    double Min = 0, Mout = 0;
    for (j = 0; j < Ny; j++) Min = rho_u[(unsigned)g->Nx_u[0] + j * Nx];
    for (j = 0; j < Ny; j++) Mout = rho_u[(unsigned)g->Nx_u[1] + j * Nx];
    for (j = 0; j < Ny; j++) {
      node = (unsigned)g->Nx_u[1] - 1 + j * Nx;
      u[node + 1] = u[node] * Min / Mout;
    }
  }
  if (g->dudy_0_yb)
    for (i = 0; i < Nx; i++) {
      node = i + (unsigned)g->Ny_u[0] * Nx;
      for (j = 0; j < (unsigned)g->Ny_u[0]; j++)
        if (m[node]) u[i + j * Nx] = u[node];
    }
  if (g->dudy_0_ye)
    for (i = 0; i < Nx; i++) {
      node = i + ((unsigned)g->Ny_u[1] - 1) * Nx;
      for (j = (unsigned)g->Ny_u[1]; j < Ny; j++)
        if (m[node]) u[i + j * Nx] = u[node];
    }
}

/* Wall temperature of the fluid seen through a face with body k (S:7762-7797). */
static double wall_T(const ora_grid *g, unsigned k, double T_fluid, double rho_fluid, double half_h,
                     int *written) {
  const ora_body *b = body_of(g, k);
  int is_jump = b ? b->is_jump : 0;
  double dTdn = b ? b->dTdn_on_wall_0 : 0.0, Tb = b ? b->T_body : 0.0;
  *written = 1;
  if (is_jump) {
    double FK = g->kn_local ? (b->F_jump * g->Kn) / rho_fluid : b->F_jump * g->Kn;
    double a = FK / half_h;
    return (a * T_fluid + Tb) / (1.0 + a);
  } else if (dTdn) {
    return T_fluid;
  }
  return Tb;
}

/* S:7744-8045 (list S:10258-10306). */
void ora_bc_temperature_jump(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const unsigned *vol = g->vol;
  const double *hx = g->hx, *hy = g->hy;
  const double *T = f->T, *rho = f->rho;
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij;
      if (!(m[ij] && ((!m[n.i_1j]) || (!m[n.i1j]) || (!m[n.ij_1]) || (!m[n.ij1])))) continue;
      int w;
      if (!m[n.i_1j]) f->Tw_u[ij] = wall_T(g, vol[n.i_1j], T[ij], rho[ij], 0.5 * hx[i], &w);
      if (!m[n.i1j]) f->Tw_u[n.i1j] = wall_T(g, vol[n.i1j], T[ij], rho[ij], 0.5 * hx[i], &w);
      if (!m[n.ij_1]) f->Tw_v[ij] = wall_T(g, vol[n.ij_1], T[ij], rho[ij], 0.5 * hy[j], &w);
      if (!m[n.ij1]) f->Tw_v[n.ij1] = wall_T(g, vol[n.ij1], T[ij], rho[ij], 0.5 * hy[j], &w);
    }
}

/* S:8048-8099 */
void ora_jump_overwrite(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  ora_bc_temperature_jump(g, f);
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij;
      if (!(m[ij] && ((!m[n.i_1j]) || (!m[n.i1j]) || (!m[n.ij_1]) || (!m[n.ij1])))) continue;
      if (!m[n.i_1j]) f->T[n.i_1j] = f->Tw_u[ij];
      if (!m[n.i1j]) f->T[n.i1j] = f->Tw_u[n.i1j];
      if (!m[n.ij_1]) f->T[n.ij_1] = f->Tw_v[ij];
      if (!m[n.ij1]) f->T[n.ij1] = f->Tw_v[n.ij1];
    }
}

static void edge_copy_x(const ora_grid *g, double *a, int b0, int b1, int lo_on, int hi_on) {
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx, Ny = (unsigned)g->Ny;
  unsigned i, j, node;
  if (lo_on)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)b0 + j * Nx;
      for (i = 0; i < (unsigned)b0; i++)
        if (m[node]) a[i + j * Nx] = a[node];
    }
  if (hi_on)
    for (j = 0; j < Ny; j++) {
      node = (unsigned)b1 - 1 + j * Nx;
      for (i = (unsigned)b1; i < Nx; i++)
        if (m[node]) a[i + j * Nx] = a[node];
    }
}

static void edge_copy_y(const ora_grid *g, double *a, int b0, int b1, int lo_on, int hi_on) {
  const unsigned char *m = g->m;
  const unsigned Nx = (unsigned)g->Nx, Ny = (unsigned)g->Ny;
  unsigned i, j, node;
  if (lo_on)
    for (i = 0; i < Nx; i++) {
      node = i + (unsigned)b0 * Nx;
      for (j = 0; j < (unsigned)b0; j++)
        if (m[node]) a[i + j * Nx] = a[node];
    }
  if (hi_on)
    for (i = 0; i < Nx; i++) {
      node = i + ((unsigned)b1 - 1) * Nx;
      for (j = (unsigned)b1; j < Ny; j++)
        if (m[node]) a[i + j * Nx] = a[node];
    }
}

/* S:7650-7741 */
void ora_bc_temperature(const ora_grid *g, ora_fields *f) {
  if (g->periodic_x) return; /* S:7655 */
  edge_copy_x(g, f->T, g->Nx_o[0], g->Nx_o[1], g->dTdx_0_xb, g->dTdx_0_xe);
  edge_copy_y(g, f->T, g->Ny_o[0], g->Ny_o[1], g->dTdy_0_yb, g->dTdy_0_ye);
}

/* S:7335-7428 */
void ora_bc_pressure(const ora_grid *g, ora_fields *f) {
  if (g->periodic_x) return; /* S:7340 */
  edge_copy_x(g, f->p, g->Nx_o[0], g->Nx_o[1], g->dpdx_0_xb, g->dpdx_0_xe);
  edge_copy_y(g, f->p, g->Ny_o[0], g->Ny_o[1], g->dpdy_0_yb, g->dpdy_0_ye);
}

/* A10 */
void ora_fluxes(const ora_grid *g, ora_fields *f) {
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      unsigned node = i + j * (unsigned)g->Nx;
      f->Fx[node] = g->hy[j] * f->rho_u[node] * f->u[node];
      f->Fy[node] = g->hx[i] * f->rho_v[node] * f->v[node];
    }
}

/* A11: S:10457-10618 */
void ora_diffusion(const ora_grid *g, ora_fields *f) {
  const unsigned char *m = g->m;
  const double *hx = g->hx, *hy = g->hy, *sT = f->sqrt_T, *Twu = f->Tw_u, *Twv = f->Tw_v;
  const double cT = g->cT, cT1 = g->cT1;
  for (unsigned j = 0; j < (unsigned)g->Ny; j++)
    for (unsigned i = 0; i < (unsigned)g->Nx; i++) {
      nb_t n;
      nbr(g, i, j, &n);
      const unsigned ij = n.ij, i_1j = n.i_1j, ij_1 = n.ij_1, i_1j_1 = n.i_1j_1, i_1 = n.i_1;
      const unsigned jm = j - 1 * (0 < j);
      if (m[i_1j_1] && m[ij_1] && m[i_1j] && m[ij])
        f->Gamma_yf[ij] = 0.5 * (LI(sT[i_1j_1], sT[i_1j], hy[jm], hy[j]) + LI(sT[ij_1], sT[ij], hy[jm], hy[j]));
      else if (m[i_1j_1] || m[ij_1] || m[i_1j] || m[ij])
        f->Gamma_yf[ij] =
            0.5 * ((double)(m[i_1j_1] && m[i_1j]) * LI(sT[i_1j_1], sT[i_1j], hy[jm], hy[j])
                   + (double)(m[i_1j_1] != m[i_1j]) * sqrt(Twv[i_1j])
                   + (double)(m[ij_1] && m[ij]) * LI(sT[ij_1], sT[ij], hy[jm], hy[j])
                   + (double)(m[ij_1] != m[ij]) * sqrt(Twv[ij])
                   + (double)((!(m[i_1j_1] && m[i_1j])) || (!(m[ij_1] && m[ij])))
                         * LI(sqrt(Twu[ij_1]), sqrt(Twu[ij]), hy[jm], hy[j]));
      f->D_ux[ij] = cT * sT[i_1j] * (hy[j] / hx[i_1]);
      if ((m[ij] && m[i_1j]) || (m[ij_1] && m[i_1j_1]))
        f->D_uy[ij] = cT * f->Gamma_yf[ij]
                      * ((hx[i] + hx[i_1]) / ((double)(m[ij] * m[i_1j]) * hy[j] + (double)(m[ij_1] * m[i_1j_1]) * hy[jm]));
      else
        f->D_uy[ij] = 0;
      if (m[i_1j_1] && m[ij_1] && m[i_1j] && m[ij])
        f->Gamma_xf[ij] = 0.5 * (LI(sT[i_1j_1], sT[ij_1], hx[i_1], hx[i]) + LI(sT[i_1j], sT[ij], hx[i_1], hx[i]));
      else if (m[i_1j_1] || m[ij_1] || m[i_1j] || m[ij])
        f->Gamma_xf[ij] =
            0.5 * ((double)(m[i_1j_1] && m[ij_1]) * LI(sT[i_1j_1], sT[ij_1], hx[i_1], hx[i])
                   + (double)(m[i_1j_1] != m[ij_1]) * sqrt(Twu[ij_1])
                   + (double)(m[i_1j] && m[ij]) * LI(sT[i_1j], sT[ij], hx[i_1], hx[i])
                   + (double)(m[i_1j] != m[ij]) * sqrt(Twu[ij])
                   + (double)((!(m[i_1j_1] && m[ij_1])) || (!(m[i_1j] && m[ij])))
                         * LI(sqrt(Twv[i_1j]), sqrt(Twv[ij]), hx[i_1], hx[i]));
      if ((m[ij] && m[ij_1]) || (m[i_1j] && m[i_1j_1]))
        f->D_vx[ij] = cT * f->Gamma_xf[ij]
                      * ((hy[j] + hy[jm]) / ((double)(m[ij] * m[ij_1]) * hx[i] + (double)(m[i_1j] * m[i_1j_1]) * hx[i_1]));
      else
        f->D_vx[ij] = 0;
      f->D_vy[ij] = cT * sT[ij_1] * (hx[i] / hy[jm]);
      if (m[ij] || m[i_1j])
        f->DTx[ij] = cT1 * (hy[j] / (0.5 * ((double)m[ij] * hx[i] + (double)m[i_1j] * hx[i_1])))
                     * ((double)((!m[ij]) || (!m[i_1j])) * sqrt(Twu[ij])
                        + (double)(m[ij] * m[i_1j]) * HI(sT[i_1j], sT[ij], hx[i_1], hx[i]));
      else
        f->DTx[ij] = 0;
      if (m[ij] || m[ij_1])
        f->DTy[ij] = cT1 * (hx[i] / (0.5 * ((double)m[ij] * hy[j] + (double)m[ij_1] * hy[jm])))
                     * ((double)((!m[ij]) || (!m[ij_1])) * sqrt(Twv[ij])
                        + (double)(m[ij] * m[ij_1]) * HI(sT[ij_1], sT[ij], hy[jm], hy[j]));
      else
        f->DTy[ij] = 0;
    }
}

/* S:6470-6584 (clamped indices in every case, S:6479) */
void ora_apply_bodies(const ora_grid *g, ora_fields *f) {
  const unsigned *vol = g->vol;
  unsigned i, j;
  for (i = 0; i < (unsigned)g->Nx; i++)
    for (j = 0; j < (unsigned)g->Ny; j++) {
      nb_t n;
      nbr_px(g, i, j, 0, &n);
      const ora_body *b = body_of(g, vol[n.ij]);
      if (!b) continue;
      f->u[n.ij] = b->u_body;
      f->u[n.i1j] = b->u_body;
    }
  for (i = 0; i < (unsigned)g->Nx; i++)
    for (j = 0; j < (unsigned)g->Ny; j++) {
      nb_t n;
      nbr_px(g, i, j, 0, &n);
      const ora_body *b = body_of(g, vol[n.ij]);
      if (!b) continue;
      f->v[n.ij] = b->v_body;
      f->v[n.ij1] = b->v_body;
    }
  for (i = 0; i < (unsigned)g->Nx; i++)
    for (j = 0; j < (unsigned)g->Ny; j++) {
      unsigned node = i + j * (unsigned)g->Nx;
      const ora_body *b = body_of(g, vol[node]);
      if (!b) continue;
      f->p[node] = b->p_body;
      f->T[node] = b->T_body;
      if (g->m[node]) f->rho[node] = f->p[node] / f->T[node];
      else f->rho[node] = 0;
    }
}

/* S:10621-10646 (fluxes recomputed first, as the reference does). */
void ora_continuity(const ora_grid *g, const ora_fields *f, double *cont_stat, double *cont) {
  const unsigned Nx = (unsigned)g->Nx;
  for (unsigned j = 0; j < (unsigned)g->Ny - 1; j++)
    for (unsigned i = 0; i < Nx - 1; i++) {
      unsigned ij = i + j * Nx;
      double Fx0 = g->hy[j] * f->rho_u[ij] * f->u[ij];
      double Fx1 = g->hy[j] * f->rho_u[ij + 1] * f->u[ij + 1];
      double Fy0 = g->hx[i] * f->rho_v[ij] * f->v[ij];
      double Fy1 = g->hx[i] * f->rho_v[ij + Nx] * f->v[ij + Nx];
      cont_stat[ij] = Fx1 - Fx0 + Fy1 - Fy0;
      cont[ij] = (f->rho[ij] - f->rho_pr[ij]) * g->hx[i] * g->hy[j] / g->ht + cont_stat[ij];
    }
}

void ora_outer_iteration(const ora_grid *g, ora_fields *f, int n_pc, double *res) {
  double rpt[2 * 64];
  double ruv[2];
  if (n_pc > 64) n_pc = 64;
  ora_assemble_u(g, f);
  ora_assemble_v(g, f);
  ora_energy_source(g, f);
  ora_pt_jacobi(g, f, n_pc, rpt);
  ora_correct_velocities(g, f, ruv);
  ora_properties(g, f);
  ora_rho_in_middle_points(g, f);
  ora_bc_slip_u(g, f);
  ora_bc_slip_v(g, f);
  ora_bc_velocities(g, f);
  ora_bc_temperature_jump(g, f);
  ora_bc_temperature(g, f);
  ora_bc_pressure(g, f);
  ora_fluxes(g, f);
  ora_diffusion(g, f);
  if (res) {
    res[0] = ruv[0];
    res[1] = ruv[1];
    res[2] = rpt[2 * (n_pc - 1)];
    res[3] = rpt[2 * (n_pc - 1) + 1];
  }
}
