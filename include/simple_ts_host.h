/*
 * simple_ts_host.h -- host-side setup of a SIMPLE-TS case (libsimple_ts_host.so).
 *
 * Replaces the setup the reference runs once in main() before loop 1
 * (SIMPLE_TS_MPI.cpp, "S:n"): reading EnteredData.txt (ReadEnteredDataFromFile,
 * S:5241-5561) and the body files (Objects_for_Solving.h:306-412 and the
 * circle reader before it), the non-dimensional constants (S:2013-2071), the
 * mesh (S:2078-2248, define_Nx_Ny S:4955-5041, hx_all and friends
 * S:4736-4952), the body mask (define_vol_inf S:6306-6466 with
 * Polyhedrons_Objects::is_point_inside_polyhedron, Polyhedrons_Objects.cpp:190,
 * and CCircles_Objects::is_point_inside_circle, Circles_Objects.h:118) and the
 * solve boxes (define_area_for_solving S:6009-6267).  Host memory only; no
 * GPU is needed.  Errors: negative return, message in sts_host_last_error().
 */
#ifndef SIMPLE_TS_HOST_H
#define SIMPLE_TS_HOST_H

#include "simple_ts.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sts_case sts_case;

const char *sts_host_last_error(void);

/* Reads dir/EnteredData.txt and the body files it asks for, builds mesh,
   mask and boxes.  Fails (STS_ERR_IO) on a missing or unparsable file, where
   the reference prints a message and continues. */
int sts_case_load(const char *dir, sts_case **out);
void sts_case_free(sts_case *c);

int sts_case_dims(const sts_case *c, int *Nx, int *Ny);

/* Fills g with the case's grid; its pointers stay valid while c lives. */
int sts_case_grid(const sts_case *c, sts_grid *g);

/* Copies of the mesh arrays: x_f[Nx+1], y_f[Ny+1], hx[Nx], hy[Ny], x_v[Nx],
   y_v[Ny] (any pointer may be NULL). */
int sts_case_mesh(const sts_case *c, double *x_f, double *y_f, double *hx, double *hy, double *x_v, double *y_v);

/* vol_inf_fb[Nx*Ny] (copy). */
int sts_case_vol(const sts_case *c, unsigned int *vol);

/* Value of an EnteredData.txt entry (or a derived constant: cT, cT1, cT2,
   cPch, c_Fr_v, u_given_xb, rho_gas_nd) by the reference's variable name;
   NaN if unknown. */
double sts_case_value(const sts_case *c, const char *name);

/* Writes EnteredData.txt in the reference's format (WriteEnteredDataToFile,
   S:5044-5238): one "value<TAB>//comment" line per entry, values with 15
   significant digits; comments are the ones read from the input file. */
int sts_case_write_config(const sts_case *c, const char *path);

#ifdef __cplusplus
}
#endif
#endif
