/*
 * sspt_cart_oracle.h -- CPU restatement of scripts_diapir/CartesianStokes_diapir.jl.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the Cartesian entry
 * points of libsspt.so.  Only tests/ and the checker legs of bench.py load it.
 *
 * Layout: the reference's Julia layout, x fastest; shapes as
 * CartesianStokes_diapir.jl:317-337 (centre (nx,ny,nz), VX (nx+1,ny,nz),
 * VY (nx,ny+1,nz), VZ (nx,ny,nz+1), tau_XY (nx-1,ny-1,nz-2),
 * tau_XZ (nx-1,ny-2,nz-1), tau_YZ (nx-2,ny-1,nz-1)).
 */
#ifndef SSPT_CART_ORACLE_H
#define SSPT_CART_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CART_BC_X_LIN = 0, ORC_CART_BC_X_0, ORC_CART_COPY_BC_X, ORC_CART_BC_Y_FACT, ORC_CART_BC_Y_0,
       ORC_CART_BC_Z_0, ORC_CART_COPY_BC_Z };

typedef struct orc_cart {
    int nx, ny, nz;
    double *P, *tXX, *tYY, *tZZ, *tXY, *tXZ, *tYZ;
    double *VX, *VY, *VZ, *ERR_VX, *ERR_VY, *ERR_VZ;
    double *rho, *eta, *eta_ini, *eta_PL, *eta_PL_old, *etaSM, *DT_R, *tII;
    const double *y1;   /* y coordinates of the cell centres (Y_i) */
} orc_cart;

typedef struct orc_cart_params {
    double _dx, _dy, _dz, dmp, vpdt, lx, tau_C, n_exp, relax, g, sh;
} orc_cart_params;

void orc_cart_equal(double *A, const double *B, size_t n);
void orc_cart_maxloc(orc_cart *s);
void orc_cart_copy_boundary(double *A, int n1, int n2, int n3);
void orc_cart_timesteps(orc_cart *s, double vpdt, double dmp, double lx);
void orc_cart_compute_P(orc_cart *s, double dmp, double vpdt, double lx, double _dx, double _dy, double _dz);
void orc_cart_compute_TII(orc_cart *s);
void orc_cart_power_law(orc_cart *s, double tau_C, double n_exp, double relax);
void orc_cart_compute_V(orc_cart *s, double _dx, double _dy, double _dz, double g);
void orc_cart_bc(double *A, int n1, int n2, int n3, int op, const double *y, double fact);
double orc_cart_check_err(orc_cart *s, double comp[3]);
void orc_cart_iteration(orc_cart *s, const orc_cart_params *p, int check);
int orc_cart_pt_solve(orc_cart *s, const orc_cart_params *p, int maxiter, int nout, double eps, double *err,
                      int *converged);

#ifdef __cplusplus
}
#endif
#endif
