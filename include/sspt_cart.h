/*
 * sspt_cart.h -- C ABI of the Cartesian twin of the diapir solver in
 * libsspt.so (gfx950): the PT Stokes iteration of
 * scripts_diapir/CartesianStokes_diapir.jl (CART_3D_Inclusion_Perf), kernel
 * for kernel, for the cross-coordinate check of the spherical solver
 * (a thin spherical sector against the Cartesian box, docs/fig_compare.png).
 *
 * Conventions as in sspt.h: the library owns the device buffers inside an
 * opaque ss_cart context; every call blocks; return codes are the
 * ss_status values of sspt.h (0 ok, > 0 hipError_t, < 0 SS_ERR_*).
 *
 * Field shapes (reference 1-based sizes, CartesianStokes_diapir.jl:317-337):
 *   centre (nx  ,ny  ,nz  )  P TAU_XX TAU_YY TAU_ZZ RHO ETA ETA_INI ETA_PL ETA_PL_OLD ETASM DT_R TII
 *   VX     (nx+1,ny  ,nz  )  VX ERR_VX
 *   VY     (nx  ,ny+1,nz  )  VY ERR_VY
 *   VZ     (nx  ,ny  ,nz+1)  VZ ERR_VZ
 *   TAU_XY (nx-1,ny-1,nz-2), TAU_XZ (nx-1,ny-2,nz-1), TAU_YZ (nx-2,ny-1,nz-1)
 * Layouts: SS_LAYOUT_JULIA (x fastest) or SS_LAYOUT_C ([ix][iy][iz], z fastest).
 */
#ifndef SSPT_CART_H
#define SSPT_CART_H
#include "sspt.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ss_cart ss_cart;

enum ss_cart_field {
    SC_P = 0, SC_TAU_XX, SC_TAU_YY, SC_TAU_ZZ, SC_TAU_XY, SC_TAU_XZ, SC_TAU_YZ,
    SC_VX, SC_VY, SC_VZ, SC_ERR_VX, SC_ERR_VY, SC_ERR_VZ,
    SC_RHO, SC_ETA, SC_ETA_INI, SC_ETA_PL, SC_ETA_PL_OLD, SC_ETASM, SC_DT_R, SC_TII,
    SC_NFIELDS
};

/* Scalars of one iteration, with the reference's names (CartesianStokes_diapir.jl:267-293). */
typedef struct ss_cart_params {
    double _dx, _dy, _dz;   /* inverse grid spacings */
    double dmp, vpdt, lx;   /* damping, PT step (dx*CFL), x extent */
    double tau_C, n_exp;    /* power law; compute_TII!/power_law! run only for n_exp > 1 (:358-363) */
    double relax;           /* log-space relaxation of eta_PL (:217) */
    double g;               /* gravity (acts along z, compute_V! :231-234) */
    double sh;              /* simple-shear rate of the x and y boundaries (:371-379) */
} ss_cart_params;

typedef struct ss_cart_result {
    int iters;      /* iter at the break (:389-391), or the iterations run */
    int converged;  /* err < eps at an iteration > 20 that is a multiple of nout */
    double err;     /* last error of the exit test (NaN before the first check) */
    double seconds; /* wall time of the loop */
} ss_cart_result;

/* Context over an nx x ny x nz cell grid on `device` (every field zero). */
int ss_cart_create(int nx, int ny, int nz, int device, ss_cart **out);
void ss_cart_destroy(ss_cart *c);
const char *ss_cart_last_error(const ss_cart *c);

/* y coordinates of the cell centres, Y_i (:296): the boundary kernels read
 * them as Y_VX / Y_VZ (:299-306), which vary along y only. */
int ss_cart_set_y(ss_cart *c, const double *y);

int ss_cart_upload(ss_cart *c, int field, const double *host, int layout);
int ss_cart_download(ss_cart *c, int field, double *host, int layout);

/* ---- the reference's kernels, one call per @parallel launch ---- */
int ss_cart_equal(ss_cart *c, int dst, int src);                          /* equal!        :41-44   */
int ss_cart_maxloc(ss_cart *c);                                           /* maxloc!(etaSM, eta) :125-128 */
int ss_cart_copy_boundary(ss_cart *c, int field);                         /* copy_boundary! :97-105 */
int ss_cart_timesteps(ss_cart *c, double vpdt, double dmp, double lx);    /* timesteps!    :132-139 */
int ss_cart_compute_P(ss_cart *c, double dmp, double vpdt, double lx,
                      double _dx, double _dy, double _dz);                /* compute_P!    :165-193 */
int ss_cart_compute_TII(ss_cart *c);                                      /* compute_TII!  :195-208 */
int ss_cart_power_law(ss_cart *c, double tau_C, double n_exp, double relax); /* power_law! :210-221 */
int ss_cart_compute_V(ss_cart *c, double _dx, double _dy, double _dz, double g); /* compute_V! :239-255 */
/* boundary kernels on one velocity field (:52-95): fact and the y
 * coordinates enter bc_x_lin! and bc_y_fact! */
int ss_cart_bc_x_lin(ss_cart *c, int field, double fact);
int ss_cart_bc_x_0(ss_cart *c, int field);
int ss_cart_copy_bc_x(ss_cart *c, int field);
int ss_cart_bc_y_fact(ss_cart *c, int field, double fact);
int ss_cart_bc_y_0(ss_cart *c, int field);
int ss_cart_bc_z_0(ss_cart *c, int field);
int ss_cart_copy_bc_z(ss_cart *c, int field);
/* error!(ERR_V, V) for the three components and the exit-test error
 * max over components of maximum(ERR_V) / maximum(|V|) (:257-260, :381-388). */
int ss_cart_check_err(ss_cart *c, double *err, double comp[3]);

/* The reference loop (:349-396): iterations with the exit test every
 * nout (err < eps and iter > 20).  maxiter iterations at most. */
int ss_cart_pt_solve(ss_cart *c, const ss_cart_params *p, int maxiter, int nout, double eps, ss_cart_result *out);
/* niter iterations without checks; *iter_ms = wall time per iteration,
 * *kernel_ms = average time of compute_P! (HIP events on the context's stream). */
int ss_cart_pt_iterate_timed(ss_cart *c, const ss_cart_params *p, int niter, double *iter_ms, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
