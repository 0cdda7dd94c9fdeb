/*
 * sspt_cyl.h -- C ABI of the cylindrical twin of the diapir solver in
 * libsspt.so (gfx950): the PT Stokes iteration of
 * scripts_diapir/CylindricalStokes_diapir.jl (CYL_3D_Inclusion), kernel for
 * kernel, for the cross-coordinate check (a thin cylindrical shell against
 * the Cartesian box and the thin spherical sector, docs/fig_compare.png).
 *
 * Conventions as in sspt.h (blocking calls, ss_status return codes, opaque
 * context, SS_LAYOUT_JULIA = r fastest or SS_LAYOUT_C).  Field shapes
 * (reference 1-based sizes, CylindricalStokes_diapir.jl:278-317):
 *   centre  (nr  ,nth  ,nz  )  P DIVV D_RR D_THTH D_ZZ TAU_RR TAU_THTH TAU_ZZ SIG_RR SIG_THTH SIG_ZZ
 *                               RHO RHOG ETA ETA_INI ETA_PL ETA_PL_OLD ETASM KBDT GSDT DT_R TII
 *   VR, VR_OLD (nr+1,nth,nz); VTH, VTH_OLD (nr,nth+1,nz); VZ, VZ_OLD (nr,nth,nz+1)
 *   DVR (nr-1,nth-2,nz-2), DVTH (nr-2,nth-1,nz-2), DVZ (nr-2,nth-2,nz-1)
 *   D_RTH, TAU_RTH, ERR_RTH (nr-1,nth-1,nz-2); D_RZ, TAU_RZ (nr-1,nth-2,nz-1);
 *   D_THZ, TAU_THZ (nr-2,nth-1,nz-1)
 */
#ifndef SSPT_CYL_H
#define SSPT_CYL_H
#include "sspt.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ss_cyl ss_cyl;

enum ss_cyl_field {
    SY_P = 0, SY_DIVV, SY_VR, SY_VTH, SY_VZ, SY_DVR, SY_DVTH, SY_DVZ, SY_VR_OLD, SY_VTH_OLD, SY_VZ_OLD,
    SY_D_RR, SY_D_THTH, SY_D_ZZ, SY_D_RTH, SY_D_RZ, SY_D_THZ,
    SY_TAU_RR, SY_TAU_THTH, SY_TAU_ZZ, SY_TAU_RTH, SY_TAU_RZ, SY_TAU_THZ, SY_SIG_RR, SY_SIG_THTH, SY_SIG_ZZ,
    SY_RHO, SY_RHOG, SY_ETA, SY_ETA_INI, SY_ETA_PL, SY_ETA_PL_OLD, SY_ETASM, SY_KBDT, SY_GSDT, SY_DT_R,
    SY_TII, SY_ERR_RTH,
    SY_NFIELDS
};

/* boundary kernels (CylindricalStokes_diapir.jl:52-101) */
enum ss_cyl_bc { SY_COPY_BC_R = 0, SY_COPY_BC_TH, SY_COPY_BC_Z, SY_BC_R_0, SY_BC_TH_FACT, SY_BC_TH_0, SY_BC_Z_0,
                 SY_BC_Z_LIN };

/* Scalars with the reference's names (:224-256); the kernels divide by the
 * spacings dr, dθ, dz. */
typedef struct ss_cyl_params {
    double dr, dth, dz, dmp, vpdt, lr;
    double tau_C, n_exp, n_exp_PL;  /* n_exp until the exit-test error < 1e-5, then n_exp_PL (:342-344) */
    double relax, g, sh, r;
} ss_cyl_params;

typedef struct ss_cyl_result {
    int iters;       /* iter at the break (:375-377), or the iterations run */
    int converged;
    double err;      /* last exit-test error */
    double n_exp;    /* exponent in force at the end */
    double seconds;
} ss_cyl_result;

int ss_cyl_create(int nr, int nth, int nz, int device, ss_cyl **out);
void ss_cyl_destroy(ss_cyl *c);
const char *ss_cyl_last_error(const ss_cyl *c);
/* R_i[nr] and th_i[nth] (:246-248): the 3D coordinate arrays R, theta_VZ,
 * theta_VR of the kernels vary along one axis only */
int ss_cyl_set_grid(ss_cyl *c, const double *R, const double *th);
int ss_cyl_upload(ss_cyl *c, int field, const double *host, int layout);
int ss_cyl_download(ss_cyl *c, int field, double *host, int layout);

/* ---- the reference's kernels, one call per @parallel launch ---- */
int ss_cyl_equal(ss_cyl *c, int dst, int src);                              /* equal!      :41-44   */
int ss_cyl_multiply(ss_cyl *c, int dst, int src, double fact);              /* multiply!   :46-49   */
int ss_cyl_maxloc(ss_cyl *c);                                               /* maxloc!     :125-128 */
int ss_cyl_timesteps(ss_cyl *c, double dmp, double vpdt, double lr);        /* timesteps!  :130-135 */
int ss_cyl_timesteps2(ss_cyl *c, double vpdt);                              /* timesteps2! :137-141 */
int ss_cyl_compute_P(ss_cyl *c, double dr, double dth, double dz);          /* compute_P!  :144-167 */
int ss_cyl_compute_sigma(ss_cyl *c);                                        /* compute_sigma! :169-175 */
int ss_cyl_compute_TII(ss_cyl *c);                                          /* compute_TII! :177-182 */
int ss_cyl_power_law(ss_cyl *c, double tau_C, double n_exp, double relax);  /* power_law!  :184-191 */
int ss_cyl_compute_dV(ss_cyl *c, double dr, double dth, double dz);         /* compute_dV! :193-206 */
int ss_cyl_compute_V(ss_cyl *c);                                            /* compute_V!  :208-214 */
int ss_cyl_bc(ss_cyl *c, int field, int op, double fact, double r);         /* :52-101 (ss_cyl_bc ops) */
int ss_cyl_err_rth(ss_cyl *c);                                              /* err_rθ!     :217-220 */
/* the exit test's error (:369-372): max over the components of
 * maximum(|V - V_OLD|) / maximum(|V|) */
int ss_cyl_exit_error(ss_cyl *c, double *err, double comp[3]);

/* the reference loop (:329-384) */
int ss_cyl_pt_solve(ss_cyl *c, const ss_cyl_params *p, int maxiter, int nout, double eps, ss_cyl_result *out);
/* niter iterations with exponent p->n_exp, no checks; per-iteration wall
 * time and the average time of compute_P! (HIP events) */
int ss_cyl_pt_iterate_timed(ss_cyl *c, const ss_cyl_params *p, int niter, double *iter_ms, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
