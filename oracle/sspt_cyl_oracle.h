/*
 * sspt_cyl_oracle.h -- CPU restatement of scripts_diapir/CylindricalStokes_diapir.jl.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the cylindrical entry
 * points of libsspt.so.  Only tests/ load it.
 *
 * Layout: the reference's Julia layout (r fastest); shapes as
 * CylindricalStokes_diapir.jl:278-318: centre (nr,nth,nz), VR/VR_OLD
 * (nr+1,nth,nz), Vth/Vth_OLD (nr,nth+1,nz), VZ/VZ_OLD (nr,nth,nz+1),
 * dVR (nr-1,nth-2,nz-2), dVth (nr-2,nth-1,nz-2), dVZ (nr-2,nth-2,nz-1),
 * D_Rth/tau_Rth/ERR_Rth (nr-1,nth-1,nz-2), D_RZ/tau_RZ (nr-1,nth-2,nz-1),
 * D_thZ/tau_thZ (nr-2,nth-1,nz-1).
 */
#ifndef SSPT_CYL_ORACLE_H
#define SSPT_CYL_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* field order of the arrays in orc_cyl.f */
enum {
    OCY_P, OCY_DIVV, OCY_VR, OCY_VT, OCY_VZ, OCY_dVR, OCY_dVT, OCY_dVZ, OCY_VR_OLD, OCY_VT_OLD, OCY_VZ_OLD,
    OCY_D_RR, OCY_D_TT, OCY_D_ZZ, OCY_D_RT, OCY_D_RZ, OCY_D_TZ,
    OCY_tRR, OCY_tTT, OCY_tZZ, OCY_tRT, OCY_tRZ, OCY_tTZ, OCY_sRR, OCY_sTT, OCY_sZZ,
    OCY_rho, OCY_rhoG, OCY_eta, OCY_eta_INI, OCY_eta_PL, OCY_eta_PL_OLD, OCY_etaSM, OCY_KBDT, OCY_GSDT, OCY_DT_R,
    OCY_tII, OCY_ERR_RT, OCY_N
};

enum { OCY_COPY_BC_R = 0, OCY_COPY_BC_TH, OCY_COPY_BC_Z, OCY_BC_R_0, OCY_BC_TH_FACT, OCY_BC_TH_0, OCY_BC_Z_0,
       OCY_BC_Z_LIN };

typedef struct orc_cyl {
    int nr, nth, nz;
    double *f[OCY_N];
    const double *R1, *th1;   /* R_i, th_i (the 3D coordinate arrays vary along one axis) */
} orc_cyl;

typedef struct orc_cyl_params {
    double dr, dth, dz, dmp, vpdt, lr, tau_C, n_exp, n_exp_PL, relax, g, sh, r;
} orc_cyl_params;

void orc_cyl_dims(const orc_cyl *s, int field, int d[3]);
void orc_cyl_equal(orc_cyl *s, int dst, int src);
void orc_cyl_multiply(orc_cyl *s, int dst, int src, double fact);
void orc_cyl_maxloc(orc_cyl *s);
void orc_cyl_timesteps(orc_cyl *s, double dmp, double vpdt, double lr);
void orc_cyl_timesteps2(orc_cyl *s, double vpdt);
void orc_cyl_compute_P(orc_cyl *s, double dr, double dth, double dz);
void orc_cyl_compute_sigma(orc_cyl *s);
void orc_cyl_compute_TII(orc_cyl *s);
void orc_cyl_power_law(orc_cyl *s, double tau_C, double n_exp, double relax);
void orc_cyl_compute_dV(orc_cyl *s, double dr, double dth, double dz);
void orc_cyl_compute_V(orc_cyl *s);
void orc_cyl_bc(orc_cyl *s, int field, int op, double fact, double r);
void orc_cyl_err_rt(orc_cyl *s);
double orc_cyl_exit_error(orc_cyl *s, double comp[3]);
/* one iteration of the loop body (:332-362); n_exp is the exponent in force */
void orc_cyl_iteration(orc_cyl *s, const orc_cyl_params *p, double n_exp);
/* the loop with the exponent switch and the exit test (:329-372) */
int orc_cyl_pt_solve(orc_cyl *s, const orc_cyl_params *p, int maxiter, int nout, double eps, double *err,
                     int *converged);

#ifdef __cplusplus
}
#endif
#endif
