/*
 * sspt_oracle.h -- CPU restatement of the SphericalStokes PT iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product library (libsspt.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * Layout: the reference's own Julia layout, column-major with r fastest:
 *   A[ix,iy,iz] (1-based)  ->  a[(ix-1) + n1*((iy-1) + n2*(iz-1))]
 * Array shapes follow scripts_diapir/SphericalStokes_diapir.jl:381-399:
 *   centre  (nr  ,nth  ,nph  )   P, tau_RR, tau_thth, tau_phph, eta, ...
 *   VR      (nr+1,nth  ,nph  )
 *   Vth     (nr  ,nth+1,nph  )
 *   Vph     (nr  ,nth  ,nph+1)
 *   tau_Rth (nr-1,nth-1,nph-2)
 *   tau_Rph (nr-1,nth-2,nph-1)
 *   tau_thph(nr-2,nth-1,nph-1)
 * Coordinates are 1D: R1[nr], th1[nth], ph1[nph]; the reference's 3D
 * coordinate arrays vary along one axis only (SphericalStokes_diapir.jl:374).
 */
#ifndef SSPT_ORACLE_H
#define SSPT_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_SHEAR = 0, ORC_BC_FREE_SLIP = 1, ORC_BC_PURE_SHEAR = 2 };

typedef struct orc_state {
    int nr, nth, nph;
    /* state, caller-owned, reference layout */
    double *P, *DT_R, *tRR, *tTT, *tPP, *tRT, *tRP, *tTP;
    double *VR, *VT, *VP, *VR_err, *VT_err, *VP_err;
    double *eta, *eta_ini, *etaSM, *tII, *rho;
    double *eta_PL;                 /* only used when relax != 1 */
    const double *R1, *th1, *ph1;   /* 1D coordinates */
} orc_state;

typedef struct orc_params {
    double _dr, _dth, _dph;
    double dmp, vpdt, lr;
    double tau_C, n_exp;
    double g;
    double r, sh;
    double relax;           /* 1.0 = reference form (no relaxation) */
    double eta_min, eta_max;/* cutoffs; <=0 / inf disable */
    int bc_mode;            /* ORC_BC_* */
    int err_drop_nan_vph;   /* plateau: ignore NaN err of Vph */
    int etasm_spread;       /* 0 = reference maxloc!; m > 0: spread (orc_spread_etasm) */
} orc_params;

/* --- one function per reference kernel --- */
void orc_equal3(double *VR_err, double *VT_err, double *VP_err,
                const double *VR, const double *VT, const double *VP,
                int nr, int nth, int nph);
void orc_maxloc(double *etaSM, const double *eta, int nr, int nth, int nph);
/* build option etasm_spread = m: etaSM <- three separable passes (r, th, phi) of
 * out(i) = max_{|a|<=m} in(clamp(i+a)) * w[|a|], w[0] = 1, w[a] = w[a-1] * 0.1;
 * tmp is a centre-shaped workspace */
void orc_spread_etasm(double *etaSM, double *tmp, int nr, int nth, int nph, int m);
void orc_copy_bc(double *A, int n1, int n2, int n3);
void orc_compute_P(orc_state *s, const orc_params *p);
void orc_compute_TII(orc_state *s);
void orc_power_law(orc_state *s, const orc_params *p);
void orc_compute_V(orc_state *s, const orc_params *p);
void orc_bc_R(orc_state *s);
void orc_bc_th(orc_state *s, const orc_params *p);
void orc_bc_ph(orc_state *s, const orc_params *p);
/* V_err <- V - V_err; returns err as the reference computes it */
double orc_check_err(orc_state *s, const orc_params *p, double out_comp[3]);

/* --- setup kernels --- */
void orc_initialize_inclusion(double *A, const double *R1, const double *th1,
                              const double *ph1, int nr, int nth, int nph,
                              double r, double radius, double in);
void orc_initialize_velocity(double *Vph, const double *th1, int nr, int nth,
                             int nph, double fact, double r);

/* --- the PT loop body (reference order) and driver --- */
void orc_pt_step(orc_state *s, const orc_params *p, int do_equal3);
/* Runs up to maxiter iterations with the reference exit test every nout.
 * Returns ittot (with the reference's quirk: on break ittot = iter-1),
 * *err_out = last evaluated err, *iter_end = iteration of the break or 0. */
int orc_pt_solve(orc_state *s, const orc_params *p, int maxiter, int nout,
                 double eps, int always_equal3, double *err_out, int *iter_end);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
