/*
 * sspt.h -- C ABI of libsspt.so, the MI355X (gfx950) pseudo-transient (PT)
 * Stokes solver for a staggered spherical-sector grid.
 *
 * It replaces the ParallelStencil kernel calls of the reference scripts
 * (PTsolvers/SphericalStokes) one for one, plus a fused PT loop.  Plain C:
 * no C++ or torch types, no exceptions across the boundary.
 *
 * Ownership: the library owns all device buffers inside an opaque ss_ctx.
 * Host arrays are caller-owned and are copied by ss_upload/ss_download.
 * Every call is blocking (it returns after the GPU work is done), as each
 * `@parallel` call of the reference is.  A context is bound to one device and
 * is not thread-safe; use one context per host thread.
 *
 * Return codes: 0 = SS_OK; > 0 = a hipError_t from the runtime; < 0 = one of
 * the SS_ERR_* codes below.  ss_last_error() returns a message.
 *
 * Field shapes (reference 1-based sizes, scripts_diapir/SphericalStokes_diapir.jl:381-399):
 *   centre  (nr  ,nth  ,nph  )  P DT_R tau_RR tau_thth tau_phph eta eta_ini etaSM tII rho eta_PL
 *   VR      (nr+1,nth  ,nph  )  VR VR_err
 *   Vth     (nr  ,nth+1,nph  )  Vth Vth_err
 *   Vph     (nr  ,nth  ,nph+1)  Vph Vph_err
 *   tau_Rth (nr-1,nth-1,nph-2), tau_Rph (nr-1,nth-2,nph-1), tau_thph (nr-2,nth-1,nph-1)
 */
#ifndef SSPT_H
#define SSPT_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ss_ctx ss_ctx;

enum ss_status {
    SS_OK = 0,
    SS_ERR_ARG = -1,        /* bad size, null pointer, unknown id */
    SS_ERR_NONFINITE = -2,  /* NaN in the convergence error (SphericalStokes_mxpu.jl:426) */
    SS_NOT_CONVERGED = -3,  /* maxiter reached without err < eps */
    SS_ERR_STATE = -4,      /* call order (e.g. no grid set) */
    SS_ERR_NOMEM = -5       /* device allocation failed */
};

enum ss_field {
    SS_P = 0, SS_DT_R, SS_TAU_RR, SS_TAU_THTH, SS_TAU_PHPH,
    SS_TAU_RTH, SS_TAU_RPH, SS_TAU_THPH,
    SS_VR, SS_VTH, SS_VPH, SS_VR_ERR, SS_VTH_ERR, SS_VPH_ERR,
    SS_ETA, SS_ETA_INI, SS_ETASM, SS_TII, SS_RHO, SS_ETA_PL,
    SS_NFIELDS
};

enum ss_layout {
    SS_LAYOUT_JULIA = 0, /* column-major, r fastest: A[ix,iy,iz] of a Julia Array */
    SS_LAYOUT_C = 1      /* row-major [ir][ith][iph], phi fastest */
};

enum ss_bc_mode {
    SS_BC_SHEAR = 0,     /* diapir: simple shear on th/ph faces (SphericalStokes_diapir.jl:61-99) */
    SS_BC_FREE_SLIP = 1, /* plateau: free slip on all faces (SphericalStokes_plateau.jl:64-92) */
    SS_BC_PURE_SHEAR = 2 /* extension: pure shear, normal velocity prescribed on th/ph faces */
};

/* Scalars of one PT iteration, with the reference's parameter names. */
typedef struct ss_params {
    double _dr, _dth, _dph;  /* inverse grid spacings (SphericalStokes_diapir.jl:370) */
    double dmp, vpdt, lr;    /* damping, PT step, radial extent (:367,400) */
    double tau_C, n_exp;     /* power law (:355-356) */
    double g;                /* gravity */
    double r, sh;            /* sector centre radius and boundary shear rate (BCs) */
    double relax;            /* 1.0 = reference power law; <1 log-space relaxation */
    double eta_min, eta_max; /* viscosity cutoffs; 0 / +inf disable */
    int bc_mode;             /* ss_bc_mode */
    int err_drop_nan_vph;    /* plateau convergence test ignores a NaN Vph error */
    int etasm_spread;        /* 0: etaSM = maxloc! + copy_BC! as the reference (SphericalStokes_diapir.jl:149-152).
                                m in 1..8: build option for sharp viscosity contrasts above ~1e3, where the
                                reference's step sizes diverge -- the maxloc'd etaSM is then spread by three
                                separable passes (r, th, phi) of out(i) = max_{|a|<=m} in(clamp(i+a)) * 0.1^|a|,
                                so the pseudo-time steps of neighbouring cells differ by at most 10x near a jump.
                                The converged solution does not depend on it. */
} ss_params;

typedef struct ss_result {
    int ittot;        /* reference's ittot (on break: iter-1, SphericalStokes_diapir.jl:439-441) */
    int iter_end;     /* iteration at which err < eps && iter > 20, else 0 */
    double err;       /* last evaluated err */
    double err_comp[3];
    double seconds;   /* wall time of iterations 11..ittot as the reference times it */
} ss_result;

/* ---- context, grid, data movement ---- */
/* One block of nr x nth x nph cells on one device.  Every field is stored in a
 * box of (nr+1) x (nth+1) x roundup(nph+1, 64) doubles, which must stay below
 * 4 GiB (about 775^3 cells); larger blocks return SS_ERR_ARG. */
int ss_create(int nr, int nth, int nph, int device, ss_ctx **out);
int ss_destroy(ss_ctx *ctx);
const char *ss_last_error(const ss_ctx *ctx);
/* Field dimensions (reference sizes) of a field id. */
int ss_field_dims(const ss_ctx *ctx, int field, int dims[3]);
/* 1D coordinates R_i[nr], th_i[nth], ph_i[nph] (SphericalStokes_diapir.jl:371-373);
 * metric tables are built from them in FP64 on the host. */
int ss_set_grid(ss_ctx *ctx, const double *R1, const double *th1, const double *ph1);
int ss_upload(ss_ctx *ctx, int field, const double *host, int layout);
int ss_download(ss_ctx *ctx, int field, double *host, int layout);
/* A .= value over the field's reference shape (@zeros / @ones / rho0*@ones). */
int ss_fill(ss_ctx *ctx, int field, double value);
/* dst .= src, both of the same shape (eta_ini .= eta, :404). */
int ss_copy_field(ss_ctx *ctx, int dst, int src);

/* ---- setup kernels ---- */
/* initialize_inclusion!(A,R,th,ph,r,radius,in), SphericalStokes_diapir.jl:127-137
 * (compares the squared distance with `radius`, as the reference does). */
int ss_initialize_inclusion(ss_ctx *ctx, int field, double r, double radius, double in);
/* initialize_velocity!(Vph, th, fact, r), SphericalStokes_diapir.jl:139-146 */
int ss_initialize_velocity(ss_ctx *ctx, double fact, double r);

/* ---- per-kernel entry points (one per reference kernel) ---- */
int ss_equal3(ss_ctx *ctx);                                   /* equal3!       :46-58   */
int ss_maxloc(ss_ctx *ctx);                                   /* maxloc!       :149-152 */
/* The etaSM spread of ss_params.etasm_spread = m (1..8) applied to etaSM in
 * place, for the per-kernel sequence after maxloc! + copy_BC!(etaSM).  No
 * reference counterpart (build option); DT_R serves as its workspace, which
 * compute_P! rewrites next in the sequence. */
int ss_spread_etasm(ss_ctx *ctx, int m);
int ss_copy_bc(ss_ctx *ctx, int field);                       /* copy_BC!      :102-124 */
int ss_compute_P(ss_ctx *ctx, double _dr, double _dth, double _dph,
                 double dmp, double vpdt, double lr);         /* compute_P!    :209-241 */
int ss_compute_TII(ss_ctx *ctx);                              /* compute_TII!  :243-257 */
int ss_power_law(ss_ctx *ctx, double tau_C, double n_exp);    /* power_law!    :259-265 */
/* power_law with log-space relaxation (CartesianStokes_diapir.jl:210-221) and cutoffs */
int ss_power_law_ex(ss_ctx *ctx, double tau_C, double n_exp, double relax,
                    double eta_min, double eta_max);
int ss_compute_V(ss_ctx *ctx, double _dr, double _dth, double _dph, double g); /* compute_V! :308-329 */
int ss_bc_R(ss_ctx *ctx);                                     /* BC_R!         :61-69   */
int ss_bc_th(ss_ctx *ctx, int bc_mode, double r, double sh);  /* BC_th!        :71-86 / plateau :74-82 */
int ss_bc_ph(ss_ctx *ctx, int bc_mode, double r, double sh);  /* BC_ph!        :88-99 / plateau :84-92 */
/* check_err! (:332-344) + the host max-reductions of :428-432.
 * err_comp[3] = max|V_err|/max|V| per component; *err = their max. */
int ss_check_err(ss_ctx *ctx, int drop_nan_vph, double err_comp[3], double *err);
/* ss_check_err without a context and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight solver or serving loops).  VR, Vth, Vph, VR_err,
 * Vth_err, Vph_err are caller-owned device arrays of the reference shapes
 * (any element order: the operation is element-wise); V_err <- V - V_err,
 * then out[0..2] = max|V_err|/max|V| per component and out[3] = err, as
 * ss_check_err computes them, into the caller's device buffer out[4].
 * Enqueued on `stream` (a hipStream_t; NULL = the null stream) without
 * synchronising.  The reduction partials live in a 3 MiB static device
 * array per GPU: at most 268 435 456 elements per component (SS_ERR_ARG
 * beyond); calls that overlap on one GPU race on it. */
int ss_check_err_noalloc(const double *VR, const double *Vth, const double *Vph, double *VR_err, double *Vth_err,
                         double *Vph_err, int nr, int nth, int nph, int drop_nan_vph, double *out, void *stream);

/* ---- fused PT loop ---- */
/* Runs `niter` PT iterations with the reference's loop semantics
 * (SphericalStokes_diapir.jl:413-425) through the fused kernels.  All fields,
 * including etaSM, DT_R and tII, equal what the per-kernel sequence leaves. */
int ss_pt_iterate(ss_ctx *ctx, const ss_params *p, int niter);
/* The PT loop with the reference exit test (err every nout, break when
 * err < eps && iter > 20), SphericalStokes_diapir.jl:409-443.
 * Returns SS_OK, SS_NOT_CONVERGED or SS_ERR_NONFINITE. */
int ss_pt_solve(ss_ctx *ctx, const ss_params *p, int maxiter, int nout, double eps,
                ss_result *out);
/* Same loop through the per-kernel entry points (reference launch sequence). */
int ss_pt_solve_unfused(ss_ctx *ctx, const ss_params *p, int maxiter, int nout,
                        double eps, ss_result *out);

/* ---- true residual, time-to-tolerance and the mixed-precision solve ---- */
/* Residual of the current state with tau replaced by the stress consistent
 * with V and eta: out[1] = max|divV|*lr/max|V|, out[2] = max|div(sigma) -
 * rho g| / max|rho g| (or / (max|sigma|/lr) without a body force),
 * out[0] = max(out[1], out[2]).  The state is not modified. */
int ss_residual(ss_ctx *ctx, const ss_params *p, double out[3]);
/* Plain FP64 fused PT iterations until ss_residual <= tol, checked every
 * ncheck iterations.  out->err = last residual, out->iter_end = iterations
 * at convergence (0 if not), out->seconds = wall time.  No reference
 * counterpart: the reference stops on the velocity change (ss_pt_solve). */
int ss_pt_solve_residual(ss_ctx *ctx, const ss_params *p, int maxiter, int ncheck, double tol, ss_result *out);

#define SS_MIXED_MAX_OUTER 64
typedef struct ss_mixed_opts {
    int max_outer;     /* FP64 outer (defect-correction) steps, <= SS_MIXED_MAX_OUTER */
    int max_inner;     /* FP32 PT sweeps per outer step */
    int nout_inner;    /* sweeps between convergence checks of the correction */
    double tol_inner;  /* stop an outer step's sweeps when max|dd|/max|d| < tol_inner */
    double tol;        /* stop when ss_residual of the FP64 iterate <= tol */
} ss_mixed_opts;
typedef struct ss_mixed_result {
    int outer, inner_total, converged;
    int inner_hist[SS_MIXED_MAX_OUTER];
    double r_hist[SS_MIXED_MAX_OUTER + 1];   /* residual before each outer step (and final) */
    double err_inner_hist[SS_MIXED_MAX_OUTER];
    double r, seconds;
    double inner_seconds;   /* part of seconds spent in the FP32 sweeps */
} ss_mixed_result;
/* FP32 PT sweeps inside an FP64 defect-correction loop (no reference
 * counterpart).  Each outer step resets tau to the stress consistent with V
 * and eta, evaluates the FP64 continuity and momentum residuals, runs FP32
 * fused sweeps on the correction (frozen eta, homogeneous BCs, residual
 * forcing), adds the correction in FP64 and, for a power law, updates eta.
 * Converges to the FP64 solution of the same equations. */
int ss_pt_solve_mixed(ss_ctx *ctx, const ss_params *p, const ss_mixed_opts *o, ss_mixed_result *res);

/* ---- adjoint of the converged PT solve (no reference counterpart) ----
 * At a converged state x* = Phi(x*) of the PT iteration Phi, the gradient of
 * J = the volume-weighted mean VR over the r-faces inside the inclusion
 * (face centre (R_f, th, ph) with (R_f - r)^2 + ((th - pi/2) r)^2 + (ph r)^2 <
 * radius, the squared-distance test of initialize_inclusion!, weight
 * R_f^2 sin(th)) with respect to the per-cell rho and eta_ini is
 * (dPhi/dm)^T lam with lam = T^T lam + dJ/dx, T = dPhi/dx at x* (step sizes
 * held at x*, exact at the fixed point).  lam is iterated with the transposed
 * PT map (the forward iteration's contraction); only x* is used, no record
 * of the forward iterations.  Call after ss_pt_iterate/ss_pt_solve (their
 * last iteration leaves etaSM, DT_R and tII of the final state). */
enum ss_adj_field {
    SS_ADJ_VR = 0, SS_ADJ_VTH, SS_ADJ_VPH, SS_ADJ_P, SS_ADJ_TAU_RR, SS_ADJ_TAU_THTH, SS_ADJ_TAU_PHPH,
    SS_ADJ_TAU_RTH, SS_ADJ_TAU_RPH, SS_ADJ_TAU_THPH, SS_ADJ_ETA, SS_ADJ_ETA_PL,
    SS_ADJ_GRAD_RHO,      /* dJ/drho (centre shape) */
    SS_ADJ_GRAD_ETA_INI,  /* dJ/deta_ini (centre shape); d/dlog10(eta_ini) = this * eta_ini * ln 10 */
    SS_ADJ_NFIELDS
};
typedef struct ss_adjoint_result {
    int iters, converged;
    double J;        /* objective at the state */
    double err;      /* last max|dlam|/max|lam| over the velocity components */
    double seconds;
} ss_adjoint_result;
/* Coefficients of T at the current state (allocates the adjoint boxes on first use). */
int ss_adjoint_prepare(ss_ctx *ctx, const ss_params *p);
/* lam = 0 and gradients = 0. */
int ss_adjoint_zero(ss_ctx *ctx);
/* J at the current state; keeps dJ/dVR for ss_adjoint_solve. */
int ss_adjoint_objective(ss_ctx *ctx, double r, double radius, double *J);
int ss_adjoint_upload(ss_ctx *ctx, int adj_field, const double *host, int layout);
int ss_adjoint_download(ss_ctx *ctx, int adj_field, double *host, int layout);
/* lam <- T^T lam; with gradients != 0 also grad += (dPhi/dm)^T lam (before the step). */
int ss_adjoint_step(ss_ctx *ctx, const ss_params *p, int gradients);
/* prepare + objective + lam = T^T lam + dJ/dx until max|dlam|/max|lam| < tol
 * (checked every ncheck iterations) + gradients.  SS_NOT_CONVERGED if maxiter
 * was reached (the gradients are those of the last lam). */
int ss_adjoint_solve(ss_ctx *ctx, const ss_params *p, double r, double radius, int maxiter, int ncheck, double tol,
                     ss_adjoint_result *out);
/* The same solve with FP32 sweeps inside an FP64 defect correction (no
 * reference counterpart; the counterpart of ss_pt_solve_mixed).  It iterates
 * the BC-folded lam (mu = F lam, F the transpose of the BC updates; the same
 * fixed point).  Each outer step applies one FP64 step Y = F(T^T mu + dJ/dx)
 * and stops when max|Y - mu|/max|Y| over the velocity components < o->tol;
 * otherwise it iterates the correction d = F T^T d + (Y - mu) with the
 * r-marching kernel in FP32 (state, tables and d rounded to FP32) until
 * max|dd|/max|d| < the inner tolerance (checked every o->nout_inner sweeps)
 * or o->max_inner sweeps, then mu += F d in FP64.  The inner tolerance is
 * o->tol_inner, loosened (to at most 1e-2) when the reduction still needed
 * is smaller than the last outer step's.  Gradients as ss_adjoint_solve.
 * Returns SS_OK, SS_NOT_CONVERGED (o->max_outer steps) or SS_ERR_NONFINITE
 * (a non-finite adjoint value at an FP64 step or an FP32 check).
 * Allocates 49 FP32 boxes on first use. */
typedef struct ss_adjoint_mixed_result {
    int outer, inner_total, converged;
    double J;              /* objective at the state */
    double err;            /* max|Y - mu|/max|Y| of the last outer step */
    double seconds;
    double inner_seconds;  /* part of seconds spent in the FP32 sweeps */
} ss_adjoint_mixed_result;
int ss_adjoint_solve_mixed(ss_ctx *ctx, const ss_params *p, double r, double radius, const ss_mixed_opts *o,
                           ss_adjoint_mixed_result *out);

/* ---- timing hooks for the benchmark (no reference counterpart) ---- */
/* Average device time (ms) per iteration of the last ss_pt_iterate, from HIP
 * events on the context's stream. */
int ss_last_kernel_ms(const ss_ctx *ctx, double *ms_per_iter);
/* ss_pt_iterate with HIP events around every launch of the r-marching main
 * kernel: *main_ms = its average duration, *iter_ms = average time of a whole
 * iteration (main kernel + boundary pass). */
int ss_pt_iterate_timed(ss_ctx *ctx, const ss_params *p, int niter, double *main_ms, double *iter_ms);

#ifdef __cplusplus
}
#endif
#endif
