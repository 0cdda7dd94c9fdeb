/*
 * sspt_oracle.c -- CPU restatement of the SphericalStokes PT iteration.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product library never links or calls this file.
 *
 * Every function restates one reference kernel and keeps its launch range,
 * guards, in-place update and floating-point association (Julia evaluates
 * a*b*c and a+b+c left to right; macros substitute whole subexpressions).
 * Build with -ffp-contract=off -fno-fast-math so no FMA changes a rounding.
 *
 * Indexing helpers take the reference's 1-based (ix,iy,iz) so that each line
 * can be checked against the Julia source it follows.
 */
#include "sspt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1-based access into a column-major (n1,n2,n3) array */
#define IDX(n1, n2, i, j, k) (((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1))

static const double PI = 3.14159265358979323846; /* Float64(pi) */

/* Julia's max/min for Float64 (NaN-propagating, +0 > -0),
 * Base.max(x::T, y::T) where T<:AbstractFloat */
static inline double jl_max(double x, double y)
{
    int pick_y = (y > x) || (signbit(y) < signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}
static inline double jl_min(double x, double y)
{
    int pick_y = (y < x) || (signbit(y) > signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the parallel loops from now on (tiny grids run faster on one) */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* SphericalStokes_diapir.jl:46-58 */
void orc_equal3(double *VR_err, double *VT_err, double *VP_err,
                const double *VR, const double *VT, const double *VP,
                int nr, int nth, int nph)
{
    size_t nVR = (size_t)(nr + 1) * nth * nph;
    size_t nVT = (size_t)nr * (nth + 1) * nph;
    size_t nVP = (size_t)nr * nth * (nph + 1);
    memcpy(VR_err, VR, nVR * sizeof(double));
    memcpy(VT_err, VT, nVT * sizeof(double));
    memcpy(VP_err, VP, nVP * sizeof(double));
}

/* SphericalStokes_diapir.jl:149-152 (@inn(etaSM) = @maxloc(eta));
 * the 7-point expansion and its max() nesting follow
 * scripts_perf/SphericalStokes_mxpu.jl:295-297 */
void orc_maxloc(double *etaSM, const double *eta, int nr, int nth, int nph)
{
    const int n1 = nr, n2 = nth;
#pragma omp parallel for schedule(static)
    for (int iz = 2; iz <= nph - 1; ++iz)
        for (int iy = 2; iy <= nth - 1; ++iy)
            for (int ix = 2; ix <= nr - 1; ++ix) {
                double m = jl_max(jl_max(jl_max(jl_max(eta[IDX(n1, n2, ix - 1, iy, iz)],
                                                       eta[IDX(n1, n2, ix + 1, iy, iz)]),
                                                eta[IDX(n1, n2, ix, iy, iz)]),
                                         jl_max(eta[IDX(n1, n2, ix, iy - 1, iz)],
                                                eta[IDX(n1, n2, ix, iy + 1, iz)])),
                                  jl_max(eta[IDX(n1, n2, ix, iy, iz - 1)],
                                         eta[IDX(n1, n2, ix, iy, iz + 1)]));
                etaSM[IDX(n1, n2, ix, iy, iz)] = m;
            }
}

/* Build option without a reference counterpart (etasm_spread = m): after
 * maxloc! + copy_BC!, etaSM is spread by three separable passes r, th, phi
 * of out(i) = max over |a| <= m of in(clamp(i + a)) * w[|a|], w[0] = 1,
 * w[a] = w[a-1] * 0.1, so that the pseudo-time steps of neighbouring cells
 * differ by at most a factor 10 within m cells of a stiff cell.  Pinned by
 * brute force (tests/test_etasm_spread.py), not by a reference. */
static inline int clampi_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
void orc_spread_etasm(double *A, double *tmp, int nr, int nth, int nph, int m)
{
    double w[9];
    w[0] = 1.0;
    for (int a = 1; a <= 8; ++a) w[a] = w[a - 1] * 0.1;
    const int n[3] = {nr, nth, nph};
    for (int axis = 0; axis < 3; ++axis) {
        const double *src = axis == 1 ? tmp : A;
        double *dst = axis == 1 ? A : tmp;
#pragma omp parallel for schedule(static)
        for (int iz = 1; iz <= nph; ++iz)
            for (int iy = 1; iy <= nth; ++iy)
                for (int ix = 1; ix <= nr; ++ix) {
                    int idx[3] = {ix, iy, iz};
                    const int c = idx[axis];
                    double v = src[IDX(nr, nth, ix, iy, iz)];
                    for (int a = 1; a <= m; ++a) {
                        idx[axis] = clampi_(c - a, 1, n[axis]);
                        v = jl_max(v, src[IDX(nr, nth, idx[0], idx[1], idx[2])] * w[a]);
                        idx[axis] = clampi_(c + a, 1, n[axis]);
                        v = jl_max(v, src[IDX(nr, nth, idx[0], idx[1], idx[2])] * w[a]);
                    }
                    dst[IDX(nr, nth, ix, iy, iz)] = v;
                }
    }
    memcpy(A, tmp, (size_t)nr * nth * nph * sizeof(double));
}

/* copy_bc_r!, copy_bc_th!, copy_bc_ph! in that order,
 * SphericalStokes_diapir.jl:102-124 */
void orc_copy_bc(double *A, int n1, int n2, int n3)
{
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= n3; ++iz)
        for (int iy = 1; iy <= n2; ++iy) {
            A[IDX(n1, n2, 1, iy, iz)] = A[IDX(n1, n2, 2, iy, iz)];
            A[IDX(n1, n2, n1, iy, iz)] = A[IDX(n1, n2, n1 - 1, iy, iz)];
        }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= n3; ++iz)
        for (int ix = 1; ix <= n1; ++ix) {
            A[IDX(n1, n2, ix, 1, iz)] = A[IDX(n1, n2, ix, 2, iz)];
            A[IDX(n1, n2, ix, n2, iz)] = A[IDX(n1, n2, ix, n2 - 1, iz)];
        }
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy <= n2; ++iy)
        for (int ix = 1; ix <= n1; ++ix) {
            A[IDX(n1, n2, ix, iy, 1)] = A[IDX(n1, n2, ix, iy, 2)];
            A[IDX(n1, n2, ix, iy, n3)] = A[IDX(n1, n2, ix, iy, n3 - 1)];
        }
}

/* compute_P!, SphericalStokes_diapir.jl:209-241 with the macros of :155-208.
 * R[ix,iy,iz] = R1[ix], th[ix,iy,iz] = th1[iy]; cot(x) = 1/tan(x) as in
 * Julia's Base.cot. */
void orc_compute_P(orc_state *s, const orc_params *p)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *R1 = s->R1 - 1, *T1 = s->th1 - 1; /* 1-based views */
    const double _dr = p->_dr, _dth = p->_dth, _dph = p->_dph;
    /* @KBDT = dmp*2.0*pi*vpdt/lr*etaSM ; @GSDT = 4.0*pi*vpdt/lr*etaSM */
    const double kb = (((p->dmp * 2.0) * PI) * p->vpdt) / p->lr;
    const double gs = ((4.0 * PI) * p->vpdt) / p->lr;
    const double vpdt2 = p->vpdt * p->vpdt; /* vpdt^2 (literal power) */
    const double third = 1.0 / 3.0;
    double *P = s->P, *DT = s->DT_R, *tRR = s->tRR, *tTT = s->tTT, *tPP = s->tPP;
    double *tRT = s->tRT, *tRP = s->tRP, *tTP = s->tTP;
    const double *VR = s->VR, *VT = s->VT, *VP = s->VP, *eta = s->eta, *SM = s->etaSM;
#define C_(A, i, j, k) A[IDX(nr, nt, i, j, k)]
#define VR_(i, j, k) VR[IDX(nr + 1, nt, i, j, k)]
#define VT_(i, j, k) VT[IDX(nr, nt + 1, i, j, k)]
#define VP_(i, j, k) VP[IDX(nr, nt, i, j, k)]
#define GS_(i, j, k) (gs * SM[IDX(nr, nt, i, j, k)])
    /* centre block: ix<=nr, iy<=nth, iz<=nph */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            const double th = T1[iy];
            const double sn = sin(th), ct = 1.0 / tan(th); /* sin(th), cot(th) */
            for (int ix = 1; ix <= nr; ++ix) {
                const double R = R1[ix];
                const double dVRr = (VR_(ix + 1, iy, iz) - VR_(ix, iy, iz)) * _dr;
                const double dVTt = ((1.0 / R) * (VT_(ix, iy + 1, iz) - VT_(ix, iy, iz))) * _dth;
                const double dVPp = (((1.0 / R) / sn) * (VP_(ix, iy, iz + 1) - VP_(ix, iy, iz))) * _dph;
                const double avVR = (VR_(ix, iy, iz) + VR_(ix + 1, iy, iz)) * 0.5;  /* @avxa_VR */
                const double avVT = (VT_(ix, iy, iz) + VT_(ix, iy + 1, iz)) * 0.5;  /* @avya_Vth */
                /* @DIVV */
                const double divV = (((dVRr + dVTt) + dVPp) + (2.0 * avVR) / R) + (avVT / R) * ct;
                const double Drr = dVRr - third * divV;                            /* @DRR */
                const double Dtt = (dVTt + avVR / R) - third * divV;               /* @Dthth */
                const double Dpp = ((dVPp + avVR / R) + (avVT / R) * ct) - third * divV; /* @Dphph */
                const size_t c = IDX(nr, nt, ix, iy, iz);
                const double KB = kb * SM[c], GS = gs * SM[c];
                P[c] = P[c] - KB * divV;
                tRR[c] = (tRR[c] + (GS * 2.0) * Drr) / (1.0 + GS / eta[c]);
                tTT[c] = (tTT[c] + (GS * 2.0) * Dtt) / (1.0 + GS / eta[c]);
                tPP[c] = (tPP[c] + (GS * 2.0) * Dpp) / (1.0 + GS / eta[c]);
                DT[c] = vpdt2 / (KB + GS / (1.0 + GS / eta[c]));
            }
        }
    /* tau_Rth: (nr-1, nth-1, nph-2) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rxy = (((R1[ix] + R1[ix]) + R1[ix + 1]) + R1[ix + 1]) * 0.25; /* @avxyi_R */
                const double G = (((GS_(ix, iy, iz + 1) + GS_(ix, iy + 1, iz + 1)) + GS_(ix + 1, iy, iz + 1)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix, iy, iz + 1) + C_(eta, ix, iy + 1, iz + 1)) +
                                   C_(eta, ix + 1, iy, iz + 1)) + C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double avVT = (VT_(ix, iy + 1, iz + 1) + VT_(ix + 1, iy + 1, iz + 1)) * 0.5; /* @avxi_Vth */
                const double D = ((VT_(ix + 1, iy + 1, iz + 1) - VT_(ix, iy + 1, iz + 1)) * _dr +
                                  ((1.0 / Rxy) * (VR_(ix + 1, iy + 1, iz + 1) - VR_(ix + 1, iy, iz + 1))) * _dth) -
                                 avVT / Rxy;
                const size_t e = IDX(nr - 1, nt - 1, ix, iy, iz);
                tRT[e] = (tRT[e] + G * D) / (1.0 + G / E);
            }
    /* tau_Rph: (nr-1, nth-2, nph-1) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy) {
            /* th-only terms evaluated once per (iy, iz) row: the same values */
            const double Txz = (((T1[iy + 1] + T1[iy + 1]) + T1[iy + 1]) + T1[iy + 1]) * 0.25; /* @avxzi_th */
            const double sTxz = sin(Txz);
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rxz = (((R1[ix] + R1[ix]) + R1[ix + 1]) + R1[ix + 1]) * 0.25;   /* @avxzi_R */
                const double G = (((GS_(ix, iy + 1, iz) + GS_(ix, iy + 1, iz + 1)) + GS_(ix + 1, iy + 1, iz)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix, iy + 1, iz) + C_(eta, ix, iy + 1, iz + 1)) +
                                   C_(eta, ix + 1, iy + 1, iz)) + C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double avVP = (VP_(ix, iy + 1, iz + 1) + VP_(ix + 1, iy + 1, iz + 1)) * 0.5; /* @avxi_Vph */
                const double D = ((VP_(ix + 1, iy + 1, iz + 1) - VP_(ix, iy + 1, iz + 1)) * _dr +
                                  (((1.0 / Rxz) / sTxz) * (VR_(ix + 1, iy + 1, iz + 1) - VR_(ix + 1, iy + 1, iz))) * _dph) -
                                 avVP / Rxz;
                const size_t e = IDX(nr - 1, nt - 2, ix, iy, iz);
                tRP[e] = (tRP[e] + G * D) / (1.0 + G / E);
            }
        }
    /* tau_thph: (nr-2, nth-1, nph-1) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy) {
            const double Tyz = (((T1[iy] + T1[iy]) + T1[iy + 1]) + T1[iy + 1]) * 0.25;         /* @avyzi_th */
            const double sTyz = sin(Tyz), cTyz = 1.0 / tan(Tyz);
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Ryz = (((R1[ix + 1] + R1[ix + 1]) + R1[ix + 1]) + R1[ix + 1]) * 0.25; /* @avyzi_R */
                const double G = (((GS_(ix + 1, iy, iz) + GS_(ix + 1, iy, iz + 1)) + GS_(ix + 1, iy + 1, iz)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix + 1, iy, iz) + C_(eta, ix + 1, iy, iz + 1)) +
                                   C_(eta, ix + 1, iy + 1, iz)) + C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double avVP = (VP_(ix + 1, iy, iz + 1) + VP_(ix + 1, iy + 1, iz + 1)) * 0.5; /* @avyi_Vph */
                const double D = (((1.0 / Ryz) * (VP_(ix + 1, iy + 1, iz + 1) - VP_(ix + 1, iy, iz + 1))) * _dth +
                                  (((1.0 / Ryz) / sTyz) * (VT_(ix + 1, iy + 1, iz + 1) - VT_(ix + 1, iy + 1, iz))) * _dph) -
                                 (avVP / Ryz) * cTyz;
                const size_t e = IDX(nr - 2, nt - 1, ix, iy, iz);
                tTP[e] = (tTP[e] + G * D) / (1.0 + G / E);
            }
        }
#undef C_
#undef VR_
#undef VT_
#undef VP_
#undef GS_
}

/* compute_TII!, SphericalStokes_diapir.jl:243-257.  x^2.0 is x*x (an
 * integer-valued exponent; LLVM and Julia both reduce it to one multiply). */
void orc_compute_TII(orc_state *s)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *tRR = s->tRR, *tTT = s->tTT, *tPP = s->tPP;
    const double *tRT = s->tRT, *tRP = s->tRP, *tTP = s->tTP;
    double *tII = s->tII;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const size_t c = IDX(nr, nt, ix + 1, iy + 1, iz + 1);
                const double a = tRR[c], b = tTT[c], d = tPP[c];
#define RT_(i, j, k) tRT[IDX(nr - 1, nt - 1, i, j, k)]
#define RP_(i, j, k) tRP[IDX(nr - 1, nt - 2, i, j, k)]
#define TP_(i, j, k) tTP[IDX(nr - 2, nt - 1, i, j, k)]
                const double x = (((RT_(ix, iy, iz) + RT_(ix, iy + 1, iz)) + RT_(ix + 1, iy, iz)) + RT_(ix + 1, iy + 1, iz)) * 0.25;
                const double y = (((RP_(ix, iy, iz) + RP_(ix, iy, iz + 1)) + RP_(ix + 1, iy, iz)) + RP_(ix + 1, iy, iz + 1)) * 0.25;
                const double z = (((TP_(ix, iy, iz) + TP_(ix, iy, iz + 1)) + TP_(ix, iy + 1, iz)) + TP_(ix, iy + 1, iz + 1)) * 0.25;
#undef RT_
#undef RP_
#undef TP_
                tII[c] = sqrt((((1.0 / 2.0) * ((a * a + b * b) + d * d) + x * x) + y * y) + z * z);
            }
}

/* power_law!, SphericalStokes_diapir.jl:259-265:
 *   eta = 2/(1/eta_ini + 1/(eta_ini*(tII/tau_C)^(1-n_exp)))
 * Options beyond the spherical scripts (defaults reproduce them bit for bit):
 *   relax != 1: log-space relaxation of the power-law viscosity against its
 *     previous value, CartesianStokes_diapir.jl:210-221 (eta_PL_OLD <- eta_PL
 *     first, :361-362);
 *   eta_min > 0 / eta_max < inf: clamp of the combined viscosity. */
void orc_power_law(orc_state *s, const orc_params *p)
{
    const size_t n = (size_t)s->nr * s->nth * s->nph;
    const double tC = p->tau_C, ex = 1.0 - p->n_exp, relax = p->relax;
    const int do_relax = (relax != 1.0) && s->eta_PL;
    const int do_lo = p->eta_min > 0.0, do_hi = p->eta_max < INFINITY;
    double *eta = s->eta, *ePL = s->eta_PL;
    const double *tII = s->tII, *ini = s->eta_ini;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        double pl = ini[c] * pow(tII[c] / tC, ex);
        if (do_relax) {
            pl = exp(log(pl) * relax + log(ePL[c]) * (1.0 - relax));
            ePL[c] = pl;
        }
        double e = 2.0 / (1.0 / ini[c] + 1.0 / pl);
        if (do_lo) e = jl_max(e, p->eta_min);
        if (do_hi) e = jl_min(e, p->eta_max);
        eta[c] = e;
    }
}

/* compute_V!, SphericalStokes_diapir.jl:308-329 with the macros of :267-307 */
void orc_compute_V(orc_state *s, const orc_params *p)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *R1 = s->R1 - 1, *T1 = s->th1 - 1;
    const double _dr = p->_dr, _dth = p->_dth, _dph = p->_dph, g = p->g;
    const double *P = s->P, *DT = s->DT_R, *tRR = s->tRR, *tTT = s->tTT, *tPP = s->tPP;
    const double *tRT = s->tRT, *tRP = s->tRP, *tTP = s->tTP, *rho = s->rho;
    double *VR = s->VR, *VT = s->VT, *VP = s->VP;
#define C_(A, i, j, k) A[IDX(nr, nt, i, j, k)]
#define SRR(i, j, k) (-C_(P, i, j, k) + C_(tRR, i, j, k))
#define STT(i, j, k) (-C_(P, i, j, k) + C_(tTT, i, j, k))
#define SPP(i, j, k) (-C_(P, i, j, k) + C_(tPP, i, j, k))
#define RT_(i, j, k) tRT[IDX(nr - 1, nt - 1, i, j, k)]
#define RP_(i, j, k) tRP[IDX(nr - 1, nt - 2, i, j, k)]
#define TP_(i, j, k) tTP[IDX(nr - 2, nt - 1, i, j, k)]
    /* VR on [2:nr, 2:nth-1, 2:nph-1] */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy) {
            const double Tx = (T1[iy + 1] + T1[iy + 1]) * 0.5;      /* @avxi_th */
            const double sTx = sin(Tx), cTx = 1.0 / tan(Tx);
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rx = (R1[ix] + R1[ix + 1]) * 0.5;          /* @avxi_R */
                const double sRR = (SRR(ix, iy + 1, iz + 1) + SRR(ix + 1, iy + 1, iz + 1)) * 0.5;
                const double sTT = (STT(ix, iy + 1, iz + 1) + STT(ix + 1, iy + 1, iz + 1)) * 0.5;
                const double sPP = (SPP(ix, iy + 1, iz + 1) + SPP(ix + 1, iy + 1, iz + 1)) * 0.5;
                const double aRT = (RT_(ix, iy, iz) + RT_(ix, iy + 1, iz)) * 0.5; /* @avya_tauRth */
                const double rg = (C_(rho, ix, iy + 1, iz + 1) * g + C_(rho, ix + 1, iy + 1, iz + 1) * g) * 0.5;
                const double dV = (((((((SRR(ix + 1, iy + 1, iz + 1) - SRR(ix, iy + 1, iz + 1)) * _dr +
                                        ((1.0 / Rx) * (RT_(ix, iy + 1, iz) - RT_(ix, iy, iz))) * _dth) +
                                       (((1.0 / Rx) / sTx) * (RP_(ix, iy, iz + 1) - RP_(ix, iy, iz))) * _dph) +
                                      (2.0 * sRR) / Rx) -
                                     (sTT + sPP) / Rx) +
                                    (aRT / Rx) * cTx) -
                                   rg);
                const double dt = (C_(DT, ix, iy + 1, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const size_t f = IDX(nr + 1, nt, ix + 1, iy + 1, iz + 1);
                VR[f] = VR[f] + dV * dt;
            }
        }
    /* Vth on [2:nr-1, 2:nth, 2:nph-1] */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy) {
            const double Ty = (T1[iy] + T1[iy + 1]) * 0.5;          /* @avyi_th */
            const double sTy = sin(Ty), cTy = 1.0 / tan(Ty);
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Ry = (R1[ix + 1] + R1[ix + 1]) * 0.5;      /* @avyi_R */
                const double sTT = (STT(ix + 1, iy, iz + 1) + STT(ix + 1, iy + 1, iz + 1)) * 0.5;
                const double sPP = (SPP(ix + 1, iy, iz + 1) + SPP(ix + 1, iy + 1, iz + 1)) * 0.5;
                const double aRT = (RT_(ix, iy, iz) + RT_(ix + 1, iy, iz)) * 0.5; /* @avxa_tauRth */
                const double dV = ((((RT_(ix + 1, iy, iz) - RT_(ix, iy, iz)) * _dr +
                                     ((1.0 / Ry) * (STT(ix + 1, iy + 1, iz + 1) - STT(ix + 1, iy, iz + 1))) * _dth) +
                                    (((1.0 / Ry) / sTy) * (TP_(ix, iy, iz + 1) - TP_(ix, iy, iz))) * _dph) +
                                   (3.0 * aRT) / Ry) +
                                  ((sTT - sPP) / Ry) * cTy;
                const double dt = (C_(DT, ix + 1, iy, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const size_t f = IDX(nr, nt + 1, ix + 1, iy + 1, iz + 1);
                VT[f] = VT[f] + dV * dt;
            }
        }
    /* Vph on [2:nr-1, 2:nth-1, 2:nph] */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy) {
            const double Tz = (T1[iy + 1] + T1[iy + 1]) * 0.5;      /* @avzi_th */
            const double sTz = sin(Tz), cTz = 1.0 / tan(Tz);
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Rz = (R1[ix + 1] + R1[ix + 1]) * 0.5;      /* @avzi_R */
                const double aRP = (RP_(ix, iy, iz) + RP_(ix + 1, iy, iz)) * 0.5; /* @avxa_tauRph */
                const double aTP = (TP_(ix, iy, iz) + TP_(ix, iy + 1, iz)) * 0.5; /* @avya_tauthph */
                const double dV = ((((RP_(ix + 1, iy, iz) - RP_(ix, iy, iz)) * _dr +
                                     ((1.0 / Rz) * (TP_(ix, iy + 1, iz) - TP_(ix, iy, iz))) * _dth) +
                                    (((1.0 / Rz) / sTz) * (SPP(ix + 1, iy + 1, iz + 1) - SPP(ix + 1, iy + 1, iz))) * _dph) +
                                   (3.0 * aRP) / Rz) +
                                  ((2.0 * aTP) / Rz) * cTz;
                const double dt = (C_(DT, ix + 1, iy + 1, iz) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const size_t f = IDX(nr, nt, ix + 1, iy + 1, iz + 1);
                VP[f] = VP[f] + dV * dt;
            }
        }
#undef C_
#undef SRR
#undef STT
#undef SPP
#undef RT_
#undef RP_
#undef TP_
}

/* BC_R!, identical in SphericalStokes_diapir.jl:61-69 and
 * SphericalStokes_plateau.jl:64-72: free slip on the r faces. */
void orc_bc_R(orc_state *s)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    double *VR = s->VR, *VT = s->VT, *VP = s->VP;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            VR[IDX(nr + 1, nt, 1, iy, iz)] = -VR[IDX(nr + 1, nt, 2, iy, iz)];
            VR[IDX(nr + 1, nt, nr + 1, iy, iz)] = -VR[IDX(nr + 1, nt, nr, iy, iz)];
        }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt + 1; ++iy) {
            VT[IDX(nr, nt + 1, 1, iy, iz)] = VT[IDX(nr, nt + 1, 2, iy, iz)];
            VT[IDX(nr, nt + 1, nr, iy, iz)] = VT[IDX(nr, nt + 1, nr - 1, iy, iz)];
        }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np + 1; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            VP[IDX(nr, nt, 1, iy, iz)] = VP[IDX(nr, nt, 2, iy, iz)];
            VP[IDX(nr, nt, nr, iy, iz)] = VP[IDX(nr, nt, nr - 1, iy, iz)];
        }
}

/* BC_th!:
 *  ORC_BC_SHEAR     SphericalStokes_diapir.jl:71-86 (VR=0, Vth antisymmetric,
 *                   Vph = (th-pi/2)*r*sh, last Vph column uses th[..,nph])
 *  ORC_BC_FREE_SLIP SphericalStokes_plateau.jl:74-82
 *  ORC_BC_PURE_SHEAR extension: normal velocity prescribed at the boundary
 *                   centre, Vth_b = -(((th_b-pi/2)*r)*sh)/sin(th_b),
 *                   tangential copied.  With Vph_b = (ph_b*r)*sh on the phi
 *                   faces (orc_bc_ph) these are the normal components of
 *                   V = (0, -sh*r*(th-pi/2)/sin(th), sh*r*ph), a
 *                   divergence-free pure shear of rate sh in the th-phi
 *                   plane, so the prescribed boundary fluxes cancel. */
void orc_bc_th(orc_state *s, const orc_params *p)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    double *VR = s->VR, *VT = s->VT, *VP = s->VP;
    const double *T1 = s->th1 - 1;
    const double r = p->r, sh = p->sh;
    const int mode = p->bc_mode;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np; ++iz)
        for (int ix = 1; ix <= nr + 1; ++ix) {
            if (mode == ORC_BC_SHEAR) {
                VR[IDX(nr + 1, nt, ix, 1, iz)] = 0.0;
                VR[IDX(nr + 1, nt, ix, nt, iz)] = 0.0;
            } else {
                VR[IDX(nr + 1, nt, ix, 1, iz)] = VR[IDX(nr + 1, nt, ix, 2, iz)];
                VR[IDX(nr + 1, nt, ix, nt, iz)] = VR[IDX(nr + 1, nt, ix, nt - 1, iz)];
            }
        }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np; ++iz)
        for (int ix = 1; ix <= nr; ++ix) {
            if (mode == ORC_BC_PURE_SHEAR) {
                const double v0 = -((((T1[1] - PI / 2.0) * r) * sh) / sin(T1[1]));
                const double v1 = -((((T1[nt] - PI / 2.0) * r) * sh) / sin(T1[nt]));
                VT[IDX(nr, nt + 1, ix, 1, iz)] = 2.0 * v0 - VT[IDX(nr, nt + 1, ix, 2, iz)];
                VT[IDX(nr, nt + 1, ix, nt + 1, iz)] = 2.0 * v1 - VT[IDX(nr, nt + 1, ix, nt, iz)];
            } else {
                VT[IDX(nr, nt + 1, ix, 1, iz)] = -VT[IDX(nr, nt + 1, ix, 2, iz)];
                VT[IDX(nr, nt + 1, ix, nt + 1, iz)] = -VT[IDX(nr, nt + 1, ix, nt, iz)];
            }
        }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= np + 1; ++iz)
        for (int ix = 1; ix <= nr; ++ix) {
            if (mode == ORC_BC_SHEAR) {
                /* th[ix,iy,iz] (iz<=nph) or th[ix,iy,iz-1] (iz=nph+1): both th1[iy] */
                VP[IDX(nr, nt, ix, 1, iz)] = ((T1[1] - PI / 2.0) * r) * sh;
                VP[IDX(nr, nt, ix, nt, iz)] = ((T1[nt] - PI / 2.0) * r) * sh;
            } else {
                VP[IDX(nr, nt, ix, 1, iz)] = VP[IDX(nr, nt, ix, 2, iz)];
                VP[IDX(nr, nt, ix, nt, iz)] = VP[IDX(nr, nt, ix, nt - 1, iz)];
            }
        }
}

/* BC_ph!:
 *  ORC_BC_SHEAR     SphericalStokes_diapir.jl:88-99 (VR=Vth=0,
 *                   Vph = 2*(th-pi/2)*r*sh - neighbour)
 *  ORC_BC_FREE_SLIP SphericalStokes_plateau.jl:84-92
 *  ORC_BC_PURE_SHEAR extension: Vph_b = (ph_b*r)*sh at the boundary centre. */
void orc_bc_ph(orc_state *s, const orc_params *p)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    double *VR = s->VR, *VT = s->VT, *VP = s->VP;
    const double *T1 = s->th1 - 1, *F1 = s->ph1 - 1;
    const double r = p->r, sh = p->sh;
    const int mode = p->bc_mode;
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy <= nt; ++iy)
        for (int ix = 1; ix <= nr + 1; ++ix) {
            if (mode == ORC_BC_SHEAR) {
                VR[IDX(nr + 1, nt, ix, iy, 1)] = 0.0;
                VR[IDX(nr + 1, nt, ix, iy, np)] = 0.0;
            } else {
                VR[IDX(nr + 1, nt, ix, iy, 1)] = VR[IDX(nr + 1, nt, ix, iy, 2)];
                VR[IDX(nr + 1, nt, ix, iy, np)] = VR[IDX(nr + 1, nt, ix, iy, np - 1)];
            }
        }
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy <= nt + 1; ++iy)
        for (int ix = 1; ix <= nr; ++ix) {
            if (mode == ORC_BC_SHEAR) {
                VT[IDX(nr, nt + 1, ix, iy, 1)] = 0.0;
                VT[IDX(nr, nt + 1, ix, iy, np)] = 0.0;
            } else {
                VT[IDX(nr, nt + 1, ix, iy, 1)] = VT[IDX(nr, nt + 1, ix, iy, 2)];
                VT[IDX(nr, nt + 1, ix, iy, np)] = VT[IDX(nr, nt + 1, ix, iy, np - 1)];
            }
        }
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy <= nt; ++iy)
        for (int ix = 1; ix <= nr; ++ix) {
            double *lo = &VP[IDX(nr, nt, ix, iy, 1)], *hi = &VP[IDX(nr, nt, ix, iy, np + 1)];
            const double in_lo = VP[IDX(nr, nt, ix, iy, 2)], in_hi = VP[IDX(nr, nt, ix, iy, np)];
            if (mode == ORC_BC_SHEAR) {
                *lo = ((2.0 * (T1[iy] - PI / 2.0)) * r) * sh - in_lo;
                *hi = ((2.0 * (T1[iy] - PI / 2.0)) * r) * sh - in_hi;
            } else if (mode == ORC_BC_FREE_SLIP) {
                *lo = -in_lo;
                *hi = -in_hi;
            } else {
                *lo = 2.0 * ((F1[1] * r) * sh) - in_lo;
                *hi = 2.0 * ((F1[np] * r) * sh) - in_hi;
            }
        }
}

/* check_err!, SphericalStokes_diapir.jl:332-344, and the host reduction of
 * :428-432 (err = max over components of max|V-V_prev| / max|V|).  The
 * plateau script drops a NaN Vph component (SphericalStokes_plateau.jl:545-549). */
double orc_check_err(orc_state *s, const orc_params *p, double out_comp[3])
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    double *E[3] = {s->VR_err, s->VT_err, s->VP_err};
    const double *V[3] = {s->VR, s->VT, s->VP};
    const size_t n[3] = {(size_t)(nr + 1) * nt * np, (size_t)nr * (nt + 1) * np, (size_t)nr * nt * (np + 1)};
    double comp[3];
    for (int a = 0; a < 3; ++a) {
        double me = 0.0, mv = 0.0;
        int nan_e = 0, nan_v = 0;
#pragma omp parallel for reduction(max : me, mv) reduction(| : nan_e, nan_v) schedule(static)
        for (size_t c = 0; c < n[a]; ++c) {
            const double d = V[a][c] - E[a][c];
            E[a][c] = d;
            const double ad = fabs(d), av = fabs(V[a][c]);
            if (isnan(ad)) nan_e = 1; else if (ad > me) me = ad;
            if (isnan(av)) nan_v = 1; else if (av > mv) mv = av;
        }
        /* maximum() propagates NaN */
        comp[a] = (nan_e || nan_v) ? NAN : me / mv;
    }
    if (out_comp) { out_comp[0] = comp[0]; out_comp[1] = comp[1]; out_comp[2] = comp[2]; }
    if (p->err_drop_nan_vph && isnan(comp[2])) return jl_max(comp[0], comp[1]);
    return jl_max(jl_max(comp[0], comp[1]), comp[2]);
}

/* initialize_inclusion!, SphericalStokes_diapir.jl:127-137.  The test
 * compares the SQUARED distance with `radius` (as the reference does). */
void orc_initialize_inclusion(double *A, const double *R1, const double *th1,
                              const double *ph1, int nr, int nth, int nph,
                              double r, double radius, double in)
{
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nph; ++iz)
        for (int iy = 1; iy <= nth; ++iy)
            for (int ix = 1; ix <= nr; ++ix) {
                const double a = R1[ix - 1] - r, b = (th1[iy - 1] - PI / 2.0) * r, c = ph1[iz - 1] * r;
                if (((a * a + b * b) + c * c) < radius) A[IDX(nr, nth, ix, iy, iz)] = in;
            }
}

/* initialize_velocity!, SphericalStokes_diapir.jl:139-146, applied to Vph
 * with COORD = th (:403). */
void orc_initialize_velocity(double *Vph, const double *th1, int nr, int nth,
                             int nph, double fact, double r)
{
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nph + 1; ++iz)
        for (int iy = 1; iy <= nth; ++iy)
            for (int ix = 1; ix <= nr; ++ix)
                Vph[IDX(nr, nth, ix, iy, iz)] = ((th1[iy - 1] - PI / 2.0) * r) * fact;
}

/* One PT iteration, loop body SphericalStokes_diapir.jl:413-425
 * (plateau :526-538 with the free-slip BCs). */
void orc_pt_step(orc_state *s, const orc_params *p, int do_equal3)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    if (do_equal3) orc_equal3(s->VR_err, s->VT_err, s->VP_err, s->VR, s->VT, s->VP, nr, nt, np);
    orc_maxloc(s->etaSM, s->eta, nr, nt, np);
    orc_copy_bc(s->etaSM, nr, nt, np);
    /* build option (no reference counterpart): DT_R serves as the
     * workspace, compute_P! overwrites every entry of it next */
    if (p->etasm_spread > 0) orc_spread_etasm(s->etaSM, s->DT_R, nr, nt, np, p->etasm_spread);
    orc_compute_P(s, p);
    orc_compute_TII(s);
    orc_copy_bc(s->tII, nr, nt, np);
    orc_power_law(s, p);
    orc_compute_V(s, p);
    orc_bc_R(s);
    orc_bc_th(s, p);
    orc_bc_ph(s, p);
}

/* PT loop with the exit test, SphericalStokes_diapir.jl:409-443 */
int orc_pt_solve(orc_state *s, const orc_params *p, int maxiter, int nout,
                 double eps, int always_equal3, double *err_out, int *iter_end)
{
    int ittot = 0;
    double err = NAN;
    if (iter_end) *iter_end = 0;
    for (int iter = 1; iter <= maxiter; ++iter) {
        const int check = (nout > 0) && (iter % nout == 0);
        orc_pt_step(s, p, always_equal3 || check);
        if (check) {
            err = orc_check_err(s, p, NULL);
            if (err < eps && iter > 20) {
                if (iter_end) *iter_end = iter;
                break;
            }
        }
        ittot = iter;
    }
    if (err_out) *err_out = err;
    return ittot;
}
