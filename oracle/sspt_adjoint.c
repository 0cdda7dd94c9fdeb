/*
 * sspt_adjoint.c -- CPU reference of the adjoint PT iteration.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint.
 *
 * One PT iteration x' = Phi(x; m) (orc_pt_step: compute_P!, compute_TII! +
 * copy_BC!, power_law!, compute_V!, BC_R!/BC_th!/BC_ph!) has at a fixed point
 * x* = Phi(x*) the Jacobian T = dPhi/dx.  Its step sizes (etaSM through KBDT,
 * GSDT and DT_R) multiply residuals that vanish at x*, so T is taken with the
 * step sizes of x* held fixed; this is exact at the fixed point.
 * orc_adj_step applies T^T (and accumulates (dPhi/dm)^T for m = rho, eta_ini)
 * by running the oracle's statements backwards: every assignment y = f(u)
 * adds (df/du)^T lam_y to lam_u and clears lam_y when y overwrote an older
 * value; every in-place update y += g(u) keeps lam_y.  Loops follow
 * sspt_oracle.c one for one, in reverse order.
 *
 * The adjoint of the converged solve is lam = T^T lam + dJ/dx (iterated from
 * 0, the same contraction as the forward PT loop), and dJ/dm = (dPhi/dm)^T
 * lam.  J = the volume-weighted mean of VR over the r-faces inside the
 * inclusion (orc_adj_objective).
 */
#include "sspt_adjoint.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(n1, n2, i, j, k) (((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1))
static const double PI = 3.14159265358979323846;

size_t orc_adj_len(int nr, int nth, int nph, int which)
{
    switch (which) {
    case ORC_A_VR: return (size_t)(nr + 1) * nth * nph;
    case ORC_A_VT: return (size_t)nr * (nth + 1) * nph;
    case ORC_A_VP: return (size_t)nr * nth * (nph + 1);
    case ORC_A_RT: return (size_t)(nr - 1) * (nth - 1) * (nph - 2);
    case ORC_A_RP: return (size_t)(nr - 1) * (nth - 2) * (nph - 1);
    case ORC_A_TP: return (size_t)(nr - 2) * (nth - 1) * (nph - 1);
    default: return (size_t)nr * nth * nph; /* P, tRR, tTT, tPP, eta, eta_PL */
    }
}

/* ---------------------------------------------------------------- BCs
 * Each BC statement is "ghost = a*source + c"; backwards:
 * lam_source += a*lam_ghost; lam_ghost = 0. */
static void adj_bc_R(orc_adj *L, int nr, int nt, int np)
{
    double *VR = L->a[ORC_A_VR], *VT = L->a[ORC_A_VT], *VP = L->a[ORC_A_VP];
    for (int iz = 1; iz <= np + 1; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            double *g1 = &VP[IDX(nr, nt, 1, iy, iz)], *g2 = &VP[IDX(nr, nt, nr, iy, iz)];
            VP[IDX(nr, nt, 2, iy, iz)] += *g1; *g1 = 0.0;
            VP[IDX(nr, nt, nr - 1, iy, iz)] += *g2; *g2 = 0.0;
        }
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt + 1; ++iy) {
            double *g1 = &VT[IDX(nr, nt + 1, 1, iy, iz)], *g2 = &VT[IDX(nr, nt + 1, nr, iy, iz)];
            VT[IDX(nr, nt + 1, 2, iy, iz)] += *g1; *g1 = 0.0;
            VT[IDX(nr, nt + 1, nr - 1, iy, iz)] += *g2; *g2 = 0.0;
        }
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            double *g1 = &VR[IDX(nr + 1, nt, 1, iy, iz)], *g2 = &VR[IDX(nr + 1, nt, nr + 1, iy, iz)];
            VR[IDX(nr + 1, nt, 2, iy, iz)] -= *g1; *g1 = 0.0;
            VR[IDX(nr + 1, nt, nr, iy, iz)] -= *g2; *g2 = 0.0;
        }
}

static void adj_bc_th(orc_adj *L, int nr, int nt, int np, int mode)
{
    double *VR = L->a[ORC_A_VR], *VT = L->a[ORC_A_VT], *VP = L->a[ORC_A_VP];
    /* Vph rows (last loop of BC_th!) */
    for (int iz = 1; iz <= np + 1; ++iz)
        for (int ix = 1; ix <= nr; ++ix) {
            double *g1 = &VP[IDX(nr, nt, ix, 1, iz)], *g2 = &VP[IDX(nr, nt, ix, nt, iz)];
            if (mode != ORC_BC_SHEAR) {
                VP[IDX(nr, nt, ix, 2, iz)] += *g1;
                VP[IDX(nr, nt, ix, nt - 1, iz)] += *g2;
            }
            *g1 = 0.0; *g2 = 0.0;
        }
    /* Vth: ghost = (2 vb) - inner (pure shear) or -inner */
    for (int iz = 1; iz <= np; ++iz)
        for (int ix = 1; ix <= nr; ++ix) {
            double *g1 = &VT[IDX(nr, nt + 1, ix, 1, iz)], *g2 = &VT[IDX(nr, nt + 1, ix, nt + 1, iz)];
            VT[IDX(nr, nt + 1, ix, 2, iz)] -= *g1;
            VT[IDX(nr, nt + 1, ix, nt, iz)] -= *g2;
            *g1 = 0.0; *g2 = 0.0;
        }
    for (int iz = 1; iz <= np; ++iz)
        for (int ix = 1; ix <= nr + 1; ++ix) {
            double *g1 = &VR[IDX(nr + 1, nt, ix, 1, iz)], *g2 = &VR[IDX(nr + 1, nt, ix, nt, iz)];
            if (mode != ORC_BC_SHEAR) {
                VR[IDX(nr + 1, nt, ix, 2, iz)] += *g1;
                VR[IDX(nr + 1, nt, ix, nt - 1, iz)] += *g2;
            }
            *g1 = 0.0; *g2 = 0.0;
        }
}

static void adj_bc_ph(orc_adj *L, int nr, int nt, int np, int mode)
{
    double *VR = L->a[ORC_A_VR], *VT = L->a[ORC_A_VT], *VP = L->a[ORC_A_VP];
    /* Vph: lo = c - in_lo (every mode) */
    for (int iy = 1; iy <= nt; ++iy)
        for (int ix = 1; ix <= nr; ++ix) {
            double *g1 = &VP[IDX(nr, nt, ix, iy, 1)], *g2 = &VP[IDX(nr, nt, ix, iy, np + 1)];
            VP[IDX(nr, nt, ix, iy, 2)] -= *g1;
            VP[IDX(nr, nt, ix, iy, np)] -= *g2;
            *g1 = 0.0; *g2 = 0.0;
        }
    for (int iy = 1; iy <= nt + 1; ++iy)
        for (int ix = 1; ix <= nr; ++ix) {
            double *g1 = &VT[IDX(nr, nt + 1, ix, iy, 1)], *g2 = &VT[IDX(nr, nt + 1, ix, iy, np)];
            if (mode != ORC_BC_SHEAR) {
                VT[IDX(nr, nt + 1, ix, iy, 2)] += *g1;
                VT[IDX(nr, nt + 1, ix, iy, np - 1)] += *g2;
            }
            *g1 = 0.0; *g2 = 0.0;
        }
    for (int iy = 1; iy <= nt; ++iy)
        for (int ix = 1; ix <= nr + 1; ++ix) {
            double *g1 = &VR[IDX(nr + 1, nt, ix, iy, 1)], *g2 = &VR[IDX(nr + 1, nt, ix, iy, np)];
            if (mode != ORC_BC_SHEAR) {
                VR[IDX(nr + 1, nt, ix, iy, 2)] += *g1;
                VR[IDX(nr + 1, nt, ix, iy, np - 1)] += *g2;
            }
            *g1 = 0.0; *g2 = 0.0;
        }
}

/* copy_BC! backwards (phi, th, r passes in reverse order) */
static void adj_copy_bc(double *A, int n1, int n2, int n3)
{
    for (int iy = 1; iy <= n2; ++iy)
        for (int ix = 1; ix <= n1; ++ix) {
            A[IDX(n1, n2, ix, iy, 2)] += A[IDX(n1, n2, ix, iy, 1)]; A[IDX(n1, n2, ix, iy, 1)] = 0.0;
            A[IDX(n1, n2, ix, iy, n3 - 1)] += A[IDX(n1, n2, ix, iy, n3)]; A[IDX(n1, n2, ix, iy, n3)] = 0.0;
        }
    for (int iz = 1; iz <= n3; ++iz)
        for (int ix = 1; ix <= n1; ++ix) {
            A[IDX(n1, n2, ix, 2, iz)] += A[IDX(n1, n2, ix, 1, iz)]; A[IDX(n1, n2, ix, 1, iz)] = 0.0;
            A[IDX(n1, n2, ix, n2 - 1, iz)] += A[IDX(n1, n2, ix, n2, iz)]; A[IDX(n1, n2, ix, n2, iz)] = 0.0;
        }
    for (int iz = 1; iz <= n3; ++iz)
        for (int iy = 1; iy <= n2; ++iy) {
            A[IDX(n1, n2, 2, iy, iz)] += A[IDX(n1, n2, 1, iy, iz)]; A[IDX(n1, n2, 1, iy, iz)] = 0.0;
            A[IDX(n1, n2, n1 - 1, iy, iz)] += A[IDX(n1, n2, n1, iy, iz)]; A[IDX(n1, n2, n1, iy, iz)] = 0.0;
        }
}

/* ---------------------------------------------------------------- compute_V! backwards
 * V_f += DT_f * (div sigma(P', tau') - rho g)_f: lam_V unchanged, and
 * w = DT_f lam_V_f is spread onto the stresses and pressure it reads. */
static void adj_compute_V(const orc_state *s, const orc_params *p, orc_adj *L, double *g_rho)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *R1 = s->R1 - 1, *T1 = s->th1 - 1;
    const double _dr = p->_dr, _dth = p->_dth, _dph = p->_dph, g = p->g;
    const double *DT = s->DT_R;
    double *lP = L->a[ORC_A_P], *lRR = L->a[ORC_A_RR], *lTT = L->a[ORC_A_TT], *lPP = L->a[ORC_A_PP];
    double *lRT = L->a[ORC_A_RT], *lRP = L->a[ORC_A_RP], *lTP = L->a[ORC_A_TP];
    const double *lVR = L->a[ORC_A_VR], *lVT = L->a[ORC_A_VT], *lVP = L->a[ORC_A_VP];
#define C_(A, i, j, k) A[IDX(nr, nt, i, j, k)]
#define RT_(i, j, k) lRT[IDX(nr - 1, nt - 1, i, j, k)]
#define RP_(i, j, k) lRP[IDX(nr - 1, nt - 2, i, j, k)]
#define TP_(i, j, k) lTP[IDX(nr - 2, nt - 1, i, j, k)]
/* sigma_nn = -P + tau_nn: a weight on sigma_nn goes to tau_nn and, negated, to P */
#define S_(T, i, j, k, w) do { C_(T, i, j, k) += (w); C_(lP, i, j, k) -= (w); } while (0)
    /* VR faces [2:nr, 2:nth-1, 2:nph-1] */
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rx = (R1[ix] + R1[ix + 1]) * 0.5;
                const double Tx = (T1[iy + 1] + T1[iy + 1]) * 0.5;
                const double dt = (C_(DT, ix, iy + 1, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const double w = dt * lVR[IDX(nr + 1, nt, ix + 1, iy + 1, iz + 1)];
                if (w == 0.0) continue;
                const double ct = 1.0 / tan(Tx), rs = (1.0 / Rx) / sin(Tx);
                /* (SRR(+1) - SRR(0))*_dr + (2 sRR)/Rx - (sTT + sPP)/Rx with sXX = avg of the two centres */
                S_(lRR, ix + 1, iy + 1, iz + 1, w * (_dr + 1.0 / Rx));
                S_(lRR, ix, iy + 1, iz + 1, w * (-_dr + 1.0 / Rx));
                S_(lTT, ix, iy + 1, iz + 1, -w * 0.5 / Rx);
                S_(lTT, ix + 1, iy + 1, iz + 1, -w * 0.5 / Rx);
                S_(lPP, ix, iy + 1, iz + 1, -w * 0.5 / Rx);
                S_(lPP, ix + 1, iy + 1, iz + 1, -w * 0.5 / Rx);
                RT_(ix, iy + 1, iz) += w * (_dth / Rx + 0.5 * ct / Rx);
                RT_(ix, iy, iz) += w * (-_dth / Rx + 0.5 * ct / Rx);
                RP_(ix, iy, iz + 1) += w * rs * _dph;
                RP_(ix, iy, iz) -= w * rs * _dph;
                if (g_rho) {
                    C_(g_rho, ix, iy + 1, iz + 1) -= w * 0.5 * g;
                    C_(g_rho, ix + 1, iy + 1, iz + 1) -= w * 0.5 * g;
                }
            }
    /* Vth faces [2:nr-1, 2:nth, 2:nph-1] */
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Ry = (R1[ix + 1] + R1[ix + 1]) * 0.5;
                const double Ty = (T1[iy] + T1[iy + 1]) * 0.5;
                const double dt = (C_(DT, ix + 1, iy, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const double w = dt * lVT[IDX(nr, nt + 1, ix + 1, iy + 1, iz + 1)];
                if (w == 0.0) continue;
                const double ct = 1.0 / tan(Ty), rs = (1.0 / Ry) / sin(Ty);
                RT_(ix + 1, iy, iz) += w * (_dr + 1.5 / Ry);
                RT_(ix, iy, iz) += w * (-_dr + 1.5 / Ry);
                S_(lTT, ix + 1, iy + 1, iz + 1, w * (_dth / Ry + 0.5 * ct / Ry));
                S_(lTT, ix + 1, iy, iz + 1, w * (-_dth / Ry + 0.5 * ct / Ry));
                S_(lPP, ix + 1, iy, iz + 1, -w * 0.5 * ct / Ry);
                S_(lPP, ix + 1, iy + 1, iz + 1, -w * 0.5 * ct / Ry);
                TP_(ix, iy, iz + 1) += w * rs * _dph;
                TP_(ix, iy, iz) -= w * rs * _dph;
            }
    /* Vph faces [2:nr-1, 2:nth-1, 2:nph] */
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Rz = (R1[ix + 1] + R1[ix + 1]) * 0.5;
                const double Tz = (T1[iy + 1] + T1[iy + 1]) * 0.5;
                const double dt = (C_(DT, ix + 1, iy + 1, iz) + C_(DT, ix + 1, iy + 1, iz + 1)) * 0.5;
                const double w = dt * lVP[IDX(nr, nt, ix + 1, iy + 1, iz + 1)];
                if (w == 0.0) continue;
                const double ct = 1.0 / tan(Tz), rs = (1.0 / Rz) / sin(Tz);
                RP_(ix + 1, iy, iz) += w * (_dr + 1.5 / Rz);
                RP_(ix, iy, iz) += w * (-_dr + 1.5 / Rz);
                TP_(ix, iy + 1, iz) += w * (_dth / Rz + ct / Rz);
                TP_(ix, iy, iz) += w * (-_dth / Rz + ct / Rz);
                S_(lPP, ix + 1, iy + 1, iz + 1, w * rs * _dph);
                S_(lPP, ix + 1, iy + 1, iz, -w * rs * _dph);
            }
#undef S_
#undef C_
#undef RT_
#undef RP_
#undef TP_
}

/* ---------------------------------------------------------------- power_law! backwards
 * eta' = f(tII, eta_ini, eta_PL) on every cell: lam_tII, the eta_ini
 * gradient and lam_eta_PL from lam_eta' (and lam_eta_PL'); eta and eta_PL
 * are overwritten, so their incoming adjoints are consumed. */
static void adj_power_law(const orc_state *s, const orc_params *p, orc_adj *L, double *ltII, double *g_ini)
{
    const size_t n = (size_t)s->nr * s->nth * s->nph;
    const double tC = p->tau_C, ex = 1.0 - p->n_exp, relax = p->relax;
    const int do_relax = relax != 1.0;
    const int do_lo = p->eta_min > 0.0, do_hi = p->eta_max < INFINITY;
    double *le = L->a[ORC_A_ETA], *lpl = L->a[ORC_A_EPL];
    for (size_t c = 0; c < n; ++c) {
        const double tii = s->tII[c], ini = s->eta_ini[c];
        const double pl0 = ini * pow(tii / tC, ex);
        /* the fixed point: eta_PL* = pl (relaxed value), eta* = 2/(1/ini + 1/pl) */
        const double pl = do_relax ? s->eta_PL[c] : pl0;
        const double e_raw = 2.0 / (1.0 / ini + 1.0 / pl);
        const int clamped = (do_lo && e_raw < p->eta_min) || (do_hi && e_raw > p->eta_max);
        const double a_e = clamped ? 0.0 : le[c];            /* lam of the unclamped eta */
        if (!isfinite(pl) || !isfinite(pl0)) {
            /* tII = 0 (n > 1): pl = inf and eta = 2 eta_ini, locally independent of tII and eta_PL */
            ltII[c] = 0.0;
            if (g_ini) g_ini[c] += a_e * 2.0;
            if (do_relax) lpl[c] = 0.0;
            le[c] = 0.0;
            continue;
        }
        /* eta = 2/(1/ini + 1/pl): d/dpl = eta^2/(2 pl^2), d/dini (direct) = eta^2/(2 ini^2) */
        double l_pl = a_e * (e_raw * e_raw) / (2.0 * pl * pl);
        double l_ini = a_e * (e_raw * e_raw) / (2.0 * ini * ini);
        if (do_relax) {
            l_pl += lpl[c];                                   /* eta_PL' = pl */
            /* pl = exp(relax log pl0 + (1-relax) log ePL_old) */
            const double l_pl0 = l_pl * relax * pl / pl0;
            lpl[c] = l_pl * (1.0 - relax) * pl / s->eta_PL[c]; /* lam of the old eta_PL */
            l_pl = l_pl0;
        }
        /* pl0 = ini (tII/tC)^ex */
        ltII[c] = (tii > 0.0) ? l_pl * ex * pl0 / tii : 0.0;
        l_ini += l_pl * pl0 / ini;
        if (g_ini) g_ini[c] += l_ini;
        le[c] = 0.0;
    }
}

/* compute_TII! backwards: interior cells only (the boundary tII are the
 * copy_BC! copies, already folded back onto the interior) */
static void adj_compute_TII(const orc_state *s, orc_adj *L, const double *ltII)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *tRR = s->tRR, *tTT = s->tTT, *tPP = s->tPP;
    const double *tRT = s->tRT, *tRP = s->tRP, *tTP = s->tTP;
    double *lRR = L->a[ORC_A_RR], *lTT = L->a[ORC_A_TT], *lPP = L->a[ORC_A_PP];
    double *lRT = L->a[ORC_A_RT], *lRP = L->a[ORC_A_RP], *lTP = L->a[ORC_A_TP];
#define RT_(A, i, j, k) A[IDX(nr - 1, nt - 1, i, j, k)]
#define RP_(A, i, j, k) A[IDX(nr - 1, nt - 2, i, j, k)]
#define TP_(A, i, j, k) A[IDX(nr - 2, nt - 1, i, j, k)]
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const size_t c = IDX(nr, nt, ix + 1, iy + 1, iz + 1);
                const double l = ltII[c];
                if (l == 0.0) continue;
                const double tii = s->tII[c];
                if (!(tii > 0.0)) continue;
                const double x = (((RT_(tRT, ix, iy, iz) + RT_(tRT, ix, iy + 1, iz)) + RT_(tRT, ix + 1, iy, iz)) +
                                  RT_(tRT, ix + 1, iy + 1, iz)) * 0.25;
                const double y = (((RP_(tRP, ix, iy, iz) + RP_(tRP, ix, iy, iz + 1)) + RP_(tRP, ix + 1, iy, iz)) +
                                  RP_(tRP, ix + 1, iy, iz + 1)) * 0.25;
                const double z = (((TP_(tTP, ix, iy, iz) + TP_(tTP, ix, iy, iz + 1)) + TP_(tTP, ix, iy + 1, iz)) +
                                  TP_(tTP, ix, iy + 1, iz + 1)) * 0.25;
                const double q = l / tii;
                lRR[c] += q * 0.5 * tRR[c];
                lTT[c] += q * 0.5 * tTT[c];
                lPP[c] += q * 0.5 * tPP[c];
                const double wx = q * x * 0.25, wy = q * y * 0.25, wz = q * z * 0.25;
                RT_(lRT, ix, iy, iz) += wx; RT_(lRT, ix, iy + 1, iz) += wx;
                RT_(lRT, ix + 1, iy, iz) += wx; RT_(lRT, ix + 1, iy + 1, iz) += wx;
                RP_(lRP, ix, iy, iz) += wy; RP_(lRP, ix, iy, iz + 1) += wy;
                RP_(lRP, ix + 1, iy, iz) += wy; RP_(lRP, ix + 1, iy, iz + 1) += wy;
                TP_(lTP, ix, iy, iz) += wz; TP_(lTP, ix, iy, iz + 1) += wz;
                TP_(lTP, ix, iy + 1, iz) += wz; TP_(lTP, ix, iy + 1, iz + 1) += wz;
            }
#undef RT_
#undef RP_
#undef TP_
}

/* ---------------------------------------------------------------- compute_P! backwards */
static void adj_compute_P(const orc_state *s, const orc_params *p, orc_adj *L)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *R1 = s->R1 - 1, *T1 = s->th1 - 1;
    const double _dr = p->_dr, _dth = p->_dth, _dph = p->_dph;
    const double kb = (((p->dmp * 2.0) * PI) * p->vpdt) / p->lr;
    const double gs = ((4.0 * PI) * p->vpdt) / p->lr;
    const double third = 1.0 / 3.0;
    const double *eta = s->eta, *SM = s->etaSM;
    double *lVR = L->a[ORC_A_VR], *lVT = L->a[ORC_A_VT], *lVP = L->a[ORC_A_VP];
    double *lP = L->a[ORC_A_P], *lRR = L->a[ORC_A_RR], *lTT = L->a[ORC_A_TT], *lPP = L->a[ORC_A_PP];
    double *lRT = L->a[ORC_A_RT], *lRP = L->a[ORC_A_RP], *lTP = L->a[ORC_A_TP], *le = L->a[ORC_A_ETA];
#define C_(A, i, j, k) A[IDX(nr, nt, i, j, k)]
#define VR_(i, j, k) lVR[IDX(nr + 1, nt, i, j, k)]
#define VT_(i, j, k) lVT[IDX(nr, nt + 1, i, j, k)]
#define VP_(i, j, k) lVP[IDX(nr, nt, i, j, k)]
#define GS_(i, j, k) (gs * SM[IDX(nr, nt, i, j, k)])
    /* centres: P' = P - KB divV; tnn' = (tnn + 2 GS Dnn)/den, den = 1 + GS/eta */
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt; ++iy) {
            const double sn = sin(T1[iy]), ct = 1.0 / tan(T1[iy]);
            for (int ix = 1; ix <= nr; ++ix) {
                const size_t c = IDX(nr, nt, ix, iy, iz);
                const double R = R1[ix];
                const double KB = kb * SM[c], GS = gs * SM[c], den = 1.0 + GS / eta[c];
                const double aRR = lRR[c], aTT = lTT[c], aPP = lPP[c], aP = lP[c];
                /* stresses are overwritten: lam of the old ones */
                lRR[c] = aRR / den;
                lTT[c] = aTT / den;
                lPP[c] = aPP / den;
                /* d tnn'/d eta = tnn' GS / (den eta^2), tnn' = tnn* at the fixed point */
                le[c] += (aRR * s->tRR[c] + aTT * s->tTT[c] + aPP * s->tPP[c]) * GS / (den * eta[c] * eta[c]);
                /* weights on the strain-rate terms */
                const double f = 2.0 * GS / den;
                const double wR = f * aRR, wT = f * aTT, wP = f * aPP;
                /* Drr = dVRr - divV/3, Dtt = dVTt + avVR/R - divV/3, Dpp = dVPp + avVR/R + avVT/R ct - divV/3 */
                const double wdiv = -KB * aP - third * (wR + wT + wP);
                const double w_dVRr = wR + wdiv, w_dVTt = wT + wdiv, w_dVPp = wP + wdiv;
                const double w_avVR = (wT + wP) / R + wdiv * 2.0 / R;
                const double w_avVT = (wP + wdiv) * ct / R;
                VR_(ix + 1, iy, iz) += w_dVRr * _dr + 0.5 * w_avVR;
                VR_(ix, iy, iz) += -w_dVRr * _dr + 0.5 * w_avVR;
                VT_(ix, iy + 1, iz) += w_dVTt * _dth / R + 0.5 * w_avVT;
                VT_(ix, iy, iz) += -w_dVTt * _dth / R + 0.5 * w_avVT;
                VP_(ix, iy, iz + 1) += w_dVPp * _dph / (R * sn);
                VP_(ix, iy, iz) -= w_dVPp * _dph / (R * sn);
            }
        }
    /* tau_Rth edges: t' = (t + G D)/(1 + G/E) */
    for (int iz = 1; iz <= np - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const size_t e = IDX(nr - 1, nt - 1, ix, iy, iz);
                const double a = lRT[e];
                const double Rxy = (((R1[ix] + R1[ix]) + R1[ix + 1]) + R1[ix + 1]) * 0.25;
                const double G = (((GS_(ix, iy, iz + 1) + GS_(ix, iy + 1, iz + 1)) + GS_(ix + 1, iy, iz + 1)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix, iy, iz + 1) + C_(eta, ix, iy + 1, iz + 1)) + C_(eta, ix + 1, iy, iz + 1)) +
                                  C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double den = 1.0 + G / E;
                lRT[e] = a / den;
                const double wE = a * s->tRT[e] * (G / (E * E)) / den * 0.25;
                C_(le, ix, iy, iz + 1) += wE; C_(le, ix, iy + 1, iz + 1) += wE;
                C_(le, ix + 1, iy, iz + 1) += wE; C_(le, ix + 1, iy + 1, iz + 1) += wE;
                const double w = a * G / den;
                /* D = (VT(i+1) - VT(i)) _dr + (1/Rxy)(VR(j+1) - VR(j)) _dth - avVT/Rxy */
                VT_(ix + 1, iy + 1, iz + 1) += w * (_dr - 0.5 / Rxy);
                VT_(ix, iy + 1, iz + 1) += w * (-_dr - 0.5 / Rxy);
                VR_(ix + 1, iy + 1, iz + 1) += w * _dth / Rxy;
                VR_(ix + 1, iy, iz + 1) -= w * _dth / Rxy;
            }
    /* tau_Rph edges */
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const size_t e = IDX(nr - 1, nt - 2, ix, iy, iz);
                const double a = lRP[e];
                const double Rxz = (((R1[ix] + R1[ix]) + R1[ix + 1]) + R1[ix + 1]) * 0.25;
                const double Txz = (((T1[iy + 1] + T1[iy + 1]) + T1[iy + 1]) + T1[iy + 1]) * 0.25;
                const double G = (((GS_(ix, iy + 1, iz) + GS_(ix, iy + 1, iz + 1)) + GS_(ix + 1, iy + 1, iz)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix, iy + 1, iz) + C_(eta, ix, iy + 1, iz + 1)) + C_(eta, ix + 1, iy + 1, iz)) +
                                  C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double den = 1.0 + G / E;
                lRP[e] = a / den;
                const double wE = a * s->tRP[e] * (G / (E * E)) / den * 0.25;
                C_(le, ix, iy + 1, iz) += wE; C_(le, ix, iy + 1, iz + 1) += wE;
                C_(le, ix + 1, iy + 1, iz) += wE; C_(le, ix + 1, iy + 1, iz + 1) += wE;
                const double w = a * G / den;
                const double rs = (1.0 / Rxz) / sin(Txz);
                VP_(ix + 1, iy + 1, iz + 1) += w * (_dr - 0.5 / Rxz);
                VP_(ix, iy + 1, iz + 1) += w * (-_dr - 0.5 / Rxz);
                VR_(ix + 1, iy + 1, iz + 1) += w * rs * _dph;
                VR_(ix + 1, iy + 1, iz) -= w * rs * _dph;
            }
    /* tau_thph edges */
    for (int iz = 1; iz <= np - 1; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const size_t e = IDX(nr - 2, nt - 1, ix, iy, iz);
                const double a = lTP[e];
                const double Ryz = (((R1[ix + 1] + R1[ix + 1]) + R1[ix + 1]) + R1[ix + 1]) * 0.25;
                const double Tyz = (((T1[iy] + T1[iy]) + T1[iy + 1]) + T1[iy + 1]) * 0.25;
                const double G = (((GS_(ix + 1, iy, iz) + GS_(ix + 1, iy, iz + 1)) + GS_(ix + 1, iy + 1, iz)) +
                                  GS_(ix + 1, iy + 1, iz + 1)) * 0.25;
                const double E = (((C_(eta, ix + 1, iy, iz) + C_(eta, ix + 1, iy, iz + 1)) + C_(eta, ix + 1, iy + 1, iz)) +
                                  C_(eta, ix + 1, iy + 1, iz + 1)) * 0.25;
                const double den = 1.0 + G / E;
                lTP[e] = a / den;
                const double wE = a * s->tTP[e] * (G / (E * E)) / den * 0.25;
                C_(le, ix + 1, iy, iz) += wE; C_(le, ix + 1, iy, iz + 1) += wE;
                C_(le, ix + 1, iy + 1, iz) += wE; C_(le, ix + 1, iy + 1, iz + 1) += wE;
                const double w = a * G / den;
                const double ct = 1.0 / tan(Tyz), rs = (1.0 / Ryz) / sin(Tyz);
                /* D = (1/Ryz)(VP(j+1)-VP(j)) _dth + rs (VT(k+1)-VT(k)) _dph - (avVP/Ryz) ct */
                VP_(ix + 1, iy + 1, iz + 1) += w * (_dth / Ryz - 0.5 * ct / Ryz);
                VP_(ix + 1, iy, iz + 1) += w * (-_dth / Ryz - 0.5 * ct / Ryz);
                VT_(ix + 1, iy + 1, iz + 1) += w * rs * _dph;
                VT_(ix + 1, iy + 1, iz) -= w * rs * _dph;
            }
#undef C_
#undef VR_
#undef VT_
#undef VP_
#undef GS_
    (void)lP;
}

void orc_adj_step(const orc_state *s, const orc_params *p, orc_adj *L, double *g_rho, double *g_ini)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const size_t nc = (size_t)nr * nt * np;
    adj_bc_ph(L, nr, nt, np, p->bc_mode);
    adj_bc_th(L, nr, nt, np, p->bc_mode);
    adj_bc_R(L, nr, nt, np);
    adj_compute_V(s, p, L, g_rho);
    double *ltII = (double *)calloc(nc, sizeof(double));
    adj_power_law(s, p, L, ltII, g_ini);
    adj_copy_bc(ltII, nr, nt, np);
    adj_compute_TII(s, L, ltII);
    free(ltII);
    adj_compute_P(s, p, L);
    if (p->relax == 1.0) memset(L->a[ORC_A_EPL], 0, nc * sizeof(double));
}

/* J = sum_f w_f VR_f / sum_f w_f over the interior r-faces whose centre
 * (R face, th, ph) lies inside the inclusion (the squared-distance test of
 * initialize_inclusion!), w_f = Rf^2 sin(th) (the cell volume up to the
 * constant dr dth dph).  Writes dJ/dVR into dJ (VR shape) if not null. */
double orc_adj_objective(const orc_state *s, double r, double radius, double *dJ)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const double *R1 = s->R1, *th1 = s->th1, *ph1 = s->ph1;
    double num = 0.0, wsum = 0.0;
    if (dJ) memset(dJ, 0, orc_adj_len(nr, nt, np, ORC_A_VR) * sizeof(double));
    for (int iz = 1; iz <= np; ++iz)
        for (int iy = 1; iy <= nt; ++iy)
            for (int ix = 2; ix <= nr; ++ix) {
                const double Rf = (R1[ix - 2] + R1[ix - 1]) * 0.5;
                const double a = Rf - r, b = (th1[iy - 1] - PI / 2.0) * r, c = ph1[iz - 1] * r;
                if (((a * a + b * b) + c * c) < radius) {
                    const double w = (Rf * Rf) * sin(th1[iy - 1]);
                    num += w * s->VR[IDX(nr + 1, nt, ix, iy, iz)];
                    wsum += w;
                    if (dJ) dJ[IDX(nr + 1, nt, ix, iy, iz)] = w;
                }
            }
    if (wsum == 0.0) return 0.0;
    if (dJ) {
        const size_t n = orc_adj_len(nr, nt, np, ORC_A_VR);
        for (size_t q = 0; q < n; ++q) dJ[q] /= wsum;
    }
    return num / wsum;
}

/* lam = T^T lam + dJ/dx until max|dlam|/max|lam| < tol (checked every
 * ncheck iterations on the velocity components); then the gradients.
 * Returns the iterations done (negative if not converged). */
int orc_adj_solve(const orc_state *s, const orc_params *p, double r, double radius, orc_adj *L, int maxiter,
                  int ncheck, double tol, double *g_rho, double *g_ini)
{
    const int nr = s->nr, nt = s->nth, np = s->nph;
    const size_t nVR = orc_adj_len(nr, nt, np, ORC_A_VR);
    double *b = (double *)malloc(nVR * sizeof(double));
    orc_adj_objective(s, r, radius, b);
    size_t len[ORC_A_N];
    double *prev[3];
    for (int q = 0; q < ORC_A_N; ++q) len[q] = orc_adj_len(nr, nt, np, q);
    for (int q = 0; q < 3; ++q) prev[q] = (double *)malloc(len[q] * sizeof(double));
    int it = 0, done = 0;
    while (it < maxiter) {
        const int check = ((it + 1) % ncheck) == 0;
        if (check)
            for (int q = 0; q < 3; ++q) memcpy(prev[q], L->a[q], len[q] * sizeof(double));
        orc_adj_step(s, p, L, NULL, NULL);
        for (size_t q = 0; q < nVR; ++q) L->a[ORC_A_VR][q] += b[q];
        ++it;
        if (check) {
            double err = 0.0;
            for (int q = 0; q < 3; ++q) {
                double md = 0.0, mv = 0.0;
                for (size_t i = 0; i < len[q]; ++i) {
                    const double d = fabs(L->a[q][i] - prev[q][i]), v = fabs(L->a[q][i]);
                    if (d > md) md = d;
                    if (v > mv) mv = v;
                }
                if (mv > 0.0 && md / mv > err) err = md / mv;
            }
            if (err < tol) { done = 1; break; }
        }
    }
    /* gradients: (dPhi/dm)^T lam, one more backward pass on a copy */
    if (g_rho || g_ini) {
        orc_adj W;
        for (int q = 0; q < ORC_A_N; ++q) {
            W.a[q] = (double *)malloc(len[q] * sizeof(double));
            memcpy(W.a[q], L->a[q], len[q] * sizeof(double));
        }
        orc_adj_step(s, p, &W, g_rho, g_ini);
        for (int q = 0; q < ORC_A_N; ++q) free(W.a[q]);
    }
    for (int q = 0; q < 3; ++q) free(prev[q]);
    free(b);
    return done ? it : -it;
}
