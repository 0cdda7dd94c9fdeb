// sspt_adjoint.hip -- adjoint of the PT iteration at a converged state.
//
// With x* = Phi(x*) the converged state of one PT iteration Phi (the
// compute_P! / compute_TII! + copy_BC! / power_law! / compute_V! / BC
// sequence, SphericalStokes_diapir.jl:413-425) and J(x) = the
// volume-weighted mean VR over the r-faces inside the inclusion,
//     dJ/dm = (dPhi/dm)^T lam,   lam = T^T lam + dJ/dx,   T = dPhi/dx (x*),
// for m = rho and eta_ini.  T is taken with the step sizes of x* fixed
// (etaSM through KBDT, GSDT and DT_R): they multiply residuals that vanish
// at x*, so this is the exact Jacobian there.  The iteration for lam is the
// transpose of the forward PT map: same spectrum, same convergence rate, and
// it needs only the converged state, no record of the forward iterations.
//
// T^T is applied in gather form, one kernel per family of outputs, with no
// atomics (each output sums the coefficients of the forward statements that
// read it); the shifted storage of the shear stresses makes every transposed
// stencil a +-1 offset in the box frame:
//   1. BC fold: ghost adjoints go to the points their values came from
//      (BC_ph!, BC_th!, BC_R! backwards);
//   2. power law: lam_tII of every cell from lam_eta (and lam_eta_PL);
//   3. interior cells: Q = (sum of lam_tII over the cells that copy_BC!
//      filled from them) / tII;
//   4. centres and edges (one kernel): adjoints of the new P and tau
//      (compute_V!'s and compute_TII!'s transposes), written scaled by
//      compute_P!'s 1/den;
//   5. faces and centres (one kernel): lam_V (compute_P!'s strain-rate
//      transposes) and lam_eta (its viscosity dependence).
#include "sspt_internal.h"
#include "sspt_launch.h"
#include "sspt_adj_fused_body.h"
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace sspt {

// adjoint boxes held by the context
enum {
    AD_SETA = 0,   // lam set A: VR VT VP P RR TT PP RT RP TP E PL (12)
    AD_SETB = 12,  // lam set B
    AD_PREV = 24,  // VR VT VP of the previous iterate (convergence check)
    AD_LT = 27, AD_Q = 28, AD_BJ = 29, AD_GRHO = 30, AD_GINI = 31,
    AD_SID = 32, AD_SGS = 33, AD_SKB = 34, AD_SAE = 35, AD_SX = 36, AD_SY = 37, AD_SZ = 38, AD_SIT = 39,
    AD_GRT = 40, AD_IDRT = 41, AD_CRT = 42, AD_GRP = 43, AD_IDRP = 44, AD_CRP = 45,
    AD_GTP = 46, AD_IDTP = 47, AD_CTP = 48,
    AD_LB = 49,    // boundary power-law sums of the cells next to the boundary (r-marching kernel)
    AD_NBOX = 50
};
static_assert(AD_NBOX == SS_ADJ_NBOX, "adjoint box count");
enum { L_VR, L_VT, L_VP, L_P, L_RR, L_TT, L_PP, L_RT, L_RP, L_TP, L_E, L_PL, L_N };

struct AdjSet {
    double *f[L_N];
};

struct AdjArgs {
    Box b;
    Tables t;
    AdjSet in, out;
    // converged forward state
    const double *tRR, *tTT, *tPP, *tRT, *tRP, *tTP, *eta, *ini, *ePL, *tII, *DT, *SM, *VR;
    // fixed coefficients of the linearisation
    double *sID, *sGS, *sKB, *sAE, *sX, *sY, *sZ, *sIT;
    double *gRT, *idRT, *cRT, *gRP, *idRP, *cRP, *gTP, *idTP, *cTP;
    double *LT, *Q, *bJ, *gRho, *gIni;
    double _dr, _dth, _dph, g, kb, gs, tC, ex, relax, emin, emax;
    int do_relax, do_lo, do_hi, grad, bc_mode;
};

#define A_LOOP(NI, NJ, NK)                                        \
    const int k = blockIdx.x * blockDim.x + threadIdx.x;          \
    const int j = blockIdx.y * blockDim.y + threadIdx.y;          \
    const int i = blockIdx.z;                                     \
    if (i >= (NI) || j >= (NJ) || k >= (NK)) return;

static const dim3 AB(64, 4, 1);
static inline dim3 agrid(int ni, int nj, int nk) { return dim3((nk + 63) / 64, (nj + 3) / 4, ni); }

// ------------------------------------------------------------------ ranges
// updated faces of compute_V! and the edge boxes, in box indices
__device__ __forceinline__ bool upd_VR(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 2;
}
__device__ __forceinline__ bool upd_VT(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 2;
}
__device__ __forceinline__ bool upd_VP(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 1;
}
__device__ __forceinline__ bool in_RT(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 2;
}
__device__ __forceinline__ bool in_RP(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 1;
}
__device__ __forceinline__ bool in_TP(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 1;
}
__device__ __forceinline__ bool inner_c(const Box &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 2;
}

// w_f = DT_f lam_V_f of the faces compute_V! updates (DT_f: average of the
// two centres), 0 elsewhere
__device__ __forceinline__ double w_VR(const AdjArgs &a, int i, int j, int k)
{
    if (!upd_VR(a.b, i, j, k)) return 0.0;
    const long long o = a.b.off(i, j, k);
    return ((a.DT[o - a.b.si] + a.DT[o]) * 0.5) * a.in.f[L_VR][o];
}
__device__ __forceinline__ double w_VT(const AdjArgs &a, int i, int j, int k)
{
    if (!upd_VT(a.b, i, j, k)) return 0.0;
    const long long o = a.b.off(i, j, k);
    return ((a.DT[o - a.b.sj] + a.DT[o]) * 0.5) * a.in.f[L_VT][o];
}
__device__ __forceinline__ double w_VP(const AdjArgs &a, int i, int j, int k)
{
    if (!upd_VP(a.b, i, j, k)) return 0.0;
    const long long o = a.b.off(i, j, k);
    return ((a.DT[o - 1] + a.DT[o]) * 0.5) * a.in.f[L_VP][o];
}

// ------------------------------------------------------------------ fixed coefficients
__global__ void __launch_bounds__(256) k_adj_prep_centre(AdjArgs a)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    const Box &b = a.b;
    const long long o = b.off(i, j, k);
    const double GS = a.gs * a.SM[o], e = a.eta[o];
    const double den = 1.0 + GS / e;
    a.sGS[o] = GS;
    a.sKB[o] = a.kb * a.SM[o];
    a.sID[o] = 1.0 / den;
    a.sAE[o] = GS / (e * e);
    double it = 0.0, x = 0.0, y = 0.0, z = 0.0;
    if (inner_c(b, i, j, k) && a.tII[o] > 0.0) {
        it = 1.0 / a.tII[o];
        x = (((a.tRT[o] + a.tRT[o + b.sj]) + a.tRT[o + b.si]) + a.tRT[o + b.si + b.sj]) * 0.25;
        y = (((a.tRP[o] + a.tRP[o + 1]) + a.tRP[o + b.si]) + a.tRP[o + b.si + 1]) * 0.25;
        z = (((a.tTP[o] + a.tTP[o + 1]) + a.tTP[o + b.sj]) + a.tTP[o + b.sj + 1]) * 0.25;
    }
    a.sIT[o] = it;
    a.sX[o] = 0.25 * x;
    a.sY[o] = 0.25 * y;
    a.sZ[o] = 0.25 * z;
}

// edge families: 0 = tau_Rth, 1 = tau_Rph, 2 = tau_thph
__global__ void __launch_bounds__(256) k_adj_prep_edge(AdjArgs a, int fam)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    const Box &b = a.b;
    const long long o = b.off(i, j, k);
    long long c00, c01, c10;
    const double *tau;
    double *G, *ID, *CC;
    if (fam == 0) {
        if (!in_RT(b, i, j, k)) return;
        c00 = o - b.si - b.sj; c01 = o - b.si; c10 = o - b.sj; tau = a.tRT; G = a.gRT; ID = a.idRT; CC = a.cRT;
    } else if (fam == 1) {
        if (!in_RP(b, i, j, k)) return;
        c00 = o - b.si - 1; c01 = o - b.si; c10 = o - 1; tau = a.tRP; G = a.gRP; ID = a.idRP; CC = a.cRP;
    } else {
        if (!in_TP(b, i, j, k)) return;
        c00 = o - b.sj - 1; c01 = o - b.sj; c10 = o - 1; tau = a.tTP; G = a.gTP; ID = a.idTP; CC = a.cTP;
    }
    const double g4 = (((a.gs * a.SM[c00] + a.gs * a.SM[c01]) + a.gs * a.SM[c10]) + a.gs * a.SM[o]) * 0.25;
    const double e4 = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]) * 0.25;
    G[o] = g4;
    ID[o] = 1.0 / (1.0 + g4 / e4);
    CC[o] = 0.25 * tau[o] * (g4 / (e4 * e4));
}

// ------------------------------------------------------------------ 1. BC fold
// Threads cover the two other directions; each thread handles both sides of
// its line (with 3 points in a direction both sides fold onto the same point).
template <typename R> __global__ void k_adj_bc_ph(Box b, int bc_mode, R *VR, R *VT, R *VP)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const bool cp = bc_mode != SS_BC_SHEAR;
    if (i < nr && j < nt) { // Vph: ghost = c - inner
        const long long lo = b.off(i, j, 0), hi = b.off(i, j, np);
        VP[lo + 1] -= VP[lo]; VP[lo] = R(0);
        VP[hi - 1] -= VP[hi]; VP[hi] = R(0);
    }
    if (i < nr && j <= nt) {
        const long long lo = b.off(i, j, 0), hi = b.off(i, j, np - 1);
        if (cp) { VT[lo + 1] += VT[lo]; VT[hi - 1] += VT[hi]; }
        VT[lo] = R(0); VT[hi] = R(0);
    }
    if (i <= nr && j < nt) {
        const long long lo = b.off(i, j, 0), hi = b.off(i, j, np - 1);
        if (cp) { VR[lo + 1] += VR[lo]; VR[hi - 1] += VR[hi]; }
        VR[lo] = R(0); VR[hi] = R(0);
    }
}

template <typename R> __global__ void k_adj_bc_th(Box b, int bc_mode, R *VR, R *VT, R *VP)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const bool cp = bc_mode != SS_BC_SHEAR;
    const long long sj = b.sj;
    if (i < nr && k <= np) {
        const long long lo = b.off(i, 0, k), hi = b.off(i, nt - 1, k);
        if (cp) { VP[lo + sj] += VP[lo]; VP[hi - sj] += VP[hi]; }
        VP[lo] = R(0); VP[hi] = R(0);
    }
    if (i < nr && k < np) { // Vth: ghost = c - inner
        const long long lo = b.off(i, 0, k), hi = b.off(i, nt, k);
        VT[lo + sj] -= VT[lo]; VT[lo] = R(0);
        VT[hi - sj] -= VT[hi]; VT[hi] = R(0);
    }
    if (i <= nr && k < np) {
        const long long lo = b.off(i, 0, k), hi = b.off(i, nt - 1, k);
        if (cp) { VR[lo + sj] += VR[lo]; VR[hi - sj] += VR[hi]; }
        VR[lo] = R(0); VR[hi] = R(0);
    }
}

template <typename R> __global__ void k_adj_bc_R(Box b, int bc_mode, R *VR, R *VT, R *VP)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const long long si = b.si;
    if (j < nt && k <= np) {
        const long long lo = b.off(0, j, k), hi = b.off(nr - 1, j, k);
        VP[lo + si] += VP[lo]; VP[lo] = R(0);
        VP[hi - si] += VP[hi]; VP[hi] = R(0);
    }
    if (j <= nt && k < np) {
        const long long lo = b.off(0, j, k), hi = b.off(nr - 1, j, k);
        VT[lo + si] += VT[lo]; VT[lo] = R(0);
        VT[hi - si] += VT[hi]; VT[hi] = R(0);
    }
    if (j < nt && k < np) { // VR: ghost = -inner
        const long long lo = b.off(0, j, k), hi = b.off(nr, j, k);
        VR[lo + si] -= VR[lo]; VR[lo] = R(0);
        VR[hi - si] -= VR[hi]; VR[hi] = R(0);
    }
}

// ------------------------------------------------------------------ 2. power law
__global__ void __launch_bounds__(256) k_adj_power_law(AdjArgs a)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    const long long o = a.b.off(i, j, k);
    const double le = a.in.f[L_E][o];
    const double tii = a.tII[o], ini = a.ini[o];
    const double pl0 = ini * pow(tii / a.tC, a.ex);
    const double pl = a.do_relax ? a.ePL[o] : pl0;
    const double e_raw = 2.0 / (1.0 / ini + 1.0 / pl);
    const bool clamped = (a.do_lo && e_raw < a.emin) || (a.do_hi && e_raw > a.emax);
    const double a_e = clamped ? 0.0 : le;
    if (!isfinite(pl) || !isfinite(pl0)) { // tII = 0: eta = 2 eta_ini
        a.LT[o] = 0.0;
        if (a.grad) a.gIni[o] += a_e * 2.0;
        if (a.do_relax) a.out.f[L_PL][o] = 0.0;
        return;
    }
    double l_pl = a_e * (e_raw * e_raw) / (2.0 * pl * pl);
    double l_ini = a_e * (e_raw * e_raw) / (2.0 * ini * ini);
    if (a.do_relax) {
        l_pl += a.in.f[L_PL][o];
        const double l_pl0 = l_pl * a.relax * pl / pl0;
        a.out.f[L_PL][o] = l_pl * (1.0 - a.relax) * pl / a.ePL[o];
        l_pl = l_pl0;
    } // without relaxation eta_PL is no state: its adjoint stays 0 in both sets
    a.LT[o] = (tii > 0.0) ? l_pl * a.ex * pl0 / tii : 0.0;
    l_ini += l_pl * pl0 / ini;
    if (a.grad) a.gIni[o] += l_ini;
}

// ------------------------------------------------------------------ 3. tII of interior cells
__global__ void __launch_bounds__(256) k_adj_tii(AdjArgs a)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    const Box &b = a.b;
    const long long o = b.off(i, j, k);
    if (!inner_c(b, i, j, k)) { a.Q[o] = 0.0; return; }
    // copy_BC! filled a boundary cell from the interior cell at its clamped
    // position: this cell collects its own lam_tII and those of the boundary
    // cells that clamp onto it
    int di[3] = {0, 0, 0}, dj[3] = {0, 0, 0}, dk[3] = {0, 0, 0};
    int ni = 1, nj = 1, nk = 1;
    if (i == 1) di[ni++] = -1;
    if (i == b.nr - 2) di[ni++] = 1;
    if (j == 1) dj[nj++] = -1;
    if (j == b.nt - 2) dj[nj++] = 1;
    if (k == 1) dk[nk++] = -1;
    if (k == b.np - 2) dk[nk++] = 1;
    double s = 0.0;
    for (int p = 0; p < ni; ++p)
        for (int q = 0; q < nj; ++q)
            for (int r = 0; r < nk; ++r) s += a.LT[o + di[p] * b.si + dj[q] * b.sj + dk[r]];
    a.Q[o] = s * a.sIT[o];
}

// ------------------------------------------------------------------ 4. new-stress adjoints
// One thread per box index (i,j,k): the centre (i,j,k) and the three edges
// stored at (i,j,k) share the face weights they gather.  The adjoints of the
// new P and tau (compute_V!'s and compute_TII!'s transposes) are written
// scaled by compute_P!'s 1/den: the out set holds lam_P' and lam_tau'/den.
__device__ __forceinline__ void adj_sig_point(const AdjArgs &a, const int i, const int j, const int k)
{
    const Box &b = a.b;
    const Tables &t = a.t;
    const int nt = b.nt;
    const long long o = b.off(i, j, k);
    const bool eRT = in_RT(b, i, j, k), eRP = in_RP(b, i, j, k), eTP = in_TP(b, i, j, k);
    // face weights (0 outside compute_V!'s ranges)
    const double wR = w_VR(a, i, j, k), wRi = w_VR(a, i + 1, j, k), wRj = w_VR(a, i, j - 1, k),
                 wRk = w_VR(a, i, j, k - 1);
    const double wT = w_VT(a, i, j, k), wTj = w_VT(a, i, j + 1, k), wTi = w_VT(a, i - 1, j, k),
                 wTk = w_VT(a, i, j, k - 1);
    const double wP = w_VP(a, i, j, k), wPk = w_VP(a, i, j, k + 1), wPi = w_VP(a, i - 1, j, k),
                 wPj = w_VP(a, i, j - 1, k);
    double vP = 0.0, vRR = 0.0, vTT = 0.0, vPP = 0.0, vRT = 0.0, vRP = 0.0, vTP = 0.0;
    { // the centre (every (i,j,k) of this grid is one)
        double sRR = 0.0, sTT = 0.0, sPP = 0.0;
        if (wR != 0.0) {
            const double iR = 1.0 / t.Rf[i];
            sRR += wR * (a._dr + iR);
            sTT -= 0.5 * wR * iR;
            sPP -= 0.5 * wR * iR;
        }
        if (wRi != 0.0) {
            const double iR = 1.0 / t.Rf[i + 1];
            sRR += wRi * (-a._dr + iR);
            sTT -= 0.5 * wRi * iR;
            sPP -= 0.5 * wRi * iR;
        }
        const double iRy = 1.0 / t.Rc2[i];
        if (wT != 0.0) {
            const double ct = t.cotf2[j];
            sTT += wT * (a._dth * iRy + 0.5 * ct * iRy);
            sPP -= 0.5 * wT * ct * iRy;
        }
        if (wTj != 0.0) {
            const double ct = t.cotf2[j + 1];
            sTT += wTj * (-a._dth * iRy + 0.5 * ct * iRy);
            sPP -= 0.5 * wTj * ct * iRy;
        }
        sPP += (wP - wPk) * (t.rs_z[i * (nt + 1) + j] * a._dph);
        double nRR = a.in.f[L_RR][o] + sRR, nTT = a.in.f[L_TT][o] + sTT, nPP = a.in.f[L_PP][o] + sPP;
        if (inner_c(b, i, j, k)) {
            const double q = 0.5 * a.Q[o];
            nRR += q * a.tRR[o];
            nTT += q * a.tTT[o];
            nPP += q * a.tPP[o];
        }
        vP = a.in.f[L_P][o] - ((sRR + sTT) + sPP);
        const double id = a.sID[o];
        vRR = nRR * id;
        vTT = nTT * id;
        vPP = nPP * id;
    }
    if (eRT) {
        const double iR = 1.0 / t.Rf[i];
        double s = wRj * (a._dth * iR + 0.5 * t.cotc2[j - 1] * iR) + wR * (-a._dth * iR + 0.5 * t.cotc2[j] * iR);
        s += wTi * (a._dr + 1.5 / t.Rc2[i - 1]) + wT * (-a._dr + 1.5 / t.Rc2[i]);
        double tq = 0.0;
        for (int di = 0; di <= 1; ++di)
            for (int dj = 0; dj <= 1; ++dj)
                if (inner_c(b, i - di, j - dj, k)) {
                    const long long c = o - di * b.si - dj * b.sj;
                    tq += a.Q[c] * a.sX[c];
                }
        vRT = ((a.in.f[L_RT][o] + s) + tq) * a.idRT[o];
    }
    if (eRP) {
        double s = (wRk - wR) * (t.rs_f[i * (nt + 1) + j] * a._dph);
        s += wPi * (a._dr + 1.5 / t.Rc2[i - 1]) + wP * (-a._dr + 1.5 / t.Rc2[i]);
        double tq = 0.0;
        for (int di = 0; di <= 1; ++di)
            for (int dk = 0; dk <= 1; ++dk)
                if (inner_c(b, i - di, j, k - dk)) {
                    const long long c = o - di * b.si - dk;
                    tq += a.Q[c] * a.sY[c];
                }
        vRP = ((a.in.f[L_RP][o] + s) + tq) * a.idRP[o];
    }
    if (eTP) {
        double s = (wTk - wT) * (t.rs_y[i * (nt + 1) + j] * a._dph);
        const double iR = 1.0 / t.Rc2[i];
        s += wPj * (a._dth * iR + t.cotc2[j - 1] * iR) + wP * (-a._dth * iR + t.cotc2[j] * iR);
        double tq = 0.0;
        for (int dj = 0; dj <= 1; ++dj)
            for (int dk = 0; dk <= 1; ++dk)
                if (inner_c(b, i, j - dj, k - dk)) {
                    const long long c = o - dj * b.sj - dk;
                    tq += a.Q[c] * a.sZ[c];
                }
        vTP = ((a.in.f[L_TP][o] + s) + tq) * a.idTP[o];
    }
    // stores last (the loads above may then be scheduled freely)
    a.out.f[L_P][o] = vP;
    a.out.f[L_RR][o] = vRR;
    a.out.f[L_TT][o] = vTT;
    a.out.f[L_PP][o] = vPP;
    if (eRT) a.out.f[L_RT][o] = vRT;
    if (eRP) a.out.f[L_RP][o] = vRP;
    if (eTP) a.out.f[L_TP][o] = vTP;
}

// ------------------------------------------------------------------ 5. lam_V and lam_eta
// Weights of compute_P!'s centre stencil from the new-stress adjoints of
// cell c: (w_dVRr, w_dVTt, w_dVPp, w_avVR, w_avVT).  The out set holds
// lam_tnn' / den, and (2 GS / den) lam_tnn' = 2 GS lam_out.
struct CW {
    double dR, dT, dP, aR, aT;
};
__device__ __forceinline__ CW centre_w(const AdjArgs &a, long long c, int i, int j)
{
    const double GS2 = 2.0 * a.sGS[c];
    const double wR = GS2 * a.out.f[L_RR][c], wT = GS2 * a.out.f[L_TT][c], wP = GS2 * a.out.f[L_PP][c];
    const double wdiv = -a.sKB[c] * a.out.f[L_P][c] - (wR + wT + wP) * (1.0 / 3.0);
    const double iR = a.t.invRc[i];
    CW w;
    w.dR = wR + wdiv;
    w.dT = wT + wdiv;
    w.dP = wP + wdiv;
    w.aR = (wT + wP) * iR + 2.0 * wdiv * iR;
    w.aT = (wP + wdiv) * a.t.cotc[j] * iR;
    return w;
}

// One thread per box index: VR, Vth and Vph faces (i,j,k) and the centre
// viscosity (i,j,k).  Each face takes compute_P!'s centre stencil from the
// cell above and below it and the shear-edge stencils around it.
__device__ __forceinline__ void adj_vel_point(const AdjArgs &a, const int i, const int j, const int k)
{
    const Box &b = a.b;
    const Tables &t = a.t;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const long long o = b.off(i, j, k);
    const bool fR = j < nt && k < np, fT = i < nr && k < np, fP = i < nr && j < nt, cen = i < nr && j < nt && k < np;
    CW w0{0, 0, 0, 0, 0};
    if (cen) w0 = centre_w(a, o, i, j);
    // shear-edge terms g * lam_out of the edges around
    const bool rt = in_RT(b, i, j, k), rp = in_RP(b, i, j, k), tp = in_TP(b, i, j, k);
    const double sRT = rt ? a.gRT[o] * a.out.f[L_RT][o] : 0.0;
    const double sRP = rp ? a.gRP[o] * a.out.f[L_RP][o] : 0.0;
    const double sTP = tp ? a.gTP[o] * a.out.f[L_TP][o] : 0.0;
    double vR = 0.0, vT = 0.0, vP = 0.0, vE = 0.0;
    if (fR) {
        double v = a.in.f[L_VR][o];
        if (i >= 1) {
            const CW w = centre_w(a, o - b.si, i - 1, j);
            v += w.dR * a._dr + 0.5 * w.aR;
        }
        if (i <= nr - 1) v += -w0.dR * a._dr + 0.5 * w0.aR;
        if (rt) v += sRT * a._dth / t.Re[i];
        if (in_RT(b, i, j + 1, k)) v -= a.gRT[o + b.sj] * a.out.f[L_RT][o + b.sj] * a._dth / t.Re[i];
        if (rp || in_RP(b, i, j, k + 1)) {
            const double re = t.rs_e[i * (nt + 1) + j] * a._dph;
            v += sRP * re;
            if (in_RP(b, i, j, k + 1)) v -= a.gRP[o + 1] * a.out.f[L_RP][o + 1] * re;
        }
        if (a.bJ) v += a.bJ[o];
        vR = v;
    }
    if (fT) {
        double v = a.in.f[L_VT][o];
        const double dth_R = a._dth * t.invRc[i];
        if (j >= 1) {
            const CW w = centre_w(a, o - b.sj, i, j - 1);
            v += w.dT * dth_R + 0.5 * w.aT;
        }
        if (j <= nt - 1) v += -w0.dT * dth_R + 0.5 * w0.aT;
        if (rt) v += sRT * (a._dr - 0.5 / t.Re[i]);
        if (in_RT(b, i + 1, j, k)) v += a.gRT[o + b.si] * a.out.f[L_RT][o + b.si] * (-a._dr - 0.5 / t.Re[i + 1]);
        if (tp || in_TP(b, i, j, k + 1)) {
            const double r4 = t.rs_c4[i * (nt + 1) + j] * a._dph;
            v += sTP * r4;
            if (in_TP(b, i, j, k + 1)) v -= a.gTP[o + 1] * a.out.f[L_TP][o + 1] * r4;
        }
        vT = v;
    }
    if (fP) {
        double v = a.in.f[L_VP][o];
        const double rc = t.rs_c[i * (nt + 1) + j] * a._dph;
        if (k >= 1) {
            const CW w = centre_w(a, o - 1, i, j);
            v += w.dP * rc;
        }
        if (k <= np - 1) v -= w0.dP * rc;
        if (rp) v += sRP * (a._dr - 0.5 / t.Re[i]);
        if (in_RP(b, i + 1, j, k)) v += a.gRP[o + b.si] * a.out.f[L_RP][o + b.si] * (-a._dr - 0.5 / t.Re[i + 1]);
        if (tp) {
            const double iR = 1.0 / t.Rc4[i];
            v += sTP * (a._dth * iR - 0.5 * t.cotf4[j] * iR);
        }
        if (in_TP(b, i, j + 1, k)) {
            const double iR = 1.0 / t.Rc4[i];
            v += a.gTP[o + b.sj] * a.out.f[L_TP][o + b.sj] * (-a._dth * iR - 0.5 * t.cotf4[j + 1] * iR);
        }
        vP = v;
    }
    if (cen) {
        double v = ((a.out.f[L_RR][o] * a.tRR[o] + a.out.f[L_TT][o] * a.tTT[o]) + a.out.f[L_PP][o] * a.tPP[o]) *
                   a.sAE[o];
        for (int d1 = 0; d1 <= 1; ++d1)
            for (int d2 = 0; d2 <= 1; ++d2) {
                if (in_RT(b, i + d1, j + d2, k)) {
                    const long long e = o + d1 * b.si + d2 * b.sj;
                    v += a.cRT[e] * a.out.f[L_RT][e];
                }
                if (in_RP(b, i + d1, j, k + d2)) {
                    const long long e = o + d1 * b.si + d2;
                    v += a.cRP[e] * a.out.f[L_RP][e];
                }
                if (in_TP(b, i, j + d1, k + d2)) {
                    const long long e = o + d1 * b.sj + d2;
                    v += a.cTP[e] * a.out.f[L_TP][e];
                }
            }
        vE = v;
    }
    if (fR) a.out.f[L_VR][o] = vR;
    if (fT) a.out.f[L_VT][o] = vT;
    if (fP) a.out.f[L_VP][o] = vP;
    if (cen) a.out.f[L_E][o] = vE;
}

// One r-plane of a 64 x 4 (phi, th) tile per workgroup, planes slowest: the
// workgroups of a few consecutive planes run together, so the i-1 / i+1
// planes the stencils read are mostly served by the last-level cache.
// (Walking several planes per workgroup, or dealing a tile's planes to one
// XCD, measured slower at 511^3.)
__global__ void __launch_bounds__(256) k_adj_sig(AdjArgs a)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    adj_sig_point(a, i, j, k);
}

__global__ void __launch_bounds__(256) k_adj_vel(AdjArgs a)
{
    A_LOOP(a.b.nr + 1, a.b.nt + 1, a.b.np + 1)
    adj_vel_point(a, i, j, k);
}

// gradient with respect to rho: compute_V!'s -rho g on VR faces
__global__ void __launch_bounds__(256) k_adj_grad_rho(AdjArgs a)
{
    A_LOOP(a.b.nr, a.b.nt, a.b.np)
    const long long o = a.b.off(i, j, k);
    a.gRho[o] += -0.5 * a.g * (w_VR(a, i, j, k) + w_VR(a, i + 1, j, k));
}

// ------------------------------------------------------------------ objective
// J = sum w VR / sum w over interior r-faces inside the inclusion; bJ = w
// (scaled by 1/sum w afterwards).  Per-block partial sums, then one block.
constexpr int OBJ_BLOCK = 256;
__global__ void __launch_bounds__(OBJ_BLOCK) k_adj_obj_partial(AdjArgs a, double r, double radius, double *part)
{
    const Box &b = a.b;
    const long long nface = (long long)(b.nr + 1) * b.nt * b.np;
    double sw = 0.0, swv = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nface; q += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(q % b.np), j = (int)((q / b.np) % b.nt), i = (int)(q / ((long long)b.np * b.nt));
        const long long o = b.off(i, j, k);
        double w = 0.0;
        if (i >= 1 && i <= b.nr - 1) {
            const double Rf = a.t.Rf[i];
            const double x = Rf - r, y = (a.t.th[j] - 3.14159265358979323846 / 2.0) * r, z = a.t.ph[k] * r;
            if (((x * x + y * y) + z * z) < radius) w = (Rf * Rf) * sin(a.t.th[j]);
        }
        a.bJ[o] = w;
        sw += w;
        swv += w * a.VR[o];
    }
    __shared__ double s1[OBJ_BLOCK], s2[OBJ_BLOCK];
    s1[threadIdx.x] = sw;
    s2[threadIdx.x] = swv;
    __syncthreads();
    for (int h = OBJ_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ void __launch_bounds__(OBJ_BLOCK) k_adj_obj_final(const double *part, int nblk, double *out)
{
    __shared__ double s1[OBJ_BLOCK], s2[OBJ_BLOCK];
    double a1 = 0.0, a2 = 0.0;
    for (int q = threadIdx.x; q < nblk; q += OBJ_BLOCK) {
        a1 += part[2 * q];
        a2 += part[2 * q + 1];
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int h = OBJ_BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

__global__ void __launch_bounds__(256) k_adj_scale(Box b, double *A, int n1, int n2, int n3, double s)
{
    A_LOOP(n1, n2, n3)
    A[b.off(i, j, k)] *= s;
}

__global__ void k_adj_zero(double *p, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = 0.0;
}

// ------------------------------------------------------------------ r-marching form
// The whole of steps 2-5 in one launch (sspt_adj_fused_body.h): a block of
// 64 x ATJ columns marches along r through its chunk of planes.  Tiles are
// dealt so that each XCD gets a contiguous run of them.
// The kernel's arguments re-read from the kernarg segment.  The pointer is
// passed through an empty asm statement at each use, so the compiler cannot
// hoist the loads of the argument fields out of the plane loop: each phase
// loads (with scalar loads) the 20-odd field bases it uses, instead of all
// 40-odd staying live in scalar registers across the loop and spilling.
template <typename R> __device__ __forceinline__ const adjf::AArgs<R> &kargs()
{
    uint64_t v = (uint64_t)(uintptr_t)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(v));
    return *(const adjf::AArgs<R> *)(const __attribute__((address_space(4))) adjf::AArgs<R> *)v;
}

// FP32: two workgroups per CU (6 wavefronts per SIMD: at most 80 VGPRs, and
// 2 x 75 KB of LDS)
template <typename R> constexpr int ADJ_WAVES_PER_EU = sizeof(R) == 4 ? 6 : 1;
template <typename R, bool RHS>
__global__ void __launch_bounds__(fused::TK * adjf::ATJ, ADJ_WAVES_PER_EU<R>)
    k_adj_fused(adjf::AArgs<R> a, int nx, int ny, int nz)
{
    using namespace adjf;
    __shared__ ALds<R> L;
    const int T_ = nx * ny * nz;
    const int per = (T_ + 7) / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    const int bx = t % nx, by = (t / nx) % ny, bz = t / (nx * ny);
    if (bz >= nz) return; // padding slots of the remap (uniform per block)
    const int kb = bx * AOK, jb = by * AOJ;
    const int s0 = bz * a.rchunk, s1 = min(s0 + a.rchunk, a.b.nr + 1);
    const int start = max(s0 - 1, 0), stop = min(s1, a.b.nr);
    AThread<R> T;
    athread_init(T, a, threadIdx.x, threadIdx.y, kb, jb);
    aprologue(T, L, a, start);
    AStepA<R> na;
    aloadA(na, a, T, start);
    __syncthreads(); // the first phase A reads the DT(start) plane the prologue wrote
    for (int q = start; q <= stop; ++q) {
        const bool wr = q >= s0 && q < s1, wr_m = q - 1 >= s0 && q - 1 < s1;
        const AStepA<R> ca = na;
        AStepB<R> cb;
        aloadB(cb, kargs<R>(), T, q);
        aphaseA<R, RHS>(T, L, kargs<R>(), ca, q, wr);
        __syncthreads();
        aphaseB<R, RHS>(T, L, kargs<R>(), cb, q, wr);
        __syncthreads();
        aloadA(na, kargs<R>(), T, q + 1); // next step's (clamped into the box at the end)
        // phase C reads only the phase-B arrays; the next phase A writes only
        // the phase-A arrays, so no barrier is needed between them
        aphaseC<R, RHS>(T, L, kargs<R>(), q, wr, wr_m);
    }
}

// LB of the interior cells next to the boundary: the power-law adjoints of
// the boundary cells copy_BC! fills from each.  Families: 0 -> i in {1,
// nr-2}; 1 -> j in {1, nt-2} (i inside); 2 -> k in {1, np-2} (i, j inside);
// each such cell once.
template <typename R> __global__ void __launch_bounds__(256) k_adj_lb(adjf::AArgs<R> a, R *LB)
{
    const adjf::BoxG &b = a.b;
    const int u = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y * blockDim.y + threadIdx.y;
    const int fam = blockIdx.z, nr = b.nr, nt = b.nt, np = b.np;
    for (int side = 0; side < 2; ++side) {
        int i, j, k;
        if (fam == 0) {
            i = side ? nr - 2 : 1; j = v; k = u;
            if (side && nr - 2 == 1) continue;
            if (j < 1 || j > nt - 2 || k < 1 || k > np - 2) continue;
        } else if (fam == 1) {
            j = side ? nt - 2 : 1; i = v; k = u;
            if (side && nt - 2 == 1) continue;
            if (i < 2 || i > nr - 3 || k < 1 || k > np - 2) continue;
        } else {
            k = side ? np - 2 : 1; i = v; j = u;
            if (side && np - 2 == 1) continue;
            if (i < 2 || i > nr - 3 || j < 2 || j > nt - 3) continue;
        }
        LB[b.off(i, j, k)] = adjf::lb_sum(a, i, j, k);
    }
}

// ------------------------------------------------------------------ FP32 correction
// (ss_adjoint_solve_mixed) elementwise over whole boxes: the FP32 copy of an
// FP64 box, the FP32 residual Y - mu, and mu += d in FP64
__global__ void __launch_bounds__(256) k_adj_to_f32(float *__restrict__ D, const double *__restrict__ S, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        D[q] = (float)S[q];
}
__global__ void __launch_bounds__(256) k_adj_resid_f32(float *__restrict__ D, const double *__restrict__ Y,
                                                       const double *__restrict__ X, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        D[q] = (float)(Y[q] - X[q]);
}
__global__ void __launch_bounds__(256) k_adj_acc_f32(double *__restrict__ X, const float *__restrict__ d, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        X[q] = X[q] + (double)d[q];
}
__global__ void __launch_bounds__(256) k_adj_zero_f32(float *__restrict__ p, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.f;
}
__global__ void k_adj_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Increment test of the FP32 correction sweeps: max|new - old| (red[0..2])
// and max|new| (red[3..5]) of lam_VR, lam_Vth, lam_Vph over the faces
// compute_V! updates (both sets BC-folded), as the bit patterns of
// non-negative doubles (ordered as integers); a non-finite value anywhere
// sets red[0] to the NaN pattern above every finite one.
// ADJ_CHECK_PLANES r-planes per thread, one set of atomics per wavefront.
constexpr int ADJ_CHECK_PLANES = 32;
__global__ void __launch_bounds__(256) k_adj_inc_f32(Box b, const float *__restrict__ n0, const float *__restrict__ n1,
                                                     const float *__restrict__ n2, const float *__restrict__ o0,
                                                     const float *__restrict__ o1, const float *__restrict__ o2,
                                                     unsigned long long *red)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    double me[3] = {0.0, 0.0, 0.0}, mv[3] = {0.0, 0.0, 0.0};
    int bad = 0; // fmax drops NaN: non-finite values are flagged apart
    const int i0 = blockIdx.z * ADJ_CHECK_PLANES, i1 = min(i0 + ADJ_CHECK_PLANES, b.nr + 1);
    for (int i = i0; i < i1 && j <= b.nt && k <= b.np; ++i) {
        const long long o = b.off(i, j, k);
        const bool in[3] = {upd_VR(b, i, j, k), upd_VT(b, i, j, k), upd_VP(b, i, j, k)};
        const float *N[3] = {n0, n1, n2}, *O[3] = {o0, o1, o2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!in[c]) continue;
            const double v = N[c][o], d = fabs(v - (double)O[c][o]);
            bad |= !isfinite(v) || !isfinite(d);
            me[c] = fmax(me[c], d);
            mv[c] = fmax(mv[c], fabs(v));
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        for (int m = 32; m >= 1; m >>= 1) {
            me[c] = fmax(me[c], __shfl_xor(me[c], m, 64));
            mv[c] = fmax(mv[c], __shfl_xor(mv[c], m, 64));
        }
    for (int m = 32; m >= 1; m >>= 1) bad |= __shfl_xor(bad, m, 64);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) {
        if (bad) atomicMax(&red[0], 0x7ff8000000000000ull);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (me[c] > 0.0) atomicMax(&red[c], (unsigned long long)__double_as_longlong(me[c]));
            if (mv[c] > 0.0) atomicMax(&red[3 + c], (unsigned long long)__double_as_longlong(mv[c]));
        }
    }
}

} // namespace sspt

// ======================================================================== API
using namespace sspt;

namespace {

constexpr int OBJ_NBLK = 1024;

int afail(ss_ctx *c, int code, const std::string &msg)
{
    c->err = msg;
    return code;
}

// Waits for the context's stream and returns the status of the work queued
// on it (the entry points are blocking, as the reference's @parallel calls).
int sync_status(ss_ctx *c, const char *where)
{
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        c->err = std::string(where) + ": " + hipGetErrorString(e);
        return (int)e;
    }
    return SS_OK;
}

int ensure_adj(ss_ctx *c)
{
    if (c->adj_ready) return SS_OK;
    for (int q = 0; q < AD_NBOX; ++q) {
        if (hipMalloc(&c->adj[q], c->box_elems * sizeof(double)) != hipSuccess)
            return afail(c, SS_ERR_NOMEM, "adjoint boxes");
        hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[q], c->box_elems);
    }
    if (hipMalloc(&c->adj_part, (2 * OBJ_NBLK + 2) * sizeof(double)) != hipSuccess)
        return afail(c, SS_ERR_NOMEM, "adjoint partial sums");
    c->adj_cur = 0;
    c->adj_ready = true;
    return SS_OK;
}

AdjSet set_of(ss_ctx *c, int which)
{
    AdjSet s;
    for (int q = 0; q < L_N; ++q) s.f[q] = c->adj[(which ? AD_SETB : AD_SETA) + q];
    return s;
}

AdjArgs make_adj_args(ss_ctx *c, const ss_params &p)
{
    static const double PI = 3.14159265358979323846;
    AdjArgs a{};
    a.b = c->box;
    a.t = c->tab;
    const FieldSet &s = c->cur;
    a.tRR = s.f[SS_TAU_RR]; a.tTT = s.f[SS_TAU_THTH]; a.tPP = s.f[SS_TAU_PHPH];
    a.tRT = s.f[SS_TAU_RTH]; a.tRP = s.f[SS_TAU_RPH]; a.tTP = s.f[SS_TAU_THPH];
    a.eta = s.f[SS_ETA]; a.ini = s.f[SS_ETA_INI]; a.ePL = s.f[SS_ETA_PL]; a.tII = s.f[SS_TII];
    a.DT = s.f[SS_DT_R]; a.SM = s.f[SS_ETASM]; a.VR = s.f[SS_VR];
    double **A = c->adj;
    a.sID = A[AD_SID]; a.sGS = A[AD_SGS]; a.sKB = A[AD_SKB]; a.sAE = A[AD_SAE];
    a.sX = A[AD_SX]; a.sY = A[AD_SY]; a.sZ = A[AD_SZ]; a.sIT = A[AD_SIT];
    a.gRT = A[AD_GRT]; a.idRT = A[AD_IDRT]; a.cRT = A[AD_CRT];
    a.gRP = A[AD_GRP]; a.idRP = A[AD_IDRP]; a.cRP = A[AD_CRP];
    a.gTP = A[AD_GTP]; a.idTP = A[AD_IDTP]; a.cTP = A[AD_CTP];
    a.LT = A[AD_LT]; a.Q = A[AD_Q]; a.bJ = nullptr; a.gRho = A[AD_GRHO]; a.gIni = A[AD_GINI];
    a._dr = p._dr; a._dth = p._dth; a._dph = p._dph; a.g = p.g;
    a.kb = (((p.dmp * 2.0) * PI) * p.vpdt) / p.lr;
    a.gs = ((4.0 * PI) * p.vpdt) / p.lr;
    a.tC = p.tau_C; a.ex = 1.0 - p.n_exp; a.relax = p.relax; a.emin = p.eta_min; a.emax = p.eta_max;
    a.do_relax = p.relax != 1.0;
    a.do_lo = p.eta_min > 0.0;
    a.do_hi = p.eta_max < INFINITY;
    a.bc_mode = p.bc_mode;
    a.in = set_of(c, c->adj_cur);
    a.out = set_of(c, 1 - c->adj_cur);
    return a;
}

// Arguments of the r-marching kernel for the sets of a (the same step)
adjf::AArgs<double> fused_args(ss_ctx *c, const AdjArgs &a, bool with_b)
{
    adjf::AArgs<double> f{};
    const Box &b = c->box;
    f.b = adjf::BoxG{b.nr, b.nt, b.np, b.sj, b.si};
    f.t = fused::TabsG<double>{c->ftr, c->ftt, c->ft2, c->tab.th, c->tab.ph};
    for (int q = 0; q < L_N; ++q) {
        f.li[q] = a.in.f[q];
        f.lo[q] = a.out.f[q];
    }
    f.tRR = a.tRR; f.tTT = a.tTT; f.tPP = a.tPP; f.tRT = a.tRT; f.tRP = a.tRP; f.tTP = a.tTP;
    f.eta = a.eta; f.ini = a.ini; f.ePL = a.ePL; f.tII = a.tII; f.DT = a.DT; f.SM = a.SM;
    f.LB = c->adj[AD_LB];
    f.bJ = with_b ? c->adj[AD_BJ] : nullptr;
    f.obj_r = c->adj_obj_r;
    f.obj_rad = c->adj_obj_rad;
    f._dr = a._dr; f._dth = a._dth; f._dph = a._dph; f.kb = a.kb; f.gs = a.gs; f.tC = a.tC; f.ex = a.ex;
    f.relax = a.relax; f.emin = a.emin; f.emax = a.emax;
    f.do_relax = a.do_relax; f.do_lo = a.do_lo; f.do_hi = a.do_hi;
    return f;
}

// r-chunk length of the marching kernel (march_chunk; 511^3: chunks of 74
// planes, 12.9 ms, against 13.1-13.6 ms for 26, 37, 47, 57 or 103);
// SSPT_ADJ_RCHUNK overrides it for measurements.
int adj_rchunk(const ss_ctx *c, int nplanes, int tiles, int per_cu)
{
    if (const char *e = std::getenv("SSPT_ADJ_RCHUNK")) {
        const int v = std::atoi(e);
        if (v >= 1) return std::min(v, nplanes);
    }
    return march_chunk(nplanes, tiles, c->n_cu * per_cu, false);
}

// The BC fold of lam_V in place (transposes of BC_ph!, BC_th!, BC_R!)
template <typename R> void launch_adj_fold(ss_ctx *c, int bc_mode, R *VR, R *VT, R *VP)
{
    const Box &b = c->box;
    const int m = std::max(std::max(b.nr, b.nt), b.np) + 1;
    const dim3 B2(64, 4), G2((m + 63) / 64, (m + 3) / 4);
    hipLaunchKernelGGL(k_adj_bc_ph<R>, G2, B2, 0, c->stream, b, bc_mode, VR, VT, VP);
    hipLaunchKernelGGL(k_adj_bc_th<R>, G2, B2, 0, c->stream, b, bc_mode, VR, VT, VP);
    hipLaunchKernelGGL(k_adj_bc_R<R>, G2, B2, 0, c->stream, b, bc_mode, VR, VT, VP);
}

// The boundary power-law sums (into LB) and the r-marching kernel for the
// sets of f (lam_V of f.li BC-folded beforehand)
template <typename R, bool RHS> void launch_adj_march(ss_ctx *c, adjf::AArgs<R> f, R *LB)
{
    const Box &b = c->box;
    const int m = std::max(std::max(b.nr, b.nt), b.np) + 1;
    f.LB = LB;
    hipLaunchKernelGGL(k_adj_lb<R>, dim3((m + 63) / 64, (m + 3) / 4, 3), dim3(64, 4), 0, c->stream, f, LB);
    const int nx = (b.np + 1 + adjf::AOK - 1) / adjf::AOK, ny = (b.nt + 1 + adjf::AOJ - 1) / adjf::AOJ;
    f.rchunk = adj_rchunk(c, b.nr + 1, nx * ny, sizeof(R) == 4 ? 2 : 1); // workgroups per CU
    const int nz = (b.nr + 1 + f.rchunk - 1) / f.rchunk;
    const int T = nx * ny * nz;
    hipLaunchKernelGGL((k_adj_fused<R, RHS>), dim3(((T + 7) / 8) * 8), dim3(fused::TK, adjf::ATJ), 0, c->stream, f, nx,
                       ny, nz);
}

// lam_out = T^T lam_in (+ bJ on VR if given); gradients when a.grad
void launch_adj_step(ss_ctx *c, AdjArgs a, bool with_b)
{
    const Box &b = c->box;
    launch_adj_fold(c, a.bc_mode, a.in.f[L_VR], a.in.f[L_VT], a.in.f[L_VP]);
    if (c->adj_fused) {
        launch_adj_march<double, false>(c, fused_args(c, a, with_b), c->adj[AD_LB]);
        if (a.grad) {
            hipLaunchKernelGGL(k_adj_power_law, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
            hipLaunchKernelGGL(k_adj_grad_rho, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
        }
        return;
    }
    hipLaunchKernelGGL(k_adj_power_law, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    if (a.grad) hipLaunchKernelGGL(k_adj_grad_rho, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    hipLaunchKernelGGL(k_adj_tii, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    hipLaunchKernelGGL(k_adj_sig, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    a.bJ = with_b ? c->adj[AD_BJ] : nullptr;
    hipLaunchKernelGGL(k_adj_vel, agrid(b.nr + 1, b.nt + 1, b.np + 1), AB, 0, c->stream, a);
}

int adj_field(int field, int *fwd, int *box)
{
    static const int map[SS_ADJ_NFIELDS] = {SS_VR, SS_VTH, SS_VPH, SS_P, SS_TAU_RR, SS_TAU_THTH, SS_TAU_PHPH,
                                            SS_TAU_RTH, SS_TAU_RPH, SS_TAU_THPH, SS_ETA, SS_ETA_PL, SS_RHO,
                                            SS_ETA_INI};
    if (field < 0 || field >= SS_ADJ_NFIELDS) return SS_ERR_ARG;
    *fwd = map[field];
    *box = field < L_N ? -1 : (field == SS_ADJ_GRAD_RHO ? AD_GRHO : AD_GINI);
    return SS_OK;
}

// Gradients from the converged lam (set adj_cur): one more backward pass on
// a BC-folded copy in the other set (its lam outputs are not needed)
int adj_gradients(ss_ctx *c, const ss_params &p)
{
    const Box &b = c->box;
    // gradients from the converged lam: one more backward pass (its outputs are discarded)
    AdjArgs a = make_adj_args(c, p);
    a.grad = 1;
    hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[AD_GRHO], c->box_elems);
    hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[AD_GINI], c->box_elems);
    // keep lam: fold and differentiate a copy in the other set
    AdjSet cur = set_of(c, c->adj_cur), oth = set_of(c, 1 - c->adj_cur);
    for (int q = 0; q < L_N; ++q)
        if (hipMemcpyAsync(oth.f[q], cur.f[q], c->box_elems * sizeof(double), hipMemcpyDeviceToDevice, c->stream) !=
            hipSuccess)
            return afail(c, SS_ERR_STATE, "adjoint gradient copy");
    a.in = oth;
    a.out = oth;
    launch_adj_fold(c, a.bc_mode, a.in.f[L_VR], a.in.f[L_VT], a.in.f[L_VP]);
    hipLaunchKernelGGL(k_adj_power_law, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    hipLaunchKernelGGL(k_adj_grad_rho, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    return sync_status(c, "adjoint gradients");
}

// FP32 boxes of the mixed-precision adjoint solve
enum { AM_D = 0, AM_RHS = 2 * L_N, AM_ST = 3 * L_N, AM_LB = 3 * L_N + 12, AM_NBOX = 3 * L_N + 13 };
static_assert(AM_NBOX == SS_ADJ32_NBOX, "FP32 adjoint box count");

int ensure_adj32(ss_ctx *c)
{
    for (int q = 0; q < AM_NBOX; ++q) {
        if (c->adj32[q]) continue;
        if (hipMalloc(&c->adj32[q], c->box_elems * sizeof(float)) != hipSuccess)
            return afail(c, SS_ERR_NOMEM, "FP32 adjoint boxes");
        hipLaunchKernelGGL(k_adj_zero_f32, dim3(2048), dim3(256), 0, c->stream, c->adj32[q], c->box_elems);
    }
    return SS_OK;
}

// Arguments of the FP32 correction sweeps: the FP32 state copies, the FP32
// tables, right-hand side set AM_RHS (lam sets filled in per sweep)
adjf::AArgs<float> f32_args(ss_ctx *c, const AdjArgs &a)
{
    adjf::AArgs<float> f{};
    const Box &b = c->box;
    float *const *F = c->adj32;
    f.b = adjf::BoxG{b.nr, b.nt, b.np, b.sj, b.si};
    f.t = fused::TabsG<float>{c->ftr32, c->ftt32, c->ft232, c->th32, c->ph32};
    for (int q = 0; q < L_N; ++q) f.rhs[q] = F[AM_RHS + q];
    const float *const *S = F + AM_ST;
    f.tRR = S[0]; f.tTT = S[1]; f.tPP = S[2]; f.tRT = S[3]; f.tRP = S[4]; f.tTP = S[5];
    f.eta = S[6]; f.ini = S[7]; f.ePL = S[8]; f.tII = S[9]; f.DT = S[10]; f.SM = S[11];
    f.bJ = nullptr;
    f._dr = (float)a._dr; f._dth = (float)a._dth; f._dph = (float)a._dph;
    f.kb = (float)a.kb; f.gs = (float)a.gs; f.tC = (float)a.tC; f.ex = (float)a.ex;
    f.relax = (float)a.relax; f.emin = (float)a.emin; f.emax = (float)a.emax;
    f.do_relax = a.do_relax; f.do_lo = a.do_lo; f.do_hi = a.do_hi;
    return f;
}

// max over the velocity components of red[c] / red[3 + c] (bit patterns of
// non-negative doubles); -1 when a value is not finite
double ratio_of(const unsigned long long *red)
{
    double err = 0.0;
    for (int q = 0; q < 3; ++q) {
        double dv, vv;
        memcpy(&dv, &red[q], 8);
        memcpy(&vv, &red[3 + q], 8);
        if (!std::isfinite(dv) || !std::isfinite(vv)) return -1.0;
        if (vv > 0.0) err = std::fmax(err, dv / vv);
    }
    return err;
}

using aclk = std::chrono::steady_clock;

} // namespace

extern "C" {

int ss_adjoint_prepare(ss_ctx *c, const ss_params *p)
{
    if (!c || !p) return SS_ERR_ARG;
    if (!c->grid_set) return afail(c, SS_ERR_STATE, "ss_set_grid was not called");
    if (hipSetDevice(c->device) != hipSuccess) return afail(c, SS_ERR_STATE, "hipSetDevice");
    int rc = ensure_adj(c);
    if (rc != SS_OK) return rc;
    // the r-marching kernel from about 255^3 cells up; below, the per-family
    // kernels (one thread per point) keep more of the GPU busy (measured:
    // 0.50 vs 0.37 ms at 127^3, 2.40 vs 2.48 ms at 255^3, 14.0 vs 19.6 ms at
    // 511^3)
    c->adj_fused = (double)c->box.nr * c->box.nt * c->box.np >= 1.6e7;
    if (const char *e = std::getenv("SSPT_ADJ_FUSED")) c->adj_fused = std::atoi(e) != 0;
    AdjArgs a = make_adj_args(c, *p);
    const Box &b = c->box;
    hipLaunchKernelGGL(k_adj_prep_centre, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a);
    for (int fam = 0; fam < 3; ++fam)
        hipLaunchKernelGGL(k_adj_prep_edge, agrid(b.nr, b.nt, b.np), AB, 0, c->stream, a, fam);
    return sync_status(c, "ss_adjoint_prepare");
}

int ss_adjoint_zero(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    if (!c->adj_ready) return afail(c, SS_ERR_STATE, "ss_adjoint_prepare was not called");
    for (int q = 0; q < AD_PREV; ++q)
        hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[q], c->box_elems);
    hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[AD_GRHO], c->box_elems);
    hipLaunchKernelGGL(k_adj_zero, dim3(2048), dim3(256), 0, c->stream, c->adj[AD_GINI], c->box_elems);
    c->adj_cur = 0;
    return sync_status(c, "ss_adjoint_zero");
}

int ss_adjoint_objective(ss_ctx *c, double r, double radius, double *J)
{
    if (!c || !J) return SS_ERR_ARG;
    if (!c->adj_ready) return afail(c, SS_ERR_STATE, "ss_adjoint_prepare was not called");
    if (hipSetDevice(c->device) != hipSuccess) return afail(c, SS_ERR_STATE, "hipSetDevice");
    ss_params dummy{};
    dummy.lr = 1.0;
    dummy.relax = 1.0;
    dummy.eta_max = INFINITY;
    AdjArgs a = make_adj_args(c, dummy);
    a.bJ = c->adj[AD_BJ];
    c->adj_obj_r = r;
    c->adj_obj_rad = radius;
    hipLaunchKernelGGL(k_adj_obj_partial, dim3(OBJ_NBLK), dim3(OBJ_BLOCK), 0, c->stream, a, r, radius, c->adj_part);
    hipLaunchKernelGGL(k_adj_obj_final, dim3(1), dim3(OBJ_BLOCK), 0, c->stream, c->adj_part, OBJ_NBLK,
                       c->adj_part + 2 * OBJ_NBLK);
    double sums[2] = {0.0, 0.0};
    if (hipMemcpyAsync(sums, c->adj_part + 2 * OBJ_NBLK, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
        hipSuccess)
        return afail(c, SS_ERR_STATE, "objective readback");
    int rc = sync_status(c, "ss_adjoint_objective");
    if (rc != SS_OK) return rc;
    if (!(sums[0] > 0.0)) return afail(c, SS_ERR_ARG, "ss_adjoint_objective: no r-face inside the inclusion");
    const Box &b = c->box;
    hipLaunchKernelGGL(k_adj_scale, agrid(b.nr + 1, b.nt, b.np), AB, 0, c->stream, b, c->adj[AD_BJ], b.nr + 1, b.nt,
                       b.np, 1.0 / sums[0]);
    *J = sums[1] / sums[0];
    return sync_status(c, "ss_adjoint_objective");
}

int ss_adjoint_upload(ss_ctx *c, int field, const double *host, int layout)
{
    if (!c || !host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return SS_ERR_ARG;
    if (!c->adj_ready) return afail(c, SS_ERR_STATE, "ss_adjoint_prepare was not called");
    int fwd, box;
    if (adj_field(field, &fwd, &box) != SS_OK) return afail(c, SS_ERR_ARG, "ss_adjoint_upload: bad field");
    int d[3];
    ss_field_dims(c, fwd, d);
    const int o1 = (fwd == SS_TAU_RTH || fwd == SS_TAU_RPH || fwd == SS_TAU_THPH) ? 1 : 0;
    double *dst = box >= 0 ? c->adj[box] : set_of(c, c->adj_cur).f[field];
    const size_t n = (size_t)d[0] * d[1] * d[2];
    if (hipMemcpyAsync(c->staging, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return afail(c, SS_ERR_STATE, "ss_adjoint_upload copy");
    launch_to_box(c, dst, c->staging, d[0], d[1], d[2], o1, layout);
    return sync_status(c, "ss_adjoint_upload");
}

int ss_adjoint_download(ss_ctx *c, int field, double *host, int layout)
{
    if (!c || !host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return SS_ERR_ARG;
    if (!c->adj_ready) return afail(c, SS_ERR_STATE, "ss_adjoint_prepare was not called");
    int fwd, box;
    if (adj_field(field, &fwd, &box) != SS_OK) return afail(c, SS_ERR_ARG, "ss_adjoint_download: bad field");
    int d[3];
    ss_field_dims(c, fwd, d);
    const int o1 = (fwd == SS_TAU_RTH || fwd == SS_TAU_RPH || fwd == SS_TAU_THPH) ? 1 : 0;
    const double *src = box >= 0 ? c->adj[box] : set_of(c, c->adj_cur).f[field];
    const size_t n = (size_t)d[0] * d[1] * d[2];
    launch_from_box(c, c->staging, src, d[0], d[1], d[2], o1, layout);
    if (hipMemcpyAsync(host, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return afail(c, SS_ERR_STATE, "ss_adjoint_download copy");
    return sync_status(c, "ss_adjoint_download");
}

int ss_adjoint_step(ss_ctx *c, const ss_params *p, int gradients)
{
    if (!c || !p) return SS_ERR_ARG;
    if (!c->adj_ready) return afail(c, SS_ERR_STATE, "ss_adjoint_prepare was not called");
    if (hipSetDevice(c->device) != hipSuccess) return afail(c, SS_ERR_STATE, "hipSetDevice");
    AdjArgs a = make_adj_args(c, *p);
    a.grad = gradients ? 1 : 0;
    launch_adj_step(c, a, false);
    c->adj_cur = 1 - c->adj_cur;
    return sync_status(c, "ss_adjoint_step");
}

int ss_adjoint_solve(ss_ctx *c, const ss_params *p, double r, double radius, int maxiter, int ncheck, double tol,
                     ss_adjoint_result *out)
{
    if (!c || !p || !out || maxiter < 1 || ncheck < 1) return SS_ERR_ARG;
    memset(out, 0, sizeof *out);
    out->err = NAN;
    const aclk::time_point t0 = aclk::now();
    int rc = ss_adjoint_prepare(c, p);
    if (rc != SS_OK) return rc;
    if ((rc = ss_adjoint_zero(c)) != SS_OK) return rc;
    if ((rc = ss_adjoint_objective(c, r, radius, &out->J)) != SS_OK) return rc;
    int it = 0;
    while (it < maxiter) {
        const bool check = ((it + 1) % ncheck) == 0;
        AdjSet cur = set_of(c, c->adj_cur);
        if (check)
            for (int q = 0; q < 3; ++q)
                if (hipMemcpyAsync(c->adj[AD_PREV + q], cur.f[q], c->box_elems * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                    return afail(c, SS_ERR_STATE, "adjoint check copy");
        AdjArgs a = make_adj_args(c, *p);
        launch_adj_step(c, a, true);
        c->adj_cur = 1 - c->adj_cur;
        ++it;
        if (check) {
            FieldSet s{};
            AdjSet nw = set_of(c, c->adj_cur);
            s.f[SS_VR] = nw.f[L_VR];
            s.f[SS_VTH] = nw.f[L_VT];
            s.f[SS_VPH] = nw.f[L_VP];
            launch_check_err(c, s, c->adj[AD_PREV], c->adj[AD_PREV + 1], c->adj[AD_PREV + 2], 0);
            if (hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess)
                return afail(c, SS_ERR_STATE, "adjoint check readback");
            if ((rc = sync_status(c, "adjoint iteration")) != SS_OK) return rc;
            double err = 0.0;
            for (int q = 0; q < 3; ++q) {
                double dv, vv;
                memcpy(&dv, &c->red_host[q], 8);
                memcpy(&vv, &c->red_host[3 + q], 8);
                if (std::isnan(dv) || std::isnan(vv)) { out->iters = it; return afail(c, SS_ERR_NONFINITE, "adjoint NaN"); }
                if (vv > 0.0) err = std::fmax(err, dv / vv);
            }
            out->err = err;
            if (err < tol) { out->converged = 1; break; }
        }
    }
    out->iters = it;
    if ((rc = adj_gradients(c, *p)) != SS_OK) return rc;
    out->seconds = std::chrono::duration<double>(aclk::now() - t0).count();
    return out->converged ? SS_OK : SS_NOT_CONVERGED;
}

int ss_adjoint_solve_mixed(ss_ctx *c, const ss_params *p, double r, double radius, const ss_mixed_opts *o,
                           ss_adjoint_mixed_result *out)
{
    if (!c || !p || !o || !out) return SS_ERR_ARG;
    if (o->max_outer < 1 || o->max_outer > SS_MIXED_MAX_OUTER || o->max_inner < 1 || o->nout_inner < 1)
        return afail(c, SS_ERR_ARG, "bad mixed-solve options");
    memset(out, 0, sizeof *out);
    out->err = NAN;
    const aclk::time_point t0 = aclk::now();
    int rc = ss_adjoint_prepare(c, p);
    if (rc != SS_OK) return rc;
    if ((rc = ss_adjoint_zero(c)) != SS_OK) return rc;
    if ((rc = ss_adjoint_objective(c, r, radius, &out->J)) != SS_OK) return rc;
    if ((rc = ensure_adj32(c)) != SS_OK) return rc;
    const size_t n = c->box_elems;
    const dim3 FG(2048), FB(256);
    float *const *F = c->adj32;
    AdjArgs a0 = make_adj_args(c, *p);
    {   // the converged state in FP32, once
        const double *st[12] = {a0.tRR, a0.tTT, a0.tPP, a0.tRT, a0.tRP, a0.tTP,
                                a0.eta, a0.ini, a0.ePL, a0.tII, a0.DT, a0.SM};
        for (int q = 0; q < 12; ++q) hipLaunchKernelGGL(k_adj_to_f32, FG, FB, 0, c->stream, F[AM_ST + q], st[q], n);
    }
    adjf::AArgs<float> f = f32_args(c, a0);
    const Box &b = c->box;
    const dim3 CB(64, 4), CG((b.np + 1 + 63) / 64, (b.nt + 1 + 3) / 4, (b.nr + 1 + ADJ_CHECK_PLANES - 1) / ADJ_CHECK_PLANES);
    double tin_prev = o->tol_inner, err_prev = NAN;
    for (int outer = 0; outer < o->max_outer; ++outer) {
        // Y = F(T^T mu + dJ/dx) into the other set; err = max|Y - mu| / max|Y|
        AdjArgs a = make_adj_args(c, *p);
        launch_adj_step(c, a, true);
        launch_adj_fold(c, p->bc_mode, a.out.f[L_VR], a.out.f[L_VT], a.out.f[L_VP]);
        FieldSet y{};
        y.f[SS_VR] = a.out.f[L_VR];
        y.f[SS_VTH] = a.out.f[L_VT];
        y.f[SS_VPH] = a.out.f[L_VP];
        launch_check_err(c, y, a.in.f[L_VR], a.in.f[L_VT], a.in.f[L_VP], 0);
        if (hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess)
            return afail(c, SS_ERR_STATE, "adjoint check readback");
        if ((rc = sync_status(c, "adjoint outer step")) != SS_OK) return rc;
        const double err = ratio_of(c->red_host);
        if (err < 0.0) return afail(c, SS_ERR_NONFINITE, "mixed adjoint: non-finite adjoint value");
        out->err = err;
        out->outer = outer;
        if (err < o->tol) { // Y is the iterate
            c->adj_cur = 1 - c->adj_cur;
            out->converged = 1;
            break;
        }
        // The reduction of an outer step scales with the inner tolerance it
        // used; when the reduction still needed is smaller than the last
        // step's, the inner tolerance is loosened in proportion (factor 2 of
        // margin, at most 1e-2), never tightened beyond o->tol_inner.
        double tin = o->tol_inner;
        if (std::isfinite(err_prev) && err_prev > 0.0 && err < err_prev) {
            const double q = err / err_prev, want = 0.5 * (o->tol / err);
            if (want > q) tin = std::fmin(1e-2, tin_prev * (want / q));
            tin = std::fmax(tin, o->tol_inner);
        }
        err_prev = err;
        tin_prev = tin;
        // right-hand side Y - mu in FP32; both correction sets zero
        for (int q = 0; q < L_N; ++q)
            hipLaunchKernelGGL(k_adj_resid_f32, FG, FB, 0, c->stream, F[AM_RHS + q], a.out.f[q], a.in.f[q], n);
        for (int q = 0; q < 2 * L_N; ++q) hipLaunchKernelGGL(k_adj_zero_f32, FG, FB, 0, c->stream, F[AM_D + q], n);
        if ((rc = sync_status(c, "adjoint correction set-up")) != SS_OK) return rc;
        const aclk::time_point ti0 = aclk::now();
        int cur = 0, it;
        for (it = 1; it <= o->max_inner; ++it) {
            float *const *in = F + AM_D + L_N * cur, *const *nw = F + AM_D + L_N * (1 - cur);
            for (int q = 0; q < L_N; ++q) {
                f.li[q] = in[q];
                f.lo[q] = nw[q];
            }
            launch_adj_fold(c, p->bc_mode, in[L_VR], in[L_VT], in[L_VP]);
            launch_adj_march<float, true>(c, f, F[AM_LB]);
            cur = 1 - cur;
            if (it % o->nout_inner == 0) {
                // compare folded with folded: the new set's own fold at the
                // next sweep then changes nothing (the fold is idempotent)
                launch_adj_fold(c, p->bc_mode, nw[L_VR], nw[L_VT], nw[L_VP]);
                hipLaunchKernelGGL(k_adj_zero_u64, dim3(1), dim3(64), 0, c->stream, c->red, 6);
                hipLaunchKernelGGL(k_adj_inc_f32, CG, CB, 0, c->stream, b, nw[L_VR], nw[L_VT], nw[L_VP], in[L_VR],
                                   in[L_VT], in[L_VP], c->red);
                if (hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   c->stream) != hipSuccess)
                    return afail(c, SS_ERR_STATE, "adjoint inner readback");
                if ((rc = sync_status(c, "adjoint FP32 sweeps")) != SS_OK) return rc;
                const double ein = ratio_of(c->red_host);
                if (ein < 0.0) return afail(c, SS_ERR_NONFINITE, "mixed adjoint: non-finite value in the FP32 sweeps");
                if (ein < tin) break;
            }
        }
        if ((rc = sync_status(c, "adjoint FP32 sweeps")) != SS_OK) return rc;
        out->inner_seconds += std::chrono::duration<double>(aclk::now() - ti0).count();
        out->inner_total += std::min(it, o->max_inner);
        // mu += F d (FP64)
        float *const *d = F + AM_D + L_N * cur;
        launch_adj_fold(c, p->bc_mode, d[L_VR], d[L_VT], d[L_VP]);
        for (int q = 0; q < L_N; ++q) hipLaunchKernelGGL(k_adj_acc_f32, FG, FB, 0, c->stream, a.in.f[q], d[q], n);
        out->outer = outer + 1;
    }
    if ((rc = adj_gradients(c, *p)) != SS_OK) return rc;
    out->seconds = std::chrono::duration<double>(aclk::now() - t0).count();
    return out->converged ? SS_OK : SS_NOT_CONVERGED;
}

} // extern "C"
