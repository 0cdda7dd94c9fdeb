// sspt_adj_fused_body.h -- one application of the transposed PT map
// (lam_out = T^T lam_in at the converged state) as an r-marching block
// schedule, the adjoint counterpart of sspt_fused_body.h.
//
// The forward iteration (SphericalStokes_diapir.jl:413-425) is, per point,
// compute_P! (new P, tau from old V, eta, tau), compute_TII! + copy_BC!,
// power_law! (new eta from tII) and compute_V! (new V from new P, tau).  Its
// transpose, in gather form (each output sums the coefficients of the forward
// statements that read it; no atomics), needs at plane q:
//   * w_f = DT_f lam_V_f on the faces compute_V! updates (r-faces q, q+1;
//     th/phi faces q-1, q);
//   * Q = lam_tII / tII of interior cells (lam_tII from the power-law
//     adjoint of the cell and of the boundary cells copy_BC! filled from it);
//   * the adjoints of the new P, tau of centre q and of the edges stored at
//     box index q ("sig", written scaled by compute_P!'s 1/den, which makes
//     them the adjoints of the old P, tau);
//   * lam_V' of faces and lam_eta' of centres from the sig values of the
//     centres and edges around them ("vel").
// Step q of a block computes the w's and Q of plane q (phase A), sig of
// plane q (phase B) and the vel contributions of plane q's sig values
// (phase C): VR(q) is complete then, while Vth, Vph and eta of plane q
// still wait for the edges of box index q+1 and are finished at step q+1.
// Values the same column needs at the next plane are carried in registers;
// values neighbours need go through LDS (two barriers per step), and so do
// each thread's accumulators and the previous plane's values that the edges
// of the next box index need (to keep the register count down).
//
// Block = 64 lanes (phi) x ATJ rows (th); lane tx covers k = kb-1+tx, row ty
// covers j = jb-1+ty; the inner 62 x (ATJ-2) positions are owned.  Outputs
// cover the face extents: j in [0, nth], k in [0, nph], r-planes [0, nr].
//
// The BC fold (the transposes of BC_R!/BC_th!/BC_ph!) has been applied to
// lam_in beforehand, and LB holds, for each interior cell next to the
// boundary, the power-law adjoints of the boundary cells copy_BC! filled
// from it.  The coefficients that depend only on the converged state are
// recomputed from it (GS, KB, 1/den, 1/tII, the 4-point averages of the
// viscosities and shear stresses).  The shear stresses of the plane ahead
// are loaded one step early and carried.
#pragma once
#include "sspt_fused_body.h"

namespace sspt {
namespace adjf {

using fused::BoxG;
using fused::TabsG;
using fused::TK;
using fused::clampi;
using fused::fat;
using fused::tload;
using fused::NTR;
using fused::NTT;
using fused::NT2;

#ifndef SSPT_ATJB
#define SSPT_ATJB 12
#endif
constexpr int ATJ = SSPT_ATJB;    // rows per block
constexpr int AOK = TK - 2;       // owned lanes per row
constexpr int AOJ = ATJ - 2;      // owned rows

// adjoint state fields, in the order of the adjoint sets
enum { A_VR, A_VT, A_VP, A_P, A_RR, A_TT, A_PP, A_RT, A_RP, A_TP, A_E, A_PL, A_N };

template <typename R> struct AArgs {
    BoxG b;
    TabsG<R> t;
    const R *li[A_N]; // lam in (BC-folded)
    R *lo[A_N];       // lam out
    const R *rhs[A_N]; // added to every output in the RHS instances (else unused)
    // converged forward state
    const R *tRR, *tTT, *tPP, *tRT, *tRP, *tTP, *eta, *ini, *ePL, *tII, *DT, *SM;
    const R *LB;
    // objective weights on the r-faces inside the inclusion (or null)
    const R *bJ;
    double obj_r, obj_rad;
    R _dr, _dth, _dph, kb, gs, tC, ex, relax, emin, emax;
    int do_relax, do_lo, do_hi;
    int rchunk;
};

// Output f at element o: v, plus the right-hand side there in the RHS
// instances (the correction sweeps delta = T^T delta + r of the
// mixed-precision solve).
template <bool RHS, typename R> SSPT_DH inline void aput(const AArgs<R> &a, int f, uint32_t o, R v)
{
    if constexpr (RHS) v += fat(a.rhs[f], o);
    fused::st_stream(fat(a.lo[f], o), v);
}

// ---------------------------------------------------------------- ranges
// faces compute_V! updates and the edge boxes, in box indices
SSPT_DH inline bool upd_VR(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 2;
}
SSPT_DH inline bool upd_VT(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 2;
}
SSPT_DH inline bool upd_VP(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 1;
}
SSPT_DH inline bool in_RT(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 2;
}
SSPT_DH inline bool in_RP(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 1 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 1;
}
SSPT_DH inline bool in_TP(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 1 && k >= 1 && k <= b.np - 1;
}
SSPT_DH inline bool inner_c(const BoxG &b, int i, int j, int k)
{
    return i >= 1 && i <= b.nr - 2 && j >= 1 && j <= b.nt - 2 && k >= 1 && k <= b.np - 2;
}
SSPT_DH inline bool is_cell(const BoxG &b, int i, int j, int k)
{
    return i >= 0 && i < b.nr && j >= 0 && j < b.nt && k >= 0 && k < b.np;
}
// an interior cell that copy_BC! copies onto a boundary cell
SSPT_DH inline bool next_to_boundary(const BoxG &b, int i, int j, int k)
{
    return i == 1 || i == b.nr - 2 || j == 1 || j == b.nt - 2 || k == 1 || k == b.np - 2;
}

// power_law!'s adjoint at a cell from lam_eta (le), tII, eta_ini and, with
// relaxation, the old eta_PL and lam_eta_PL: returns lam_tII; *pl_out = the
// adjoint of the old eta_PL (relaxation only).
template <typename R> SSPT_DH inline R adj_power_law_v(const AArgs<R> &a, R le, R tii, R ini, R ePL, R lPL,
                                      R *pl_out)
{
    const R pl0 = ini * std::pow(tii / a.tC, a.ex);
    const R pl = a.do_relax ? ePL : pl0;
    const R e_raw = R(2.0) / (R(1.0) / ini + R(1.0) / pl);
    const bool clamped = (a.do_lo && e_raw < a.emin) || (a.do_hi && e_raw > a.emax);
    const R a_e = clamped ? R(0.0) : le;
    *pl_out = R(0.0);
    if (!std::isfinite(pl) || !std::isfinite(pl0)) return R(0.0); // tII = 0: eta = 2 eta_ini
    R l_pl = a_e * (e_raw * e_raw) / (R(2.0) * pl * pl);
    if (a.do_relax) {
        l_pl += lPL;
        const R l_pl0 = l_pl * a.relax * pl / pl0;
        *pl_out = l_pl * (R(1.0) - a.relax) * pl / ePL;
        l_pl = l_pl0;
    }
    return (tii > R(0.0)) ? l_pl * a.ex * pl0 / tii : R(0.0);
}
template <typename R> SSPT_DH inline R adj_power_law(const AArgs<R> &a, uint32_t o, R *pl_out)
{
    return adj_power_law_v(a, fat(a.li[A_E], o), fat(a.tII, o), fat(a.ini, o), a.do_relax ? fat(a.ePL, o) : R(0.0),
                           a.do_relax ? fat(a.li[A_PL], o) : R(0.0), pl_out);
}

// objective weight test of the r-face (i, j, k): inside the inclusion
template <typename R> SSPT_DH inline bool in_objective(const AArgs<R> &a, int i, int j, int k)
{
    if (i < 1 || i > a.b.nr - 1) return false;
    const double Rf = tload(a.t.tr, i * NTR + fused::TR_Rf);
    const double x = Rf - a.obj_r, y = (a.t.th[j] - fused::PI_ / 2.0) * a.obj_r, z = a.t.ph[k] * a.obj_r;
    return ((x * x + y * y) + z * z) < a.obj_rad;
}

template <typename R> struct ALds {
    // phase A -> B; QX, QY, GSM, ETA for planes q and q-1 (index = plane & 1)
    R wVR[ATJ][TK], wVT[ATJ][TK], wVP[ATJ][TK], QZ[ATJ][TK];
    R QX[2][ATJ][TK], QY[2][ATJ][TK], GSM[2][ATJ][TK], ETA[2][ATJ][TK];
    // DT_R of planes q and q+1 (index = plane & 1), written one step ahead:
    // the j-1 / k-1 neighbours' values for the face weights of Vth / Vph
    R DTL[2][ATJ][TK];
    // phase B -> C (values read only from the lane next door, gRP, cRP and
    // the centre term of Vph, come by lane shuffles instead)
    R tV[ATJ][TK]; // centre term of Vth at j+1
    R gRT[ATJ][TK], gTP[ATJ][TK], cRT[ATJ][TK], cTP[ATJ][TK];
    // each thread's own accumulators (Vth, Vph, eta of plane q-1; VR(q)'s
    // term of centre q-1) and w of its th/phi faces of plane q-1: kept here
    // rather than in registers
    R ACC[6][ATJ][TK];
};
enum { ACC_VT, ACC_VP, ACC_E, ACC_VRN, ACC_WVT, ACC_WVP };

// Member m of the thread d = +1 / -1 lanes away in the same row when
// `in_row` (the neighbour lies in the row), else 0.  On the device a row is
// one wavefront and the value comes by a lane shuffle that every lane
// executes; the host emulation keeps a row's threads contiguous in memory
// and reads the neighbour's state.
template <typename S, typename V> SSPT_DH inline V lane_nb(const S &self, V S::*m, int d, bool in_row)
{
#ifdef __HIP_DEVICE_COMPILE__
    const V v = d > 0 ? __shfl_down(self.*m, 1, TK) : __shfl_up(self.*m, 1, TK);
    return in_row ? v : V(0);
#else
    return in_row ? (&self)[d].*m : V(0);
#endif
}

template <typename R> struct AThread {
    int tx, ty, j, k;
    int jc, kc; // clamped into [0, nth] x [0, nph] for loads
    bool own;
    // carries from plane q-1 (or q, loaded ahead)
    R DTq, wVRq, lVRq;          // DT(q), w_VR(q), lam_VR(q)
    R RTq, RTqj, RPq, RPqk;     // tau_Rth at (q, j), (q, j+1); tau_Rph at (q, k), (q, k+1)
    R RTn, RTnj, RPn, RPnk;     // the same of box index q+1 (loaded this step)
    // values of the current step
    R DTq1, lVR1, wVR1, lVT, lVP, wVT, wVP, Q;
    R GS, KB, e, tRR, tTT, tPP, tRT, tRP, tTP;
    R lP, lRR, lTT, lPP, lRT, lRP, lTP;
    R cw_dR, cw_dT, cw_dP, cw_aR, cw_aT, eC;
    R gRT, gRP, gTP, cRT, cRP, cTP;
};

template <typename R> SSPT_DH inline void athread_init(AThread<R> &T, const AArgs<R> &a, int tx, int ty, int kb, int jb)
{
    T.tx = tx;
    T.ty = SSPT_UNIFORM(ty);
    T.k = kb - 1 + tx;
    T.j = SSPT_UNIFORM(jb - 1 + ty);
    T.jc = clampi(T.j, 0, a.b.nt);
    T.kc = clampi(T.k, 0, a.b.np);
    T.own = tx >= 1 && tx <= TK - 2 && ty >= 1 && ty <= ATJ - 2 && T.j <= a.b.nt && T.k <= a.b.np;
    T.DTq = T.wVRq = T.lVRq = R(0.0);
}

// DT(q), lam_VR(q) and w_VR(q) of the first plane of the march (DT(q) also
// into the LDS plane that the first phase A reads: a barrier follows)
template <typename R> SSPT_DH inline void aprologue(AThread<R> &T, ALds<R> &L, const AArgs<R> &a, int q)
{
    for (int m = 0; m < 6; ++m) L.ACC[m][T.ty][T.tx] = R(0.0);
    const BoxG &b = a.b;
    const int qc = clampi(q, 0, b.nr);
    T.DTq = fat(a.DT, b.o32(qc, T.jc, T.kc));
    L.DTL[q & 1][T.ty][T.tx] = T.DTq;
    T.lVRq = fat(a.li[A_VR], b.o32(qc, T.jc, T.kc));
    const R DTm = fat(a.DT, b.o32(clampi(q - 1, 0, b.nr), T.jc, T.kc));
    T.wVRq = upd_VR(b, q, T.j, T.k) ? ((DTm + T.DTq) * R(0.5)) * T.lVRq : R(0.0);
    const int jc1 = clampi(T.jc + 1, 0, b.nt), kc1 = clampi(T.kc + 1, 0, b.np);
    T.RTq = fat(a.tRT, b.o32(qc, T.jc, T.kc));
    T.RTqj = fat(a.tRT, b.o32(qc, jc1, T.kc));
    T.RPq = fat(a.tRP, b.o32(qc, T.jc, T.kc));
    T.RPqk = fat(a.tRP, b.o32(qc, T.jc, kc1));
}

// Field values one step reads (plane q, r-face q+1 for VR), in two halves:
// those of phase A (loaded during phase B of the step before) and those of
// phase B (loaded at the start of phase A).  Each half's latency hides
// behind a phase and a barrier, with one half in flight at a time.
template <typename R> struct AStepA {
    R DT1, lVR1, lVT, lVP;
    R DTjm, DTkm;               // DT(q) at j-1 (first row) / k-1 (first lane); the others read LDS
    R le, tii, ini, ePL, lPL, lb, sm, e;
    R RT1, RT1j, RP1, RP1k;     // tau_Rth / tau_Rph of box index q+1 at (j, j+1) / (k, k+1)
    R TP, TPk, TPj, TPjk;       // tau_thph of box index q at (j,k), (j,k+1), (j+1,k), (j+1,k+1)
    // (the k+1 values RP1k, TPk, TPjk are loaded by the last lane only; the
    // others take them from the lane next door)
};
template <typename R> struct AStepB {
    R tRR, tTT, tPP;
    R lP, lRR, lTT, lPP, lRT, lRP, lTP;
};

// The loads of step q (every position clamped into the box, so all are
// unconditional).
template <typename R> SSPT_DH inline void aloadA(AStepA<R> &S, const AArgs<R> &a, const AThread<R> &T, int q)
{
    const BoxG &b = a.b;
    const int qc = clampi(q, 0, b.nr);
    const uint32_t o = b.o32(qc, T.jc, T.kc);
    const uint32_t o1 = b.o32(clampi(q + 1, 0, b.nr), T.jc, T.kc);
    S.DT1 = fat(a.DT, o1);
    S.lVR1 = fat(a.li[A_VR], o1);
    S.lVT = fat(a.li[A_VT], o);
    S.lVP = fat(a.li[A_VP], o);
    S.DTjm = T.ty == 0 ? fat(a.DT, b.o32(qc, clampi(T.jc - 1, 0, b.nt), T.kc)) : R(0.0);
    S.DTkm = T.tx == 0 ? fat(a.DT, b.o32(qc, T.jc, clampi(T.kc - 1, 0, b.np))) : R(0.0);
    S.le = fat(a.li[A_E], o);
    S.tii = fat(a.tII, o);
    S.ini = fat(a.ini, o);
    S.ePL = a.do_relax ? fat(a.ePL, o) : R(0.0);
    S.lPL = a.do_relax ? fat(a.li[A_PL], o) : R(0.0);
    const int jc1 = clampi(T.jc + 1, 0, b.nt), kc1 = clampi(T.kc + 1, 0, b.np);
    S.RT1 = fat(a.tRT, o1);
    S.RT1j = fat(a.tRT, b.o32(clampi(q + 1, 0, b.nr), jc1, T.kc));
    S.RP1 = fat(a.tRP, o1);
    const bool lastl = T.tx == TK - 1;
    S.RP1k = lastl ? fat(a.tRP, b.o32(clampi(q + 1, 0, b.nr), T.jc, kc1)) : R(0.0);
    S.TP = fat(a.tTP, o);
    S.TPk = lastl ? fat(a.tTP, b.o32(qc, T.jc, kc1)) : R(0.0);
    S.TPj = fat(a.tTP, b.o32(qc, jc1, T.kc));
    S.TPjk = lastl ? fat(a.tTP, b.o32(qc, jc1, kc1)) : R(0.0);
    // LB is read only next to the boundary (the condition depends on the
    // position alone, so the load still issues with the others)
    S.lb = next_to_boundary(b, q, T.j, T.k) ? fat(a.LB, o) : R(0.0);
    S.sm = fat(a.SM, o);
    S.e = fat(a.eta, o);
}
template <typename R> SSPT_DH inline void aloadB(AStepB<R> &S, const AArgs<R> &a, const AThread<R> &T, int q)
{
    const BoxG &b = a.b;
    const uint32_t o = b.o32(clampi(q, 0, b.nr), T.jc, T.kc);
    S.tRR = fat(a.tRR, o);
    S.tTT = fat(a.tTT, o);
    S.tPP = fat(a.tPP, o);
    S.lP = fat(a.li[A_P], o);
    S.lRR = fat(a.li[A_RR], o);
    S.lTT = fat(a.li[A_TT], o);
    S.lPP = fat(a.li[A_PP], o);
    S.lRT = fat(a.li[A_RT], o);
    S.lRP = fat(a.li[A_RP], o);
    S.lTP = fat(a.li[A_TP], o);
}

// ---- phase A: w's of plane q (and w_VR of q+1), power-law adjoint and Q of
// the cells of plane q, gs SM and eta of plane q, from the values S of step q
template <typename R, bool RHS> SSPT_DH inline void aphaseA(AThread<R> &T, ALds<R> &L, const AArgs<R> &a, const AStepA<R> &S, int q, bool wr)
{
    const BoxG &b = a.b;
    const int nr = b.nr;
    const int j = T.j, k = T.k;
    const int qc = clampi(q, 0, nr);
    const uint32_t o = b.o32(qc, T.jc, T.kc);
    T.DTq1 = S.DT1;
    T.lVR1 = S.lVR1;
    T.wVR1 = upd_VR(b, q + 1, j, k) ? ((T.DTq + T.DTq1) * R(0.5)) * T.lVR1 : R(0.0);
    T.lVT = S.lVT;
    T.lVP = S.lVP;
    const R le = S.le, tii = S.tii, ini = S.ini, ePL = S.ePL, lPL = S.lPL;
    const R lb = S.lb, sm = S.sm;
    // the k+1 neighbours' shear stresses: lane tx+1 loaded them as its own
    // (where they are used, at inner cells, k+1 is inside the box, so the
    // clamped positions agree); the last lane loaded its own
    const bool nk = T.tx + 1 < TK;
    const R nRP1 = lane_nb(S, &AStepA<R>::RP1, +1, nk), nTP = lane_nb(S, &AStepA<R>::TP, +1, nk);
    const R nTPj = lane_nb(S, &AStepA<R>::TPj, +1, nk);
    const R RP1k = nk ? nRP1 : S.RP1k, TPk = nk ? nTP : S.TPk, TPjk = nk ? nTPj : S.TPjk;
    // compute_TII!'s 4-point shear averages (times 1/4) of the cell
    const R sX = R(0.25) * ((((T.RTq + T.RTqj) + S.RT1) + S.RT1j) * R(0.25));
    const R sY = R(0.25) * ((((T.RPq + T.RPqk) + S.RP1) + RP1k) * R(0.25));
    const R sZ = R(0.25) * ((((S.TP + TPk) + S.TPj) + TPjk) * R(0.25));
    T.tRT = T.RTq;
    T.tRP = T.RPq;
    T.tTP = S.TP;
    T.RTn = S.RT1;
    T.RTnj = S.RT1j;
    T.RPn = S.RP1;
    T.RPnk = RP1k;
    T.e = S.e;
    // DT(q) of the j-1 / k-1 neighbours, written by them one step ahead (where
    // the weights are used, j-1 and k-1 are inside the box, so the clamped
    // positions of neighbour and thread agree); DT(q+1) for the next step
    const int c = q & 1;
    const R DTjm = T.ty >= 1 ? L.DTL[c][T.ty - 1][T.tx] : S.DTjm;
    const R DTkm = T.tx >= 1 ? L.DTL[c][T.ty][T.tx - 1] : S.DTkm;
    L.DTL[c ^ 1][T.ty][T.tx] = T.DTq1;
    T.wVT = upd_VT(b, q, j, k) ? ((DTjm + T.DTq) * R(0.5)) * T.lVT : R(0.0);
    T.wVP = upd_VP(b, q, j, k) ? ((DTkm + T.DTq) * R(0.5)) * T.lVP : R(0.0);
    // power-law adjoint and Q of the cell
    R Q = R(0.0), qx = R(0.0), qy = R(0.0), qz = R(0.0);
    if (is_cell(b, q, j, k)) {
        R plo;
        const R lt = adj_power_law_v(a, le, tii, ini, ePL, lPL, &plo);
        if (a.do_relax && wr && T.own) aput<RHS>(a, A_PL, o, plo);
        if (inner_c(b, q, j, k)) {
            // Q = lam_tII / tII (compute_TII!'s transpose; 0 where tII = 0)
            const R s = lt + (next_to_boundary(b, q, j, k) ? lb : R(0.0));
            Q = (tii > R(0.0)) ? s * (R(1.0) / tii) : R(0.0);
            qx = Q * sX;
            qy = Q * sY;
            qz = Q * sZ;
        }
    }
    T.Q = Q;
    T.GS = a.gs * sm;
    T.KB = a.kb * sm;
    L.wVR[T.ty][T.tx] = T.wVRq;
    L.wVT[T.ty][T.tx] = T.wVT;
    L.wVP[T.ty][T.tx] = T.wVP;
    L.QX[c][T.ty][T.tx] = qx;
    L.QY[c][T.ty][T.tx] = qy;
    L.QZ[T.ty][T.tx] = qz;
    L.GSM[c][T.ty][T.tx] = T.GS;
    L.ETA[c][T.ty][T.tx] = T.e;
}

// ---- phase B: adjoints of the new P and tau of centre q and of the edges
// stored at box index q (the old-state adjoints: scaled by 1/den)
template <typename R, bool RHS> SSPT_DH inline void aphaseB(AThread<R> &T, ALds<R> &L, const AArgs<R> &a, const AStepB<R> &S, int q, bool wr)
{
    T.tRR = S.tRR;
    T.tTT = S.tTT;
    T.tPP = S.tPP;
    T.lP = S.lP;
    T.lRR = S.lRR;
    T.lTT = S.lTT;
    T.lPP = S.lPP;
    T.lRT = S.lRT;
    T.lRP = S.lRP;
    T.lTP = S.lTP;
    const BoxG &b = a.b;
    const TabsG<R> &t = a.t;
    const int nr = b.nr, nt = b.nt;
    const int j = T.j, k = T.k, tx = T.tx, ty = T.ty;
    const int qt = clampi(q, 0, nr), qt1 = clampi(q + 1, 0, nr), qtm = clampi(q - 1, 0, nr);
    const int jt = clampi(j, 0, nt), jt1 = clampi(j + 1, 0, nt), jtm = clampi(j - 1, 0, nt);
    const uint32_t o = b.o32(qt, T.jc, T.kc);
    const bool up = ty >= 1, lf = tx >= 1, dn = ty + 1 < ATJ, rt_ = tx + 1 < TK;
    const int c = q & 1, pv = c ^ 1; // LDS planes q and q-1
    const R wR = T.wVRq, wT = T.wVT, wP = T.wVP;
    const R wRj = up ? L.wVR[ty - 1][tx] : R(0.0), wRk = lf ? L.wVR[ty][tx - 1] : R(0.0);
    const R wTj = dn ? L.wVT[ty + 1][tx] : R(0.0), wTk = lf ? L.wVT[ty][tx - 1] : R(0.0);
    const R wPk = rt_ ? L.wVP[ty][tx + 1] : R(0.0), wPj = up ? L.wVP[ty - 1][tx] : R(0.0);
    // ---- centre q
    T.cw_dR = T.cw_dT = T.cw_dP = T.cw_aR = T.cw_aT = T.eC = R(0.0);
    if (is_cell(b, q, j, k)) {
        const R wRi = T.wVR1;
        const R iR = tload(t.tr, qt * NTR + fused::TR_invRf), iR1 = tload(t.tr, qt1 * NTR + fused::TR_invRf);
        R sRR = R(0.0), sTT = R(0.0), sPP = R(0.0);
        sRR += wR * (a._dr + iR);
        sTT -= R(0.5) * wR * iR;
        sPP -= R(0.5) * wR * iR;
        sRR += wRi * (-a._dr + iR1);
        sTT -= R(0.5) * wRi * iR1;
        sPP -= R(0.5) * wRi * iR1;
        const R iRy = tload(t.tr, qt * NTR + fused::TR_invRc2);
        const R ct = tload(t.tt, jt * NTT + fused::TT_cotf2), ct1 = tload(t.tt, jt1 * NTT + fused::TT_cotf2);
        sTT += wT * (a._dth * iRy + R(0.5) * ct * iRy);
        sPP -= R(0.5) * wT * ct * iRy;
        sTT += wTj * (-a._dth * iRy + R(0.5) * ct1 * iRy);
        sPP -= R(0.5) * wTj * ct1 * iRy;
        sPP += (wP - wPk) * (tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_z) * a._dph);
        R nRR = T.lRR + sRR, nTT = T.lTT + sTT, nPP = T.lPP + sPP;
        if (inner_c(b, q, j, k)) {
            const R qq = R(0.5) * T.Q;
            nRR += qq * T.tRR;
            nTT += qq * T.tTT;
            nPP += qq * T.tPP;
        }
        const R vP = T.lP - ((sRR + sTT) + sPP);
        const R den = R(1.0) + T.GS / T.e;
        const R id = R(1.0) / den;
        const R vRR = nRR * id, vTT = nTT * id, vPP = nPP * id;
        if (wr && T.own) {
            aput<RHS>(a, A_P, o, vP);
            aput<RHS>(a, A_RR, o, vRR);
            aput<RHS>(a, A_TT, o, vTT);
            aput<RHS>(a, A_PP, o, vPP);
        }
        // compute_P!'s centre stencil weights for vel
        const R GS2 = R(2.0) * T.GS;
        const R xR = GS2 * vRR, xT = GS2 * vTT, xP = GS2 * vPP;
        const R wdiv = -T.KB * vP - (xR + xT + xP) * (R(1.0) / R(3.0));
        const R iRc = tload(t.tr, qt * NTR + fused::TR_invRc);
        T.cw_dR = xR + wdiv;
        T.cw_dT = xT + wdiv;
        T.cw_dP = xP + wdiv;
        T.cw_aR = (xT + xP) * iRc + R(2.0) * wdiv * iRc;
        T.cw_aT = (xP + wdiv) * tload(t.tt, jt * NTT + fused::TT_cotc) * iRc;
        T.eC = ((vRR * T.tRR + vTT * T.tTT) + vPP * T.tPP) * (T.GS / (T.e * T.e));
    }
    // ---- edges of box index q (only rows/lanes with a j-1 / k-1 neighbour)
    T.gRT = T.gRP = T.gTP = T.cRT = T.cRP = T.cTP = R(0.0);
    const R r15m = R(1.5) * tload(t.tr, qtm * NTR + fused::TR_invRc2), r15 = R(1.5) * tload(t.tr, qt * NTR + fused::TR_invRc2);
    if (up && in_RT(b, q, j, k)) {
        const R iR = tload(t.tr, qt * NTR + fused::TR_invRf);
        R s = wRj * (a._dth * iR + R(0.5) * tload(t.tt, jtm * NTT + fused::TT_cotc2) * iR) +
                   wR * (-a._dth * iR + R(0.5) * tload(t.tt, jt * NTT + fused::TT_cotc2) * iR);
        s += L.ACC[ACC_WVT][ty][tx] * (a._dr + r15m) + wT * (-a._dr + r15);
        const R tq = (L.QX[c][ty][tx] + L.QX[c][ty - 1][tx]) + (L.QX[pv][ty][tx] + L.QX[pv][ty - 1][tx]);
        const R g4 = (((L.GSM[pv][ty - 1][tx] + L.GSM[pv][ty][tx]) + L.GSM[c][ty - 1][tx]) + L.GSM[c][ty][tx]) * R(0.25);
        const R e4 = (((L.ETA[pv][ty - 1][tx] + L.ETA[pv][ty][tx]) + L.ETA[c][ty - 1][tx]) + L.ETA[c][ty][tx]) * R(0.25);
        const R v = ((T.lRT + s) + tq) * (R(1.0) / (R(1.0) + g4 / e4));
        if (wr && T.own) aput<RHS>(a, A_RT, o, v);
        T.gRT = g4 * v;
        T.cRT = (R(0.25) * T.tRT * (g4 / (e4 * e4))) * v;
    }
    if (lf && in_RP(b, q, j, k)) {
        R s = (wRk - wR) * (tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_f) * a._dph);
        s += L.ACC[ACC_WVP][ty][tx] * (a._dr + r15m) + wP * (-a._dr + r15);
        const R tq = (L.QY[c][ty][tx] + L.QY[c][ty][tx - 1]) + (L.QY[pv][ty][tx] + L.QY[pv][ty][tx - 1]);
        const R g4 = (((L.GSM[pv][ty][tx - 1] + L.GSM[pv][ty][tx]) + L.GSM[c][ty][tx - 1]) + L.GSM[c][ty][tx]) * R(0.25);
        const R e4 = (((L.ETA[pv][ty][tx - 1] + L.ETA[pv][ty][tx]) + L.ETA[c][ty][tx - 1]) + L.ETA[c][ty][tx]) * R(0.25);
        const R v = ((T.lRP + s) + tq) * (R(1.0) / (R(1.0) + g4 / e4));
        if (wr && T.own) aput<RHS>(a, A_RP, o, v);
        T.gRP = g4 * v;
        T.cRP = (R(0.25) * T.tRP * (g4 / (e4 * e4))) * v;
    }
    if (up && lf && in_TP(b, q, j, k)) {
        R s = (wTk - wT) * (tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_y) * a._dph);
        const R iR = tload(t.tr, qt * NTR + fused::TR_invRc2);
        s += wPj * (a._dth * iR + tload(t.tt, jtm * NTT + fused::TT_cotc2) * iR) +
             wP * (-a._dth * iR + tload(t.tt, jt * NTT + fused::TT_cotc2) * iR);
        const R tq = ((L.QZ[ty][tx] + L.QZ[ty][tx - 1]) + L.QZ[ty - 1][tx]) + L.QZ[ty - 1][tx - 1];
        const R g4 = (((L.GSM[c][ty - 1][tx - 1] + L.GSM[c][ty - 1][tx]) + L.GSM[c][ty][tx - 1]) + L.GSM[c][ty][tx]) * R(0.25);
        const R e4 = (((L.ETA[c][ty - 1][tx - 1] + L.ETA[c][ty - 1][tx]) + L.ETA[c][ty][tx - 1]) + L.ETA[c][ty][tx]) * R(0.25);
        const R v = ((T.lTP + s) + tq) * (R(1.0) / (R(1.0) + g4 / e4));
        if (wr && T.own) aput<RHS>(a, A_TP, o, v);
        T.gTP = g4 * v;
        T.cTP = (R(0.25) * T.tTP * (g4 / (e4 * e4))) * v;
    }
    L.tV[ty][tx] = T.cw_dT * (a._dth * tload(t.tr, qt * NTR + fused::TR_invRc)) + R(0.5) * T.cw_aT;
    L.gRT[ty][tx] = T.gRT;
    L.gTP[ty][tx] = T.gTP;
    L.cRT[ty][tx] = T.cRT;
    L.cTP[ty][tx] = T.cTP;
}

// ---- phase C: lam_V' and lam_eta' contributions of plane q's sig values;
// writes VR(q) and the finished Vth, Vph, eta of plane q-1
template <typename R, bool RHS> SSPT_DH inline void aphaseC(AThread<R> &T, ALds<R> &L, const AArgs<R> &a, int q, bool wr, bool wr_m)
{
    const BoxG &b = a.b;
    const TabsG<R> &t = a.t;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const int j = T.j, k = T.k, tx = T.tx, ty = T.ty;
    const int qt = clampi(q, 0, nr), jt = clampi(j, 0, nt), jt1 = clampi(j + 1, 0, nt);
    const bool up = ty >= 1, lf = tx >= 1, dn = ty + 1 < ATJ, rt_ = tx + 1 < TK;
    const R gRPk = lane_nb(T, &AThread<R>::gRP, +1, rt_), cRPk = lane_nb(T, &AThread<R>::cRP, +1, rt_);
    const R dPm = lane_nb(T, &AThread<R>::cw_dP, -1, lf); // centre k-1's term of Vph(k)
    const R gRTj = dn ? L.gRT[ty + 1][tx] : R(0.0);
    const R gTPk = rt_ ? L.gTP[ty][tx + 1] : R(0.0), gTPj = dn ? L.gTP[ty + 1][tx] : R(0.0);
    const R cRTj = dn ? L.cRT[ty + 1][tx] : R(0.0);
    const R iRe = tload(t.tr, qt * NTR + fused::TR_invRe);
    const bool inj = j >= 0 && j < nt, ink = k >= 0 && k < np;
    // ---- the finished Vth, Vph, eta of plane q-1 (edges of box index q)
    if (wr_m && T.own && q >= 1) {
        const uint32_t om = b.o32(q - 1, T.jc, T.kc);
        const R mR = -a._dr - R(0.5) * iRe;
        if (j >= 0 && j <= nt && ink) aput<RHS>(a, A_VT, om, L.ACC[ACC_VT][ty][tx] + T.gRT * mR);
        if (inj && k >= 0 && k <= np) aput<RHS>(a, A_VP, om, L.ACC[ACC_VP][ty][tx] + T.gRP * mR);
        if (inj && ink) aput<RHS>(a, A_E, om, ((L.ACC[ACC_E][ty][tx] + T.cRT) + cRTj) + (T.cRP + cRPk));
    }
    // ---- VR(q): complete
    if (q <= nr && inj && ink) {
        R v = T.lVRq + L.ACC[ACC_VRN][ty][tx];
        if (q <= nr - 1) v += -T.cw_dR * a._dr + R(0.5) * T.cw_aR;
        const R dthR = a._dth * iRe;
        v += T.gRT * dthR;
        v -= gRTj * dthR;
        const R re = tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_e) * a._dph;
        v += T.gRP * re;
        v -= gRPk * re;
        if (a.bJ && in_objective(a, q, j, k)) v += fat(a.bJ, b.o32(qt, T.jc, T.kc));
        if (wr && T.own) aput<RHS>(a, A_VR, b.o32(qt, T.jc, T.kc), v);
    }
    L.ACC[ACC_VRN][ty][tx] = is_cell(b, q, j, k) ? T.cw_dR * a._dr + R(0.5) * T.cw_aR : R(0.0);
    // ---- Vth(q), Vph(q), eta(q) without the edges of box index q+1
    R vT = R(0.0), vP = R(0.0), vE = R(0.0);
    if (q < nr) {
        const R pR = a._dr - R(0.5) * iRe;
        if (j >= 0 && j <= nt && ink) {
            R v = T.lVT;
            const R dth_R = a._dth * tload(t.tr, qt * NTR + fused::TR_invRc);
            if (up) v += L.tV[ty - 1][tx];
            v += -T.cw_dT * dth_R + R(0.5) * T.cw_aT;
            v += T.gRT * pR;
            const R r4 = tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_c4) * a._dph;
            v += T.gTP * r4;
            v -= gTPk * r4;
            vT = v;
        }
        if (inj && k >= 0 && k <= np) {
            R v = T.lVP;
            const R rc = tload(t.t2, (qt * (nt + 1) + jt) * NT2 + fused::T2_c) * a._dph;
            if (lf) v += dPm * rc;
            v -= T.cw_dP * rc;
            v += T.gRP * pR;
            const R iR4 = tload(t.tr, qt * NTR + fused::TR_invRc4);
            v += T.gTP * (a._dth * iR4 - R(0.5) * tload(t.tt, jt * NTT + fused::TT_cotf4) * iR4);
            v += gTPj * (-a._dth * iR4 - R(0.5) * tload(t.tt, jt1 * NTT + fused::TT_cotf4) * iR4);
            vP = v;
        }
        if (inj && ink) {
            const R cTPk = rt_ ? L.cTP[ty][tx + 1] : R(0.0), cTPj = dn ? L.cTP[ty + 1][tx] : R(0.0);
            const R cTPjk = (dn && rt_) ? L.cTP[ty + 1][tx + 1] : R(0.0);
            vE = ((((T.eC + T.cRT) + T.cRP) + T.cTP) + (cRTj + cRPk)) + ((cTPk + cTPj) + cTPjk);
        }
    }
    L.ACC[ACC_VT][ty][tx] = vT;
    L.ACC[ACC_VP][ty][tx] = vP;
    L.ACC[ACC_E][ty][tx] = vE;
    // ---- carries for step q+1
    T.DTq = T.DTq1;
    T.RTq = T.RTn;
    T.RTqj = T.RTnj;
    T.RPq = T.RPn;
    T.RPqk = T.RPnk;
    T.wVRq = T.wVR1;
    T.lVRq = T.lVR1;
    L.ACC[ACC_WVT][ty][tx] = T.wVT;
    L.ACC[ACC_WVP][ty][tx] = T.wVP;
    (void)up;
}

// Sum of lam_tII over the boundary cells copy_BC! fills from interior cell
// (i, j, k) (the cells at its clamped preimage other than itself).
template <typename R> SSPT_DH inline R lb_sum(const AArgs<R> &a, int i, int j, int k)
{
    const BoxG &b = a.b;
    int di[3] = {0, 0, 0}, dj[3] = {0, 0, 0}, dk[3] = {0, 0, 0};
    int ni = 1, nj = 1, nk = 1;
    if (i == 1) di[ni++] = -1;
    if (i == b.nr - 2) di[ni++] = 1;
    if (j == 1) dj[nj++] = -1;
    if (j == b.nt - 2) dj[nj++] = 1;
    if (k == 1) dk[nk++] = -1;
    if (k == b.np - 2) dk[nk++] = 1;
    R s = R(0.0);
    for (int p = 0; p < ni; ++p)
        for (int r = 0; r < nj; ++r)
            for (int w = 0; w < nk; ++w) {
                if (p == 0 && r == 0 && w == 0) continue;
                R plo;
                s += adj_power_law(a, b.o32(i + di[p], j + dj[r], k + dk[w]), &plo);
            }
    return s;
}

} // namespace adjf
} // namespace sspt
