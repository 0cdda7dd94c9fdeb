// sspt_kernels.hip -- one HIP kernel per reference kernel (the per-kernel
// drop-in path).  One thread per output point, phi (k) mapped to threadIdx.x
// so a wavefront reads 64 consecutive doubles of one phi row.
//
// Each kernel keeps the reference's floating-point association (compiled with
// -ffp-contract=off) so that results match the CPU restatement bit for bit,
// except power_law's pow() (device libm, within an ulp).
#include "sspt_internal.h"
#include "sspt_launch.h"
#include "sspt_fused_body.h"
#include <cmath>

namespace sspt {

static constexpr double PI = 3.14159265358979323846;

// Julia's Base.max / Base.min for Float64 (NaN-propagating, +0 > -0)
__device__ __forceinline__ double jl_max(double x, double y)
{
    bool pick_y = (y > x) || (signbit(y) < signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}
__device__ __forceinline__ double jl_min(double x, double y)
{
    bool pick_y = (y < x) || (signbit(y) > signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

#define BOX_LOOP_3D(NI, NJ, NK)                                   \
    const int k = blockIdx.x * blockDim.x + threadIdx.x;          \
    const int j = blockIdx.y * blockDim.y + threadIdx.y;          \
    const int i = blockIdx.z;                                     \
    if (i >= (NI) || j >= (NJ) || k >= (NK)) return;

// ---------------------------------------------------------------- equal3!
__global__ void __launch_bounds__(256) k_equal3(Box b, double *__restrict__ VRe, double *__restrict__ VTe,
                                                double *__restrict__ VPe, const double *__restrict__ VR,
                                                const double *__restrict__ VT, const double *__restrict__ VP)
{
    BOX_LOOP_3D(b.nr + 1, b.nt + 1, b.np + 1)
    const long long o = b.off(i, j, k);
    if (j < b.nt && k < b.np) VRe[o] = VR[o];
    if (i < b.nr && k < b.np) VTe[o] = VT[o];
    if (i < b.nr && j < b.nt) VPe[o] = VP[o];
}

// ---------------------------------------------------------------- maxloc!
__global__ void __launch_bounds__(256) k_maxloc(Box b, double *__restrict__ SM, const double *__restrict__ eta)
{
    BOX_LOOP_3D(b.nr - 1, b.nt - 1, b.np - 1)
    if (i < 1 || j < 1 || k < 1) return;
    const long long o = b.off(i, j, k);
    SM[o] = jl_max(jl_max(jl_max(jl_max(eta[o - b.si], eta[o + b.si]), eta[o]),
                          jl_max(eta[o - b.sj], eta[o + b.sj])),
                   jl_max(eta[o - 1], eta[o + 1]));
}

// ---------------------------------------------------------------- copy_bc_{r,th,ph}!
// dir 0: faces i=0,n1-1 over (j,k); dir 1: j faces over (i,k); dir 2: k faces over (i,j)
__global__ void __launch_bounds__(256) k_copy_bc(Box b, double *__restrict__ A, int n1, int n2, int n3, int dir)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (dir == 0) {
        if (x >= n3 || y >= n2) return;
        A[b.off(0, y, x)] = A[b.off(1, y, x)];
        A[b.off(n1 - 1, y, x)] = A[b.off(n1 - 2, y, x)];
    } else if (dir == 1) {
        if (x >= n3 || y >= n1) return;
        A[b.off(y, 0, x)] = A[b.off(y, 1, x)];
        A[b.off(y, n2 - 1, x)] = A[b.off(y, n2 - 2, x)];
    } else {
        if (x >= n2 || y >= n1) return;
        A[b.off(y, x, 0)] = A[b.off(y, x, 1)];
        A[b.off(y, x, n3 - 1)] = A[b.off(y, x, n3 - 2)];
    }
}

// ---------------------------------------------------------------- compute_P!
struct PArgs {
    double *P, *DT, *tRR, *tTT, *tPP, *tRT, *tRP, *tTP;
    const double *VR, *VT, *VP, *eta, *SM;
    double _dr, _dth, _dph, kb, gs, vpdt2;
};

__global__ void __launch_bounds__(256) k_compute_P(Box b, Tables t, PArgs a)
{
    BOX_LOOP_3D(b.nr, b.nt + 1, b.np + 1)
    const int nr = b.nr, nt = b.nt, np = b.np;
    const long long o = b.off(i, j, k);
    const double third = 1.0 / 3.0;
    if (j < nt && k < np) { // centre (i,j,k)
        const double R = t.Rc[i];
        const double dVRr = (a.VR[o + b.si] - a.VR[o]) * a._dr;
        const double dVTt = (t.invRc[i] * (a.VT[o + b.sj] - a.VT[o])) * a._dth;
        const double dVPp = (t.rs_c[i * (nt + 1) + j] * (a.VP[o + 1] - a.VP[o])) * a._dph;
        const double avVR = (a.VR[o] + a.VR[o + b.si]) * 0.5;
        const double avVT = (a.VT[o] + a.VT[o + b.sj]) * 0.5;
        const double ct = t.cotc[j];
        const double vr_R = avVR / R, vt_R = avVT / R;
        const double divV = (((dVRr + dVTt) + dVPp) + (2.0 * avVR) / R) + vt_R * ct;
        const double Drr = dVRr - third * divV;
        const double Dtt = (dVTt + vr_R) - third * divV;
        const double Dpp = ((dVPp + vr_R) + vt_R * ct) - third * divV;
        const double sm = a.SM[o], e = a.eta[o];
        const double KB = a.kb * sm, GS = a.gs * sm;
        const double den = 1.0 + GS / e;
        a.P[o] = a.P[o] - KB * divV;
        a.tRR[o] = (a.tRR[o] + (GS * 2.0) * Drr) / den;
        a.tTT[o] = (a.tTT[o] + (GS * 2.0) * Dtt) / den;
        a.tPP[o] = (a.tPP[o] + (GS * 2.0) * Dpp) / den;
        a.DT[o] = a.vpdt2 / (KB + GS / den);
    }
    // tau_Rth at edge (i,j,k): i in [1,nr-1], j in [1,nt-1], k in [1,np-2]
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 2) {
        const long long c00 = o - b.si - b.sj, c01 = o - b.si, c10 = o - b.sj, c11 = o;
        const double G = (((a.gs * a.SM[c00] + a.gs * a.SM[c01]) + a.gs * a.SM[c10]) + a.gs * a.SM[c11]) * 0.25;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[c11]) * 0.25;
        const double avVT = (a.VT[o - b.si] + a.VT[o]) * 0.5;
        const double D = ((a.VT[o] - a.VT[o - b.si]) * a._dr + (t.invRe[i] * (a.VR[o] - a.VR[o - b.sj])) * a._dth) -
                         avVT / t.Re[i];
        a.tRT[o] = (a.tRT[o] + G * D) / (1.0 + G / E);
    }
    // tau_Rph at edge (i,j,k): i in [1,nr-1], j in [1,nt-2], k in [1,np-1]
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 1) {
        const long long c00 = o - b.si - 1, c01 = o - b.si, c10 = o - 1, c11 = o;
        const double G = (((a.gs * a.SM[c00] + a.gs * a.SM[c01]) + a.gs * a.SM[c10]) + a.gs * a.SM[c11]) * 0.25;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[c11]) * 0.25;
        const double avVP = (a.VP[o - b.si] + a.VP[o]) * 0.5;
        const double D = ((a.VP[o] - a.VP[o - b.si]) * a._dr + (t.rs_e[i * (nt + 1) + j] * (a.VR[o] - a.VR[o - 1])) * a._dph) -
                         avVP / t.Re[i];
        a.tRP[o] = (a.tRP[o] + G * D) / (1.0 + G / E);
    }
    // tau_thph at edge (i,j,k): i in [1,nr-2], j in [1,nt-1], k in [1,np-1]
    if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 1) {
        const long long c00 = o - b.sj - 1, c01 = o - b.sj, c10 = o - 1, c11 = o;
        const double G = (((a.gs * a.SM[c00] + a.gs * a.SM[c01]) + a.gs * a.SM[c10]) + a.gs * a.SM[c11]) * 0.25;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[c11]) * 0.25;
        const double avVP = (a.VP[o - b.sj] + a.VP[o]) * 0.5;
        const double D = ((t.invRc4[i] * (a.VP[o] - a.VP[o - b.sj])) * a._dth +
                          (t.rs_c4[i * (nt + 1) + j] * (a.VT[o] - a.VT[o - 1])) * a._dph) -
                         (avVP / t.Rc4[i]) * t.cotf4[j];
        a.tTP[o] = (a.tTP[o] + G * D) / (1.0 + G / E);
    }
}

// ---------------------------------------------------------------- compute_TII!
__global__ void __launch_bounds__(256) k_compute_TII(Box b, double *__restrict__ tII, const double *__restrict__ tRR,
                                                     const double *__restrict__ tTT, const double *__restrict__ tPP,
                                                     const double *__restrict__ tRT, const double *__restrict__ tRP,
                                                     const double *__restrict__ tTP)
{
    BOX_LOOP_3D(b.nr - 1, b.nt - 1, b.np - 1)
    if (i < 1 || j < 1 || k < 1) return;
    const long long o = b.off(i, j, k);
    const double A = tRR[o], B = tTT[o], C = tPP[o];
    const double x = (((tRT[o] + tRT[o + b.sj]) + tRT[o + b.si]) + tRT[o + b.si + b.sj]) * 0.25;
    const double y = (((tRP[o] + tRP[o + 1]) + tRP[o + b.si]) + tRP[o + b.si + 1]) * 0.25;
    const double z = (((tTP[o] + tTP[o + 1]) + tTP[o + b.sj]) + tTP[o + b.sj + 1]) * 0.25;
    tII[o] = sqrt((((1.0 / 2.0) * ((A * A + B * B) + C * C) + x * x) + y * y) + z * z);
}

// ---------------------------------------------------------------- power_law!
__device__ __forceinline__ double power_law_point(double tii, double ini, double *ePL, double tC, double ex,
                                                  double relax, bool do_relax, double emin, double emax)
{
    double pl = ini * fused::pow_ex(tii / tC, ex);
    if (do_relax) {
        pl = exp(log(pl) * relax + log(*ePL) * (1.0 - relax));
        *ePL = pl;
    }
    double e = 2.0 / (1.0 / ini + 1.0 / pl);
    if (emin > 0.0) e = jl_max(e, emin);
    if (emax < INFINITY) e = jl_min(e, emax);
    return e;
}

__global__ void __launch_bounds__(256) k_power_law(Box b, double *__restrict__ eta, const double *__restrict__ tII,
                                                   const double *__restrict__ ini, double *__restrict__ ePL, double tC,
                                                   double ex, double relax, int do_relax, double emin, double emax)
{
    BOX_LOOP_3D(b.nr, b.nt, b.np)
    const long long o = b.off(i, j, k);
    eta[o] = power_law_point(tII[o], ini[o], do_relax ? &ePL[o] : nullptr, tC, ex, relax, do_relax != 0, emin, emax);
}

// ---------------------------------------------------------------- compute_V!
struct VArgs {
    double *VR, *VT, *VP;
    const double *DT, *tRR, *tTT, *tPP, *tRT, *tRP, *tTP, *P, *rho;
    double _dr, _dth, _dph, g;
};

__global__ void __launch_bounds__(256) k_compute_V(Box b, Tables t, VArgs a)
{
    BOX_LOOP_3D(b.nr, b.nt, b.np)
    const int nr = b.nr, nt = b.nt, np = b.np;
    const long long o = b.off(i, j, k);
#define SRR(q) (-a.P[q] + a.tRR[q])
#define STT(q) (-a.P[q] + a.tTT[q])
#define SPP(q) (-a.P[q] + a.tPP[q])
    // VR face (i,j,k): i in [1,nr-1], j in [1,nt-2], k in [1,np-2]
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 2) {
        const long long m = o - b.si;
        const double Rx = t.Rf[i];
        const double sRR = (SRR(m) + SRR(o)) * 0.5;
        const double sTT = (STT(m) + STT(o)) * 0.5;
        const double sPP = (SPP(m) + SPP(o)) * 0.5;
        const double aRT = (a.tRT[o] + a.tRT[o + b.sj]) * 0.5;
        const double rg = (a.rho[m] * a.g + a.rho[o] * a.g) * 0.5;
        const double dV = ((((((SRR(o) - SRR(m)) * a._dr + (t.invRf[i] * (a.tRT[o + b.sj] - a.tRT[o])) * a._dth) +
                              (t.rs_f[i * (nt + 1) + j] * (a.tRP[o + 1] - a.tRP[o])) * a._dph) +
                             (2.0 * sRR) / Rx) -
                            (sTT + sPP) / Rx) +
                           (aRT / Rx) * t.cotc2[j]) -
                          rg;
        const double dt = (a.DT[m] + a.DT[o]) * 0.5;
        a.VR[o] = a.VR[o] + dV * dt;
    }
    // Vth face (i,j,k): i in [1,nr-2], j in [1,nt-1], k in [1,np-2]
    if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 2) {
        const long long m = o - b.sj;
        const double Ry = t.Rc2[i];
        const double sTT = (STT(m) + STT(o)) * 0.5;
        const double sPP = (SPP(m) + SPP(o)) * 0.5;
        const double aRT = (a.tRT[o] + a.tRT[o + b.si]) * 0.5;
        const double dV = ((((a.tRT[o + b.si] - a.tRT[o]) * a._dr + (t.invRc2[i] * (STT(o) - STT(m))) * a._dth) +
                            (t.rs_y[i * (nt + 1) + j] * (a.tTP[o + 1] - a.tTP[o])) * a._dph) +
                           (3.0 * aRT) / Ry) +
                          ((sTT - sPP) / Ry) * t.cotf2[j];
        const double dt = (a.DT[m] + a.DT[o]) * 0.5;
        a.VT[o] = a.VT[o] + dV * dt;
    }
    // Vph face (i,j,k): i in [1,nr-2], j in [1,nt-2], k in [1,np-1]
    if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 1) {
        const long long m = o - 1;
        const double Rz = t.Rc2[i];
        const double aRP = (a.tRP[o] + a.tRP[o + b.si]) * 0.5;
        const double aTP = (a.tTP[o] + a.tTP[o + b.sj]) * 0.5;
        const double dV = ((((a.tRP[o + b.si] - a.tRP[o]) * a._dr + (t.invRc2[i] * (a.tTP[o + b.sj] - a.tTP[o])) * a._dth) +
                            (t.rs_z[i * (nt + 1) + j] * (SPP(o) - SPP(m))) * a._dph) +
                           (3.0 * aRP) / Rz) +
                          ((2.0 * aTP) / Rz) * t.cotc2[j];
        const double dt = (a.DT[m] + a.DT[o]) * 0.5;
        a.VP[o] = a.VP[o] + dV * dt;
    }
#undef SRR
#undef STT
#undef SPP
}

// ---------------------------------------------------------------- BC_R!
__global__ void __launch_bounds__(256) k_bc_R(Box b, double *__restrict__ VR, double *__restrict__ VT,
                                              double *__restrict__ VP)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x; // 0..np
    const int j = blockIdx.y * blockDim.y + threadIdx.y; // 0..nt
    const int nr = b.nr, nt = b.nt, np = b.np;
    if (j < nt && k < np) {
        VR[b.off(0, j, k)] = -VR[b.off(1, j, k)];
        VR[b.off(nr, j, k)] = -VR[b.off(nr - 1, j, k)];
    }
    if (j <= nt && k < np) {
        VT[b.off(0, j, k)] = VT[b.off(1, j, k)];
        VT[b.off(nr - 1, j, k)] = VT[b.off(nr - 2, j, k)];
    }
    if (j < nt && k <= np) {
        VP[b.off(0, j, k)] = VP[b.off(1, j, k)];
        VP[b.off(nr - 1, j, k)] = VP[b.off(nr - 2, j, k)];
    }
}

// ---------------------------------------------------------------- BC_th!
__global__ void __launch_bounds__(256) k_bc_th(Box b, double *__restrict__ VR, double *__restrict__ VT,
                                               double *__restrict__ VP, const double *__restrict__ th, int mode,
                                               double r, double sh, double v0, double v1)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x; // 0..np
    const int i = blockIdx.y * blockDim.y + threadIdx.y; // 0..nr
    const int nr = b.nr, nt = b.nt, np = b.np;
    if (i <= nr && k < np) {
        if (mode == SS_BC_SHEAR) {
            VR[b.off(i, 0, k)] = 0.0;
            VR[b.off(i, nt - 1, k)] = 0.0;
        } else {
            VR[b.off(i, 0, k)] = VR[b.off(i, 1, k)];
            VR[b.off(i, nt - 1, k)] = VR[b.off(i, nt - 2, k)];
        }
    }
    if (i < nr && k < np) {
        if (mode == SS_BC_PURE_SHEAR) { // v0, v1: prescribed Vth at th[0], th[nt-1]
            VT[b.off(i, 0, k)] = 2.0 * v0 - VT[b.off(i, 1, k)];
            VT[b.off(i, nt, k)] = 2.0 * v1 - VT[b.off(i, nt - 1, k)];
        } else {
            VT[b.off(i, 0, k)] = -VT[b.off(i, 1, k)];
            VT[b.off(i, nt, k)] = -VT[b.off(i, nt - 1, k)];
        }
    }
    if (i < nr && k <= np) {
        if (mode == SS_BC_SHEAR) {
            VP[b.off(i, 0, k)] = ((th[0] - PI / 2.0) * r) * sh;
            VP[b.off(i, nt - 1, k)] = ((th[nt - 1] - PI / 2.0) * r) * sh;
        } else {
            VP[b.off(i, 0, k)] = VP[b.off(i, 1, k)];
            VP[b.off(i, nt - 1, k)] = VP[b.off(i, nt - 2, k)];
        }
    }
}

// ---------------------------------------------------------------- BC_ph!
__global__ void __launch_bounds__(256) k_bc_ph(Box b, double *__restrict__ VR, double *__restrict__ VT,
                                               double *__restrict__ VP, const double *__restrict__ th,
                                               const double *__restrict__ ph, int mode, double r, double sh)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x; // 0..nt
    const int i = blockIdx.y * blockDim.y + threadIdx.y; // 0..nr
    const int nr = b.nr, nt = b.nt, np = b.np;
    if (i <= nr && j < nt) {
        if (mode == SS_BC_SHEAR) {
            VR[b.off(i, j, 0)] = 0.0;
            VR[b.off(i, j, np - 1)] = 0.0;
        } else {
            VR[b.off(i, j, 0)] = VR[b.off(i, j, 1)];
            VR[b.off(i, j, np - 1)] = VR[b.off(i, j, np - 2)];
        }
    }
    if (i < nr && j <= nt) {
        if (mode == SS_BC_SHEAR) {
            VT[b.off(i, j, 0)] = 0.0;
            VT[b.off(i, j, np - 1)] = 0.0;
        } else {
            VT[b.off(i, j, 0)] = VT[b.off(i, j, 1)];
            VT[b.off(i, j, np - 1)] = VT[b.off(i, j, np - 2)];
        }
    }
    if (i < nr && j < nt) {
        const double in_lo = VP[b.off(i, j, 1)], in_hi = VP[b.off(i, j, np - 1)];
        double lo, hi;
        if (mode == SS_BC_SHEAR) {
            const double s = ((2.0 * (th[j] - PI / 2.0)) * r) * sh;
            lo = s - in_lo;
            hi = s - in_hi;
        } else if (mode == SS_BC_FREE_SLIP) {
            lo = -in_lo;
            hi = -in_hi;
        } else {
            lo = 2.0 * ((ph[0] * r) * sh) - in_lo;
            hi = 2.0 * ((ph[np - 1] * r) * sh) - in_hi;
        }
        VP[b.off(i, j, 0)] = lo;
        VP[b.off(i, j, np)] = hi;
    }
}

// ---------------------------------------------------------------- check_err! + max reductions
// red[0..2] = max|V-V_err| bits, red[3..5] = max|V| bits (non-negative doubles
// order like their unsigned bit patterns; NaN sorts above +inf, so a NaN
// anywhere propagates, as Julia's maximum does).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    for (int s = 32; s > 0; s >>= 1) {
        unsigned long long o = __shfl_xor(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

constexpr int CHECK_PLANES = 32;

__global__ void __launch_bounds__(256) k_check_err(Box b, double *__restrict__ VRe, double *__restrict__ VTe,
                                                   double *__restrict__ VPe, const double *__restrict__ VR,
                                                   const double *__restrict__ VT, const double *__restrict__ VP,
                                                   unsigned long long *red, int write_err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long me[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
    // CHECK_PLANES r-planes per thread: one set of atomics per wavefront
    const int i0 = blockIdx.z * CHECK_PLANES, i1 = min(i0 + CHECK_PLANES, b.nr + 1);
    for (int i = i0; i < i1 && j <= b.nt && k <= b.np; ++i) {
        const long long o = b.off(i, j, k);
        const bool in[3] = {j < b.nt && k < b.np, i < b.nr && k < b.np, i < b.nr && j < b.nt};
        double *E[3] = {VRe, VTe, VPe};
        const double *V[3] = {VR, VT, VP};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!in[c]) continue;
            const double v = V[c][o];
            const double d = v - E[c][o];
            if (write_err) E[c][o] = d;
            const unsigned long long ud = (unsigned long long)__double_as_longlong(fabs(d));
            const unsigned long long uv = (unsigned long long)__double_as_longlong(fabs(v));
            me[c] = ud > me[c] ? ud : me[c];
            mv[c] = uv > mv[c] ? uv : mv[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        me[c] = wave_max_u64(me[c]);
        mv[c] = wave_max_u64(mv[c]);
    }
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (me[c]) atomicMax(&red[c], me[c]);
            if (mv[c]) atomicMax(&red[3 + c], mv[c]);
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int t = threadIdx.x;
    if (t < n) p[t] = 0ull;
}

// ---------------------------------------------------------------- setup kernels
__global__ void __launch_bounds__(256) k_fill(Box b, double *__restrict__ A, int n1, int n2, int n3, int o1, double v)
{
    BOX_LOOP_3D(n1, n2, n3)
    A[b.off(i + o1, j + o1, k + o1)] = v;
}

__global__ void __launch_bounds__(256) k_copy_field(Box b, double *__restrict__ D, const double *__restrict__ S, int n1,
                                                    int n2, int n3, int o1)
{
    BOX_LOOP_3D(n1, n2, n3)
    const long long o = b.off(i + o1, j + o1, k + o1);
    D[o] = S[o];
}

__global__ void __launch_bounds__(256) k_init_inclusion(Box b, Tables t, double *__restrict__ A, double r, double radius,
                                                        double in)
{
    BOX_LOOP_3D(b.nr, b.nt, b.np)
    const double x = t.Rc[i] - r, y = (t.th[j] - PI / 2.0) * r, z = t.ph[k] * r;
    if (((x * x + y * y) + z * z) < radius) A[b.off(i, j, k)] = in;
}

__global__ void __launch_bounds__(256) k_init_velocity(Box b, Tables t, double *__restrict__ VP, double fact, double r)
{
    BOX_LOOP_3D(b.nr, b.nt, b.np + 1)
    VP[b.off(i, j, k)] = ((t.th[j] - PI / 2.0) * r) * fact;
}

// ---------------------------------------------------------------- layout permutation
// Julia layout (n1 fastest, r) <-> box (phi fastest).  For each (j), the
// (i,k) plane is transposed through an LDS tile so that both sides access
// memory contiguously.
static constexpr int TT = 32;
__global__ void __launch_bounds__(256) k_julia_to_box(Box b, double *__restrict__ dst, const double *__restrict__ src,
                                                      int n1, int n2, int n3, int o1)
{
    __shared__ double tile[TT][TT + 1];
    const int j = blockIdx.z;
    const int i0 = blockIdx.x * TT, k0 = blockIdx.y * TT;
    for (int y = threadIdx.y; y < TT; y += blockDim.y) {
        const int k = k0 + y, i = i0 + threadIdx.x;
        if (i < n1 && k < n3) tile[y][threadIdx.x] = src[((size_t)k * n2 + j) * n1 + i];
    }
    __syncthreads();
    for (int y = threadIdx.y; y < TT; y += blockDim.y) {
        const int i = i0 + y, k = k0 + threadIdx.x;
        if (i < n1 && k < n3) dst[b.off(i + o1, j + o1, k + o1)] = tile[threadIdx.x][y];
    }
}

__global__ void __launch_bounds__(256) k_box_to_julia(Box b, double *__restrict__ dst, const double *__restrict__ src,
                                                      int n1, int n2, int n3, int o1)
{
    __shared__ double tile[TT][TT + 1];
    const int j = blockIdx.z;
    const int i0 = blockIdx.x * TT, k0 = blockIdx.y * TT;
    for (int y = threadIdx.y; y < TT; y += blockDim.y) {
        const int i = i0 + y, k = k0 + threadIdx.x;
        if (i < n1 && k < n3) tile[threadIdx.x][y] = src[b.off(i + o1, j + o1, k + o1)];
    }
    __syncthreads();
    for (int y = threadIdx.y; y < TT; y += blockDim.y) {
        const int k = k0 + y, i = i0 + threadIdx.x;
        if (i < n1 && k < n3) dst[((size_t)k * n2 + j) * n1 + i] = tile[y][threadIdx.x];
    }
}

__global__ void __launch_bounds__(256) k_c_to_box(Box b, double *__restrict__ dst, const double *__restrict__ src, int n1,
                                                  int n2, int n3, int o1)
{
    BOX_LOOP_3D(n1, n2, n3)
    dst[b.off(i + o1, j + o1, k + o1)] = src[((size_t)i * n2 + j) * n3 + k];
}

__global__ void __launch_bounds__(256) k_box_to_c(Box b, double *__restrict__ dst, const double *__restrict__ src, int n1,
                                                  int n2, int n3, int o1)
{
    BOX_LOOP_3D(n1, n2, n3)
    dst[((size_t)i * n2 + j) * n3 + k] = src[b.off(i + o1, j + o1, k + o1)];
}

// ================================================================ launchers
static inline dim3 grid3(int ni, int nj, int nk, dim3 blk)
{
    return dim3((nk + blk.x - 1) / blk.x, (nj + blk.y - 1) / blk.y, ni);
}
static const dim3 B3(64, 4, 1);
static const dim3 B2(64, 4, 1);

void launch_equal3(ss_ctx *c, const FieldSet &s)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_equal3, grid3(b.nr + 1, b.nt + 1, b.np + 1, B3), B3, 0, c->stream, b, s.f[SS_VR_ERR],
                       s.f[SS_VTH_ERR], s.f[SS_VPH_ERR], s.f[SS_VR], s.f[SS_VTH], s.f[SS_VPH]);
}

void launch_maxloc(ss_ctx *c, double *SM, const double *eta)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_maxloc, grid3(b.nr - 1, b.nt - 1, b.np - 1, B3), B3, 0, c->stream, b, SM, eta);
}

void launch_copy_bc(ss_ctx *c, double *A, int n1, int n2, int n3)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_copy_bc, dim3((n3 + 63) / 64, (n2 + 3) / 4), B2, 0, c->stream, b, A, n1, n2, n3, 0);
    hipLaunchKernelGGL(k_copy_bc, dim3((n3 + 63) / 64, (n1 + 3) / 4), B2, 0, c->stream, b, A, n1, n2, n3, 1);
    hipLaunchKernelGGL(k_copy_bc, dim3((n2 + 63) / 64, (n1 + 3) / 4), B2, 0, c->stream, b, A, n1, n2, n3, 2);
}

void launch_compute_P(ss_ctx *c, const FieldSet &s, double _dr, double _dth, double _dph, double dmp, double vpdt,
                      double lr)
{
    const Box &b = c->box;
    PArgs a;
    a.P = s.f[SS_P];
    a.DT = s.f[SS_DT_R];
    a.tRR = s.f[SS_TAU_RR];
    a.tTT = s.f[SS_TAU_THTH];
    a.tPP = s.f[SS_TAU_PHPH];
    a.tRT = s.f[SS_TAU_RTH];
    a.tRP = s.f[SS_TAU_RPH];
    a.tTP = s.f[SS_TAU_THPH];
    a.VR = s.f[SS_VR];
    a.VT = s.f[SS_VTH];
    a.VP = s.f[SS_VPH];
    a.eta = s.f[SS_ETA];
    a.SM = s.f[SS_ETASM];
    a._dr = _dr;
    a._dth = _dth;
    a._dph = _dph;
    a.kb = (((dmp * 2.0) * PI) * vpdt) / lr; // @KBDT prefactor
    a.gs = ((4.0 * PI) * vpdt) / lr;         // @GSDT prefactor
    a.vpdt2 = vpdt * vpdt;
    hipLaunchKernelGGL(k_compute_P, grid3(b.nr, b.nt + 1, b.np + 1, B3), B3, 0, c->stream, b, c->tab, a);
}

void launch_compute_TII(ss_ctx *c, const FieldSet &s)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_compute_TII, grid3(b.nr - 1, b.nt - 1, b.np - 1, B3), B3, 0, c->stream, b, s.f[SS_TII],
                       s.f[SS_TAU_RR], s.f[SS_TAU_THTH], s.f[SS_TAU_PHPH], s.f[SS_TAU_RTH], s.f[SS_TAU_RPH],
                       s.f[SS_TAU_THPH]);
}

void launch_power_law(ss_ctx *c, const FieldSet &s, double tau_C, double n_exp, double relax, double emin, double emax)
{
    const Box &b = c->box;
    const int do_relax = (relax != 1.0) ? 1 : 0;
    hipLaunchKernelGGL(k_power_law, grid3(b.nr, b.nt, b.np, B3), B3, 0, c->stream, b, s.f[SS_ETA], s.f[SS_TII],
                       s.f[SS_ETA_INI], s.f[SS_ETA_PL], tau_C, 1.0 - n_exp, relax, do_relax, emin, emax);
}

void launch_compute_V(ss_ctx *c, const FieldSet &s, double _dr, double _dth, double _dph, double g)
{
    const Box &b = c->box;
    VArgs a;
    a.VR = s.f[SS_VR];
    a.VT = s.f[SS_VTH];
    a.VP = s.f[SS_VPH];
    a.DT = s.f[SS_DT_R];
    a.tRR = s.f[SS_TAU_RR];
    a.tTT = s.f[SS_TAU_THTH];
    a.tPP = s.f[SS_TAU_PHPH];
    a.tRT = s.f[SS_TAU_RTH];
    a.tRP = s.f[SS_TAU_RPH];
    a.tTP = s.f[SS_TAU_THPH];
    a.P = s.f[SS_P];
    a.rho = s.f[SS_RHO];
    a._dr = _dr;
    a._dth = _dth;
    a._dph = _dph;
    a.g = g;
    hipLaunchKernelGGL(k_compute_V, grid3(b.nr, b.nt, b.np, B3), B3, 0, c->stream, b, c->tab, a);
}

void launch_bc_R(ss_ctx *c, const FieldSet &s)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_bc_R, dim3((b.np + 1 + 63) / 64, (b.nt + 1 + 3) / 4), B2, 0, c->stream, b, s.f[SS_VR],
                       s.f[SS_VTH], s.f[SS_VPH]);
}

void launch_bc_th(ss_ctx *c, const FieldSet &s, int mode, double r, double sh)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_bc_th, dim3((b.np + 1 + 63) / 64, (b.nr + 1 + 3) / 4), B2, 0, c->stream, b, s.f[SS_VR],
                       s.f[SS_VTH], s.f[SS_VPH], c->tab.th, mode, r, sh, fused::pure_shear_vth(c->th_lo, r, sh),
                       fused::pure_shear_vth(c->th_hi, r, sh));
}

void launch_bc_ph(ss_ctx *c, const FieldSet &s, int mode, double r, double sh)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_bc_ph, dim3((b.nt + 1 + 63) / 64, (b.nr + 1 + 3) / 4), B2, 0, c->stream, b, s.f[SS_VR],
                       s.f[SS_VTH], s.f[SS_VPH], c->tab.th, c->tab.ph, mode, r, sh);
}

void launch_check_err(ss_ctx *c, const FieldSet &s, const double *VRprev, const double *VTprev, const double *VPprev,
                      int write_err)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->red, 6);
    // when prev != V_err the kernel only reads (write_err = 0) or writes V_err
    hipLaunchKernelGGL(k_check_err, grid3((b.nr + 1 + CHECK_PLANES - 1) / CHECK_PLANES, b.nt + 1, b.np + 1, B3), B3, 0,
                       c->stream, b,
                       const_cast<double *>(VRprev), const_cast<double *>(VTprev), const_cast<double *>(VPprev),
                       s.f[SS_VR], s.f[SS_VTH], s.f[SS_VPH], c->red, write_err);
}

void launch_fill(ss_ctx *c, double *A, int n1, int n2, int n3, int o1, double v)
{
    hipLaunchKernelGGL(k_fill, grid3(n1, n2, n3, B3), B3, 0, c->stream, c->box, A, n1, n2, n3, o1, v);
}

void launch_copy_field(ss_ctx *c, double *D, const double *S, int n1, int n2, int n3, int o1)
{
    hipLaunchKernelGGL(k_copy_field, grid3(n1, n2, n3, B3), B3, 0, c->stream, c->box, D, S, n1, n2, n3, o1);
}

void launch_init_inclusion(ss_ctx *c, double *A, double r, double radius, double in)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_init_inclusion, grid3(b.nr, b.nt, b.np, B3), B3, 0, c->stream, b, c->tab, A, r, radius, in);
}

void launch_init_velocity(ss_ctx *c, double *VP, double fact, double r)
{
    const Box &b = c->box;
    hipLaunchKernelGGL(k_init_velocity, grid3(b.nr, b.nt, b.np + 1, B3), B3, 0, c->stream, b, c->tab, VP, fact, r);
}

void launch_to_box(ss_ctx *c, double *dst, const double *src, int n1, int n2, int n3, int o1, int layout)
{
    const Box &b = c->box;
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_julia_to_box, dim3((n1 + TT - 1) / TT, (n3 + TT - 1) / TT, n2), dim3(TT, 8), 0, c->stream,
                           b, dst, src, n1, n2, n3, o1);
    else
        hipLaunchKernelGGL(k_c_to_box, grid3(n1, n2, n3, B3), B3, 0, c->stream, b, dst, src, n1, n2, n3, o1);
}

void launch_from_box(ss_ctx *c, double *dst, const double *src, int n1, int n2, int n3, int o1, int layout)
{
    const Box &b = c->box;
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_box_to_julia, dim3((n1 + TT - 1) / TT, (n3 + TT - 1) / TT, n2), dim3(TT, 8), 0, c->stream,
                           b, dst, src, n1, n2, n3, o1);
    else
        hipLaunchKernelGGL(k_box_to_c, grid3(n1, n2, n3, B3), B3, 0, c->stream, b, dst, src, n1, n2, n3, o1);
}

} // namespace sspt
