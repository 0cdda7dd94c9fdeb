// sspt_mixed.hip -- FP32 inner sweeps inside an FP64 defect-correction loop,
// and the FP64 residual used to measure time-to-tolerance.
//
// Outer step (FP64):
//   1. tau := the stress consistent with V and the current eta (the fixed
//      point of the PT stress update: tau_nn = 2*eta*D_nn on centres,
//      tau_sh = eta_avg*D_sh on edges), so the constitutive residual is zero;
//   2. residual: gP = divV on centres, fV = div(sigma) - rho g on interior
//      faces (the compute_P!/compute_V! operators), stored in FP32 as the
//      forcing of the correction equations, with their max-norms;
//   3. FP32 sweeps of the fused PT iteration on the correction
//      (V, P, tau) -> (V, P, tau) + d with frozen eta, homogeneous BCs
//      (sh = 0) and the forcing in place of rho g and the continuity RHS;
//   4. V, P, tau += d in FP64; for a power law, eta from the new stresses.
// The fixed point satisfies the FP64 equations, so the answer is the FP64
// solution; FP32 only limits how much one outer step can gain.
#include "sspt_internal.h"
#include "sspt_launch.h"
#include "sspt_fused_body.h"
#include "sspt_tables.h"
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace sspt {

using fused::Args;

__device__ __forceinline__ unsigned long long abits(double x)
{
    return (unsigned long long)__double_as_longlong(fabs(x));
}
__device__ __forceinline__ unsigned long long wmax(unsigned long long v)
{
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned long long o = __shfl_xor(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long umax(unsigned long long x, unsigned long long y) { return x > y ? x : y; }
constexpr int RES_PLANES = 32;

#define M_LOOP(NI, NJ, NK)                                        \
    const int k = blockIdx.x * blockDim.x + threadIdx.x;          \
    const int j = blockIdx.y * blockDim.y + threadIdx.y;          \
    const int i = blockIdx.z;                                     \
    if (i >= (NI) || j >= (NJ) || k >= (NK)) return;

struct ConsArgs {
    const double *VR, *VT, *VP, *eta;
    double *tRR, *tTT, *tPP, *tRT, *tRP, *tTP;
    double _dr, _dth, _dph;
};

// tau consistent with V and eta (strain rates as in compute_P!)
__global__ void __launch_bounds__(256) k_consistent_stress(Box b, Tables t, ConsArgs a)
{
    M_LOOP(b.nr, b.nt, b.np)
    const int nr = b.nr, nt = b.nt, np = b.np;
    const long long o = b.off(i, j, k);
    const double third = 1.0 / 3.0;
    {
        const double R = t.Rc[i];
        const double dVRr = (a.VR[o + b.si] - a.VR[o]) * a._dr;
        const double dVTt = (t.invRc[i] * (a.VT[o + b.sj] - a.VT[o])) * a._dth;
        const double dVPp = (t.rs_c[i * (nt + 1) + j] * (a.VP[o + 1] - a.VP[o])) * a._dph;
        const double avVR = (a.VR[o] + a.VR[o + b.si]) * 0.5;
        const double avVT = (a.VT[o] + a.VT[o + b.sj]) * 0.5;
        const double ct = t.cotc[j];
        const double vr_R = avVR / R, vt_R = avVT / R;
        const double divV = (((dVRr + dVTt) + dVPp) + (2.0 * avVR) / R) + vt_R * ct;
        const double e2 = 2.0 * a.eta[o];
        a.tRR[o] = e2 * (dVRr - third * divV);
        a.tTT[o] = e2 * ((dVTt + vr_R) - third * divV);
        a.tPP[o] = e2 * (((dVPp + vr_R) + vt_R * ct) - third * divV);
    }
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 2) {
        const long long c00 = o - b.si - b.sj, c01 = o - b.si, c10 = o - b.sj;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]) * 0.25;
        const double D = ((a.VT[o] - a.VT[o - b.si]) * a._dr + (t.invRe[i] * (a.VR[o] - a.VR[o - b.sj])) * a._dth) -
                         ((a.VT[o - b.si] + a.VT[o]) * 0.5) / t.Re[i];
        a.tRT[o] = E * D;
    }
    if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 1) {
        const long long c00 = o - b.si - 1, c01 = o - b.si, c10 = o - 1;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]) * 0.25;
        const double D = ((a.VP[o] - a.VP[o - b.si]) * a._dr + (t.rs_e[i * (nt + 1) + j] * (a.VR[o] - a.VR[o - 1])) * a._dph) -
                         ((a.VP[o - b.si] + a.VP[o]) * 0.5) / t.Re[i];
        a.tRP[o] = E * D;
    }
    if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 1) {
        const long long c00 = o - b.sj - 1, c01 = o - b.sj, c10 = o - 1;
        const double E = (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]) * 0.25;
        const double D = ((t.invRc4[i] * (a.VP[o] - a.VP[o - b.sj])) * a._dth +
                          (t.rs_c4[i * (nt + 1) + j] * (a.VT[o] - a.VT[o - 1])) * a._dph) -
                         ((((a.VP[o - b.sj] + a.VP[o]) * 0.5) / t.Rc4[i]) * t.cotf4[j]);
        a.tTP[o] = E * D;
    }
}

struct ResArgs {
    const double *VR, *VT, *VP, *P, *tRR, *tTT, *tPP, *tRT, *tRP, *tTP, *rho;
    float *gP, *fVR, *fVT, *fVP; // may be null (norms only)
    double _dr, _dth, _dph, g;
    unsigned long long *red;     // [0] max|divV| [1..3] max|fV| [4..6] max|V| [7] max|rho g| [8..10] max|tau_nn| [11] max|P|
};

// Continuity and momentum residuals of the FP64 state (compute_P!'s divV,
// compute_V!'s dV), written as FP32 forcing, and their max-norms.
__global__ void __launch_bounds__(256) k_residual(Box b, Tables t, ResArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int nr = b.nr, nt = b.nt, np = b.np;
    unsigned long long m[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    // each thread marches RES_PLANES planes and keeps running maxima, so
    // that a wavefront issues its atomics once, not once per plane
    const int i0 = blockIdx.z * RES_PLANES, i1 = min(i0 + RES_PLANES, nr + 1);
    for (int i = i0; i < i1; ++i) {
        if (j > nt || k > np) break;
        const long long o = b.off(i, j, k);
        if (i < nr && j < nt && k < np) { // centre
            const double R = t.Rc[i];
            const double dVRr = (a.VR[o + b.si] - a.VR[o]) * a._dr;
            const double dVTt = (t.invRc[i] * (a.VT[o + b.sj] - a.VT[o])) * a._dth;
            const double dVPp = (t.rs_c[i * (nt + 1) + j] * (a.VP[o + 1] - a.VP[o])) * a._dph;
            const double avVR = (a.VR[o] + a.VR[o + b.si]) * 0.5;
            const double avVT = (a.VT[o] + a.VT[o + b.sj]) * 0.5;
            const double divV = (((dVRr + dVTt) + dVPp) + (2.0 * avVR) / R) + (avVT / R) * t.cotc[j];
            if (a.gP) a.gP[o] = (float)divV;
            m[0] = umax(m[0], abits(divV));
            m[7] = umax(m[7], abits(a.rho[o] * a.g));
            m[8] = umax(m[8], abits(a.tRR[o]));
            m[9] = umax(m[9], abits(a.tTT[o]));
            m[10] = umax(m[10], abits(a.tPP[o]));
            m[11] = umax(m[11], abits(a.P[o]));
        }
        if (j < nt && k < np) m[4] = umax(m[4], abits(a.VR[o]));
        if (i < nr && k < np) m[5] = umax(m[5], abits(a.VT[o]));
        if (i < nr && j < nt) m[6] = umax(m[6], abits(a.VP[o]));
#define SRR(q) (-a.P[q] + a.tRR[q])
#define STT(q) (-a.P[q] + a.tTT[q])
#define SPP(q) (-a.P[q] + a.tPP[q])
        float fr = 0.f, ft = 0.f, fp = 0.f;
        if (i >= 1 && i <= nr - 1 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 2) {
            const long long q = o - b.si;
            const double Rx = t.Rf[i];
            const double sRR = (SRR(q) + SRR(o)) * 0.5, sTT = (STT(q) + STT(o)) * 0.5, sPP = (SPP(q) + SPP(o)) * 0.5;
            const double aRT = (a.tRT[o] + a.tRT[o + b.sj]) * 0.5;
            const double rg = (a.rho[q] * a.g + a.rho[o] * a.g) * 0.5;
            const double dV = ((((((SRR(o) - SRR(q)) * a._dr + (t.invRf[i] * (a.tRT[o + b.sj] - a.tRT[o])) * a._dth) +
                                  (t.rs_f[i * (nt + 1) + j] * (a.tRP[o + 1] - a.tRP[o])) * a._dph) +
                                 (2.0 * sRR) / Rx) - (sTT + sPP) / Rx) + (aRT / Rx) * t.cotc2[j]) - rg;
            fr = (float)dV;
            m[1] = umax(m[1], abits(dV));
        }
        if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 1 && k >= 1 && k <= np - 2) {
            const long long q = o - b.sj;
            const double Ry = t.Rc2[i];
            const double sTT = (STT(q) + STT(o)) * 0.5, sPP = (SPP(q) + SPP(o)) * 0.5;
            const double aRT = (a.tRT[o] + a.tRT[o + b.si]) * 0.5;
            const double dV = ((((a.tRT[o + b.si] - a.tRT[o]) * a._dr + (t.invRc2[i] * (STT(o) - STT(q))) * a._dth) +
                                (t.rs_y[i * (nt + 1) + j] * (a.tTP[o + 1] - a.tTP[o])) * a._dph) + (3.0 * aRT) / Ry) +
                              ((sTT - sPP) / Ry) * t.cotf2[j];
            ft = (float)dV;
            m[2] = umax(m[2], abits(dV));
        }
        if (i >= 1 && i <= nr - 2 && j >= 1 && j <= nt - 2 && k >= 1 && k <= np - 1) {
            const long long q = o - 1;
            const double Rz = t.Rc2[i];
            const double aRP = (a.tRP[o] + a.tRP[o + b.si]) * 0.5, aTP = (a.tTP[o] + a.tTP[o + b.sj]) * 0.5;
            const double dV = ((((a.tRP[o + b.si] - a.tRP[o]) * a._dr + (t.invRc2[i] * (a.tTP[o + b.sj] - a.tTP[o])) * a._dth) +
                                (t.rs_z[i * (nt + 1) + j] * (SPP(o) - SPP(q))) * a._dph) + (3.0 * aRP) / Rz) +
                              ((2.0 * aTP) / Rz) * t.cotc2[j];
            fp = (float)dV;
            m[3] = umax(m[3], abits(dV));
        }
#undef SRR
#undef STT
#undef SPP
        if (a.fVR) {
            a.fVR[o] = fr;
            a.fVT[o] = ft;
            a.fVP[o] = fp;
        }
    }
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const unsigned long long v = wmax(m[c]);
        if (((threadIdx.x + threadIdx.y * blockDim.x) & 63) == 0 && v) atomicMax(&a.red[c], v);
    }
}

// state (FP64) += correction (FP32) over each field's reference shape
__global__ void __launch_bounds__(256) k_accumulate(Box b, double *__restrict__ A, const float *__restrict__ d, int n1,
                                                    int n2, int n3, int o1)
{
    M_LOOP(n1, n2, n3)
    const long long o = b.off(i + o1, j + o1, k + o1);
    A[o] = A[o] + (double)d[o];
}

__global__ void __launch_bounds__(256) k_to_f32(Box b, float *__restrict__ D, const double *__restrict__ S, int n1,
                                                int n2, int n3)
{
    M_LOOP(n1, n2, n3)
    const long long o = b.off(i, j, k);
    D[o] = (float)S[o];
}

__global__ void __launch_bounds__(256) k_zero_f32(float *__restrict__ p, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.f;
}

__global__ void k_zero_slots(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

static const dim3 MB(64, 4, 1);
static inline dim3 mgrid(int ni, int nj, int nk) { return dim3((nk + 63) / 64, (nj + 3) / 4, ni); }

void launch_consistent_stress(ss_ctx *c, const FieldSet &s, double *const tau[6], double _dr, double _dth,
                              double _dph)
{
    ConsArgs a{s.f[SS_VR], s.f[SS_VTH], s.f[SS_VPH], s.f[SS_ETA], tau[0], tau[1], tau[2], tau[3], tau[4], tau[5],
               _dr, _dth, _dph};
    const Box &b = c->box;
    hipLaunchKernelGGL(k_consistent_stress, mgrid(b.nr, b.nt, b.np), MB, 0, c->stream, b, c->tab, a);
}

void launch_residual(ss_ctx *c, const FieldSet &s, double *const tau[6], float *gP, float *fV[3], double _dr,
                     double _dth, double _dph, double g)
{
    ResArgs a{s.f[SS_VR], s.f[SS_VTH], s.f[SS_VPH], s.f[SS_P], tau[0], tau[1], tau[2], tau[3], tau[4], tau[5],
              s.f[SS_RHO], gP, fV ? fV[0] : nullptr, fV ? fV[1] : nullptr, fV ? fV[2] : nullptr,
              _dr, _dth, _dph, g, c->red};
    const Box &b = c->box;
    hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(64), 0, c->stream, c->red, 16);
    hipLaunchKernelGGL(k_residual, mgrid((b.nr + 1 + RES_PLANES - 1) / RES_PLANES, b.nt + 1, b.np + 1), MB, 0, c->stream,
                       b, c->tab, a);
}

template <typename R> void launch_fused_args(ss_ctx *c, Args<R> a, hipEvent_t ev_main0, hipEvent_t ev_main1);
template <typename R> void launch_etasm_spread_box(ss_ctx *c, const fused::BoxG &b, const R *eta, R *sm, R *tmp, int m);

} // namespace sspt

// ======================================================================== API
using namespace sspt;

namespace {

int mfail(ss_ctx *c, int code, const char *msg)
{
    c->err = msg;
    return code;
}

int msync(ss_ctx *c, const char *where)
{
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        c->err = std::string(where) + ": " + hipGetErrorString(e);
        return (int)e;
    }
    return SS_OK;
}

double ubits(unsigned long long u)
{
    double d;
    memcpy(&d, &u, 8);
    return d;
}

// Scratch FP64 stress boxes for residual evaluation of a state that must
// not be modified (the plain solve), allocated on first use.
int ensure_tau_scratch(ss_ctx *c)
{
    if (c->tau_scr[0]) return SS_OK;
    for (int q = 0; q < 6; ++q) {
        if (hipMalloc(&c->tau_scr[q], c->box_elems * sizeof(double)) != hipSuccess) return mfail(c, SS_ERR_NOMEM, "scratch");
        if (hipMemsetAsync(c->tau_scr[q], 0, c->box_elems * sizeof(double), c->stream) != hipSuccess)
            return mfail(c, SS_ERR_NOMEM, "scratch");
    }
    return SS_OK;
}

// FP32 boxes of the correction solve: two sets of (VR,Vth,Vph,P,tau x6),
// eta, the forcing (gP, fVR, fVth, fVph) and the FP32 metric tables.
int ensure_f32(ss_ctx *c)
{
    if (c->f32_ready) return SS_OK;
    const size_t n = c->box_elems;
    for (int q = 0; q < SS_F32_NBOX; ++q) {
        if (c->pair_state && q >= 10 && q < 20) {
            c->f32[q] = c->pair_base32[q - 10] + n; // upper half of the pair (set A | set B)
        } else if (c->pair_state && q < 10) {
            if (hipMalloc(&c->pair_base32[q], 2 * n * sizeof(float)) != hipSuccess)
                return mfail(c, SS_ERR_NOMEM, "FP32 boxes");
            c->f32[q] = c->pair_base32[q];
        } else if (hipMalloc(&c->f32[q], n * sizeof(float)) != hipSuccess) {
            return mfail(c, SS_ERR_NOMEM, "FP32 boxes");
        }
        hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, c->stream, c->f32[q], n);
    }
    c->f32_ready = true;
    return SS_OK;
}

void tau_of(const FieldSet &s, double *tau[6])
{
    tau[0] = s.f[SS_TAU_RR]; tau[1] = s.f[SS_TAU_THTH]; tau[2] = s.f[SS_TAU_PHPH];
    tau[3] = s.f[SS_TAU_RTH]; tau[4] = s.f[SS_TAU_RPH]; tau[5] = s.f[SS_TAU_THPH];
}

// Residual norms from the reduction slots of launch_residual.
void residual_norms(ss_ctx *c, double *r, double *r_div, double *r_mom, const ss_params &p)
{
    const unsigned long long *h = c->red_host;
    const double mdiv = ubits(h[0]);
    const double mf = std::fmax(std::fmax(ubits(h[1]), ubits(h[2])), ubits(h[3]));
    const double mv = std::fmax(std::fmax(ubits(h[4]), ubits(h[5])), ubits(h[6]));
    const double mrg = ubits(h[7]);
    const double mtau = std::fmax(std::fmax(std::fmax(ubits(h[8]), ubits(h[9])), ubits(h[10])), ubits(h[11]));
    // momentum residual relative to the body force (or, without one, to the
    // stress divided by the radial extent); divergence relative to |V|/lr
    const double fscale = mrg > 0.0 ? mrg : (mtau > 0.0 ? mtau / p.lr : 1.0);
    *r_mom = mf / fscale;
    *r_div = mv > 0.0 ? mdiv * p.lr / mv : mdiv;
    *r = std::fmax(*r_mom, *r_div);
}

int eval_residual(ss_ctx *c, const ss_params &p, double *const tau[6], float *gP, float *fV[3], double *r, double *rdiv,
                  double *rmom)
{
    launch_consistent_stress(c, c->cur, tau, p._dr, p._dth, p._dph);
    launch_residual(c, c->cur, tau, gP, fV, p._dr, p._dth, p._dph, p.g);
    hipError_t e = hipMemcpyAsync(c->red_host, c->red, 12 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return mfail(c, (int)e, "residual readback");
    int rc = msync(c, "residual");
    if (rc != SS_OK) return rc;
    residual_norms(c, r, rdiv, rmom, p);
    return SS_OK;
}

using clk = std::chrono::steady_clock;

} // namespace

extern "C" {

int ss_residual(ss_ctx *c, const ss_params *p, double out[3])
{
    if (!c || !p || !out) return SS_ERR_ARG;
    if (!c->grid_set) return mfail(c, SS_ERR_STATE, "ss_set_grid was not called");
    if (hipSetDevice(c->device) != hipSuccess) return mfail(c, SS_ERR_STATE, "hipSetDevice");
    int rc = ensure_tau_scratch(c);
    if (rc != SS_OK) return rc;
    return eval_residual(c, *p, c->tau_scr, nullptr, nullptr, &out[0], &out[1], &out[2]);
}

int ss_pt_solve_residual(ss_ctx *c, const ss_params *p, int maxiter, int ncheck, double tol, ss_result *out)
{
    if (!c || !p || !out || maxiter < 0 || ncheck < 1) return SS_ERR_ARG;
    if (!c->grid_set) return mfail(c, SS_ERR_STATE, "ss_set_grid was not called");
    if (hipSetDevice(c->device) != hipSuccess) return mfail(c, SS_ERR_STATE, "hipSetDevice");
    int rc = ensure_tau_scratch(c);
    if (rc != SS_OK) return rc;
    memset(out, 0, sizeof *out);
    out->err = NAN;
    const clk::time_point t0 = clk::now();
    int done = 0;
    while (done < maxiter) {
        const int n = std::min(ncheck, maxiter - done);
        if ((rc = ss_pt_iterate(c, p, n)) != SS_OK) return rc;
        done += n;
        double r, rd, rm;
        if ((rc = eval_residual(c, *p, c->tau_scr, nullptr, nullptr, &r, &rd, &rm)) != SS_OK) return rc;
        out->err = r;
        out->err_comp[0] = rd;
        out->err_comp[1] = rm;
        if (std::isnan(r)) { out->ittot = done; return mfail(c, SS_ERR_NONFINITE, "NaN residual"); }
        if (r <= tol) { out->iter_end = done; break; }
    }
    out->ittot = done;
    out->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return out->iter_end ? SS_OK : SS_NOT_CONVERGED;
}

int ss_pt_solve_mixed(ss_ctx *c, const ss_params *p, const ss_mixed_opts *o, ss_mixed_result *res)
{
    if (!c || !p || !o || !res) return SS_ERR_ARG;
    if (!c->grid_set) return mfail(c, SS_ERR_STATE, "ss_set_grid was not called");
    if (o->max_outer < 1 || o->max_outer > SS_MIXED_MAX_OUTER || o->max_inner < 1 || o->nout_inner < 1)
        return mfail(c, SS_ERR_ARG, "bad mixed-solve options");
    if (p->etasm_spread < 0 || p->etasm_spread > SS_SPREAD_MAX)
        return mfail(c, SS_ERR_ARG, "etasm_spread must be in [0, 8]");
    if (hipSetDevice(c->device) != hipSuccess) return mfail(c, SS_ERR_STATE, "hipSetDevice");
    int rc = ensure_f32(c);
    if (rc != SS_OK) return rc;
    memset(res, 0, sizeof *res);
    const Box &b = c->box;
    const clk::time_point t0 = clk::now();
    float *const *F = c->f32;
    float *fV[3] = {F[SS_F32_FVR], F[SS_F32_FVT], F[SS_F32_FVP]};
    const bool nonlinear = p->n_exp != 1.0;
    // the FP64 iterate satisfies the (inhomogeneous) BCs; corrections satisfy the homogeneous ones
    launch_bc_R(c, c->cur);
    launch_bc_th(c, c->cur, p->bc_mode, p->r, p->sh);
    launch_bc_ph(c, c->cur, p->bc_mode, p->r, p->sh);
    double *taus[6];
    tau_of(c->cur, taus);
    // FP32 correction arguments (set A = F[0..9], set B = F[10..19])
    Args<float> a{};
    a.b = fused::BoxG{b.nr, b.nt, b.np, b.sj, b.si};
    a.t = fused::TabsG<float>{c->ftr32, c->ftt32, c->ft232, c->th32, c->ph32};
    a.eta = F[SS_F32_ETA];
    a.gP = F[SS_F32_GP];
    a.fVR = fV[0];
    a.fVT = fV[1];
    a.fVP = fV[2];
    a.red = c->red;
    const double PI = 3.14159265358979323846;
    a._dr = (float)p->_dr; a._dth = (float)p->_dth; a._dph = (float)p->_dph;
    a.kb = (float)((((p->dmp * 2.0) * PI) * p->vpdt) / p->lr);
    a.gs = (float)(((4.0 * PI) * p->vpdt) / p->lr);
    a.vpdt2 = (float)(p->vpdt * p->vpdt);
    a.g = 0.f; a.tC = 1.f; a.ex = 0.f; a.relax = 1.f; a.emin = 0.f; a.emax = INFINITY;
    a.bc_mode = p->bc_mode; a.r = (float)p->r; a.sh = 0.f;
    // The residual reduction of an outer step scales with the inner
    // tolerance it used (the error left in the correction).  When the
    // reduction still needed is smaller than the last step's, the inner
    // tolerance is loosened in proportion (with a factor 2 of margin); it is
    // never tightened beyond o->tol_inner.
    double tin_prev = o->tol_inner, r_prev = NAN;
    for (int outer = 0; outer < o->max_outer; ++outer) {
        double r, rd, rm;
        // tau := consistent stress (in place), residual -> FP32 forcing
        if ((rc = eval_residual(c, *p, taus, F[SS_F32_GP], fV, &r, &rd, &rm)) != SS_OK) return rc;
        res->r_hist[outer] = r;
        res->outer = outer;
        res->r = r;
        if (std::isnan(r)) return mfail(c, SS_ERR_NONFINITE, "NaN residual");
        if (r <= o->tol) { res->converged = 1; break; }
        double tin = o->tol_inner;
        if (std::isfinite(r_prev) && r_prev > 0.0 && r < r_prev) {
            const double q = r / r_prev, want = 0.5 * (o->tol / r);
            if (want > q) tin = std::fmin(1e-2, tin_prev * (want / q));
            tin = std::fmax(tin, o->tol_inner);
        }
        r_prev = r;
        tin_prev = tin;
        // frozen FP32 viscosity; zero correction (both buffer sets)
        hipLaunchKernelGGL(k_to_f32, mgrid(b.nr, b.nt, b.np), MB, 0, c->stream, b, F[SS_F32_ETA], c->cur.f[SS_ETA], b.nr,
                           b.nt, b.np);
        // spread etaSM of the frozen viscosity, once for all sweeps of the step
        const bool wide = p->etasm_spread > 0;
        if (wide) launch_etasm_spread_box<float>(c, a.b, F[SS_F32_ETA], F[SS_F32_SM], F[SS_F32_TMP], p->etasm_spread);
        a.SM = wide ? F[SS_F32_SM] : nullptr;
        // the first sweep reads set A (= 0) and rewrites every point of set B
        for (int q = 0; q < 10; ++q)
            hipLaunchKernelGGL(k_zero_f32, dim3(2048), dim3(256), 0, c->stream, F[q], c->box_elems);
        int cur = 0, it = 0;
        double err_in = NAN;
        const clk::time_point ti0 = clk::now(); // eval_residual ended with a synchronisation
        for (it = 1; it <= o->max_inner; ++it) {
            float *const *in = F + 10 * cur, *const *out = F + 10 * (1 - cur);
            a.VR = in[0]; a.VT = in[1]; a.VP = in[2]; a.P = in[3];
            a.tRR = in[4]; a.tTT = in[5]; a.tPP = in[6]; a.tRT = in[7]; a.tRP = in[8]; a.tTP = in[9];
            a.oVR = out[0]; a.oVT = out[1]; a.oVP = out[2]; a.oP = out[3];
            a.oRR = out[4]; a.oTT = out[5]; a.oPP = out[6]; a.oRT = out[7]; a.oRP = out[8]; a.oTP = out[9];
            const bool check = it % o->nout_inner == 0;
            a.flags = fused::F_CORR | (check ? fused::F_CHECK : 0) | (wide ? fused::F_SMBUF : 0);
            launch_fused_args<float>(c, a, nullptr, nullptr);
            cur = 1 - cur;
            if (check) {
                hipError_t e = hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                              c->stream);
                if (e != hipSuccess) return mfail(c, (int)e, "inner readback");
                if ((rc = msync(c, "inner sweep")) != SS_OK) return rc;
                double em = 0.0;
                for (int q = 0; q < 3; ++q) {
                    const double dv = ubits(c->red_host[q]), vv = ubits(c->red_host[3 + q]);
                    if (vv > 0.0) em = std::fmax(em, dv / vv);
                }
                err_in = em;
                if (err_in < tin) break;
            }
        }
        if ((rc = msync(c, "inner sweeps")) != SS_OK) return rc;
        res->inner_seconds += std::chrono::duration<double>(clk::now() - ti0).count();
        res->inner_hist[outer] = std::min(it, o->max_inner);
        res->inner_total += res->inner_hist[outer];
        res->err_inner_hist[outer] = err_in;
        // FP64 accumulation of the correction
        float *const *d = F + 10 * cur;
        double *dst[10] = {c->cur.f[SS_VR], c->cur.f[SS_VTH], c->cur.f[SS_VPH], c->cur.f[SS_P],
                           c->cur.f[SS_TAU_RR], c->cur.f[SS_TAU_THTH], c->cur.f[SS_TAU_PHPH],
                           c->cur.f[SS_TAU_RTH], c->cur.f[SS_TAU_RPH], c->cur.f[SS_TAU_THPH]};
        const int fid[10] = {SS_VR, SS_VTH, SS_VPH, SS_P, SS_TAU_RR, SS_TAU_THTH, SS_TAU_PHPH, SS_TAU_RTH, SS_TAU_RPH,
                             SS_TAU_THPH};
        for (int q = 0; q < 10; ++q) {
            int dm[3], o1;
            ss_field_dims(c, fid[q], dm);
            o1 = (fid[q] == SS_TAU_RTH || fid[q] == SS_TAU_RPH || fid[q] == SS_TAU_THPH) ? 1 : 0;
            hipLaunchKernelGGL(k_accumulate, mgrid(dm[0], dm[1], dm[2]), MB, 0, c->stream, b, dst[q], d[q], dm[0], dm[1],
                               dm[2], o1);
        }
        if (nonlinear) { // eta from the accumulated stresses (FP64)
            launch_compute_TII(c, c->cur);
            launch_copy_bc(c, c->cur.f[SS_TII], b.nr, b.nt, b.np);
            launch_power_law(c, c->cur, p->tau_C, p->n_exp, p->relax, p->eta_min, p->eta_max);
        }
        res->outer = outer + 1;
    }
    if ((rc = msync(c, "mixed solve")) != SS_OK) return rc;
    if (!res->converged) { // residual of the final state
        double r, rd, rm;
        if ((rc = eval_residual(c, *p, taus, F[SS_F32_GP], fV, &r, &rd, &rm)) != SS_OK) return rc;
        res->r = r;
        if (res->outer < SS_MIXED_MAX_OUTER) res->r_hist[res->outer] = r;
        res->converged = r <= o->tol;
    }
    res->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return res->converged ? SS_OK : SS_NOT_CONVERGED;
}

} // extern "C"
