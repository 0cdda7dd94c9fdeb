// adj_emu.cpp -- host emulation of the r-marching adjoint kernel
// (sphericalstokes_amd/csrc/sspt_adj_fused_body.h), test support only.
//
// Runs the same phase functions as the device kernel, looping over the
// threads of a block between the barriers, on box-layout host arrays.  The
// BC fold and the boundary power-law sums that the
// library computes in separate small kernels are restated here as loops.
#define SSPT_DH
#ifndef SSPT_ACCESS_REGISTER
#define SSPT_ACCESS_REGISTER(p, n) ((void)0)
#define SSPT_ACCESS_UNREGISTER(p) ((void)0)
#endif
#define SSPT_UNIFORM(x) (x)
#define SSPT_CONSTANT_PTR(T, p) (p)
#include "../../sphericalstokes_amd/csrc/sspt_adj_fused_body.h"
#include "../../sphericalstokes_amd/csrc/sspt_tables.h"
#include "../../include/sspt.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using namespace sspt::adjf;

namespace {

struct HostBox {
    int nr, nt, np;
    long long sj, si;
};

// transposes of BC_ph!, BC_th!, BC_R! (in this order) on lam_V in place
template <typename R> void bc_fold(const BoxG &b, int bc_mode, R *VR, R *VT, R *VP)
{
    const int nr = b.nr, nt = b.nt, np = b.np;
    const bool cp = bc_mode != SS_BC_SHEAR;
    const long long sj = b.sj, si = b.si;
    for (int i = 0; i <= nr; ++i)
        for (int j = 0; j <= nt; ++j) {
            if (i < nr && j < nt) {
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
    for (int i = 0; i <= nr; ++i)
        for (int k = 0; k <= np; ++k) {
            if (i < nr && k <= np) {
                const long long lo = b.off(i, 0, k), hi = b.off(i, nt - 1, k);
                if (cp) { VP[lo + sj] += VP[lo]; VP[hi - sj] += VP[hi]; }
                VP[lo] = R(0); VP[hi] = R(0);
            }
            if (i < nr && k < np) {
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
    for (int j = 0; j <= nt; ++j)
        for (int k = 0; k <= np; ++k) {
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
            if (j < nt && k < np) {
                const long long lo = b.off(0, j, k), hi = b.off(nr, j, k);
                VR[lo + si] -= VR[lo]; VR[lo] = R(0);
                VR[hi - si] -= VR[hi]; VR[hi] = R(0);
            }
        }
}

// One adjoint step in precision R; with RHS, every output also gets rhs[f]
// at the same element (the correction sweeps of the mixed-precision solve).
template <typename R, bool RHS>
int adj_step(int nr, int nt, int np, int pitch, const double *R1, const double *th1, const double *ph1,
             const ss_params *p, R **state, R **lam_in, R **lam_out, int rchunk, const R *bJ, double obj_r,
             double obj_rad, R *const *rhs)
{
    HostBox hb{nr, nt, np, pitch, (long long)(nt + 1) * pitch};
    const BoxG b{nr, nt, np, hb.sj, hb.si};
    std::vector<double> tab, ftab;
    sspt::build_tables(hb, R1, th1, ph1, tab);
    sspt::build_fused_tables(hb, tab, ftab);
    const size_t n1 = nr + 1, n2 = nt + 1;
    // the kernel's tables in its precision: the fused records, then th, ph
    std::vector<R> tR(ftab.begin(), ftab.end());
    const double *th_d = tab.data() + 10 * n1 + 4 * n2 + 6 * n1 * n2;
    tR.insert(tR.end(), th_d, th_d + n2 + (np + 1));
    AArgs<R> a{};
    a.b = b;
    a.t.tr = tR.data();
    a.t.tt = a.t.tr + sspt::FT_NTR * n1;
    a.t.t2 = a.t.tt + sspt::FT_NTT * n2;
    a.t.th = tR.data() + ftab.size();
    a.t.ph = a.t.th + n2;
    static constexpr double PI = 3.14159265358979323846;
    a._dr = R(p->_dr); a._dth = R(p->_dth); a._dph = R(p->_dph);
    a.kb = R((((p->dmp * 2.0) * PI) * p->vpdt) / p->lr);
    a.gs = R(((4.0 * PI) * p->vpdt) / p->lr);
    a.tC = R(p->tau_C); a.ex = R(1.0 - p->n_exp); a.relax = R(p->relax); a.emin = R(p->eta_min); a.emax = R(p->eta_max);
    a.do_relax = p->relax != 1.0;
    a.do_lo = p->eta_min > 0.0;
    a.do_hi = p->eta_max < INFINITY;
    a.tRR = state[0]; a.tTT = state[1]; a.tPP = state[2]; a.tRT = state[3]; a.tRP = state[4]; a.tTP = state[5];
    a.eta = state[6]; a.ini = state[7]; a.ePL = state[8]; a.tII = state[9]; a.DT = state[10]; a.SM = state[11];
    for (int q = 0; q < A_N; ++q) {
        a.li[q] = lam_in[q];
        a.lo[q] = lam_out[q];
        a.rhs[q] = rhs ? rhs[q] : nullptr;
    }
    bc_fold(b, p->bc_mode, lam_in[A_VR], lam_in[A_VT], lam_in[A_VP]);
    const size_t nbox = (size_t)(nr + 1) * (nt + 1) * pitch;
    std::vector<R> LB(nbox, R(std::nan("")));
    SSPT_ACCESS_REGISTER(LB.data(), LB.size() * sizeof(R));
    for (int i = 1; i <= nr - 2; ++i)
        for (int j = 1; j <= nt - 2; ++j)
            for (int k = 1; k <= np - 2; ++k)
                if (next_to_boundary(b, i, j, k)) LB[b.off(i, j, k)] = lb_sum(a, i, j, k);
    a.LB = LB.data();
    a.bJ = bJ;   // objective weights on the r-faces (box layout) or null
    a.obj_r = obj_r;
    a.obj_rad = obj_rad;
    const int rc = rchunk > 0 ? rchunk : nr + 1;
    a.rchunk = rc;
    const int nx = (np + 1 + AOK - 1) / AOK, ny = (nt + 1 + AOJ - 1) / AOJ, nz = (nr + 1 + rc - 1) / rc;
    std::unique_ptr<ALds<R>> L(new ALds<R>());
    std::vector<AThread<R>> th(TK * ATJ);
    std::vector<AStepA<R>> na(TK * ATJ), ca(TK * ATJ);
    std::vector<AStepB<R>> cb(TK * ATJ);
    for (int bz = 0; bz < nz; ++bz)
        for (int by = 0; by < ny; ++by)
            for (int bx = 0; bx < nx; ++bx) {
                memset(L.get(), 0xff, sizeof(ALds<R>)); // NaN garbage: unset LDS reads must not reach outputs
                const int kb = bx * AOK, jb = by * AOJ;
                const int s0 = bz * rc, s1 = std::min(s0 + rc, nr + 1);
                const int start = std::max(s0 - 1, 0), stop = std::min(s1, nr);
                for (int ty = 0; ty < ATJ; ++ty)
                    for (int tx = 0; tx < TK; ++tx) {
                        AThread<R> &T = th[ty * TK + tx];
                        athread_init(T, a, tx, ty, kb, jb);
                        aprologue(T, *L, a, start);
                        aloadA(na[ty * TK + tx], a, T, start);
                    }
                for (int q = start; q <= stop; ++q) {
                    const bool wr = q >= s0 && q < s1, wr_m = q - 1 >= s0 && q - 1 < s1;
                    for (size_t t = 0; t < th.size(); ++t) {
                        ca[t] = na[t];
                        aloadB(cb[t], a, th[t], q);
                    }
                    // phase A reads the loaded values of the lane next door
                    for (size_t t = 0; t < th.size(); ++t) aphaseA<R, RHS>(th[t], *L, a, ca[t], q, wr);
                    for (size_t t = 0; t < th.size(); ++t) {
                        aloadA(na[t], a, th[t], q + 1);
                        aphaseB<R, RHS>(th[t], *L, a, cb[t], q, wr);
                    }
                    for (auto &T : th) aphaseC<R, RHS>(T, *L, a, q, wr, wr_m);
                }
            }
    SSPT_ACCESS_UNREGISTER(LB.data());
    return 0;
}

} // namespace

extern "C" {

// One adjoint step lam_out = T^T lam_in on box-layout buffers (shape
// (nr+1) x (nth+1) x pitch, phi fastest).  state: tau_RR tau_thth tau_phph
// tau_Rth tau_Rph tau_thph eta eta_ini eta_PL tII DT_R etaSM; lam_in is
// BC-folded in place; lam_out must be zero-initialised.  With bJ (box layout,
// the objective weights: zero outside the inclusion) the step adds bJ on the
// r-faces that the objective test (centre obj_r, squared radius obj_rad) finds
// inside.
int emu_adj_step(int nr, int nt, int np, int pitch, const double *R1, const double *th1, const double *ph1,
                 const ss_params *p, double **state, double **lam_in, double **lam_out, int rchunk,
                 const double *bJ, double obj_r, double obj_rad)
{
    return adj_step<double, false>(nr, nt, np, pitch, R1, th1, ph1, p, state, lam_in, lam_out, rchunk, bJ, obj_r,
                                   obj_rad, nullptr);
}

// lam_out = T^T lam_in + rhs (every output), in FP64 (f32 = 0: the arrays
// are double) or FP32 (f32 = 1: state, lam and rhs are float arrays)
int emu_adj_step_rhs(int f32, int nr, int nt, int np, int pitch, const double *R1, const double *th1,
                     const double *ph1, const ss_params *p, void **state, void **lam_in, void **lam_out, int rchunk,
                     void **rhs)
{
    if (f32)
        return adj_step<float, true>(nr, nt, np, pitch, R1, th1, ph1, p, (float **)state, (float **)lam_in,
                                     (float **)lam_out, rchunk, nullptr, 0.0, 0.0, (float *const *)rhs);
    return adj_step<double, true>(nr, nt, np, pitch, R1, th1, ph1, p, (double **)state, (double **)lam_in,
                                  (double **)lam_out, rchunk, nullptr, 0.0, 0.0, (double *const *)rhs);
}

} // extern "C"
