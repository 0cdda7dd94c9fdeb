// fused_emu.cpp -- host emulation of the fused PT iteration (test support).
//
// Compiles the device phase code of sphericalstokes_amd/csrc/sspt_fused_body.h
// with the host compiler and runs it workgroup by workgroup: every thread of a
// block executes a phase before any thread starts the next one, which is what
// the barriers between phases guarantee on the GPU.  Tests use it to check the
// fused schedule (tiling, halos, r-march, carries, boundary pass) against the
// CPU oracle without a GPU.  It is never part of the product library.
#define SSPT_DH
#ifndef SSPT_ACCESS_REGISTER
#define SSPT_ACCESS_REGISTER(p, n) ((void)0)
#define SSPT_ACCESS_UNREGISTER(p) ((void)0)
#endif
#define SSPT_UNIFORM(x) (x)
#define SSPT_CONSTANT_PTR(T, p) (p)
#include "../../sphericalstokes_amd/csrc/sspt_fused_body.h"
#include "../../sphericalstokes_amd/csrc/sspt_tables.h"
#include "../../include/sspt.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

using namespace sspt::fused;

namespace {

struct HostBox {
    int nr, nt, np;
    long long sj, si;
};

template <typename R, int FL>
void run_main_t(const Args<R> &a, int rchunk, unsigned long long red[6])
{
    const int nx = (a.b.np + OKX<R> - 1) / OKX<R>;
    const int ny = (a.b.nt + OJ<R> - 1) / OJ<R>;
    const int nz = (a.b.nr + rchunk - 1) / rchunk;
    std::unique_ptr<Lds<R>> L(new Lds<R>());
    std::vector<Thread<R>> th(TKX<R> * TJB<R>);
    for (int bz = 0; bz < nz; ++bz)
        for (int by = 0; by < ny; ++by)
            for (int bx = 0; bx < nx; ++bx) {
                memset(L.get(), 0xff, sizeof(Lds<R>)); // NaN garbage: unset LDS reads must not reach outputs
                const int kb = bx * OKX<R>, jb = by * OJ<R>;
                const int s0 = bz * rchunk, s1 = std::min(s0 + rchunk, a.b.nr);
                const int start = std::max(s0 - 2, -1);
                for (int ty = 0; ty < TJB<R>; ++ty)
                    for (int tx = 0; tx < TKX<R>; ++tx) {
                        Thread<R> &T = th[ty * TKX<R> + tx];
                        thread_init(T, a, tx, ty, kb, jb);
                        thread_prologue<R, FL>(T, a, start);
                    }
                for (int s = start; s < s1; ++s) {
                    const bool wr = s >= s0;
                    for (auto &T : th) phase1<R, FL>(T, *L, a, s);
                    for (auto &T : th) phase2<R, FL>(T, *L, a, s, wr);
                    for (auto &T : th) phase3<R, FL>(T, *L, a, s, wr);
                    for (auto &T : th) phase4<R, FL>(T, *L, a, s, wr);
                }
                for (auto &T : th)
                    for (int c = 0; c < 3; ++c) {
                        red[c] = std::max(red[c], T.mdv[c]);
                        red[3 + c] = std::max(red[3 + c], T.mv[c]);
                    }
            }
}

template <typename R> void run_main(const Args<R> &a, int rchunk, unsigned long long red[6])
{
    switch (a.flags) {
#define SSPT_FL(v) case v: run_main_t<R, v>(a, rchunk, red); break;
    SSPT_FL(0) SSPT_FL(1) SSPT_FL(3) SSPT_FL(4) SSPT_FL(5) SSPT_FL(7)
    SSPT_FL(8) SSPT_FL(9) SSPT_FL(11) SSPT_FL(12) SSPT_FL(13) SSPT_FL(15) SSPT_FL(16) SSPT_FL(18)
    SSPT_FL(32) SSPT_FL(33) SSPT_FL(35) SSPT_FL(36) SSPT_FL(37) SSPT_FL(39)
    SSPT_FL(40) SSPT_FL(41) SSPT_FL(43) SSPT_FL(44) SSPT_FL(45) SSPT_FL(47)
    SSPT_FL(64) SSPT_FL(65) SSPT_FL(67) SSPT_FL(68) SSPT_FL(69) SSPT_FL(71)
    SSPT_FL(72) SSPT_FL(73) SSPT_FL(75) SSPT_FL(76) SSPT_FL(77) SSPT_FL(79)
    SSPT_FL(96) SSPT_FL(97) SSPT_FL(99) SSPT_FL(100) SSPT_FL(101) SSPT_FL(103)
    SSPT_FL(104) SSPT_FL(105) SSPT_FL(107) SSPT_FL(108) SSPT_FL(109) SSPT_FL(111)
#undef SSPT_FL
    default: std::abort(); // a flag combination without an instantiation
    }
}

// The etaSM passes before an F_SMBUF iteration (k_star_bc and k_spread on the
// GPU, same order of buffers): star eta -> tmp, r tmp -> sm, th sm -> tmp,
// phi tmp -> sm.
template <typename R> void run_etasm_spread(const BoxG &b, const R *eta, R *sm, R *tmp, int m)
{
    const SpreadW<R> w = spread_weights<R>();
    for (int i = 0; i < b.nr; ++i)
        for (int j = 0; j < b.nt; ++j)
            for (int k = 0; k < b.np; ++k) tmp[b.off(i, j, k)] = star_bc_at(b, eta, i, j, k);
    for (int axis = 0; axis < 3; ++axis) {
        const R *src = axis == 1 ? sm : tmp;
        R *dst = axis == 1 ? tmp : sm;
        for (int i = 0; i < b.nr; ++i)
            for (int j = 0; j < b.nt; ++j)
                for (int k = 0; k < b.np; ++k) dst[b.off(i, j, k)] = spread_at(b, src, i, j, k, axis, m, w);
    }
}

template <typename R> void run_boundary(const Args<R> &a, unsigned long long red[6])
{
    const int nr = a.b.nr, nt = a.b.nt, np = a.b.np;
    const int m = std::max(std::max(nr, nt), np) + 1;
    for (int fam = 0; fam < 3; ++fam)
        for (int v = 0; v < m; ++v)
            for (int u = 0; u < m; ++u) {
                unsigned long long mdv[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
                for (int c = 0; c < 3; ++c) {
                    const int n0 = c == 0 ? nr + 1 : nr, n1 = c == 1 ? nt + 1 : nt, n2 = c == 2 ? np + 1 : np;
                    if (fam == 0 && u < n2 && v < n1) {
                        ghost_v_point(a, c, 0, v, u, &mdv[c], &mv[c]);
                        ghost_v_point(a, c, n0 - 1, v, u, &mdv[c], &mv[c]);
                    } else if (fam == 1 && u < n2 && v < n0) {
                        ghost_v_point(a, c, v, 0, u, &mdv[c], &mv[c]);
                        ghost_v_point(a, c, v, n1 - 1, u, &mdv[c], &mv[c]);
                    } else if (fam == 2 && u < n1 && v < n0) {
                        ghost_v_point(a, c, v, u, 0, &mdv[c], &mv[c]);
                        ghost_v_point(a, c, v, u, n2 - 1, &mdv[c], &mv[c]);
                    }
                }
                if (a.flags & F_CORR) {
                } else if (fam == 0 && u < np && v < nt) {
                    boundary_eta_point(a, 0, v, u);
                    boundary_eta_point(a, nr - 1, v, u);
                } else if (fam == 1 && u < np && v >= 1 && v < nr - 1) {
                    boundary_eta_point(a, v, 0, u);
                    boundary_eta_point(a, v, nt - 1, u);
                } else if (fam == 2 && u >= 1 && u < nt - 1 && v >= 1 && v < nr - 1) {
                    boundary_eta_point(a, v, u, 0, 0);
                    boundary_eta_point(a, v, u, np - 1, 1);
                }
                for (int c = 0; c < 3; ++c) {
                    red[c] = std::max(red[c], mdv[c]);
                    red[3 + c] = std::max(red[3 + c], mv[c]);
                }
            }
}

} // namespace

extern "C" {

// x^ex through the kernels' pow_ex for n values (tests of its rounding)
void emu_pow_ex(const double *x, int n, double ex, double *out)
{
    for (int q = 0; q < n; ++q) out[q] = pow_ex(x[q], ex);
}

// Runs `niter` fused iterations on box-layout buffers (shape (nr+1) x (nth+1)
// x pitch, phi fastest).  cur/alt: SS_NFIELDS pointers each; non-state fields
// must alias between the two sets.  The last iteration writes the auxiliary
// fields; with check_last it is also a convergence-check iteration whose
// max-reductions land in red[6].  Returns 0 if the final state is in `cur`,
// 1 if it is in `alt`.  pair_bytes > 0: the state buffers of cur and alt are
// halves of paired allocations pair_bytes apart and the main kernel runs
// with F_PAIRED (as the library does); the boundary pass keeps the plain
// pointers.
int emu_fused_iterations(int nr, int nt, int np, int pitch, const double *R1, const double *th1, const double *ph1,
                         const ss_params *p, double **cur, double **alt, int niter, int check_last, int rchunk,
                         unsigned long long red[6], long long pair_bytes)
{
    HostBox hb{nr, nt, np, pitch, (long long)(nt + 1) * pitch};
    std::vector<double> tab, ftab;
    sspt::build_tables(hb, R1, th1, ph1, tab);
    sspt::build_fused_tables(hb, tab, ftab);
    const size_t n1 = nr + 1, n2 = nt + 1;
    TabsG<double> t;
    t.tr = ftab.data();
    t.tt = t.tr + sspt::FT_NTR * n1;
    t.t2 = t.tt + sspt::FT_NTT * n2;
    // th and ph follow the 20 separate tables in build_tables' layout
    t.th = tab.data() + 10 * n1 + 4 * n2 + 6 * n1 * n2;
    t.ph = t.th + n2;
    static constexpr double PI = 3.14159265358979323846;
    for (int c = 0; c < 6; ++c) red[c] = 0ull;
    double **in = cur, **out = alt;
    std::vector<double> tbp(2 * (size_t)nr * nt, std::nan(""));
    SSPT_ACCESS_REGISTER(tbp.data(), tbp.size() * sizeof(double));
    int which = 0;
    for (int it = 0; it < niter; ++it) {
        const bool last = it == niter - 1, check = last && check_last;
        Args<double> a{};
        a.b = BoxG{nr, nt, np, hb.sj, hb.si};
        a.t = t;
        a.VR = in[SS_VR]; a.VT = in[SS_VTH]; a.VP = in[SS_VPH]; a.eta = in[SS_ETA];
        a.P = in[SS_P]; a.tRR = in[SS_TAU_RR]; a.tTT = in[SS_TAU_THTH]; a.tPP = in[SS_TAU_PHPH];
        a.tRT = in[SS_TAU_RTH]; a.tRP = in[SS_TAU_RPH]; a.tTP = in[SS_TAU_THPH];
        a.rho = in[SS_RHO]; a.ini = in[SS_ETA_INI]; a.ePL = in[SS_ETA_PL];
        a.oVR = out[SS_VR]; a.oVT = out[SS_VTH]; a.oVP = out[SS_VPH]; a.oEta = out[SS_ETA];
        a.oP = out[SS_P]; a.oRR = out[SS_TAU_RR]; a.oTT = out[SS_TAU_THTH]; a.oPP = out[SS_TAU_PHPH];
        a.oRT = out[SS_TAU_RTH]; a.oRP = out[SS_TAU_RPH]; a.oTP = out[SS_TAU_THPH]; a.oPL = out[SS_ETA_PL];
        a.SM = out[SS_ETASM]; a.DT = out[SS_DT_R]; a.TII = out[SS_TII];
        a.TBP = tbp.data();
        a.VRe = out[SS_VR_ERR]; a.VTe = out[SS_VTH_ERR]; a.VPe = out[SS_VPH_ERR];
        a.red = nullptr;
        a._dr = p->_dr; a._dth = p->_dth; a._dph = p->_dph;
        a.kb = (((p->dmp * 2.0) * PI) * p->vpdt) / p->lr;
        a.gs = ((4.0 * PI) * p->vpdt) / p->lr;
        a.vpdt2 = p->vpdt * p->vpdt;
        a.g = p->g; a.tC = p->tau_C; a.ex = 1.0 - p->n_exp; a.relax = p->relax;
        a.emin = p->eta_min; a.emax = p->eta_max;
        a.flags = (last ? F_LAST : 0) | (check ? F_CHECK : 0) | (p->relax != 1.0 ? F_RELAX : 0) |
                  ((p->eta_min > 0.0 || p->eta_max < INFINITY) ? F_CLAMP : 0) |
                  (p->etasm_spread > 0 ? F_SMBUF : 0);
        a.bc_mode = p->bc_mode; a.r = p->r; a.sh = p->sh;
        a.spread = p->etasm_spread;
        a.vth0 = pure_shear_vth(th1[0], p->r, p->sh);
        a.vth1 = pure_shear_vth(th1[nt - 1], p->r, p->sh);
        a.rchunk = rchunk > 0 ? rchunk : nr;
        Args<double> am = a;
        if (pair_bytes > 0 && !pair_args(am, pair_bytes)) {
            SSPT_ACCESS_UNREGISTER(tbp.data());
            return -1;
        }
        if (a.flags & F_SMBUF) run_etasm_spread<double>(a.b, a.eta, a.SM, a.DT, a.spread);
        run_main(am, a.rchunk, red);
        run_boundary(a, red);
        std::swap(in, out);
        which ^= 1;
    }
    SSPT_ACCESS_UNREGISTER(tbp.data());
    return which;
}

} // extern "C"
