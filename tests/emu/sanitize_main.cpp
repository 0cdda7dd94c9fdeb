// sanitize_main.cpp -- the CPU oracle and the host emulations of the fused
// forward and r-marching adjoint kernels under AddressSanitizer and
// UndefinedBehaviorSanitizer (test support only; built and run by
// tests/test_sanitize.py, never on a GPU).
//
// Every field box is a separate allocation of exactly (nr+1) x (nth+1) x
// pitch doubles (plus the paired form: two boxes in one allocation), so an
// access past a box is a heap overflow for ASan; in addition every access
// through the kernels' 32-bit byte offsets (fat/fat_pp) is checked against
// the box it is based on (SSPT_ACCESS_CHECK).  The shapes cover the tiling
// edges of the fused kernels (phi around the 62 owned lanes of a 64-lane
// row, th around the owned rows of a block, short and chunked r-marches).
// For each shape the emulated forward iterations are compared with the
// oracle bit for bit (n_exp = 1) or to 1e-12 (power law), and one emulated
// adjoint step with the oracle's adjoint to 1e-12.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>

static std::map<const char *, uint64_t> g_regions;
static long long g_checks = 0;

extern "C" void sspt_access_check(const void *p, unsigned long long off, unsigned long long n)
{
    ++g_checks;
    const char *q = static_cast<const char *>(p) + off;
    auto it = g_regions.upper_bound(static_cast<const char *>(p));
    if (it != g_regions.begin()) {
        --it;
        if (static_cast<const char *>(p) < it->first + it->second && q >= it->first && q + n <= it->first + it->second)
            return;
    }
    std::fprintf(stderr, "out-of-box access: base %p byte offset %llu\n", p, off);
    std::abort();
}
extern "C" void sspt_access_register(const void *p, unsigned long long bytes)
{
    g_regions[static_cast<const char *>(p)] = bytes;
}
extern "C" void sspt_access_unregister(const void *p) { g_regions.erase(static_cast<const char *>(p)); }

#include "../../include/sspt.h"
#include "../../oracle/sspt_adjoint.h"
#include "../../oracle/sspt_oracle.h"

// the host emulations (fused_emu.cpp, adj_emu.cpp)
extern "C" int emu_fused_iterations(int nr, int nt, int np, int pitch, const double *R1, const double *th1,
                                    const double *ph1, const ss_params *p, double **cur, double **alt, int niter,
                                    int check_last, int rchunk, unsigned long long red[6], long long pair_bytes);
extern "C" int emu_adj_step(int nr, int nt, int np, int pitch, const double *R1, const double *th1, const double *ph1,
                            const ss_params *p, double **state, double **lam_in, double **lam_out, int rchunk,
                            const double *bJ, double obj_r, double obj_rad);
extern "C" int emu_adj_step_rhs(int f32, int nr, int nt, int np, int pitch, const double *R1, const double *th1,
                                const double *ph1, const ss_params *p, void **state, void **lam_in, void **lam_out,
                                int rchunk, void **rhs);
#include <cmath>
#include <cstring>
#include <vector>

namespace {

const int NF = SS_NFIELDS;

struct Shape { int n1, n2, n3, off; };

Shape shape_of(int f, int nr, int nt, int np)
{
    switch (f) {
    case SS_VR: case SS_VR_ERR: return {nr + 1, nt, np, 0};
    case SS_VTH: case SS_VTH_ERR: return {nr, nt + 1, np, 0};
    case SS_VPH: case SS_VPH_ERR: return {nr, nt, np + 1, 0};
    case SS_TAU_RTH: return {nr - 1, nt - 1, np - 2, 1};
    case SS_TAU_RPH: return {nr - 1, nt - 2, np - 1, 1};
    case SS_TAU_THPH: return {nr - 2, nt - 1, np - 1, 1};
    default: return {nr, nt, np, 0};
    }
}

bool is_state(int f)
{
    return f == SS_P || f == SS_TAU_RR || f == SS_TAU_THTH || f == SS_TAU_PHPH || f == SS_TAU_RTH ||
           f == SS_TAU_RPH || f == SS_TAU_THPH || f == SS_VR || f == SS_VTH || f == SS_VPH || f == SS_ETA;
}

uint64_t g_rng = 88172645463325252ull;
double urand() // xorshift64, uniform in [-1, 1)
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

double *alloc_region(size_t n)
{
    double *p = new double[n]();
    g_regions[reinterpret_cast<const char *>(p)] = n * sizeof(double);
    return p;
}
void free_region(double *p)
{
    g_regions.erase(reinterpret_cast<const char *>(p));
    delete[] p;
}
float *alloc_region_f(size_t n)
{
    float *p = new float[n]();
    g_regions[reinterpret_cast<const char *>(p)] = n * sizeof(float);
    return p;
}
void free_region_f(float *p)
{
    g_regions.erase(reinterpret_cast<const char *>(p));
    delete[] p;
}

struct Case {
    int nr, nt, np, bc_mode;
    double n_exp, relax, emin, emax;
    int rchunk, paired;
    int spread = 0;
};

double worst_rel(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0.0, den = 0.0;
    for (size_t q = 0; q < a.size(); ++q) {
        num += (a[q] - b[q]) * (a[q] - b[q]);
        den += b[q] * b[q];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

int run_case(const Case &c)
{
    const int nr = c.nr, nt = c.nt, np = c.np, niter = 3;
    const double PI = 3.14159265358979323846;
    std::vector<double> R(nr), th(nt), ph(np);
    for (int i = 0; i < nr; ++i) R[i] = 1.0 + (double)i / (nr - 1);
    for (int j = 0; j < nt; ++j) th[j] = PI / 3.0 + (PI / 3.0) * j / (nt - 1);
    for (int k = 0; k < np; ++k) ph[k] = -PI / 6.0 + (PI / 3.0) * k / (np - 1);
    // reference-layout fields
    std::vector<std::vector<double>> F(NF);
    for (int f = 0; f < NF; ++f) {
        const Shape s = shape_of(f, nr, nt, np);
        F[f].resize((size_t)s.n1 * s.n2 * s.n3);
        for (double &v : F[f]) {
            if (f == SS_ETA || f == SS_ETA_INI || f == SS_ETA_PL || f == SS_ETASM) v = std::exp(urand());
            else if (f == SS_DT_R) v = 1e-5 * (1.5 + urand());
            else if (f == SS_TII) v = 1e-2 * (1.5 + urand());
            else v = 1e-2 * urand();
        }
    }
    ss_params p{};
    const double lr = 1.0, lth = PI / 3.0, lph = PI / 3.0;
    p._dr = (nr - 1) / lr; p._dth = (nt - 1) / lth; p._dph = (np - 1) / lph;
    p.dmp = 4.5; p.vpdt = (lr / (nr - 1)) / (2.0 + 4.5 * 3.0); p.lr = lr;
    p.tau_C = 1.0; p.n_exp = c.n_exp; p.g = 1.0; p.r = 1.5; p.sh = 0.7;
    p.relax = c.relax; p.eta_min = c.emin; p.eta_max = c.emax; p.bc_mode = c.bc_mode;
    p.etasm_spread = c.spread;
    // oracle
    std::vector<std::vector<double>> O = F;
    orc_state os{};
    os.nr = nr; os.nth = nt; os.nph = np;
    double **op[NF] = {&os.P, &os.DT_R, &os.tRR, &os.tTT, &os.tPP, &os.tRT, &os.tRP, &os.tTP, &os.VR, &os.VT,
                       &os.VP, &os.VR_err, &os.VT_err, &os.VP_err, &os.eta, &os.eta_ini, &os.etaSM, &os.tII,
                       &os.rho, &os.eta_PL};
    for (int f = 0; f < NF; ++f) *op[f] = O[f].data();
    os.R1 = R.data(); os.th1 = th.data(); os.ph1 = ph.data();
    orc_params pp{p._dr, p._dth, p._dph, p.dmp, p.vpdt, p.lr, p.tau_C, p.n_exp, p.g, p.r, p.sh, p.relax,
                  p.eta_min, p.eta_max, p.bc_mode, 0, c.spread};
    for (int it = 0; it < niter; ++it) orc_pt_step(&os, &pp, 1);
    // emulation on exact-size box allocations
    const int pitch = ((np + 1) + 63) / 64 * 64;
    const size_t box = (size_t)(nr + 1) * (nt + 1) * pitch;
    double *cur[NF], *alt[NF];
    std::vector<double *> owned;
    for (int f = 0; f < NF; ++f) {
        if (c.paired && is_state(f)) {
            double *b = alloc_region(2 * box);
            cur[f] = b;
            alt[f] = b + box;
            owned.push_back(b);
        } else {
            cur[f] = alloc_region(box);
            owned.push_back(cur[f]);
            if (is_state(f)) {
                alt[f] = alloc_region(box);
                owned.push_back(alt[f]);
            } else {
                alt[f] = cur[f];
            }
        }
        const Shape s = shape_of(f, nr, nt, np);
        for (int k = 0; k < s.n3; ++k)
            for (int j = 0; j < s.n2; ++j)
                for (int i = 0; i < s.n1; ++i)
                    cur[f][(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)] =
                        F[f][((size_t)k * s.n2 + j) * s.n1 + i];
    }
    unsigned long long red[6];
    const int which = emu_fused_iterations(nr, nt, np, pitch, R.data(), th.data(), ph.data(), &p, cur, alt, niter, 1,
                                           c.rchunk, red, c.paired ? (long long)(box * sizeof(double)) : 0);
    if (which < 0) { std::fprintf(stderr, "pairing failed\n"); return 1; }
    double **res = which == 0 ? cur : alt;
    int bad = 0;
    const bool exact = c.n_exp == 1.0;
    for (int f = 0; f < NF; ++f) {
        const Shape s = shape_of(f, nr, nt, np);
        std::vector<double> g(F[f].size());
        for (int k = 0; k < s.n3; ++k)
            for (int j = 0; j < s.n2; ++j)
                for (int i = 0; i < s.n1; ++i)
                    g[((size_t)k * s.n2 + j) * s.n1 + i] =
                        res[f][(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)];
        // V_err: the emulation's check iteration holds V_new - V_old, as check_err! leaves it
        if (f == SS_VR_ERR || f == SS_VTH_ERR || f == SS_VPH_ERR) continue;
        const bool eq = std::memcmp(g.data(), O[f].data(), g.size() * sizeof(double)) == 0;
        const double e = worst_rel(g, O[f]);
        if ((exact && !eq) || !(e <= 1e-12)) {
            std::fprintf(stderr, "forward mismatch: case %dx%dx%d field %d rel %.3e\n", nr, nt, np, f, e);
            ++bad;
        }
    }
    // one adjoint step at the iterated state
    const int NA = ORC_A_N;
    std::vector<std::vector<double>> lam(NA);
    const int adj_field[NA] = {SS_VR, SS_VTH, SS_VPH, SS_P, SS_TAU_RR, SS_TAU_THTH, SS_TAU_PHPH, SS_TAU_RTH,
                               SS_TAU_RPH, SS_TAU_THPH, SS_ETA, SS_ETA_PL};
    orc_adj oa{};
    for (int q = 0; q < NA; ++q) {
        lam[q].resize(O[adj_field[q]].size());
        for (double &v : lam[q]) v = urand();
        if (q == ORC_A_EPL && c.relax == 1.0)
            for (double &v : lam[q]) v = 0.0;
    }
    std::vector<std::vector<double>> lam_ref = lam;
    for (int q = 0; q < NA; ++q) oa.a[q] = lam_ref[q].data();
    orc_adj_step(&os, &pp, &oa, nullptr, nullptr);
    // with the objective (mean VR in the sphere of squared radius 0.01 at r = 1.5):
    // its weights are added on the r-faces inside
    std::vector<double> dJ(O[SS_VR].size(), 0.0);
    (void)orc_adj_objective(&os, 1.5, 0.01, dJ.data());
    for (size_t q = 0; q < dJ.size(); ++q) lam_ref[ORC_A_VR][q] += dJ[q];
    const int st_field[12] = {SS_TAU_RR, SS_TAU_THTH, SS_TAU_PHPH, SS_TAU_RTH, SS_TAU_RPH, SS_TAU_THPH,
                              SS_ETA, SS_ETA_INI, SS_ETA_PL, SS_TII, SS_DT_R, SS_ETASM};
    double *sb[12], *li[NA], *lo[NA];
    std::vector<double *> owned2;
    auto to_box = [&](const std::vector<double> &src, int f) {
        double *b = alloc_region(box);
        owned2.push_back(b);
        const Shape s = shape_of(f, nr, nt, np);
        for (int k = 0; k < s.n3; ++k)
            for (int j = 0; j < s.n2; ++j)
                for (int i = 0; i < s.n1; ++i)
                    b[(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)] =
                        src[((size_t)k * s.n2 + j) * s.n1 + i];
        return b;
    };
    for (int q = 0; q < 12; ++q) sb[q] = to_box(O[st_field[q]], st_field[q]);
    for (int q = 0; q < NA; ++q) {
        li[q] = to_box(lam[q], adj_field[q]);
        lo[q] = alloc_region(box);
        owned2.push_back(lo[q]);
    }
    double *bJ = to_box(dJ, SS_VR);
    emu_adj_step(nr, nt, np, pitch, R.data(), th.data(), ph.data(), &p, sb, li, lo, c.rchunk, bJ, 1.5, 0.01);
    for (int q = 0; q < NA; ++q) {
        const Shape s = shape_of(adj_field[q], nr, nt, np);
        double num = 0.0, mx = 0.0;
        for (int k = 0; k < s.n3; ++k)
            for (int j = 0; j < s.n2; ++j)
                for (int i = 0; i < s.n1; ++i) {
                    const double a = lo[q][(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)];
                    const double b = lam_ref[q][((size_t)k * s.n2 + j) * s.n1 + i];
                    num = std::fmax(num, std::fabs(a - b));
                    mx = std::fmax(mx, std::fabs(b));
                }
        if (!(num <= 1e-12 * std::fmax(mx, 1e-300))) {
            std::fprintf(stderr, "adjoint mismatch: case %dx%dx%d field %d err %.3e of %.3e\n", nr, nt, np, q, num, mx);
            ++bad;
        }
    }
    // the FP32 right-hand-side instance on the same case: rhs = dJ on VR and 0
    // elsewhere, so T^T lam + rhs is lam_ref again (to FP32 rounding)
    {
        std::vector<float *> owned3;
        auto to_box_f = [&](const std::vector<double> &src, int f) {
            float *b = alloc_region_f(box);
            owned3.push_back(b);
            const Shape s = shape_of(f, nr, nt, np);
            for (int k = 0; k < s.n3; ++k)
                for (int j = 0; j < s.n2; ++j)
                    for (int i = 0; i < s.n1; ++i)
                        b[(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)] =
                            (float)src[((size_t)k * s.n2 + j) * s.n1 + i];
            return b;
        };
        void *sf[12], *lif[NA], *lof[NA], *rf[NA];
        for (int q = 0; q < 12; ++q) sf[q] = to_box_f(O[st_field[q]], st_field[q]);
        for (int q = 0; q < NA; ++q) {
            lif[q] = to_box_f(lam[q], adj_field[q]);
            lof[q] = alloc_region_f(box);
            owned3.push_back((float *)lof[q]);
            rf[q] = q == ORC_A_VR ? to_box_f(dJ, SS_VR) : alloc_region_f(box);
            if (q != ORC_A_VR) owned3.push_back((float *)rf[q]);
        }
        emu_adj_step_rhs(1, nr, nt, np, pitch, R.data(), th.data(), ph.data(), &p, sf, lif, lof, c.rchunk, rf);
        for (int q = 0; q < NA; ++q) {
            const Shape s = shape_of(adj_field[q], nr, nt, np);
            const float *lo32 = (const float *)lof[q];
            double num = 0.0, mx = 0.0;
            for (int k = 0; k < s.n3; ++k)
                for (int j = 0; j < s.n2; ++j)
                    for (int i = 0; i < s.n1; ++i) {
                        const double a = lo32[(size_t)(i + s.off) * (nt + 1) * pitch + (size_t)(j + s.off) * pitch + (k + s.off)];
                        const double b = lam_ref[q][((size_t)k * s.n2 + j) * s.n1 + i];
                        num = std::fmax(num, std::fabs(a - b));
                        mx = std::fmax(mx, std::fabs(b));
                    }
            if (!(num <= 2e-6 * std::fmax(mx, 1e-300))) {
                std::fprintf(stderr, "FP32 adjoint mismatch: case %dx%dx%d field %d err %.3e of %.3e\n", nr, nt, np, q,
                             num, mx);
                ++bad;
            }
        }
        for (float *b : owned3) free_region_f(b);
    }
    for (double *b : owned) free_region(b);
    for (double *b : owned2) free_region(b);
    return bad;
}

} // namespace

int main()
{
    const double inf = INFINITY;
    const Case cases[] = {
        // nr nt  np  bc  n_exp relax emin emax rchunk paired
        {5, 5, 5, SS_BC_FREE_SLIP, 1.0, 1.0, 0.0, inf, 0, 0},
        {7, 11, 62, SS_BC_SHEAR, 1.0, 1.0, 0.0, inf, 0, 1},
        {9, 12, 63, SS_BC_FREE_SLIP, 1.0, 1.0, 0.0, inf, 4, 1},
        {6, 13, 64, SS_BC_PURE_SHEAR, 1.0, 1.0, 0.0, inf, 3, 0},
        {11, 21, 65, SS_BC_SHEAR, 3.0, 0.5, 0.3, 3.0, 5, 1},
        {13, 10, 126, SS_BC_PURE_SHEAR, 3.0, 1.0, 0.0, inf, 0, 1},
        {8, 22, 127, SS_BC_FREE_SLIP, 3.0, 0.5, 0.0, inf, 3, 0},
        {12, 31, 17, SS_BC_SHEAR, 5.0, 1.0, 1e-2, 1e2, 7, 1},
        // etaSM spread (F_SMBUF): window wider than the short axes
        {5, 13, 64, SS_BC_FREE_SLIP, 1.0, 1.0, 0.0, inf, 0, 1, 8},
        {9, 12, 65, SS_BC_PURE_SHEAR, 3.0, 0.5, 1e-4, 1e4, 4, 0, 3},
    };
    int bad = 0;
    for (const Case &c : cases) {
        const int b = run_case(c);
        std::printf("case %dx%dx%d bc %d n_exp %g relax %g rchunk %d paired %d: %s\n", c.nr, c.nt, c.np, c.bc_mode,
                    c.n_exp, c.relax, c.rchunk, c.paired, b ? "MISMATCH" : "ok");
        bad += b;
    }
    std::printf("%lld checked accesses\n", g_checks);
    if (bad) return 1;
    std::printf("sanitized run: all cases ok\n");
    return 0;
}
