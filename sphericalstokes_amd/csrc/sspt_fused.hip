// sspt_fused.hip -- one PT iteration in two launches: the r-marching fused
// kernel (all interior updates) and the boundary pass (ghost faces of V and
// boundary-cell viscosities).  The arithmetic lives in sspt_fused_body.h.
#include "sspt_internal.h"
#include "sspt_launch.h"
#include "sspt_fused_body.h"
#include <algorithm>
#include <cstdlib>

namespace sspt {

using namespace fused;

__device__ __forceinline__ unsigned long long wave_max(unsigned long long v)
{
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned long long o = __shfl_xor(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

template <typename R>
__device__ __forceinline__ void flush_max(const Args<R> &a, const unsigned long long *mdv, const unsigned long long *mv,
                                          int lane)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const unsigned long long d = wave_max(mdv[c]), v = wave_max(mv[c]);
        if (lane == 0) {
            if (d) atomicMax(&a.red[c], d);
            if (v) atomicMax(&a.red[3 + c], v);
        }
    }
}

// Tiles are numbered (z, y, x) with x fastest.  Consecutive workgroups are
// dealt round-robin to the 8 XCDs; the remap gives each XCD a contiguous run
// of tiles so that th/phi neighbours (which re-read each other's edges) share
// that XCD's L2.
__device__ __forceinline__ void tile_of(int nx, int ny, int nz, int &bx, int &by, int &bz)
{
    const int T = nx * ny * nz;
    const int per = (T + 7) / 8;
    const int b = blockIdx.x;
    const int t = (b % 8) * per + b / 8;
    bx = t % nx;
    by = (t / nx) % ny;
    bz = t / (nx * ny);
}

// SSPT_MAIN_WAVES64: optional waves-per-SIMD floor for the FP64 instantiations
// (register budget 512 / waves), for tile-shape measurements.
#ifdef SSPT_MAIN_WAVES64
#define SSPT_MAIN_ATTR __attribute__((amdgpu_waves_per_eu(SSPT_MAIN_WAVES64)))
#else
#define SSPT_MAIN_ATTR
#endif

template <typename R, int FL>
__global__ void __launch_bounds__(TKX<R> * TJB<R>) SSPT_MAIN_ATTR k_fused_main(Args<R> a, int nx, int ny, int nz)
{
    __shared__ Lds<R> L;
    int bx, by, bz;
    tile_of(nx, ny, nz, bx, by, bz);
    if (bz >= nz) return; // padding slots of the remap (uniform per block)
    static_assert(TKX<R> % 64 == 0 && TKX<R> * TJB<R> <= 1024, "rows of whole wavefronts, at most 1024 threads");
    const int kb = bx * OKX<R>, jb = by * OJ<R>;
    const int s0 = bz * a.rchunk, s1 = min(s0 + a.rchunk, a.b.nr);
    const int start = max(s0 - 2, -1);
    Thread<R> T;
    thread_init(T, a, threadIdx.x, threadIdx.y, kb, jb);
    thread_prologue<R, FL>(T, a, start);
    for (int s = start; s < s1; ++s) {
        const bool wr = s >= s0;
        phase1<R, FL>(T, L, a, s);
        __syncthreads();
        phase2<R, FL>(T, L, a, s, wr);
        __syncthreads();
        phase3<R, FL>(T, L, a, s, wr);
        __syncthreads();
        // phase4 reads only STT/SPP/DT/TP/RT/RP; the next phase1 writes only
        // VR/VT/VP/ETA, so no barrier is needed between them.
        phase4<R, FL>(T, L, a, s, wr);
    }
    if (FL & F_CHECK) flush_max(a, T.mdv, T.mv, threadIdx.x & 63);   // lane within the wavefront
}

// Boundary pass.  blockIdx.z = family: 0 -> r faces over (j,k),
// 1 -> th faces over (i,k), 2 -> phi faces over (i,j).
template <typename R>
__global__ void __launch_bounds__(256) k_fused_boundary(Args<R> a)
{
    const BoxG &b = a.b;
    const int u = blockIdx.x * blockDim.x + threadIdx.x; // fastest coordinate
    const int v = blockIdx.y * blockDim.y + threadIdx.y;
    const int fam = blockIdx.z;
    unsigned long long mdv[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
    const int nr = b.nr, nt = b.nt, np = b.np;
    for (int c = 0; c < 3; ++c) {
        const int n0 = c == 0 ? nr + 1 : nr, n1 = c == 1 ? nt + 1 : nt, n2 = c == 2 ? np + 1 : np;
        if (fam == 0) {
            if (u < n2 && v < n1) {
                ghost_v_point(a, c, 0, v, u, &mdv[c], &mv[c]);
                ghost_v_point(a, c, n0 - 1, v, u, &mdv[c], &mv[c]);
            }
        } else if (fam == 1) {
            if (u < n2 && v < n0) {
                ghost_v_point(a, c, v, 0, u, &mdv[c], &mv[c]);
                ghost_v_point(a, c, v, n1 - 1, u, &mdv[c], &mv[c]);
            }
        } else {
            if (u < n1 && v < n0) {
                ghost_v_point(a, c, v, u, 0, &mdv[c], &mv[c]);
                ghost_v_point(a, c, v, u, n2 - 1, &mdv[c], &mv[c]);
            }
        }
    }
    if (a.flags & F_CORR) {
        // eta is frozen in a correction sweep
    } else if (fam == 0) {
        if (u < np && v < nt) {
            boundary_eta_point(a, 0, v, u);
            boundary_eta_point(a, nr - 1, v, u);
        }
    } else if (fam == 1) {
        // each boundary cell exactly once: eta_PL is updated in place
        if (u < np && v >= 1 && v < nr - 1) {
            boundary_eta_point(a, v, 0, u);
            boundary_eta_point(a, v, nt - 1, u);
        }
    } else {
        if (u >= 1 && u < nt - 1 && v >= 1 && v < nr - 1) {
            boundary_eta_point(a, v, u, 0, 0);
            boundary_eta_point(a, v, u, np - 1, 1);
        }
    }
    if (a.flags & F_CHECK) flush_max(a, mdv, mv, (threadIdx.x + threadIdx.y * blockDim.x) & 63);
}

// The etaSM spread (spread_at's three separable passes over the cells, after
// the star of maxloc! + copy_BC!) in two launches.
//
// k_spread_r: the r pass, fused with the star when STAR.  A thread owns one
// (j, k) column of a chunk of rch planes and keeps the 2M+1 source values
// around its plane in registers (source = star_bc_at(src) or src itself, at
// the clamped plane), loading one plane ahead per step.
template <typename R, int M, bool STAR>
__global__ void __launch_bounds__(256) k_spread_r(BoxG b, const R *__restrict__ src, R *__restrict__ dst, int rch)
{
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    if (k >= b.np || j >= b.nt) return;
    const int i0 = blockIdx.z * rch, i1 = min(i0 + rch, b.nr);
    const SpreadW<R> w = spread_weights<R>();
    auto S = [&](int i) -> R {
        const int ic = clampi(i, 0, b.nr - 1);
        if constexpr (STAR) return star_bc_at(b, src, ic, j, k);
        else return src[b.off(ic, j, k)];
    };
    R win[2 * M + 1];
#pragma unroll
    for (int q = 0; q < 2 * M + 1; ++q) win[q] = S(i0 - M + q);
#pragma unroll 2
    for (int i = i0; i < i1; ++i) {
        const R ahead = S(i + M + 1);
        R v = win[M];
#pragma unroll
        for (int a = 1; a <= M; ++a) {
            v = std::fmax(v, win[M - a] * w.w[a]);
            v = std::fmax(v, win[M + a] * w.w[a]);
        }
        dst[b.off(i, j, k)] = v;
#pragma unroll
        for (int q = 0; q < 2 * M; ++q) win[q] = win[q + 1];
        win[2 * M] = ahead;
    }
}

// k_spread_thph: the th and phi passes of one r-plane tile (SPT_J rows x 64
// columns of outputs) through LDS: the tile and a halo of M clamped rows
// and columns of the r-pass result are staged, the th pass runs over the
// tile's rows and all staged columns, the phi pass over the outputs.  A
// staged row / column holds the value at the clamped index, so each pass
// reads src(clamp(x + a)) as spread_at does.
constexpr int SPT_J = 32;
template <typename R, int M>
__global__ void __launch_bounds__(256) k_spread_thph(BoxG b, const R *__restrict__ src, R *__restrict__ dst)
{
    constexpr int W = 64 + 2 * M, H = SPT_J + 2 * M;
    __shared__ R in[H][W];
    __shared__ R mid[SPT_J][W];
    const int k0 = blockIdx.x * 64, j0 = blockIdx.y * SPT_J, i = blockIdx.z;
    const int t = threadIdx.x + 64 * threadIdx.y;
    const SpreadW<R> w = spread_weights<R>();
    for (int e = t; e < H * W; e += 256) {
        const int r = e / W, c = e - r * W;
        in[r][c] = src[b.off(i, clampi(j0 - M + r, 0, b.nt - 1), clampi(k0 - M + c, 0, b.np - 1))];
    }
    __syncthreads();
    for (int e = t; e < SPT_J * W; e += 256) {
        const int r = e / W, c = e - r * W;
        R v = in[r + M][c];
#pragma unroll
        for (int a = 1; a <= M; ++a) {
            v = std::fmax(v, in[r + M - a][c] * w.w[a]);
            v = std::fmax(v, in[r + M + a][c] * w.w[a]);
        }
        mid[r][c] = v;
    }
    __syncthreads();
    for (int e = t; e < SPT_J * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        const int j = j0 + r, k = k0 + c;
        if (j >= b.nt || k >= b.np) continue;
        R v = mid[r][c + M];
#pragma unroll
        for (int a = 1; a <= M; ++a) {
            v = std::fmax(v, mid[r][c + M - a] * w.w[a]);
            v = std::fmax(v, mid[r][c + M + a] * w.w[a]);
        }
        dst[b.off(i, j, k)] = v;
    }
}

// r-chunk of k_spread_r: enough chunks to fill the device at small sizes,
// few enough that the 2M warm-up planes per chunk stay a small share
static int spread_rchunk(int nr) { return std::max(8, std::min(32, nr / 8)); }

template <typename R, int M> static void spread_two_pass(ss_ctx *c, const BoxG &b, const R *eta, R *sm, R *tmp)
{
    const int rch = spread_rchunk(b.nr);
    const dim3 Gr((b.np + 63) / 64, (b.nt + 3) / 4, (b.nr + rch - 1) / rch), Br(64, 4);
    if (eta) hipLaunchKernelGGL((k_spread_r<R, M, true>), Gr, Br, 0, c->stream, b, eta, tmp, rch);
    else hipLaunchKernelGGL((k_spread_r<R, M, false>), Gr, Br, 0, c->stream, b, (const R *)sm, tmp, rch);
    const dim3 Gt((b.np + 63) / 64, (b.nt + SPT_J - 1) / SPT_J, b.nr), Bt(64, 4);
    hipLaunchKernelGGL((k_spread_thph<R, M>), Gt, Bt, 0, c->stream, b, (const R *)tmp, sm);
}

// sm <- spread(maxloc! + copy_BC! of eta) with eta != null, sm <- spread(sm)
// with eta == null; tmp (a box of the same shape) holds the r-pass result.
// m is checked against SPREAD_MAX by every entry point that accepts it.
template <typename R> void launch_etasm_spread_box(ss_ctx *c, const BoxG &b, const R *eta, R *sm, R *tmp, int m)
{
    switch (m) {
    case 1: spread_two_pass<R, 1>(c, b, eta, sm, tmp); break;
    case 2: spread_two_pass<R, 2>(c, b, eta, sm, tmp); break;
    case 3: spread_two_pass<R, 3>(c, b, eta, sm, tmp); break;
    case 4: spread_two_pass<R, 4>(c, b, eta, sm, tmp); break;
    case 5: spread_two_pass<R, 5>(c, b, eta, sm, tmp); break;
    case 6: spread_two_pass<R, 6>(c, b, eta, sm, tmp); break;
    case 7: spread_two_pass<R, 7>(c, b, eta, sm, tmp); break;
    case 8: spread_two_pass<R, 8>(c, b, eta, sm, tmp); break;
    default: break;
    }
}
static_assert(SPREAD_MAX == 8, "launch_etasm_spread_box instantiates m = 1..8");
template void launch_etasm_spread_box<float>(ss_ctx *, const BoxG &, const float *, float *, float *, int);

void launch_etasm_spread(ss_ctx *c, double *SM, double *tmp, int m)
{
    const BoxG b{c->box.nr, c->box.nt, c->box.np, c->box.sj, c->box.si};
    launch_etasm_spread_box<double>(c, b, nullptr, SM, tmp, m);
}

__global__ void k_zero_red(unsigned long long *p)
{
    if (threadIdx.x < 6) p[threadIdx.x] = 0ull;
}

static Args<double> make_args(ss_ctx *c, const FieldSet &in, const FieldSet &out, const ss_params &p, bool last, bool check)
{
    static constexpr double PI = 3.14159265358979323846;
    Args<double> a{};
    a.b.nr = c->box.nr;
    a.b.nt = c->box.nt;
    a.b.np = c->box.np;
    a.b.sj = c->box.sj;
    a.b.si = c->box.si;
    a.t = TabsG<double>{c->ftr, c->ftt, c->ft2, c->tab.th, c->tab.ph};
    a.VR = in.f[SS_VR];
    a.VT = in.f[SS_VTH];
    a.VP = in.f[SS_VPH];
    a.eta = in.f[SS_ETA];
    a.P = in.f[SS_P];
    a.tRR = in.f[SS_TAU_RR];
    a.tTT = in.f[SS_TAU_THTH];
    a.tPP = in.f[SS_TAU_PHPH];
    a.tRT = in.f[SS_TAU_RTH];
    a.tRP = in.f[SS_TAU_RPH];
    a.tTP = in.f[SS_TAU_THPH];
    a.rho = in.f[SS_RHO];
    a.ini = in.f[SS_ETA_INI];
    a.ePL = in.f[SS_ETA_PL];
    a.oVR = out.f[SS_VR];
    a.oVT = out.f[SS_VTH];
    a.oVP = out.f[SS_VPH];
    a.oEta = out.f[SS_ETA];
    a.oP = out.f[SS_P];
    a.oRR = out.f[SS_TAU_RR];
    a.oTT = out.f[SS_TAU_THTH];
    a.oPP = out.f[SS_TAU_PHPH];
    a.oRT = out.f[SS_TAU_RTH];
    a.oRP = out.f[SS_TAU_RPH];
    a.oTP = out.f[SS_TAU_THPH];
    a.oPL = out.f[SS_ETA_PL];
    a.SM = out.f[SS_ETASM];
    a.DT = out.f[SS_DT_R];
    a.TII = out.f[SS_TII];
    a.TBP = c->tbp;
    a.VRe = out.f[SS_VR_ERR];
    a.VTe = out.f[SS_VTH_ERR];
    a.VPe = out.f[SS_VPH_ERR];
    a.red = c->red;
    a._dr = p._dr;
    a._dth = p._dth;
    a._dph = p._dph;
    a.kb = (((p.dmp * 2.0) * PI) * p.vpdt) / p.lr;
    a.gs = ((4.0 * PI) * p.vpdt) / p.lr;
    a.vpdt2 = p.vpdt * p.vpdt;
    a.g = p.g;
    a.tC = p.tau_C;
    a.ex = 1.0 - p.n_exp;
    a.relax = p.relax;
    a.emin = p.eta_min;
    a.emax = p.eta_max;
    a.flags = (last ? F_LAST : 0) | (check ? F_CHECK : 0) | (p.relax != 1.0 ? F_RELAX : 0) |
              ((p.eta_min > 0.0 || p.eta_max < INFINITY) ? F_CLAMP : 0) | (p.etasm_spread > 0 ? F_SMBUF : 0);
    a.spread = p.etasm_spread;
    a.bc_mode = p.bc_mode;
    a.r = p.r;
    a.sh = p.sh;
    a.vth0 = pure_shear_vth(c->th_lo, p.r, p.sh);
    a.vth1 = pure_shear_vth(c->th_hi, p.r, p.sh);
    a.rchunk = c->box.nr;
    return a;
}

// r-chunk length (march_chunk); SSPT_RCHUNK overrides it for measurements.
static int pick_rchunk(const ss_ctx *c, int nr, int tiles_xy, bool fp32)
{
    if (const char *e = std::getenv("SSPT_RCHUNK")) {
        const int v = std::atoi(e);
        if (v >= 1) return std::min(v, nr);
    }
    return march_chunk(nr, tiles_xy, c->n_cu, !fp32);
}

// Launches one fused iteration (main kernel + boundary pass) for prepared
// arguments; the flag combination selects the kernel instantiation.
template <typename R>
void launch_fused_args(ss_ctx *c, Args<R> a, hipEvent_t ev_main0, hipEvent_t ev_main1)
{
    a.pp_in = a.pp_out = a.ppe_in = a.ppe_out = 0;
    const int nx = (a.b.np + OKX<R> - 1) / OKX<R>;
    const int ny = (a.b.nt + OJ<R> - 1) / OJ<R>;
    a.rchunk = pick_rchunk(c, a.b.nr, nx * ny, sizeof(R) == 4);
    const int nz = (a.b.nr + a.rchunk - 1) / a.rchunk;
    if (a.flags & F_CHECK) hipLaunchKernelGGL(k_zero_red, dim3(1), dim3(64), 0, c->stream, a.red);
    // spread etaSM of the old viscosity into SM (DT_R as the temporary: the
    // main kernel writes DT_R only in a last iteration, over every cell); a
    // correction sweep's caller fills its SM field once per outer step
    if ((a.flags & F_SMBUF) && !(a.flags & F_CORR)) launch_etasm_spread_box<R>(c, a.b, a.eta, a.SM, a.DT, a.spread);
    const int T = nx * ny * nz;
    const int grid = ((T + 7) / 8) * 8;
    if (ev_main0) (void)hipEventRecord(ev_main0, c->stream);
    const dim3 G(grid), B(TKX<R>, TJB<R>);
    Args<R> am = a; // main-kernel view; the boundary pass keeps the plain pointers
    if (c->pair_state) pair_args(am, (long long)(c->box_elems * sizeof(R)));
    if constexpr (sizeof(R) == 8) {
        switch (am.flags) {
#define SSPT_FL(v) case v: hipLaunchKernelGGL((k_fused_main<R, v>), G, B, 0, c->stream, am, nx, ny, nz); break;
        SSPT_FL(0) SSPT_FL(1) SSPT_FL(3) SSPT_FL(4) SSPT_FL(5) SSPT_FL(7)
        SSPT_FL(8) SSPT_FL(9) SSPT_FL(11) SSPT_FL(12) SSPT_FL(13) SSPT_FL(15)
        SSPT_FL(32) SSPT_FL(33) SSPT_FL(35) SSPT_FL(36) SSPT_FL(37) SSPT_FL(39)
        SSPT_FL(40) SSPT_FL(41) SSPT_FL(43) SSPT_FL(44) SSPT_FL(45) SSPT_FL(47)
        SSPT_FL(64) SSPT_FL(65) SSPT_FL(67) SSPT_FL(68) SSPT_FL(69) SSPT_FL(71)
        SSPT_FL(72) SSPT_FL(73) SSPT_FL(75) SSPT_FL(76) SSPT_FL(77) SSPT_FL(79)
        SSPT_FL(96) SSPT_FL(97) SSPT_FL(99) SSPT_FL(100) SSPT_FL(101) SSPT_FL(103)
        SSPT_FL(104) SSPT_FL(105) SSPT_FL(107) SSPT_FL(108) SSPT_FL(109) SSPT_FL(111)
#undef SSPT_FL
        default: break; // CHECK without LAST is never requested in FP64
        }
    } else {
        switch (am.flags) {
#define SSPT_FL(v) case v: hipLaunchKernelGGL((k_fused_main<R, v>), G, B, 0, c->stream, am, nx, ny, nz); break;
        SSPT_FL(16) SSPT_FL(18) SSPT_FL(48) SSPT_FL(50) SSPT_FL(80) SSPT_FL(82) SSPT_FL(112) SSPT_FL(114)
#undef SSPT_FL
        default: break; // FP32 runs correction sweeps only
        }
    }
    if (ev_main1) (void)hipEventRecord(ev_main1, c->stream);
    const int m = std::max(std::max(a.b.nr, a.b.nt), a.b.np) + 1;
    hipLaunchKernelGGL((k_fused_boundary<R>), dim3((m + 63) / 64, (m + 3) / 4, 3), dim3(64, 4), 0, c->stream, a);
}

template void launch_fused_args<float>(ss_ctx *, Args<float>, hipEvent_t, hipEvent_t);

void launch_fused_iteration(ss_ctx *c, const FieldSet &in, const FieldSet &out, const ss_params &p, bool last,
                            bool check, hipEvent_t ev_main0, hipEvent_t ev_main1)
{
    launch_fused_args<double>(c, make_args(c, in, out, p, last, check), ev_main0, ev_main1);
}

} // namespace sspt
