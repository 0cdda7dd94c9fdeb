// sspt_cyl.hip -- the cylindrical twin of the diapir solver
// (scripts_diapir/CylindricalStokes_diapir.jl), kernel for kernel.
//
// The script's kernels are ParallelStencil @parallel kernels: each statement
// assigns where its left-hand array is defined (all points, or the inner
// points for @inn), with FiniteDifferences3D differences and averages over
// "a"ll points or the "i"nner points (index + 1) of the other directions;
// the averages sum their terms in the order (i,j) (i+1,j) (i,j+1) (i+1,j+1)
// of their two directions.  Fields live in boxes of (nr+1) x (nth+1) x pitch
// doubles (z fastest, pitch a multiple of 64 for coalesced wavefront rows).
// The kernels address them by the reference's 1-based indices through a
// per-field shift (box index = reference index - 1 for centre and face
// fields, = reference index for the edge and interior-face fields), so
// every statement below reads like the script's and one thread runs the
// statements of one (ix, iy, iz) in the script's order.
#include "sspt_internal.h"
#include "../../include/sspt_cyl.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace sspt {
namespace cyl {

constexpr double PI_ = 3.14159265358979323846;

struct YBox {
    int nr, nt, nz, pitch;
    long long sj, si;
};

// reference sizes of a field and its box shift
inline void dims_of(const YBox &b, int f, int d[3], int *sh)
{
    int a = 0, c = 0, e = 0;
    *sh = 0;
    switch (f) {
    case SY_VR: case SY_VR_OLD: a = 1; break;
    case SY_VTH: case SY_VTH_OLD: c = 1; break;
    case SY_VZ: case SY_VZ_OLD: e = 1; break;
    case SY_DVR: a = -1; c = -2; e = -2; *sh = 1; break;
    case SY_DVTH: a = -2; c = -1; e = -2; *sh = 1; break;
    case SY_DVZ: a = -2; c = -2; e = -1; *sh = 1; break;
    case SY_D_RTH: case SY_TAU_RTH: case SY_ERR_RTH: a = -1; c = -1; e = -2; *sh = 1; break;
    case SY_D_RZ: case SY_TAU_RZ: a = -1; c = -2; e = -1; *sh = 1; break;
    case SY_D_THZ: case SY_TAU_THZ: a = -2; c = -1; e = -1; *sh = 1; break;
    default: break;
    }
    d[0] = b.nr + a; d[1] = b.nt + c; d[2] = b.nz + e;
}

// the fields of a context as the kernels see them
struct F {
    YBox b;
    double *p[SY_NFIELDS];
    int sh[SY_NFIELDS];
    const double *R, *th; // R_i[nr], th_i[nth]
    // reference element A[ix, iy, iz] (1-based)
    __device__ __forceinline__ double &operator()(int f, int ix, int iy, int iz) const
    {
        const int s = sh[f] - 1;
        return p[f][(long long)(ix + s) * b.si + (long long)(iy + s) * b.sj + (iz + s)];
    }
};

static const dim3 B3(64, 4, 1);
static inline dim3 grid3(int n1, int n2, int n3) { return dim3((n3 + 63) / 64, (n2 + 3) / 4, n1); }

// thread -> 1-based (ix, iy, iz) over n1 x n2 x n3 (z fastest)
#define Y_LOOP(N1, N2, N3)                                             \
    const int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;          \
    const int iy = blockIdx.y * blockDim.y + threadIdx.y + 1;          \
    const int ix = blockIdx.z + 1;                                     \
    if (ix > (N1) || iy > (N2) || iz > (N3)) return;

__device__ __forceinline__ double jl_max(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}

// ---------------------------------------------------------------- transposes
__global__ void k_julia_to_box(YBox b, double *dst, const double *src, int n1, int n2, int n3, int sh)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n1 * n2 * n3) return;
    const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long long)n1 * n2));
    dst[(long long)(i + sh) * b.si + (long long)(j + sh) * b.sj + (k + sh)] = src[t];
}
__global__ void k_box_to_julia(YBox b, double *dst, const double *src, int n1, int n2, int n3, int sh)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n1 * n2 * n3) return;
    const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long long)n1 * n2));
    dst[t] = src[(long long)(i + sh) * b.si + (long long)(j + sh) * b.sj + (k + sh)];
}
__global__ void k_c_to_box(YBox b, double *dst, const double *src, int n1, int n2, int n3, int sh)
{
    Y_LOOP(n1, n2, n3)
    dst[(long long)(ix - 1 + sh) * b.si + (long long)(iy - 1 + sh) * b.sj + (iz - 1 + sh)] =
        src[((long long)(ix - 1) * n2 + (iy - 1)) * n3 + (iz - 1)];
}
__global__ void k_box_to_c(YBox b, double *dst, const double *src, int n1, int n2, int n3, int sh)
{
    Y_LOOP(n1, n2, n3)
    dst[((long long)(ix - 1) * n2 + (iy - 1)) * n3 + (iz - 1)] =
        src[(long long)(ix - 1 + sh) * b.si + (long long)(iy - 1 + sh) * b.sj + (iz - 1 + sh)];
}
__global__ void k_zero(double *p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

// ---------------------------------------------------------------- kernels
// equal! / multiply! (:41-49) over dst's size
__global__ void k_equal(F a, int dst, int src, int n1, int n2, int n3, int mul, double fact)
{
    Y_LOOP(n1, n2, n3)
    a(dst, ix, iy, iz) = mul ? a(src, ix, iy, iz) * fact : a(src, ix, iy, iz);
}

// maxloc! (:125-128): @inn(etaSM) = @maxloc(eta)
__global__ void k_maxloc(F a)
{
    Y_LOOP(a.b.nr - 2, a.b.nt - 2, a.b.nz - 2)
    const int e = SY_ETA;
    a(SY_ETASM, ix + 1, iy + 1, iz + 1) =
        jl_max(jl_max(jl_max(jl_max(a(e, ix, iy + 1, iz + 1), a(e, ix + 2, iy + 1, iz + 1)), a(e, ix + 1, iy + 1, iz + 1)),
                      jl_max(a(e, ix + 1, iy, iz + 1), a(e, ix + 1, iy + 2, iz + 1))),
               jl_max(a(e, ix + 1, iy + 1, iz), a(e, ix + 1, iy + 1, iz + 2)));
}

// timesteps! (:130-135): kb = dmp*2.0*pi*vpdt/lr, gs = 4.0*pi*vpdt/lr (left folds)
__global__ void k_timesteps(F a, double kb, double gs)
{
    Y_LOOP(a.b.nr, a.b.nt, a.b.nz)
    const double sm = a(SY_ETASM, ix, iy, iz);
    a(SY_KBDT, ix, iy, iz) = kb * sm;
    a(SY_GSDT, ix, iy, iz) = gs * sm;
}

// timesteps2! (:137-141), vpdt^2 a single product
__global__ void k_timesteps2(F a, double vpdt2)
{
    Y_LOOP(a.b.nr, a.b.nt, a.b.nz)
    const double G = a(SY_GSDT, ix, iy, iz);
    a(SY_DT_R, ix, iy, iz) = vpdt2 / (a(SY_KBDT, ix, iy, iz) + G / (1.0 + G / a(SY_ETA, ix, iy, iz)));
}

__device__ __forceinline__ double av4(double a0, double a1, double a2, double a3)
{
    return (((a0 + a1) + a2) + a3) * 0.25;
}

// compute_P! (:144-167): the statements in order at one point; R varies
// along r only (R[ix])
__global__ void k_compute_P(F a, double dr, double dth, double dz)
{
    Y_LOOP(a.b.nr, a.b.nt, a.b.nz)
    const int nr = a.b.nr, nt = a.b.nt, nz = a.b.nz;
    const double *R = a.R;
    const double Rc = R[ix - 1];
    // centres: DIVV, D_RR, D_thth, D_ZZ
    const double dvr = a(SY_VR, ix + 1, iy, iz) - a(SY_VR, ix, iy, iz);
    const double dvt = a(SY_VTH, ix, iy + 1, iz) - a(SY_VTH, ix, iy, iz);
    const double dvz = a(SY_VZ, ix, iy, iz + 1) - a(SY_VZ, ix, iy, iz);
    const double avr = (a(SY_VR, ix, iy, iz) + a(SY_VR, ix + 1, iy, iz)) * 0.5;
    const double div = ((dvr / dr + ((1.0 / Rc) * dvt) / dth) + avr / Rc) + dvz / dz;
    const double drr = dvr / dr - (1.0 / 3.0) * div;
    const double dtt = (((1.0 / Rc) * dvt) / dth + avr / Rc) - (1.0 / 3.0) * div;
    const double dzz = dvz / dz - (1.0 / 3.0) * div;
    a(SY_DIVV, ix, iy, iz) = div;
    a(SY_D_RR, ix, iy, iz) = drr;
    a(SY_D_THTH, ix, iy, iz) = dtt;
    a(SY_D_ZZ, ix, iy, iz) = dzz;
    double drt = 0.0, drz = 0.0, dtz = 0.0;
    const bool in_rt = ix <= nr - 1 && iy <= nt - 1 && iz <= nz - 2;
    const bool in_rz = ix <= nr - 1 && iy <= nt - 2 && iz <= nz - 1;
    const bool in_tz = ix <= nr - 2 && iy <= nt - 1 && iz <= nz - 1;
    if (in_rt) {
        const double Rxy = (((R[ix - 1] + R[ix]) + R[ix - 1]) + R[ix]) * 0.25;               // @av_xyi(R)
        const double dxi = a(SY_VTH, ix + 1, iy + 1, iz + 1) - a(SY_VTH, ix, iy + 1, iz + 1);
        const double dyi = a(SY_VR, ix + 1, iy + 1, iz + 1) - a(SY_VR, ix + 1, iy, iz + 1);
        const double axi = (a(SY_VTH, ix, iy + 1, iz + 1) + a(SY_VTH, ix + 1, iy + 1, iz + 1)) * 0.5;
        drt = (dxi / dr + ((1.0 / Rxy) * dyi) / dth) - axi / Rxy;
        a(SY_D_RTH, ix, iy, iz) = drt;
    }
    if (in_rz) {
        const double dxi = a(SY_VZ, ix + 1, iy + 1, iz + 1) - a(SY_VZ, ix, iy + 1, iz + 1);
        const double dzi = a(SY_VR, ix + 1, iy + 1, iz + 1) - a(SY_VR, ix + 1, iy + 1, iz);
        drz = dxi / dr + dzi / dz;
        a(SY_D_RZ, ix, iy, iz) = drz;
    }
    if (in_tz) {
        const double Ryz = (((R[ix] + R[ix]) + R[ix]) + R[ix]) * 0.25;                       // @av_yzi(R)
        const double dyi = a(SY_VZ, ix + 1, iy + 1, iz + 1) - a(SY_VZ, ix + 1, iy, iz + 1);
        const double dzi = a(SY_VTH, ix + 1, iy + 1, iz + 1) - a(SY_VTH, ix + 1, iy + 1, iz);
        dtz = ((1.0 / Ryz) * dyi) / dth + dzi / dz;
        a(SY_D_THZ, ix, iy, iz) = dtz;
    }
    {
        const double K = a(SY_KBDT, ix, iy, iz), G = a(SY_GSDT, ix, iy, iz), e = a(SY_ETA, ix, iy, iz);
        a(SY_P, ix, iy, iz) = a(SY_P, ix, iy, iz) - K * div;
        a(SY_TAU_RR, ix, iy, iz) = (a(SY_TAU_RR, ix, iy, iz) + (G * 2.0) * drr) / (1.0 + G / e);
        a(SY_TAU_THTH, ix, iy, iz) = (a(SY_TAU_THTH, ix, iy, iz) + (G * 2.0) * dtt) / (1.0 + G / e);
        a(SY_TAU_ZZ, ix, iy, iz) = (a(SY_TAU_ZZ, ix, iy, iz) + (G * 2.0) * dzz) / (1.0 + G / e);
    }
    const int g = SY_GSDT, e = SY_ETA;
    if (in_rt) { // @av_xyi
        const double G = av4(a(g, ix, iy, iz + 1), a(g, ix + 1, iy, iz + 1), a(g, ix, iy + 1, iz + 1), a(g, ix + 1, iy + 1, iz + 1));
        const double E = av4(a(e, ix, iy, iz + 1), a(e, ix + 1, iy, iz + 1), a(e, ix, iy + 1, iz + 1), a(e, ix + 1, iy + 1, iz + 1));
        a(SY_TAU_RTH, ix, iy, iz) = (a(SY_TAU_RTH, ix, iy, iz) + G * drt) / (1.0 + G / E);
    }
    if (in_rz) { // @av_xzi
        const double G = av4(a(g, ix, iy + 1, iz), a(g, ix + 1, iy + 1, iz), a(g, ix, iy + 1, iz + 1), a(g, ix + 1, iy + 1, iz + 1));
        const double E = av4(a(e, ix, iy + 1, iz), a(e, ix + 1, iy + 1, iz), a(e, ix, iy + 1, iz + 1), a(e, ix + 1, iy + 1, iz + 1));
        a(SY_TAU_RZ, ix, iy, iz) = (a(SY_TAU_RZ, ix, iy, iz) + G * drz) / (1.0 + G / E);
    }
    if (in_tz) { // @av_yzi
        const double G = av4(a(g, ix + 1, iy, iz), a(g, ix + 1, iy + 1, iz), a(g, ix + 1, iy, iz + 1), a(g, ix + 1, iy + 1, iz + 1));
        const double E = av4(a(e, ix + 1, iy, iz), a(e, ix + 1, iy + 1, iz), a(e, ix + 1, iy, iz + 1), a(e, ix + 1, iy + 1, iz + 1));
        a(SY_TAU_THZ, ix, iy, iz) = (a(SY_TAU_THZ, ix, iy, iz) + G * dtz) / (1.0 + G / E);
    }
}

// compute_sigma! (:169-175)
__global__ void k_compute_sigma(F a)
{
    Y_LOOP(a.b.nr, a.b.nt, a.b.nz)
    const double P = a(SY_P, ix, iy, iz);
    a(SY_SIG_RR, ix, iy, iz) = -P + a(SY_TAU_RR, ix, iy, iz);
    a(SY_SIG_THTH, ix, iy, iz) = -P + a(SY_TAU_THTH, ix, iy, iz);
    a(SY_SIG_ZZ, ix, iy, iz) = -P + a(SY_TAU_ZZ, ix, iy, iz);
}

// compute_TII! (:177-182)
__global__ void k_compute_TII(F a)
{
    Y_LOOP(a.b.nr - 2, a.b.nt - 2, a.b.nz - 2)
    const double x0 = a(SY_TAU_RR, ix + 1, iy + 1, iz + 1), y0 = a(SY_TAU_THTH, ix + 1, iy + 1, iz + 1),
                 z0 = a(SY_TAU_ZZ, ix + 1, iy + 1, iz + 1);
    const int t = SY_TAU_RTH, u = SY_TAU_RZ, w = SY_TAU_THZ;
    const double x = av4(a(t, ix, iy, iz), a(t, ix + 1, iy, iz), a(t, ix, iy + 1, iz), a(t, ix + 1, iy + 1, iz));
    const double y = av4(a(u, ix, iy, iz), a(u, ix + 1, iy, iz), a(u, ix, iy, iz + 1), a(u, ix + 1, iy, iz + 1));
    const double z = av4(a(w, ix, iy, iz), a(w, ix, iy + 1, iz), a(w, ix, iy, iz + 1), a(w, ix, iy + 1, iz + 1));
    a(SY_TII, ix + 1, iy + 1, iz + 1) = sqrt((((0.5 * ((x0 * x0 + y0 * y0) + z0 * z0)) + x * x) + y * y) + z * z);
}

// power_law! (:184-191)
__global__ void k_power_law(F a, double tC, double ex, double relax)
{
    Y_LOOP(a.b.nr, a.b.nt, a.b.nz)
    const double ini = a(SY_ETA_INI, ix, iy, iz);
    double pl = ini * pow(a(SY_TII, ix, iy, iz) / tC, ex);
    pl = exp(log(pl) * relax + log(a(SY_ETA_PL_OLD, ix, iy, iz)) * (1.0 - relax));
    a(SY_ETA_PL, ix, iy, iz) = pl;
    a(SY_ETA, ix, iy, iz) = 2.0 / (1.0 / ini + 1.0 / pl);
}

// compute_dV! (:193-206) over the union of the three ranges
__global__ void k_compute_dV(F a, double dr, double dth, double dz)
{
    Y_LOOP(a.b.nr - 1, a.b.nt - 1, a.b.nz - 1)
    const int nr = a.b.nr, nt = a.b.nt, nz = a.b.nz;
    const double *R = a.R;
    if (ix <= nr - 1 && iy <= nt - 2 && iz <= nz - 2) {
        const double Rx = (R[ix - 1] + R[ix]) * 0.5;                                          // @av_xi(R)
        const double t = ((((a(SY_SIG_RR, ix + 1, iy + 1, iz + 1) - a(SY_SIG_RR, ix, iy + 1, iz + 1)) / dr +
                            ((1.0 / Rx) * (a(SY_TAU_RTH, ix, iy + 1, iz) - a(SY_TAU_RTH, ix, iy, iz))) / dth) +
                           (a(SY_TAU_RZ, ix, iy, iz + 1) - a(SY_TAU_RZ, ix, iy, iz)) / dz) +
                          ((a(SY_SIG_RR, ix, iy + 1, iz + 1) + a(SY_SIG_RR, ix + 1, iy + 1, iz + 1)) * 0.5) / Rx) -
                         ((a(SY_SIG_THTH, ix, iy + 1, iz + 1) + a(SY_SIG_THTH, ix + 1, iy + 1, iz + 1)) * 0.5) / Rx;
        a(SY_DVR, ix, iy, iz) = t - (a(SY_RHOG, ix, iy + 1, iz + 1) + a(SY_RHOG, ix + 1, iy + 1, iz + 1)) * 0.5;
    }
    if (ix <= nr - 2 && iy <= nt - 1 && iz <= nz - 2) {
        const double Ry = (R[ix] + R[ix]) * 0.5;                                              // @av_yi(R)
        a(SY_DVTH, ix, iy, iz) =
            (((a(SY_TAU_RTH, ix + 1, iy, iz) - a(SY_TAU_RTH, ix, iy, iz)) / dr +
              ((1.0 / Ry) * (a(SY_SIG_THTH, ix + 1, iy + 1, iz + 1) - a(SY_SIG_THTH, ix + 1, iy, iz + 1))) / dth) +
             (a(SY_TAU_THZ, ix, iy, iz + 1) - a(SY_TAU_THZ, ix, iy, iz)) / dz) +
            (2 * ((a(SY_TAU_RTH, ix, iy, iz) + a(SY_TAU_RTH, ix + 1, iy, iz)) * 0.5)) / Ry;
    }
    if (ix <= nr - 2 && iy <= nt - 2 && iz <= nz - 1) {
        const double Rz = (R[ix] + R[ix]) * 0.5;                                              // @av_zi(R)
        a(SY_DVZ, ix, iy, iz) =
            (((a(SY_TAU_RZ, ix + 1, iy, iz) - a(SY_TAU_RZ, ix, iy, iz)) / dr +
              ((1.0 / Rz) * (a(SY_TAU_THZ, ix, iy + 1, iz) - a(SY_TAU_THZ, ix, iy, iz))) / dth) +
             (a(SY_SIG_ZZ, ix + 1, iy + 1, iz + 1) - a(SY_SIG_ZZ, ix + 1, iy + 1, iz)) / dz) +
            ((a(SY_TAU_RZ, ix, iy, iz) + a(SY_TAU_RZ, ix + 1, iy, iz)) * 0.5) / Rz;
    }
}

// compute_V! (:208-214)
__global__ void k_compute_V(F a)
{
    Y_LOOP(a.b.nr - 1, a.b.nt - 1, a.b.nz - 1)
    const int nr = a.b.nr, nt = a.b.nt, nz = a.b.nz, D = SY_DT_R;
    if (ix <= nr - 1 && iy <= nt - 2 && iz <= nz - 2)
        a(SY_VR, ix + 1, iy + 1, iz + 1) = a(SY_VR, ix + 1, iy + 1, iz + 1) +
            a(SY_DVR, ix, iy, iz) * ((a(D, ix, iy + 1, iz + 1) + a(D, ix + 1, iy + 1, iz + 1)) * 0.5);
    if (ix <= nr - 2 && iy <= nt - 1 && iz <= nz - 2)
        a(SY_VTH, ix + 1, iy + 1, iz + 1) = a(SY_VTH, ix + 1, iy + 1, iz + 1) +
            a(SY_DVTH, ix, iy, iz) * ((a(D, ix + 1, iy, iz + 1) + a(D, ix + 1, iy + 1, iz + 1)) * 0.5);
    if (ix <= nr - 2 && iy <= nt - 2 && iz <= nz - 1)
        a(SY_VZ, ix + 1, iy + 1, iz + 1) = a(SY_VZ, ix + 1, iy + 1, iz + 1) +
            a(SY_DVZ, ix, iy, iz) * ((a(D, ix + 1, iy + 1, iz) + a(D, ix + 1, iy + 1, iz + 1)) * 0.5);
}

// boundary kernels (:52-101) on field f of size n1 x n2 x n3: threads over
// the two other directions, both sides per thread
__global__ void k_bc(F a, int f, int n1, int n2, int n3, int op, double fact, double r)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1, v = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const double *th = a.th;
    if (op == SY_COPY_BC_R || op == SY_BC_R_0) {
        const int iy = v, iz = u;
        if (iy > n2 || iz > n3) return;
        const double s = op == SY_BC_R_0 ? -1.0 : 1.0;
        a(f, 1, iy, iz) = op == SY_BC_R_0 ? -a(f, 2, iy, iz) : a(f, 2, iy, iz);
        a(f, n1, iy, iz) = op == SY_BC_R_0 ? -a(f, n1 - 1, iy, iz) : a(f, n1 - 1, iy, iz);
        (void)s;
    } else if (op == SY_COPY_BC_TH || op == SY_BC_TH_FACT || op == SY_BC_TH_0) {
        const int ix = v, iz = u;
        if (ix > n1 || iz > n3) return;
        if (op == SY_COPY_BC_TH) {
            a(f, ix, 1, iz) = a(f, ix, 2, iz);
            a(f, ix, n2, iz) = a(f, ix, n2 - 1, iz);
        } else if (op == SY_BC_TH_0) {
            a(f, ix, 1, iz) = -a(f, ix, 2, iz);
            a(f, ix, n2, iz) = -a(f, ix, n2 - 1, iz);
        } else {
            a(f, ix, 1, iz) = (fact * th[0]) * r;
            a(f, ix, n2, iz) = (fact * th[n2 - 1]) * r;
        }
    } else {
        const int ix = v, iy = u;
        if (ix > n1 || iy > n2) return;
        if (op == SY_COPY_BC_Z) {
            a(f, ix, iy, 1) = a(f, ix, iy, 2);
            a(f, ix, iy, n3) = a(f, ix, iy, n3 - 1);
        } else if (op == SY_BC_Z_0) {
            a(f, ix, iy, 1) = 0.0;
            a(f, ix, iy, n3) = 0.0;
        } else {
            a(f, ix, iy, 1) = ((2 * th[iy - 1]) * r) * fact - a(f, ix, iy, 2);
            a(f, ix, iy, n3) = ((2 * th[iy - 1]) * r) * fact - a(f, ix, iy, n3 - 1);
        }
    }
}

// err_rθ! (:217-220)
__global__ void k_err_rt(F a)
{
    Y_LOOP(a.b.nr - 1, a.b.nt - 1, a.b.nz - 2)
    const int e = SY_ETA;
    a(SY_ERR_RTH, ix, iy, iz) = a(SY_TAU_RTH, ix, iy, iz) -
        a(SY_D_RTH, ix, iy, iz) * av4(a(e, ix, iy, iz + 1), a(e, ix + 1, iy, iz + 1), a(e, ix, iy + 1, iz + 1),
                                      a(e, ix + 1, iy + 1, iz + 1));
}

// max |V - V_OLD| and max |V| as bit patterns (non-negative doubles order
// like their bits; NaN sorts above +inf, so it propagates as in maximum())
__global__ void __launch_bounds__(256) k_exit_max(F a, int v, int o, int n1, int n2, int n3, unsigned long long *red)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int ix = blockIdx.z + 1;
    unsigned long long me = 0, mv = 0;
    if (ix <= n1 && iy <= n2 && iz <= n3) {
        const double V = a(v, ix, iy, iz);
        me = (unsigned long long)__double_as_longlong(fabs(V - a(o, ix, iy, iz)));
        mv = (unsigned long long)__double_as_longlong(fabs(V));
    }
    for (int s = 32; s >= 1; s >>= 1) {
        const unsigned long long oe = __shfl_xor(me, s, 64), ov = __shfl_xor(mv, s, 64);
        me = oe > me ? oe : me;
        mv = ov > mv ? ov : mv;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&red[0], me);
        atomicMax(&red[1], mv);
    }
}

__global__ void k_zero_red(unsigned long long *red, int n)
{
    if ((int)threadIdx.x < n) red[threadIdx.x] = 0ull;
}

} // namespace cyl
} // namespace sspt

using namespace sspt::cyl;

struct ss_cyl {
    int device = 0;
    YBox box{};
    size_t box_elems = 0;
    hipStream_t stream = nullptr;
    double *p[SY_NFIELDS] = {};
    double *R = nullptr, *th = nullptr;
    bool grid_set = false;
    double *staging = nullptr;
    unsigned long long *red = nullptr, *red_host = nullptr;
    hipEvent_t ev[2] = {};
    std::string err;
};

namespace {

int yfail(ss_cyl *c, int code, const std::string &m)
{
    c->err = m;
    return code;
}

int ysync(ss_cyl *c, const char *where)
{
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return yfail(c, (int)e, std::string(where) + ": " + hipGetErrorString(e));
    return SS_OK;
}

#define YHIP(c, expr)                                                                                           \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess) return yfail((c), (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

F fields(const ss_cyl *c)
{
    F a{};
    a.b = c->box;
    for (int f = 0; f < SY_NFIELDS; ++f) {
        int d[3];
        a.p[f] = c->p[f];
        dims_of(c->box, f, d, &a.sh[f]);
    }
    a.R = c->R;
    a.th = c->th;
    return a;
}

void l_equal(ss_cyl *c, int dst, int src, bool mul = false, double fact = 1.0)
{
    int d[3], sh;
    dims_of(c->box, dst, d, &sh);
    hipLaunchKernelGGL(k_equal, grid3(d[0], d[1], d[2]), B3, 0, c->stream, fields(c), dst, src, d[0], d[1], d[2],
                       mul ? 1 : 0, fact);
}

void l_bc(ss_cyl *c, int f, int op, double fact, double r)
{
    int d[3], sh;
    dims_of(c->box, f, d, &sh);
    int u, v;
    if (op == SY_COPY_BC_R || op == SY_BC_R_0) { u = d[2]; v = d[1]; }
    else if (op == SY_COPY_BC_TH || op == SY_BC_TH_FACT || op == SY_BC_TH_0) { u = d[2]; v = d[0]; }
    else { u = d[1]; v = d[0]; }
    hipLaunchKernelGGL(k_bc, dim3((u + 63) / 64, (v + 3) / 4), B3, 0, c->stream, fields(c), f, d[0], d[1], d[2], op,
                       fact, r);
}

void l_maxloc(ss_cyl *c)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_maxloc, grid3(b.nr - 2, b.nt - 2, b.nz - 2), B3, 0, c->stream, fields(c));
}
void l_timesteps(ss_cyl *c, double dmp, double vpdt, double lr)
{
    const YBox &b = c->box;
    const double kb = (((dmp * 2.0) * PI_) * vpdt) / lr, gs = ((4.0 * PI_) * vpdt) / lr;
    hipLaunchKernelGGL(k_timesteps, grid3(b.nr, b.nt, b.nz), B3, 0, c->stream, fields(c), kb, gs);
}
void l_timesteps2(ss_cyl *c, double vpdt)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_timesteps2, grid3(b.nr, b.nt, b.nz), B3, 0, c->stream, fields(c), vpdt * vpdt);
}
void l_compute_P(ss_cyl *c, double dr, double dth, double dz)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_compute_P, grid3(b.nr, b.nt, b.nz), B3, 0, c->stream, fields(c), dr, dth, dz);
}
void l_compute_sigma(ss_cyl *c)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_compute_sigma, grid3(b.nr, b.nt, b.nz), B3, 0, c->stream, fields(c));
}
void l_compute_TII(ss_cyl *c)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_compute_TII, grid3(b.nr - 2, b.nt - 2, b.nz - 2), B3, 0, c->stream, fields(c));
}
void l_power_law(ss_cyl *c, double tC, double n_exp, double relax)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_power_law, grid3(b.nr, b.nt, b.nz), B3, 0, c->stream, fields(c), tC, 1.0 - n_exp, relax);
}
void l_compute_dV(ss_cyl *c, double dr, double dth, double dz)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_compute_dV, grid3(b.nr - 1, b.nt - 1, b.nz - 1), B3, 0, c->stream, fields(c), dr, dth, dz);
}
void l_compute_V(ss_cyl *c)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_compute_V, grid3(b.nr - 1, b.nt - 1, b.nz - 1), B3, 0, c->stream, fields(c));
}
void l_err_rt(ss_cyl *c)
{
    const YBox &b = c->box;
    hipLaunchKernelGGL(k_err_rt, grid3(b.nr - 1, b.nt - 1, b.nz - 2), B3, 0, c->stream, fields(c));
}

int exit_error(ss_cyl *c, double *err, double comp[3])
{
    hipLaunchKernelGGL(k_zero_red, dim3(1), dim3(64), 0, c->stream, c->red, 6);
    static const int V[3] = {SY_VR, SY_VTH, SY_VZ}, O[3] = {SY_VR_OLD, SY_VTH_OLD, SY_VZ_OLD};
    for (int q = 0; q < 3; ++q) {
        int d[3], sh;
        dims_of(c->box, V[q], d, &sh);
        hipLaunchKernelGGL(k_exit_max, grid3(d[0], d[1], d[2]), B3, 0, c->stream, fields(c), V[q], O[q], d[0], d[1],
                           d[2], c->red + 2 * q);
    }
    YHIP(c, hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    int rc = ysync(c, "exit test");
    if (rc != SS_OK) return rc;
    double e = 0.0;
    bool nan = false;
    for (int q = 0; q < 3; ++q) {
        double me, mv;
        memcpy(&me, &c->red_host[2 * q], 8);
        memcpy(&mv, &c->red_host[2 * q + 1], 8);
        const double r = me / mv;
        if (comp) comp[q] = r;
        if (std::isnan(r)) nan = true;
        else if (q == 0 || r > e) e = r;
    }
    *err = nan ? NAN : e;
    return SS_OK;
}

// the loop body (:331-366) with the exponent n_exp
void l_iteration(ss_cyl *c, const ss_cyl_params &p, double n_exp, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    l_equal(c, SY_ETASM, SY_ETA);
    l_maxloc(c);
    l_bc(c, SY_ETASM, SY_COPY_BC_R, 0.0, 0.0);
    l_bc(c, SY_ETASM, SY_COPY_BC_TH, 0.0, 0.0);
    l_bc(c, SY_ETASM, SY_COPY_BC_Z, 0.0, 0.0);
    l_timesteps(c, p.dmp, p.vpdt, p.lr);
    l_timesteps2(c, p.vpdt);
    if (e0) (void)hipEventRecord(e0, c->stream);
    l_compute_P(c, p.dr, p.dth, p.dz);
    if (e1) (void)hipEventRecord(e1, c->stream);
    l_compute_sigma(c);
    if (n_exp > 1.0) {
        l_compute_TII(c);
        l_bc(c, SY_TII, SY_COPY_BC_R, 0.0, 0.0);
        l_bc(c, SY_TII, SY_COPY_BC_TH, 0.0, 0.0);
        l_bc(c, SY_TII, SY_COPY_BC_Z, 0.0, 0.0);
        l_equal(c, SY_ETA_PL_OLD, SY_ETA_PL);
        l_power_law(c, p.tau_C, n_exp, p.relax);
    }
    l_equal(c, SY_VR_OLD, SY_VR);
    l_equal(c, SY_VTH_OLD, SY_VTH);
    l_equal(c, SY_VZ_OLD, SY_VZ);
    l_compute_dV(c, p.dr, p.dth, p.dz);
    l_compute_V(c);
    l_bc(c, SY_VZ, SY_COPY_BC_R, 0.0, 0.0);
    l_bc(c, SY_VTH, SY_COPY_BC_R, 0.0, 0.0);
    l_bc(c, SY_VR, SY_BC_R_0, 0.0, 0.0);
    l_bc(c, SY_VZ, SY_BC_TH_FACT, p.sh, p.r);
    l_bc(c, SY_VR, SY_BC_TH_FACT, 0.0, p.r);
    l_bc(c, SY_VTH, SY_BC_TH_0, 0.0, 0.0);
    l_bc(c, SY_VTH, SY_BC_Z_0, 0.0, 0.0);
    l_bc(c, SY_VR, SY_BC_Z_0, 0.0, 0.0);
    l_bc(c, SY_VZ, SY_BC_Z_LIN, p.sh, p.r);
}

bool valid_bc(const ss_cyl *c, int f, int op)
{
    if (op < SY_COPY_BC_R || op > SY_BC_Z_LIN) return false;
    // bc_th_fact! / bc_z_lin! take theta at cell centres (theta_VZ, theta_VR):
    // fields whose th extent is nth
    if ((op == SY_BC_TH_FACT || op == SY_BC_Z_LIN) && (f == SY_VTH || f == SY_VTH_OLD || f == SY_DVTH || f == SY_D_RTH ||
                                                       f == SY_TAU_RTH || f == SY_ERR_RTH || f == SY_D_THZ || f == SY_TAU_THZ))
        return false;
    int d[3], sh;
    dims_of(c->box, f, d, &sh);
    return d[0] >= 2 && d[1] >= 2 && d[2] >= 2;
}

using yclk = std::chrono::steady_clock;

} // namespace

#define YCTX(c)                                    \
    do {                                           \
        if (!(c)) return SS_ERR_ARG;               \
        YHIP((c), hipSetDevice((c)->device));      \
    } while (0)
#define YFIELD(c, f)                                                                       \
    do {                                                                                   \
        if ((f) < 0 || (f) >= SY_NFIELDS) return yfail((c), SS_ERR_ARG, "bad field id"); \
    } while (0)
#define YGRID(c)                                                                               \
    do {                                                                                       \
        if (!(c)->grid_set) return yfail((c), SS_ERR_STATE, "ss_cyl_set_grid was not called"); \
    } while (0)

extern "C" {

int ss_cyl_create(int nr, int nth, int nz, int device, ss_cyl **out)
{
    if (!out) return SS_ERR_ARG;
    *out = nullptr;
    if (nr < 4 || nth < 4 || nz < 4 || nr > 4096 || nth > 4096 || nz > 4096) return SS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return (int)hipErrorNoDevice;
    if (device < 0 || device >= ndev) return SS_ERR_ARG;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    ss_cyl *c = new ss_cyl();
    c->device = device;
    YBox &b = c->box;
    b.nr = nr; b.nt = nth; b.nz = nz;
    b.pitch = ((nz + 1) + 63) / 64 * 64;
    b.sj = b.pitch;
    b.si = (long long)(nth + 1) * b.pitch;
    c->box_elems = (size_t)(nr + 1) * (size_t)b.si;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SS_ERR_NOMEM; }
    for (int f = 0; f < SY_NFIELDS; ++f) {
        if (hipMalloc(&c->p[f], c->box_elems * sizeof(double)) != hipSuccess) { ss_cyl_destroy(c); return SS_ERR_NOMEM; }
        hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, c->stream, c->p[f], c->box_elems);
    }
    const size_t st = (size_t)(nr + 1) * (nth + 1) * (nz + 1);
    if (hipMalloc(&c->staging, st * sizeof(double)) != hipSuccess || hipMalloc(&c->R, nr * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->th, nth * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->red, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&c->red_host, 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        ss_cyl_destroy(c);
        return SS_ERR_NOMEM;
    }
    for (auto &ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess) { ss_cyl_destroy(c); return SS_ERR_NOMEM; }
    int rc = ysync(c, "ss_cyl_create");
    if (rc != SS_OK) { ss_cyl_destroy(c); return rc; }
    *out = c;
    return SS_OK;
}

void ss_cyl_destroy(ss_cyl *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto q : c->p)
        if (q) (void)hipFree(q);
    if (c->staging) (void)hipFree(c->staging);
    if (c->R) (void)hipFree(c->R);
    if (c->th) (void)hipFree(c->th);
    if (c->red) (void)hipFree(c->red);
    if (c->red_host) (void)hipHostFree(c->red_host);
    for (auto ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *ss_cyl_last_error(const ss_cyl *c) { return c ? c->err.c_str() : "null context"; }

int ss_cyl_set_grid(ss_cyl *c, const double *R, const double *th)
{
    YCTX(c);
    if (!R || !th) return yfail(c, SS_ERR_ARG, "ss_cyl_set_grid: null");
    YHIP(c, hipMemcpyAsync(c->R, R, c->box.nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    YHIP(c, hipMemcpyAsync(c->th, th, c->box.nt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->grid_set = true;
    return ysync(c, "ss_cyl_set_grid");
}

int ss_cyl_upload(ss_cyl *c, int field, const double *host, int layout)
{
    YCTX(c);
    YFIELD(c, field);
    if (!host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return yfail(c, SS_ERR_ARG, "ss_cyl_upload");
    int d[3], sh;
    dims_of(c->box, field, d, &sh);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    YHIP(c, hipMemcpyAsync(c->staging, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_julia_to_box, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->box,
                           c->p[field], (const double *)c->staging, d[0], d[1], d[2], sh);
    else
        hipLaunchKernelGGL(k_c_to_box, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->p[field],
                           (const double *)c->staging, d[0], d[1], d[2], sh);
    return ysync(c, "ss_cyl_upload");
}

int ss_cyl_download(ss_cyl *c, int field, double *host, int layout)
{
    YCTX(c);
    YFIELD(c, field);
    if (!host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return yfail(c, SS_ERR_ARG, "ss_cyl_download");
    int d[3], sh;
    dims_of(c->box, field, d, &sh);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_box_to_julia, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->box,
                           c->staging, (const double *)c->p[field], d[0], d[1], d[2], sh);
    else
        hipLaunchKernelGGL(k_box_to_c, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->staging,
                           (const double *)c->p[field], d[0], d[1], d[2], sh);
    YHIP(c, hipMemcpyAsync(host, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return ysync(c, "ss_cyl_download");
}

static int same_shape(const ss_cyl *c, int a, int b)
{
    int d1[3], d2[3], s1, s2;
    dims_of(c->box, a, d1, &s1);
    dims_of(c->box, b, d2, &s2);
    return d1[0] == d2[0] && d1[1] == d2[1] && d1[2] == d2[2] && s1 == s2;
}

int ss_cyl_equal(ss_cyl *c, int dst, int src)
{
    YCTX(c);
    YFIELD(c, dst);
    YFIELD(c, src);
    if (!same_shape(c, dst, src)) return yfail(c, SS_ERR_ARG, "ss_cyl_equal: shapes differ");
    l_equal(c, dst, src);
    return ysync(c, "ss_cyl_equal");
}

int ss_cyl_multiply(ss_cyl *c, int dst, int src, double fact)
{
    YCTX(c);
    YFIELD(c, dst);
    YFIELD(c, src);
    if (!same_shape(c, dst, src)) return yfail(c, SS_ERR_ARG, "ss_cyl_multiply: shapes differ");
    l_equal(c, dst, src, true, fact);
    return ysync(c, "ss_cyl_multiply");
}

int ss_cyl_maxloc(ss_cyl *c) { YCTX(c); l_maxloc(c); return ysync(c, "ss_cyl_maxloc"); }
int ss_cyl_timesteps(ss_cyl *c, double dmp, double vpdt, double lr)
{
    YCTX(c);
    l_timesteps(c, dmp, vpdt, lr);
    return ysync(c, "ss_cyl_timesteps");
}
int ss_cyl_timesteps2(ss_cyl *c, double vpdt) { YCTX(c); l_timesteps2(c, vpdt); return ysync(c, "ss_cyl_timesteps2"); }
int ss_cyl_compute_P(ss_cyl *c, double dr, double dth, double dz)
{
    YCTX(c);
    YGRID(c);
    l_compute_P(c, dr, dth, dz);
    return ysync(c, "ss_cyl_compute_P");
}
int ss_cyl_compute_sigma(ss_cyl *c) { YCTX(c); l_compute_sigma(c); return ysync(c, "ss_cyl_compute_sigma"); }
int ss_cyl_compute_TII(ss_cyl *c) { YCTX(c); l_compute_TII(c); return ysync(c, "ss_cyl_compute_TII"); }
int ss_cyl_power_law(ss_cyl *c, double tau_C, double n_exp, double relax)
{
    YCTX(c);
    l_power_law(c, tau_C, n_exp, relax);
    return ysync(c, "ss_cyl_power_law");
}
int ss_cyl_compute_dV(ss_cyl *c, double dr, double dth, double dz)
{
    YCTX(c);
    YGRID(c);
    l_compute_dV(c, dr, dth, dz);
    return ysync(c, "ss_cyl_compute_dV");
}
int ss_cyl_compute_V(ss_cyl *c) { YCTX(c); l_compute_V(c); return ysync(c, "ss_cyl_compute_V"); }
int ss_cyl_bc(ss_cyl *c, int field, int op, double fact, double r)
{
    YCTX(c);
    YFIELD(c, field);
    if (!valid_bc(c, field, op)) return yfail(c, SS_ERR_ARG, "ss_cyl_bc: operation not defined on this field");
    if ((op == SY_BC_TH_FACT || op == SY_BC_Z_LIN) && !c->grid_set) return yfail(c, SS_ERR_STATE, "ss_cyl_set_grid was not called");
    l_bc(c, field, op, fact, r);
    return ysync(c, "ss_cyl_bc");
}
int ss_cyl_err_rth(ss_cyl *c) { YCTX(c); l_err_rt(c); return ysync(c, "ss_cyl_err_rth"); }
int ss_cyl_exit_error(ss_cyl *c, double *err, double comp[3])
{
    YCTX(c);
    if (!err) return yfail(c, SS_ERR_ARG, "ss_cyl_exit_error: null");
    return exit_error(c, err, comp);
}

int ss_cyl_pt_solve(ss_cyl *c, const ss_cyl_params *p, int maxiter, int nout, double eps, ss_cyl_result *out)
{
    YCTX(c);
    YGRID(c);
    if (!p || !out || maxiter < 1 || nout < 1) return yfail(c, SS_ERR_ARG, "ss_cyl_pt_solve: bad argument");
    memset(out, 0, sizeof *out);
    out->err = NAN;
    const yclk::time_point t0 = yclk::now();
    double err = 1.0, n_exp = p->n_exp;   // err starts at 1.0 (:325)
    for (int iter = 1; iter <= maxiter; ++iter) {
        if (err < 1e-5) n_exp = p->n_exp_PL;  // :342-344, from the last exit-test error
        l_iteration(c, *p, n_exp);
        if (iter % nout == 0) {
            int rc = exit_error(c, &err, nullptr);
            if (rc != SS_OK) return rc;
            l_err_rt(c);
            out->err = err;
            out->n_exp = n_exp;
            if (std::isnan(err)) {
                out->iters = iter;
                out->seconds = std::chrono::duration<double>(yclk::now() - t0).count();
                return yfail(c, SS_ERR_NONFINITE, "NaN in the exit test");
            }
            if (err < eps && iter > 20) {   // :375-377 (iters = iter)
                out->iters = iter;
                out->converged = 1;
                break;
            }
        }
    }
    if (!out->converged) out->iters = maxiter;
    out->n_exp = n_exp;
    int rc = ysync(c, "ss_cyl_pt_solve");
    out->seconds = std::chrono::duration<double>(yclk::now() - t0).count();
    if (rc != SS_OK) return rc;
    return out->converged ? SS_OK : SS_NOT_CONVERGED;
}

int ss_cyl_pt_iterate_timed(ss_cyl *c, const ss_cyl_params *p, int niter, double *iter_ms, double *kernel_ms)
{
    YCTX(c);
    YGRID(c);
    if (!p || niter < 1) return yfail(c, SS_ERR_ARG, "ss_cyl_pt_iterate_timed: bad argument");
    int rc = ysync(c, "ss_cyl_pt_iterate_timed");
    if (rc != SS_OK) return rc;
    std::vector<hipEvent_t> ev(2 * (size_t)niter, nullptr);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            for (auto x : ev)
                if (x) (void)hipEventDestroy(x);
            return yfail(c, SS_ERR_NOMEM, "ss_cyl_pt_iterate_timed: events");
        }
    (void)hipEventRecord(c->ev[0], c->stream);
    for (int it = 0; it < niter; ++it) l_iteration(c, *p, p->n_exp, ev[2 * it], ev[2 * it + 1]);
    (void)hipEventRecord(c->ev[1], c->stream);
    rc = ysync(c, "ss_cyl_pt_iterate_timed");
    double kms = 0.0;
    float total = 0.0f;
    if (rc == SS_OK) {
        for (int it = 0; it < niter; ++it) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[2 * it], ev[2 * it + 1]);
            kms += ms;
        }
        (void)hipEventElapsedTime(&total, c->ev[0], c->ev[1]);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    if (rc != SS_OK) return rc;
    if (iter_ms) *iter_ms = total / niter;
    if (kernel_ms) *kernel_ms = kms / niter;
    return SS_OK;
}

} // extern "C"
