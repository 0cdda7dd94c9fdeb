// sspt_cart.hip -- the Cartesian twin of the diapir solver
// (scripts_diapir/CartesianStokes_diapir.jl), kernel for kernel.
//
// Same device layout as the spherical solver (sspt_internal.h): every field
// in a box of (nx+1) x (ny+1) x pitch doubles, z fastest, pitch a multiple
// of 64, so a wavefront reads 64 consecutive z points.  Centre fields and
// the velocities use box index = reference index - 1; the shear stresses
// tau_XY, tau_XZ, tau_YZ use box index = reference index, which puts an
// edge at the face index of its staggered directions and makes every stencil
// below a +-1 offset in one frame.  One thread per output point, with the
// reference's guard ranges; each reference statement keeps its association
// (n-ary + and * of Julia are left folds), FP contraction is off, and x^2 /
// vpdt^2 are single products as Julia's literal powers are.
#include "sspt_internal.h"
#include "../../include/sspt_cart.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace sspt {
namespace cart {

constexpr double PI_ = 3.14159265358979323846;

struct CBox {
    int nx, ny, nz, pitch;
    long long sj, si;
    __host__ __device__ inline long long off(int i, int j, int k) const
    {
        return (long long)i * si + (long long)j * sj + k;
    }
};

struct Fields {
    double *f[SC_NFIELDS];
};

// reference sizes of a field and the box offset of its first entry
inline void dims_of(const CBox &b, int f, int d[3], int *o1)
{
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    *o1 = 0;
    switch (f) {
    case SC_VX: case SC_ERR_VX: d[0] = nx + 1; d[1] = ny; d[2] = nz; break;
    case SC_VY: case SC_ERR_VY: d[0] = nx; d[1] = ny + 1; d[2] = nz; break;
    case SC_VZ: case SC_ERR_VZ: d[0] = nx; d[1] = ny; d[2] = nz + 1; break;
    case SC_TAU_XY: d[0] = nx - 1; d[1] = ny - 1; d[2] = nz - 2; *o1 = 1; break;
    case SC_TAU_XZ: d[0] = nx - 1; d[1] = ny - 2; d[2] = nz - 1; *o1 = 1; break;
    case SC_TAU_YZ: d[0] = nx - 2; d[1] = ny - 1; d[2] = nz - 1; *o1 = 1; break;
    default: d[0] = nx; d[1] = ny; d[2] = nz; break;
    }
}

static const dim3 B3(64, 4, 1);
static inline dim3 grid3(int n1, int n2, int n3) { return dim3((n3 + 63) / 64, (n2 + 3) / 4, n1); }

#define C_LOOP(N1, N2, N3)                                        \
    const int k = blockIdx.x * blockDim.x + threadIdx.x;          \
    const int j = blockIdx.y * blockDim.y + threadIdx.y;          \
    const int i = blockIdx.z;                                     \
    if (i >= (N1) || j >= (N2) || k >= (N3)) return;

// Julia's max: NaN if either argument is NaN
__device__ __forceinline__ double jl_max(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}

// ---------------------------------------------------------------- transposes
__global__ void k_julia_to_box(CBox b, double *dst, const double *src, int n1, int n2, int n3, int o1)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)n1 * n2 * n3;
    if (t >= n) return;
    const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long long)n1 * n2));
    dst[b.off(i + o1, j + o1, k + o1)] = src[t];
}
__global__ void k_box_to_julia(CBox b, double *dst, const double *src, int n1, int n2, int n3, int o1)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)n1 * n2 * n3;
    if (t >= n) return;
    const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long long)n1 * n2));
    dst[t] = src[b.off(i + o1, j + o1, k + o1)];
}
__global__ void k_c_to_box(CBox b, double *dst, const double *src, int n1, int n2, int n3, int o1)
{
    C_LOOP(n1, n2, n3)
    dst[b.off(i + o1, j + o1, k + o1)] = src[((long long)i * n2 + j) * n3 + k];
}
__global__ void k_box_to_c(CBox b, double *dst, const double *src, int n1, int n2, int n3, int o1)
{
    C_LOOP(n1, n2, n3)
    dst[((long long)i * n2 + j) * n3 + k] = src[b.off(i + o1, j + o1, k + o1)];
}
__global__ void k_zero(double *p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

// ---------------------------------------------------------------- kernels
// equal!(A, B) (CartesianStokes_diapir.jl:41-44), over A's size
__global__ void k_equal(CBox b, double *A, const double *B, int n1, int n2, int n3, int o1)
{
    C_LOOP(n1, n2, n3)
    const long long o = b.off(i + o1, j + o1, k + o1);
    A[o] = B[o];
}

// maxloc!(etaSM, eta) (:125-128): @inn(etaSM) = @maxloc(eta), the 7-point
// star with ParallelStencil's nesting (as written out in
// SphericalStokes_mxpu.jl:295-297)
__global__ void k_maxloc(CBox b, double *SM, const double *eta)
{
    C_LOOP(b.nx - 2, b.ny - 2, b.nz - 2)
    const long long o = b.off(i + 1, j + 1, k + 1);
    SM[o] = jl_max(jl_max(jl_max(jl_max(eta[o - b.si], eta[o + b.si]), eta[o]), jl_max(eta[o - b.sj], eta[o + b.sj])),
                   jl_max(eta[o - 1], eta[o + 1]));
}

// copy_boundary!(A) (:97-105): the six faces of a centre array without
// their edges; family f = 0 (x faces over (j, k)), 1 (y faces over (i, k)),
// 2 (z faces over (i, j)), both sides per thread
__global__ void k_copy_boundary(CBox b, double *A)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y * blockDim.y + threadIdx.y;
    const int f = blockIdx.z, nx = b.nx, ny = b.ny, nz = b.nz;
    if (f == 0) {
        const int j = v + 1, k = u + 1;
        if (j > ny - 2 || k > nz - 2) return;
        A[b.off(0, j, k)] = A[b.off(1, j, k)];
        A[b.off(nx - 1, j, k)] = A[b.off(nx - 2, j, k)];
    } else if (f == 1) {
        const int i = v + 1, k = u + 1;
        if (i > nx - 2 || k > nz - 2) return;
        A[b.off(i, 0, k)] = A[b.off(i, 1, k)];
        A[b.off(i, ny - 1, k)] = A[b.off(i, ny - 2, k)];
    } else {
        const int i = v + 1, j = u + 1;
        if (i > nx - 2 || j > ny - 2) return;
        A[b.off(i, j, 0)] = A[b.off(i, j, 1)];
        A[b.off(i, j, nz - 1)] = A[b.off(i, j, nz - 2)];
    }
}

// @KBDT and @GSDT without the etaSM factor (:130-131): dmp*2.0*pi*vpdt/lx
// and 4.0*pi*vpdt/lx, left-folded as Julia evaluates them
struct Steps {
    double kb, gs;
};
inline Steps steps_of(double dmp, double vpdt, double lx)
{
    return Steps{(((dmp * 2.0) * PI_) * vpdt) / lx, ((4.0 * PI_) * vpdt) / lx};
}

// timesteps! (:132-139)
__global__ void k_timesteps(CBox b, double *DT, const double *eta, const double *SM, double vpdt2, Steps st)
{
    C_LOOP(b.nx, b.ny, b.nz)
    const long long o = b.off(i, j, k);
    const double kb = st.kb * SM[o], gs = st.gs * SM[o];
    DT[o] = vpdt2 / (kb + gs / (1.0 + gs / eta[o]));
}

struct PArgs {
    double *P, *tXX, *tYY, *tZZ, *tXY, *tXZ, *tYZ;
    const double *VX, *VY, *VZ, *eta, *SM;
    Steps st;
    double _dx, _dy, _dz;
};

// compute_P! (:165-193): one thread per box point; each of the four output
// families under its own guard
__global__ void k_compute_P(CBox b, PArgs a)
{
    C_LOOP(b.nx, b.ny, b.nz)
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    const long long o = b.off(i, j, k), si = b.si, sj = b.sj;
    {   // centres: P, tau_XX, tau_YY, tau_ZZ
        const double kb = a.st.kb * a.SM[o], gs = a.st.gs * a.SM[o];
        const double dxx = (a.VX[o + si] - a.VX[o]) * a._dx;
        const double dyy = (a.VY[o + sj] - a.VY[o]) * a._dy;
        const double dzz = (a.VZ[o + 1] - a.VZ[o]) * a._dz;
        const double den = 1.0 + gs / a.eta[o];
        a.P[o] = a.P[o] - kb * ((dxx + dyy) + dzz);
        a.tXX[o] = (a.tXX[o] + (gs * 2.0) * dxx) / den;
        a.tYY[o] = (a.tYY[o] + (gs * 2.0) * dyy) / den;
        a.tZZ[o] = (a.tZZ[o] + (gs * 2.0) * dzz) / den;
    }
    // edges (box index = reference index): tau_XY on [1,nx-1]x[1,ny-1]x[1,nz-2]
    if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 2) {
        const long long c00 = o - si - sj, c01 = o - si, c10 = o - sj;
        const double g4 = a.st.gs * (0.25 * (((a.SM[c00] + a.SM[c01]) + a.SM[c10]) + a.SM[o]));
        const double e4 = 0.25 * (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]);
        const double d = (a.VX[o] - a.VX[o - sj]) * a._dy + (a.VY[o] - a.VY[o - si]) * a._dx;
        a.tXY[o] = (a.tXY[o] + g4 * d) / (1.0 + g4 / e4);
    }
    // tau_XZ on [1,nx-1]x[1,ny-2]x[1,nz-1]
    if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 1) {
        const long long c00 = o - si - 1, c01 = o - si, c10 = o - 1;
        const double g4 = a.st.gs * (0.25 * (((a.SM[c00] + a.SM[c01]) + a.SM[c10]) + a.SM[o]));
        const double e4 = 0.25 * (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]);
        const double d = (a.VX[o] - a.VX[o - 1]) * a._dz + (a.VZ[o] - a.VZ[o - si]) * a._dx;
        a.tXZ[o] = (a.tXZ[o] + g4 * d) / (1.0 + g4 / e4);
    }
    // tau_YZ on [1,nx-2]x[1,ny-1]x[1,nz-1]
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 1) {
        const long long c00 = o - sj - 1, c01 = o - sj, c10 = o - 1;
        const double g4 = a.st.gs * (0.25 * (((a.SM[c00] + a.SM[c01]) + a.SM[c10]) + a.SM[o]));
        const double e4 = 0.25 * (((a.eta[c00] + a.eta[c01]) + a.eta[c10]) + a.eta[o]);
        const double d = (a.VY[o] - a.VY[o - 1]) * a._dz + (a.VZ[o] - a.VZ[o - sj]) * a._dy;
        a.tYZ[o] = (a.tYZ[o] + g4 * d) / (1.0 + g4 / e4);
    }
}

// compute_TII! (:195-208) on the inner centres
__global__ void k_compute_TII(CBox b, double *TII, const double *tXX, const double *tYY, const double *tZZ,
                              const double *tXY, const double *tXZ, const double *tYZ)
{
    C_LOOP(b.nx - 2, b.ny - 2, b.nz - 2)
    const long long o = b.off(i + 1, j + 1, k + 1), si = b.si, sj = b.sj;
    const double a = tXX[o], c = tYY[o], d = tZZ[o];
    const double x = 0.25 * (((tXY[o] + tXY[o + sj]) + tXY[o + si]) + tXY[o + si + sj]);
    const double y = 0.25 * (((tXZ[o] + tXZ[o + 1]) + tXZ[o + si]) + tXZ[o + si + 1]);
    const double z = 0.25 * (((tYZ[o] + tYZ[o + 1]) + tYZ[o + sj]) + tYZ[o + sj + 1]);
    TII[o] = sqrt((((0.5 * ((a * a + c * c) + d * d)) + x * x) + y * y) + z * z);
}

// power_law! (:210-221): eta_PL = eta_INI (tII/tau_C)^(1-n_exp), relaxed in
// log space against eta_PL_OLD, then the harmonic combination
__global__ void k_power_law(CBox b, double *ePL, double *eta, const double *TII, const double *old,
                            const double *ini, double tC, double ex, double relax)
{
    C_LOOP(b.nx, b.ny, b.nz)
    const long long o = b.off(i, j, k);
    double pl = ini[o] * pow(TII[o] / tC, ex);
    pl = exp(log(pl) * relax + log(old[o]) * (1.0 - relax));
    ePL[o] = pl;
    eta[o] = 2.0 / (1.0 / ini[o] + 1.0 / pl);
}

struct VArgs {
    double *VX, *VY, *VZ;
    const double *tXX, *tYY, *tZZ, *tXY, *tXZ, *tYZ, *DT, *P, *rho;
    double _dx, _dy, _dz, g;
};

// compute_V! (:223-255): VX on [1,nx-1]x[1,ny-2]x[1,nz-2], VY on
// [1,nx-2]x[1,ny-1]x[1,nz-2], VZ on [1,nx-2]x[1,ny-2]x[1,nz-1] (box indices)
__global__ void k_compute_V(CBox b, VArgs a)
{
    C_LOOP(b.nx, b.ny, b.nz)
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    const long long o = b.off(i, j, k), si = b.si, sj = b.sj;
    if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
        const double d = (((-(a.P[o] - a.P[o - si])) * a._dx + (a.tXX[o] - a.tXX[o - si]) * a._dx) +
                          (a.tXY[o + sj] - a.tXY[o]) * a._dy) +
                         (a.tXZ[o + 1] - a.tXZ[o]) * a._dz;
        a.VX[o] = a.VX[o] + d * (0.5 * (a.DT[o - si] + a.DT[o]));
    }
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 2) {
        const double d = (((a.tXY[o + si] - a.tXY[o]) * a._dx - (a.P[o] - a.P[o - sj]) * a._dy) +
                          (a.tYY[o] - a.tYY[o - sj]) * a._dy) +
                         (a.tYZ[o + 1] - a.tYZ[o]) * a._dz;
        a.VY[o] = a.VY[o] + d * (0.5 * (a.DT[o - sj] + a.DT[o]));
    }
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 1) {
        const double d = ((((a.tXZ[o + si] - a.tXZ[o]) * a._dx + (a.tYZ[o + sj] - a.tYZ[o]) * a._dy) -
                           (a.P[o] - a.P[o - 1]) * a._dz) +
                          (a.tZZ[o] - a.tZZ[o - 1]) * a._dz) -
                         a.g * (0.5 * (a.rho[o - 1] + a.rho[o]));
        a.VZ[o] = a.VZ[o] + d * (0.5 * (a.DT[o - 1] + a.DT[o]));
    }
}

// boundary kernels (:52-95) on a velocity field of reference size n1 x n2 x
// n3; ghost layers at both ends of one axis, both sides per thread
enum { BC_X_LIN, BC_X_0, COPY_BC_X, BC_Y_FACT, BC_Y_0, BC_Z_0, COPY_BC_Z };
__global__ void k_bc(CBox b, double *A, int n1, int n2, int n3, int op, const double *y, double fact)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y * blockDim.y + threadIdx.y;
    if (op == BC_X_LIN || op == BC_X_0 || op == COPY_BC_X) {
        const int j = v, k = u;
        if (j >= n2 || k >= n3) return;
        const long long lo = b.off(0, j, k), hi = b.off(n1 - 1, j, k);
        if (op == BC_X_LIN) {
            A[lo] = (2.0 * y[j]) * fact - A[lo + b.si];
            A[hi] = (2.0 * y[j]) * fact - A[hi - b.si];
        } else if (op == BC_X_0) {
            A[lo] = 0.0;
            A[hi] = 0.0;
        } else {
            A[lo] = A[lo + b.si];
            A[hi] = A[hi - b.si];
        }
    } else if (op == BC_Y_FACT || op == BC_Y_0) {
        const int i = v, k = u;
        if (i >= n1 || k >= n3) return;
        const long long lo = b.off(i, 0, k), hi = b.off(i, n2 - 1, k);
        if (op == BC_Y_FACT) {
            A[lo] = fact * y[0];
            A[hi] = fact * y[n2 - 1];
        } else {
            A[lo] = -A[lo + b.sj];
            A[hi] = -A[hi - b.sj];
        }
    } else {
        const int i = v, j = u;
        if (i >= n1 || j >= n2) return;
        const long long lo = b.off(i, j, 0), hi = b.off(i, j, n3 - 1);
        if (op == BC_Z_0) {
            A[lo] = -A[lo + 1];
            A[hi] = -A[hi - 1];
        } else {
            A[lo] = A[lo + 1];
            A[hi] = A[hi - 1];
        }
    }
}

// error!(A, B) (:257-260) with the two maxima of the exit test: A = |A - B|,
// then max A and max |B| as bit patterns (non-negative doubles order like
// their bits; NaN sorts above +inf, so a NaN propagates as in maximum())
__global__ void __launch_bounds__(256) k_error(CBox b, double *A, const double *B, int n1, int n2, int n3,
                                               unsigned long long *red)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    unsigned long long me = 0, mv = 0;
    if (i < n1 && j < n2 && k < n3) {
        const long long o = b.off(i, j, k);
        const double e = fabs(A[o] - B[o]);
        A[o] = e;
        me = (unsigned long long)__double_as_longlong(e);
        mv = (unsigned long long)__double_as_longlong(fabs(B[o]));
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

} // namespace cart
} // namespace sspt

using namespace sspt::cart;

struct ss_cart {
    int device = 0;
    CBox box{};
    size_t box_elems = 0;
    hipStream_t stream = nullptr;
    Fields cur{};
    double *y = nullptr;            // y coordinates of the cell centres
    bool y_set = false;
    double *staging = nullptr;
    unsigned long long *red = nullptr;
    unsigned long long *red_host = nullptr;
    hipEvent_t ev[4] = {};
    std::string err;
};

namespace {

int cfail(ss_cart *c, int code, const std::string &msg)
{
    c->err = msg;
    return code;
}

int csync(ss_cart *c, const char *where)
{
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return cfail(c, (int)e, std::string(where) + ": " + hipGetErrorString(e));
    return SS_OK;
}

#define CHIP_TRY(c, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return cfail((c), (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool is_velocity(int f)
{
    return f == SC_VX || f == SC_VY || f == SC_VZ || f == SC_ERR_VX || f == SC_ERR_VY || f == SC_ERR_VZ;
}

bool is_centre(int f)
{
    return f == SC_P || f == SC_TAU_XX || f == SC_TAU_YY || f == SC_TAU_ZZ || (f >= SC_RHO && f < SC_NFIELDS);
}

void launch_bc(ss_cart *c, int f, int op, double fact)
{
    int d[3], o1;
    dims_of(c->box, f, d, &o1);
    int a, bdim;
    if (op == BC_X_LIN || op == BC_X_0 || op == COPY_BC_X) { a = d[2]; bdim = d[1]; }
    else if (op == BC_Y_FACT || op == BC_Y_0) { a = d[2]; bdim = d[0]; }
    else { a = d[1]; bdim = d[0]; }
    hipLaunchKernelGGL(k_bc, dim3((a + 63) / 64, (bdim + 3) / 4), B3, 0, c->stream, c->box, c->cur.f[f], d[0], d[1],
                       d[2], op, (const double *)c->y, fact);
}

void launch_equal(ss_cart *c, int dst, int src)
{
    int d[3], o1;
    dims_of(c->box, dst, d, &o1);
    hipLaunchKernelGGL(k_equal, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->cur.f[dst],
                       (const double *)c->cur.f[src], d[0], d[1], d[2], o1);
}

void launch_maxloc(ss_cart *c)
{
    const CBox &b = c->box;
    hipLaunchKernelGGL(k_maxloc, grid3(b.nx - 2, b.ny - 2, b.nz - 2), B3, 0, c->stream, b, c->cur.f[SC_ETASM],
                       (const double *)c->cur.f[SC_ETA]);
}

void launch_copy_boundary(ss_cart *c, int f)
{
    const CBox &b = c->box;
    const int m = std::max(std::max(b.nx, b.ny), b.nz);
    hipLaunchKernelGGL(k_copy_boundary, dim3((m + 63) / 64, (m + 3) / 4, 3), B3, 0, c->stream, b, c->cur.f[f]);
}

void launch_timesteps(ss_cart *c, double vpdt, double dmp, double lx)
{
    const CBox &b = c->box;
    hipLaunchKernelGGL(k_timesteps, grid3(b.nx, b.ny, b.nz), B3, 0, c->stream, b, c->cur.f[SC_DT_R],
                       (const double *)c->cur.f[SC_ETA], (const double *)c->cur.f[SC_ETASM], vpdt * vpdt,
                       steps_of(dmp, vpdt, lx));
}

void launch_compute_P(ss_cart *c, double dmp, double vpdt, double lx, double _dx, double _dy, double _dz)
{
    const CBox &b = c->box;
    double **f = c->cur.f;
    PArgs a{f[SC_P], f[SC_TAU_XX], f[SC_TAU_YY], f[SC_TAU_ZZ], f[SC_TAU_XY], f[SC_TAU_XZ], f[SC_TAU_YZ],
            f[SC_VX], f[SC_VY], f[SC_VZ], f[SC_ETA], f[SC_ETASM], steps_of(dmp, vpdt, lx), _dx, _dy, _dz};
    hipLaunchKernelGGL(k_compute_P, grid3(b.nx, b.ny, b.nz), B3, 0, c->stream, b, a);
}

void launch_compute_TII(ss_cart *c)
{
    const CBox &b = c->box;
    double **f = c->cur.f;
    hipLaunchKernelGGL(k_compute_TII, grid3(b.nx - 2, b.ny - 2, b.nz - 2), B3, 0, c->stream, b, f[SC_TII],
                       (const double *)f[SC_TAU_XX], (const double *)f[SC_TAU_YY], (const double *)f[SC_TAU_ZZ],
                       (const double *)f[SC_TAU_XY], (const double *)f[SC_TAU_XZ], (const double *)f[SC_TAU_YZ]);
}

void launch_power_law(ss_cart *c, double tC, double n_exp, double relax)
{
    const CBox &b = c->box;
    double **f = c->cur.f;
    hipLaunchKernelGGL(k_power_law, grid3(b.nx, b.ny, b.nz), B3, 0, c->stream, b, f[SC_ETA_PL], f[SC_ETA],
                       (const double *)f[SC_TII], (const double *)f[SC_ETA_PL_OLD], (const double *)f[SC_ETA_INI], tC,
                       1.0 - n_exp, relax);
}

void launch_compute_V(ss_cart *c, double _dx, double _dy, double _dz, double g)
{
    const CBox &b = c->box;
    double **f = c->cur.f;
    VArgs a{f[SC_VX], f[SC_VY], f[SC_VZ], f[SC_TAU_XX], f[SC_TAU_YY], f[SC_TAU_ZZ], f[SC_TAU_XY], f[SC_TAU_XZ],
            f[SC_TAU_YZ], f[SC_DT_R], f[SC_P], f[SC_RHO], _dx, _dy, _dz, g};
    hipLaunchKernelGGL(k_compute_V, grid3(b.nx, b.ny, b.nz), B3, 0, c->stream, b, a);
}

// error! on the three components; c->red[2q], c->red[2q+1] = max ERR, max |V|
void launch_error(ss_cart *c)
{
    hipLaunchKernelGGL(k_zero_red, dim3(1), dim3(64), 0, c->stream, c->red, 6);
    static const int E[3] = {SC_ERR_VX, SC_ERR_VY, SC_ERR_VZ}, V[3] = {SC_VX, SC_VY, SC_VZ};
    for (int q = 0; q < 3; ++q) {
        int d[3], o1;
        dims_of(c->box, V[q], d, &o1);
        hipLaunchKernelGGL(k_error, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->cur.f[E[q]],
                           (const double *)c->cur.f[V[q]], d[0], d[1], d[2], c->red + 2 * q);
    }
}

int read_err(ss_cart *c, double *err, double comp[3])
{
    CHIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    int rc = csync(c, "exit test");
    if (rc != SS_OK) return rc;
    double e = 0.0;
    bool nan = false;
    for (int q = 0; q < 3; ++q) {
        double me, mv;
        memcpy(&me, &c->red_host[2 * q], 8);
        memcpy(&mv, &c->red_host[2 * q + 1], 8);
        const double r = me / mv;   // maximum(ERR)./maximum(abs.(V))
        if (comp) comp[q] = r;
        if (std::isnan(r)) nan = true;
        else if (q == 0 || r > e) e = r;
    }
    *err = nan ? NAN : e;  // max(err_vx, err_vy, err_vz) propagates NaN
    return SS_OK;
}

// the iteration body of the reference loop (:351-379), with the copies of
// the velocities before compute_V! on check iterations
void launch_iteration(ss_cart *c, const ss_cart_params &p, bool check, hipEvent_t e0 = nullptr,
                      hipEvent_t e1 = nullptr)
{
    launch_equal(c, SC_ETASM, SC_ETA);
    launch_maxloc(c);
    launch_copy_boundary(c, SC_ETASM);
    launch_timesteps(c, p.vpdt, p.dmp, p.lx);
    if (e0) (void)hipEventRecord(e0, c->stream);
    launch_compute_P(c, p.dmp, p.vpdt, p.lx, p._dx, p._dy, p._dz);
    if (e1) (void)hipEventRecord(e1, c->stream);
    if (p.n_exp > 1.0) {
        launch_compute_TII(c);
        launch_copy_boundary(c, SC_TII);
        launch_equal(c, SC_ETA_PL_OLD, SC_ETA_PL);
        launch_power_law(c, p.tau_C, p.n_exp, p.relax);
    }
    if (check) {
        launch_equal(c, SC_ERR_VX, SC_VX);
        launch_equal(c, SC_ERR_VY, SC_VY);
        launch_equal(c, SC_ERR_VZ, SC_VZ);
    }
    launch_compute_V(c, p._dx, p._dy, p._dz, p.g);
    launch_bc(c, SC_VX, BC_X_LIN, p.sh);
    launch_bc(c, SC_VY, BC_X_0, 0.0);
    launch_bc(c, SC_VZ, BC_X_0, 0.0);
    launch_bc(c, SC_VX, BC_Y_FACT, p.sh);
    launch_bc(c, SC_VY, BC_Y_0, 0.0);
    launch_bc(c, SC_VZ, BC_Y_FACT, 0.0);
    launch_bc(c, SC_VX, COPY_BC_Z, 0.0);
    launch_bc(c, SC_VY, COPY_BC_Z, 0.0);
    launch_bc(c, SC_VZ, BC_Z_0, 0.0);
}

using cclk = std::chrono::steady_clock;

} // namespace

#define CTX_OK(c)                                       \
    do {                                                \
        if (!(c)) return SS_ERR_ARG;                    \
        CHIP_TRY((c), hipSetDevice((c)->device));       \
    } while (0)
#define FIELD_OK(c, f)                                                                    \
    do {                                                                                  \
        if ((f) < 0 || (f) >= SC_NFIELDS) return cfail((c), SS_ERR_ARG, "bad field id"); \
    } while (0)

extern "C" {

int ss_cart_create(int nx, int ny, int nz, int device, ss_cart **out)
{
    if (!out) return SS_ERR_ARG;
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3 || nx > 4096 || ny > 4096 || nz > 4096) return SS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return (int)hipErrorNoDevice;
    if (device < 0 || device >= ndev) return SS_ERR_ARG;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    ss_cart *c = new ss_cart();
    c->device = device;
    CBox &b = c->box;
    b.nx = nx; b.ny = ny; b.nz = nz;
    b.pitch = ((nz + 1) + 63) / 64 * 64;
    b.sj = b.pitch;
    b.si = (long long)(ny + 1) * b.pitch;
    c->box_elems = (size_t)(nx + 1) * (size_t)b.si;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SS_ERR_NOMEM; }
    for (int f = 0; f < SC_NFIELDS; ++f) {
        if (hipMalloc(&c->cur.f[f], c->box_elems * sizeof(double)) != hipSuccess) {
            ss_cart_destroy(c);
            return SS_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, c->stream, c->cur.f[f], c->box_elems);
    }
    const size_t st = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    if (hipMalloc(&c->staging, st * sizeof(double)) != hipSuccess || hipMalloc(&c->y, (ny + 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->red, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipHostMalloc(&c->red_host, 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        ss_cart_destroy(c);
        return SS_ERR_NOMEM;
    }
    for (auto &ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess) { ss_cart_destroy(c); return SS_ERR_NOMEM; }
    int rc = csync(c, "ss_cart_create");
    if (rc != SS_OK) { ss_cart_destroy(c); return rc; }
    *out = c;
    return SS_OK;
}

void ss_cart_destroy(ss_cart *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto p : c->cur.f)
        if (p) (void)hipFree(p);
    if (c->staging) (void)hipFree(c->staging);
    if (c->y) (void)hipFree(c->y);
    if (c->red) (void)hipFree(c->red);
    if (c->red_host) (void)hipHostFree(c->red_host);
    for (auto ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *ss_cart_last_error(const ss_cart *c) { return c ? c->err.c_str() : "null context"; }

int ss_cart_set_y(ss_cart *c, const double *y)
{
    CTX_OK(c);
    if (!y) return cfail(c, SS_ERR_ARG, "ss_cart_set_y: null");
    CHIP_TRY(c, hipMemcpyAsync(c->y, y, c->box.ny * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->y_set = true;
    return csync(c, "ss_cart_set_y");
}

int ss_cart_upload(ss_cart *c, int field, const double *host, int layout)
{
    CTX_OK(c);
    FIELD_OK(c, field);
    if (!host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return cfail(c, SS_ERR_ARG, "ss_cart_upload");
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    CHIP_TRY(c, hipMemcpyAsync(c->staging, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_julia_to_box, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->box,
                           c->cur.f[field], (const double *)c->staging, d[0], d[1], d[2], o1);
    else
        hipLaunchKernelGGL(k_c_to_box, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->cur.f[field],
                           (const double *)c->staging, d[0], d[1], d[2], o1);
    return csync(c, "ss_cart_upload");
}

int ss_cart_download(ss_cart *c, int field, double *host, int layout)
{
    CTX_OK(c);
    FIELD_OK(c, field);
    if (!host || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C)) return cfail(c, SS_ERR_ARG, "ss_cart_download");
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    if (layout == SS_LAYOUT_JULIA)
        hipLaunchKernelGGL(k_box_to_julia, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->box,
                           c->staging, (const double *)c->cur.f[field], d[0], d[1], d[2], o1);
    else
        hipLaunchKernelGGL(k_box_to_c, grid3(d[0], d[1], d[2]), B3, 0, c->stream, c->box, c->staging,
                           (const double *)c->cur.f[field], d[0], d[1], d[2], o1);
    CHIP_TRY(c, hipMemcpyAsync(host, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return csync(c, "ss_cart_download");
}

int ss_cart_equal(ss_cart *c, int dst, int src)
{
    CTX_OK(c);
    FIELD_OK(c, dst);
    FIELD_OK(c, src);
    int d1[3], d2[3], o1, o2;
    dims_of(c->box, dst, d1, &o1);
    dims_of(c->box, src, d2, &o2);
    if (d1[0] != d2[0] || d1[1] != d2[1] || d1[2] != d2[2] || o1 != o2)
        return cfail(c, SS_ERR_ARG, "ss_cart_equal: shapes differ");
    launch_equal(c, dst, src);
    return csync(c, "ss_cart_equal");
}

int ss_cart_maxloc(ss_cart *c)
{
    CTX_OK(c);
    launch_maxloc(c);
    return csync(c, "ss_cart_maxloc");
}

int ss_cart_copy_boundary(ss_cart *c, int field)
{
    CTX_OK(c);
    FIELD_OK(c, field);
    if (!is_centre(field)) return cfail(c, SS_ERR_ARG, "ss_cart_copy_boundary: centre fields only");
    launch_copy_boundary(c, field);
    return csync(c, "ss_cart_copy_boundary");
}

int ss_cart_timesteps(ss_cart *c, double vpdt, double dmp, double lx)
{
    CTX_OK(c);
    launch_timesteps(c, vpdt, dmp, lx);
    return csync(c, "ss_cart_timesteps");
}

int ss_cart_compute_P(ss_cart *c, double dmp, double vpdt, double lx, double _dx, double _dy, double _dz)
{
    CTX_OK(c);
    launch_compute_P(c, dmp, vpdt, lx, _dx, _dy, _dz);
    return csync(c, "ss_cart_compute_P");
}

int ss_cart_compute_TII(ss_cart *c)
{
    CTX_OK(c);
    launch_compute_TII(c);
    return csync(c, "ss_cart_compute_TII");
}

int ss_cart_power_law(ss_cart *c, double tau_C, double n_exp, double relax)
{
    CTX_OK(c);
    launch_power_law(c, tau_C, n_exp, relax);
    return csync(c, "ss_cart_power_law");
}

int ss_cart_compute_V(ss_cart *c, double _dx, double _dy, double _dz, double g)
{
    CTX_OK(c);
    launch_compute_V(c, _dx, _dy, _dz, g);
    return csync(c, "ss_cart_compute_V");
}

static int bc_call(ss_cart *c, int field, int op, double fact, const char *what)
{
    CTX_OK(c);
    FIELD_OK(c, field);
    if (!is_velocity(field)) return cfail(c, SS_ERR_ARG, std::string(what) + ": velocity fields only");
    if (op == BC_X_LIN || op == BC_Y_FACT) {
        // the coordinate array is y at cell centres (Y_VX, Y_VZ): fields
        // whose y extent is ny
        if (field == SC_VY || field == SC_ERR_VY)
            return cfail(c, SS_ERR_ARG, std::string(what) + ": VX, VZ (y at cell centres) only");
        if (!c->y_set) return cfail(c, SS_ERR_STATE, "ss_cart_set_y was not called");
    }
    launch_bc(c, field, op, fact);
    return csync(c, what);
}

int ss_cart_bc_x_lin(ss_cart *c, int field, double fact) { return bc_call(c, field, BC_X_LIN, fact, "ss_cart_bc_x_lin"); }
int ss_cart_bc_x_0(ss_cart *c, int field) { return bc_call(c, field, BC_X_0, 0.0, "ss_cart_bc_x_0"); }
int ss_cart_copy_bc_x(ss_cart *c, int field) { return bc_call(c, field, COPY_BC_X, 0.0, "ss_cart_copy_bc_x"); }
int ss_cart_bc_y_fact(ss_cart *c, int field, double fact) { return bc_call(c, field, BC_Y_FACT, fact, "ss_cart_bc_y_fact"); }
int ss_cart_bc_y_0(ss_cart *c, int field) { return bc_call(c, field, BC_Y_0, 0.0, "ss_cart_bc_y_0"); }
int ss_cart_bc_z_0(ss_cart *c, int field) { return bc_call(c, field, BC_Z_0, 0.0, "ss_cart_bc_z_0"); }
int ss_cart_copy_bc_z(ss_cart *c, int field) { return bc_call(c, field, COPY_BC_Z, 0.0, "ss_cart_copy_bc_z"); }

int ss_cart_check_err(ss_cart *c, double *err, double comp[3])
{
    CTX_OK(c);
    if (!err) return cfail(c, SS_ERR_ARG, "ss_cart_check_err: null");
    launch_error(c);
    return read_err(c, err, comp);
}

int ss_cart_pt_solve(ss_cart *c, const ss_cart_params *p, int maxiter, int nout, double eps, ss_cart_result *out)
{
    CTX_OK(c);
    if (!p || !out || maxiter < 1 || nout < 1) return cfail(c, SS_ERR_ARG, "ss_cart_pt_solve: bad argument");
    if (!c->y_set) return cfail(c, SS_ERR_STATE, "ss_cart_set_y was not called");
    memset(out, 0, sizeof *out);
    out->err = NAN;
    const cclk::time_point t0 = cclk::now();
    int iter = 1;
    for (; iter <= maxiter; ++iter) {
        const bool check = iter % nout == 0;
        launch_iteration(c, *p, check);
        if (check) {
            launch_error(c);
            double err, comp[3];
            int rc = read_err(c, &err, comp);
            if (rc != SS_OK) return rc;
            out->err = err;
            if (std::isnan(err)) {
                out->iters = iter;
                out->seconds = std::chrono::duration<double>(cclk::now() - t0).count();
                return cfail(c, SS_ERR_NONFINITE, "NaN in the exit test");
            }
            if (err < eps && iter > 20) {   // :389-391 (iters = iter)
                out->iters = iter;
                out->converged = 1;
                break;
            }
        }
    }
    if (!out->converged) out->iters = maxiter;
    int rc = csync(c, "ss_cart_pt_solve");
    out->seconds = std::chrono::duration<double>(cclk::now() - t0).count();
    if (rc != SS_OK) return rc;
    return out->converged ? SS_OK : SS_NOT_CONVERGED;
}

int ss_cart_pt_iterate_timed(ss_cart *c, const ss_cart_params *p, int niter, double *iter_ms, double *kernel_ms)
{
    CTX_OK(c);
    if (!p || niter < 1) return cfail(c, SS_ERR_ARG, "ss_cart_pt_iterate_timed: bad argument");
    if (!c->y_set) return cfail(c, SS_ERR_STATE, "ss_cart_set_y was not called");
    int rc = csync(c, "ss_cart_pt_iterate_timed");
    if (rc != SS_OK) return rc;
    // one event pair around compute_P! per iteration, read after the run
    std::vector<hipEvent_t> ev(2 * (size_t)niter, nullptr);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            for (auto x : ev)
                if (x) (void)hipEventDestroy(x);
            return cfail(c, SS_ERR_NOMEM, "ss_cart_pt_iterate_timed: events");
        }
    (void)hipEventRecord(c->ev[2], c->stream);
    for (int it = 0; it < niter; ++it) launch_iteration(c, *p, false, ev[2 * it], ev[2 * it + 1]);
    (void)hipEventRecord(c->ev[3], c->stream);
    rc = csync(c, "ss_cart_pt_iterate_timed");
    double kms = 0.0;
    float total = 0.0f;
    if (rc == SS_OK) {
        for (int it = 0; it < niter; ++it) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[2 * it], ev[2 * it + 1]);
            kms += ms;
        }
        (void)hipEventElapsedTime(&total, c->ev[2], c->ev[3]);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    if (rc != SS_OK) return rc;
    if (iter_ms) *iter_ms = total / niter;
    if (kernel_ms) *kernel_ms = kms / niter;
    return csync(c, "ss_cart_pt_iterate_timed");
}

} // extern "C"
