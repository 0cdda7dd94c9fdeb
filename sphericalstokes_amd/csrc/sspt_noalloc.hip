// sspt_noalloc.hip -- ss_check_err_noalloc: the convergence check of
// ss_check_err (check_err! + the host max-reductions) on caller-owned device
// arrays, with its reduction scratch in a static device array instead of
// context memory.  It makes no allocation of any kind, enqueues on the
// caller's stream and does not synchronise, so it can be called inside HIP
// graph capture or a tight loop that must not allocate.
//
// Two passes, no atomics: pass 1 has block b reduce the fixed segment
// [b*SEG, (b+1)*SEG) of each component (V_err <- V - V_err, max|V_err|,
// max|V| as FP64 bit patterns, which order like the values for non-negative
// doubles; a NaN sorts above +inf and propagates like Julia's maximum) into
// scratch[b][0..5]; pass 2 (one block) reduces the partials and writes
// err_comp[3] and err to the caller's device buffer.  Every partial that
// pass 2 reads is written by pass 1 of the same call, so nothing has to be
// cleared between calls; two calls that overlap on one device race on the
// scratch and must be ordered by the caller (one stream, or events).
#include "sspt_internal.h"
#include <cmath>
#include <mutex>

namespace sspt {
namespace noalloc {

constexpr int BLOCK = 256;
constexpr int PER_THREAD = 16;
constexpr long long SEG = (long long)BLOCK * PER_THREAD; // elements per block and component
constexpr int MAX_BLOCKS = 65536;                        // scratch holds this many partial sets
constexpr long long MAX_ELEMS = SEG * MAX_BLOCKS;        // 268 435 456 elements per component
constexpr int MAX_DEVICES = 64;

// per-block partial maxima [block][6]: 3 MiB, held by every device the
// library runs on until the process ends
static __device__ __align__(256) unsigned long long scratch[(size_t)MAX_BLOCKS * 6];

static std::mutex g_mu;
static unsigned long long *g_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.
static hipError_t scratch_ptr(unsigned long long **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

__device__ __forceinline__ unsigned long long umax(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

__device__ __forceinline__ unsigned long long wave_umax(unsigned long long v)
{
    for (int s = 32; s > 0; s >>= 1) v = umax(v, (unsigned long long)__shfl_xor(v, s, 64));
    return v;
}

struct Comp {
    const double *V[3];
    double *E[3];
    long long n[3];
};

__global__ void __launch_bounds__(BLOCK) k_partials(Comp c, unsigned long long *part)
{
    __shared__ unsigned long long red[BLOCK / 64][6];
    const long long base = (long long)blockIdx.x * SEG;
    unsigned long long me[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const long long end = min(base + SEG, c.n[q]);
        for (long long o = base + threadIdx.x; o < end; o += BLOCK) {
            const double v = c.V[q][o];
            const double d = v - c.E[q][o];
            c.E[q][o] = d;
            me[q] = umax(me[q], (unsigned long long)__double_as_longlong(fabs(d)));
            mv[q] = umax(mv[q], (unsigned long long)__double_as_longlong(fabs(v)));
        }
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        me[q] = wave_umax(me[q]);
        mv[q] = wave_umax(mv[q]);
        if (lane == 0) {
            red[w][q] = me[q];
            red[w][3 + q] = mv[q];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        unsigned long long m = 0;
        for (int k = 0; k < BLOCK / 64; ++k) m = umax(m, red[k][threadIdx.x]);
        part[(size_t)blockIdx.x * 6 + threadIdx.x] = m;
    }
}

// Julia's max for Float64 (NaN-propagating, +0 > -0)
__device__ __forceinline__ double jl_max_d(double x, double y)
{
    const bool pick_y = (y > x) || (signbit(y) < signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

__global__ void __launch_bounds__(BLOCK) k_final(const unsigned long long *part, int nblocks, int drop_nan_vph,
                                                 double *out)
{
    __shared__ unsigned long long red[BLOCK / 64][6];
    unsigned long long m[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += BLOCK)
#pragma unroll
        for (int q = 0; q < 6; ++q) m[q] = umax(m[q], part[(size_t)b * 6 + q]);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        m[q] = wave_umax(m[q]);
        if (lane == 0) red[w][q] = m[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double comp[3];
        for (int q = 0; q < 3; ++q) {
            unsigned long long a = 0, b = 0;
            for (int k = 0; k < BLOCK / 64; ++k) {
                a = umax(a, red[k][q]);
                b = umax(b, red[k][3 + q]);
            }
            comp[q] = __longlong_as_double((long long)a) / __longlong_as_double((long long)b);
            out[q] = comp[q];
        }
        out[3] = (drop_nan_vph && isnan(comp[2])) ? jl_max_d(comp[0], comp[1])
                                                   : jl_max_d(jl_max_d(comp[0], comp[1]), comp[2]);
    }
}

} // namespace noalloc
} // namespace sspt

using namespace sspt::noalloc;

extern "C" int ss_check_err_noalloc(const double *VR, const double *Vth, const double *Vph, double *VR_err,
                                    double *Vth_err, double *Vph_err, int nr, int nth, int nph, int drop_nan_vph,
                                    double *out, void *stream)
{
    if (!VR || !Vth || !Vph || !VR_err || !Vth_err || !Vph_err || !out) return SS_ERR_ARG;
    if (nr < 1 || nth < 1 || nph < 1) return SS_ERR_ARG;
    Comp c{{VR, Vth, Vph}, {VR_err, Vth_err, Vph_err},
           {(long long)(nr + 1) * nth * nph, (long long)nr * (nth + 1) * nph, (long long)nr * nth * (nph + 1)}};
    long long nmax = 0;
    for (int q = 0; q < 3; ++q) nmax = c.n[q] > nmax ? c.n[q] : nmax;
    if (nmax > MAX_ELEMS) return SS_ERR_ARG; // beyond the scratch's 65 536 partial sets of 4 096 elements
    unsigned long long *part = nullptr;
    hipError_t e = scratch_ptr(&part);
    if (e != hipSuccess) return (int)e;
    const int nblocks = (int)((nmax + SEG - 1) / SEG);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_partials, dim3(nblocks), dim3(BLOCK), 0, s, c, part);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(BLOCK), 0, s, part, nblocks, drop_nan_vph, out);
    e = hipGetLastError();
    return e == hipSuccess ? SS_OK : (int)e;
}
