// march_pattern.hip -- the memory-access pattern of the fused PT kernel without
// its arithmetic (a ceiling for that kernel at its tiling; measurement only,
// not part of the library).
// Box layout as the library: (nr+1) x (nt+1) x pitch doubles, phi fastest.
// Each workgroup (TKW x TJ threads, TKW a multiple of 64: TKW/64 wavefronts
// per row) marches a chunk of r-planes; per step every
// thread loads NIN values of plane s+1 at its (j,k) from NIN input fields and
// stores NOUT values of plane s into NOUT output fields (values kept one step
// in registers), with NB block barriers per step.  All indices are clamped
// into the box, so every access is in bounds by construction.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>

#ifndef TJ
#define TJ 12
#endif
#ifndef TKW
#define TKW 64
#endif
#ifndef NIN
#define NIN 13
#endif
#ifndef NOUT
#define NOUT 11
#endif
#ifndef NB
#define NB 3
#endif

struct Args {
    const double *in[NIN];
    double *out[NOUT];
    int nr, nt, np, pitch, rchunk, nx, ny, nz;
};

__global__ void __launch_bounds__(TKW * TJ) k_pattern(Args a)
{
    __shared__ double L[TJ][TKW];
    const int T = a.nx * a.ny * a.nz;
    const int per = (T + 7) / 8;
    const int b = blockIdx.x;
    const int t = (b % 8) * per + b / 8;
    if (t >= T) return;
    const int bx = t % a.nx, by = (t / a.nx) % a.ny, bz = t / (a.nx * a.ny);
    const int k = min(bx * (TKW - 2) - 1 + (int)threadIdx.x, a.np);
    const int j = min(max(by * (TJ - 2) - 1 + (int)threadIdx.y, 0), a.nt);
    const int kk = max(k, 0);
    const int s0 = bz * a.rchunk, s1 = min(s0 + a.rchunk, a.nr);
    const int start = max(s0 - 2, 0);
    const long long sj = a.pitch, si = (long long)(a.nt + 1) * a.pitch;
    double cur[NIN], nxt[NIN];
    for (int q = 0; q < NIN; ++q) cur[q] = a.in[q][start * si + j * sj + kk];
    const bool own = threadIdx.x >= 1 && threadIdx.x <= TKW - 2 && threadIdx.y >= 1 && threadIdx.y <= TJ - 2;
    for (int s = start; s < s1; ++s) {
        const int sn = min(s + 1, a.nr - 1);
        for (int q = 0; q < NIN; ++q) nxt[q] = a.in[q][sn * si + j * sj + kk];
        double acc = 0.0;
        for (int q = 0; q < NIN; ++q) acc += cur[q];
        L[threadIdx.y][threadIdx.x] = acc;
        for (int bb = 0; bb < NB; ++bb) __syncthreads();
        const double nb = L[(threadIdx.y + 1) % TJ][threadIdx.x];
        if (own && s >= s0)
            for (int q = 0; q < NOUT; ++q) a.out[q][s * si + j * sj + kk] = acc + nb * q;
        for (int q = 0; q < NIN; ++q) cur[q] = nxt[q];
        __syncthreads();
    }
}

int main(int argc, char **argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 511;
    const int rchunk = argc > 2 ? atoi(argv[2]) : 73;
    const int nr = n, nt = n, np = n, pitch = ((np + 1) + 63) / 64 * 64;
    const size_t box = (size_t)(nr + 1) * (nt + 1) * pitch;
    Args a;
    a.nr = nr; a.nt = nt; a.np = np; a.pitch = pitch; a.rchunk = rchunk;
    a.nx = (np + TKW - 3) / (TKW - 2); a.ny = (nt + TJ - 3) / (TJ - 2); a.nz = (nr + rchunk - 1) / rchunk;
    for (int q = 0; q < NIN; ++q) {
        double *p; if (hipMalloc(&p, box * 8) != hipSuccess) { printf("alloc fail\n"); return 1; }
        hipMemset(p, 0, box * 8); a.in[q] = p;
    }
    for (int q = 0; q < NOUT; ++q) {
        double *p; if (hipMalloc(&p, box * 8) != hipSuccess) { printf("alloc fail\n"); return 1; }
        hipMemset(p, 0, box * 8); a.out[q] = p;
    }
    const int T = a.nx * a.ny * a.nz, grid = (T + 7) / 8 * 8;
    hipEvent_t e0, e1; hipEventCreate(&e0); hipEventCreate(&e1);
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_pattern, dim3(grid), dim3(TKW, TJ), 0, 0, a);
    hipDeviceSynchronize();
    const int reps = 20;
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_pattern, dim3(grid), dim3(TKW, TJ), 0, 0, a);
    hipEventRecord(e1, 0); hipEventSynchronize(e1);
    float ms; hipEventElapsedTime(&ms, e0, e1); ms /= reps;
    const double cells = (double)nr * nt * np;
    const double bytes = cells * (NIN + NOUT) * 8;
    printf("n=%d TKW=%d TJ=%d NIN=%d NOUT=%d NB=%d rchunk=%d: %.3f ms  %.1f GB/s (algorithmic %d arrays)  A_eff-equiv %.1f GB/s\n",
           n, TKW, TJ, NIN, NOUT, NB, rchunk, ms, bytes / (ms * 1e-3) / 1e9, NIN + NOUT, cells * 26 * 8 / (ms * 1e-3) / 1e9);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
