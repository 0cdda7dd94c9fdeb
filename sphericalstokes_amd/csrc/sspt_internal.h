// sspt_internal.h -- device layout, metric tables and context of libsspt.so.
//
// Device layout ("box"): every field lives in a box of (nr+1) x (nth+1) x pitch
// doubles, phi fastest, so all fields share the same strides:
//     off(i,j,k) = i*si + j*sj + k,   sj = pitch, si = (nth+1)*pitch.
// Index conventions (0-based):
//   centre fields, VR, Vth, Vph: box (i,j,k) = reference (ix-1, iy-1, iz-1);
//   shear stresses tau_Rth, tau_Rph, tau_thph: box (i,j,k) = reference (ix,iy,iz),
//   i.e. an edge is indexed by the face index in its staggered directions and
//   by the centre index otherwise (tau_Rth(i,j,k) sits on r-face i, th-face j,
//   phi-centre k).  With this shift every stencil below is written in the same
//   natural (i,j,k) frame.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/sspt.h"

namespace sspt {

struct Box {
    int nr, nt, np;   // cells in r, th, phi
    int pitch;        // phi row length (>= np+1)
    long long sj, si; // strides of j and i
    __host__ __device__ inline long long off(int i, int j, int k) const
    {
        return (long long)i * si + (long long)j * sj + k;
    }
};

// 1D/2D FP64 metric tables, built on the host with the reference's exact
// expression order (build_tables in sspt_tables.h), stored in one device
// allocation.  r-tables have nr+1 entries, th-tables nth+1, 2D tables are
// (nr+1) x (nth+1) with th fastest.
struct Tables {
    // r direction
    const double *Rc;     // R_i                                        (centre)
    const double *invRc;  // 1.0/R_i
    const double *Rf;     // (R[i-1]+R[i])*0.5                          (@avxi_R, r face)
    const double *invRf;
    const double *Re;     // (((R[i-1]+R[i-1])+R[i])+R[i])*0.25         (@avxyi_R/@avxzi_R, r-edge)
    const double *invRe;
    const double *Rc4;    // (((R[i]+R[i])+R[i])+R[i])*0.25             (@avyzi_R)
    const double *invRc4;
    const double *Rc2;    // (R[i]+R[i])*0.5                            (@avyi_R/@avzi_R)
    const double *invRc2;
    // th direction
    const double *cotc;   // cot(th_j) = 1/tan(th_j)
    const double *cotf4;  // cot((((th[j-1]+th[j-1])+th[j])+th[j])*0.25)  (@avyzi_th)
    const double *cotc2;  // cot((th[j]+th[j])*0.5)                        (@avxi_th/@avzi_th)
    const double *cotf2;  // cot((th[j-1]+th[j])*0.5)                      (@avyi_th)
    // 2D (1/R)/sin(th) combinations, indexed [i*(nth+1)+j]
    const double *rs_c;   // (1/R_i)/sin(th_j)                compute_P centre
    const double *rs_e;   // (1/Re_i)/sin(th4c_j)             tau_Rph edge
    const double *rs_c4;  // (1/Rc4_i)/sin(th4f_j)            tau_thph edge
    const double *rs_f;   // (1/Rf_i)/sin(th2c_j)             VR face
    const double *rs_y;   // (1/Rc2_i)/sin(th2f_j)            Vth face
    const double *rs_z;   // (1/Rc2_i)/sin(th2c_j)            Vph face
    // coordinates for BCs / setup
    const double *th;     // th_j
    const double *ph;     // ph_k (nph+1 entries)
};

struct FieldSet {
    double *f[SS_NFIELDS];
};

// FP32 boxes of the defect-correction solve: two correction sets
// (VR, Vth, Vph, P, tau x6 at 0..9 and 10..19), the frozen viscosity and the
// forcing (divV and the momentum residual of the FP64 iterate).
// (SM, TMP: the radius-2 etaSM of the frozen viscosity and its temporary)
enum { SS_F32_ETA = 20, SS_F32_GP = 21, SS_F32_FVR = 22, SS_F32_FVT = 23, SS_F32_FVP = 24, SS_F32_SM = 25, SS_F32_TMP = 26,
       SS_F32_NBOX = 27 };

// FP64 boxes of the adjoint solve (sspt_adjoint.hip lists them)
enum { SS_ADJ_NBOX = 50, SS_ADJ32_NBOX = 49 };

// Length of the r-chunks of a marching kernel (grid z).  A workgroup marches
// its chunk plus two warm-up planes, and one workgroup fits a CU at a time,
// so a launch over `tiles` column tiles costs about
//     ceil(tiles * ceil(nplanes / ch) / slots) * (ch + 2)
// plane steps of one CU.  The chunk minimising that count is taken (ties to
// the shorter chunk), between 3 and 80 planes: longer chunks measured slower
// at 383^3-767^3 whatever the count says (767^3 FP64: 59 planes 21.6 ms,
// 257 planes 24.0 ms, 767 planes 24.9 ms), and the count picks the measured
// best below (FP64 63^3: 4, 127^3: 23, 191^3: 65, 255^3: 17, 383^3: 35,
// 511^3: 73; FP32 127^3: 16, 767^3: 77).  odd_long keeps chunks above 8
// planes odd (FP64 at 511^3: chunks of 32 and 64 planes, concurrent
// workgroups a power-of-two number of planes apart, measured 4-6 % slower).
inline int march_chunk(int nplanes, int tiles, int slots, bool odd_long)
{
    int best = nplanes < 3 ? nplanes : 3;
    long long best_cost = -1, best_wg = 0;
    const int top = nplanes < 80 ? nplanes : 80;
    for (int ch = 3; ch <= top; ++ch) {
        if (odd_long && ch > 8 && ch % 2 == 0) continue;
        const long long wg = (long long)tiles * ((nplanes + ch - 1) / ch);
        const long long cost = ((wg + slots - 1) / slots) * (ch + 2);
        if (best_cost < 0 || cost < best_cost || (cost == best_cost && wg > best_wg)) {
            best = ch;
            best_cost = cost;
            best_wg = wg;
        }
    }
    return best;
}

} // namespace sspt

struct ss_ctx {
    int device = 0;
    int n_cu = 256;                 // compute units of the device (workgroup slots of the marching kernels)
    sspt::Box box{};
    size_t box_elems = 0;
    hipStream_t stream = nullptr;
    sspt::FieldSet cur{};   // current state
    sspt::FieldSet alt{};   // ping-pong partners used by the fused loop
    bool alt_ready = false;
    bool pair_state = true;         // allocate in/out halves of state fields together (fused kernel F_PAIRED)
    double *pair_base[SS_NFIELDS] = {}; // lower halves of the paired FP64 state fields
    float *pair_base32[10] = {};        // FP32 correction pairs (set A | set B)
    double *tab_mem = nullptr;
    sspt::Tables tab{};
    double *ftab_mem = nullptr;     // interleaved tables of the fused kernel
    const double *ftr = nullptr, *ftt = nullptr, *ft2 = nullptr;
    float *ftab32_mem = nullptr;    // the same tables (and th, ph) rounded to FP32
    const float *ftr32 = nullptr, *ftt32 = nullptr, *ft232 = nullptr, *th32 = nullptr, *ph32 = nullptr;
    double *tau_scr[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}; // residual scratch
    float *f32[sspt::SS_F32_NBOX] = {};
    bool f32_ready = false;
    double *adj[sspt::SS_ADJ_NBOX] = {}; // adjoint: two lam sets, check copies, temporaries, fixed coefficients
    double *adj_part = nullptr;     // partial sums of the objective
    bool adj_ready = false;
    int adj_cur = 0;                // lam set holding the current iterate
    double adj_obj_r = 0.0, adj_obj_rad = 0.0; // inclusion of the objective (ss_adjoint_objective)
    float *adj32[sspt::SS_ADJ32_NBOX] = {}; // FP32 adjoint correction: two sets, right-hand side, state, LB
    bool adj_fused = true;          // r-marching adjoint kernel, else the per-family kernels (set by ss_adjoint_prepare)
    bool grid_set = false;
    double th_lo = 0.0, th_hi = 0.0; // th[0], th[nth-1] (host copies for the pure-shear BC)
    double *staging = nullptr;      // upload/download staging (largest field)
    size_t staging_elems = 0;
    unsigned long long *red = nullptr; // reduction slots (device)
    double *tbp = nullptr; // tII next to the phi faces, [2][nr][nth] (fused iteration)
    unsigned long long *red_host = nullptr; // pinned mirror
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> ev_pool;       // per-launch timing events
    double last_kernel_ms = 0.0;
    std::string err;
};
