// sspt_fused_body.h -- the fused PT iteration as a sequence of block phases.
//
// One launch of the main kernel performs a whole PT iteration
// (SphericalStokes_diapir.jl:413-425: maxloc! + copy_BC!, compute_P!,
// compute_TII! + copy_BC!, power_law!, compute_V!) for every interior point,
// reading the old state and writing the new state into separate (ping-pong)
// buffers.  A second, small kernel (boundary pass) then writes the ghost
// faces of V (the BC_R!/BC_th!/BC_ph! sequence) and the boundary-cell
// viscosities (power law of the copied tII).
//
// Block = TKX lanes along phi x TJB rows along th (64 x 12 in FP64, 64 x 16 in
// FP32; a row is one wavefront at TKX = 64, several at a wider build option).
// Lane tx covers k = kb-1+tx, row ty covers j = jb-1+ty; the (TKX-2) x (TJB-2)
// inner points are owned (written), lane 0 / row 0 are a low halo and lane
// TKX-1 / row TJB-1 a high halo that are recomputed from the old state.  The block marches along r: step s handles
// centre plane s, r-face s (VR) and r-face s+1 (tau_Rth, tau_Rph).  Values of
// the current plane that neighbours need go through LDS; values the same
// column needs at the next plane stay in registers (the Carry struct).
//
// The phases are plain functions of (thread state, LDS, args) separated by
// block barriers, so the same code also runs in a host emulation that loops
// over the threads of a block between barriers.
#pragma once
#include <cmath>
#include <cstdint>

#ifndef SSPT_DH
#define SSPT_DH __host__ __device__
#endif
// A value that is the same in every lane of a wavefront (blockDim.x == 64, so
// threadIdx.y and everything derived from it is wave-uniform).
#ifndef SSPT_UNIFORM
#define SSPT_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#endif

namespace sspt {
namespace fused {

constexpr int TK = 64;        // lanes along phi of the adjoint's blocks (one wavefront)
// lanes along phi per row of the PT-iteration blocks, a multiple of 64 (a row
// is one or more whole wavefronts; every exchange between lanes goes through
// LDS across a barrier, so a row may span several)
#ifndef SSPT_TK64
#define SSPT_TK64 64
#endif
#ifndef SSPT_TK32
#define SSPT_TK32 64
#endif
template <typename R> constexpr int TKX = sizeof(R) == 8 ? SSPT_TK64 : SSPT_TK32;
// rows along th: 12 for FP64 (3 wavefronts per SIMD at <= 168 VGPRs), 16 for
// FP32 (4 per SIMD)
#ifndef SSPT_TJB64
#define SSPT_TJB64 12
#endif
#ifndef SSPT_TJB32
#define SSPT_TJB32 16
#endif
template <typename R> constexpr int TJB = sizeof(R) == 8 ? SSPT_TJB64 : SSPT_TJB32;
template <typename R> constexpr int OKX = TKX<R> - 2;   // owned points per row
template <typename R> constexpr int OJ = TJB<R> - 2; // owned rows
constexpr double PI_ = 3.14159265358979323846;

// F_CORR: correction sweep of the FP32/FP64 defect-correction solve (forcing
// arrays instead of rho*g and the continuity right-hand side, frozen eta, no
// power law; homogeneous BCs come from sh = 0).
// F_PAIRED: the in and out buffers of each state field are the two halves of
// one allocation (one field base serves both, see fat_pp / wout).
// F_SMBUF: etaSM is read from the SM field (filled before the launch, with
// the copy_BC'd values at every cell) instead of being computed from the
// old eta plane by plane; the wide (radius-2) etaSM option uses it.
enum Flags { F_LAST = 1, F_CHECK = 2, F_RELAX = 4, F_CLAMP = 8, F_CORR = 16, F_PAIRED = 32, F_SMBUF = 64 };

struct BoxG {
    int nr, nt, np;
    long long sj, si;
    SSPT_DH long long off(int i, int j, int k) const { return (long long)i * si + (long long)j * sj + k; }
    // the same offset in 32 bits; a field box is below 4 GiB (checked when
    // the context is created), so byte offsets fit too
    SSPT_DH uint32_t o32(int i, int j, int k) const
    {
        return (uint32_t)i * (uint32_t)si + (uint32_t)j * (uint32_t)sj + (uint32_t)k;
    }
};

// Element index in the phi-face tII buffer ([2][nr][nt], side 0 at k = 1,
// side 1 at k = np-2).
SSPT_DH inline uint32_t tbp_off(const BoxG &b, int side, int i, int j)
{
    return ((uint32_t)side * (uint32_t)b.nr + (uint32_t)i) * (uint32_t)b.nt + (uint32_t)j;
}

// SSPT_ACCESS_CHECK(base, byte_offset, bytes): a hook on every field access
// through the 32-bit offsets below; empty in the library, a bounds check in
// the host emulation's sanitizer build.
#ifndef SSPT_ACCESS_CHECK
#define SSPT_ACCESS_CHECK(p, off, n) ((void)0)
#endif

// Element o of a field through a 32-bit byte offset from the field's base:
// with a wave-uniform base this is the SGPR-base + 32-bit VGPR-offset form of
// global memory instructions, with no 64-bit address arithmetic per access.
template <typename T> SSPT_DH inline T &fat(T *p, uint32_t o)
{
    SSPT_ACCESS_CHECK(p, (uint32_t)(o * (uint32_t)sizeof(T)), sizeof(T));
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + (uint32_t)(o * (uint32_t)sizeof(T)));
}
// Element o of a field at byte offset `off` from base p (the in or out half
// of a paired allocation; off + 8 o stays below 4 GiB).
template <typename T> SSPT_DH inline const T &fat_pp(const T *p, uint32_t o, uint32_t off)
{
    SSPT_ACCESS_CHECK(p, (uint32_t)(o * (uint32_t)sizeof(T) + off), sizeof(T));
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (uint32_t)(o * (uint32_t)sizeof(T) + off));
}
// the in half of a state field (offset first, for reading at call sites)
template <typename T> SSPT_DH inline const T &fat_in(const T *p, uint32_t off, uint32_t o) { return fat_pp(p, o, off); }
template <typename T> SSPT_DH inline const T &fat(const T *p, uint32_t o)
{
    SSPT_ACCESS_CHECK(p, (uint32_t)(o * (uint32_t)sizeof(T)), sizeof(T));
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (uint32_t)(o * (uint32_t)sizeof(T)));
}

// Interleaved FP64 metric tables (one record per index, see
// build_fused_tables in sspt_tables.h): a record is read with a few scalar
// loads because s and the wavefront's j are uniform.
enum { TR_Rc, TR_invRc, TR_Rf, TR_invRf, TR_Re, TR_invRe, TR_Rc4, TR_invRc4, TR_Rc2, TR_invRc2, NTR = 16 };
enum { TT_cotc, TT_cotf4, TT_cotc2, TT_cotf2, NTT = 4 };
enum { T2_c, T2_e, T2_c4, T2_f, T2_y, T2_z, NT2 = 8 };

// Table reads go through the constant address space: the compiler may then
// assume that no store of the kernel changes them and, the index being
// wave-uniform, read them with scalar loads.  Those are counted apart from
// the vector memory operations, so a table read never waits for the
// prefetches and stores of the field data in flight.
#ifndef SSPT_CONSTANT_PTR
#define SSPT_CONSTANT_PTR(T, p) ((const __attribute__((address_space(4))) T *)(p))
#endif
template <typename R> SSPT_DH inline R tload(const R *p, int i) { return SSPT_CONSTANT_PTR(R, p)[i]; }

template <typename R> struct TabsG {
    const R *tr; // [i*NTR + TR_*], i in [0, nr]
    const R *tt; // [j*NTT + TT_*], j in [0, nt]
    const R *t2; // [(i*(nt+1)+j)*NT2 + T2_*]
    const R *th, *ph;
};

template <typename R> struct Args {
    BoxG b;
    TabsG<R> t;
    // old state
    const R *VR, *VT, *VP, *eta, *P, *tRR, *tTT, *tPP, *tRT, *tRP, *tTP, *rho, *ini, *ePL;
    // new state
    R *oVR, *oVT, *oVP, *oEta, *oP, *oRR, *oTT, *oPP, *oRT, *oRP, *oTP, *oPL;
    // auxiliary outputs (last iteration / convergence check)
    R *SM, *DT, *TII, *VRe, *VTe, *VPe;
    // tII of the interior cells next to the phi faces, [side][i][j] with
    // side 0 at k = 1 and side 1 at k = np-2: written by the main kernel,
    // read by the boundary pass for the viscosity of the phi-face cells
    R *TBP;
    // forcing of a correction sweep (F_CORR): divV of the FP64 iterate at
    // centres and its momentum residual at faces
    const R *gP, *fVR, *fVT, *fVP;
    unsigned long long *red; // [0..2] max|dV|, [3..5] max|V| as bit patterns
    R _dr, _dth, _dph, kb, gs, vpdt2, g, tC, ex, relax, emin, emax;
    int flags;
    int bc_mode;
    R r, sh;
    // pure shear: prescribed Vth at the th = th[0] and th = th[nt-1] boundaries
    // (pure_shear_vth, evaluated on the host)
    R vth0, vth1;
    int spread; // etaSM spread (F_SMBUF; filled before the launch)
    int rchunk; // r-steps per block along grid z
    // byte offsets of the in and out halves of the state fields (viscosity
    // apart: it is frozen in a correction sweep); 0 unless F_PAIRED, where
    // the state pointers above are the lower halves
    uint32_t pp_in, pp_out, ppe_in, ppe_out;
};

// Arguments of the main kernel for paired state buffers: when the in and
// out buffers of every state field are the two halves of one allocation of
// 2 x box bytes (below 4 GiB), the lower half becomes the field's only base
// and the halves are addressed by byte offsets (F_PAIRED); this frees the
// scalar registers of eleven out pointers.  FS = bytes of one box.  Returns
// false (and changes nothing) when the buffers are not paired so.
template <typename R> SSPT_DH inline bool pair_args(Args<R> &a, long long FS)
{
    if (2 * FS > 0xFFFFFFFFll) return false;
    const R **in[10] = {&a.P, &a.tRR, &a.tTT, &a.tPP, &a.tRT, &a.tRP, &a.tTP, &a.VR, &a.VT, &a.VP};
    R **out[10] = {&a.oP, &a.oRR, &a.oTT, &a.oPP, &a.oRT, &a.oRP, &a.oTP, &a.oVR, &a.oVT, &a.oVP};
    const long long d0 = (const char *)*out[0] - (const char *)*in[0];
    if (d0 != FS && d0 != -FS) return false;
    for (int q = 1; q < 10; ++q)
        if ((const char *)*out[q] - (const char *)*in[q] != d0) return false;
    const bool corr = (a.flags & F_CORR) != 0;
    if (!corr && (const char *)a.oEta - (const char *)a.eta != d0) return false;
    const uint32_t pin = d0 > 0 ? 0u : (uint32_t)FS, pout = d0 > 0 ? (uint32_t)FS : 0u;
    for (int q = 0; q < 10; ++q) {
        const R *lo = d0 > 0 ? *in[q] : *out[q];
        *in[q] = lo;
        *out[q] = const_cast<R *>(lo);
    }
    a.pp_in = pin;
    a.pp_out = pout;
    if (!corr) {
        const R *lo = d0 > 0 ? a.eta : a.oEta;
        a.eta = lo;
        a.oEta = const_cast<R *>(lo);
        a.ppe_in = pin;
        a.ppe_out = pout;
    }
    a.flags |= F_PAIRED;
    return true;
}

// Destination of a state write: with F_PAIRED the out half of the pair
// based at `in`, otherwise the separate out buffer.
template <typename R, int FL> SSPT_DH inline R &wout(R *out, const R *in, uint32_t o, uint32_t pp)
{
    if constexpr ((FL & F_PAIRED) != 0)
        return const_cast<R &>(fat_pp(in, o, pp));
    else
        return fat(out, o);
}
// Store of an output that only the next launch reads.  SSPT_NT_STORE adds the
// non-temporal hint (global_store ... nt), a build option for measurements:
// at 511^3 it measured no gain on the FP64 iteration and made the FP32 sweep
// 10 % and the adjoint step 1.6 % slower (both builds in one process,
// alternating repetitions).
template <typename T> SSPT_DH inline void st_stream(T &dst, T v)
{
#if defined(__HIP_DEVICE_COMPILE__) && defined(SSPT_NT_STORE)
    __builtin_nontemporal_store(v, &dst);
#else
    dst = v;
#endif
}
// Stores a new state value.
template <typename R, int FL> SSPT_DH inline void st_out(R *out, const R *in, uint32_t o, uint32_t pp, R v)
{
    st_stream(wout<R, FL>(out, in, o, pp), v);
}

template <typename R> struct Lds {
    R VR[TJB<R>][TKX<R>], VT[TJB<R>][TKX<R>], VP[TJB<R>][TKX<R>];
    R ETA[TJB<R> + 2][TKX<R> + 2];
    R SM[TJB<R>][TKX<R>];
    R STT[TJB<R>][TKX<R>], SPP[TJB<R>][TKX<R>], DT[TJB<R>][TKX<R>];
    R RT[TJB<R>][TKX<R>], RP[TJB<R>][TKX<R>], TP[TJB<R>][TKX<R>];
};

// Normal velocity that the pure-shear BC prescribes on a th boundary at th_b:
// the th component of V = (0, -sh*r*(th-pi/2)/sin(th), sh*r*ph), a
// divergence-free pure shear of rate sh in the th-phi plane whose phi
// component is the phi faces' prescribed Vph_b = (ph_b*r)*sh.  Host code
// only: sin comes from the host math library, as in the CPU restatement.
inline double pure_shear_vth(double th_b, double r, double sh)
{
    return -((((th_b - PI_ / 2.0) * r) * sh) / std::sin(th_b));
}

// Julia Base.max/min for Float64
template <typename R> SSPT_DH inline R jmax(R x, R y)
{
    bool pick_y = (y > x) || (std::signbit(y) < std::signbit(x));
    return pick_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}
template <typename R> SSPT_DH inline R jmin(R x, R y)
{
    bool pick_y = (y < x) || (std::signbit(y) > std::signbit(x));
    return pick_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}
SSPT_DH inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
SSPT_DH inline unsigned long long absbits(double x)
{
    union { double d; unsigned long long u; } c;
    c.d = std::fabs(x);
    return c.u;
}
// the max-reductions work on FP64 bit patterns in both precisions
SSPT_DH inline unsigned long long absbits(float x) { return absbits((double)x); }

// x / d for d with a correctly rounded reciprocal inv_d = RN(1/d): one FMA
// correction of x*inv_d gives the correctly rounded quotient (Markstein's
// theorem), i.e. exactly what IEEE division returns, for a product and two FMAs
// instead of a full division sequence.
template <typename R> SSPT_DH inline R div_rcp(R x, R d, R inv_d)
{
    const R q = x * inv_d;
    const R r = std::fma(-q, d, x);
    return std::fma(r, inv_d, q);
}

// x^ex of the power law.  ex = 0 (n_exp = 1) is 1 exactly, as pow returns.
// Small integer exponents (n_exp = 2..9 -> ex = -1..-8): x^m is formed as
// an unevaluated sum hi + lo (each product's rounding error recovered with
// an FMA), then 1/(hi + lo) as q = RN(1/hi) plus one Newton correction
// q + q*e, e = 1 - q*(hi + lo), whose relative error before the last
// rounding is about 2^-100: the result is the correctly rounded x^-m except
// within ~2^-47 of a rounding tie.  A zero, infinite or NaN q (x = 0,
// overflow, NaN) is returned as is, as pow does (pow(0, -m) = +inf).  Other
// exponents use pow.
template <typename R> SSPT_DH inline R pow_ex(R x, R ex)
{
    if (ex == R(0)) return R(1);
    if (ex < R(0) && ex >= R(-8) && ex == (R)(int)ex) {
        const int m = -(int)ex;
        R hi = x, lo = R(0);
        for (int i = 1; i < m; ++i) {
            const R h = hi * x;
            lo = std::fma(hi, x, -h) + lo * x;
            hi = h;
        }
        const R q = R(1) / hi;
        if (!(std::fabs(q) < R(INFINITY)) || q == R(0)) return q;
        const R e = std::fma(-q, hi, R(1)) - q * lo;
        return std::fma(q, e, q);
    }
    return std::pow(x, ex);
}

// power_law! (SphericalStokes_diapir.jl:259-265) with the optional log-space
// relaxation (CartesianStokes_diapir.jl:210-221) and cutoffs.
template <typename R>
SSPT_DH inline R power_law_pt(R tii, R ini, R ePL_old, R *ePL_new, const Args<R> &a, int flags)
{
    R pl = ini * pow_ex(tii / a.tC, a.ex);
    if (flags & F_RELAX) pl = std::exp(std::log(pl) * a.relax + std::log(ePL_old) * (R(1) - a.relax));
    *ePL_new = pl;
    R e = R(2) / (R(1) / ini + R(1) / pl);
    if (flags & F_CLAMP) {
        e = jmax(e, a.emin);   // eta_min = 0 / eta_max = +inf leave eta unchanged
        e = jmin(e, a.emax);
    }
    return e;
}

template <typename R> struct Thread {
    int tx, ty, j, k;
    int jc, kc;       // j,k clamped into the box for loads
    bool own;         // owned (written) position
    // registers carried along r
    R vr0, vr1, vt0, vt1, vp0, vp1;     // old V at planes s, s+1 (own column)
    R e0, e1, e2;                       // old eta at s, s+1, s+2
    R nvr, nvt, nvp, ne, ringJ, ringK;  // prefetched: V(s+2), eta(s+3), eta ring of plane s+2
    R sm0, sm1;                         // etaSM at s, s+1
    R nsm;                              // F_SMBUF: prefetched etaSM of the next step's plane
    R cRR, cTT, cPP, cDT, crho;         // sigma/DT/rho of plane s-1 (VR face s)
    R rt0, rt0n, rp0, rp0n;             // new tau_Rth/tau_Rph of face s (own, +1 neighbour)
    R pG_rt, pE_rt, pG_rp, pE_rp;       // partial 4-point sums of plane s for face s+1
    R gtp, etp;                         // 4-point G / eta averages for tau_thph(s)
    // values of the current step
    R oP, oRR, oTT, oPP, oRTold, oRPold, oTPold, orho, oini, oPL;
    R ogP, ofR, ofT, ofP;               // forcing of a correction sweep (step s+1 after phase 2)
    R cfT, cfP;                         // th/ph forcing of step s, kept for phase 4
    R nRR, nTT, nPP, nSRR, nSTT, nSPP, nDT;
    R nrt, nrp, ntp;
    R pG_rt2, pE_rt2, pG_rp2, pE_rp2, gtp2, etp2;
    unsigned long long mdv[3], mv[3];
};

template <typename R> SSPT_DH inline void thread_init(Thread<R> &T, const Args<R> &a, int tx, int ty, int kb, int jb)
{
    T.tx = tx;
    T.ty = SSPT_UNIFORM(ty);
    T.k = kb - 1 + tx;
    T.j = SSPT_UNIFORM(jb - 1 + ty);
    T.kc = clampi(T.k, 0, a.b.np);
    T.jc = clampi(T.j, 0, a.b.nt);
    T.own = tx >= 1 && tx <= TKX<R> - 2 && ty >= 1 && ty <= TJB<R> - 2 && T.j < a.b.nt && T.k < a.b.np;
    for (int c = 0; c < 3; ++c) T.mdv[c] = T.mv[c] = 0ull;
    T.vr0 = T.vr1 = T.vt0 = T.vt1 = T.vp0 = T.vp1 = R(0);
    T.e0 = T.e1 = T.e2 = R(1);
    T.sm0 = T.sm1 = R(1);
    T.cRR = T.cTT = T.cPP = T.cDT = T.crho = R(0);
    T.rt0 = T.rt0n = T.rp0 = T.rp0n = R(0);
    T.pG_rt = T.pE_rt = T.pG_rp = T.pE_rp = T.gtp = T.etp = R(1);
    T.nrt = T.nrp = T.ntp = R(0);
    T.ogP = T.ofR = T.ofT = T.ofP = T.cfT = T.cfP = R(0);
}

// Row/column ring of the eta plane p for the star of the block's edge threads.
template <typename R> SSPT_DH inline void load_ring(Thread<R> &T, const Args<R> &a, int p)
{
    const BoxG &b = a.b;
    const int ie = clampi(p, 0, b.nr - 1);
    const int jcc = clampi(T.jc, 0, b.nt - 1), kcc = clampi(T.kc, 0, b.np - 1);
    if (T.ty == 0) T.ringJ = fat_in(a.eta, a.ppe_in, b.o32(ie, clampi(T.j - 1, 0, b.nt - 1), kcc));
    if (T.ty == TJB<R> - 1) T.ringJ = fat_in(a.eta, a.ppe_in, b.o32(ie, clampi(T.j + 1, 0, b.nt - 1), kcc));
    if (T.tx == 0) T.ringK = fat_in(a.eta, a.ppe_in, b.o32(ie, jcc, clampi(T.k - 1, 0, b.np - 1)));
    if (T.tx == TKX<R> - 1) T.ringK = fat_in(a.eta, a.ppe_in, b.o32(ie, jcc, clampi(T.k + 1, 0, b.np - 1)));
}

template <typename R> SSPT_DH inline uint32_t own_centre(const Thread<R> &T, const BoxG &b, int s)
{
    return b.o32(clampi(s, 0, b.nr - 1), clampi(T.jc, 0, b.nt - 1), clampi(T.kc, 0, b.np - 1));
}

// Own old values of step s.  They are issued one step ahead (right after
// their previous use), so their latency hides behind a whole step.
template <typename R, int FL> SSPT_DH inline void load_centre_old(Thread<R> &T, const Args<R> &a, int s)
{
    const uint32_t oc = own_centre(T, a.b, s);
    T.oP = fat_in(a.P, a.pp_in, oc);
    T.oRR = fat_in(a.tRR, a.pp_in, oc);
    T.oTT = fat_in(a.tTT, a.pp_in, oc);
    T.oPP = fat_in(a.tPP, a.pp_in, oc);
    if (FL & F_CORR) {
        T.ogP = fat(a.gP, oc);
        T.ofR = fat(a.fVR, oc);
        T.ofT = fat(a.fVT, oc);
        T.ofP = fat(a.fVP, oc);
    } else {
        T.orho = fat(a.rho, oc);
    }
    T.oTPold = fat_in(a.tTP, a.pp_in, a.b.o32(clampi(s, 0, a.b.nr - 1), T.jc, T.kc));
}
template <typename R> SSPT_DH inline void load_face_old(Thread<R> &T, const Args<R> &a, int s)
{
    const uint32_t oe = a.b.o32(clampi(s + 1, 0, a.b.nr), T.jc, T.kc);
    T.oRTold = fat_in(a.tRT, a.pp_in, oe);
    T.oRPold = fat_in(a.tRP, a.pp_in, oe);
}
template <typename R, int FL> SSPT_DH inline void load_visc_old(Thread<R> &T, const Args<R> &a, int s)
{
    if (!(FL & F_CORR) && T.own) {
        const uint32_t oc = own_centre(T, a.b, s);
        T.oini = fat(a.ini, oc);
        T.oPL = (FL & F_RELAX) ? fat(a.ePL, oc) : R(0);
    }
}

// Loads before the march (start = first step): the V/eta window, the
// prefetch registers and the own old values of step `start`.
template <typename R, int FL> SSPT_DH inline void thread_prologue(Thread<R> &T, const Args<R> &a, int start)
{
    const BoxG &b = a.b;
    const int jcc = clampi(T.jc, 0, b.nt - 1), kcc = clampi(T.kc, 0, b.np - 1);
    T.e0 = fat_in(a.eta, a.ppe_in, b.o32(clampi(start, 0, b.nr - 1), jcc, kcc));
    T.e1 = fat_in(a.eta, a.ppe_in, b.o32(clampi(start + 1, 0, b.nr - 1), jcc, kcc));
    T.ne = fat_in(a.eta, a.ppe_in, b.o32(clampi(start + 2, 0, b.nr - 1), jcc, kcc));
    T.vr1 = fat_in(a.VR, a.pp_in, b.o32(clampi(start, 0, b.nr), T.jc, T.kc));
    T.vt1 = fat_in(a.VT, a.pp_in, b.o32(clampi(start, 0, b.nr - 1), T.jc, T.kc));
    T.vp1 = fat_in(a.VP, a.pp_in, b.o32(clampi(start, 0, b.nr - 1), T.jc, T.kc));
    T.nvr = fat_in(a.VR, a.pp_in, b.o32(clampi(start + 1, 0, b.nr), T.jc, T.kc));
    T.nvt = fat_in(a.VT, a.pp_in, b.o32(clampi(start + 1, 0, b.nr - 1), T.jc, T.kc));
    T.nvp = fat_in(a.VP, a.pp_in, b.o32(clampi(start + 1, 0, b.nr - 1), T.jc, T.kc));
    T.ringJ = T.ringK = R(0);
    T.nsm = (FL & F_SMBUF) ? fat(a.SM, b.o32(clampi(start + 1, 0, b.nr - 1), jcc, kcc)) : R(1);
    load_ring(T, a, start + 1);
    load_centre_old<R, FL>(T, a, start);
    load_face_old(T, a, start);
    load_visc_old<R, FL>(T, a, start);
}

// ---- phase 1: stage old V and eta of the coming plane in LDS (values
// prefetched during the previous step) and prefetch the plane after it
template <typename R, int FL>
SSPT_DH inline void phase1(Thread<R> &T, Lds<R> &L, const Args<R> &a, int s)
{
    const BoxG &b = a.b;
    const int nr = b.nr;
    // shift the r window: plane s+1 values were prefetched at step s-1
    T.vr0 = T.vr1;
    T.vt0 = T.vt1;
    T.vp0 = T.vp1;
    T.vr1 = T.nvr;
    T.vt1 = T.nvt;
    T.vp1 = T.nvp;
    T.e2 = T.ne;
    L.VR[T.ty][T.tx] = T.vr1;
    L.VT[T.ty][T.tx] = T.vt0;
    L.VP[T.ty][T.tx] = T.vp0;
    // eta plane s+1 with a one-point ring for the 7-point star of the halo
    L.ETA[T.ty + 1][T.tx + 1] = T.e1;
    if (T.ty == 0) L.ETA[0][T.tx + 1] = T.ringJ;
    if (T.ty == TJB<R> - 1) L.ETA[TJB<R> + 1][T.tx + 1] = T.ringJ;
    if (T.tx == 0) L.ETA[T.ty + 1][0] = T.ringK;
    if (T.tx == TKX<R> - 1) L.ETA[T.ty + 1][TKX<R> + 1] = T.ringK;
    // prefetch for step s+1
    const int jcc = clampi(T.jc, 0, b.nt - 1), kcc = clampi(T.kc, 0, b.np - 1);
    T.nvr = fat_in(a.VR, a.pp_in, b.o32(clampi(s + 2, 0, nr), T.jc, T.kc));
    T.nvt = fat_in(a.VT, a.pp_in, b.o32(clampi(s + 2, 0, nr - 1), T.jc, T.kc));
    T.nvp = fat_in(a.VP, a.pp_in, b.o32(clampi(s + 2, 0, nr - 1), T.jc, T.kc));
    T.ne = fat_in(a.eta, a.ppe_in, b.o32(clampi(s + 3, 0, nr - 1), jcc, kcc));
    load_ring(T, a, s + 2);
}

// 7-point star max of eta at (p, cj, ck) with in-plane values from LDS
// (@maxloc expansion, SphericalStokes_mxpu.jl:295-297).  The hardware max
// (fmax) equals Julia's max for the positive, finite viscosities of a valid
// state; it does not propagate a NaN viscosity into etaSM (a NaN still
// reaches V through eta itself, so the NaN exit test still fires).
template <typename R> SSPT_DH inline R star(const Lds<R> &L, int ty, int tx, R em, R ep)
{
    const R c = L.ETA[ty + 1][tx + 1];
    return std::fmax(std::fmax(std::fmax(std::fmax(em, ep), c), std::fmax(L.ETA[ty][tx + 1], L.ETA[ty + 2][tx + 1])),
                     std::fmax(L.ETA[ty + 1][tx], L.ETA[ty + 1][tx + 2]));
}

// maxloc! followed by copy_BC! at cell (i, j, k) of a centre field src: the
// 7-point star maximum at the cell clamped into the interior, with the
// nesting of the star above.  Run over every cell it fills the SM field that
// F_SMBUF reads.
template <typename R> SSPT_DH inline R star_bc_at(const BoxG &b, const R *src, int i, int j, int k)
{
    const int ci = clampi(i, 1, b.nr - 2), cj = clampi(j, 1, b.nt - 2), ck = clampi(k, 1, b.np - 2);
    const long long o = b.off(ci, cj, ck);
    return std::fmax(std::fmax(std::fmax(std::fmax(src[o - b.si], src[o + b.si]), src[o]),
                               std::fmax(src[o - b.sj], src[o + b.sj])),
                     std::fmax(src[o - 1], src[o + 1]));
}

// Spread etaSM (the build option ss_params.etasm_spread = m): one of three
// separable passes (axis 0 = r, 1 = th, 2 = phi) of
//     out(i) = max over |a| <= m of src(clamp(i + a e_axis)) * w[|a|],
// w[0] = 1, w[a] = w[a-1] * 0.1.  After the three passes a cell's etaSM is
// at least a tenth of its neighbour's within m cells of a stiff cell, so
// the pseudo-time steps of adjacent cells never differ by more than that
// factor, whatever the viscosity jump.
constexpr int SPREAD_MAX = 8;
constexpr double SPREAD_DECAY = 0.1;
template <typename R> struct SpreadW {
    R w[SPREAD_MAX + 1];
};
template <typename R> SSPT_DH inline SpreadW<R> spread_weights()
{
    SpreadW<R> s;
    s.w[0] = R(1);
    for (int a = 1; a <= SPREAD_MAX; ++a) s.w[a] = s.w[a - 1] * R(SPREAD_DECAY);
    return s;
}
template <typename R>
SSPT_DH inline R spread_at(const BoxG &b, const R *src, int i, int j, int k, int axis, int m, const SpreadW<R> &w)
{
    const int n = axis == 0 ? b.nr : (axis == 1 ? b.nt : b.np);
    const int c = axis == 0 ? i : (axis == 1 ? j : k);
    const long long st = axis == 0 ? b.si : (axis == 1 ? b.sj : 1);
    const long long o = b.off(i, j, k) - (long long)c * st;   // the line's start
    R v = src[o + (long long)c * st];
    for (int a = 1; a <= m; ++a) {
        v = std::fmax(v, src[o + (long long)clampi(c - a, 0, n - 1) * st] * w.w[a]);
        v = std::fmax(v, src[o + (long long)clampi(c + a, 0, n - 1) * st] * w.w[a]);
    }
    return v;
}

// ---- phase 2: etaSM(s+1); C(s) = P, tau_nn, DT at centres; tau_thph(s); VR face s
template <typename R, int FL>
SSPT_DH inline void phase2(Thread<R> &T, Lds<R> &L, const Args<R> &a, int s, bool wr)
{
    const BoxG &b = a.b;
    const TabsG<R> &t = a.t;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const int p = s + 1;
    // etaSM of plane p after copy_BC!: star at the clamped position
    if (FL & F_SMBUF) {
        T.sm1 = T.nsm;
        T.nsm = fat(a.SM, b.o32(clampi(p + 1, 0, nr - 1), clampi(T.jc, 0, nt - 1), clampi(T.kc, 0, np - 1)));
    } else if (p >= 1 && p <= nr - 2) {
        const int cj = clampi(T.j, 1, nt - 2), ck = clampi(T.k, 1, np - 2);
        if (cj == T.j && ck == T.k) {
            T.sm1 = star(L, T.ty, T.tx, T.e0, T.e2);
        } else {
            const int dty = cj - T.j, dtx = ck - T.k;
            const int ty = clampi(T.ty + dty, 0, TJB<R> - 1), tx = clampi(T.tx + dtx, 0, TKX<R> - 1);
            const R em = fat_in(a.eta, a.ppe_in, b.o32(p - 1, cj, ck)), ep = fat_in(a.eta, a.ppe_in, b.o32(p + 1, cj, ck));
            T.sm1 = star(L, ty, tx, em, ep);
        }
    } else if (p == nr - 1) {
        T.sm1 = T.sm0;
    }
    // etaSM(0) = etaSM(1) after copy_BC!; a filled SM field (F_SMBUF) holds
    // plane 0's own values, which sm0 already has
    if (!(FL & F_SMBUF) && s == 0) T.sm0 = T.sm1;
    L.SM[T.ty][T.tx] = T.sm1;

    const bool centre = T.j >= 0 && T.j < nt && T.k >= 0 && T.k < np;
    // C(s): compute_P! centre part
    if (s >= 0 && s <= nr - 1 && centre && T.tx < TKX<R> - 1 && T.ty < TJB<R> - 1) {
        const R Rv = tload(t.tr, s * NTR + TR_Rc), iR = tload(t.tr, s * NTR + TR_invRc);
        const R dVRr = (T.vr1 - T.vr0) * a._dr;
        const R dVTt = (iR * (L.VT[T.ty + 1][T.tx] - T.vt0)) * a._dth;
        const R dVPp = (tload(t.t2, (s * (nt + 1) + T.j) * NT2 + T2_c) * (L.VP[T.ty][T.tx + 1] - T.vp0)) * a._dph;
        const R avVR = (T.vr0 + T.vr1) * R(0.5);
        const R avVT = (T.vt0 + L.VT[T.ty + 1][T.tx]) * R(0.5);
        const R ct = tload(t.tt, T.j * NTT + TT_cotc);
        const R vr_R = div_rcp(avVR, Rv, iR), vt_R = div_rcp(avVT, Rv, iR);
        const R divV = (((dVRr + dVTt) + dVPp) + div_rcp(R(2) * avVR, Rv, iR)) + vt_R * ct;
        const R third = R(1) / R(3);
        const R Drr = dVRr - third * divV;
        const R Dtt = (dVTt + vr_R) - third * divV;
        const R Dpp = ((dVPp + vr_R) + vt_R * ct) - third * divV;
        const R KB = a.kb * T.sm0, GS = a.gs * T.sm0;
        const R den = R(1) + GS / T.e0;
        const R iden = R(1) / den;
        const R nP = (FL & F_CORR) ? T.oP - KB * (divV + T.ogP) : T.oP - KB * divV;
        T.nRR = div_rcp(T.oRR + (GS * R(2)) * Drr, den, iden);
        T.nTT = div_rcp(T.oTT + (GS * R(2)) * Dtt, den, iden);
        T.nPP = div_rcp(T.oPP + (GS * R(2)) * Dpp, den, iden);
        T.nDT = a.vpdt2 / (KB + div_rcp(GS, den, iden));
        T.nSRR = -nP + T.nRR;
        T.nSTT = -nP + T.nTT;
        T.nSPP = -nP + T.nPP;
        if (wr && T.own) {
            const uint32_t o = b.o32(s, T.j, T.k);
            st_out<R, FL>(a.oP, a.P, o, a.pp_out, nP);
            st_out<R, FL>(a.oRR, a.tRR, o, a.pp_out, T.nRR);
            st_out<R, FL>(a.oTT, a.tTT, o, a.pp_out, T.nTT);
            st_out<R, FL>(a.oPP, a.tPP, o, a.pp_out, T.nPP);
            if ((FL & F_LAST) && !(FL & F_CORR)) {
                fat(a.DT, o) = T.nDT;
                if (!(FL & F_SMBUF)) fat(a.SM, o) = T.sm0;   // F_SMBUF: the SM field already holds it
            }
        }
    } else {
        T.nRR = T.nTT = T.nPP = R(0);
        T.nSRR = T.nSTT = T.nSPP = R(0);
        T.nDT = R(0);
    }
    L.STT[T.ty][T.tx] = T.nSTT;
    L.SPP[T.ty][T.tx] = T.nSPP;
    L.DT[T.ty][T.tx] = T.nDT;

    // tau_thph(s): compute_P! edge part, plane s, th-face j, phi-face k
    T.ntp = R(0);
    if (s >= 1 && s <= nr - 2 && T.tx >= 1 && T.ty >= 1 && T.j >= 1 && T.j <= nt - 1 && T.k >= 1 && T.k <= np - 1) {
        const R dVP = T.vp0 - L.VP[T.ty - 1][T.tx];
        const R dVT = T.vt0 - L.VT[T.ty][T.tx - 1];
        const R avVP = (L.VP[T.ty - 1][T.tx] + T.vp0) * R(0.5);
        const R D = ((tload(t.tr, s * NTR + TR_invRc4) * dVP) * a._dth + (tload(t.t2, (s * (nt + 1) + T.j) * NT2 + T2_c4) * dVT) * a._dph) -
                         div_rcp(avVP, tload(t.tr, s * NTR + TR_Rc4), tload(t.tr, s * NTR + TR_invRc4)) * tload(t.tt, T.j * NTT + TT_cotf4);
        T.ntp = (T.oTPold + T.gtp * D) / (R(1) + T.gtp / T.etp);
        if (wr && T.own) st_out<R, FL>(a.oTP, a.tTP, b.o32(s, T.j, T.k), a.pp_out, T.ntp);
    }
    L.TP[T.ty][T.tx] = T.ntp;

    // VR at r-face s (compute_V!): needs C(s-1), C(s), tau_Rth/tau_Rph of face s
    if (wr && T.own && s >= 1 && s <= nr - 1 && T.j >= 1 && T.j <= nt - 2 && T.k >= 1 && T.k <= np - 2) {
        const R Rx = tload(t.tr, s * NTR + TR_Rf), iRx = tload(t.tr, s * NTR + TR_invRf);
        const R sRR = (T.cRR + T.nSRR) * R(0.5);
        const R sTT = (T.cTT + T.nSTT) * R(0.5);
        const R sPP = (T.cPP + T.nSPP) * R(0.5);
        const R aRT = (T.rt0 + T.rt0n) * R(0.5);
        const R dsig = (((((T.nSRR - T.cRR) * a._dr + (iRx * (T.rt0n - T.rt0)) * a._dth) +
                            (tload(t.t2, (s * (nt + 1) + T.j) * NT2 + T2_f) * (T.rp0n - T.rp0)) * a._dph) +
                           div_rcp(R(2) * sRR, Rx, iRx)) -
                          div_rcp(sTT + sPP, Rx, iRx)) +
                         div_rcp(aRT, Rx, iRx) * tload(t.tt, T.j * NTT + TT_cotc2);
        R dV;
        if (FL & F_CORR) {
            dV = dsig + T.ofR;
        } else {
            const R rg = (T.crho * a.g + T.orho * a.g) * R(0.5);
            dV = dsig - rg;
        }
        const R dt = (T.cDT + T.nDT) * R(0.5);
        const R v = T.vr0 + dV * dt;
        const uint32_t o = b.o32(s, T.j, T.k);
        st_out<R, FL>(a.oVR, a.VR, o, a.pp_out, v);
        if (FL & F_CHECK) {
            const R d = v - T.vr0;
            if (!(FL & F_CORR)) fat(a.VRe, o) = d;
            const unsigned long long ud = absbits(d), uv = absbits(v);
            T.mdv[0] = ud > T.mdv[0] ? ud : T.mdv[0];
            T.mv[0] = uv > T.mv[0] ? uv : T.mv[0];
        } else if ((FL & F_LAST) && !(FL & F_CORR)) {
            fat(a.VRe, o) = T.vr0;
        }
    }
    T.crho = T.orho;
    if (FL & F_CORR) {
        T.cfT = T.ofT;
        T.cfP = T.ofP;
    }
    load_centre_old<R, FL>(T, a, s + 1);
}

// ---- phase 3: tau_Rth / tau_Rph on r-face s+1; 4-point sums of plane s+1
template <typename R, int FL>
SSPT_DH inline void phase3(Thread<R> &T, Lds<R> &L, const Args<R> &a, int s, bool wr)
{
    const BoxG &b = a.b;
    const TabsG<R> &t = a.t;
    const int nr = b.nr, nt = b.nt, np = b.np;
    const int f = s + 1;
    const R gs = a.gs;
    const bool inner = T.tx >= 1 && T.ty >= 1;
    // sums over plane s+1 (etaSM(s+1) in L.SM, eta(s+1) in L.ETA)
    if (inner) {
        const R smJ = L.SM[T.ty - 1][T.tx], smK = L.SM[T.ty][T.tx - 1], smJK = L.SM[T.ty - 1][T.tx - 1];
        const R eJ = L.ETA[T.ty][T.tx + 1], eK = L.ETA[T.ty + 1][T.tx], eJK = L.ETA[T.ty][T.tx];
        T.pG_rt2 = gs * smJ + gs * T.sm1;
        T.pE_rt2 = eJ + T.e1;
        T.pG_rp2 = gs * smK + gs * T.sm1;
        T.pE_rp2 = eK + T.e1;
        T.gtp2 = (((gs * smJK + gs * smJ) + gs * smK) + gs * T.sm1) * R(0.25);
        T.etp2 = (((eJK + eJ) + eK) + T.e1) * R(0.25);
        // etaSM(0,:,:) = etaSM(1,:,:) after copy_BC!: plane-0 sums of gs*etaSM
        // are those of plane 1 (with F_SMBUF plane 0 was read as it is)
        if (!(FL & F_SMBUF) && s == 0) {
            T.pG_rt = T.pG_rt2;
            T.pG_rp = T.pG_rp2;
        }
    }
    T.nrt = R(0);
    T.nrp = R(0);
    if (f >= 1 && f <= nr - 1 && inner) {
        const R Re = tload(t.tr, f * NTR + TR_Re), iRe = tload(t.tr, f * NTR + TR_invRe);
        if (T.j >= 1 && T.j <= nt - 1 && T.k >= 1 && T.k <= np - 2) {
            const R G = ((T.pG_rt + gs * L.SM[T.ty - 1][T.tx]) + gs * T.sm1) * R(0.25);
            const R E = ((T.pE_rt + L.ETA[T.ty][T.tx + 1]) + T.e1) * R(0.25);
            const R D = ((T.vt1 - T.vt0) * a._dr + (iRe * (T.vr1 - L.VR[T.ty - 1][T.tx])) * a._dth) -
                             div_rcp((T.vt0 + T.vt1) * R(0.5), Re, iRe);
            T.nrt = (T.oRTold + G * D) / (R(1) + G / E);
            if (wr && T.own) st_out<R, FL>(a.oRT, a.tRT, b.o32(f, T.j, T.k), a.pp_out, T.nrt);
        }
        if (T.j >= 1 && T.j <= nt - 2 && T.k >= 1 && T.k <= np - 1) {
            const R G = ((T.pG_rp + gs * L.SM[T.ty][T.tx - 1]) + gs * T.sm1) * R(0.25);
            const R E = ((T.pE_rp + L.ETA[T.ty + 1][T.tx]) + T.e1) * R(0.25);
            const R D = ((T.vp1 - T.vp0) * a._dr + (tload(t.t2, (f * (nt + 1) + T.j) * NT2 + T2_e) * (T.vr1 - L.VR[T.ty][T.tx - 1])) * a._dph) -
                             div_rcp((T.vp0 + T.vp1) * R(0.5), Re, iRe);
            T.nrp = (T.oRPold + G * D) / (R(1) + G / E);
            if (wr && T.own) st_out<R, FL>(a.oRP, a.tRP, b.o32(f, T.j, T.k), a.pp_out, T.nrp);
        }
    }
    L.RT[T.ty][T.tx] = T.nrt;
    L.RP[T.ty][T.tx] = T.nrp;
    load_face_old(T, a, s + 1);
}

// ---- phase 4: Vth(s), Vph(s), tII(s) -> eta(s); shift carries
template <typename R, int FL>
SSPT_DH inline void phase4(Thread<R> &T, Lds<R> &L, const Args<R> &a, int s, bool wr)
{
    const BoxG &b = a.b;
    const TabsG<R> &t = a.t;
    const int nr = b.nr, nt = b.nt, np = b.np;
    if (wr && T.own && s >= 1 && s <= nr - 2) {
        const R Ry = tload(t.tr, s * NTR + TR_Rc2), iRy = tload(t.tr, s * NTR + TR_invRc2);
        // Vth at (s, j, k)
        if (T.j >= 1 && T.j <= nt - 1 && T.k >= 1 && T.k <= np - 2) {
            const R sTTm = L.STT[T.ty - 1][T.tx], sPPm = L.SPP[T.ty - 1][T.tx];
            const R sTT = (sTTm + T.nSTT) * R(0.5), sPP = (sPPm + T.nSPP) * R(0.5);
            const R aRT = (T.rt0 + T.nrt) * R(0.5);
            const R tpK = L.TP[T.ty][T.tx + 1];
            R dV = ((((T.nrt - T.rt0) * a._dr + (iRy * (T.nSTT - sTTm)) * a._dth) +
                                (tload(t.t2, (s * (nt + 1) + T.j) * NT2 + T2_y) * (tpK - T.ntp)) * a._dph) +
                               div_rcp(R(3) * aRT, Ry, iRy)) +
                              div_rcp(sTT - sPP, Ry, iRy) * tload(t.tt, T.j * NTT + TT_cotf2);
            if (FL & F_CORR) dV = dV + T.cfT;
            const R dt = (L.DT[T.ty - 1][T.tx] + T.nDT) * R(0.5);
            const R v = T.vt0 + dV * dt;
            const uint32_t o = b.o32(s, T.j, T.k);
            st_out<R, FL>(a.oVT, a.VT, o, a.pp_out, v);
            if (FL & F_CHECK) {
                const R d = v - T.vt0;
                if (!(FL & F_CORR)) fat(a.VTe, o) = d;
                const unsigned long long ud = absbits(d), uv = absbits(v);
                T.mdv[1] = ud > T.mdv[1] ? ud : T.mdv[1];
                T.mv[1] = uv > T.mv[1] ? uv : T.mv[1];
            } else if ((FL & F_LAST) && !(FL & F_CORR)) {
                fat(a.VTe, o) = T.vt0;
            }
        }
        // Vph at (s, j, k)
        if (T.j >= 1 && T.j <= nt - 2 && T.k >= 1 && T.k <= np - 1) {
            const R tpJ = L.TP[T.ty + 1][T.tx];
            const R sPPm = L.SPP[T.ty][T.tx - 1];
            const R aRP = (T.rp0 + T.nrp) * R(0.5);
            const R aTP = (T.ntp + tpJ) * R(0.5);
            R dV = ((((T.nrp - T.rp0) * a._dr + (iRy * (tpJ - T.ntp)) * a._dth) +
                                (tload(t.t2, (s * (nt + 1) + T.j) * NT2 + T2_z) * (T.nSPP - sPPm)) * a._dph) +
                               div_rcp(R(3) * aRP, Ry, iRy)) +
                              div_rcp(R(2) * aTP, Ry, iRy) * tload(t.tt, T.j * NTT + TT_cotc2);
            if (FL & F_CORR) dV = dV + T.cfP;
            const R dt = (L.DT[T.ty][T.tx - 1] + T.nDT) * R(0.5);
            const R v = T.vp0 + dV * dt;
            const uint32_t o = b.o32(s, T.j, T.k);
            st_out<R, FL>(a.oVP, a.VP, o, a.pp_out, v);
            if (FL & F_CHECK) {
                const R d = v - T.vp0;
                if (!(FL & F_CORR)) fat(a.VPe, o) = d;
                const unsigned long long ud = absbits(d), uv = absbits(v);
                T.mdv[2] = ud > T.mdv[2] ? ud : T.mdv[2];
                T.mv[2] = uv > T.mv[2] ? uv : T.mv[2];
            } else if ((FL & F_LAST) && !(FL & F_CORR)) {
                fat(a.VPe, o) = T.vp0;
            }
        }
        // tII and the new viscosity of interior cells (eta is frozen in a correction sweep)
        if (!(FL & F_CORR) && T.j >= 1 && T.j <= nt - 2 && T.k >= 1 && T.k <= np - 2) {
            const R x = (((T.rt0 + T.rt0n) + T.nrt) + L.RT[T.ty + 1][T.tx]) * R(0.25);
            const R y = (((T.rp0 + T.rp0n) + T.nrp) + L.RP[T.ty][T.tx + 1]) * R(0.25);
            const R z = (((T.ntp + L.TP[T.ty][T.tx + 1]) + L.TP[T.ty + 1][T.tx]) + L.TP[T.ty + 1][T.tx + 1]) * R(0.25);
            const R tii =
                std::sqrt((((R(1) / R(2)) * ((T.nRR * T.nRR + T.nTT * T.nTT) + T.nPP * T.nPP) + x * x) + y * y) + z * z);
            R pl;
            const R e = power_law_pt(tii, T.oini, T.oPL, &pl, a, FL);
            const uint32_t o = b.o32(s, T.j, T.k);
            st_out<R, FL>(a.oEta, a.eta, o, a.ppe_out, e);
            if (FL & F_RELAX) fat(a.oPL, o) = pl;
            if (FL & F_LAST) fat(a.TII, o) = tii;
            if (T.k == 1) fat(a.TBP, tbp_off(b, 0, s, T.j)) = tii;
            if (T.k == np - 2) fat(a.TBP, tbp_off(b, 1, s, T.j)) = tii;
        }
    }
    // carries for step s+1
    if (T.ty + 1 < TJB<R>) T.rt0n = L.RT[T.ty + 1][T.tx];
    if (T.tx + 1 < TKX<R>) T.rp0n = L.RP[T.ty][T.tx + 1];
    T.rt0 = T.nrt;
    T.rp0 = T.nrp;
    T.cRR = T.nSRR;
    T.cTT = T.nSTT;
    T.cPP = T.nSPP;
    T.cDT = T.nDT;
    T.e0 = T.e1;
    T.e1 = T.e2;
    T.sm0 = T.sm1;
    T.pG_rt = T.pG_rt2;
    T.pE_rt = T.pE_rt2;
    T.pG_rp = T.pG_rp2;
    T.pE_rp = T.pE_rp2;
    T.gtp = T.gtp2;
    T.etp = T.etp2;
    load_visc_old<R, FL>(T, a, s + 1);
}

// ---------------------------------------------------------------- boundary pass
// Final value of a velocity component at a ghost position after the
// BC_R! -> BC_th! -> BC_ph! sequence: the last BC among the directions in
// which the position is a ghost decides it, from the final value at the
// inward neighbour (which is a ghost only in earlier directions).
// comp: 0 = VR, 1 = Vth, 2 = Vph.
SSPT_DH inline bool ghost_dir(const BoxG &b, int comp, int dir, int idx, int *inward)
{
    int n;
    if (dir == 0) n = (comp == 0) ? b.nr + 1 : b.nr;
    else if (dir == 1) n = (comp == 1) ? b.nt + 1 : b.nt;
    else n = (comp == 2) ? b.np + 1 : b.np;
    if (idx == 0) { *inward = 1; return true; }
    if (idx == n - 1) { *inward = n - 2; return true; }
    return false;
}

// affine BC map: value = a*v + c  (a in {-1,0,1})
template <typename R> SSPT_DH inline R bc_apply(const Args<R> &A, int comp, int dir, int side, int j, R v)
{
    const BoxG &b = A.b;
    const int mode = A.bc_mode; // 0 shear, 1 free slip, 2 pure shear
    if (dir == 0) return comp == 0 ? -v : v;
    if (dir == 1) {
        if (comp == 1) {
            if (mode == 2) return R(2) * (side ? A.vth1 : A.vth0) - v;
            return -v;
        }
        if (mode == 0) {
            if (comp == 0) return R(0);
            const R th = side ? A.t.th[b.nt - 1] : A.t.th[0];
            return ((th - R(PI_ / 2.0)) * A.r) * A.sh;
        }
        return v;
    }
    // dir == 2
    if (comp == 2) {
        if (mode == 0) return ((R(2) * (A.t.th[j] - R(PI_ / 2.0))) * A.r) * A.sh - v;
        if (mode == 1) return -v;
        const R ph = side ? A.t.ph[b.np - 1] : A.t.ph[0];
        return R(2) * ((ph * A.r) * A.sh) - v;
    }
    if (mode == 0) return R(0);
    return v;
}

template <typename R> SSPT_DH inline bool bc_is_const(const Args<R> &A, int comp, int dir)
{
    if (A.bc_mode != 0) return false;
    if (dir == 1) return comp != 1;
    if (dir == 2) return comp != 2;
    return false;
}

template <typename R> SSPT_DH inline R ghost_value(const Args<R> &A, const R *V, int comp, int i, int j, int k)
{
    const BoxG &b = A.b;
    int ii[3] = {i, j, k};
    int dirs[3], sides[3], nd = 0;
    // collect directions from last (phi) to first (r); stop at a constant BC
    bool konst = false;
    for (int d = 2; d >= 0; --d) {
        int inw;
        if (ghost_dir(b, comp, d, ii[d], &inw)) {
            dirs[nd] = d;
            sides[nd] = ii[d] == 0 ? 0 : 1;
            ++nd;
            ii[d] = inw;
            if (bc_is_const(A, comp, d)) { konst = true; break; }
        }
    }
    R v = konst ? R(0) : V[b.off(ii[0], ii[1], ii[2])];
    for (int q = nd - 1; q >= 0; --q) {
        // j of the position at which this BC is applied (ii[1] restored up to this level)
        v = bc_apply(A, comp, dirs[q], sides[q], j, v);
    }
    return v;
}

// tII of interior cell (i,j,k) from the new stresses (compute_TII!)
template <typename R> SSPT_DH inline R tii_inner(const Args<R> &a, int i, int j, int k)
{
    const BoxG &b = a.b;
    const long long o = b.off(i, j, k);
    const R A = a.oRR[o], B = a.oTT[o], C = a.oPP[o];
    const R *RT = a.oRT, *RP = a.oRP, *TP = a.oTP;
    const R x = (((RT[o] + RT[o + b.sj]) + RT[o + b.si]) + RT[o + b.si + b.sj]) * R(0.25);
    const R y = (((RP[o] + RP[o + 1]) + RP[o + b.si]) + RP[o + b.si + 1]) * R(0.25);
    const R z = (((TP[o] + TP[o + 1]) + TP[o + b.sj]) + TP[o + b.sj + 1]) * R(0.25);
    return std::sqrt((((R(1) / R(2)) * ((A * A + B * B) + C * C) + x * x) + y * y) + z * z);
}

template <typename R> SSPT_DH inline void ghost_v_point(const Args<R> &a, int comp, int i, int j, int k, unsigned long long *mdv,
                                  unsigned long long *mv)
{
    const BoxG &b = a.b;
    const R *Vn = comp == 0 ? a.oVR : (comp == 1 ? a.oVT : a.oVP);
    R *Vw = comp == 0 ? a.oVR : (comp == 1 ? a.oVT : a.oVP);
    const R *Vo = comp == 0 ? a.VR : (comp == 1 ? a.VT : a.VP);
    R *Ve = comp == 0 ? a.VRe : (comp == 1 ? a.VTe : a.VPe);
    const R v = ghost_value(a, Vn, comp, i, j, k);
    const long long o = b.off(i, j, k);
    Vw[o] = v;
    if (a.flags & F_CHECK) {
        const R d = v - Vo[o];
        if (!(a.flags & F_CORR)) Ve[o] = d;
        const unsigned long long ud = absbits(d), uv = absbits(v);
        *mdv = ud > *mdv ? ud : *mdv;
        *mv = uv > *mv ? uv : *mv;
    } else if ((a.flags & F_LAST) && !(a.flags & F_CORR)) {
        Ve[o] = Vo[o];
    }
}

// Viscosity of boundary cell (i,j,k) from tII at the clamped interior cell
// (copy_BC!(tII) then power_law!).  phi_side >= 0: (i,j,k) lies on that phi
// face with 1 <= i <= nr-2, 1 <= j <= nt-2, and its clamped cell's tII was
// stored by the main kernel; otherwise it is evaluated from the stresses.
template <typename R> SSPT_DH inline void boundary_eta_point(const Args<R> &a, int i, int j, int k, int phi_side = -1)
{
    const BoxG &b = a.b;
    const int ci = clampi(i, 1, b.nr - 2), cj = clampi(j, 1, b.nt - 2), ck = clampi(k, 1, b.np - 2);
    const R tii = phi_side >= 0 ? fat(a.TBP, tbp_off(b, phi_side, i, j)) : tii_inner(a, ci, cj, ck);
    const long long o = b.off(i, j, k);
    R pl;
    const R oldpl = (a.flags & F_RELAX) ? a.ePL[o] : R(0);
    const R e = power_law_pt(tii, a.ini[o], oldpl, &pl, a, a.flags);
    a.oEta[o] = e;
    if (a.flags & F_RELAX) a.oPL[o] = pl;
    if (a.flags & F_LAST) a.TII[o] = tii;
}

} // namespace fused
} // namespace sspt
