// sspt_tables.h -- host construction of the FP64 metric tables.
//
// Each entry is computed with the association of the reference macros
// (SphericalStokes_diapir.jl:162-183, 271-276), e.g. @avxyi_R =
// (((R[i]+R[i])+R[i+1])+R[i+1])*0.25, so that the device kernels, which read
// these values instead of recomputing them, round exactly as the reference
// expressions do.  Entries outside a table's valid range are 0 and unread.
#pragma once
#include <cmath>
#include <cstddef>
#include <vector>

namespace sspt {

template <class BoxT> inline size_t tables_elems(const BoxT &b)
{
    const size_t n1 = b.nr + 1, n2 = b.nt + 1;
    return 10 * n1 + 4 * n2 + 6 * n1 * n2 + n2 + (b.np + 1);
}

template <class BoxT>
inline void build_tables(const BoxT &b, const double *R, const double *th, const double *ph, std::vector<double> &h)
{
    const int nr = b.nr, nt = b.nt, np = b.np;
    const size_t n1 = nr + 1, n2 = nt + 1;
    h.assign(tables_elems(b), 0.0);
    double *p = h.data();
    auto take = [&](size_t n) { double *q = p; p += n; return q; };
    double *Rc = take(n1), *invRc = take(n1), *Rf = take(n1), *invRf = take(n1);
    double *Re = take(n1), *invRe = take(n1), *Rc4 = take(n1), *invRc4 = take(n1);
    double *Rc2 = take(n1), *invRc2 = take(n1);
    double *cotc = take(n2), *cotf4 = take(n2), *cotc2 = take(n2), *cotf2 = take(n2);
    double *rs_c = take(n1 * n2), *rs_e = take(n1 * n2), *rs_c4 = take(n1 * n2);
    double *rs_f = take(n1 * n2), *rs_y = take(n1 * n2), *rs_z = take(n1 * n2);
    double *thc = take(n2), *phc = take(np + 1);
    for (int i = 0; i < nr; ++i) {
        Rc[i] = R[i];
        invRc[i] = 1.0 / R[i];
        Rc4[i] = (((R[i] + R[i]) + R[i]) + R[i]) * 0.25;
        invRc4[i] = 1.0 / Rc4[i];
        Rc2[i] = (R[i] + R[i]) * 0.5;
        invRc2[i] = 1.0 / Rc2[i];
    }
    for (int i = 1; i < nr; ++i) {
        Rf[i] = (R[i - 1] + R[i]) * 0.5;
        invRf[i] = 1.0 / Rf[i];
        Re[i] = (((R[i - 1] + R[i - 1]) + R[i]) + R[i]) * 0.25;
        invRe[i] = 1.0 / Re[i];
    }
    std::vector<double> sc(n2, 0.0), s4c(n2, 0.0), s4f(n2, 0.0), s2c(n2, 0.0), s2f(n2, 0.0);
    for (int j = 0; j < nt; ++j) {
        thc[j] = th[j];
        sc[j] = sin(th[j]);
        cotc[j] = 1.0 / tan(th[j]);
        const double t4 = (((th[j] + th[j]) + th[j]) + th[j]) * 0.25;
        s4c[j] = sin(t4);
        const double t2 = (th[j] + th[j]) * 0.5;
        s2c[j] = sin(t2);
        cotc2[j] = 1.0 / tan(t2);
    }
    for (int j = 1; j < nt; ++j) {
        const double t4 = (((th[j - 1] + th[j - 1]) + th[j]) + th[j]) * 0.25;
        s4f[j] = sin(t4);
        cotf4[j] = 1.0 / tan(t4);
        const double t2 = (th[j - 1] + th[j]) * 0.5;
        s2f[j] = sin(t2);
        cotf2[j] = 1.0 / tan(t2);
    }
    for (int k = 0; k < np; ++k) phc[k] = ph[k];
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nt; ++j) {
            const size_t q = (size_t)i * n2 + j;
            rs_c[q] = (1.0 / Rc[i]) / sc[j];
            rs_c4[q] = (j >= 1) ? (1.0 / Rc4[i]) / s4f[j] : 0.0;
            rs_y[q] = (j >= 1) ? (1.0 / Rc2[i]) / s2f[j] : 0.0;
            rs_z[q] = (1.0 / Rc2[i]) / s2c[j];
            if (i >= 1) {
                rs_e[q] = (1.0 / Re[i]) / s4c[j];
                rs_f[q] = (1.0 / Rf[i]) / s2c[j];
            }
        }
}

// Interleaved records for the fused kernel: per r index 16 doubles
// (Rc invRc Rf invRf Re invRe Rc4 invRc4 Rc2 invRc2, padding), per th index 4
// (cotc cotf4 cotc2 cotf2), per (i,j) 8 ((1/R)/sin(th) for the centre, the
// tau_Rph edge, the tau_thph edge and the VR, Vth, Vph faces, padding).  The
// values are copied from build_tables' arrays, bit for bit.
constexpr int FT_NTR = 16, FT_NTT = 4, FT_NT2 = 8;

template <class BoxT> inline size_t fused_tables_elems(const BoxT &b)
{
    const size_t n1 = b.nr + 1, n2 = b.nt + 1;
    return FT_NTR * n1 + FT_NTT * n2 + FT_NT2 * n1 * n2;
}

template <class BoxT>
inline void build_fused_tables(const BoxT &b, const std::vector<double> &h, std::vector<double> &f)
{
    const size_t n1 = b.nr + 1, n2 = b.nt + 1;
    f.assign(fused_tables_elems(b), 0.0);
    const double *src = h.data();
    for (int q = 0; q < 10; ++q, src += n1)
        for (size_t i = 0; i < n1; ++i) f[i * FT_NTR + q] = src[i];
    double *tt = f.data() + FT_NTR * n1;
    for (int q = 0; q < 4; ++q, src += n2)
        for (size_t j = 0; j < n2; ++j) tt[j * FT_NTT + q] = src[j];
    double *t2 = tt + FT_NTT * n2;
    for (int q = 0; q < 6; ++q, src += n1 * n2)
        for (size_t ij = 0; ij < n1 * n2; ++ij) t2[ij * FT_NT2 + q] = src[ij];
}

} // namespace sspt
