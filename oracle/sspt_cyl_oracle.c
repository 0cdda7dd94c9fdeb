/*
 * sspt_cyl_oracle.c -- CPU restatement of the cylindrical twin of the
 * diapir solver, scripts_diapir/CylindricalStokes_diapir.jl.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the cylindrical entry
 * points of libsspt.so.  The product never links or calls this file.
 * Pinned bit for bit against vectors evaluated from the script's own text
 * (tests/test_cyl_reference_text.py).
 *
 * The script's kernels are ParallelStencil @parallel kernels written with
 * the FiniteDifferences3D macros (@all, @inn, @d_xa, @d_xi, @av_xa,
 * @av_xi, @av_xya, @av_xyi, ...): each statement assigns where its
 * left-hand array is defined (all points, or the inner points for @inn),
 * differences/averages over "a"ll points or the "i"nner points (index + 1)
 * of the other directions.  Below they are written out at 1-based
 * (ix, iy, iz); the averages sum their terms in the order
 * (ix,iy) (ix+1,iy) (ix,iy+1) (ix+1,iy+1) of the pair of directions.
 * ParallelStencil (a dependency, not part of the reference) defines those
 * macros; the vectors are evaluated with the same restatement of them
 * (tests/golden/jl_subset.py), so the term order inside a macro is pinned
 * only by that restatement ("parity unpinned" for it); the statements, their
 * association outside the macros, the guards and the loop are the script's.
 */
#include "sspt_cyl_oracle.h"
#include <math.h>
#include <stdlib.h>

static const double PI = 3.14159265358979323846;

static inline double jl_max(double x, double y)
{
    int pick_y = (y > x) || (signbit(y) < signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

void orc_cyl_dims(const orc_cyl *s, int f, int d[3])
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    int a = 0, b = 0, c = 0;
    switch (f) {
    case OCY_VR: case OCY_VR_OLD: a = 1; break;
    case OCY_VT: case OCY_VT_OLD: b = 1; break;
    case OCY_VZ: case OCY_VZ_OLD: c = 1; break;
    case OCY_dVR: a = -1; b = -2; c = -2; break;
    case OCY_dVT: a = -2; b = -1; c = -2; break;
    case OCY_dVZ: a = -2; b = -2; c = -1; break;
    case OCY_D_RT: case OCY_tRT: case OCY_ERR_RT: a = -1; b = -1; c = -2; break;
    case OCY_D_RZ: case OCY_tRZ: a = -1; b = -2; c = -1; break;
    case OCY_D_TZ: case OCY_tTZ: a = -2; b = -1; c = -1; break;
    default: break;
    }
    d[0] = nr + a; d[1] = nt + b; d[2] = nz + c;
}

/* 1-based access into field f */
#define D3(F) int F##_d[3]; orc_cyl_dims(s, F, F##_d)
#define AT(F, i, j, k) s->f[F][(((size_t)(k) - 1) * (size_t)F##_d[1] + ((size_t)(j) - 1)) * (size_t)F##_d[0] + ((size_t)(i) - 1)]
#define R_(i) s->R1[(i) - 1]

static size_t count(const orc_cyl *s, int f)
{
    int d[3];
    orc_cyl_dims(s, f, d);
    return (size_t)d[0] * d[1] * d[2];
}

void orc_cyl_equal(orc_cyl *s, int dst, int src)
{
    const size_t n = count(s, dst);
    for (size_t c = 0; c < n; ++c) s->f[dst][c] = s->f[src][c];
}

void orc_cyl_multiply(orc_cyl *s, int dst, int src, double fact)
{
    const size_t n = count(s, dst);
    for (size_t c = 0; c < n; ++c) s->f[dst][c] = s->f[src][c] * fact;
}

/* maxloc!(etaSM, eta) :125-128 */
void orc_cyl_maxloc(orc_cyl *s)
{
    D3(OCY_etaSM); D3(OCY_eta);
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix)
                AT(OCY_etaSM, ix + 1, iy + 1, iz + 1) =
                    jl_max(jl_max(jl_max(jl_max(AT(OCY_eta, ix, iy + 1, iz + 1), AT(OCY_eta, ix + 2, iy + 1, iz + 1)),
                                         AT(OCY_eta, ix + 1, iy + 1, iz + 1)),
                                  jl_max(AT(OCY_eta, ix + 1, iy, iz + 1), AT(OCY_eta, ix + 1, iy + 2, iz + 1))),
                           jl_max(AT(OCY_eta, ix + 1, iy + 1, iz), AT(OCY_eta, ix + 1, iy + 1, iz + 2)));
}

/* timesteps! :130-135 */
void orc_cyl_timesteps(orc_cyl *s, double dmp, double vpdt, double lr)
{
    const size_t n = count(s, OCY_KBDT);
    const double kb = (((dmp * 2.0) * PI) * vpdt) / lr, gs = ((4.0 * PI) * vpdt) / lr;
    for (size_t c = 0; c < n; ++c) {
        s->f[OCY_KBDT][c] = kb * s->f[OCY_etaSM][c];
        s->f[OCY_GSDT][c] = gs * s->f[OCY_etaSM][c];
    }
}

/* timesteps2! :137-141 */
void orc_cyl_timesteps2(orc_cyl *s, double vpdt)
{
    const size_t n = count(s, OCY_DT_R);
    const double v2 = vpdt * vpdt;
    for (size_t c = 0; c < n; ++c) {
        const double G = s->f[OCY_GSDT][c];
        s->f[OCY_DT_R][c] = v2 / (s->f[OCY_KBDT][c] + G / (1.0 + G / s->f[OCY_eta][c]));
    }
}

/* compute_P! :144-167 */
void orc_cyl_compute_P(orc_cyl *s, double dr, double dth, double dz)
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    D3(OCY_VR); D3(OCY_VT); D3(OCY_VZ); D3(OCY_P); D3(OCY_DIVV); D3(OCY_D_RR); D3(OCY_D_TT); D3(OCY_D_ZZ);
    D3(OCY_D_RT); D3(OCY_D_RZ); D3(OCY_D_TZ); D3(OCY_tRR); D3(OCY_tTT); D3(OCY_tZZ); D3(OCY_tRT); D3(OCY_tRZ);
    D3(OCY_tTZ); D3(OCY_KBDT); D3(OCY_GSDT); D3(OCY_eta);
    /* the kernel's statements in order; each at every point of its array */
    for (int iz = 1; iz <= nz; ++iz)
        for (int iy = 1; iy <= nt; ++iy)
            for (int ix = 1; ix <= nr; ++ix) {
                const double R = R_(ix);
                const double dvr = AT(OCY_VR, ix + 1, iy, iz) - AT(OCY_VR, ix, iy, iz);
                const double dvt = AT(OCY_VT, ix, iy + 1, iz) - AT(OCY_VT, ix, iy, iz);
                const double dvz = AT(OCY_VZ, ix, iy, iz + 1) - AT(OCY_VZ, ix, iy, iz);
                const double avr = (AT(OCY_VR, ix, iy, iz) + AT(OCY_VR, ix + 1, iy, iz)) * 0.5;
                const double div = ((dvr / dr + ((1.0 / R) * dvt) / dth) + avr / R) + dvz / dz;
                AT(OCY_DIVV, ix, iy, iz) = div;
                AT(OCY_D_RR, ix, iy, iz) = dvr / dr - (1.0 / 3.0) * div;
                AT(OCY_D_TT, ix, iy, iz) = (((1.0 / R) * dvt) / dth + avr / R) - (1.0 / 3.0) * div;
                AT(OCY_D_ZZ, ix, iy, iz) = dvz / dz - (1.0 / 3.0) * div;
            }
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rxy = (((R_(ix) + R_(ix + 1)) + R_(ix)) + R_(ix + 1)) * 0.25;          /* @av_xyi(R) */
                const double dxi = AT(OCY_VT, ix + 1, iy + 1, iz + 1) - AT(OCY_VT, ix, iy + 1, iz + 1); /* @d_xi(Vth) */
                const double dyi = AT(OCY_VR, ix + 1, iy + 1, iz + 1) - AT(OCY_VR, ix + 1, iy, iz + 1); /* @d_yi(VR) */
                const double axi = (AT(OCY_VT, ix, iy + 1, iz + 1) + AT(OCY_VT, ix + 1, iy + 1, iz + 1)) * 0.5;
                AT(OCY_D_RT, ix, iy, iz) = (dxi / dr + ((1.0 / Rxy) * dyi) / dth) - axi / Rxy;
            }
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double dxi = AT(OCY_VZ, ix + 1, iy + 1, iz + 1) - AT(OCY_VZ, ix, iy + 1, iz + 1);     /* @d_xi(VZ) */
                const double dzi = AT(OCY_VR, ix + 1, iy + 1, iz + 1) - AT(OCY_VR, ix + 1, iy + 1, iz);     /* @d_zi(VR) */
                AT(OCY_D_RZ, ix, iy, iz) = dxi / dr + dzi / dz;
            }
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Ryz = (((R_(ix + 1) + R_(ix + 1)) + R_(ix + 1)) + R_(ix + 1)) * 0.25;  /* @av_yzi(R) */
                const double dyi = AT(OCY_VZ, ix + 1, iy + 1, iz + 1) - AT(OCY_VZ, ix + 1, iy, iz + 1);     /* @d_yi(VZ) */
                const double dzi = AT(OCY_VT, ix + 1, iy + 1, iz + 1) - AT(OCY_VT, ix + 1, iy + 1, iz);     /* @d_zi(Vth) */
                AT(OCY_D_TZ, ix, iy, iz) = ((1.0 / Ryz) * dyi) / dth + dzi / dz;
            }
    for (int iz = 1; iz <= nz; ++iz)
        for (int iy = 1; iy <= nt; ++iy)
            for (int ix = 1; ix <= nr; ++ix) {
                const double K = AT(OCY_KBDT, ix, iy, iz), G = AT(OCY_GSDT, ix, iy, iz), e = AT(OCY_eta, ix, iy, iz);
                AT(OCY_P, ix, iy, iz) = AT(OCY_P, ix, iy, iz) - K * AT(OCY_DIVV, ix, iy, iz);
                AT(OCY_tRR, ix, iy, iz) = (AT(OCY_tRR, ix, iy, iz) + (G * 2.0) * AT(OCY_D_RR, ix, iy, iz)) / (1.0 + G / e);
                AT(OCY_tTT, ix, iy, iz) = (AT(OCY_tTT, ix, iy, iz) + (G * 2.0) * AT(OCY_D_TT, ix, iy, iz)) / (1.0 + G / e);
                AT(OCY_tZZ, ix, iy, iz) = (AT(OCY_tZZ, ix, iy, iz) + (G * 2.0) * AT(OCY_D_ZZ, ix, iy, iz)) / (1.0 + G / e);
            }
#define AV4(a0, a1, a2, a3) (((((a0) + (a1)) + (a2)) + (a3)) * 0.25)
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) { /* @av_xyi */
                const double G = AV4(AT(OCY_GSDT, ix, iy, iz + 1), AT(OCY_GSDT, ix + 1, iy, iz + 1),
                                     AT(OCY_GSDT, ix, iy + 1, iz + 1), AT(OCY_GSDT, ix + 1, iy + 1, iz + 1));
                const double e = AV4(AT(OCY_eta, ix, iy, iz + 1), AT(OCY_eta, ix + 1, iy, iz + 1),
                                     AT(OCY_eta, ix, iy + 1, iz + 1), AT(OCY_eta, ix + 1, iy + 1, iz + 1));
                AT(OCY_tRT, ix, iy, iz) = (AT(OCY_tRT, ix, iy, iz) + G * AT(OCY_D_RT, ix, iy, iz)) / (1.0 + G / e);
            }
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) { /* @av_xzi */
                const double G = AV4(AT(OCY_GSDT, ix, iy + 1, iz), AT(OCY_GSDT, ix + 1, iy + 1, iz),
                                     AT(OCY_GSDT, ix, iy + 1, iz + 1), AT(OCY_GSDT, ix + 1, iy + 1, iz + 1));
                const double e = AV4(AT(OCY_eta, ix, iy + 1, iz), AT(OCY_eta, ix + 1, iy + 1, iz),
                                     AT(OCY_eta, ix, iy + 1, iz + 1), AT(OCY_eta, ix + 1, iy + 1, iz + 1));
                AT(OCY_tRZ, ix, iy, iz) = (AT(OCY_tRZ, ix, iy, iz) + G * AT(OCY_D_RZ, ix, iy, iz)) / (1.0 + G / e);
            }
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) { /* @av_yzi */
                const double G = AV4(AT(OCY_GSDT, ix + 1, iy, iz), AT(OCY_GSDT, ix + 1, iy + 1, iz),
                                     AT(OCY_GSDT, ix + 1, iy, iz + 1), AT(OCY_GSDT, ix + 1, iy + 1, iz + 1));
                const double e = AV4(AT(OCY_eta, ix + 1, iy, iz), AT(OCY_eta, ix + 1, iy + 1, iz),
                                     AT(OCY_eta, ix + 1, iy, iz + 1), AT(OCY_eta, ix + 1, iy + 1, iz + 1));
                AT(OCY_tTZ, ix, iy, iz) = (AT(OCY_tTZ, ix, iy, iz) + G * AT(OCY_D_TZ, ix, iy, iz)) / (1.0 + G / e);
            }
}

/* compute_sigma! :169-175 */
void orc_cyl_compute_sigma(orc_cyl *s)
{
    const size_t n = count(s, OCY_P);
    for (size_t c = 0; c < n; ++c) {
        const double P = s->f[OCY_P][c];
        s->f[OCY_sRR][c] = -P + s->f[OCY_tRR][c];
        s->f[OCY_sTT][c] = -P + s->f[OCY_tTT][c];
        s->f[OCY_sZZ][c] = -P + s->f[OCY_tZZ][c];
    }
}

/* compute_TII! :177-182 */
void orc_cyl_compute_TII(orc_cyl *s)
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    D3(OCY_tII); D3(OCY_tRR); D3(OCY_tTT); D3(OCY_tZZ); D3(OCY_tRT); D3(OCY_tRZ); D3(OCY_tTZ);
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double a = AT(OCY_tRR, ix + 1, iy + 1, iz + 1), b = AT(OCY_tTT, ix + 1, iy + 1, iz + 1),
                             c = AT(OCY_tZZ, ix + 1, iy + 1, iz + 1);
                const double x = AV4(AT(OCY_tRT, ix, iy, iz), AT(OCY_tRT, ix + 1, iy, iz), AT(OCY_tRT, ix, iy + 1, iz),
                                     AT(OCY_tRT, ix + 1, iy + 1, iz));                         /* @av_xya */
                const double y = AV4(AT(OCY_tRZ, ix, iy, iz), AT(OCY_tRZ, ix + 1, iy, iz), AT(OCY_tRZ, ix, iy, iz + 1),
                                     AT(OCY_tRZ, ix + 1, iy, iz + 1));                         /* @av_xza */
                const double z = AV4(AT(OCY_tTZ, ix, iy, iz), AT(OCY_tTZ, ix, iy + 1, iz), AT(OCY_tTZ, ix, iy, iz + 1),
                                     AT(OCY_tTZ, ix, iy + 1, iz + 1));                         /* @av_yza */
                AT(OCY_tII, ix + 1, iy + 1, iz + 1) = sqrt((((1.0 / 2.0) * ((a * a + b * b) + c * c) + x * x) + y * y) + z * z);
            }
}

/* power_law! :184-191 */
void orc_cyl_power_law(orc_cyl *s, double tau_C, double n_exp, double relax)
{
    const size_t n = count(s, OCY_eta);
    const double ex = 1.0 - n_exp;
    for (size_t c = 0; c < n; ++c) {
        const double ini = s->f[OCY_eta_INI][c];
        double pl = ini * pow(s->f[OCY_tII][c] / tau_C, ex);
        pl = exp(log(pl) * relax + log(s->f[OCY_eta_PL_OLD][c]) * (1.0 - relax));
        s->f[OCY_eta_PL][c] = pl;
        s->f[OCY_eta][c] = 2.0 / (1.0 / ini + 1.0 / pl);
    }
}

/* compute_dV! :193-206 */
void orc_cyl_compute_dV(orc_cyl *s, double dr, double dth, double dz)
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    D3(OCY_dVR); D3(OCY_dVT); D3(OCY_dVZ); D3(OCY_sRR); D3(OCY_sTT); D3(OCY_sZZ); D3(OCY_tRT); D3(OCY_tRZ);
    D3(OCY_tTZ); D3(OCY_rhoG);
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix) {
                const double Rx = (R_(ix) + R_(ix + 1)) * 0.5;                               /* @av_xi(R) */
                const double t = ((((AT(OCY_sRR, ix + 1, iy + 1, iz + 1) - AT(OCY_sRR, ix, iy + 1, iz + 1)) / dr +
                                    ((1.0 / Rx) * (AT(OCY_tRT, ix, iy + 1, iz) - AT(OCY_tRT, ix, iy, iz))) / dth) +
                                   (AT(OCY_tRZ, ix, iy, iz + 1) - AT(OCY_tRZ, ix, iy, iz)) / dz) +
                                  ((AT(OCY_sRR, ix, iy + 1, iz + 1) + AT(OCY_sRR, ix + 1, iy + 1, iz + 1)) * 0.5) / Rx) -
                                 ((AT(OCY_sTT, ix, iy + 1, iz + 1) + AT(OCY_sTT, ix + 1, iy + 1, iz + 1)) * 0.5) / Rx;
                AT(OCY_dVR, ix, iy, iz) = t - (AT(OCY_rhoG, ix, iy + 1, iz + 1) + AT(OCY_rhoG, ix + 1, iy + 1, iz + 1)) * 0.5;
            }
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Ry = (R_(ix + 1) + R_(ix + 1)) * 0.5;                           /* @av_yi(R) */
                AT(OCY_dVT, ix, iy, iz) =
                    (((AT(OCY_tRT, ix + 1, iy, iz) - AT(OCY_tRT, ix, iy, iz)) / dr +
                      ((1.0 / Ry) * (AT(OCY_sTT, ix + 1, iy + 1, iz + 1) - AT(OCY_sTT, ix + 1, iy, iz + 1))) / dth) +
                     (AT(OCY_tTZ, ix, iy, iz + 1) - AT(OCY_tTZ, ix, iy, iz)) / dz) +
                    (2 * ((AT(OCY_tRT, ix, iy, iz) + AT(OCY_tRT, ix + 1, iy, iz)) * 0.5)) / Ry;
            }
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix) {
                const double Rz = (R_(ix + 1) + R_(ix + 1)) * 0.5;                           /* @av_zi(R) */
                AT(OCY_dVZ, ix, iy, iz) =
                    (((AT(OCY_tRZ, ix + 1, iy, iz) - AT(OCY_tRZ, ix, iy, iz)) / dr +
                      ((1.0 / Rz) * (AT(OCY_tTZ, ix, iy + 1, iz) - AT(OCY_tTZ, ix, iy, iz))) / dth) +
                     (AT(OCY_sZZ, ix + 1, iy + 1, iz + 1) - AT(OCY_sZZ, ix + 1, iy + 1, iz)) / dz) +
                    ((AT(OCY_tRZ, ix, iy, iz) + AT(OCY_tRZ, ix + 1, iy, iz)) * 0.5) / Rz;
            }
}

/* compute_V! :208-214 */
void orc_cyl_compute_V(orc_cyl *s)
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    D3(OCY_VR); D3(OCY_VT); D3(OCY_VZ); D3(OCY_dVR); D3(OCY_dVT); D3(OCY_dVZ); D3(OCY_DT_R);
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix)
                AT(OCY_VR, ix + 1, iy + 1, iz + 1) = AT(OCY_VR, ix + 1, iy + 1, iz + 1) +
                    AT(OCY_dVR, ix, iy, iz) * ((AT(OCY_DT_R, ix, iy + 1, iz + 1) + AT(OCY_DT_R, ix + 1, iy + 1, iz + 1)) * 0.5);
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix)
                AT(OCY_VT, ix + 1, iy + 1, iz + 1) = AT(OCY_VT, ix + 1, iy + 1, iz + 1) +
                    AT(OCY_dVT, ix, iy, iz) * ((AT(OCY_DT_R, ix + 1, iy, iz + 1) + AT(OCY_DT_R, ix + 1, iy + 1, iz + 1)) * 0.5);
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= nt - 2; ++iy)
            for (int ix = 1; ix <= nr - 2; ++ix)
                AT(OCY_VZ, ix + 1, iy + 1, iz + 1) = AT(OCY_VZ, ix + 1, iy + 1, iz + 1) +
                    AT(OCY_dVZ, ix, iy, iz) * ((AT(OCY_DT_R, ix + 1, iy + 1, iz) + AT(OCY_DT_R, ix + 1, iy + 1, iz + 1)) * 0.5);
}

/* boundary kernels :52-101 on field f; the coordinate arrays theta_VZ /
 * theta_VR of bc_th_fact! and bc_z_lin! vary along th only (th1) */
void orc_cyl_bc(orc_cyl *s, int f, int op, double fact, double r)
{
    int d[3];
    orc_cyl_dims(s, f, d);
    const int n1 = d[0], n2 = d[1], n3 = d[2];
    double *A = s->f[f];
#define A_(i, j, k) A[(((size_t)(k) - 1) * (size_t)n2 + ((size_t)(j) - 1)) * (size_t)n1 + ((size_t)(i) - 1)]
    for (int iz = 1; iz <= n3; ++iz)
        for (int iy = 1; iy <= n2; ++iy)
            for (int ix = 1; ix <= n1; ++ix) {
                switch (op) {
                case OCY_COPY_BC_R:
                    if (ix == 1) A_(ix, iy, iz) = A_(ix + 1, iy, iz);
                    if (ix == n1) A_(ix, iy, iz) = A_(ix - 1, iy, iz);
                    break;
                case OCY_BC_R_0:
                    if (ix == 1) A_(ix, iy, iz) = -A_(ix + 1, iy, iz);
                    if (ix == n1) A_(ix, iy, iz) = -A_(ix - 1, iy, iz);
                    break;
                case OCY_COPY_BC_TH:
                    if (iy == 1) A_(ix, iy, iz) = A_(ix, iy + 1, iz);
                    if (iy == n2) A_(ix, iy, iz) = A_(ix, iy - 1, iz);
                    break;
                case OCY_BC_TH_FACT:
                    if (iy == 1 || iy == n2) A_(ix, iy, iz) = (fact * s->th1[iy - 1]) * r;
                    break;
                case OCY_BC_TH_0:
                    if (iy == 1) A_(ix, iy, iz) = -A_(ix, iy + 1, iz);
                    if (iy == n2) A_(ix, iy, iz) = -A_(ix, iy - 1, iz);
                    break;
                case OCY_COPY_BC_Z:
                    if (iz == 1) A_(ix, iy, iz) = A_(ix, iy, iz + 1);
                    if (iz == n3) A_(ix, iy, iz) = A_(ix, iy, iz - 1);
                    break;
                case OCY_BC_Z_0:
                    if (iz == 1 || iz == n3) A_(ix, iy, iz) = 0.0;
                    break;
                case OCY_BC_Z_LIN:
                    if (iz == 1) A_(ix, iy, iz) = ((2 * s->th1[iy - 1]) * r) * fact - A_(ix, iy, iz + 1);
                    if (iz == n3) A_(ix, iy, iz) = ((2 * s->th1[iy - 1]) * r) * fact - A_(ix, iy, iz - 1);
                    break;
                default:
                    break;
                }
            }
#undef A_
}

/* err_rth! :217-220 */
void orc_cyl_err_rt(orc_cyl *s)
{
    const int nr = s->nr, nt = s->nth, nz = s->nz;
    D3(OCY_ERR_RT); D3(OCY_tRT); D3(OCY_D_RT); D3(OCY_eta);
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= nt - 1; ++iy)
            for (int ix = 1; ix <= nr - 1; ++ix)
                AT(OCY_ERR_RT, ix, iy, iz) = AT(OCY_tRT, ix, iy, iz) -
                    AT(OCY_D_RT, ix, iy, iz) * AV4(AT(OCY_eta, ix, iy, iz + 1), AT(OCY_eta, ix + 1, iy, iz + 1),
                                                   AT(OCY_eta, ix, iy + 1, iz + 1), AT(OCY_eta, ix + 1, iy + 1, iz + 1));
}

/* the exit test's error (host code :369-372):
 * max over components of maximum(abs.(V .- V_OLD)) ./ maximum(abs.(V)) */
double orc_cyl_exit_error(orc_cyl *s, double comp[3])
{
    static const int V[3] = {OCY_VR, OCY_VT, OCY_VZ}, O[3] = {OCY_VR_OLD, OCY_VT_OLD, OCY_VZ_OLD};
    double err = 0.0;
    int any_nan = 0;
    for (int a = 0; a < 3; ++a) {
        const size_t n = count(s, V[a]);
        double me = 0.0, mv = 0.0;
        int nan = 0;
        for (size_t c = 0; c < n; ++c) {
            const double d = fabs(s->f[V[a]][c] - s->f[O[a]][c]), v = fabs(s->f[V[a]][c]);
            if (isnan(d) || isnan(v)) nan = 1;
            if (d > me) me = d;
            if (v > mv) mv = v;
        }
        const double r = nan ? NAN : me / mv;
        if (comp) comp[a] = r;
        if (isnan(r)) any_nan = 1;
        else if (a == 0 || r > err) err = r;
    }
    return any_nan ? NAN : err;
}

/* one iteration of the loop body :331-366 with the exponent n_exp */
void orc_cyl_iteration(orc_cyl *s, const orc_cyl_params *p, double n_exp)
{
    orc_cyl_equal(s, OCY_etaSM, OCY_eta);
    orc_cyl_maxloc(s);
    orc_cyl_bc(s, OCY_etaSM, OCY_COPY_BC_R, 0.0, 0.0);
    orc_cyl_bc(s, OCY_etaSM, OCY_COPY_BC_TH, 0.0, 0.0);
    orc_cyl_bc(s, OCY_etaSM, OCY_COPY_BC_Z, 0.0, 0.0);
    orc_cyl_timesteps(s, p->dmp, p->vpdt, p->lr);
    orc_cyl_timesteps2(s, p->vpdt);
    orc_cyl_compute_P(s, p->dr, p->dth, p->dz);
    orc_cyl_compute_sigma(s);
    if (n_exp > 1.0) {
        orc_cyl_compute_TII(s);
        orc_cyl_bc(s, OCY_tII, OCY_COPY_BC_R, 0.0, 0.0);
        orc_cyl_bc(s, OCY_tII, OCY_COPY_BC_TH, 0.0, 0.0);
        orc_cyl_bc(s, OCY_tII, OCY_COPY_BC_Z, 0.0, 0.0);
        orc_cyl_equal(s, OCY_eta_PL_OLD, OCY_eta_PL);
        orc_cyl_power_law(s, p->tau_C, n_exp, p->relax);
    }
    orc_cyl_equal(s, OCY_VR_OLD, OCY_VR);
    orc_cyl_equal(s, OCY_VT_OLD, OCY_VT);
    orc_cyl_equal(s, OCY_VZ_OLD, OCY_VZ);
    orc_cyl_compute_dV(s, p->dr, p->dth, p->dz);
    orc_cyl_compute_V(s);
    orc_cyl_bc(s, OCY_VZ, OCY_COPY_BC_R, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VT, OCY_COPY_BC_R, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VR, OCY_BC_R_0, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VZ, OCY_BC_TH_FACT, p->sh, p->r);
    orc_cyl_bc(s, OCY_VR, OCY_BC_TH_FACT, 0.0, p->r);
    orc_cyl_bc(s, OCY_VT, OCY_BC_TH_0, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VT, OCY_BC_Z_0, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VR, OCY_BC_Z_0, 0.0, 0.0);
    orc_cyl_bc(s, OCY_VZ, OCY_BC_Z_LIN, p->sh, p->r);
}

/* the loop :329-384: the exponent switches to n_exp_PL once the last
 * exit-test error is below 1e-5 (err starts at 1.0, :325); err_rth! on
 * every check; iters = iter at the break */
int orc_cyl_pt_solve(orc_cyl *s, const orc_cyl_params *p, int maxiter, int nout, double eps, double *err_out,
                     int *converged)
{
    double err = 1.0, n_exp = p->n_exp;
    *converged = 0;
    *err_out = NAN;
    for (int iter = 1; iter <= maxiter; ++iter) {
        if (err < 1e-5) n_exp = p->n_exp_PL;
        orc_cyl_iteration(s, p, n_exp);
        if (iter % nout == 0) {
            err = orc_cyl_exit_error(s, NULL);
            orc_cyl_err_rt(s);
            *err_out = err;
            if (isnan(err)) return iter;
            if (err < eps && iter > 20) {
                *converged = 1;
                return iter;
            }
        }
    }
    return maxiter;
}
