/*
 * sspt_cart_oracle.c -- CPU restatement of the Cartesian twin of the diapir
 * solver, scripts_diapir/CartesianStokes_diapir.jl.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker of the Cartesian entry points
 * of libsspt.so (and their CPU baseline).  The product never links or calls
 * this file.  Pinned bit for bit against vectors evaluated from the
 * reference's own kernel text (tests/test_reference_text.py).
 *
 * Arrays in the reference's Julia layout (x fastest), 1-based helpers, each
 * statement with the reference's association (Julia's n-ary + and * are left
 * folds); -ffp-contract=off -fno-fast-math.
 */
#include "sspt_cart_oracle.h"
#include <math.h>
#include <stdlib.h>

#define IDX(n1, n2, i, j, k) (((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1))

static const double PI = 3.14159265358979323846; /* Float64(pi) */

static inline double jl_max(double x, double y)
{
    int pick_y = (y > x) || (signbit(y) < signbit(x));
    return pick_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

/* equal!(A, B) :41-44 */
void orc_cart_equal(double *A, const double *B, size_t n)
{
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) A[c] = B[c];
}

/* maxloc!(etaSM, eta) :125-128 (@inn = @maxloc, SphericalStokes_mxpu.jl:295-297 nesting) */
void orc_cart_maxloc(orc_cart *s)
{
    const int n1 = s->nx, n2 = s->ny, n3 = s->nz;
    const double *e = s->eta;
#pragma omp parallel for schedule(static)
    for (int iz = 2; iz <= n3 - 1; ++iz)
        for (int iy = 2; iy <= n2 - 1; ++iy)
            for (int ix = 2; ix <= n1 - 1; ++ix)
                s->etaSM[IDX(n1, n2, ix, iy, iz)] =
                    jl_max(jl_max(jl_max(jl_max(e[IDX(n1, n2, ix - 1, iy, iz)], e[IDX(n1, n2, ix + 1, iy, iz)]),
                                         e[IDX(n1, n2, ix, iy, iz)]),
                                  jl_max(e[IDX(n1, n2, ix, iy - 1, iz)], e[IDX(n1, n2, ix, iy + 1, iz)])),
                           jl_max(e[IDX(n1, n2, ix, iy, iz - 1)], e[IDX(n1, n2, ix, iy, iz + 1)]));
}

/* copy_boundary!(A) :97-105, A of size (nx, ny, nz) */
void orc_cart_copy_boundary(double *A, int n1, int n2, int n3)
{
    for (int iz = 1; iz <= n3 - 2; ++iz)
        for (int iy = 1; iy <= n2 - 2; ++iy) {
            A[IDX(n1, n2, 1, iy + 1, iz + 1)] = A[IDX(n1, n2, 2, iy + 1, iz + 1)];
            A[IDX(n1, n2, n1, iy + 1, iz + 1)] = A[IDX(n1, n2, n1 - 1, iy + 1, iz + 1)];
        }
    for (int iz = 1; iz <= n3 - 2; ++iz)
        for (int ix = 1; ix <= n1 - 2; ++ix) {
            A[IDX(n1, n2, ix + 1, 1, iz + 1)] = A[IDX(n1, n2, ix + 1, 2, iz + 1)];
            A[IDX(n1, n2, ix + 1, n2, iz + 1)] = A[IDX(n1, n2, ix + 1, n2 - 1, iz + 1)];
        }
    for (int iy = 1; iy <= n2 - 2; ++iy)
        for (int ix = 1; ix <= n1 - 2; ++ix) {
            A[IDX(n1, n2, ix + 1, iy + 1, 1)] = A[IDX(n1, n2, ix + 1, iy + 1, 2)];
            A[IDX(n1, n2, ix + 1, iy + 1, n3)] = A[IDX(n1, n2, ix + 1, iy + 1, n3 - 1)];
        }
}

/* timesteps! :130-139 */
void orc_cart_timesteps(orc_cart *s, double vpdt, double dmp, double lx)
{
    const size_t n = (size_t)s->nx * s->ny * s->nz;
    const double kb = (((dmp * 2.0) * PI) * vpdt) / lx, gs = ((4.0 * PI) * vpdt) / lx;
    const double vpdt2 = vpdt * vpdt;  /* vpdt^2, a literal power */
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        const double K = kb * s->etaSM[c], G = gs * s->etaSM[c];
        s->DT_R[c] = vpdt2 / (K + G / (1.0 + G / s->eta[c]));
    }
}

/* compute_P! :141-193 (macros :130-164) */
void orc_cart_compute_P(orc_cart *s, double dmp, double vpdt, double lx, double _dx, double _dy, double _dz)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double kb = (((dmp * 2.0) * PI) * vpdt) / lx, gs = ((4.0 * PI) * vpdt) / lx;
    const double *VX = s->VX, *VY = s->VY, *VZ = s->VZ, *eta = s->eta, *SM = s->etaSM;
#define C_(A, i, j, k) A[IDX(nx, ny, i, j, k)]
#define VX_(i, j, k) VX[IDX(nx + 1, ny, i, j, k)]
#define VY_(i, j, k) VY[IDX(nx, ny + 1, i, j, k)]
#define VZ_(i, j, k) VZ[IDX(nx, ny, i, j, k)]
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz; ++iz)
        for (int iy = 1; iy <= ny; ++iy)
            for (int ix = 1; ix <= nx; ++ix) {
                const double K = kb * C_(SM, ix, iy, iz), G = gs * C_(SM, ix, iy, iz);
                const double dxx = (VX_(ix + 1, iy, iz) - VX_(ix, iy, iz)) * _dx;
                const double dyy = (VY_(ix, iy + 1, iz) - VY_(ix, iy, iz)) * _dy;
                const double dzz = (VZ[IDX(nx, ny, ix, iy, iz + 1)] - VZ_(ix, iy, iz)) * _dz;
                const size_t c = IDX(nx, ny, ix, iy, iz);
                s->P[c] = s->P[c] - K * ((dxx + dyy) + dzz);
                s->tXX[c] = (s->tXX[c] + (G * 2.0) * dxx) / (1.0 + G / eta[c]);
                s->tYY[c] = (s->tYY[c] + (G * 2.0) * dyy) / (1.0 + G / eta[c]);
                s->tZZ[c] = (s->tZZ[c] + (G * 2.0) * dzz) / (1.0 + G / eta[c]);
            }
    /* tau_XY (nx-1, ny-1, nz-2) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= ny - 1; ++iy)
            for (int ix = 1; ix <= nx - 1; ++ix) {
                const double g4 = gs * (0.25 * (((C_(SM, ix, iy, iz + 1) + C_(SM, ix, iy + 1, iz + 1)) +
                                                 C_(SM, ix + 1, iy, iz + 1)) + C_(SM, ix + 1, iy + 1, iz + 1)));
                const double e4 = 0.25 * (((C_(eta, ix, iy, iz + 1) + C_(eta, ix, iy + 1, iz + 1)) +
                                           C_(eta, ix + 1, iy, iz + 1)) + C_(eta, ix + 1, iy + 1, iz + 1));
                const double d = (VX_(ix + 1, iy + 1, iz + 1) - VX_(ix + 1, iy, iz + 1)) * _dy +
                                 (VY_(ix + 1, iy + 1, iz + 1) - VY_(ix, iy + 1, iz + 1)) * _dx;
                double *t = &s->tXY[IDX(nx - 1, ny - 1, ix, iy, iz)];
                *t = (*t + g4 * d) / (1.0 + g4 / e4);
            }
    /* tau_XZ (nx-1, ny-2, nz-1) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= ny - 2; ++iy)
            for (int ix = 1; ix <= nx - 1; ++ix) {
                const double g4 = gs * (0.25 * (((C_(SM, ix, iy + 1, iz) + C_(SM, ix, iy + 1, iz + 1)) +
                                                 C_(SM, ix + 1, iy + 1, iz)) + C_(SM, ix + 1, iy + 1, iz + 1)));
                const double e4 = 0.25 * (((C_(eta, ix, iy + 1, iz) + C_(eta, ix, iy + 1, iz + 1)) +
                                           C_(eta, ix + 1, iy + 1, iz)) + C_(eta, ix + 1, iy + 1, iz + 1));
                const double d = (VX_(ix + 1, iy + 1, iz + 1) - VX_(ix + 1, iy + 1, iz)) * _dz +
                                 (VZ_(ix + 1, iy + 1, iz + 1) - VZ_(ix, iy + 1, iz + 1)) * _dx;
                double *t = &s->tXZ[IDX(nx - 1, ny - 2, ix, iy, iz)];
                *t = (*t + g4 * d) / (1.0 + g4 / e4);
            }
    /* tau_YZ (nx-2, ny-1, nz-1) */
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 1; ++iz)
        for (int iy = 1; iy <= ny - 1; ++iy)
            for (int ix = 1; ix <= nx - 2; ++ix) {
                const double g4 = gs * (0.25 * (((C_(SM, ix + 1, iy, iz) + C_(SM, ix + 1, iy, iz + 1)) +
                                                 C_(SM, ix + 1, iy + 1, iz)) + C_(SM, ix + 1, iy + 1, iz + 1)));
                const double e4 = 0.25 * (((C_(eta, ix + 1, iy, iz) + C_(eta, ix + 1, iy, iz + 1)) +
                                           C_(eta, ix + 1, iy + 1, iz)) + C_(eta, ix + 1, iy + 1, iz + 1));
                const double d = (VY_(ix + 1, iy + 1, iz + 1) - VY_(ix + 1, iy + 1, iz)) * _dz +
                                 (VZ_(ix + 1, iy + 1, iz + 1) - VZ_(ix + 1, iy, iz + 1)) * _dy;
                double *t = &s->tYZ[IDX(nx - 2, ny - 1, ix, iy, iz)];
                *t = (*t + g4 * d) / (1.0 + g4 / e4);
            }
}

/* compute_TII! :195-208 */
void orc_cart_compute_TII(orc_cart *s)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
#define XY_(i, j, k) s->tXY[IDX(nx - 1, ny - 1, i, j, k)]
#define XZ_(i, j, k) s->tXZ[IDX(nx - 1, ny - 2, i, j, k)]
#define YZ_(i, j, k) s->tYZ[IDX(nx - 2, ny - 1, i, j, k)]
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= ny - 2; ++iy)
            for (int ix = 1; ix <= nx - 2; ++ix) {
                const size_t c = IDX(nx, ny, ix + 1, iy + 1, iz + 1);
                const double a = s->tXX[c], b = s->tYY[c], d = s->tZZ[c];
                const double x = 0.25 * (((XY_(ix, iy, iz) + XY_(ix, iy + 1, iz)) + XY_(ix + 1, iy, iz)) +
                                         XY_(ix + 1, iy + 1, iz));
                const double y = 0.25 * (((XZ_(ix, iy, iz) + XZ_(ix, iy, iz + 1)) + XZ_(ix + 1, iy, iz)) +
                                         XZ_(ix + 1, iy, iz + 1));
                const double z = 0.25 * (((YZ_(ix, iy, iz) + YZ_(ix, iy, iz + 1)) + YZ_(ix, iy + 1, iz)) +
                                         YZ_(ix, iy + 1, iz + 1));
                s->tII[c] = sqrt((((1.0 / 2.0) * ((a * a + b * b) + d * d) + x * x) + y * y) + z * z);
            }
}

/* power_law!(eta_PL, eta, tII, eta_PL_OLD, eta_INI, tau_C, n_exp, relax) :210-221 */
void orc_cart_power_law(orc_cart *s, double tau_C, double n_exp, double relax)
{
    const size_t n = (size_t)s->nx * s->ny * s->nz;
    const double ex = 1.0 - n_exp;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        double pl = s->eta_ini[c] * pow(s->tII[c] / tau_C, ex);
        pl = exp(log(pl) * relax + log(s->eta_PL_old[c]) * (1.0 - relax));
        s->eta_PL[c] = pl;
        s->eta[c] = 2.0 / (1.0 / s->eta_ini[c] + 1.0 / pl);
    }
}

/* compute_V! :223-255 */
void orc_cart_compute_V(orc_cart *s, double _dx, double _dy, double _dz, double g)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double *P = s->P, *DT = s->DT_R, *rho = s->rho;
    double *VX = s->VX, *VY = s->VY, *VZ = s->VZ;
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= ny - 2; ++iy)
            for (int ix = 1; ix <= nx + 1 - 2; ++ix) {
                const double d = (((-(C_(P, ix + 1, iy + 1, iz + 1) - C_(P, ix, iy + 1, iz + 1))) * _dx +
                                   (C_(s->tXX, ix + 1, iy + 1, iz + 1) - C_(s->tXX, ix, iy + 1, iz + 1)) * _dx) +
                                  (XY_(ix, iy + 1, iz) - XY_(ix, iy, iz)) * _dy) +
                                 (XZ_(ix, iy, iz + 1) - XZ_(ix, iy, iz)) * _dz;
                double *v = &VX[IDX(nx + 1, ny, ix + 1, iy + 1, iz + 1)];
                *v = *v + d * (0.5 * (C_(DT, ix, iy + 1, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)));
            }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz - 2; ++iz)
        for (int iy = 1; iy <= ny + 1 - 2; ++iy)
            for (int ix = 1; ix <= nx - 2; ++ix) {
                const double d = (((XY_(ix + 1, iy, iz) - XY_(ix, iy, iz)) * _dx -
                                   (C_(P, ix + 1, iy + 1, iz + 1) - C_(P, ix + 1, iy, iz + 1)) * _dy) +
                                  (C_(s->tYY, ix + 1, iy + 1, iz + 1) - C_(s->tYY, ix + 1, iy, iz + 1)) * _dy) +
                                 (YZ_(ix, iy, iz + 1) - YZ_(ix, iy, iz)) * _dz;
                double *v = &VY[IDX(nx, ny + 1, ix + 1, iy + 1, iz + 1)];
                *v = *v + d * (0.5 * (C_(DT, ix + 1, iy, iz + 1) + C_(DT, ix + 1, iy + 1, iz + 1)));
            }
#pragma omp parallel for schedule(static)
    for (int iz = 1; iz <= nz + 1 - 2; ++iz)
        for (int iy = 1; iy <= ny - 2; ++iy)
            for (int ix = 1; ix <= nx - 2; ++ix) {
                const double d = ((((XZ_(ix + 1, iy, iz) - XZ_(ix, iy, iz)) * _dx +
                                    (YZ_(ix, iy + 1, iz) - YZ_(ix, iy, iz)) * _dy) -
                                   (C_(P, ix + 1, iy + 1, iz + 1) - C_(P, ix + 1, iy + 1, iz)) * _dz) +
                                  (C_(s->tZZ, ix + 1, iy + 1, iz + 1) - C_(s->tZZ, ix + 1, iy + 1, iz)) * _dz) -
                                 g * (0.5 * (C_(rho, ix + 1, iy + 1, iz) + C_(rho, ix + 1, iy + 1, iz + 1)));
                double *v = &VZ[IDX(nx, ny, ix + 1, iy + 1, iz + 1)];
                *v = *v + d * (0.5 * (C_(DT, ix + 1, iy + 1, iz) + C_(DT, ix + 1, iy + 1, iz + 1)));
            }
}

/* boundary kernels :52-95 on A (n1, n2, n3); y = the y coordinates Y_i */
void orc_cart_bc(double *A, int n1, int n2, int n3, int op, const double *y, double fact)
{
    switch (op) {
    case ORC_CART_BC_X_LIN:
        for (int iz = 1; iz <= n3; ++iz)
            for (int iy = 1; iy <= n2; ++iy) {
                A[IDX(n1, n2, 1, iy, iz)] = (2 * y[iy - 1]) * fact - A[IDX(n1, n2, 2, iy, iz)];
                A[IDX(n1, n2, n1, iy, iz)] = (2 * y[iy - 1]) * fact - A[IDX(n1, n2, n1 - 1, iy, iz)];
            }
        break;
    case ORC_CART_BC_X_0:
        for (int iz = 1; iz <= n3; ++iz)
            for (int iy = 1; iy <= n2; ++iy) {
                A[IDX(n1, n2, 1, iy, iz)] = 0.0;
                A[IDX(n1, n2, n1, iy, iz)] = 0.0;
            }
        break;
    case ORC_CART_COPY_BC_X:
        for (int iz = 1; iz <= n3; ++iz)
            for (int iy = 1; iy <= n2; ++iy) {
                A[IDX(n1, n2, 1, iy, iz)] = A[IDX(n1, n2, 2, iy, iz)];
                A[IDX(n1, n2, n1, iy, iz)] = A[IDX(n1, n2, n1 - 1, iy, iz)];
            }
        break;
    case ORC_CART_BC_Y_FACT:
        for (int iz = 1; iz <= n3; ++iz)
            for (int ix = 1; ix <= n1; ++ix) {
                A[IDX(n1, n2, ix, 1, iz)] = fact * y[0];
                A[IDX(n1, n2, ix, n2, iz)] = fact * y[n2 - 1];
            }
        break;
    case ORC_CART_BC_Y_0:
        for (int iz = 1; iz <= n3; ++iz)
            for (int ix = 1; ix <= n1; ++ix) {
                A[IDX(n1, n2, ix, 1, iz)] = -A[IDX(n1, n2, ix, 2, iz)];
                A[IDX(n1, n2, ix, n2, iz)] = -A[IDX(n1, n2, ix, n2 - 1, iz)];
            }
        break;
    case ORC_CART_BC_Z_0:
        for (int iy = 1; iy <= n2; ++iy)
            for (int ix = 1; ix <= n1; ++ix) {
                A[IDX(n1, n2, ix, iy, 1)] = -A[IDX(n1, n2, ix, iy, 2)];
                A[IDX(n1, n2, ix, iy, n3)] = -A[IDX(n1, n2, ix, iy, n3 - 1)];
            }
        break;
    case ORC_CART_COPY_BC_Z:
        for (int iy = 1; iy <= n2; ++iy)
            for (int ix = 1; ix <= n1; ++ix) {
                A[IDX(n1, n2, ix, iy, 1)] = A[IDX(n1, n2, ix, iy, 2)];
                A[IDX(n1, n2, ix, iy, n3)] = A[IDX(n1, n2, ix, iy, n3 - 1)];
            }
        break;
    default:
        break;
    }
}

/* error!(ERR_V, V) and maximum(ERR_V) ./ maximum(abs.(V)) :257-260, :382-388 */
double orc_cart_check_err(orc_cart *s, double comp[3])
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    double *E[3] = {s->ERR_VX, s->ERR_VY, s->ERR_VZ};
    const double *V[3] = {s->VX, s->VY, s->VZ};
    const size_t n[3] = {(size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1)};
    double err = 0.0;
    int any_nan = 0;
    for (int a = 0; a < 3; ++a) {
        double me = 0.0, mv = 0.0;
        int nan_e = 0, nan_v = 0;
        for (size_t c = 0; c < n[a]; ++c) {
            const double d = fabs(E[a][c] - V[a][c]);
            E[a][c] = d;
            const double av = fabs(V[a][c]);
            if (isnan(d)) nan_e = 1; else if (d > me) me = d;
            if (isnan(av)) nan_v = 1; else if (av > mv) mv = av;
        }
        const double r = (nan_e || nan_v) ? NAN : me / mv;
        if (comp) comp[a] = r;
        if (isnan(r)) any_nan = 1;
        else if (a == 0 || r > err) err = r;
    }
    return any_nan ? NAN : err;
}

/* one iteration of the reference loop :351-379 (check: copy V into ERR_V
 * before compute_V!, as on every nout-th iteration) */
void orc_cart_iteration(orc_cart *s, const orc_cart_params *p, int check)
{
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const size_t nc = (size_t)nx * ny * nz;
    orc_cart_equal(s->etaSM, s->eta, nc);
    orc_cart_maxloc(s);
    orc_cart_copy_boundary(s->etaSM, nx, ny, nz);
    orc_cart_timesteps(s, p->vpdt, p->dmp, p->lx);
    orc_cart_compute_P(s, p->dmp, p->vpdt, p->lx, p->_dx, p->_dy, p->_dz);
    if (p->n_exp > 1.0) {
        orc_cart_compute_TII(s);
        orc_cart_copy_boundary(s->tII, nx, ny, nz);
        orc_cart_equal(s->eta_PL_old, s->eta_PL, nc);
        orc_cart_power_law(s, p->tau_C, p->n_exp, p->relax);
    }
    if (check) {
        orc_cart_equal(s->ERR_VX, s->VX, (size_t)(nx + 1) * ny * nz);
        orc_cart_equal(s->ERR_VY, s->VY, (size_t)nx * (ny + 1) * nz);
        orc_cart_equal(s->ERR_VZ, s->VZ, (size_t)nx * ny * (nz + 1));
    }
    orc_cart_compute_V(s, p->_dx, p->_dy, p->_dz, p->g);
    orc_cart_bc(s->VX, nx + 1, ny, nz, ORC_CART_BC_X_LIN, s->y1, p->sh);
    orc_cart_bc(s->VY, nx, ny + 1, nz, ORC_CART_BC_X_0, s->y1, 0.0);
    orc_cart_bc(s->VZ, nx, ny, nz + 1, ORC_CART_BC_X_0, s->y1, 0.0);
    orc_cart_bc(s->VX, nx + 1, ny, nz, ORC_CART_BC_Y_FACT, s->y1, p->sh);
    orc_cart_bc(s->VY, nx, ny + 1, nz, ORC_CART_BC_Y_0, s->y1, 0.0);
    orc_cart_bc(s->VZ, nx, ny, nz + 1, ORC_CART_BC_Y_FACT, s->y1, 0.0);
    orc_cart_bc(s->VX, nx + 1, ny, nz, ORC_CART_COPY_BC_Z, s->y1, 0.0);
    orc_cart_bc(s->VY, nx, ny + 1, nz, ORC_CART_COPY_BC_Z, s->y1, 0.0);
    orc_cart_bc(s->VZ, nx, ny, nz + 1, ORC_CART_BC_Z_0, s->y1, 0.0);
}

/* the reference loop with its exit test :349-391; returns iters (iter at
 * the break, or maxiter), *err the last error, *converged */
int orc_cart_pt_solve(orc_cart *s, const orc_cart_params *p, int maxiter, int nout, double eps, double *err,
                      int *converged)
{
    *err = NAN;
    *converged = 0;
    for (int iter = 1; iter <= maxiter; ++iter) {
        const int check = iter % nout == 0;
        orc_cart_iteration(s, p, check);
        if (check) {
            *err = orc_cart_check_err(s, NULL);
            if (isnan(*err)) return iter;
            if (*err < eps && iter > 20) {
                *converged = 1;
                return iter;
            }
        }
    }
    return maxiter;
}
