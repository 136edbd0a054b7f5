/*
 * CPU oracle for the FDFK2D time-stepping path -- TEST INFRASTRUCTURE ONLY.
 * See fdfk_oracle.h for what is a literal restatement of reference source
 * (pinned bit-for-bit against oracle/_ref) and what is this project's own
 * formulation of the binary-only parts (pinned by invariants in tests/).
 *
 * Floating-point semantics: Fortran evaluates `mysum + coef(k,n)*(a-b)` with
 * mysum, a, b REAL(4) and coef REAL(8) as: (a-b) in FP32, product and sum in
 * FP64, result rounded to FP32 on assignment.  The helpers below reproduce
 * that exactly; this file must be compiled without FP contraction
 * (-ffp-contract=off) so no FMA changes the roundings.
 */
#include "fdfk_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define CO(coef, j, i) (coef)[((j) - 1) + norder * ((i) - 1)]
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
/* Fortran REAL(4) accumulate with a REAL(8) coefficient */
static inline orc_real acc(orc_real s, double c, orc_real d) { return (orc_real)((double)s + c * (double)d); }

/* ------------------------------------------------------------------ */
/* FD coefficients, src/Submain.f90:1111-1243                          */
/* ------------------------------------------------------------------ */
void orc_fd_coef0(int norder, double *coef)
{
    memset(coef, 0, sizeof(double) * norder * norder);
    for (int i = 1; i <= norder; ++i)
        for (int j = 1; j <= i; ++j) {
            double t1 = 1.0, t2 = 1.0, t3 = 1.0;
            for (int k = 1; k <= j - 1; ++k) {
                t1 = t1 * (double)((2 * k - 1) * (2 * k - 1));
                t2 = t2 * ((double)((2 * j - 1) * (2 * j - 1)) - (double)((2 * k - 1) * (2 * k - 1)));
            }
            for (int k = j + 1; k <= i; ++k) {
                t1 = t1 * (double)((2 * k - 1) * (2 * k - 1));
                t3 = t3 * ((double)((2 * k - 1) * (2 * k - 1)) - (double)((2 * j - 1) * (2 * j - 1)));
            }
            t1 = t1 * (double)((j + 1) % 2 == 0 ? 1 : -1);
            t2 = t2 * (double)(2 * j - 1);
            CO(coef, j, i) = t1 / t2 / t3;
        }
}

void orc_fd_coef1(int norder, double *coef)
{
    memset(coef, 0, sizeof(double) * norder * norder);
    for (int i = 1; i <= norder; ++i)
        for (int j = 1; j <= i; ++j) {
            double t1 = 1.0, t2 = 1.0, t3 = 1.0;
            for (int k = 1; k <= j - 1; ++k) {
                t1 = t1 * (double)(k * k);
                t2 = t2 * ((double)(j * j) - (double)(k * k));
            }
            for (int k = j + 1; k <= i; ++k) {
                t1 = t1 * (double)(k * k);
                t3 = t3 * ((double)(k * k) - (double)(j * j));
            }
            t1 = t1 * (double)((j + 1) % 2 == 0 ? 1 : -1);
            t2 = t2 * (2.0 * (double)j);
            CO(coef, j, i) = t1 / t2 / t3;
        }
}

void orc_fd_coef2(int norder, double *coef)
{
    memset(coef, 0, sizeof(double) * norder * norder);
    for (int i = 1; i <= norder; ++i)
        for (int j = 1; j <= i; ++j) {
            double t1 = 1.0, t2 = 1.0, t3 = 1.0;
            for (int k = 1; k <= j - 1; ++k) {
                t1 = t1 * (double)(k * k);
                t2 = t2 * ((double)(j * j) - (double)(k * k));
            }
            for (int k = j + 1; k <= i; ++k) {
                t1 = t1 * (double)(k * k);
                t3 = t3 * ((double)(k * k) - (double)(j * j));
            }
            t1 = t1 * (double)((j + 1) % 2 == 0 ? 1 : -1);
            t2 = t2 * (double)(j * j);
            CO(coef, j, i) = t1 / t2 / t3;
        }
}

/* src/Submain.f90:118-164 */
int orc_check_stability(int nx, int nz, int npml, int izt, int norder, orc_real dx, orc_real dz,
                        double dt, const double *coef2, const orc_real *c, orc_real *dt_stable)
{
    orc_real h = dx > dz ? dx : dz;
    orc_real cx = (h / dx) * (h / dx);
    orc_real cz = (h / dz) * (h / dz);
    double sa = 0.0, s = 0.0;
    for (int j = 1; j <= norder; ++j) sa += fabs(CO(coef2, j, norder));
    for (int j = 1; j <= norder; ++j) s += CO(coef2, j, norder);
    orc_real cfl = (orc_real)sqrt(2.0 / ((double)(cx + cz) * (sa + s)));
    long ncell = (long)(nz + npml - izt + 1) * (long)(nx + 2 * npml);
    orc_real vpmax = -INFINITY;
    for (long i = 0; i < ncell; ++i) {
        orc_real v = sqrtf((c[3 * i + 1] + 2.0f * c[3 * i + 2]) / c[3 * i + 0]);
        if (v > vpmax) vpmax = v;
    }
    orc_real dts = cfl * h / vpmax;
    *dt_stable = dts;
    if (dt > (double)dts) return 1;
    if (dt < (double)(0.4f * dts)) return 2;
    return 0;
}

/* 2-D array helpers: a(izt:nz+npml, 1-npml:nx+npml), z fastest */
#define DIMS int NZ = nz + npml - izt + 1; (void)NZ
#define A2(a, iz, ix) (a)[(size_t)((iz) - izt) + (size_t)NZ * (size_t)((ix) + npml - 1)]

/* src/Submain.f90:167-212 */
orc_real orc_diff1(int norder, int dire, int nx, int nz, int npml, int izt, int ix, int iz,
                orc_real d, const double *coef, const orc_real *u)
{
    DIMS;
    int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
    int noz = imin(imin(iz - izt, nz + npml - iz), norder);
    int kx = 2 - dire, kz = dire - 1;
    int node = kx * nox + kz * noz;
    orc_real s = 0.0f;
    for (int k = 1; k <= node; ++k) {
        int lx2 = ix + kx * k, lx1 = ix - kx * k, lz2 = iz + kz * k, lz1 = iz - kz * k;
        s = acc(s, CO(coef, k, node), A2(u, lz2, lx2) - A2(u, lz1, lx1));
    }
    return s / d;
}

/* src/Submain.f90:215-249; u(1-npml:nx+npml) */
orc_real orc_diffx1(int norder, int nx, int npml, int ix, orc_real dx, const double *coef, const orc_real *u)
{
    int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= nox; ++k)
        s = acc(s, CO(coef, k, nox), u[ix + k + npml - 1] - u[ix - k + npml - 1]);
    return s / dx;
}

/* ------------------------------------------------------------------ */
/* mixed operators, src/Submain.f90:251-762                            */
/* ------------------------------------------------------------------ */
void orc_diffzx_reg(int norder, int nx, int nz, int npml, int izt, orc_real dx, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    size_t n = (size_t)NZ * (size_t)(nx + 2 * npml);
    orc_real *v = (orc_real *)calloc(n, sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int iz = 2; iz <= nz; ++iz) {
        int noz = imin(imin(iz - izt, nz + npml - iz), norder);
        for (int ix = 1 - norder; ix <= nx + norder; ++ix) {
            orc_real s = 0.0f;
            for (int k = 1; k <= noz; ++k) s = acc(s, CO(coef, k, noz), A2(u, iz + k, ix) - A2(u, iz - k, ix));
            A2(v, iz, ix) = A2(mat, iz, ix) * s / dz;
        }
    }
    memset(out, 0, n * sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int ix = 1; ix <= nx; ++ix) {
        int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
        for (int iz = 2; iz <= nz; ++iz) {
            orc_real s = 0.0f;
            for (int k = 1; k <= nox; ++k) s = acc(s, CO(coef, k, nox), A2(v, iz, ix + k) - A2(v, iz, ix - k));
            A2(out, iz, ix) = s / dx;
        }
    }
    free(v);
}

void orc_diffzx_pml(int norder, int nx, int nz, int npml, int izt, orc_real dx, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    size_t n = (size_t)NZ * (size_t)(nx + 2 * npml);
    orc_real *v = (orc_real *)calloc(n, sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int iz = izt + 1; iz <= nz + npml - 1; ++iz) {
        int noz = imin(imin(iz - izt, nz + npml - iz), norder);
        for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix) {
            /* mask: 1 except (1:nz, norder+1:nx-norder) and row izt */
            if (iz >= 1 && iz <= nz && ix >= norder + 1 && ix <= nx - norder) continue;
            orc_real s = 0.0f;
            for (int k = 1; k <= noz; ++k) s = acc(s, CO(coef, k, noz), A2(u, iz + k, ix) - A2(u, iz - k, ix));
            A2(v, iz, ix) = A2(mat, iz, ix) * s / dz;
        }
    }
    memset(out, 0, n * sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix) {
        int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
        for (int iz = izt + 1; iz <= nz + npml - 1; ++iz) {
            if (iz >= 1 && iz <= nz && ix >= 1 && ix <= nx) continue;
            orc_real s = 0.0f;
            for (int k = 1; k <= nox; ++k) s = acc(s, CO(coef, k, nox), A2(v, iz, ix + k) - A2(v, iz, ix - k));
            A2(out, iz, ix) = s / dx;
        }
    }
    free(v);
}

void orc_diffxz_reg(int norder, int nx, int nz, int npml, int izt, orc_real dx, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    size_t n = (size_t)NZ * (size_t)(nx + 2 * npml);
    orc_real *v = (orc_real *)calloc(n, sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int ix = 1; ix <= nx; ++ix) {
        int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
        for (int iz = 1; iz <= nz + norder; ++iz) {
            orc_real s = 0.0f;
            for (int k = 1; k <= nox; ++k) s = acc(s, CO(coef, k, nox), A2(u, iz, ix + k) - A2(u, iz, ix - k));
            A2(v, iz, ix) = A2(mat, iz, ix) * s / dx;
        }
    }
    memset(out, 0, n * sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int iz = 2; iz <= nz; ++iz) {
        int noz = imin(imin(iz - izt, nz + npml - iz), norder);
        for (int ix = 1; ix <= nx; ++ix) {
            orc_real s = 0.0f;
            for (int k = 1; k <= noz; ++k) s = acc(s, CO(coef, k, noz), A2(v, iz + k, ix) - A2(v, iz - k, ix));
            A2(out, iz, ix) = s / dz;
        }
    }
    free(v);
}

void orc_diffxz_pml(int norder, int nx, int nz, int npml, int izt, orc_real dx, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    size_t n = (size_t)NZ * (size_t)(nx + 2 * npml);
    orc_real *v = (orc_real *)calloc(n, sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix) {
        int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
        for (int iz = izt; iz <= nz + npml - 1; ++iz) {
            if (iz >= 1 && iz <= nz - norder && ix >= 1 && ix <= nx) continue;
            orc_real s = 0.0f;
            for (int k = 1; k <= nox; ++k) s = acc(s, CO(coef, k, nox), A2(u, iz, ix + k) - A2(u, iz, ix - k));
            A2(v, iz, ix) = A2(mat, iz, ix) * s / dx;
        }
    }
    memset(out, 0, n * sizeof(orc_real));
#pragma omp parallel for schedule(dynamic)
    for (int iz = izt + 1; iz <= nz + npml - 1; ++iz) {
        int noz = imin(imin(iz - izt, nz + npml - iz), norder);
        for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix) {
            if (iz >= 1 && iz <= nz && ix >= 1 && ix <= nx) continue;
            orc_real s = 0.0f;
            for (int k = 1; k <= noz; ++k) s = acc(s, CO(coef, k, noz), A2(v, iz + k, ix) - A2(v, iz - k, ix));
            A2(out, iz, ix) = s / dz;
        }
    }
    free(v);
}

/* ------------------------------------------------------------------ */
/* second-derivative operators, src/Submain.f90:764-1109               */
/* ------------------------------------------------------------------ */
static inline orc_real hmean(orc_real a, orc_real b) { return 2.0f / (1.0f / a + 1.0f / b); }

/* one output of diffx2 at (iz, ix): the body of the loops at :815-837 */
static inline orc_real x2_point(int norder, int nx, int npml, int izt, int NZ, const double *coef,
                             const orc_real *mat, const orc_real *u, int iz, int ix)
{
    int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= nox; ++k) {
        int i = (2 * k - 1) / 2;
        int n1 = imin(imin(ix + i + npml - 1, nx + npml - ix - i), norder);
        orc_real s1 = 0.0f;
        for (int j = 1; j <= n1; ++j) s1 = acc(s1, CO(coef, j, n1), A2(u, iz, ix + j + k - 1) - A2(u, iz, ix - j + k));
        int n2 = imin(imin(ix - i + npml - 1, nx + npml - ix + i), norder);
        orc_real s2 = 0.0f;
        for (int j = 1; j <= n2; ++j) s2 = acc(s2, CO(coef, j, n2), A2(u, iz, ix + j - k) - A2(u, iz, ix - j - k + 1));
        orc_real t = hmean(A2(mat, iz, ix + k), A2(mat, iz, ix + k - 1)) * s1 -
                  hmean(A2(mat, iz, ix - k + 1), A2(mat, iz, ix - k)) * s2;
        s = acc(s, CO(coef, k, nox), t);
    }
    return s;
}

static inline orc_real z2_point(int norder, int nz, int npml, int izt, int NZ, const double *coef,
                             const orc_real *mat, const orc_real *u, int iz, int ix)
{
    int noz = imin(imin(iz - izt, nz + npml - iz), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= noz; ++k) {
        int i = (2 * k - 1) / 2;
        int n1 = imin(imin(iz + i - izt, nz + npml - iz - i), norder);
        orc_real s1 = 0.0f;
        for (int j = 1; j <= n1; ++j) s1 = acc(s1, CO(coef, j, n1), A2(u, iz + j + k - 1, ix) - A2(u, iz - j + k, ix));
        int n2 = imin(imin(iz - i - izt, nz + npml - iz + i), norder);
        orc_real s2 = 0.0f;
        for (int j = 1; j <= n2; ++j) s2 = acc(s2, CO(coef, j, n2), A2(u, iz + j - k, ix) - A2(u, iz - j - k + 1, ix));
        orc_real t = hmean(A2(mat, iz + k, ix), A2(mat, iz + k - 1, ix)) * s1 -
                  hmean(A2(mat, iz - k + 1, ix), A2(mat, iz - k, ix)) * s2;
        s = acc(s, CO(coef, k, noz), t);
    }
    return s;
}

void orc_diffx2_reg(int norder, int nx, int nz, int npml, int izt, orc_real dx,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    memset(out, 0, (size_t)NZ * (size_t)(nx + 2 * npml) * sizeof(orc_real));
    orc_real dx2 = (orc_real)(1.0 / (double)(dx * dx));
#pragma omp parallel for schedule(dynamic)
    for (int ix = 1; ix <= nx; ++ix)
        for (int iz = izt; iz <= nz; ++iz) {
            if (iz < 1) continue; /* mask(1:nz,1:nx) */
            A2(out, iz, ix) = x2_point(norder, nx, npml, izt, NZ, coef, mat, u, iz, ix) * dx2;
        }
}

void orc_diffx2_pml(int norder, int nx, int nz, int npml, int izt, orc_real dx,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    memset(out, 0, (size_t)NZ * (size_t)(nx + 2 * npml) * sizeof(orc_real));
    orc_real dx2 = 1.0f / (dx * dx);
#pragma omp parallel for schedule(dynamic)
    for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix)
        for (int iz = izt; iz <= nz + npml - 1; ++iz) {
            if (iz >= 1 && iz <= nz && ix >= 1 && ix <= nx) continue;
            A2(out, iz, ix) = x2_point(norder, nx, npml, izt, NZ, coef, mat, u, iz, ix) * dx2;
        }
}

void orc_diffz2_reg(int norder, int nx, int nz, int npml, int izt, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    memset(out, 0, (size_t)NZ * (size_t)(nx + 2 * npml) * sizeof(orc_real));
    orc_real dz2 = 1.0f / (dz * dz);
#pragma omp parallel for schedule(dynamic)
    for (int iz = 2; iz <= nz; ++iz)
        for (int ix = 1; ix <= nx; ++ix)
            A2(out, iz, ix) = z2_point(norder, nz, npml, izt, NZ, coef, mat, u, iz, ix) * dz2;
}

void orc_diffz2_pml(int norder, int nx, int nz, int npml, int izt, orc_real dz,
                    const double *coef, const orc_real *mat, const orc_real *u, orc_real *out)
{
    DIMS;
    memset(out, 0, (size_t)NZ * (size_t)(nx + 2 * npml) * sizeof(orc_real));
    orc_real dz2 = 1.0f / (dz * dz);
#pragma omp parallel for schedule(dynamic)
    for (int iz = izt + 1; iz <= nz + npml - 1; ++iz)
        for (int ix = 1 - npml; ix <= nx + npml; ++ix) {
            if (iz >= 1 && iz <= nz && ix >= 1 && ix <= nx) continue;
            A2(out, iz, ix) = z2_point(norder, nz, npml, izt, NZ, coef, mat, u, iz, ix) * dz2;
        }
}

/* ------------------------------------------------------------------ */
/* receiver interpolation, src/Submain.f90:3-116                       */
/* ------------------------------------------------------------------ */
static double dirac_w(double x)
{
    const double C1_0 = 1.0, C1_2 = -5.0 / 4.0, C1_3 = -35.0 / 12.0, C1_4 = 21.0 / 4.0, C1_5 = -25.0 / 12.0;
    const double C2_0 = -4.0, C2_1 = 75.0 / 4.0, C2_2 = -245.0 / 8.0, C2_3 = 545.0 / 24.0,
                 C2_4 = -63.0 / 8.0, C2_5 = 25.0 / 24.0;
    const double C3_0 = 18.0, C3_1 = -153.0 / 4.0, C3_2 = 255.0 / 8.0, C3_3 = -313.0 / 24.0,
                 C3_4 = 21.0 / 8.0, C3_5 = -5.0 / 24.0;
    double x2 = x * x, x3 = x2 * x;
    if (x <= 1.0) return C1_0 + x2 * (C1_2 + x2 * C1_4) + x3 * (C1_3 + x2 * C1_5);
    if (x <= 2.0) return C2_0 + C2_1 * x + x2 * (C2_2 + x2 * C2_4) + x3 * (C2_3 + x2 * C2_5);
    if (x <= 3.0) return C3_0 + C3_1 * x + x2 * (C3_2 + x2 * C3_4) + x3 * (C3_3 + x2 * C3_5);
    return 0.0;
}

void orc_loadseis(int nx, int nz, int npml, int izt, orc_real x0, orc_real z0, orc_real dx, orc_real dz,
                  orc_real rx, orc_real rz, const orc_real *ux, const orc_real *uz, orc_real *dsx, orc_real *dsz)
{
    int NZ = nz + npml - izt + 1;
    int irx = (int)floorf((rx - x0) / dx) + 1;
    int irz = (int)floorf((rz - z0) / dz) + 1;
    orc_real sx = 0.0f, sz = 0.0f;
    for (int ix = imax(1 - npml, irx - 3); ix <= imin(nx + npml, irx + 3); ++ix) {
        double x = fabs((double)(x0 + (orc_real)(ix - 1) * dx) - (double)rx) / (double)dx;
        double bax = dirac_w(x);
        for (int iz = imax(izt, irz - 3); iz <= imin(nz + npml, irz + 3); ++iz) {
            double z = fabs((double)(z0 + (orc_real)(iz - 1) * dz) - (double)rz) / (double)dz;
            double baz = dirac_w(z);
            double ba = bax * baz;
            size_t off = 2 * ((size_t)(iz - izt) + (size_t)NZ * (size_t)(ix + npml - 1));
            sx = (orc_real)((double)sx + ba * (double)ux[off]);
            sz = (orc_real)((double)sz + ba * (double)uz[off]);
        }
    }
    *dsx = sx;
    *dsz = sz;
}

/* src/Input.f90:364-378 */
void orc_material_law(long ncell, orc_real *c)
{
    for (long i = 0; i < ncell; ++i) {
        orc_real vp = c[3 * i + 1], vs = c[3 * i + 2];
        orc_real rho = (vp + 980.0f) / 2.760f;
        c[3 * i + 0] = rho;
        c[3 * i + 1] = rho * (vp * vp - 2.0f * vs * vs);
        c[3 * i + 2] = rho * vs * vs;
    }
}

/* ------------------------------------------------------------------ */
/* geometry: src/calc_backazimuth.f90, coord_transform.f90,            */
/* define_rotation_matrix.f90                                          */
/* ------------------------------------------------------------------ */
static const double O_PI = 3.14159265358979323846;
void orc_backaz(double evla, double evlo, double stla, double stlo, double *baz)
{
    const double deg2rad = O_PI / 180.0, rad2deg = 180.0 / O_PI, twoPI = 2.0 * O_PI, PIo2 = 0.5 * O_PI;
    double ell = 1.0 / 298.257;
    double coeff = (1.0 - ell) * (1.0 - ell);
    double scola = PIo2 - atan(coeff * tan(stla * deg2rad));
    double ecola = PIo2 - atan(coeff * tan(evla * deg2rad));
    double slon = stlo * deg2rad, elon = evlo * deg2rad;
    double a = sin(scola) * cos(slon), b = sin(scola) * sin(slon), c = cos(scola);
    double d = sin(slon), e = -cos(slon), g = -c * e, h = c * d, k = -sin(scola);
    double aa = sin(ecola) * cos(elon), bb = sin(ecola) * sin(elon), cc = cos(ecola);
    (void)aa; (void)bb; (void)cc; (void)k;
    double rhs1 = (aa - d) * (aa - d) + (bb - e) * (bb - e) + cc * cc - 2.0;
    double rhs2 = (aa - g) * (aa - g) + (bb - h) * (bb - h) + (cc - k) * (cc - k) - 2.0;
    double dbaz = atan2(rhs1, rhs2);
    if (dbaz < 0.0) dbaz = dbaz + twoPI;
    double bz = dbaz * rad2deg;
    if (fabs(bz - 360.0) < 1.0e-5) bz = 0.0;
    *baz = bz;
    (void)a; (void)b;
}

void orc_backazimuth_grids(double lat_org, double lon_org, double az_org, int n,
                           const orc_real *coordx, double evla, double evlo, double *bazs)
{
    const double deg2rad = O_PI / 180.0, rad2deg = 180.0 / O_PI;
    double az = az_org * deg2rad, lat = lat_org * deg2rad, lon = lon_org * deg2rad;
    double st = sin(lat), ct = cos(lat), sp = sin(lon), cp = cos(lon), sa = sin(az), ca = cos(az);
    double f1 = st * cp;
    /* first column of inv_rotate_matrix (define_rotation_matrix.f90:79-87) */
    double i11 = -sa * sp + ca * f1;
    double i21 = sa * cp + ca * f1;
    double i31 = -ca * ct;
    /* geogr2cart(lat_org, lon_org) with Earth_raduis = 6371.e3 (REAL(4)) */
    double R = (double)6371.e3f;
    double colat = (90.0 - lat_org) * deg2rad, lo = lon_org * deg2rad;
    double xo = R * sin(colat) * cos(lo), yo = R * sin(colat) * sin(lo), zo = R * cos(colat);
    for (int i = 0; i < n; ++i) {
        double cx = (double)coordx[i];
        double x = i11 * cx + xo, y = i21 * cx + yo, z = i31 * cx + zo;
        double r = hypot(x, y), stla, stlo;
        if (r < 1.0e-6) {
            stla = z > 0.0 ? 90.0 : -90.0;
            stlo = 0.0;
        } else {
            stla = atan2(z, r) * rad2deg;
            stlo = atan2(y, x) * rad2deg;
            if (fabs(stla) < (double)1.e-9f) stla = 0.0;
            if (fabs(stlo) < (double)1.e-9f) stlo = 0.0;
        }
        double baz;
        orc_backaz(evla, evlo, stla, stlo, &baz);
        bazs[i] = baz * deg2rad;
    }
}

/* ================================================================== */
/* (2) this project's formulation of updatedispl                       */
/* ================================================================== */
/*
 * State: u = level 1 (time n), level 2 (time n-1) in the (1:2,...) arrays.
 * One call advances to n+1 and leaves level 1 = u^{n+1}, level 2 = u^n.
 *
 *  rho u_tt = X + Z + M, per component, where (Submain.f90 operators)
 *    ux: X = diffx2[l+2m](ux), Z = diffz2[m](ux), M = diffzx[l](uz) + diffxz[m](uz)
 *    uz: X = diffx2[m](uz),    Z = diffz2[l+2m](uz), M = diffzx[m](ux) + diffxz[l](ux)
 *  Free surface (izt == 1), row iz = 1: half-cell closure derived from the
 *  strain energy of the top layer (rows 1-2, trapezoidal rule in z), so the
 *  zero traction is the natural boundary condition (lb, mb = layer means):
 *    ux: X = diffx2[l+2m](ux), Z = 2/dz^2 H(m1,m2)(ux2-ux1),
 *        M = [Dx(lb (uz2-uz1)) + mb (Dx uz1 + Dx uz2)]/dz
 *    uz: X = diffx2[m](uz),    Z = 2/dz^2 H(M1,M2)(uz2-uz1),
 *        M = [Dx(mb (ux2-ux1)) + lb (Dx ux1 + Dx ux2)]/dz
 *  PML (stretched coordinates, split u = u1 + u2 + u3 in the padding strips):
 *    rho (u1_tt + dX u1_t)               = X - Dx(Mx psi_x)
 *    rho (u2_tt + dZ u2_t)               = Z - Dz(Mz psi_z)
 *    rho (u3_tt + (dX+dZ) u3_t + dX dZ u3) = M
 *    psi_t + d psi = d * D u   (first-derivative memory, the bp* arrays)
 *  Damping profiles d = d0 ((q-nbd)/(npml-nbd))^npower for q > nbd cells
 *  into the left/right/bottom strips (the first nbd cells are an undamped
 *  buffer that holds the hybrid ring), d0 = -(npower+1) vpml ln(R0) /
 *  (2 (npml-nbd) h); the optional top strip (is_PML_top, izt = 1-npml) has
 *  no buffer: d = d0t ((1-iz)/npml)^npower for its rows iz <= 0.
 *  Top strip: row 1 is then an ordinary cell (no free-surface closure), the
 *  strip keeps x and z memories like the bottom one, and it holds the field
 *  of the region beneath it without a hybrid correction: there is no ring
 *  across the top (the reference's twelve background arrays cover the
 *  sides and the bottom only), so the background must be one without a
 *  free surface (the FK background of izt < 1), whose waves leave through
 *  the top and are absorbed there.
 *  Multiaxial damping (Meza-Fajardo & Papageorgiou 2008, BSSA 98:1811):
 *  the split parts of a strip cell are damped along both axes,
 *    dX = dx(ix) + r dz(iz),   dZ = dz(iz) + r dx(ix)   (the dX, dZ above),
 *  with the ratio r (orc_set_pml_ratio), while the memories psi keep the
 *  classical profiles (psi_x decays with dx, psi_z with dz).  The classical
 *  PML (r = 0) lets guided modes of a strong low-velocity surface layer grow
 *  in the side strips at late time; the cross damping removes that growth.
 *  Driving the memories with dX, dZ too makes this second-order split system
 *  unstable, so the strips keep exactly the reference's memory arrays.  The damped cells use the material
 *  of the outermost undamped cell along the normal (normal-invariant layers;
 *  heterogeneous damped corners are unstable otherwise).
 *  Hybrid TF/SF: the physical region 1..nx x 1..nz holds the total field,
 *  the padding holds the scattered field.  Inside-ring cells add
 *  L[u0 on outside strips], outside-ring cells subtract L[u0 on inside strips].
 */

typedef struct {
    int norder, nx, nz, npml, izt, pml_top, npower, nbd, it, nt;
    int top_ring;   /* the top ring's background arrays are present */
    orc_real dx, dz;
    double dt;
    int NZ, NX;
} gdims;

static inline size_t g2(const gdims *g, int iz, int ix)
{
    return (size_t)(iz - g->izt) + (size_t)g->NZ * (size_t)(ix + g->npml - 1);
}

void orc_damping(int nx, int nz, int npml, int nbd, int is_pml_top, int npower, orc_real dx,
                 orc_real dz, orc_real vpml, double R0, orc_real *dpx, orc_real *dpz)
{
    int izt = is_pml_top ? 1 - npml : 1;
    int NX = nx + 2 * npml, NZ = nz + npml - izt + 1;
    int leff = npml - nbd;
    double lr = log(R0);
    double d0x = leff > 0 ? -(double)(npower + 1) * (double)vpml * lr / (2.0 * leff * (double)dx) : 0.0;
    double d0z = leff > 0 ? -(double)(npower + 1) * (double)vpml * lr / (2.0 * leff * (double)dz) : 0.0;

    for (int j = 0; j < NX; ++j) {
        int ix = j + 1 - npml;
        int q = ix <= 0 ? 1 - ix : (ix > nx ? ix - nx : 0);
        double d = 0.0;
        if (q > nbd && leff > 0) {
            double r = (double)(q - nbd) / (double)leff, p = 1.0;
            for (int e = 0; e < npower; ++e) p *= r;
            d = d0x * p;
        }
        dpx[j] = (orc_real)d;
    }
    for (int j = 0; j < NZ; ++j) {
        int iz = j + izt;
        double d = 0.0;
        if (iz > nz) {
            int q = iz - nz;
            if (q > nbd && leff > 0) {
                double r = (double)(q - nbd) / (double)leff, p = 1.0;
                for (int e = 0; e < npower; ++e) p *= r;
                d = d0z * p;
            }
        } else if (iz <= 0) {
            int q = 1 - iz;
            if (q > nbd && leff > 0) {
                double r = (double)(q - nbd) / (double)leff, p = 1.0;
                for (int e = 0; e < npower; ++e) p *= r;
                d = d0z * p;
            }
        }
        dpz[j] = (orc_real)d;
    }
}

/* multiaxial ratio r of the damping (see above) */
static double g_pml_ratio = ORC_PML_RATIO_DEFAULT;
void orc_set_pml_ratio(double r) { g_pml_ratio = r; }
double orc_pml_ratio(void) { return g_pml_ratio; }

/* central first derivative along x / z with the reduced-order rule of
 * diff1 (Submain.f90:192-209), times nothing; returns mysum (not divided) */
static inline orc_real cdx_sum(const gdims *g, const double *coef1, const orc_real *f, int iz, int ix)
{
    int norder = g->norder;
    int nox = imin(imin(ix + g->npml - 1, g->nx + g->npml - ix), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= nox; ++k) s = acc(s, CO(coef1, k, nox), f[g2(g, iz, ix + k)] - f[g2(g, iz, ix - k)]);
    return s;
}
static inline orc_real cdz_sum(const gdims *g, const double *coef1, const orc_real *f, int iz, int ix)
{
    int norder = g->norder;
    int noz = imin(imin(iz - g->izt, g->nz + g->npml - iz), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= noz; ++k) s = acc(s, CO(coef1, k, noz), f[g2(g, iz + k, ix)] - f[g2(g, iz - k, ix)]);
    return s;
}

/* modulus of the second-derivative terms: lam+2mu (mode 0) or mu (mode 1) */
static inline orc_real matval(int mode, orc_real lam, orc_real mu)
{
    return mode == 0 ? lam + 2.0f * mu : mu;
}

static orc_real x2_point_mode(const gdims *g, const double *coef0, const orc_real *lam, const orc_real *mu,
                           int mode, const orc_real *u, int iz, int ix)
{
    int norder = g->norder, nx = g->nx, npml = g->npml;
    int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
    orc_real s = 0.0f;
#define MAT(jx) matval(mode, lam[g2(g, iz, (jx))], mu[g2(g, iz, (jx))])
#define U(jx) u[g2(g, iz, (jx))]
    for (int k = 1; k <= nox; ++k) {
        int i = (2 * k - 1) / 2;
        int n1 = imin(imin(ix + i + npml - 1, nx + npml - ix - i), norder);
        orc_real s1 = 0.0f;
        for (int j = 1; j <= n1; ++j) s1 = acc(s1, CO(coef0, j, n1), U(ix + j + k - 1) - U(ix - j + k));
        int n2 = imin(imin(ix - i + npml - 1, nx + npml - ix + i), norder);
        orc_real s2 = 0.0f;
        for (int j = 1; j <= n2; ++j) s2 = acc(s2, CO(coef0, j, n2), U(ix + j - k) - U(ix - j - k + 1));
        orc_real t = hmean(MAT(ix + k), MAT(ix + k - 1)) * s1 - hmean(MAT(ix - k + 1), MAT(ix - k)) * s2;
        s = acc(s, CO(coef0, k, nox), t);
    }
#undef MAT
#undef U
    int in_reg = (iz >= 1 && iz <= g->nz && ix >= 1 && ix <= nx);
    orc_real dx2 = in_reg ? (orc_real)(1.0 / (double)(g->dx * g->dx)) : 1.0f / (g->dx * g->dx);
    return s * dx2;
}

static orc_real z2_point_mode(const gdims *g, const double *coef0, const orc_real *lam, const orc_real *mu,
                           int mode, const orc_real *u, int iz, int ix)
{
    int norder = g->norder, nz = g->nz, npml = g->npml, izt = g->izt;
    int noz = imin(imin(iz - izt, nz + npml - iz), norder);
    orc_real s = 0.0f;
#define MAT(jz) matval(mode, lam[g2(g, (jz), ix)], mu[g2(g, (jz), ix)])
#define U(jz) u[g2(g, (jz), ix)]
    for (int k = 1; k <= noz; ++k) {
        int i = (2 * k - 1) / 2;
        int n1 = imin(imin(iz + i - izt, nz + npml - iz - i), norder);
        orc_real s1 = 0.0f;
        for (int j = 1; j <= n1; ++j) s1 = acc(s1, CO(coef0, j, n1), U(iz + j + k - 1) - U(iz - j + k));
        int n2 = imin(imin(iz - i - izt, nz + npml - iz + i), norder);
        orc_real s2 = 0.0f;
        for (int j = 1; j <= n2; ++j) s2 = acc(s2, CO(coef0, j, n2), U(iz + j - k) - U(iz - j - k + 1));
        orc_real t = hmean(MAT(iz + k), MAT(iz + k - 1)) * s1 - hmean(MAT(iz - k + 1), MAT(iz - k)) * s2;
        s = acc(s, CO(coef0, k, noz), t);
    }
#undef MAT
#undef U
    return s * (1.0f / (g->dz * g->dz));
}

/* Dx( mat * Dz u ): diffzx at one output cell */
static orc_real zx_point(const gdims *g, const double *coef1, const orc_real *mat, const orc_real *u, int iz, int ix)
{
    int norder = g->norder;
    int nox = imin(imin(ix + g->npml - 1, g->nx + g->npml - ix), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= nox; ++k) {
        orc_real vp = mat[g2(g, iz, ix + k)] * cdz_sum(g, coef1, u, iz, ix + k) / g->dz;
        orc_real vm = mat[g2(g, iz, ix - k)] * cdz_sum(g, coef1, u, iz, ix - k) / g->dz;
        s = acc(s, CO(coef1, k, nox), vp - vm);
    }
    return s / g->dx;
}
/* Dz( mat * Dx u ): diffxz at one output cell */
static orc_real xz_point(const gdims *g, const double *coef1, const orc_real *mat, const orc_real *u, int iz, int ix)
{
    int norder = g->norder;
    int noz = imin(imin(iz - g->izt, g->nz + g->npml - iz), norder);
    orc_real s = 0.0f;
    for (int k = 1; k <= noz; ++k) {
        orc_real vp = mat[g2(g, iz + k, ix)] * cdx_sum(g, coef1, u, iz + k, ix) / g->dx;
        orc_real vm = mat[g2(g, iz - k, ix)] * cdx_sum(g, coef1, u, iz - k, ix) / g->dx;
        s = acc(s, CO(coef1, k, noz), vp - vm);
    }
    return s / g->dz;
}

/* the six RHS parts at one updated cell; lam/mu 2-D planes */
static void eval_parts_g(const gdims *g, const double *coef0, const double *coef1, const orc_real *lam,
                         const orc_real *mu, const orc_real *ux, const orc_real *uz, int iz, int ix, orc_real p[6])
{
    if (iz == g->izt && g->izt == 1) {
        /* free-surface row: half-cell closure from the strain energy of the
         * top layer (rows 1-2, trapezoidal in z), traction free naturally */
        size_t a = g2(g, 1, ix), b = g2(g, 2, ix);
        orc_real sdz2 = 2.0f / (g->dz * g->dz);
        orc_real M1 = lam[a] + 2.0f * mu[a], M2 = lam[b] + 2.0f * mu[b];
        int nox = imin(imin(ix + g->npml - 1, g->nx + g->npml - ix), g->norder);
        int norder = g->norder;
        /* Dx( lbar (uz2-uz1)/dz ) and Dx( mbar (ux2-ux1)/dz ) */
        orc_real sl = 0.0f, sm = 0.0f;
        for (int k = 1; k <= nox; ++k) {
            size_t a1 = g2(g, 1, ix + k), b1 = g2(g, 2, ix + k), a2 = g2(g, 1, ix - k), b2 = g2(g, 2, ix - k);
            orc_real lp = 0.5f * (lam[a1] + lam[b1]), lm = 0.5f * (lam[a2] + lam[b2]);
            orc_real mp = 0.5f * (mu[a1] + mu[b1]), mm = 0.5f * (mu[a2] + mu[b2]);
            sl = acc(sl, CO(coef1, k, nox), lp * (uz[b1] - uz[a1]) - lm * (uz[b2] - uz[a2]));
            sm = acc(sm, CO(coef1, k, nox), mp * (ux[b1] - ux[a1]) - mm * (ux[b2] - ux[a2]));
        }
        orc_real lbar = 0.5f * (lam[a] + lam[b]), mbar = 0.5f * (mu[a] + mu[b]);
        orc_real dxuz = (cdx_sum(g, coef1, uz, 1, ix) + cdx_sum(g, coef1, uz, 2, ix)) / g->dx;
        orc_real dxux = (cdx_sum(g, coef1, ux, 1, ix) + cdx_sum(g, coef1, ux, 2, ix)) / g->dx;
        p[0] = x2_point_mode(g, coef0, lam, mu, 0, ux, iz, ix);
        p[1] = (hmean(mu[a], mu[b]) * (ux[b] - ux[a])) * sdz2;
        p[2] = (sl / g->dx + mbar * dxuz) / g->dz;
        p[3] = x2_point_mode(g, coef0, lam, mu, 1, uz, iz, ix);
        p[4] = (hmean(M1, M2) * (uz[b] - uz[a])) * sdz2;
        p[5] = (sm / g->dx + lbar * dxux) / g->dz;
        return;
    }
    p[0] = x2_point_mode(g, coef0, lam, mu, 0, ux, iz, ix);
    p[1] = z2_point_mode(g, coef0, lam, mu, 1, ux, iz, ix);
    p[2] = zx_point(g, coef1, lam, uz, iz, ix) + xz_point(g, coef1, mu, uz, iz, ix);
    p[3] = x2_point_mode(g, coef0, lam, mu, 1, uz, iz, ix);
    p[4] = z2_point_mode(g, coef0, lam, mu, 0, uz, iz, ix);
    p[5] = zx_point(g, coef1, mu, ux, iz, ix) + xz_point(g, coef1, lam, ux, iz, ix);
}

void orc_eval_parts(int norder, int nx, int nz, int npml, int izt, orc_real dx, orc_real dz,
                    const double *coef0, const double *coef1, const orc_real *lam2d,
                    const orc_real *mu2d, const orc_real *ux2d, const orc_real *uz2d,
                    int iz, int ix, orc_real parts[6])
{
    gdims g;
    memset(&g, 0, sizeof g);
    g.norder = norder; g.nx = nx; g.nz = nz; g.npml = npml; g.izt = izt; g.dx = dx; g.dz = dz;
    g.NZ = nz + npml - izt + 1; g.NX = nx + 2 * npml;
    eval_parts_g(&g, coef0, coef1, lam2d, mu2d, ux2d, uz2d, iz, ix, parts);
}

/* --- strip storage --- */
enum { S_NONE = -1, S_L = 0, S_R = 1, S_B = 2, S_T = 3 };
static inline int strip_of(const gdims *g, int iz, int ix)
{
    if (iz >= 1 && iz <= g->nz) {
        if (ix <= 0) return S_L;
        if (ix > g->nx) return S_R;
        return S_NONE;
    }
    if (iz > g->nz) return S_B;
    return S_T;
}
/* cell offset (without level) inside a strip array */
static inline size_t strip_off(const gdims *g, int s, int iz, int ix)
{
    switch (s) {
    case S_L: return (size_t)(iz - 1) + (size_t)g->nz * (size_t)(ix - (1 - g->npml));
    case S_R: return (size_t)(iz - 1) + (size_t)g->nz * (size_t)(ix - (g->nx + 1));
    case S_B: return (size_t)(iz - (g->nz + 1)) + (size_t)g->npml * (size_t)(ix - (1 - g->npml));
    default:  return (size_t)(iz - (1 - g->npml)) + (size_t)g->npml * (size_t)(ix - (1 - g->npml));
    }
}

/* background ring: returns pointer offset in the right array or -1;
 * which: 0 = inside (reg) arrays, 1 = outside (pml) arrays */
static inline long bg_off(const gdims *g, int iz, int ix, int it, int *arr, int *which)
{
    int nx = g->nx, nz = g->nz, nbd = g->nbd;
    if (iz >= 1 && iz <= nz - nbd && ix >= 1 && ix <= nbd) {
        *arr = 0; *which = 0;
        return (long)((size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - 1) + (size_t)nbd * it));
    }
    if (iz >= 1 && iz <= nz - nbd && ix >= nx - nbd + 1 && ix <= nx) {
        *arr = 1; *which = 0;
        return (long)((size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - (nx - nbd + 1)) + (size_t)nbd * it));
    }
    if (iz >= nz - nbd + 1 && iz <= nz && ix >= 1 && ix <= nx) {
        *arr = 2; *which = 0;
        return (long)((size_t)(iz - (nz - nbd + 1)) + (size_t)nbd * ((size_t)(ix - 1) + (size_t)nx * it));
    }
    if (iz >= 1 && iz <= nz && ix >= 1 - nbd && ix <= 0) {
        *arr = 0; *which = 1;
        return (long)((size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (1 - nbd)) + (size_t)nbd * it));
    }
    if (iz >= 1 && iz <= nz && ix >= nx + 1 && ix <= nx + nbd) {
        *arr = 1; *which = 1;
        return (long)((size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (nx + 1)) + (size_t)nbd * it));
    }
    if (iz >= nz + 1 && iz <= nz + nbd && ix >= 1 - nbd && ix <= nx + nbd) {
        *arr = 2; *which = 1;
        return (long)((size_t)(iz - (nz + 1)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it));
    }
    if (g->top_ring) {
        /* top ring of an absorbing top layer: inside rows 1..nbd between the
         * side strips, outside rows 1-nbd..0 over cols 1-nbd..nx+nbd */
        if (iz >= 1 && iz <= nbd && ix >= nbd + 1 && ix <= nx - nbd) {
            *arr = 3; *which = 0;
            return (long)((size_t)(iz - 1) + (size_t)nbd * ((size_t)(ix - (nbd + 1)) + (size_t)(nx - 2 * nbd) * it));
        }
        if (iz >= 1 - nbd && iz <= 0 && ix >= 1 - nbd && ix <= nx + nbd) {
            *arr = 3; *which = 1;
            return (long)((size_t)(iz - (1 - nbd)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it));
        }
    }
    return -1;
}

static void updatedispl_body(const orc_real *const top[4], const int *norder_, const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *is_pml_top_, const int *npower_, const int *nbd_, const int *it_,
    const int *nt_, const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *vpml_,
    const double *R0_,
    const orc_real *uxl_reg, const orc_real *uzl_reg, const orc_real *uxr_reg, const orc_real *uzr_reg,
    const orc_real *uxb_reg, const orc_real *uzb_reg, const orc_real *uxl_pml, const orc_real *uzl_pml,
    const orc_real *uxr_pml, const orc_real *uzr_pml, const orc_real *uxb_pml, const orc_real *uzb_pml,
    const orc_real *c, const double *coef0, const double *coef1, orc_real *ux, orc_real *uz,
    orc_real *bpxlx, orc_real *bpzlx, orc_real *bpxrx, orc_real *bpzrx, orc_real *bpxtx, orc_real *bpxtz,
    orc_real *bpztx, orc_real *bpztz, orc_real *bpxbx, orc_real *bpxbz, orc_real *bpzbx, orc_real *bpzbz,
    orc_real *bux1l, orc_real *bux2l, orc_real *bux3l, orc_real *buz1l, orc_real *buz2l, orc_real *buz3l,
    orc_real *bux1r, orc_real *bux2r, orc_real *bux3r, orc_real *buz1r, orc_real *buz2r, orc_real *buz3r,
    orc_real *bux1b, orc_real *bux2b, orc_real *bux3b, orc_real *buz1b, orc_real *buz2b, orc_real *buz3b,
    orc_real *bux1t, orc_real *bux2t, orc_real *bux3t, orc_real *buz1t, orc_real *buz2t, orc_real *buz3t)
{
    gdims G, *g = &G;
    memset(&G, 0, sizeof G);
    G.norder = *norder_; G.nx = *nx_; G.nz = *nz_; G.npml = *npml_; G.izt = *izt_;
    G.pml_top = *is_pml_top_; G.npower = *npower_; G.nbd = *nbd_; G.it = *it_; G.nt = *nt_;
    G.dx = *dx_; G.dz = *dz_; G.dt = *dt_;
    G.NZ = G.nz + G.npml - G.izt + 1; G.NX = G.nx + 2 * G.npml;
    const int norder = G.norder, nx = G.nx, nz = G.nz, npml = G.npml, izt = G.izt, nbd = G.nbd;
    const size_t ncell = (size_t)G.NZ * (size_t)G.NX;
    (void)norder;

    /* copy-in of the level-1 fields and the material planes */
    orc_real *u1x = malloc(ncell * sizeof(orc_real)), *u1z = malloc(ncell * sizeof(orc_real));
    orc_real *rho = malloc(ncell * sizeof(orc_real)), *lam = malloc(ncell * sizeof(orc_real));
    orc_real *mu = malloc(ncell * sizeof(orc_real)), *lp2m = malloc(ncell * sizeof(orc_real));
    /* material seen by the stencils: the damped part of the PML (more than
     * nbd cells into the padding) repeats the outermost undamped cell along
     * the normal, so the layers are normal-invariant (heterogeneous damped
     * corners are unstable otherwise) */
    for (int ix = 1 - npml; ix <= nx + npml; ++ix)
        for (int iz = izt; iz <= nz + npml; ++iz) {
            int sx = ix < 1 - nbd ? 1 - nbd : (ix > nx + nbd ? nx + nbd : ix);
            int sz = iz > nz + nbd ? nz + nbd : (iz < 1 - nbd ? 1 - nbd : iz);
            size_t i = g2(g, iz, ix), j = g2(g, sz, sx);
            u1x[i] = ux[2 * i]; u1z[i] = uz[2 * i];
            rho[i] = c[3 * j]; lam[i] = c[3 * j + 1]; mu[i] = c[3 * j + 2];
            lp2m[i] = lam[i] + 2.0f * mu[i];
        }
    orc_real *dpx = malloc(G.NX * sizeof(orc_real)), *dpz = malloc(G.NZ * sizeof(orc_real));
    orc_damping(nx, nz, npml, nbd, G.pml_top, G.npower, G.dx, G.dz, *vpml_, *R0_, dpx, dpz);

    orc_real *bpx[4][2] = {{bpxlx, bpzlx}, {bpxrx, bpzrx}, {bpxbx, bpzbx}, {bpxtx, bpztx}};
    orc_real *bpz[4][2] = {{NULL, NULL}, {NULL, NULL}, {bpxbz, bpzbz}, {bpxtz, bpztz}};
    const orc_real ratio = (orc_real)g_pml_ratio;
    /* damping of the split parts of a strip cell along x and z (multiaxial) */
#define DAMP_X(iz, ix) (dpx[(ix) + npml - 1] + ratio * dpz[(iz) - izt])
#define DAMP_Z(iz, ix) (dpz[(iz) - izt] + ratio * dpx[(ix) + npml - 1])
    orc_real *bu[4][2][3] = {{{bux1l, bux2l, bux3l}, {buz1l, buz2l, buz3l}},
                          {{bux1r, bux2r, bux3r}, {buz1r, buz2r, buz3r}},
                          {{bux1b, bux2b, bux3b}, {buz1b, buz2b, buz3b}},
                          {{bux1t, bux2t, bux3t}, {buz1t, buz2t, buz3t}}};
    const int nstrip = G.pml_top ? 4 : 3;

    /* ---- phase A: first-derivative memories psi (level n) ---- */
    for (int s = 0; s < nstrip; ++s) {
        int iz0, iz1, ix0, ix1;
        if (s == S_L) { iz0 = 1; iz1 = nz; ix0 = 1 - npml; ix1 = 0; }
        else if (s == S_R) { iz0 = 1; iz1 = nz; ix0 = nx + 1; ix1 = nx + npml; }
        else if (s == S_B) { iz0 = nz + 1; iz1 = nz + npml; ix0 = 1 - npml; ix1 = nx + npml; }
        else { iz0 = 1 - npml; iz1 = 0; ix0 = 1 - npml; ix1 = nx + npml; }
#pragma omp parallel for schedule(dynamic)
        for (int ix = ix0; ix <= ix1; ++ix)
            for (int iz = iz0; iz <= iz1; ++iz) {
                size_t o = strip_off(g, s, iz, ix);
                /* memories: classical profiles */
                orc_real ddx = dpx[ix + npml - 1], ddz = dpz[iz - izt];
                if (ddx > 0.0f) {
                    orc_real b = (orc_real)exp(-(double)ddx * G.dt), a = 1.0f - b;
                    orc_real gx = cdx_sum(g, coef1, u1x, iz, ix) / G.dx;
                    orc_real gz = cdx_sum(g, coef1, u1z, iz, ix) / G.dx;
                    bpx[s][0][o] = b * bpx[s][0][o] + a * gx;
                    bpx[s][1][o] = b * bpx[s][1][o] + a * gz;
                }
                if (bpz[s][0] && ddz > 0.0f) {
                    orc_real b = (orc_real)exp(-(double)ddz * G.dt), a = 1.0f - b;
                    orc_real gx = cdz_sum(g, coef1, u1x, iz, ix) / G.dz;
                    orc_real gz = cdz_sum(g, coef1, u1z, iz, ix) / G.dz;
                    bpz[s][0][o] = b * bpz[s][0][o] + a * gx;
                    bpz[s][1][o] = b * bpz[s][1][o] + a * gz;
                }
            }
    }

    /* full-grid psi views (zero outside the strips) for the correction stencils */
    orc_real *psx[2], *psz[2];
    for (int q = 0; q < 2; ++q) {
        psx[q] = calloc(ncell, sizeof(orc_real));
        psz[q] = calloc(ncell, sizeof(orc_real));
    }
    for (int s = 0; s < nstrip; ++s) {
        int iz0, iz1, ix0, ix1;
        if (s == S_L) { iz0 = 1; iz1 = nz; ix0 = 1 - npml; ix1 = 0; }
        else if (s == S_R) { iz0 = 1; iz1 = nz; ix0 = nx + 1; ix1 = nx + npml; }
        else if (s == S_B) { iz0 = nz + 1; iz1 = nz + npml; ix0 = 1 - npml; ix1 = nx + npml; }
        else { iz0 = 1 - npml; iz1 = 0; ix0 = 1 - npml; ix1 = nx + npml; }
        for (int ix = ix0; ix <= ix1; ++ix)
            for (int iz = iz0; iz <= iz1; ++iz) {
                size_t o = strip_off(g, s, iz, ix), gi = g2(g, iz, ix);
                for (int q = 0; q < 2; ++q) {
                    psx[q][gi] = bpx[s][q][o];
                    if (bpz[s][q]) psz[q][gi] = bpz[s][q][o];
                }
            }
    }

    /* ---- phase B: operator passes (the Submain operators, reg + pml) ---- */
    orc_real *t1 = malloc(ncell * sizeof(orc_real)), *t2 = malloc(ncell * sizeof(orc_real));
    orc_real *P[6];
    for (int q = 0; q < 6; ++q) P[q] = calloc(ncell, sizeof(orc_real));
    orc_real dxf = G.dx, dzf = G.dz;
#define ADD2(dst) for (size_t i = 0; i < ncell; ++i) dst[i] = dst[i] + (t1[i] + t2[i])
    orc_diffx2_reg(norder, nx, nz, npml, izt, dxf, coef0, lp2m, u1x, t1);
    orc_diffx2_pml(norder, nx, nz, npml, izt, dxf, coef0, lp2m, u1x, t2); ADD2(P[0]);
    orc_diffz2_reg(norder, nx, nz, npml, izt, dzf, coef0, mu, u1x, t1);
    orc_diffz2_pml(norder, nx, nz, npml, izt, dzf, coef0, mu, u1x, t2); ADD2(P[1]);
    orc_diffzx_reg(norder, nx, nz, npml, izt, dxf, dzf, coef1, lam, u1z, t1);
    orc_diffzx_pml(norder, nx, nz, npml, izt, dxf, dzf, coef1, lam, u1z, t2); ADD2(P[2]);
    orc_diffxz_reg(norder, nx, nz, npml, izt, dxf, dzf, coef1, mu, u1z, t1);
    orc_diffxz_pml(norder, nx, nz, npml, izt, dxf, dzf, coef1, mu, u1z, t2); ADD2(P[2]);
    orc_diffx2_reg(norder, nx, nz, npml, izt, dxf, coef0, mu, u1z, t1);
    orc_diffx2_pml(norder, nx, nz, npml, izt, dxf, coef0, mu, u1z, t2); ADD2(P[3]);
    orc_diffz2_reg(norder, nx, nz, npml, izt, dzf, coef0, lp2m, u1z, t1);
    orc_diffz2_pml(norder, nx, nz, npml, izt, dzf, coef0, lp2m, u1z, t2); ADD2(P[4]);
    orc_diffzx_reg(norder, nx, nz, npml, izt, dxf, dzf, coef1, mu, u1x, t1);
    orc_diffzx_pml(norder, nx, nz, npml, izt, dxf, dzf, coef1, mu, u1x, t2); ADD2(P[5]);
    orc_diffxz_reg(norder, nx, nz, npml, izt, dxf, dzf, coef1, lam, u1x, t1);
    orc_diffxz_pml(norder, nx, nz, npml, izt, dxf, dzf, coef1, lam, u1x, t2); ADD2(P[5]);
#undef ADD2

    /* masked background fields for the hybrid correction; the top ring
     * (arrays top[0..3] = uxt_reg, uzt_reg, uxt_pml, uzt_pml) with a top PML */
    const orc_real *bgx[2][4] = {{uxl_reg, uxr_reg, uxb_reg, top[0]}, {uxl_pml, uxr_pml, uxb_pml, top[2]}};
    const orc_real *bgz[2][4] = {{uzl_reg, uzr_reg, uzb_reg, top[1]}, {uzl_pml, uzr_pml, uzb_pml, top[3]}};
    int have_bg = uxl_reg && uzl_reg && uxr_reg && uzr_reg && uxb_reg && uzb_reg && uxl_pml &&
                  uzl_pml && uxr_pml && uzr_pml && uxb_pml && uzb_pml && G.it >= 0 && G.it <= G.nt;
    G.top_ring = G.pml_top && top[0] && top[1] && top[2] && top[3];
    orc_real *wTx = NULL, *wTz = NULL, *wSx = NULL, *wSz = NULL;
    if (have_bg) {
        wTx = calloc(ncell, sizeof(orc_real)); wTz = calloc(ncell, sizeof(orc_real));
        wSx = calloc(ncell, sizeof(orc_real)); wSz = calloc(ncell, sizeof(orc_real));
        for (int ix = 1 - nbd; ix <= nx + nbd; ++ix)
            for (int iz = G.top_ring ? 1 - nbd : 1; iz <= nz + nbd; ++iz) {
                int arr, which;
                long o = bg_off(g, iz, ix, G.it, &arr, &which);
                if (o < 0) continue;
                size_t gi = g2(g, iz, ix);
                if (which == 0) { wTx[gi] = bgx[0][arr][o]; wTz[gi] = bgz[0][arr][o]; }
                else { wSx[gi] = bgx[1][arr][o]; wSz[gi] = bgz[1][arr][o]; }
            }
    }

    const orc_real dt2 = (orc_real)(G.dt * G.dt);
    const int iz_first = (izt == 1) ? 1 : izt + 1;
    orc_real *newx = malloc(ncell * sizeof(orc_real)), *newz = malloc(ncell * sizeof(orc_real));
    memcpy(newx, u1x, ncell * sizeof(orc_real)); /* cells not updated keep their value */
    memcpy(newz, u1z, ncell * sizeof(orc_real));

#pragma omp parallel for schedule(dynamic)
    for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix) {
        for (int iz = iz_first; iz <= nz + npml - 1; ++iz) {
            size_t gi = g2(g, iz, ix);
            orc_real p[6];
            for (int q = 0; q < 6; ++q) p[q] = P[q][gi];
            /* Row 1: with a free surface the closure replaces the passes.
             * Under an absorbing top layer (izt < 1) every cell is evaluated
             * pointwise: the reference's passes are written for izt = 1 near
             * the top (diffz2_reg and diffz2_pml leave row 1 of columns 1..nx
             * without z-terms, Submain.f90:971,1062-1063; diffxz_reg builds
             * its intermediate from row 1 down, :616, so the full-order Dz of
             * rows 2..norder reads zeros above it).  The pointwise operators
             * equal the passes bit for bit wherever those are complete. */
            if ((iz == 1 && izt == 1) || izt != 1) eval_parts_g(g, coef0, coef1, lam, mu, u1x, u1z, iz, ix, p);

            int s = strip_of(g, iz, ix);
            /* PML memory corrections (nonzero only next to damped cells) */
            {
                int surf = (iz == 1 && izt == 1);
                int nox = imin(imin(ix + npml - 1, nx + npml - ix), G.norder);
                int noz = imin(imin(iz - izt, nz + npml - iz), G.norder);
                orc_real sx0 = 0.0f, sx1 = 0.0f, sz0 = 0.0f, sz1 = 0.0f;
                for (int k = 1; k <= nox; ++k) {
                    size_t a = g2(g, iz, ix + k), b = g2(g, iz, ix - k);
                    orc_real ma = lp2m[a], mb = lp2m[b];
                    sx0 = acc(sx0, CO(coef1, k, nox), ma * psx[0][a] - mb * psx[0][b]);
                    sx1 = acc(sx1, CO(coef1, k, nox), mu[a] * psx[1][a] - mu[b] * psx[1][b]);
                }
                for (int k = 1; k <= noz; ++k) {
                    size_t a = g2(g, iz + k, ix), b = g2(g, iz - k, ix);
                    sz0 = acc(sz0, CO(coef1, k, noz), mu[a] * psz[0][a] - mu[b] * psz[0][b]);
                    sz1 = acc(sz1, CO(coef1, k, noz), lp2m[a] * psz[1][a] - lp2m[b] * psz[1][b]);
                }
                p[0] = p[0] - sx0 / G.dx;
                p[3] = p[3] - sx1 / G.dx;
                if (!surf) { p[1] = p[1] - sz0 / G.dz; p[4] = p[4] - sz1 / G.dz; }
            }
            /* hybrid TF/SF correction */
            if (have_bg) {
                int arr, which;
                long o = bg_off(g, iz, ix, G.it, &arr, &which);
                if (o >= 0) {
                    orc_real q[6];
                    if (which == 0) {
                        eval_parts_g(g, coef0, coef1, lam, mu, wSx, wSz, iz, ix, q);
                        for (int m = 0; m < 6; ++m) p[m] = p[m] + q[m];
                    } else {
                        eval_parts_g(g, coef0, coef1, lam, mu, wTx, wTz, iz, ix, q);
                        for (int m = 0; m < 6; ++m) p[m] = p[m] - q[m];
                    }
                }
            }

            orc_real bf = dt2 / rho[gi];
            if (s == S_NONE) {
                orc_real rx = (p[0] + p[1]) + p[2], rz = (p[3] + p[4]) + p[5];
                newx[gi] = (2.0f * u1x[gi] - ux[2 * gi + 1]) + bf * rx;
                newz[gi] = (2.0f * u1z[gi] - uz[2 * gi + 1]) + bf * rz;
            } else {
                size_t o = strip_off(g, s, iz, ix);
                orc_real ax = 0.5f * DAMP_X(iz, ix) * (orc_real)G.dt;
                orc_real az = 0.5f * DAMP_Z(iz, ix) * (orc_real)G.dt;
                orc_real c3 = 2.0f - 4.0f * ax * az;
                for (int comp = 0; comp < 2; ++comp) {
                    orc_real *b1 = bu[s][comp][0], *b2 = bu[s][comp][1], *b3 = bu[s][comp][2];
                    orc_real X = p[3 * comp], Z = p[3 * comp + 1], Mm = p[3 * comp + 2];
                    orc_real n1 = (2.0f * b1[2 * o] - (1.0f - ax) * b1[2 * o + 1] + bf * X) / (1.0f + ax);
                    orc_real n2 = (2.0f * b2[2 * o] - (1.0f - az) * b2[2 * o + 1] + bf * Z) / (1.0f + az);
                    orc_real n3 = (c3 * b3[2 * o] - (1.0f - ax - az) * b3[2 * o + 1] + bf * Mm) / (1.0f + ax + az);
                    b1[2 * o + 1] = b1[2 * o]; b1[2 * o] = n1;
                    b2[2 * o + 1] = b2[2 * o]; b2[2 * o] = n2;
                    b3[2 * o + 1] = b3[2 * o]; b3[2 * o] = n3;
                    orc_real tot = (n1 + n2) + n3;
                    if (comp == 0) newx[gi] = tot; else newz[gi] = tot;
                }
            }
        }
    }
#undef DAMP_X
#undef DAMP_Z
    /* level shift: level 2 <- level 1, level 1 <- new */
    for (size_t i = 0; i < ncell; ++i) {
        ux[2 * i + 1] = ux[2 * i]; ux[2 * i] = newx[i];
        uz[2 * i + 1] = uz[2 * i]; uz[2 * i] = newz[i];
    }

    free(newx); free(newz);
    free(wTx); free(wTz); free(wSx); free(wSz);
    for (int q = 0; q < 6; ++q) free(P[q]);
    free(t1); free(t2);
    for (int q = 0; q < 2; ++q) { free(psx[q]); free(psz[q]); }
    free(dpx); free(dpz);
    free(u1x); free(u1z); free(rho); free(lam); free(mu); free(lp2m);
}

/* the reference's 68-argument call: no top ring (a top PML then has an open
 * hybrid contour at the top) */
void orc_updatedispl(const int *norder_, const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *is_pml_top_, const int *npower_, const int *nbd_, const int *it_,
    const int *nt_, const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *vpml_,
    const double *R0_,
    const orc_real *uxl_reg, const orc_real *uzl_reg, const orc_real *uxr_reg, const orc_real *uzr_reg,
    const orc_real *uxb_reg, const orc_real *uzb_reg, const orc_real *uxl_pml, const orc_real *uzl_pml,
    const orc_real *uxr_pml, const orc_real *uzr_pml, const orc_real *uxb_pml, const orc_real *uzb_pml,
    const orc_real *c, const double *coef0, const double *coef1, orc_real *ux, orc_real *uz,
    orc_real *bpxlx, orc_real *bpzlx, orc_real *bpxrx, orc_real *bpzrx, orc_real *bpxtx, orc_real *bpxtz,
    orc_real *bpztx, orc_real *bpztz, orc_real *bpxbx, orc_real *bpxbz, orc_real *bpzbx, orc_real *bpzbz,
    orc_real *bux1l, orc_real *bux2l, orc_real *bux3l, orc_real *buz1l, orc_real *buz2l, orc_real *buz3l,
    orc_real *bux1r, orc_real *bux2r, orc_real *bux3r, orc_real *buz1r, orc_real *buz2r, orc_real *buz3r,
    orc_real *bux1b, orc_real *bux2b, orc_real *bux3b, orc_real *buz1b, orc_real *buz2b, orc_real *buz3b,
    orc_real *bux1t, orc_real *bux2t, orc_real *bux3t, orc_real *buz1t, orc_real *buz2t, orc_real *buz3t)
{
    const orc_real *const none[4] = {NULL, NULL, NULL, NULL};
    updatedispl_body(none, norder_, nx_, nz_, npml_, izt_, is_pml_top_, npower_, nbd_, it_, nt_, dx_, dz_, dt_,
        vpml_, R0_, uxl_reg, uzl_reg, uxr_reg, uzr_reg, uxb_reg, uzb_reg, uxl_pml, uzl_pml, uxr_pml, uzr_pml,
        uxb_pml, uzb_pml, c, coef0, coef1, ux, uz, bpxlx, bpzlx, bpxrx, bpzrx, bpxtx, bpxtz, bpztx, bpztz,
        bpxbx, bpxbz, bpzbx, bpzbz, bux1l, bux2l, bux3l, buz1l, buz2l, buz3l, bux1r, bux2r, bux3r, buz1r,
        buz2r, buz3r, bux1b, bux2b, bux3b, buz1b, buz2b, buz3b, bux1t, bux2t, bux3t, buz1t, buz2t, buz3t);
}

/* the same with the top ring's four background arrays (uxt_reg, uzt_reg,
 * uxt_pml, uzt_pml of orc_comput_background_wavefield_ext), which close the
 * hybrid contour across the top when is_pml_top */
void orc_updatedispl_ext(const orc_real *uxt_reg, const orc_real *uzt_reg, const orc_real *uxt_pml,
    const orc_real *uzt_pml, const int *norder_, const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *is_pml_top_, const int *npower_, const int *nbd_, const int *it_,
    const int *nt_, const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *vpml_,
    const double *R0_,
    const orc_real *uxl_reg, const orc_real *uzl_reg, const orc_real *uxr_reg, const orc_real *uzr_reg,
    const orc_real *uxb_reg, const orc_real *uzb_reg, const orc_real *uxl_pml, const orc_real *uzl_pml,
    const orc_real *uxr_pml, const orc_real *uzr_pml, const orc_real *uxb_pml, const orc_real *uzb_pml,
    const orc_real *c, const double *coef0, const double *coef1, orc_real *ux, orc_real *uz,
    orc_real *bpxlx, orc_real *bpzlx, orc_real *bpxrx, orc_real *bpzrx, orc_real *bpxtx, orc_real *bpxtz,
    orc_real *bpztx, orc_real *bpztz, orc_real *bpxbx, orc_real *bpxbz, orc_real *bpzbx, orc_real *bpzbz,
    orc_real *bux1l, orc_real *bux2l, orc_real *bux3l, orc_real *buz1l, orc_real *buz2l, orc_real *buz3l,
    orc_real *bux1r, orc_real *bux2r, orc_real *bux3r, orc_real *buz1r, orc_real *buz2r, orc_real *buz3r,
    orc_real *bux1b, orc_real *bux2b, orc_real *bux3b, orc_real *buz1b, orc_real *buz2b, orc_real *buz3b,
    orc_real *bux1t, orc_real *bux2t, orc_real *bux3t, orc_real *buz1t, orc_real *buz2t, orc_real *buz3t)
{
    const orc_real *const top[4] = {uxt_reg, uzt_reg, uxt_pml, uzt_pml};
    updatedispl_body(top, norder_, nx_, nz_, npml_, izt_, is_pml_top_, npower_, nbd_, it_, nt_, dx_, dz_, dt_,
        vpml_, R0_, uxl_reg, uzl_reg, uxr_reg, uzr_reg, uxb_reg, uzb_reg, uxl_pml, uzl_pml, uxr_pml, uzr_pml,
        uxb_pml, uzb_pml, c, coef0, coef1, ux, uz, bpxlx, bpzlx, bpxrx, bpzrx, bpxtx, bpxtz, bpztx, bpztz,
        bpxbx, bpxbz, bpzbx, bpzbz, bux1l, bux2l, bux3l, buz1l, buz2l, buz3l, bux1r, bux2r, bux3r, buz1r,
        buz2r, buz3r, bux1b, bux2b, bux3b, buz1b, buz2b, buz3b, bux1t, bux2t, bux3t, buz1t, buz2t, buz3t);
}

/* ================================================================== */
/* (3) SH (transverse) motion: this project's formulation, no         */
/*     counterpart in the reference (it keeps the in-plane P-SV part)  */
/* ================================================================== */
/*
 *  rho uy_tt = Dx(mu Dx uy) + Dz(mu Dz uy): the X and Z terms of the uz
 *  equation with modulus mu (diffx2 / diffz2 of Submain.f90 with mat = mu).
 *  Free surface (izt = 1), row 1: Z = 2/dz^2 H(mu1, mu2)(uy2 - uy1) (zero
 *  traction mu duy/dz from the same half-cell energy form).  PML: split
 *  uy = u1 + u2, rho(u1_tt + dX u1_t) = X - Dx(mu psi_x),
 *  rho(u2_tt + dZ u2_t) = Z - Dz(mu psi_z), the memories and the multiaxial
 *  damping of the P-SV formulation.  Hybrid ring: as for P-SV, one
 *  component.  Every cell is evaluated with the pointwise operators.
 *  bg: uyl_reg, uyr_reg, uyb_reg, uyt_reg, uyl_pml, uyr_pml, uyb_pml,
 *  uyt_pml (the reference's strip shapes; the top two with a top layer,
 *  NULL: the contour stays open there).  bp: x-memories of the L, R, B, T
 *  strips, then z-memories of B and T.  bu[s][0/1]: the split parts u1, u2
 *  of strip s (1:2, cells).
 */
void orc_updatedispl_sh(int norder, int nx, int nz, int npml, int pml_top, int npower, int it, int nt,
                        orc_real dx, orc_real dz, double dt, orc_real vpml, double R0,
                        const orc_real *const bg[8], const orc_real *c, const double *coef0,
                        const double *coef1, orc_real *uy, orc_real *const bp[6], orc_real *const bu[4][2])
{
    gdims G, *g = &G;
    memset(&G, 0, sizeof G);
    const int izt = pml_top ? 1 - npml : 1, nbd = 2 * norder - 1;
    G.norder = norder; G.nx = nx; G.nz = nz; G.npml = npml; G.izt = izt; G.pml_top = pml_top;
    G.npower = npower; G.nbd = nbd; G.it = it; G.nt = nt; G.dx = dx; G.dz = dz; G.dt = dt;
    G.NZ = nz + npml - izt + 1; G.NX = nx + 2 * npml;
    const size_t ncell = (size_t)G.NZ * (size_t)G.NX;
    orc_real *u1 = malloc(ncell * sizeof(orc_real)), *rho = malloc(ncell * sizeof(orc_real));
    orc_real *mu = malloc(ncell * sizeof(orc_real)), *lam = malloc(ncell * sizeof(orc_real));
    for (int ix = 1 - npml; ix <= nx + npml; ++ix)
        for (int iz = izt; iz <= nz + npml; ++iz) {
            int sx = ix < 1 - nbd ? 1 - nbd : (ix > nx + nbd ? nx + nbd : ix);
            int sz = iz > nz + nbd ? nz + nbd : (iz < 1 - nbd ? 1 - nbd : iz);
            size_t i = g2(g, iz, ix), j = g2(g, sz, sx);
            u1[i] = uy[2 * i];
            rho[i] = c[3 * j]; lam[i] = c[3 * j + 1]; mu[i] = c[3 * j + 2];
        }
    orc_real *dpx = malloc(G.NX * sizeof(orc_real)), *dpz = malloc(G.NZ * sizeof(orc_real));
    orc_damping(nx, nz, npml, nbd, pml_top, npower, dx, dz, vpml, R0, dpx, dpz);
    const int nstrip = pml_top ? 4 : 3;
    orc_real *const bpx[4] = {bp[0], bp[1], bp[2], bp[3]};
    orc_real *const bpz[4] = {NULL, NULL, bp[4], bp[5]};
    /* memories (level n) */
    for (int s = 0; s < nstrip; ++s) {
        int iz0, iz1, ix0, ix1;
        if (s == S_L) { iz0 = 1; iz1 = nz; ix0 = 1 - npml; ix1 = 0; }
        else if (s == S_R) { iz0 = 1; iz1 = nz; ix0 = nx + 1; ix1 = nx + npml; }
        else if (s == S_B) { iz0 = nz + 1; iz1 = nz + npml; ix0 = 1 - npml; ix1 = nx + npml; }
        else { iz0 = 1 - npml; iz1 = 0; ix0 = 1 - npml; ix1 = nx + npml; }
        for (int ix = ix0; ix <= ix1; ++ix)
            for (int iz = iz0; iz <= iz1; ++iz) {
                size_t o = strip_off(g, s, iz, ix);
                orc_real ddx = dpx[ix + npml - 1], ddz = dpz[iz - izt];
                if (ddx > 0.0f) {
                    orc_real b = (orc_real)exp(-(double)ddx * dt), a = 1.0f - b;
                    bpx[s][o] = b * bpx[s][o] + a * (cdx_sum(g, coef1, u1, iz, ix) / dx);
                }
                if (bpz[s] && ddz > 0.0f) {
                    orc_real b = (orc_real)exp(-(double)ddz * dt), a = 1.0f - b;
                    bpz[s][o] = b * bpz[s][o] + a * (cdz_sum(g, coef1, u1, iz, ix) / dz);
                }
            }
    }
    orc_real *psx = calloc(ncell, sizeof(orc_real)), *psz = calloc(ncell, sizeof(orc_real));
    for (int s = 0; s < nstrip; ++s) {
        int iz0, iz1, ix0, ix1;
        if (s == S_L) { iz0 = 1; iz1 = nz; ix0 = 1 - npml; ix1 = 0; }
        else if (s == S_R) { iz0 = 1; iz1 = nz; ix0 = nx + 1; ix1 = nx + npml; }
        else if (s == S_B) { iz0 = nz + 1; iz1 = nz + npml; ix0 = 1 - npml; ix1 = nx + npml; }
        else { iz0 = 1 - npml; iz1 = 0; ix0 = 1 - npml; ix1 = nx + npml; }
        for (int ix = ix0; ix <= ix1; ++ix)
            for (int iz = iz0; iz <= iz1; ++iz) {
                size_t o = strip_off(g, s, iz, ix), gi = g2(g, iz, ix);
                psx[gi] = bpx[s][o];
                if (bpz[s]) psz[gi] = bpz[s][o];
            }
    }
    /* masked background (inside: wT, outside: wS) */
    int have_bg = bg && bg[0] && bg[1] && bg[2] && bg[4] && bg[5] && bg[6] && it >= 0 && it <= nt;
    G.top_ring = pml_top && have_bg && bg[3] && bg[7];
    orc_real *wT = calloc(ncell, sizeof(orc_real)), *wS = calloc(ncell, sizeof(orc_real));
    if (have_bg)
        for (int ix = 1 - nbd; ix <= nx + nbd; ++ix)
            for (int iz = G.top_ring ? 1 - nbd : 1; iz <= nz + nbd; ++iz) {
                int arr, which;
                long o = bg_off(g, iz, ix, it, &arr, &which);
                if (o < 0) continue;
                (which == 0 ? wT : wS)[g2(g, iz, ix)] = bg[which * 4 + arr][o];
            }
    const orc_real dt2 = (orc_real)(dt * dt);
    const int iz_first = izt == 1 ? 1 : izt + 1;
    orc_real *unew = malloc(ncell * sizeof(orc_real));
    memcpy(unew, u1, ncell * sizeof(orc_real));
    const orc_real ratio = (orc_real)g_pml_ratio;
#pragma omp parallel for schedule(dynamic)
    for (int ix = 2 - npml; ix <= nx + npml - 1; ++ix)
        for (int iz = iz_first; iz <= nz + npml - 1; ++iz) {
            size_t gi = g2(g, iz, ix);
            const int surf = iz == 1 && izt == 1;
            orc_real X = x2_point_mode(g, coef0, lam, mu, 1, u1, iz, ix), Z;
            if (surf) {
                size_t a = g2(g, 1, ix), b = g2(g, 2, ix);
                Z = (hmean(mu[a], mu[b]) * (u1[b] - u1[a])) * (2.0f / (dz * dz));
            } else {
                Z = z2_point_mode(g, coef0, lam, mu, 1, u1, iz, ix);
            }
            {
                int nox = imin(imin(ix + npml - 1, nx + npml - ix), norder);
                int noz = imin(imin(iz - izt, nz + npml - iz), norder);
                orc_real sx = 0.0f, sz = 0.0f;
                for (int k = 1; k <= nox; ++k) {
                    size_t a = g2(g, iz, ix + k), b = g2(g, iz, ix - k);
                    sx = acc(sx, CO(coef1, k, nox), mu[a] * psx[a] - mu[b] * psx[b]);
                }
                for (int k = 1; k <= noz; ++k) {
                    size_t a = g2(g, iz + k, ix), b = g2(g, iz - k, ix);
                    sz = acc(sz, CO(coef1, k, noz), mu[a] * psz[a] - mu[b] * psz[b]);
                }
                X = X - sx / dx;
                if (!surf) Z = Z - sz / dz;
            }
            if (have_bg) {
                int arr, which;
                if (bg_off(g, iz, ix, it, &arr, &which) >= 0) {
                    /* inside: the operator of the background on the padding;
                     * outside: minus that of the background inside */
                    const orc_real *w = which == 0 ? wS : wT;
                    orc_real qx = x2_point_mode(g, coef0, lam, mu, 1, w, iz, ix), qz;
                    if (surf) {
                        size_t a = g2(g, 1, ix), b = g2(g, 2, ix);
                        qz = (hmean(mu[a], mu[b]) * (w[b] - w[a])) * (2.0f / (dz * dz));
                    } else {
                        qz = z2_point_mode(g, coef0, lam, mu, 1, w, iz, ix);
                    }
                    if (which == 0) { X = X + qx; Z = Z + qz; }
                    else { X = X - qx; Z = Z - qz; }
                }
            }
            orc_real bf = dt2 / rho[gi];
            int s = strip_of(g, iz, ix);
            if (s == S_NONE) {
                unew[gi] = (2.0f * u1[gi] - uy[2 * gi + 1]) + bf * (X + Z);
            } else {
                size_t o = strip_off(g, s, iz, ix);
                orc_real ax = 0.5f * (dpx[ix + npml - 1] + ratio * dpz[iz - izt]) * (orc_real)dt;
                orc_real az = 0.5f * (dpz[iz - izt] + ratio * dpx[ix + npml - 1]) * (orc_real)dt;
                orc_real *b1 = bu[s][0], *b2 = bu[s][1];
                orc_real n1 = (2.0f * b1[2 * o] - (1.0f - ax) * b1[2 * o + 1] + bf * X) / (1.0f + ax);
                orc_real n2 = (2.0f * b2[2 * o] - (1.0f - az) * b2[2 * o + 1] + bf * Z) / (1.0f + az);
                b1[2 * o + 1] = b1[2 * o]; b1[2 * o] = n1;
                b2[2 * o + 1] = b2[2 * o]; b2[2 * o] = n2;
                unew[gi] = n1 + n2;
            }
        }
    for (size_t i = 0; i < ncell; ++i) { uy[2 * i + 1] = uy[2 * i]; uy[2 * i] = unew[i]; }
    free(unew); free(wT); free(wS); free(psx); free(psz); free(dpx); free(dpz);
    free(u1); free(rho); free(mu); free(lam);
}
