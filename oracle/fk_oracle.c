/*
 * CPU oracle, FK background (libFK3D role) -- TEST INFRASTRUCTURE ONLY.
 *
 * The reference's comput_background_wavefield is binary-only; this is this
 * project's formulation (FP64 throughout), pinned in tests/ by closed forms:
 * the homogeneous half-space free-surface response, layer-over-half-space
 * conversion delays (Ps, PpPs, PsPs), causality and translation invariance.
 *
 * Formulation (z down, e^{i w (p x + q z - t)} plane waves, forward transform
 * F(w) = int f(t) e^{iwt} dt):
 *  - each layer j (top z_j) carries amplitudes a_j = (P down, S down, P up,
 *    S up) referenced at z_j; displacement-traction vector
 *    b = (ux, uz, szz/(iw), sxz/(iw)) = E_j Lambda_j(z - z_j) a_j;
 *  - b is continuous across interfaces; in the half-space the only up-going
 *    wave is the incident one (P or SV); the free surface z = 0 is traction
 *    free (izt = 1), or, with an absorbing top layer (izt < 1), the top
 *    layer extends upward and carries no down-going wave;
 *  - the incident pulse s(t) = ds exp(-(pi f0 (t - t0))^2), t0 = 1.5/f0, passes
 *    the entry column at depth z_ref (the deeper of the ring bottom and the
 *    half-space top) at t = t0;
 *  - complex frequency w + i eps (Phinney) suppresses FFT wrap-around;
 *  - per depth: IFFT to an oversampled time grid (dt/2), then each ring column
 *    is the same trace delayed by p |x - x_entry| (plane-wave translation
 *    invariance), read with 4-point cubic Lagrange interpolation.
 *  The wave travels away from the event: its along-profile slowness is
 *  -p cos(baz) (propagation azimuth = back-azimuth + 180 deg), so the
 *  direction sign is s = -sign(rayp) * sign(sum cos(bazs)); s > 0 is +x, and
 *  ux carries s (mirror image of the +x solution).  With the driver's radial
 *  projection seisx * cos(baz) (src/FDFK2D.f90:540-542) and z down, R/Z of the
 *  direct P is positive, as in the reference's published P receiver function.  Left strips use the left
 *  model, right strips the right model, the bottom strip blends them linearly
 *  in x (identical when the models agree).
 */
#include "fdfk_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;
static const double K_PI = 3.14159265358979323846;

static void fft_inplace(cplx *a, int n, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * K_PI / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                cplx w = cexp(I * (ang * k));
                cplx u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

/* E matrix (rows ux, uz, szz/iw, sxz/iw; columns Pd, Sd, Pu, Su) */
static void emat(double vp, double vs, double rho, double p, cplx ea, cplx eb, cplx E[4][4])
{
    double g = 1.0 - 2.0 * vs * vs * p * p;
    E[0][0] = vp * p;       E[0][1] = vs * eb;       E[0][2] = vp * p;       E[0][3] = vs * eb;
    E[1][0] = vp * ea;      E[1][1] = -vs * p;       E[1][2] = -vp * ea;     E[1][3] = vs * p;
    E[2][0] = rho * vp * g; E[2][1] = -2.0 * rho * vs * vs * vs * p * eb;
    E[2][2] = rho * vp * g; E[2][3] = -2.0 * rho * vs * vs * vs * p * eb;
    E[3][0] = 2.0 * rho * vs * vs * vp * p * ea; E[3][1] = rho * vs * g;
    E[3][2] = -2.0 * rho * vs * vs * vp * p * ea; E[3][3] = -rho * vs * g;
}

/* solve A X = B (4x4, nrhs columns) by Gaussian elimination with partial pivoting */
static int solve4(cplx A[4][4], cplx *B, int nrhs)
{
    int n = 4;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (cabs(A[r][c]) > cabs(A[piv][c])) piv = r;
        if (cabs(A[piv][c]) == 0.0) return -1;
        if (piv != c) {
            for (int k = 0; k < n; ++k) { cplx t = A[c][k]; A[c][k] = A[piv][k]; A[piv][k] = t; }
            for (int k = 0; k < nrhs; ++k) { cplx t = B[c * nrhs + k]; B[c * nrhs + k] = B[piv * nrhs + k]; B[piv * nrhs + k] = t; }
        }
        for (int r = c + 1; r < n; ++r) {
            cplx f = A[r][c] / A[c][c];
            for (int k = c; k < n; ++k) A[r][k] -= f * A[c][k];
            for (int k = 0; k < nrhs; ++k) B[r * nrhs + k] -= f * B[c * nrhs + k];
        }
    }
    for (int c = n - 1; c >= 0; --c)
        for (int k = 0; k < nrhs; ++k) {
            cplx s = B[c * nrhs + k];
            for (int j = c + 1; j < n; ++j) s -= A[c][j] * B[j * nrhs + k];
            B[c * nrhs + k] = s / A[c][c];
        }
    return 0;
}

static cplx vslow(double v, double p)
{
    double a = 1.0 / (v * v) - p * p;
    return a >= 0.0 ? csqrt(a + 0.0 * I) : I * sqrt(-a);
}

typedef struct {
    int nl;
    double *zt, *vp, *vs, *rho;
    cplx *ea, *eb;
    cplx (*E)[4][4];
    cplx (*Q)[4][4]; /* Q_j = E_j^{-1} E_{j+1} */
} fkmodel;

static int fk_setup(fkmodel *m, int nlayer, const orc_real *fk, double p)
{
    m->nl = nlayer;
    m->zt = malloc(nlayer * sizeof(double)); m->vp = malloc(nlayer * sizeof(double));
    m->vs = malloc(nlayer * sizeof(double)); m->rho = malloc(nlayer * sizeof(double));
    m->ea = malloc(nlayer * sizeof(cplx)); m->eb = malloc(nlayer * sizeof(cplx));
    m->E = malloc(nlayer * sizeof(*m->E)); m->Q = malloc(nlayer * sizeof(*m->Q));
    for (int j = 0; j < nlayer; ++j) {
        m->zt[j] = fk[4 * j + 0]; m->vp[j] = fk[4 * j + 1]; m->vs[j] = fk[4 * j + 2]; m->rho[j] = fk[4 * j + 3];
        if (!(m->vs[j] > 0.0) || !(m->vp[j] > m->vs[j])) return -1;
        m->ea[j] = vslow(m->vp[j], p); m->eb[j] = vslow(m->vs[j], p);
        emat(m->vp[j], m->vs[j], m->rho[j], p, m->ea[j], m->eb[j], m->E[j]);
    }
    for (int j = 0; j + 1 < nlayer; ++j) {
        cplx A[4][4], B[16];
        memcpy(A, m->E[j], sizeof A);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) B[r * 4 + c] = m->E[j + 1][r][c];
        if (solve4(A, B, 4)) return -1;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) m->Q[j][r][c] = B[r * 4 + c];
    }
    return 0;
}
static void fk_free(fkmodel *m)
{
    free(m->zt); free(m->vp); free(m->vs); free(m->rho); free(m->ea); free(m->eb); free(m->E); free(m->Q);
}

/* layer amplitudes a[j][0..3] for complex frequency w and incident amplitude
 * ainc.  free_surface: the top z_0 is traction free; otherwise the top layer
 * extends upward without end and carries no down-going wave (the
 * configuration of an absorbing top layer) */
static void fk_amplitudes(const fkmodel *m, int src_is_p, int free_surface, cplx w, cplx ainc, cplx (*a)[4])
{
    int N = m->nl;
    cplx basis[3][4] = {{0}};
    basis[0][src_is_p ? 2 : 3] = 1.0;
    basis[1][0] = 1.0;
    basis[2][1] = 1.0;
    cplx cur[3][4];
    memcpy(cur, basis, sizeof cur);
    /* store per-layer amplitudes of each basis vector */
    cplx (*ab)[3][4] = malloc(N * sizeof(*ab));
    memcpy(ab[N - 1], cur, sizeof cur);
    for (int j = N - 2; j >= 0; --j) {
        double h = m->zt[j + 1] - m->zt[j];
        cplx inv[4] = {cexp(-I * w * m->ea[j] * h), cexp(-I * w * m->eb[j] * h),
                       cexp(I * w * m->ea[j] * h), cexp(I * w * m->eb[j] * h)};
        for (int b = 0; b < 3; ++b) {
            cplx t[4];
            for (int r = 0; r < 4; ++r) {
                cplx s = 0;
                for (int c = 0; c < 4; ++c) s += m->Q[j][r][c] * ab[j + 1][b][c];
                t[r] = inv[r] * s;
            }
            memcpy(ab[j][b], t, sizeof t);
        }
    }
    /* the two top conditions of each basis vector: the traction at the
     * surface (layer 0, z = z_0 reference), or the down-going P and S
     * amplitudes of layer 0 */
    cplx T[3][2];
    for (int b = 0; b < 3; ++b)
        for (int r = 0; r < 2; ++r) {
            cplx s = 0;
            if (free_surface)
                for (int c = 0; c < 4; ++c) s += m->E[0][2 + r][c] * ab[0][b][c];
            else
                s = ab[0][b][r];
            T[b][r] = s;
        }
    /* T0*ainc + x1*T1 + x2*T2 = 0 */
    cplx det = T[1][0] * T[2][1] - T[2][0] * T[1][1];
    cplx r0 = -T[0][0] * ainc, r1 = -T[0][1] * ainc;
    cplx x1 = (r0 * T[2][1] - T[2][0] * r1) / det;
    cplx x2 = (T[1][0] * r1 - r0 * T[1][1]) / det;
    for (int j = 0; j < N; ++j)
        for (int c = 0; c < 4; ++c) a[j][c] = ainc * ab[j][0][c] + x1 * ab[j][1][c] + x2 * ab[j][2][c];
    free(ab);
}

static int layer_of(const fkmodel *m, double z)
{
    int j = 0;
    while (j + 1 < m->nl && z >= m->zt[j + 1]) ++j;
    return j;
}

/* traces g(z, t) on the oversampled grid: out[(iz*nt_f + n)*2 + comp] */
typedef struct {
    int nfft, nzq;
    double dtf, eps;
    double *g; /* nzq x nfft x 2 */
} fktraces;

static int fk_compute(const fkmodel *m, int src_is_p, int free_surface, double p, double f0, double ds, double dtf,
                      int nfft, int nzq, const double *zq, double zref, double t0, fktraces *out)
{
    int N = m->nl;
    double T = nfft * dtf;
    double eps = 27.6 / T;
    double zN = m->zt[N - 1];
    cplx etaN = src_is_p ? m->ea[N - 1] : m->eb[N - 1];
    double tref = t0 + creal(etaN) * (zref - zN);
    out->nfft = nfft; out->nzq = nzq; out->dtf = dtf; out->eps = eps;
    out->g = calloc((size_t)nzq * nfft * 2, sizeof(double));
    cplx *spec = calloc((size_t)nzq * 2 * (nfft / 2 + 1), sizeof(cplx));
    cplx (*a)[4] = malloc(N * sizeof(*a));
    int *lay = malloc(nzq * sizeof(int));
    for (int q = 0; q < nzq; ++q) lay[q] = layer_of(m, zq[q]);
    for (int k = 0; k < nfft / 2; ++k) {
        double wr = 2.0 * K_PI * k / T;
        cplx w = wr + I * eps;
        cplx x = w / (2.0 * K_PI * f0);
        cplx S = ds / (f0 * sqrt(K_PI)) * cexp(-x * x);
        cplx ainc = S * cexp(I * w * tref);
        fk_amplitudes(m, src_is_p, free_surface, w, ainc, a);
        for (int q = 0; q < nzq; ++q) {
            int j = lay[q];
            double dz = zq[q] - m->zt[j];
            cplx ph[4] = {cexp(I * w * m->ea[j] * dz), cexp(I * w * m->eb[j] * dz),
                          cexp(-I * w * m->ea[j] * dz), cexp(-I * w * m->eb[j] * dz)};
            cplx ux = 0, uz = 0;
            for (int c = 0; c < 4; ++c) {
                ux += m->E[j][0][c] * ph[c] * a[j][c];
                uz += m->E[j][1][c] * ph[c] * a[j][c];
            }
            spec[((size_t)q * 2 + 0) * (nfft / 2 + 1) + k] = ux;
            spec[((size_t)q * 2 + 1) * (nfft / 2 + 1) + k] = uz;
        }
    }
    cplx *buf = malloc(nfft * sizeof(cplx));
    for (int q = 0; q < nzq; ++q)
        for (int c = 0; c < 2; ++c) {
            const cplx *sp = spec + ((size_t)q * 2 + c) * (nfft / 2 + 1);
            buf[0] = sp[0];
            for (int k = 1; k < nfft / 2; ++k) { buf[k] = sp[k]; buf[nfft - k] = conj(sp[k]); }
            buf[nfft / 2] = 0.0;
            fft_inplace(buf, nfft, -1);
            for (int n = 0; n < nfft; ++n)
                out->g[((size_t)q * nfft + n) * 2 + c] = creal(buf[n]) * exp(eps * n * dtf) / T;
        }
    free(buf); free(spec); free(a); free(lay);
    return 0;
}

/* cubic Lagrange read of trace (stride 2) at time t; zero before t = 0 */
static double interp(const double *g, int nfft, double dtf, double t)
{
    if (t < 0.0) return 0.0;
    double u = t / dtf;
    int i = (int)floor(u);
    double f = u - i;
    if (i + 2 >= nfft / 2) return 0.0; /* beyond the usable half of the window */
    double gm1 = i >= 1 ? g[2 * (i - 1)] : 0.0, g0 = g[2 * i], g1 = g[2 * (i + 1)], g2 = g[2 * (i + 2)];
    double wm1 = -f * (f - 1.0) * (f - 2.0) / 6.0, w0 = (f + 1.0) * (f - 1.0) * (f - 2.0) / 2.0;
    double w1 = -(f + 1.0) * f * (f - 2.0) / 2.0, w2 = (f + 1.0) * f * (f - 1.0) / 6.0;
    return wm1 * gm1 + w0 * g0 + w1 * g1 + w2 * g2;
}

static int next_pow2(long n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

int orc_fk_traces(int nlayer, const orc_real *fkm, int src_is_p, double p, orc_real f0, orc_real ds,
                  double dt, int ntq, int nzq, const double *z, double zref, double t0,
                  double *ux, double *uz)
{
    return orc_fk_traces_fs(nlayer, fkm, src_is_p, 1, p, f0, ds, dt, ntq, nzq, z, zref, t0, ux, uz);
}

int orc_fk_traces_fs(int nlayer, const orc_real *fkm, int src_is_p, int free_surface, double p, orc_real f0,
                     orc_real ds, double dt, int ntq, int nzq, const double *z, double zref, double t0,
                     double *ux, double *uz)
{
    fkmodel m;
    if (fk_setup(&m, nlayer, fkm, p)) { fk_free(&m); return -1; }
    double dtf = dt / 2.0;
    double tneed = ntq * dt + 2.0 * t0 + 10.0 / f0;
    int nfft = next_pow2((long)ceil(2.0 * tneed / dtf));
    fktraces tr;
    fk_compute(&m, src_is_p, free_surface, p, f0, ds, dtf, nfft, nzq, z, zref, t0, &tr);
    for (int q = 0; q < nzq; ++q)
        for (int n = 0; n < ntq; ++n) {
            const double *g = tr.g + (size_t)q * nfft * 2;
            ux[(size_t)q * ntq + n] = interp(g, nfft, dtf, n * dt);
            uz[(size_t)q * ntq + n] = interp(g + 1, nfft, dtf, n * dt);
        }
    free(tr.g);
    fk_free(&m);
    return 0;
}

static void background_body(orc_real *const top[4], const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *nbd_, const int *nt_, const int *nlayer_left,
    const int *nlayer_right, const int *src_type, const orc_real *x0_, const orc_real *z0_,
    const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *ds_, const orc_real *f0_,
    const double *rayp_, const double *bazs, const orc_real *fk_left, const orc_real *fk_right,
    orc_real *uxl_reg, orc_real *uzl_reg, orc_real *uxr_reg, orc_real *uzr_reg, orc_real *uxb_reg,
    orc_real *uzb_reg, orc_real *uxl_pml, orc_real *uzl_pml, orc_real *uxr_pml, orc_real *uzr_pml,
    orc_real *uxb_pml, orc_real *uzb_pml)
{
    const int nx = *nx_, nz = *nz_, npml = *npml_, nbd = *nbd_, nt = *nt_;
    const double x0 = *x0_, z0 = *z0_, dx = *dx_, dz = *dz_, dt = *dt_;
    const int src_is_p = *src_type != 0;
    /* izt = 1: free surface at row 1; izt < 1 (a top PML): no free surface */
    const int free_surface = *izt_ == 1;
    double cs = 0.0;
    for (int i = 0; i < nx + 2 * npml; ++i) cs += cos(bazs[i]);
    double sgn = ((*rayp_ < 0.0) ? -1.0 : 1.0) * (cs < 0.0 ? 1.0 : -1.0);
    double p = fabs(*rayp_);
    double t0 = 1.5 / (double)*f0_;

    /* the top ring (rows 1-nbd..nbd) with an absorbing top layer */
    const int top_ring = !free_surface && top[0] && top[1] && top[2] && top[3];
    /* depths of the ring rows zf..nz+nbd (above the model top, the top
     * layer continues upward) */
    const int zf = top_ring ? 1 - nbd : 1;
    int nzq = nz + nbd - zf + 1;
    double *zq = malloc(nzq * sizeof(double));
    for (int q = 0; q < nzq; ++q) zq[q] = z0 + (double)(q + zf - 1) * dz;
    double xl = x0 + (double)(1 - nbd - 1) * dx, xr = x0 + (double)(nx + nbd - 1) * dx;
    double xe = sgn > 0 ? xl : xr;

    fkmodel ml, mr;
    int okl = fk_setup(&ml, *nlayer_left, fk_left, p) == 0;
    int okr = fk_setup(&mr, *nlayer_right, fk_right, p) == 0;
    if (!okl || !okr) { /* invalid model: leave the caller's (zeroed) arrays alone */
        fk_free(&ml); fk_free(&mr); free(zq);
        return;
    }
    double zref_l = fmax(zq[nzq - 1], ml.zt[ml.nl - 1]);
    double zref_r = fmax(zq[nzq - 1], mr.zt[mr.nl - 1]);
    double zref = fmax(zref_l, zref_r);
    double dtf = dt / 2.0;
    double tneed = nt * dt + p * (xr - xl) + 2.0 * t0 + 10.0 / *f0_;
    int nfft = next_pow2((long)ceil(2.0 * tneed / dtf));
    fktraces tl = {0}, trr = {0};
    if (okl) fk_compute(&ml, src_is_p, free_surface, p, *f0_, *ds_, dtf, nfft, nzq, zq, zref, t0, &tl);
    if (okr) fk_compute(&mr, src_is_p, free_surface, p, *f0_, *ds_, dtf, nfft, nzq, zq, zref, t0, &trr);

#define TRACE(tr, q) ((tr).g + (size_t)(q) * nfft * 2)
    for (int it = 0; it <= nt; ++it) {
        double t = it * dt;
        for (int ix = 1 - nbd; ix <= nx + nbd; ++ix) {
            double x = x0 + (double)(ix - 1) * dx;
            double tau = p * (sgn > 0 ? (x - xe) : (xe - x));
            double w = (double)(ix - 1) / (double)(nx - 1);
            if (w < 0.0) w = 0.0;
            if (w > 1.0) w = 1.0;
            for (int iz = zf; iz <= nz + nbd; ++iz) {
                int q = iz - zf;
                int inL = ix <= nbd, inR = ix >= nx - nbd + 1, inB = iz >= nz - nbd + 1, inT = iz <= nbd && top_ring;
                if (!inL && !inR && !inB && !inT) continue;
                double vx[2] = {0, 0}, vz[2] = {0, 0};
                if (okl) { vx[0] = interp(TRACE(tl, q), nfft, dtf, t - tau); vz[0] = interp(TRACE(tl, q) + 1, nfft, dtf, t - tau); }
                if (okr) { vx[1] = interp(TRACE(trr, q), nfft, dtf, t - tau); vz[1] = interp(TRACE(trr, q) + 1, nfft, dtf, t - tau); }
                double bx, bz;
                if (iz <= nz - nbd || (iz <= nz && !inB)) {
                    int side = inL ? 0 : 1;
                    bx = vx[side]; bz = vz[side];
                } else if (iz <= nz) {
                    /* bottom inside strip */
                    bx = (1.0 - w) * vx[0] + w * vx[1]; bz = (1.0 - w) * vz[0] + w * vz[1];
                } else {
                    bx = (1.0 - w) * vx[0] + w * vx[1]; bz = (1.0 - w) * vz[0] + w * vz[1];
                }
                /* left/right strips at rows 1..nz use the side model */
                if (iz >= 1 && iz <= nz && (ix <= 0 || ix > nx)) {
                    int side = ix <= 0 ? 0 : 1;
                    bx = vx[side]; bz = vz[side];
                }
                /* the top ring blends the two models in x, like the bottom */
                if (top_ring && (iz <= 0 || (iz <= nbd && ix >= nbd + 1 && ix <= nx - nbd))) {
                    bx = (1.0 - w) * vx[0] + w * vx[1]; bz = (1.0 - w) * vz[0] + w * vz[1];
                }
                orc_real fx = (orc_real)(sgn * bx), fz = (orc_real)bz;
                /* scatter into the twelve arrays (FDFK2D.f90:330-355 shapes) */
                if (iz >= 1 && iz <= nz - nbd && ix >= 1 && ix <= nbd) {
                    size_t o = (size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - 1) + (size_t)nbd * it);
                    uxl_reg[o] = fx; uzl_reg[o] = fz;
                }
                if (iz >= 1 && iz <= nz - nbd && ix >= nx - nbd + 1 && ix <= nx) {
                    size_t o = (size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - (nx - nbd + 1)) + (size_t)nbd * it);
                    uxr_reg[o] = fx; uzr_reg[o] = fz;
                }
                if (iz >= nz - nbd + 1 && iz <= nz && ix >= 1 && ix <= nx) {
                    size_t o = (size_t)(iz - (nz - nbd + 1)) + (size_t)nbd * ((size_t)(ix - 1) + (size_t)nx * it);
                    uxb_reg[o] = fx; uzb_reg[o] = fz;
                }
                if (iz >= 1 && iz <= nz && ix >= 1 - nbd && ix <= 0) {
                    size_t o = (size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (1 - nbd)) + (size_t)nbd * it);
                    uxl_pml[o] = fx; uzl_pml[o] = fz;
                }
                if (iz >= 1 && iz <= nz && ix >= nx + 1 && ix <= nx + nbd) {
                    size_t o = (size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (nx + 1)) + (size_t)nbd * it);
                    uxr_pml[o] = fx; uzr_pml[o] = fz;
                }
                if (iz >= nz + 1 && iz <= nz + nbd && ix >= 1 - nbd && ix <= nx + nbd) {
                    size_t o = (size_t)(iz - (nz + 1)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it);
                    uxb_pml[o] = fx; uzb_pml[o] = fz;
                }
                /* the top ring: inside rows 1..nbd between the side strips,
                 * outside rows 1-nbd..0; the two models blended in x as in
                 * the bottom strips */
                if (top_ring && iz >= 1 && iz <= nbd && ix >= nbd + 1 && ix <= nx - nbd) {
                    size_t o = (size_t)(iz - 1) + (size_t)nbd * ((size_t)(ix - (nbd + 1)) + (size_t)(nx - 2 * nbd) * it);
                    top[0][o] = fx; top[1][o] = fz;
                }
                if (top_ring && iz >= 1 - nbd && iz <= 0) {
                    size_t o = (size_t)(iz - (1 - nbd)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it);
                    top[2][o] = fx; top[3][o] = fz;
                }
            }
        }
    }
#undef TRACE
    free(tl.g); free(trr.g);
    fk_free(&ml); fk_free(&mr);
    free(zq);
}

void orc_comput_background_wavefield(const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *nbd_, const int *nt_, const int *nlayer_left,
    const int *nlayer_right, const int *src_type, const orc_real *x0_, const orc_real *z0_,
    const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *ds_, const orc_real *f0_,
    const double *rayp_, const double *bazs, const orc_real *fk_left, const orc_real *fk_right,
    orc_real *uxl_reg, orc_real *uzl_reg, orc_real *uxr_reg, orc_real *uzr_reg, orc_real *uxb_reg,
    orc_real *uzb_reg, orc_real *uxl_pml, orc_real *uzl_pml, orc_real *uxr_pml, orc_real *uzr_pml,
    orc_real *uxb_pml, orc_real *uzb_pml)
{
    orc_real *const none[4] = {NULL, NULL, NULL, NULL};
    background_body(none, nx_, nz_, npml_, izt_, nbd_, nt_, nlayer_left, nlayer_right, src_type, x0_, z0_, dx_,
                    dz_, dt_, ds_, f0_, rayp_, bazs, fk_left, fk_right, uxl_reg, uzl_reg, uxr_reg, uzr_reg,
                    uxb_reg, uzb_reg, uxl_pml, uzl_pml, uxr_pml, uzr_pml, uxb_pml, uzb_pml);
}

/* the same plus the top ring (izt < 1): uxt_reg/uzt_reg (nbd, nx-2nbd, nt+1)
 * for rows 1..nbd x cols nbd+1..nx-nbd, uxt_pml/uzt_pml (nbd, nx+2nbd, nt+1)
 * for rows 1-nbd..0 x cols 1-nbd..nx+nbd, z fastest */
void orc_comput_background_wavefield_ext(orc_real *uxt_reg, orc_real *uzt_reg, orc_real *uxt_pml,
    orc_real *uzt_pml, const int *nx_, const int *nz_, const int *npml_,
    const int *izt_, const int *nbd_, const int *nt_, const int *nlayer_left,
    const int *nlayer_right, const int *src_type, const orc_real *x0_, const orc_real *z0_,
    const orc_real *dx_, const orc_real *dz_, const double *dt_, const orc_real *ds_, const orc_real *f0_,
    const double *rayp_, const double *bazs, const orc_real *fk_left, const orc_real *fk_right,
    orc_real *uxl_reg, orc_real *uzl_reg, orc_real *uxr_reg, orc_real *uzr_reg, orc_real *uxb_reg,
    orc_real *uzb_reg, orc_real *uxl_pml, orc_real *uzl_pml, orc_real *uxr_pml, orc_real *uzr_pml,
    orc_real *uxb_pml, orc_real *uzb_pml)
{
    orc_real *const top[4] = {uxt_reg, uzt_reg, uxt_pml, uzt_pml};
    background_body(top, nx_, nz_, npml_, izt_, nbd_, nt_, nlayer_left, nlayer_right, src_type, x0_, z0_, dx_,
                    dz_, dt_, ds_, f0_, rayp_, bazs, fk_left, fk_right, uxl_reg, uzl_reg, uxr_reg, uzr_reg,
                    uxb_reg, uzb_reg, uxl_pml, uzl_pml, uxr_pml, uzr_pml, uxb_pml, uzb_pml);
}

/* ------------------------------------------------------------------ */
/* SH plane wave (transverse motion uy): b = (uy, syz/(iw)) = E a with   */
/* a = (S down, S up), E = [[1, 1], [mu eta, -mu eta]] per layer; the  */
/* incident S up-going in the half-space; top: traction free, or no    */
/* down-going wave in the top layer (absorbing top layer).  No         */
/* counterpart in the reference.                                       */
/* ------------------------------------------------------------------ */
static void sh_emat(const fkmodel *m, int j, cplx E[2][2])
{
    cplx mh = m->rho[j] * m->vs[j] * m->vs[j] * m->eb[j];
    E[0][0] = 1.0; E[0][1] = 1.0;
    E[1][0] = mh;  E[1][1] = -mh;
}

static void sh_amplitudes(const fkmodel *m, int free_surface, cplx w, cplx ainc, cplx (*a)[2])
{
    int N = m->nl;
    cplx (*ab)[2][2] = malloc(N * sizeof(*ab));   /* [layer][basis][comp] */
    ab[N - 1][0][0] = 0.0; ab[N - 1][0][1] = 1.0;   /* incident up-going S */
    ab[N - 1][1][0] = 1.0; ab[N - 1][1][1] = 0.0;   /* down-going S leaving the stack */
    for (int j = N - 2; j >= 0; --j) {
        cplx Ej[2][2], Ek[2][2];
        sh_emat(m, j, Ej);
        sh_emat(m, j + 1, Ek);
        /* Q = Ej^-1 Ek */
        cplx det = Ej[0][0] * Ej[1][1] - Ej[0][1] * Ej[1][0];
        cplx inv[2][2] = {{Ej[1][1] / det, -Ej[0][1] / det}, {-Ej[1][0] / det, Ej[0][0] / det}};
        double h = m->zt[j + 1] - m->zt[j];
        cplx ph[2] = {cexp(-I * w * m->eb[j] * h), cexp(I * w * m->eb[j] * h)};
        for (int b = 0; b < 2; ++b) {
            cplx v[2];
            for (int r = 0; r < 2; ++r) v[r] = Ek[r][0] * ab[j + 1][b][0] + Ek[r][1] * ab[j + 1][b][1];
            for (int r = 0; r < 2; ++r) ab[j][b][r] = ph[r] * (inv[r][0] * v[0] + inv[r][1] * v[1]);
        }
    }
    cplx E0[2][2];
    sh_emat(m, 0, E0);
    cplx T[2];
    for (int b = 0; b < 2; ++b)
        T[b] = free_surface ? E0[1][0] * ab[0][b][0] + E0[1][1] * ab[0][b][1] : ab[0][b][0];
    cplx x1 = -T[0] * ainc / T[1];
    for (int j = 0; j < N; ++j)
        for (int c = 0; c < 2; ++c) a[j][c] = ainc * ab[j][0][c] + x1 * ab[j][1][c];
    free(ab);
}

/* traces uy(z, t) on the oversampled grid, out->g[(q*nfft + n)*2] (the
 * second slot of each pair left zero) */
static int sh_compute(const fkmodel *m, int free_surface, double p, double f0, double ds, double dtf, int nfft,
                      int nzq, const double *zq, double zref, double t0, fktraces *out)
{
    int N = m->nl;
    double T = nfft * dtf;
    double eps = 27.6 / T;
    double tref = t0 + creal(m->eb[N - 1]) * (zref - m->zt[N - 1]);
    (void)p;
    out->nfft = nfft; out->nzq = nzq; out->dtf = dtf; out->eps = eps;
    out->g = calloc((size_t)nzq * nfft * 2, sizeof(double));
    cplx *spec = calloc((size_t)nzq * (nfft / 2 + 1), sizeof(cplx));
    cplx (*a)[2] = malloc(N * sizeof(*a));
    for (int k = 0; k < nfft / 2; ++k) {
        double wr = 2.0 * K_PI * k / T;
        cplx w = wr + I * eps;
        cplx x = w / (2.0 * K_PI * f0);
        cplx S = ds / (f0 * sqrt(K_PI)) * cexp(-x * x);
        cplx ainc = S * cexp(I * w * tref);
        sh_amplitudes(m, free_surface, w, ainc, a);
        for (int q = 0; q < nzq; ++q) {
            int j = layer_of(m, zq[q]);
            double dz = zq[q] - m->zt[j];
            spec[(size_t)q * (nfft / 2 + 1) + k] = cexp(I * w * m->eb[j] * dz) * a[j][0] + cexp(-I * w * m->eb[j] * dz) * a[j][1];
        }
    }
    cplx *buf = malloc(nfft * sizeof(cplx));
    for (int q = 0; q < nzq; ++q) {
        const cplx *sp = spec + (size_t)q * (nfft / 2 + 1);
        buf[0] = sp[0];
        for (int k = 1; k < nfft / 2; ++k) { buf[k] = sp[k]; buf[nfft - k] = conj(sp[k]); }
        buf[nfft / 2] = 0.0;
        fft_inplace(buf, nfft, -1);
        for (int n = 0; n < nfft; ++n) out->g[((size_t)q * nfft + n) * 2] = creal(buf[n]) * exp(eps * n * dtf) / T;
    }
    free(buf); free(spec); free(a);
    return 0;
}

int orc_fk_traces_sh(int nlayer, const orc_real *fkm, int free_surface, double p, orc_real f0, orc_real ds,
                     double dt, int ntq, int nzq, const double *z, double zref, double t0, double *uy)
{
    fkmodel m;
    if (fk_setup(&m, nlayer, fkm, p)) { fk_free(&m); return -1; }
    double dtf = dt / 2.0;
    double tneed = ntq * dt + 2.0 * t0 + 10.0 / f0;
    int nfft = next_pow2((long)ceil(2.0 * tneed / dtf));
    fktraces tr;
    sh_compute(&m, free_surface, p, f0, ds, dtf, nfft, nzq, z, zref, t0, &tr);
    for (int q = 0; q < nzq; ++q)
        for (int n = 0; n < ntq; ++n) uy[(size_t)q * ntq + n] = interp(tr.g + (size_t)q * nfft * 2, nfft, dtf, n * dt);
    free(tr.g);
    fk_free(&m);
    return 0;
}

/* SH background on the ring, the P-SV background's geometry and
 * conventions (translation by p |x - x_entry|, side models on the side
 * strips, blended in x on the bottom and top strips, the direction sign
 * applied to uy as to ux); out[8]: uyl_reg, uyr_reg, uyb_reg, uyt_reg,
 * uyl_pml, uyr_pml, uyb_pml, uyt_pml (the top two with izt < 1 only) */
void orc_background_sh(int nx, int nz, int npml, int izt, int nbd, int nt, orc_real x0, orc_real z0,
                       orc_real dx, orc_real dz, double dt, orc_real ds, orc_real f0, double rayp,
                       const double *bazs, int nlayer_left, const orc_real *fk_left, int nlayer_right,
                       const orc_real *fk_right, orc_real *const out[8])
{
    const int free_surface = izt == 1, top = !free_surface && out[3] && out[7];
    double cs = 0.0;
    for (int i = 0; i < nx + 2 * npml; ++i) cs += cos(bazs[i]);
    double sgn = ((rayp < 0.0) ? -1.0 : 1.0) * (cs < 0.0 ? 1.0 : -1.0);
    double p = fabs(rayp), t0 = 1.5 / (double)f0;
    const int zf = top ? 1 - nbd : 1;
    int nzq = nz + nbd - zf + 1;
    double *zq = malloc(nzq * sizeof(double));
    for (int q = 0; q < nzq; ++q) zq[q] = z0 + (double)(q + zf - 1) * dz;
    double xl = x0 + (double)(1 - nbd - 1) * dx, xr = x0 + (double)(nx + nbd - 1) * dx;
    double xe = sgn > 0 ? xl : xr;
    fkmodel ml, mr;
    if (fk_setup(&ml, nlayer_left, fk_left, p) || fk_setup(&mr, nlayer_right, fk_right, p)) {
        fk_free(&ml); fk_free(&mr); free(zq);
        return;
    }
    double zref = fmax(fmax(zq[nzq - 1], ml.zt[ml.nl - 1]), fmax(zq[nzq - 1], mr.zt[mr.nl - 1]));
    double dtf = dt / 2.0;
    double tneed = nt * dt + p * (xr - xl) + 2.0 * t0 + 10.0 / f0;
    int nfft = next_pow2((long)ceil(2.0 * tneed / dtf));
    fktraces tl, trr;
    sh_compute(&ml, free_surface, p, f0, ds, dtf, nfft, nzq, zq, zref, t0, &tl);
    sh_compute(&mr, free_surface, p, f0, ds, dtf, nfft, nzq, zq, zref, t0, &trr);
    for (int it = 0; it <= nt; ++it)
        for (int ix = 1 - nbd; ix <= nx + nbd; ++ix) {
            double x = x0 + (double)(ix - 1) * dx;
            double tau = p * (sgn > 0 ? (x - xe) : (xe - x));
            double w = (double)(ix - 1) / (double)(nx - 1);
            w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
            for (int iz = zf; iz <= nz + nbd; ++iz) {
                int q = iz - zf;
                double vl = interp(tl.g + (size_t)q * nfft * 2, nfft, dtf, it * dt - tau);
                double vr = interp(trr.g + (size_t)q * nfft * 2, nfft, dtf, it * dt - tau);
                const int side = (iz >= 1 && iz <= nz) && (ix <= nbd || ix >= nx - nbd + 1) && !(iz > nz - nbd && ix >= 1 && ix <= nx);
                double v = side ? (ix <= nbd ? vl : vr) : (1.0 - w) * vl + w * vr;
                orc_real f = (orc_real)(sgn * v);
                if (iz >= 1 && iz <= nz - nbd && ix >= 1 && ix <= nbd)
                    out[0][(size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - 1) + (size_t)nbd * it)] = f;
                if (iz >= 1 && iz <= nz - nbd && ix >= nx - nbd + 1 && ix <= nx)
                    out[1][(size_t)(iz - 1) + (size_t)(nz - nbd) * ((size_t)(ix - (nx - nbd + 1)) + (size_t)nbd * it)] = f;
                if (iz >= nz - nbd + 1 && iz <= nz && ix >= 1 && ix <= nx)
                    out[2][(size_t)(iz - (nz - nbd + 1)) + (size_t)nbd * ((size_t)(ix - 1) + (size_t)nx * it)] = f;
                if (top && iz >= 1 && iz <= nbd && ix >= nbd + 1 && ix <= nx - nbd)
                    out[3][(size_t)(iz - 1) + (size_t)nbd * ((size_t)(ix - (nbd + 1)) + (size_t)(nx - 2 * nbd) * it)] = f;
                if (iz >= 1 && iz <= nz && ix >= 1 - nbd && ix <= 0)
                    out[4][(size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (1 - nbd)) + (size_t)nbd * it)] = f;
                if (iz >= 1 && iz <= nz && ix >= nx + 1 && ix <= nx + nbd)
                    out[5][(size_t)(iz - 1) + (size_t)nz * ((size_t)(ix - (nx + 1)) + (size_t)nbd * it)] = f;
                if (iz >= nz + 1 && iz <= nz + nbd && ix >= 1 - nbd && ix <= nx + nbd)
                    out[6][(size_t)(iz - (nz + 1)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it)] = f;
                if (top && iz >= 1 - nbd && iz <= 0)
                    out[7][(size_t)(iz - (1 - nbd)) + (size_t)nbd * ((size_t)(ix - (1 - nbd)) + (size_t)(nx + 2 * nbd) * it)] = f;
            }
        }
    free(tl.g); free(trr.g);
    fk_free(&ml); fk_free(&mr);
    free(zq);
}
