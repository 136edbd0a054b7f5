// ---------------------------------------------------------------------
// Frame kernel: every updated cell outside the interior rectangle (the
// surface rows of the total-field region, the PML strips and the cells
// next to the padded-grid boundary), in 2-D tiles.  A tile lies on one side
// of the hybrid interface, so the view of the ring (total field inside,
// scattered field outside, the corrected reads of libHybrid) is applied
// once when the tile and its halo of 2*NO-1 cells are staged in LDS (cells
// beyond the grid staged as zeros).  Phase 1 evaluates, for the tile, the
// shared intermediate quantities with the reduced-order rule of
// src/Submain.f90:192-195:
//   FX1/FX2  x-fluxes h*D1 at faces (f, f+1), orders n1(f) / n2(f)
//   FZ1/FZ2  z-fluxes
//   DZ, DX   collocated first derivatives of u (cdz_sum / cdx_sum)
// Phase 2 combines them per cell into the six RHS parts (with the
// free-surface closure on row 1), then frame_update applies the memory
// corrections and the (split) update.
// ---------------------------------------------------------------------
namespace fdfk {

constexpr int kFrameThreads = 256;
constexpr int kFrameLds = 12288;   // max floats of LDS per block (host keeps tiles within it)

__host__ __device__ inline int frame_lds_floats(int tx, int tz, int NO)
{
    const int H = 2 * NO - 1;
    return 4 * (tx + 2 * H) * (tz + 2 * H) + 4 * (tx + 2 * NO - 1) * tz + 4 * tx * (tz + 2 * NO - 1) +
           4 * (tx + 2 * NO) * tz + 4 * tx * (tz + 2 * NO);
}

// e / d and e % d for 0 <= e < 2^20, 1 <= d <= 4096 through a float
// reciprocal: (e + 1/2) / d lies at least 1/(2d) away from an integer, far
// beyond the float rounding error, so the truncation is exact
struct FDiv {
    int d;
    float inv;
    __device__ __forceinline__ explicit FDiv(int d_) : d(d_), inv(1.0f / (float)d_) {}
    __device__ __forceinline__ int quo(int e) const { return (int)(((float)e + 0.5f) * inv); }
};

// coefficient j of order n from a staged (NO x NO) table; 0 for j > n
template <int NO>
__device__ __forceinline__ float coefs(const float *tab, int j, int n)
{
    const int nn = n < 1 ? 1 : n;
    const float c = tab[(j - 1) + NO * (nn - 1)];
    return j <= n ? c : 0.0f;
}

// FULL: every face, collocated derivative and cell of the tile carries the
// full order NO (the tile and its reach stay NO cells from the grid's edges
// and from the surface row), so the two face fluxes of each face coincide
// (n1 = n2 = NO: fx2 = fx1, fz2 = fz1, bit for bit) and the coefficients are
// block constants held in registers
template <int NO, bool FULL>
__device__ __forceinline__ void frame_tile(const StepArgs &a, const ETiles &et, long bid)
{
    extern __shared__ float sm[];   // frame_lds_floats(tx, tz, NO) of the largest tile of the launch
    const Geo &g = a.g;
    const Coefs *cf = a.cf;
    constexpr int H = 2 * NO - 1;
    int q = 0;
    while (q + 1 < et.n && bid >= et.off[q + 1]) ++q;
    const int loc = (int)(bid - et.off[q]);
    const int bx = loc / et.nbz[q], bz = loc % et.nbz[q];
    const int cx0 = et.x0[q] + bx * et.tx[q], cx1 = min(cx0 + et.tx[q] - 1, et.x1[q]);
    const int cz0 = et.z0[q] + bz * et.tz[q], cz1 = min(cz0 + et.tz[q] - 1, et.z1[q]);
    const int view = et.view[q];
    const int tx = cx1 - cx0 + 1, tz = cz1 - cz0 + 1;
    const int SX = tx + 2 * H, SZ = tz + 2 * H, lx0 = cx0 - H, lz0 = cz0 - H;
    const int NFX = tx + 2 * NO - 1, fx0 = cx0 - NO;   // x faces f..f+1, f in fx0..fx0+NFX-1
    const int NFZ = tz + 2 * NO - 1, fz0 = cz0 - NO;   // z faces
    const int NDZ = tx + 2 * NO, dz0 = cx0 - NO;       // DZ columns
    const int NDX = tz + 2 * NO, dx0 = cz0 - NO;       // DX rows
    float *sux = sm, *suz = sux + SX * SZ, *slam = suz + SX * SZ, *smu = slam + SX * SZ;
    float *fx1x = smu + SX * SZ, *fx1z = fx1x + NFX * tz, *fx2x = fx1z + NFX * tz, *fx2z = fx2x + NFX * tz;
    float *fz1x = fx2z + NFX * tz, *fz1z = fz1x + tx * NFZ, *fz2x = fz1z + tx * NFZ, *fz2z = fz2x + tx * NFZ;
    float *dzx = fz2z + tx * NFZ, *dzz = dzx + NDZ * tz;
    float *dxx = dzz + NDZ * tz, *dxz = dxx + tx * NDX;
    float *mpx0 = dxz + tx * NDX, *mpx1 = mpx0 + NDZ * tz;   // M psi_x on the DZ index space
    float *mpz0 = mpx1 + NDZ * tz, *mpz1 = mpz0 + tx * NDX;  // M psi_z on the DX index space
    if constexpr (FULL) { fx2x = fx1x; fx2z = fx1z; fz2x = fz1x; fz2z = fz1z; }
    // the band's PML strip (uniform per band): none in the total-field region
    const int sb = view == 0 ? S_NONE : (cz0 > g.nz ? S_B : (cx1 <= 0 ? S_L : S_R));
    // the two coefficient tables, staged once per block
    __shared__ float c0s[NO * NO], c1s[NO * NO];
    if (threadIdx.x < NO * NO) {
        c0s[threadIdx.x] = cf->c0f[threadIdx.x];
        c1s[threadIdx.x] = cf->c1f[threadIdx.x];
    }
    float k0[NO], k1[NO];
#pragma unroll
    for (int j = 0; j < NO; ++j) {
        k0[j] = FULL ? cf->c0f[j + NO * (NO - 1)] : 0.0f;
        k1[j] = FULL ? cf->c1f[j + NO * (NO - 1)] : 0.0f;
    }
    // coefficient j of order n (n = NO throughout a FULL tile)
    auto C0 = [&](int j, int n) { if constexpr (FULL) return k0[j - 1]; else return coefs<NO>(c0s, j, n); };
    auto C1 = [&](int j, int n) { if constexpr (FULL) return k1[j - 1]; else return coefs<NO>(c1s, j, n); };
    const FDiv dSZ(SZ), dtz(tz), dNFZ(NFZ), dNDX(NDX);
    // 1/dx and 1/dz are folded into the collocated derivatives and the
    // memory corrections (products instead of divisions)
    const float rdx = 1.0f / g.dx, rdz = 1.0f / g.dz;
    auto S = [&](int z, int x) { return (x - lx0) * SZ + (z - lz0); };
    const int nxp = g.nx + g.npml, NZ = g.NZ;
    // Global loads whose addresses do not depend on LDS are issued first so
    // that they overlap the staging and phase 1: the raw psi around the tile
    // (the first PF entries of each thread) and the update inputs of the
    // thread's own cell.
    constexpr int PF = 3;
    const Strips *st = a.st;
    float rpx[PF][2], rpz[PF][2];
    if (sb != S_NONE) {
        const float *ps0 = st->psx[sb][0], *ps1 = st->psx[sb][1];
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int e = threadIdx.x + i * kFrameThreads;
            const int qq = dtz.quo(e);
            const int x = dz0 + qq, z = cz0 + (e - qq * tz);
            rpx[i][0] = 0.0f; rpx[i][1] = 0.0f;
            if (e < NDZ * tz && strip_of(g, z, x) == sb && x >= 1 - g.npml && x <= nxp && z <= NZ) {
                const size_t o = strip_off(g, sb, z, x);
                rpx[i][0] = ps0[o]; rpx[i][1] = ps1[o];
            }
        }
        if (sb == S_B) {
            const float *pz0 = st->psz[S_B][0], *pz1 = st->psz[S_B][1];
#pragma unroll
            for (int i = 0; i < PF; ++i) {
                const int e = threadIdx.x + i * kFrameThreads;
                const int qq = dNDX.quo(e);
                const int x = cx0 + qq, z = dx0 + (e - qq * NDX);
                rpz[i][0] = 0.0f; rpz[i][1] = 0.0f;
                if (e < tx * NDX && z > g.nz && z <= NZ) {
                    const size_t o = strip_off(g, S_B, z, x);
                    rpz[i][0] = pz0[o]; rpz[i][1] = pz1[o];
                }
            }
        }
    }
    const int t = threadIdx.x;
    const bool own = t < tx * tz;
    const int tq = dtz.quo(t);
    const int ix = cx0 + tq, iz = cz0 + (t - tq * tz);
    float ubf = 0.0f, uax = 0.0f, uaz = 0.0f, ucur[2][3] = {}, uprv[2][3] = {};
    size_t uc = 0, uo = 0;
    if (own) {
        uc = cell(g, iz, ix);
        ubf = a.bf[uc];
        if (sb == S_NONE) {
            ucur[0][0] = a.ux[uc]; ucur[1][0] = a.uz[uc];
            uprv[0][0] = a.uxp[uc]; uprv[1][0] = a.uzp[uc];
        } else {
            uo = strip_off(g, sb, iz, ix);
            const float fdt = (float)g.dt;
            uax = 0.5f * damp_x(st, ix + g.npml - 1, iz - 1) * fdt;
            uaz = 0.5f * damp_z(st, ix + g.npml - 1, iz - 1) * fdt;
#pragma unroll
            for (int comp = 0; comp < 2; ++comp)
#pragma unroll
                for (int q2 = 0; q2 < 3; ++q2) {
                    ucur[comp][q2] = st->cur[sb][comp][q2][uo];
                    uprv[comp][q2] = st->prev[sb][comp][q2][uo];
                }
        }
    }
    // stage u (in the tile's view), lambda, mu; four elements per thread
    // per pass so that their loads are in flight together
    constexpr int SU = 4;
    for (int e0 = threadIdx.x; e0 < SX * SZ; e0 += SU * kFrameThreads) {
        float vx[SU], vz[SU], vl[SU], vm[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int e = e0 + u * kFrameThreads;
            const int q = dSZ.quo(e);
            const int x = lx0 + q, z = lz0 + (e - q * SZ);
            vx[u] = 0.0f; vz[u] = 0.0f; vl[u] = 0.0f; vm[u] = 0.0f;
            if (e < SX * SZ && x >= 1 - g.npml && x <= g.nx + g.npml && z >= 1 && z <= g.NZ) {
                const size_t c = cell(g, z, x);
                vx[u] = a.ux[c]; vz[u] = a.uz[c]; vl[u] = a.lam[c]; vm[u] = a.mu[c];
                if (a.ring_valid) {
                    const int side = tf_side(g, z, x);
                    if (view == 1 && side == 0) { vx[u] -= a.ringx[c]; vz[u] -= a.ringz[c]; }
                    else if (view == 0 && side == 1) { vx[u] += a.ringx[c]; vz[u] += a.ringz[c]; }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int e = e0 + u * kFrameThreads;
            if (e < SX * SZ) { sux[e] = vx[u]; suz[e] = vz[u]; slam[e] = vl[u]; smu[e] = vm[u]; }
        }
    }
    __syncthreads();
    // phase 1a: x-fluxes (mode 0 on ux, mode 1 on uz)
    for (int e = threadIdx.x; e < NFX * tz; e += kFrameThreads) {
        const int q = dtz.quo(e);
        const int f = fx0 + q, z = cz0 + (e - q * tz);
        const int n1 = imin(imin(f + g.npml - 1, nxp - f), NO), n2 = imin(imin(f + g.npml, nxp - f - 1), NO);
        const int a0 = S(z, f), a1 = S(z, f + 1);
        const float ha = hmean(slam[a1] + 2.0f * smu[a1], slam[a0] + 2.0f * smu[a0]);
        const float hm = hmean(smu[a1], smu[a0]);
        float s1x = 0.0f, s2x = 0.0f, s1z = 0.0f, s2z = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            const int p = S(z, f + j), m = S(z, f + 1 - j);
            const float dux = sux[p] - sux[m], duz = suz[p] - suz[m];
            const float c1 = C0(j, n1);
            s1x = accf(s1x, c1, dux);
            s1z = accf(s1z, c1, duz);
            if constexpr (!FULL) {
                const float c2 = C0(j, n2);
                s2x = accf(s2x, c2, dux);
                s2z = accf(s2z, c2, duz);
            }
        }
        fx1x[e] = ha * s1x;
        fx1z[e] = hm * s1z;
        if constexpr (!FULL) { fx2x[e] = ha * s2x; fx2z[e] = hm * s2z; }
    }
    // phase 1b: z-fluxes (mode 1 on ux, mode 0 on uz)
    for (int e = threadIdx.x; e < tx * NFZ; e += kFrameThreads) {
        const int q = dNFZ.quo(e);
        const int x = cx0 + q, f = fz0 + (e - q * NFZ);
        const int n1 = imin(imin(f - 1, NZ - f), NO), n2 = imin(imin(f, NZ - f - 1), NO);
        const int a0 = S(f, x), a1 = S(f + 1, x);
        const float ha = hmean(slam[a1] + 2.0f * smu[a1], slam[a0] + 2.0f * smu[a0]);
        const float hm = hmean(smu[a1], smu[a0]);
        float s1x = 0.0f, s2x = 0.0f, s1z = 0.0f, s2z = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            const int p = S(f + j, x), m = S(f + 1 - j, x);
            const float dux = sux[p] - sux[m], duz = suz[p] - suz[m];
            const float c1 = C0(j, n1);
            s1x = accf(s1x, c1, dux);
            s1z = accf(s1z, c1, duz);
            if constexpr (!FULL) {
                const float c2 = C0(j, n2);
                s2x = accf(s2x, c2, dux);
                s2z = accf(s2z, c2, duz);
            }
        }
        fz1x[e] = hm * s1x;
        fz1z[e] = ha * s1z;
        if constexpr (!FULL) { fz2x[e] = hm * s2x; fz2z[e] = ha * s2z; }
    }
    // phase 1c: collocated Dz (cdz_sum) at columns dz0.., rows of the tile
    for (int e = threadIdx.x; e < NDZ * tz; e += kFrameThreads) {
        const int q = dtz.quo(e);
        const int x = dz0 + q, z = cz0 + (e - q * tz);
        const int n = imin(imin(z - 1, NZ - z), NO);
        float sx = 0.0f, sz = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            const int p = S(z + j, x), m = S(z - j, x);
            const float c = C1(j, n);
            sx = accf(sx, c, sux[p] - sux[m]);
            sz = accf(sz, c, suz[p] - suz[m]);
        }
        dzx[e] = sx * rdz; dzz[e] = sz * rdz;
    }
    // phase 1d: collocated Dx (cdx_sum) at the tile columns, rows dx0..
    for (int e = threadIdx.x; e < tx * NDX; e += kFrameThreads) {
        const int q = dNDX.quo(e);
        const int x = cx0 + q, z = dx0 + (e - q * NDX);
        const int n = imin(imin(x + g.npml - 1, nxp - x), NO);
        float sx = 0.0f, sz = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            const int p = S(z, x + j), m = S(z, x - j);
            const float c = C1(j, n);
            sx = accf(sx, c, sux[p] - sux[m]);
            sz = accf(sz, c, suz[p] - suz[m]);
        }
        dxx[e] = sx * rdx; dxz[e] = sz * rdx;
    }
    // phase 1e: M psi of this step (k_prep) around the tile, as the memory
    // corrections read them: zero outside the strips, psi_z only in the
    // bottom strip
    if (sb != S_NONE) {
        // around a band's tile a cell is either in the band's strip or in
        // the total-field region (psi 0): one strip, uniform pointers
        const float *ps0 = st->psx[sb][0], *ps1 = st->psx[sb][1];
        for (int i = 0, e = threadIdx.x; e < NDZ * tz; ++i, e += kFrameThreads) {
            const int qq = dtz.quo(e);
            const int x = dz0 + qq, z = cz0 + (e - qq * tz);
            float r0 = 0.0f, r1 = 0.0f;
            if (i < PF) {
#pragma unroll
                for (int j = 0; j < PF; ++j)
                    if (j == i) { r0 = rpx[j][0]; r1 = rpx[j][1]; }
            } else if (strip_of(g, z, x) == sb && x >= 1 - g.npml && x <= nxp && z <= NZ) {
                const size_t o = strip_off(g, sb, z, x);
                r0 = ps0[o]; r1 = ps1[o];
            }
            const int k = S(z, x);
            mpx0[e] = (slam[k] + 2.0f * smu[k]) * r0;
            mpx1[e] = smu[k] * r1;
        }
        if (sb == S_B)
            for (int i = 0, e = threadIdx.x; e < tx * NDX; ++i, e += kFrameThreads) {
                const int qq = dNDX.quo(e);
                const int x = cx0 + qq, z = dx0 + (e - qq * NDX);
                float r0 = 0.0f, r1 = 0.0f;
                if (i < PF) {
#pragma unroll
                    for (int j = 0; j < PF; ++j)
                        if (j == i) { r0 = rpz[j][0]; r1 = rpz[j][1]; }
                } else if (z > g.nz && z <= NZ) {
                    const size_t o = strip_off(g, S_B, z, x);
                    r0 = st->psz[S_B][0][o]; r1 = st->psz[S_B][1][o];
                }
                const int k = S(z, x);
                mpz0[e] = smu[k] * r0;
                mpz1[e] = (slam[k] + 2.0f * smu[k]) * r1;
            }
    }
    __syncthreads();
    // phase 2: one thread per cell
    if (!own) return;
    const int nox = imin(imin(ix + g.npml - 1, nxp - ix), NO);
    const int noz = imin(imin(iz - 1, NZ - iz), NO);
    auto FX = [&](const float *b, int f) { return b[(f - fx0) * tz + (iz - cz0)]; };
    auto FZ = [&](const float *b, int f) { return b[(ix - cx0) * NFZ + (f - fz0)]; };
    auto DZ = [&](const float *b, int x) { return b[(x - dz0) * tz + (iz - cz0)]; };
    auto DX = [&](const float *b, int z) { return b[(ix - cx0) * NDX + (z - dx0)]; };
    const bool in_reg = (iz >= 1 && iz <= g.nz && ix >= 1 && ix <= g.nx);
    const float dx2 = in_reg ? (float)(1.0 / (double)(g.dx * g.dx)) : 1.0f / (g.dx * g.dx);
    float Xx = 0.0f, Xz = 0.0f, Zx = 0.0f, Zz = 0.0f, zxl = 0.0f, zxm = 0.0f, xzm = 0.0f, xzl = 0.0f;
#pragma unroll
    for (int k = 1; k <= NO; ++k) {
        const float c0x = C0(k, nox), c0z = C0(k, noz);
        const float c1x = C1(k, nox), c1z = C1(k, noz);
        Xx = accf(Xx, c0x, FX(fx1x, ix + k - 1) - FX(fx2x, ix - k));
        Xz = accf(Xz, c0x, FX(fx1z, ix + k - 1) - FX(fx2z, ix - k));
        Zx = accf(Zx, c0z, FZ(fz1x, iz + k - 1) - FZ(fz2x, iz - k));
        Zz = accf(Zz, c0z, FZ(fz1z, iz + k - 1) - FZ(fz2z, iz - k));
        const int xp = S(iz, ix + k), xm = S(iz, ix - k), zp = S(iz + k, ix), zm = S(iz - k, ix);
        // Dx(l Dz uz), Dx(m Dz ux), Dz(m Dx uz), Dz(l Dx ux)
        zxl = accf(zxl, c1x, slam[xp] * DZ(dzz, ix + k) - slam[xm] * DZ(dzz, ix - k));
        zxm = accf(zxm, c1x, smu[xp] * DZ(dzx, ix + k) - smu[xm] * DZ(dzx, ix - k));
        xzm = accf(xzm, c1z, smu[zp] * DX(dxz, iz + k) - smu[zm] * DX(dxz, iz - k));
        xzl = accf(xzl, c1z, slam[zp] * DX(dxx, iz + k) - slam[zm] * DX(dxx, iz - k));
    }
    float p[6];
    p[0] = Xx * dx2;
    p[3] = Xz * dx2;
    if (iz == 1) {
        // free-surface row: half-cell closure derived from the strain energy
        // of the top layer (rows 1-2, trapezoidal in z); traction free naturally
        const int r1 = S(1, ix), r2 = S(2, ix);
        const float sdz2 = 2.0f / (g.dz * g.dz);
        p[1] = (hmean(smu[r1], smu[r2]) * (sux[r2] - sux[r1])) * sdz2;
        p[4] = (hmean(slam[r1] + 2.0f * smu[r1], slam[r2] + 2.0f * smu[r2]) * (suz[r2] - suz[r1])) * sdz2;
        float sa = 0.0f, sb = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            const int p1 = S(1, ix + k), p2 = S(2, ix + k), m1 = S(1, ix - k), m2 = S(2, ix - k);
            const float c = C1(k, nox);
            const float lp = 0.5f * (slam[p1] + slam[p2]), lm = 0.5f * (slam[m1] + slam[m2]);
            const float mp = 0.5f * (smu[p1] + smu[p2]), mm = 0.5f * (smu[m1] + smu[m2]);
            sa = accf(sa, c, lp * (suz[p2] - suz[p1]) - lm * (suz[m2] - suz[m1]));
            sb = accf(sb, c, mp * (sux[p2] - sux[p1]) - mm * (sux[m2] - sux[m1]));
        }
        const float lbar = 0.5f * (slam[r1] + slam[r2]), mbar = 0.5f * (smu[r1] + smu[r2]);
        const float dxuz = DX(dxz, 1) + DX(dxz, 2);
        const float dxux = DX(dxx, 1) + DX(dxx, 2);
        p[2] = (sa * rdx + mbar * dxuz) * rdz;
        p[5] = (sb * rdx + lbar * dxux) * rdz;
    } else {
        p[1] = Zx * (1.0f / (g.dz * g.dz));
        p[4] = Zz * (1.0f / (g.dz * g.dz));
        p[2] = zxl * rdx + xzm * rdz;
        p[5] = zxm * rdx + xzl * rdz;
    }
    // memory corrections -Dx(M psi_x), -Dz(M psi_z) (the bottom strip only)
    if (sb != S_NONE) {
        float sx0 = 0.0f, sx1 = 0.0f, sz0 = 0.0f, sz1 = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            const float c = C1(k, nox);
            sx0 = accf(sx0, c, DZ(mpx0, ix + k) - DZ(mpx0, ix - k));
            sx1 = accf(sx1, c, DZ(mpx1, ix + k) - DZ(mpx1, ix - k));
        }
        p[0] = p[0] - sx0 * rdx;
        p[3] = p[3] - sx1 * rdx;
        if (sb == S_B) {
#pragma unroll
            for (int k = 1; k <= NO; ++k) {
                const float c = C1(k, noz);
                sz0 = accf(sz0, c, DX(mpz0, iz + k) - DX(mpz0, iz - k));
                sz1 = accf(sz1, c, DX(mpz1, iz + k) - DX(mpz1, iz - k));
            }
            p[1] = p[1] - sz0 * rdz;
            p[4] = p[4] - sz1 * rdz;
        }
    }
    // update: plain leapfrog in the total-field region, split parts in a strip
    if (sb == S_NONE) {
        const float rxs = (p[0] + p[1]) + p[2], rzs = (p[3] + p[4]) + p[5];
        a.uxo[uc] = (2.0f * ucur[0][0] - uprv[0][0]) + ubf * rxs;
        a.uzo[uc] = (2.0f * ucur[1][0] - uprv[1][0]) + ubf * rzs;
        return;
    }
    const float ax = uax, az = uaz;
    const float c3 = 2.0f - 4.0f * ax * az;
    const float r1 = 1.0f / (1.0f + ax), r2 = 1.0f / (1.0f + az), r3 = 1.0f / (1.0f + ax + az);
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        const float X = p[3 * comp], Z = p[3 * comp + 1], Mm = p[3 * comp + 2];
        const float n1 = (2.0f * ucur[comp][0] - (1.0f - ax) * uprv[comp][0] + ubf * X) * r1;
        const float n2 = (2.0f * ucur[comp][1] - (1.0f - az) * uprv[comp][1] + ubf * Z) * r2;
        const float n3 = (c3 * ucur[comp][2] - (1.0f - ax - az) * uprv[comp][2] + ubf * Mm) * r3;
        st->prev[sb][comp][0][uo] = n1; st->prev[sb][comp][1][uo] = n2; st->prev[sb][comp][2][uo] = n3;
        const float tot = (n1 + n2) + n3;
        if (comp == 0) a.uxo[uc] = tot; else a.uzo[uc] = tot;
    }
}

// A tile is FULL when the faces and derivatives it evaluates (its cells
// plus NO on each side, the faces of its reach 2NO) keep the full order:
// x within [2NO+1-npml, nx+npml-2NO], z within [2NO+1, NZ-2NO]
template <int NO>
__device__ __forceinline__ void frame_body(const StepArgs &a, const ETiles &et, long bid)
{
    const Geo &g = a.g;
    int q = 0;
    while (q + 1 < et.n && bid >= et.off[q + 1]) ++q;
    const int loc = (int)(bid - et.off[q]);
    const int bx = loc / et.nbz[q], bz = loc % et.nbz[q];
    const int cx0 = et.x0[q] + bx * et.tx[q], cx1 = min(cx0 + et.tx[q] - 1, et.x1[q]);
    const int cz0 = et.z0[q] + bz * et.tz[q], cz1 = min(cz0 + et.tz[q] - 1, et.z1[q]);
    const bool full = a.frame_full && cx0 >= 2 * NO + 1 - g.npml && cx1 <= g.nx + g.npml - 2 * NO && cz0 >= 2 * NO + 1 &&
                      cz1 <= g.NZ - 2 * NO;
    if (full)
        frame_tile<NO, true>(a, et, bid);
    else
        frame_tile<NO, false>(a, et, bid);
}

template <int NO>
__global__ __launch_bounds__(kFrameThreads) void k_frame(StepArgs a, ETiles t)
{
    frame_body<NO>(a, t, xcd_tile(blockIdx.x, gridDim.x));
}

// The interior and the frame of one step in one launch: blocks
// [0, nint) are interior tiles, the rest frame tiles.  Both read level n
// only and write disjoint cells of level n+1, so any interleaving is valid.
// Workgroups are dispatched in id order, so the interior tiles take every
// resident slot first and the frame tiles fill the slots that interior
// tiles release at the end of their march, instead of running after the
// slowest of them.  Dynamic LDS: the larger of the two needs.
static_assert(kFrameThreads == 256, "the fused step launches 256-lane blocks");
template <int NO, int M16 = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kIntWaves<NO>))) void k_step(IntArgs<NO> ia, StepArgs sa,
                                                                                      ETiles ft, int nint)
{
    extern __shared__ float sm[];
    if ((int)blockIdx.x < nint)
        IntKernel2<NO, 256, M16>::run(ia, xcd_tile(blockIdx.x, nint),
                                      *reinterpret_cast<typename IntKernel2<NO, 256, M16>::Smem *>(sm));
    else   // frame tiles grouped per XCD too: neighbouring tiles share halo lines in one L2
        frame_body<NO>(sa, ft, xcd_tile((int)blockIdx.x - nint, (int)gridDim.x - nint));
}

}  // namespace fdfk
