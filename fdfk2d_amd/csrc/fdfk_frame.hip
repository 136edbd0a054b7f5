// Full-order padding cells (the PML strips beyond the hybrid interface):
// a marching kernel with the interior kernel's shared-flux structure (lanes
// along z, one column per step along x, register rings for x-neighbours,
// LDS for z-neighbours) plus what the strips need:
//   * scattered-field view: every read of a cell inside the total-field
//     region subtracts the background u0 there (the reads of an outside
//     hybrid-ring cell, src/FDFK2D.f90:516-522 / libHybrid);
//   * the stretched-coordinate memory corrections -Dx(M psi_x) and, in the
//     bottom strip, -Dz(M psi_z) with the psi of this step (k_psi);
//   * the split-field leapfrog update of the three parts per component.
// The reduced-order cells next to the padded-grid boundary and the surface
// rows stay with k_frame_edge.
//
// Every output cell of a tile is outside the total-field region, so one
// view serves the whole tile.

namespace fdfk {

template <int NO, int NT>
struct OutKernel {
    static constexpr int HZ = 2 * NO - 1;   // z halo lanes on each side
    static constexpr int TZ = NT - 2 * HZ;  // output rows per block
    static constexpr int P0 = 3 * NO + 1;
    static constexpr int P = P0 + (P0 & 1); // even ring period: LDS parity is compile-time
    static constexpr int PAD = NO + 1;      // LDS guard entries on each side
    static constexpr int WARM = 4 * NO - 2; // warm-up columns before the first output
    static constexpr int LW = NT + 2 * PAD;

    struct Smem {                 // [parity][comp][lane]
        float uA[2][2][LW];       // u at column ix+NO
        float uB[2][2][LW];       // u at column ix
        float rM[2][2][LW];       // 1/(l+2m), 1/m at column ix
        float fz[2][2][LW];       // z-fluxes at column ix
        float vz[2][2][LW];       // m*Dx uz, l*Dx ux at column ix
        float pz[2][2][LW];       // m*psi_z(ux), (l+2m)*psi_z(uz) at column ix (bottom strip, else 0)
    };

    // per-lane constants: row, masks, and the lane's strip (the side strip
    // above the bottom strip, the bottom strip below it) with its arrays
    struct Lane {
        int lane, r;
        unsigned ofs;    // clamped row byte offset inside a grid column
        bool outrow, isB;
        float az;
        int cb;          // first column of the lane's strip
        int stride;      // strip column pitch (nz or npml)
        int rb;          // row offset inside a strip column
        const float *psx[2], *psz[2];
        float *cur[2][3], *prev[2][3];
        __device__ __forceinline__ size_t off(int col) const
        {
            return (size_t)(col - cb) * (size_t)stride + (size_t)rb;
        }
    };

    struct Regs {
        float U[2][P], FX[2][P], VX[2][P], PX[2][P];
        float rMa_prev, rMm_prev;
        float pu[2][2];   // [parity][comp] next new u column (ix+2NO-1 of the next step)
        float pl[2][2];   // [parity][l,m] next lead materials (ix+NO)
        float p0[2][2];   // [parity][l,m] next column materials (ix)
        float pq[2][2];   // [parity][comp] next psi_x at (r, ix+NO)
        float pr[2][2];   // [parity][comp] next psi_z at (r, ix)
        float ps[2][2];   // [parity][comp] z memory sums of column ix (used one step later)
    };

    static __device__ __forceinline__ int sl(int ph, int j) { return ((ph + j) % P + P) % P; }

    static __device__ __forceinline__ __amdgpu_buffer_rsrc_t colrsrc(const Geo &g, const float *base, int c)
    {
        const float *cb = base + (size_t)(c + g.npml - 1) * (size_t)g.ld;
        return __builtin_amdgcn_make_buffer_rsrc((void *)cb, 0, g.ld * 4, kBufferFlags);
    }
    static __device__ __forceinline__ float at(const Geo &g, const float *base, int c, unsigned o)
    {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(colrsrc(g, base, c), o, 0, 0));
    }
    static __device__ __forceinline__ void put(const Geo &g, float *base, int c, unsigned o, float v)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), colrsrc(g, base, c), o, 0, 0);
    }

    // the loads one step consumes, issued one step ahead (column ix)
    static __device__ __forceinline__ void prefetch(const FrameArgs<NO> &a, const Lane &L, Regs &R, int q, int ix,
                                                    int xs, int xe)
    {
        const Geo &g = a.g;
        const int lo = 1 - g.npml, hi = g.nx + g.npml;
        // columns of the first warm-up and the last (unused) prefetch may
        // lie beyond the grid; their values are never used, so the grid
        // addresses are clamped and the strip reads skipped
        auto cc = [&](int c) { return c < lo ? lo : (c > hi ? hi : c); };
        auto inside = [&](int c) { return c >= 1 && c <= g.nx && L.r <= g.nz; };
        const int cn = cc(ix + 2 * NO - 1), cm = ix + NO;
        float vx = at(g, a.ux, cn, L.ofs), vz = at(g, a.uz, cn, L.ofs);
        if (a.corr && inside(cn)) { vx -= at(g, a.ringx, cn, L.ofs); vz -= at(g, a.ringz, cn, L.ofs); }
        R.pu[q][0] = vx; R.pu[q][1] = vz;
        R.pl[q][0] = at(g, a.lam, cc(cm), L.ofs); R.pl[q][1] = at(g, a.mu, cc(cm), L.ofs);
        R.p0[q][0] = at(g, a.lam, cc(ix), L.ofs); R.p0[q][1] = at(g, a.mu, cc(ix), L.ofs);
        float q0 = 0.0f, q1 = 0.0f, z0 = 0.0f, z1 = 0.0f;
        if (L.outrow && cm >= xs - NO && cm < xe + NO && !inside(cm)) {
            const size_t o = L.off(cm);
            q0 = L.psx[0][o]; q1 = L.psx[1][o];
        }
        if (L.isB && ix >= xs && ix < xe) {
            const size_t o = L.off(ix);
            z0 = L.psz[0][o]; z1 = L.psz[1][o];
        }
        R.pq[q][0] = q0; R.pq[q][1] = q1;
        R.pr[q][0] = z0; R.pr[q][1] = z1;
    }

    template <int PH>
    static __device__ __forceinline__ void step(const FrameArgs<NO> &a, Smem &S, Regs &R, const Lane &L, int ix,
                                                int xs, int xe)
    {
        constexpr int PA = PH & 1, PB = 1 - PA;
        const Geo &g = a.g;
        const int lane = L.lane;
        // A. consume this step's prefetch, prefetch the next step, load the
        //    update inputs of the output column ix-1
        R.U[0][sl(PH, 2 * NO - 1)] = R.pu[PA][0];
        R.U[1][sl(PH, 2 * NO - 1)] = R.pu[PA][1];
        const float l1 = R.pl[PA][0], m1 = R.pl[PA][1];
        const float l0 = R.p0[PA][0], m0 = R.p0[PA][1];
        const float q0 = R.pq[PA][0], q1 = R.pq[PA][1];
        const float zr0 = R.pr[PA][0], zr1 = R.pr[PA][1];
        prefetch(a, L, R, PB, ix + 1, xs, xe);
        const int co = ix - 1;
        const bool out = L.outrow && co >= xs && co < xe;
        float bf = 0.0f, ax = 0.0f, sc[2][3] = {}, sp[2][3] = {};
        if (out) {
            const size_t o = L.off(co);
            bf = at(g, a.bf, co, L.ofs);
            ax = a.st->ax[co + g.npml - 1];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int p = 0; p < 3; ++p) { sc[c][p] = L.cur[c][p][o]; sp[c][p] = L.prev[c][p][o]; }
        }
        // x-flux at the face (ix+NO-1, ix+NO), M psi_x at ix+NO, LDS staging
        const float rMa = frcp(l1 + 2.0f * m1), rMm = frcp(m1);
        {
            const float ha = frcp(R.rMa_prev + rMa), hm = frcp(R.rMm_prev + rMm);
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                s0 += a.k.ai[j - 1] * (R.U[0][sl(PH, NO - 1 + j)] - R.U[0][sl(PH, NO - j)]);
                s1 += a.k.ai[j - 1] * (R.U[1][sl(PH, NO - 1 + j)] - R.U[1][sl(PH, NO - j)]);
            }
            R.FX[0][sl(PH, NO - 1)] = ha * s0;
            R.FX[1][sl(PH, NO - 1)] = hm * s1;
        }
        R.rMa_prev = rMa;
        R.rMm_prev = rMm;
        R.PX[0][sl(PH, NO)] = (l1 + 2.0f * m1) * q0;
        R.PX[1][sl(PH, NO)] = m1 * q1;
        S.uA[PA][0][PAD + lane] = R.U[0][sl(PH, NO)];
        S.uA[PA][1][PAD + lane] = R.U[1][sl(PH, NO)];
        S.uB[PA][0][PAD + lane] = R.U[0][sl(PH, 0)];
        S.uB[PA][1][PAD + lane] = R.U[1][sl(PH, 0)];
        S.rM[PA][0][PAD + lane] = frcp(l0 + 2.0f * m0);
        S.rM[PA][1][PAD + lane] = frcp(m0);
        S.pz[PA][0][PAD + lane] = m0 * zr0;
        S.pz[PA][1][PAD + lane] = (l0 + 2.0f * m0) * zr1;
        // B.
        lds_barrier();
        // C. Vx at ix+NO; z-fluxes, Vz and the z memory sums at ix
        {
            float dzz = 0.0f, dzx = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                dzz += a.k.bi[j - 1] * (S.uA[PA][1][PAD + lane + j] - S.uA[PA][1][PAD + lane - j]);
                dzx += a.k.bi[j - 1] * (S.uA[PA][0][PAD + lane + j] - S.uA[PA][0][PAD + lane - j]);
            }
            R.VX[0][sl(PH, NO)] = l1 * dzz;
            R.VX[1][sl(PH, NO)] = m1 * dzx;
        }
        {
            float sfx = 0.0f, sfz = 0.0f, pzx = 0.0f, pzz = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                sfx += a.k.ai[j - 1] * (S.uB[PA][0][PAD + lane + j] - S.uB[PA][0][PAD + lane + 1 - j]);
                sfz += a.k.ai[j - 1] * (S.uB[PA][1][PAD + lane + j] - S.uB[PA][1][PAD + lane + 1 - j]);
                pzx += a.k.bi[j - 1] * (S.pz[PA][0][PAD + lane + j] - S.pz[PA][0][PAD + lane - j]);
                pzz += a.k.bi[j - 1] * (S.pz[PA][1][PAD + lane + j] - S.pz[PA][1][PAD + lane - j]);
            }
            const float rMa0 = S.rM[PA][0][PAD + lane], rMm0 = S.rM[PA][1][PAD + lane];
            const float fzx = sfx * frcp(rMm0 + S.rM[PA][1][PAD + lane + 1]);
            const float fzz = sfz * frcp(rMa0 + S.rM[PA][0][PAD + lane + 1]);
            float dxz = 0.0f, dxx = 0.0f;
#pragma unroll
            for (int k = 1; k <= NO; ++k) {
                dxz += a.k.bi[k - 1] * (R.U[1][sl(PH, k)] - R.U[1][sl(PH, -k)]);
                dxx += a.k.bi[k - 1] * (R.U[0][sl(PH, k)] - R.U[0][sl(PH, -k)]);
            }
            S.fz[PA][0][PAD + lane] = fzx;
            S.fz[PA][1][PAD + lane] = fzz;
            S.vz[PA][0][PAD + lane] = m0 * dxz;
            S.vz[PA][1][PAD + lane] = l0 * dxx;
            R.ps[PA][0] = pzx;
            R.ps[PA][1] = pzz;
        }
        // D. column ix-1: six parts, memory corrections, split update
        if (!out) return;
        float Xx = 0.0f, Xz = 0.0f, Zx = 0.0f, Zz = 0.0f, Mx = 0.0f, Mz = 0.0f, sx0 = 0.0f, sx1 = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            Xx += a.k.aox[k - 1] * (R.FX[0][sl(PH, k - 2)] - R.FX[0][sl(PH, -k - 1)]);
            Xz += a.k.aox[k - 1] * (R.FX[1][sl(PH, k - 2)] - R.FX[1][sl(PH, -k - 1)]);
            Zx += a.k.aoz[k - 1] * (S.fz[PB][0][PAD + lane + k - 1] - S.fz[PB][0][PAD + lane - k]);
            Zz += a.k.aoz[k - 1] * (S.fz[PB][1][PAD + lane + k - 1] - S.fz[PB][1][PAD + lane - k]);
            Mx += a.k.bo[k - 1] * ((R.VX[0][sl(PH, k - 1)] - R.VX[0][sl(PH, -k - 1)]) +
                                  (S.vz[PB][0][PAD + lane + k] - S.vz[PB][0][PAD + lane - k]));
            Mz += a.k.bo[k - 1] * ((R.VX[1][sl(PH, k - 1)] - R.VX[1][sl(PH, -k - 1)]) +
                                  (S.vz[PB][1][PAD + lane + k] - S.vz[PB][1][PAD + lane - k]));
            sx0 += a.k.bi[k - 1] * (R.PX[0][sl(PH, k - 1)] - R.PX[0][sl(PH, -k - 1)]);
            sx1 += a.k.bi[k - 1] * (R.PX[1][sl(PH, k - 1)] - R.PX[1][sl(PH, -k - 1)]);
        }
        Xx = Xx - sx0 / g.dx;
        Xz = Xz - sx1 / g.dx;
        if (L.isB) {
            Zx = Zx - R.ps[PB][0] / g.dz;
            Zz = Zz - R.ps[PB][1] / g.dz;
        }
        const float az = L.az;
        const float c3 = 2.0f - 4.0f * ax * az;
        const float pa[2][3] = {{Xx, Zx, Mx}, {Xz, Zz, Mz}};
        const size_t o = L.off(co);
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            float n1 = (2.0f * sc[comp][0] - (1.0f - ax) * sp[comp][0] + bf * pa[comp][0]) / (1.0f + ax);
            float n2 = (2.0f * sc[comp][1] - (1.0f - az) * sp[comp][1] + bf * pa[comp][1]) / (1.0f + az);
            float n3 = (c3 * sc[comp][2] - (1.0f - ax - az) * sp[comp][2] + bf * pa[comp][2]) / (1.0f + ax + az);
            L.prev[comp][0][o] = n1; L.prev[comp][1][o] = n2; L.prev[comp][2][o] = n3;
            put(g, comp == 0 ? a.uxp : a.uzp, co, L.ofs, (n1 + n2) + n3);
        }
    }

    static __device__ void run(const FrameArgs<NO> &a, int bid)
    {
        __shared__ Smem S;
        const Geo &g = a.g;
        const FTiles &t = a.t;
        int q = 0;
        while (q + 1 < t.n && bid >= t.off[q + 1]) ++q;
        const int loc = bid - t.off[q];
        const int bx = loc / t.nbz[q], bz = loc % t.nbz[q];
        const int xs = t.x0[q] + bx * t.lx[q];
        const int xe = min(xs + t.lx[q], t.x1[q] + 1);
        const int zs = t.z0[q] + bz * TZ;
        const int ze = min(zs + TZ - 1, t.z1[q]);
        Lane L;
        L.lane = threadIdx.x;
        for (int i = L.lane; i < LW; i += NT)
            if (i < PAD || i >= NT + PAD)
                for (int p = 0; p < 2; ++p)
                    for (int c = 0; c < 2; ++c) {
                        S.uA[p][c][i] = 0.0f; S.uB[p][c][i] = 0.0f; S.rM[p][c][i] = 1.0f;
                        S.fz[p][c][i] = 0.0f; S.vz[p][c][i] = 0.0f; S.pz[p][c][i] = 0.0f;
                    }
        L.r = zs - HZ + L.lane;
        const bool rowok = L.r >= 1 && L.r <= g.NZ;
        const int rc = L.r < 1 ? 1 : (L.r > g.NZ ? g.NZ : L.r);
        L.ofs = (unsigned)(rc - 1) * 4u;
        L.outrow = L.lane >= HZ && L.lane < HZ + TZ && L.r <= ze;
        L.isB = rowok && L.r > g.nz;
        L.az = rowok ? a.st->az[L.r - 1] : 0.0f;
        // the side strip of this tile's rows above the bottom strip
        const int sside = xs <= 0 ? S_L : S_R;
        const int s = L.isB ? S_B : sside;
        L.cb = (s == S_R) ? g.nx + 1 : 1 - g.npml;
        L.stride = L.isB ? g.npml : g.nz;
        L.rb = L.isB ? L.r - g.nz - 1 : L.r - 1;
        const Strips *st = a.st;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            L.psx[c] = L.isB ? st->psx[S_B][c] : st->psx[sside][c];
            L.psz[c] = st->psz[S_B][c];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                L.cur[c][p] = L.isB ? st->cur[S_B][c][p] : st->cur[sside][c][p];
                L.prev[c][p] = L.isB ? st->prev[S_B][c][p] : st->prev[sside][c][p];
            }
        }
        Regs R;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int i = 0; i < P; ++i) { R.U[c][i] = 0.0f; R.FX[c][i] = 0.0f; R.VX[c][i] = 0.0f; R.PX[c][i] = 0.0f; }
            R.pu[1][c] = 0.0f; R.pl[1][c] = 1.0f; R.p0[1][c] = 1.0f; R.pq[1][c] = 0.0f; R.pr[1][c] = 0.0f;
            R.ps[0][c] = 0.0f; R.ps[1][c] = 0.0f;
        }
        R.rMa_prev = 1.0f; R.rMm_prev = 1.0f;
        const int x0 = xs - WARM;
        prefetch(a, L, R, 0, x0, xs, xe);
        __syncthreads();
        for (int base = x0; base <= xe; base += P) {
#define OK_STEP(PH) if constexpr (P > PH) { if (base + PH <= xe) step<PH>(a, S, R, L, base + PH, xs, xe); }
            OK_STEP(0) OK_STEP(1) OK_STEP(2) OK_STEP(3) OK_STEP(4) OK_STEP(5) OK_STEP(6)
            OK_STEP(7) OK_STEP(8) OK_STEP(9) OK_STEP(10) OK_STEP(11) OK_STEP(12) OK_STEP(13)
#undef OK_STEP
        }
    }
};

template <int NO>
__global__ __launch_bounds__(64) void k_frame_out(FrameArgs<NO> a)
{
    OutKernel<NO, 64>::run(a, blockIdx.x);
}

}  // namespace fdfk
