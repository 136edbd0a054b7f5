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
    static constexpr int P = 3 * NO;        // register-ring period
    static constexpr int PAD = NO + 1;      // LDS guard entries on each side
    static constexpr int WARM = 4 * NO - 2; // warm-up columns before the first output
    static constexpr int LW = NT + 2 * PAD;

    struct Smem {
        float uA[2][LW];   // u at column ix+NO
        float uB[2][LW];   // u at column ix
        float rM[2][LW];   // 1/(l+2m), 1/m at column ix
        float fz[2][LW];   // z-fluxes at column ix
        float vz[2][LW];   // m*Dx uz, l*Dx ux at column ix
        float pz[2][LW];   // m*psi_z(ux), (l+2m)*psi_z(uz) at column ix (bottom strip, else 0)
    };

    // per-lane constants: row, masks, and the lane's strip (the side strip
    // above the bottom strip, the bottom strip below it) with its arrays
    struct Lane {
        int lane, r;
        bool rowok, outrow, isB;
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

    static __device__ __forceinline__ int sl(int ph, int j) { return ((ph + j) % P + P) % P; }

    template <int PH>
    static __device__ __forceinline__ void step(const FrameArgs<NO> &a, Smem &S, const Lane &L, int ix, int xs,
                                                int xe, float (&U)[2][P], float (&FX)[2][P], float (&VX)[2][P],
                                                float (&PX)[2][P], float &rMa_prev, float &rMm_prev)
    {
        if (ix >= xe) return;  // block-uniform
        const Geo &g = a.g;
        const int lane = L.lane, r = L.r;
        const size_t rowoff = (size_t)(r - 1);
        auto at = [&](int c) { return (size_t)(c + g.npml - 1) * (size_t)g.ld + rowoff; };
        // material columns of the first warm-up steps may lie left of the
        // grid; their values are never used, so the address is clamped
        auto atm = [&](int c) { return at(c < 1 - g.npml ? 1 - g.npml : c); };
        auto inside = [&](int c) { return c >= 1 && c <= g.nx && r <= g.nz; };
        const int cnew = ix + 2 * NO - 1, cmx = ix + NO;
        const bool warm = ix < xs - 2 * NO;   // block-uniform
        const bool outc = ix >= xs;           // block-uniform
        // ---- every global load of this step, issued together
        float nx_ = 0.0f, nz_ = 0.0f, l1 = 1.0f, m1 = 1.0f, l0 = 1.0f, m0 = 1.0f;
        float q0 = 0.0f, q1 = 0.0f, z0 = 0.0f, z1 = 0.0f, bf = 0.0f, ax = 0.0f;
        float sc[2][3] = {}, sp[2][3] = {};
        if (L.rowok) {
            const size_t cn = at(cnew), cm = atm(cmx);
            nx_ = a.ux[cn]; nz_ = a.uz[cn];
            if (a.corr && inside(cnew)) { nx_ -= a.ringx[cn]; nz_ -= a.ringz[cn]; }
            l1 = a.lam[cm]; m1 = a.mu[cm];
            if (!warm) {
                const size_t c0 = atm(ix);
                l0 = a.lam[c0]; m0 = a.mu[c0];
            }
            if (outc && L.isB) {
                const size_t o = L.off(ix);
                z0 = L.psz[0][o]; z1 = L.psz[1][o];
            }
        }
        if (!warm && L.outrow && !inside(cmx)) {
            const size_t o = L.off(cmx);
            q0 = L.psx[0][o]; q1 = L.psx[1][o];
        }
        if (outc && L.outrow) {
            const size_t o = L.off(ix);
            bf = a.bf[at(ix)];
            ax = a.st->ax[ix + g.npml - 1];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int p = 0; p < 3; ++p) { sc[c][p] = L.cur[c][p][o]; sp[c][p] = L.prev[c][p][o]; }
        }
        // ---- 1. new u column (scattered view) into the ring, x-fluxes
        U[0][sl(PH, 2 * NO - 1)] = nx_;
        U[1][sl(PH, 2 * NO - 1)] = nz_;
        const float rMa = 1.0f / (l1 + 2.0f * m1), rMm = 1.0f / m1;
        if (!warm) {
            float ha = 1.0f / (rMa_prev + rMa), hm = 1.0f / (rMm_prev + rMm);
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                s0 += a.k.ai[j - 1] * (U[0][sl(PH, NO - 1 + j)] - U[0][sl(PH, NO - j)]);
                s1 += a.k.ai[j - 1] * (U[1][sl(PH, NO - 1 + j)] - U[1][sl(PH, NO - j)]);
            }
            FX[0][sl(PH, NO - 1)] = ha * s0;
            FX[1][sl(PH, NO - 1)] = hm * s1;
            // M * psi_x at column ix+NO for this lane's own row
            PX[0][sl(PH, NO)] = (l1 + 2.0f * m1) * q0;
            PX[1][sl(PH, NO)] = m1 * q1;
            S.uA[0][PAD + lane] = U[0][sl(PH, NO)];
            S.uA[1][PAD + lane] = U[1][sl(PH, NO)];
            S.uB[0][PAD + lane] = U[0][sl(PH, 0)];
            S.uB[1][PAD + lane] = U[1][sl(PH, 0)];
        }
        rMa_prev = rMa;
        rMm_prev = rMm;
        if (warm) return;
        // ---- 2. 1/M and M * psi_z at column ix
        const float rMa0 = 1.0f / (l0 + 2.0f * m0), rMm0 = 1.0f / m0;
        S.rM[0][PAD + lane] = rMa0;
        S.rM[1][PAD + lane] = rMm0;
        S.pz[0][PAD + lane] = m0 * z0;
        S.pz[1][PAD + lane] = (l0 + 2.0f * m0) * z1;
        __syncthreads();
        // ---- 3. Vx at column ix+NO: l*Dz uz and m*Dz ux
        {
            float dzz = 0.0f, dzx = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                dzz += a.k.bi[j - 1] * (S.uA[1][PAD + lane + j] - S.uA[1][PAD + lane - j]);
                dzx += a.k.bi[j - 1] * (S.uA[0][PAD + lane + j] - S.uA[0][PAD + lane - j]);
            }
            VX[0][sl(PH, NO)] = l1 * dzz;
            VX[1][sl(PH, NO)] = m1 * dzx;
        }
        if (!outc) {
            __syncthreads();  // keep the barrier cadence uniform
            return;
        }
        // ---- 4. z-fluxes at (r+1/2, ix), Vz at (r, ix), z memory sums at (r, ix)
        float sfx = 0.0f, sfz = 0.0f, pzx = 0.0f, pzz = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            sfx += a.k.ai[j - 1] * (S.uB[0][PAD + lane + j] - S.uB[0][PAD + lane + 1 - j]);
            sfz += a.k.ai[j - 1] * (S.uB[1][PAD + lane + j] - S.uB[1][PAD + lane + 1 - j]);
            pzx += a.k.bi[j - 1] * (S.pz[0][PAD + lane + j] - S.pz[0][PAD + lane - j]);
            pzz += a.k.bi[j - 1] * (S.pz[1][PAD + lane + j] - S.pz[1][PAD + lane - j]);
        }
        const float fzx = sfx / (rMm0 + S.rM[1][PAD + lane + 1]);
        const float fzz = sfz / (rMa0 + S.rM[0][PAD + lane + 1]);
        float dxz = 0.0f, dxx = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            dxz += a.k.bi[k - 1] * (U[1][sl(PH, k)] - U[1][sl(PH, -k)]);
            dxx += a.k.bi[k - 1] * (U[0][sl(PH, k)] - U[0][sl(PH, -k)]);
        }
        __syncthreads();  // all reads of uA/uB/rM/pz for this column are done
        S.fz[0][PAD + lane] = fzx;
        S.fz[1][PAD + lane] = fzz;
        S.vz[0][PAD + lane] = m0 * dxz;
        S.vz[1][PAD + lane] = l0 * dxx;
        __syncthreads();
        if (!L.outrow) return;
        // ---- 5. the six parts, memory corrections, split update
        float Xx = 0.0f, Xz = 0.0f, Zx = 0.0f, Zz = 0.0f, Mx = 0.0f, Mz = 0.0f, sx0 = 0.0f, sx1 = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            Xx += a.k.aox[k - 1] * (FX[0][sl(PH, k - 1)] - FX[0][sl(PH, -k)]);
            Xz += a.k.aox[k - 1] * (FX[1][sl(PH, k - 1)] - FX[1][sl(PH, -k)]);
            Zx += a.k.aoz[k - 1] * (S.fz[0][PAD + lane + k - 1] - S.fz[0][PAD + lane - k]);
            Zz += a.k.aoz[k - 1] * (S.fz[1][PAD + lane + k - 1] - S.fz[1][PAD + lane - k]);
            Mx += a.k.bo[k - 1] * ((VX[0][sl(PH, k)] - VX[0][sl(PH, -k)]) +
                                  (S.vz[0][PAD + lane + k] - S.vz[0][PAD + lane - k]));
            Mz += a.k.bo[k - 1] * ((VX[1][sl(PH, k)] - VX[1][sl(PH, -k)]) +
                                  (S.vz[1][PAD + lane + k] - S.vz[1][PAD + lane - k]));
            sx0 += a.k.bi[k - 1] * (PX[0][sl(PH, k)] - PX[0][sl(PH, -k)]);
            sx1 += a.k.bi[k - 1] * (PX[1][sl(PH, k)] - PX[1][sl(PH, -k)]);
        }
        Xx = Xx - sx0 / g.dx;
        Xz = Xz - sx1 / g.dx;
        if (L.isB) {
            Zx = Zx - pzx / g.dz;
            Zz = Zz - pzz / g.dz;
        }
        const float az = L.az;
        const float c3 = 2.0f - 4.0f * ax * az;
        const float pa[2][3] = {{Xx, Zx, Mx}, {Xz, Zz, Mz}};
        const size_t o = L.off(ix);
        const size_t c = at(ix);
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
            float n1 = (2.0f * sc[comp][0] - (1.0f - ax) * sp[comp][0] + bf * pa[comp][0]) / (1.0f + ax);
            float n2 = (2.0f * sc[comp][1] - (1.0f - az) * sp[comp][1] + bf * pa[comp][1]) / (1.0f + az);
            float n3 = (c3 * sc[comp][2] - (1.0f - ax - az) * sp[comp][2] + bf * pa[comp][2]) / (1.0f + ax + az);
            L.prev[comp][0][o] = n1; L.prev[comp][1][o] = n2; L.prev[comp][2][o] = n3;
            float tot = (n1 + n2) + n3;
            if (comp == 0) a.uxp[c] = tot; else a.uzp[c] = tot;
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
                for (int c = 0; c < 2; ++c) {
                    S.uA[c][i] = 0.0f; S.uB[c][i] = 0.0f; S.rM[c][i] = 1.0f;
                    S.fz[c][i] = 0.0f; S.vz[c][i] = 0.0f; S.pz[c][i] = 0.0f;
                }
        L.r = zs - HZ + L.lane;
        L.rowok = L.r >= 1 && L.r <= g.NZ;
        L.outrow = L.lane >= HZ && L.lane < HZ + TZ && L.r <= ze;
        L.isB = L.r > g.nz;
        L.az = L.rowok ? a.st->az[L.r - 1] : 0.0f;
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
        float U[2][P], FX[2][P], VX[2][P], PX[2][P];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < P; ++i) { U[c][i] = 0.0f; FX[c][i] = 0.0f; VX[c][i] = 0.0f; PX[c][i] = 0.0f; }
        float rMa_prev = 1.0f, rMm_prev = 1.0f;
        __syncthreads();
        for (int base = xs - WARM; base < xe; base += P) {
#define OK_STEP(PH) if constexpr (P > PH) step<PH>(a, S, L, base + PH, xs, xe, U, FX, VX, PX, rMa_prev, rMm_prev);
            OK_STEP(0) OK_STEP(1) OK_STEP(2) OK_STEP(3) OK_STEP(4) OK_STEP(5)
            OK_STEP(6) OK_STEP(7) OK_STEP(8) OK_STEP(9) OK_STEP(10) OK_STEP(11)
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
