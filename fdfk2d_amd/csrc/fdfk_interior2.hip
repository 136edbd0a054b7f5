// ---------------------------------------------------------------------
// Two time steps per pass over HBM for the deep interior (included by
// fdfk_api.hip after fdfk_kernels.hip and fdfk_frame.hip).
//
// A block marches along x like IntKernel2, with two stencil engines in one
// pipeline.  Engine 1 is IntKernel2's step: it takes u^n columns and
// produces u^{n+1}.  Engine 2 is the same step one stage later: its new
// column is the u^{n+1} column engine 1 finished the step before, and it
// produces u^{n+2}.  Engine 2 reads the materials engine 1 loaded through
// register delay lines, and its leapfrog takes u^n of its column from a
// delay line as well.  So a cell's u^{n-1}, u^n and materials are read from
// HBM once for two steps, and u^{n+1} and u^{n+2} are written once: 44
// bytes per cell for two cell-updates instead of 72.
//
// The deep region lies at least two stencil reaches (2 (2NO-1) cells)
// inside the regular interior (full-order stencils, no ring correction, no
// free-surface closure), so the u^{n+1} values engine 1 computes in a
// tile's halo (needed by engine 2) are exactly what the single-step
// kernels compute there.  The tile stores both levels for its own cells
// only.  Every value is formed with the same arithmetic, in the same
// order, as IntKernel2's step, so the results equal two single steps bit
// for bit.
//
// Lanes run along z: lane l holds row z0 - 2HZ + l; engine 1's results are
// valid on lanes HZ..NT-1-HZ, engine 2's on 2HZ..NT-1-2HZ.  Rows beyond the
// tile's own (tz of them from z0) are not stored.
// Engine 1 at march step s: new u^n column s+2NO-1, update of column s-1.
// Engine 2 at step s: new u^{n+1} column s-2 (engine 1's previous update),
// its ix2 = s-2NO-1, update of column s-LAG2, LAG2 = 2NO+2.  Ring slots of
// both engines are addressed relative to their own ix through the same
// phase PH = s mod P; the delay lines relative to s.
// ---------------------------------------------------------------------
namespace fdfk {

template <int NO>
struct DeepArgs {
    Geo g;          // ix0..ix1, iz0..iz1: the deep rectangle
    ICoef<NO> k;
    const float *ux, *uz;      // u^n
    const float *uxp, *uzp;    // u^{n-1}
    float *o1x, *o1z;          // u^{n+1} out
    float *o2x, *o2z;          // u^{n+2} out
    const float *bf, *lam, *mu;
    int lx;     // output columns per tile
    int nbx;    // tiles along x
    int nbz;    // tiles along z
    int tz;     // output rows per tile (the last tile may hold fewer)
};

template <int NO>
struct IntKernel22 {
    using K1 = IntKernel2<NO, 256>;
    static constexpr int NT = 256;
    static constexpr int HZ = 2 * NO - 1;
    static constexpr int TZMAX = NT - 4 * HZ;
    static constexpr int P = K1::P;
    static constexpr int PAD = NO + 1;
    static constexpr int LW = NT + 2 * PAD;
    static constexpr int LAG2 = 2 * NO + 2;
    static constexpr int WARM = 6 * NO - 3;
    static_assert(P >= LAG2, "the delay lines must fit the ring period");
    static_assert(P <= 14, "the step lists of run() cover ring periods up to 14 (NO <= 4)");

    // LDS: per row 20 words per engine (the layout of IntKernel2, one block
    // per engine) and one word of padding: an odd stride keeps the lanes of
    // a wavefront on distinct banks
    enum { UA = 0, UB = 2, RM = 4, FZ = 6, VZ = 8 };
    static constexpr int SW = 41;
    struct Smem {
        float m[LW * SW];
        __device__ __forceinline__ float &at(int e, int par, int q, int c, int i)
        {
            return m[i * SW + e * 20 + par * 10 + q + c];
        }
    };

    static __device__ __forceinline__ int sl(int ph, int j) { return ((ph + j) % P + P) % P; }

    // Stores of cells a tile does not own go to a lane offset at or above
    // 2^31, beyond the 0x7fffffff bytes of every buffer descriptor: the
    // buffer unit's range check discards them.  The march step then has no
    // branches, and the compiler schedules the two engines as one block.
    static constexpr unsigned kDropOffset = 0x80000000u;

    struct Ring {
        float U[2][P], FX[2][P], VX[2][P];
        float rMa_prev, rMm_prev;
    };
    struct Regs {
        Ring E1, E2;           // engine 1 on u^n, engine 2 on u^{n+1}
        float ML[P], MM[P];    // lam, mu of column s at slot sl(PH, 0)
        float BF[P];           // dt^2/rho of column s-1 at slot sl(PH, -1)
        float UN[2][P];        // u^n of column s-1 at slot sl(PH, -1)
        float pu[2][2], pl[2][2], p0[2][2], po[2][3];
        float nu[2];           // engine 1's last update (u^{n+1} of column s-2 at step s)
    };

    struct Rsrc {
        __amdgpu_buffer_rsrc_t ux, uz, uxp, uzp, bf, lam, mu, o1x, o1z, o2x, o2z;
        int xb, pitch;
        __device__ __forceinline__ int soff(int c) const { return (c - xb) * pitch; }
    };

    // phase A of one engine after its new column entered the ring: the
    // x-flux at face (ix+NO-1, ix+NO) and the LDS staging of u(ix+NO), u(ix),
    // 1/M(ix) (IntKernel2's step, phase A)
    template <int PH>
    static __device__ __forceinline__ void phase_a(const DeepArgs<NO> &a, Smem &S, Ring &E, int e, int lane,
                                                   float l1, float m1, float l0, float m0)
    {
        constexpr int PA = PH & 1;
        const float rMa = frcp(l1 + 2.0f * m1), rMm = frcp(m1);
        {
            const float ha = frcp(E.rMa_prev + rMa), hm = frcp(E.rMm_prev + rMm);
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                s0 += a.k.ai[j - 1] * (E.U[0][sl(PH, NO - 1 + j)] - E.U[0][sl(PH, NO - j)]);
                s1 += a.k.ai[j - 1] * (E.U[1][sl(PH, NO - 1 + j)] - E.U[1][sl(PH, NO - j)]);
            }
            E.FX[0][sl(PH, NO - 1)] = ha * s0;
            E.FX[1][sl(PH, NO - 1)] = hm * s1;
        }
        E.rMa_prev = rMa;
        E.rMm_prev = rMm;
        S.at(e, PA, UA, 0, PAD + lane) = E.U[0][sl(PH, NO)];
        S.at(e, PA, UA, 1, PAD + lane) = E.U[1][sl(PH, NO)];
        S.at(e, PA, UB, 0, PAD + lane) = E.U[0][sl(PH, 0)];
        S.at(e, PA, UB, 1, PAD + lane) = E.U[1][sl(PH, 0)];
        S.at(e, PA, RM, 0, PAD + lane) = frcp(l0 + 2.0f * m0);
        S.at(e, PA, RM, 1, PAD + lane) = frcp(m0);
    }

    // phase C of one engine: Vx at ix+NO into the ring, z-fluxes and Vz at ix
    // into LDS (IntKernel2's step, phase C)
    template <int PH>
    static __device__ __forceinline__ void phase_c(const DeepArgs<NO> &a, Smem &S, Ring &E, int e, int lane,
                                                   float l1, float m1, float l0, float m0)
    {
        constexpr int PA = PH & 1;
        {
            float dzz = 0.0f, dzx = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                dzz += a.k.bi[j - 1] * (S.at(e, PA, UA, 1, PAD + lane + j) - S.at(e, PA, UA, 1, PAD + lane - j));
                dzx += a.k.bi[j - 1] * (S.at(e, PA, UA, 0, PAD + lane + j) - S.at(e, PA, UA, 0, PAD + lane - j));
            }
            E.VX[0][sl(PH, NO)] = l1 * dzz;
            E.VX[1][sl(PH, NO)] = m1 * dzx;
        }
        float sfx = 0.0f, sfz = 0.0f;
#pragma unroll
        for (int j = 1; j <= NO; ++j) {
            sfx += a.k.ai[j - 1] * (S.at(e, PA, UB, 0, PAD + lane + j) - S.at(e, PA, UB, 0, PAD + lane + 1 - j));
            sfz += a.k.ai[j - 1] * (S.at(e, PA, UB, 1, PAD + lane + j) - S.at(e, PA, UB, 1, PAD + lane + 1 - j));
        }
        const float rMa0 = S.at(e, PA, RM, 0, PAD + lane), rMm0 = S.at(e, PA, RM, 1, PAD + lane);
        const float fzx = sfx * frcp(rMm0 + S.at(e, PA, RM, 1, PAD + lane + 1));
        const float fzz = sfz * frcp(rMa0 + S.at(e, PA, RM, 0, PAD + lane + 1));
        float dxz = 0.0f, dxx = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            dxz += a.k.bi[k - 1] * (E.U[1][sl(PH, k)] - E.U[1][sl(PH, -k)]);
            dxx += a.k.bi[k - 1] * (E.U[0][sl(PH, k)] - E.U[0][sl(PH, -k)]);
        }
        S.at(e, PA, FZ, 0, PAD + lane) = fzx;
        S.at(e, PA, FZ, 1, PAD + lane) = fzz;
        S.at(e, PA, VZ, 0, PAD + lane) = m0 * dxz;
        S.at(e, PA, VZ, 1, PAD + lane) = l0 * dxx;
    }

    // phase D of one engine: the right-hand side of the update of column
    // ix-1 (its z-terms were staged in the previous step's LDS parity)
    template <int PH>
    static __device__ __forceinline__ void phase_d(const DeepArgs<NO> &a, Smem &S, const Ring &E, int e, int lane,
                                                   float &rx, float &rz)
    {
        constexpr int PB = 1 - (PH & 1);
        float Xx = 0.0f, Xz = 0.0f, Zx = 0.0f, Zz = 0.0f, Mx = 0.0f, Mz = 0.0f;
#pragma unroll
        for (int k = 1; k <= NO; ++k) {
            Xx += a.k.aox[k - 1] * (E.FX[0][sl(PH, k - 2)] - E.FX[0][sl(PH, -k - 1)]);
            Xz += a.k.aox[k - 1] * (E.FX[1][sl(PH, k - 2)] - E.FX[1][sl(PH, -k - 1)]);
            Zx += a.k.aoz[k - 1] * (S.at(e, PB, FZ, 0, PAD + lane + k - 1) - S.at(e, PB, FZ, 0, PAD + lane - k));
            Zz += a.k.aoz[k - 1] * (S.at(e, PB, FZ, 1, PAD + lane + k - 1) - S.at(e, PB, FZ, 1, PAD + lane - k));
            Mx += a.k.bo[k - 1] * ((E.VX[0][sl(PH, k - 1)] - E.VX[0][sl(PH, -k - 1)]) +
                                  (S.at(e, PB, VZ, 0, PAD + lane + k) - S.at(e, PB, VZ, 0, PAD + lane - k)));
            Mz += a.k.bo[k - 1] * ((E.VX[1][sl(PH, k - 1)] - E.VX[1][sl(PH, -k - 1)]) +
                                  (S.at(e, PB, VZ, 1, PAD + lane + k) - S.at(e, PB, VZ, 1, PAD + lane - k)));
        }
        rx = (Xx + Zx) + Mx;
        rz = (Xz + Zz) + Mz;
    }

    template <int PH>
    static __device__ __forceinline__ void step(const DeepArgs<NO> &a, const Rsrc &B, Smem &S, Regs &R, int lane,
                                                unsigned o, bool own, int s, int xs, int xe)
    {
        constexpr int PA = PH & 1, PB = 1 - PA;
        const Geo &g = a.g;
        const int lo = 1 - g.npml, hi = g.nx + g.npml;
        auto colc = [&](int c) { return c < lo ? lo : (c > hi ? hi : c); };
        // columns whose loads matter: engine 1 updates the columns within the
        // x-reach 2NO-1 of the tile, so its inputs reach twice as far
        auto cu = [&](int c) {
            return colc(c < xs - (4 * NO - 2) ? xs - (4 * NO - 2) : (c > xe + 4 * NO - 3 ? xe + 4 * NO - 3 : c));
        };
        auto cl = [&](int c) {
            return colc(c < xs - (3 * NO - 1) ? xs - (3 * NO - 1) : (c > xe + 3 * NO - 2 ? xe + 3 * NO - 2 : c));
        };
        auto ct = [&](int c) {
            return colc(c < xs - (2 * NO - 1) ? xs - (2 * NO - 1) : (c > xe + 2 * NO - 2 ? xe + 2 * NO - 2 : c));
        };
        // ---- A: engine 1 (new u^n column, prefetch of the next step)
        R.E1.U[0][sl(PH, 2 * NO - 1)] = R.pu[PA][0];
        R.E1.U[1][sl(PH, 2 * NO - 1)] = R.pu[PA][1];
        const float l1 = R.pl[PA][0], m1 = R.pl[PA][1];
        const float l0 = R.p0[PA][0], m0 = R.p0[PA][1];
        {
            const int cn = cu(s + 2 * NO), cm = cl(s + 1 + NO), c0 = ct(s + 1), cx = ct(s);
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0), sx = B.soff(cx);
            R.pu[PB][0] = K1::ld(B.ux, o, sn);
            R.pu[PB][1] = K1::ld(B.uz, o, sn);
            R.pl[PB][0] = K1::ld(B.lam, o, sm); R.pl[PB][1] = K1::ld(B.mu, o, sm);
            R.p0[PB][0] = K1::ld(B.lam, o, s0); R.p0[PB][1] = K1::ld(B.mu, o, s0);
            R.po[PA][0] = K1::ld(B.bf, o, sx);
            R.po[PA][1] = K1::ld(B.uxp, o, sx);
            R.po[PA][2] = K1::ld(B.uzp, o, sx);
        }
        phase_a<PH>(a, S, R.E1, 0, lane, l1, m1, l0, m0);
        // delay lines: materials of column s, dt^2/rho and u^n of column s-1
        R.ML[sl(PH, 0)] = l0;
        R.MM[sl(PH, 0)] = m0;
        R.BF[sl(PH, -1)] = R.po[PB][0];
        R.UN[0][sl(PH, -1)] = R.E1.U[0][sl(PH, -1)];
        R.UN[1][sl(PH, -1)] = R.E1.U[1][sl(PH, -1)];
        // ---- A: engine 2 (new u^{n+1} column s-2; lead materials at
        // ix2+NO = s-NO-1, column materials at ix2 = s-2NO-1)
        R.E2.U[0][sl(PH, 2 * NO - 1)] = R.nu[0];
        R.E2.U[1][sl(PH, 2 * NO - 1)] = R.nu[1];
        const float l21 = R.ML[sl(PH, -NO - 1)], m21 = R.MM[sl(PH, -NO - 1)];
        const float l20 = R.ML[sl(PH, -2 * NO - 1)], m20 = R.MM[sl(PH, -2 * NO - 1)];
        phase_a<PH>(a, S, R.E2, 1, lane, l21, m21, l20, m20);
        // ---- B: one barrier for both engines
        lds_barrier();
        // ---- C
        phase_c<PH>(a, S, R.E1, 0, lane, l1, m1, l0, m0);
        phase_c<PH>(a, S, R.E2, 1, lane, l21, m21, l20, m20);
        // ---- D: engine 1's update of column s-1 (u^{n+1})
        {
            float rx, rz;
            phase_d<PH>(a, S, R.E1, 0, lane, rx, rz);
            const float bf = R.po[PB][0];
            const float nx_ = (2.0f * R.E1.U[0][sl(PH, -1)] - R.po[PB][1]) + bf * rx;
            const float nz_ = (2.0f * R.E1.U[1][sl(PH, -1)] - R.po[PB][2]) + bf * rz;
            const unsigned vo = (own && s - 1 >= xs && s - 1 < xe) ? o : (o | kDropOffset);
            K1::st(B.o1x, vo, B.soff(s - 1), nx_);
            K1::st(B.o1z, vo, B.soff(s - 1), nz_);
            R.nu[0] = nx_;
            R.nu[1] = nz_;
        }
        // ---- D: engine 2's update of column s-LAG2 (u^{n+2})
        {
            float rx, rz;
            phase_d<PH>(a, S, R.E2, 1, lane, rx, rz);
            const float bf = R.BF[sl(PH, -LAG2)];
            const float nx_ = (2.0f * R.E2.U[0][sl(PH, -1)] - R.UN[0][sl(PH, -LAG2)]) + bf * rx;
            const float nz_ = (2.0f * R.E2.U[1][sl(PH, -1)] - R.UN[1][sl(PH, -LAG2)]) + bf * rz;
            const unsigned vo = (own && s - LAG2 >= xs && s - LAG2 < xe) ? o : (o | kDropOffset);
            K1::st(B.o2x, vo, B.soff(s - LAG2), nx_);
            K1::st(B.o2z, vo, B.soff(s - LAG2), nz_);
        }
    }

    static __device__ __forceinline__ __amdgpu_buffer_rsrc_t mk(const Geo &g, const float *base, int xb)
    {
        return K1::mk(g, base, xb);
    }

    static __device__ __forceinline__ void run(const DeepArgs<NO> &a, int bid, Smem &S)
    {
        const Geo &g = a.g;
        const int lane = threadIdx.x;
        for (int i = lane; i < LW; i += NT)
            if (i < PAD || i >= NT + PAD)
                for (int e = 0; e < 2; ++e)
                    for (int q = 0; q < 2; ++q)
                        for (int c = 0; c < 2; ++c) {
                            S.at(e, q, UA, c, i) = 0.0f; S.at(e, q, UB, c, i) = 0.0f; S.at(e, q, RM, c, i) = 1.0f;
                            S.at(e, q, FZ, c, i) = 0.0f; S.at(e, q, VZ, c, i) = 0.0f;
                        }
        // tile id -> tile: consecutive ids walk down z first (z neighbours
        // share an XCD through xcd_tile)
        const int bzid = bid % a.nbz, bxid = bid / a.nbz;
        const int z0 = g.iz0 + bzid * a.tz;
        const int zl = min(z0 + a.tz - 1, g.iz1);       // the tile's last output row
        const int xs = g.ix0 + bxid * a.lx;
        const int xe = min(xs + a.lx, g.ix1 + 1);
        const int r = z0 - 2 * HZ + lane;
        const int rc = r < 1 ? 1 : (r > g.NZ ? g.NZ : r);
        const unsigned o = (unsigned)(rc - 1) * 4u;
        const bool own = lane >= 2 * HZ && r <= zl;
        Regs R;
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int i = 0; i < P; ++i) {
                R.E1.U[q][i] = 0.0f; R.E1.FX[q][i] = 0.0f; R.E1.VX[q][i] = 0.0f;
                R.E2.U[q][i] = 0.0f; R.E2.FX[q][i] = 0.0f; R.E2.VX[q][i] = 0.0f;
                R.UN[q][i] = 0.0f;
            }
#pragma unroll
        for (int i = 0; i < P; ++i) { R.ML[i] = 1.0f; R.MM[i] = 1.0f; R.BF[i] = 0.0f; }
        R.E1.rMa_prev = 1.0f; R.E1.rMm_prev = 1.0f;
        R.E2.rMa_prev = 1.0f; R.E2.rMm_prev = 1.0f;
        R.nu[0] = 0.0f; R.nu[1] = 0.0f;
        const int x0 = xs - WARM;
        Rsrc B;
        {
            const int lo = 1 - g.npml;
            B.xb = x0 < lo ? lo : x0;
            B.pitch = g.ld * 4;
            B.ux = mk(g, a.ux, B.xb); B.uz = mk(g, a.uz, B.xb);
            B.uxp = mk(g, a.uxp, B.xb); B.uzp = mk(g, a.uzp, B.xb);
            B.bf = mk(g, a.bf, B.xb); B.lam = mk(g, a.lam, B.xb); B.mu = mk(g, a.mu, B.xb);
            B.o1x = mk(g, a.o1x, B.xb); B.o1z = mk(g, a.o1z, B.xb);
            B.o2x = mk(g, a.o2x, B.xb); B.o2z = mk(g, a.o2z, B.xb);
        }
        // prologue: the loads the first step consumes
        {
            const int lo = 1 - g.npml;
            const int cn = x0 + 2 * NO - 1;   // = xs - (4NO-2), the first u^n column the tile reaches
            const int cm = xs - (3 * NO - 1) > lo ? xs - (3 * NO - 1) : lo;
            const int c0 = xs - (2 * NO - 1) > lo ? xs - (2 * NO - 1) : lo;
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0);
            R.pu[0][0] = K1::ld(B.ux, o, sn); R.pu[0][1] = K1::ld(B.uz, o, sn);
            R.pl[0][0] = K1::ld(B.lam, o, sm); R.pl[0][1] = K1::ld(B.mu, o, sm);
            R.p0[0][0] = K1::ld(B.lam, o, s0); R.p0[0][1] = K1::ld(B.mu, o, s0);
#pragma unroll
            for (int q = 0; q < 2; ++q) { R.pu[1][q] = 0.0f; R.pl[1][q] = 1.0f; R.p0[1][q] = 1.0f; }
#pragma unroll
            for (int q = 0; q < 3; ++q) { R.po[0][q] = 0.0f; R.po[1][q] = 0.0f; }
        }
        __syncthreads();
        const int xl = xe - 1 + LAG2;     // the step that updates engine 2's last column
        int base = x0;
        for (; base + P - 1 <= xl; base += P) {
#define IK22_STEP(PH) if constexpr (P > PH) step<PH>(a, B, S, R, lane, o, own, base + PH, xs, xe);
            IK22_STEP(0) IK22_STEP(1) IK22_STEP(2) IK22_STEP(3) IK22_STEP(4) IK22_STEP(5) IK22_STEP(6)
            IK22_STEP(7) IK22_STEP(8) IK22_STEP(9) IK22_STEP(10) IK22_STEP(11) IK22_STEP(12) IK22_STEP(13)
#undef IK22_STEP
        }
#define IK22_STEP(PH) if constexpr (P > PH) { if (base + PH <= xl) step<PH>(a, B, S, R, lane, o, own, base + PH, xs, xe); }
        IK22_STEP(0) IK22_STEP(1) IK22_STEP(2) IK22_STEP(3) IK22_STEP(4) IK22_STEP(5) IK22_STEP(6)
        IK22_STEP(7) IK22_STEP(8) IK22_STEP(9) IK22_STEP(10) IK22_STEP(11) IK22_STEP(12) IK22_STEP(13)
#undef IK22_STEP
    }
};

// The regular interior around the deep rectangle, in four bands (the rows
// above and below it across the whole width, the columns left and right of
// it beside it), each stepped once by IntKernel2 tiles.
template <int NO>
struct Bands {
    IntArgs<NO> b[4];
    int off[5];   // block ids of band k: [off[k], off[k+1])
};

// One launch of a two-step pair.  DEEP (the first step): blocks [0, ndeep)
// are deep tiles writing level n+1 and n+2 of their cells, then the bands
// and the frame step n -> n+1.  Without DEEP (the second step): the bands
// and the frame step n+1 -> n+2.  Every block writes cells no other block
// of the launch writes and reads only levels the launch does not write.
// Waves per SIMD of a launch with deep tiles: the two engines' rings and
// delay lines fit 256 registers up to order 3 (two waves), order 4 takes
// one wave and the accumulation registers rather than spill.
template <int NO>
constexpr int kDeepWaves = NO <= 3 ? 2 : 1;

template <int NO, bool DEEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DEEP ? kDeepWaves<NO> : kIntWaves<NO>))) void k_step2(
    DeepArgs<NO> da, Bands<NO> bs, StepArgs sa, ETiles ft, int ndeep)
{
    extern __shared__ float sm[];
    int bid = (int)blockIdx.x;
    if constexpr (DEEP) {
        if (bid < ndeep) {
            IntKernel22<NO>::run(da, xcd_tile(bid, ndeep), *reinterpret_cast<typename IntKernel22<NO>::Smem *>(sm));
            return;
        }
        bid -= ndeep;
    }
    const int nb = bs.off[4];
    if (bid < nb) {
        const int q = xcd_tile(bid, nb);
        const int k = q < bs.off[1] ? 0 : (q < bs.off[2] ? 1 : (q < bs.off[3] ? 2 : 3));
        IntKernel2<NO, 256>::run(bs.b[k], q - bs.off[k], *reinterpret_cast<typename IntKernel2<NO, 256>::Smem *>(sm));
        return;
    }
    const int nf = (int)gridDim.x - (DEEP ? ndeep : 0) - nb;
    frame_body<NO>(sa, ft, xcd_tile(bid - nb, nf));
}

}  // namespace fdfk
