// SH interior kernel (included by fdfk_api.hip after fdfk_kernels.hip): the
// transverse motion rho uy_tt = Dx(mu Dx uy) + Dz(mu Dz uy) over the deep
// total-field region of a general-path solver (FP32), with the staggered
// full-order operators of src/Submain.f90 (diffx2 / diffz2 with the
// harmonic mean of mu at the faces) and the leapfrog update.
//
// Same organisation as the P-SV interior kernel IntKernel2 with one
// component: 64-wide wavefronts along z (one row per lane, HZ halo rows on
// either side of the TZ output rows), the block marching along x with the
// x-neighbourhood of uy and the x-face fluxes in register rings, the
// z-face fluxes exchanged through LDS.  Per step:
//   A  take the prefetched column ix+2NO-1 into the ring, prefetch column
//      ix+2NO and the materials of the next step, form the x-flux at the
//      face (ix+NO-1, ix+NO) and stage uy(ix), 1/mu(ix) in LDS;
//   B  one barrier (LDS double-buffered by column parity);
//   C  the z-flux at column ix (to LDS);
//   D  the update of column ix-1, whose z-flux was staged one step before.
// It reads uy^n, uy^{n-1}, dt^2/rho and mu once and writes uy^{n+1} in place
// of uy^{n-1}: 20 algorithmic bytes per cell-update.  The field, ring and
// material arguments of IntArgs carry uy in the x slots (uz, lam unused).
namespace fdfk {

template <int NO, int NT>
struct IntKernelSH {
    static constexpr int HZ = 2 * NO - 1;
    static constexpr int TZ = NT - 2 * HZ;
    static constexpr int P0 = 3 * NO + 1;
    static constexpr int P = P0 + (P0 & 1);
    static constexpr int PAD = NO + 1;
    static constexpr int WARM = 4 * NO - 2;
    static constexpr int LW = NT + 2 * PAD;
    static_assert(HZ == IntKernel2<NO, NT>::HZ && TZ == IntKernel2<NO, NT>::TZ, "same tiling as the P-SV interior");

    // LDS, lane-major, two words per column parity: uy(ix), 1/mu(ix), the
    // z-flux at ix; the odd stride keeps a wavefront's lanes on distinct banks
    enum { UB = 0, RM = 2, FZ = 4 };
    static constexpr int SW = 7;
    struct Smem {
        float m[LW * SW];
        __device__ __forceinline__ float &at(int par, int q, int i) { return m[i * SW + q + par]; }
    };

    static __device__ __forceinline__ int sl(int ph, int j) { return ((ph + j) % P + P) % P; }

    struct Regs {
        float U[P], FX[P];
        float rM_prev;
        float pu[2];     // [parity] next new uy column
        float pl[2];     // [parity] next lead mu (ix+NO)
        float p0[2];     // [parity] next column mu (ix)
        float po[2][2];  // [parity][bf, uyp] of column ix (used one step later)
        float pr[2];     // [parity] u0 of the next new column on the outer ring
    };

    struct Lane {
        int lane, r;
        unsigned ofs;
        bool outrow;
        bool rb, rr;
    };
    static __device__ __forceinline__ bool ring_cell(const Geo &g, const Lane &L, int c)
    {
        if (c >= 1 && c <= g.nx) return L.rb;
        return c >= 1 - g.nbd && c <= g.nx + g.nbd && L.rr;
    }
    struct Rsrc {
        __amdgpu_buffer_rsrc_t u, up, uo, bf, mu, ring;
        int xb, pitch;
        __device__ __forceinline__ int soff(int c) const { return (c - xb) * pitch; }
    };
    static __device__ __forceinline__ __amdgpu_buffer_rsrc_t mk(const Geo &g, const float *base, int xb)
    {
        const float *cb = base + (size_t)(xb + g.npml - 1) * (size_t)g.ld;
        return __builtin_amdgcn_make_buffer_rsrc((void *)cb, 0, 0x7fffffff, kBufferFlags);
    }
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, unsigned o, int so)
    {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0));
    }
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, unsigned o, int so, float v)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, o, so, 0);
    }

    template <int PH>
    static __device__ __forceinline__ void step(const IntArgs<NO> &a, const Rsrc &B, Smem &S, Regs &R, const Lane &L,
                                                bool bc, int ix, int xs, int xe)
    {
        constexpr int PA = PH & 1, PB = 1 - PA;
        const Geo &g = a.g;
        const int lane = L.lane;
        const int lo = 1 - g.npml, hi = g.nx + g.npml;
        auto colc = [&](int c) { return c < lo ? lo : (c > hi ? hi : c); };
        // loads outside the columns a stored update uses are redirected to
        // the nearest such column (values never used, a cache hit)
        auto cu = [&](int c) { return colc(c < xs - (2 * NO - 1) ? xs - (2 * NO - 1) : (c > xe + 2 * NO - 2 ? xe + 2 * NO - 2 : c)); };
        auto cl = [&](int c) { return colc(c < xs - NO ? xs - NO : (c > xe + NO - 1 ? xe + NO - 1 : c)); };
        auto ct = [&](int c) { return c < xs ? xs : (c > xe - 1 ? xe - 1 : c); };
        const unsigned o = L.ofs;
        // A
        if (bc) {
            const bool rg = ring_cell(g, L, cu(ix + 2 * NO - 1));
            R.U[sl(PH, 2 * NO - 1)] = rg ? R.pu[PA] + R.pr[PA] : R.pu[PA];
        } else {
            R.U[sl(PH, 2 * NO - 1)] = R.pu[PA];
        }
        const float m1 = R.pl[PA], m0 = R.p0[PA];
        {
            const int cn = cu(ix + 2 * NO), cm = cl(ix + 1 + NO), c0 = ct(ix + 1), cx = ct(ix);
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0), sx = B.soff(cx);
            R.pu[PB] = ld(B.u, o, sn);
            if (bc && ring_cell(g, L, cn)) R.pr[PB] = ld(B.ring, o, sn);
            R.pl[PB] = ld(B.mu, o, sm);
            R.p0[PB] = ld(B.mu, o, s0);
            R.po[PA][0] = ld(B.bf, o, sx);
            R.po[PA][1] = ld(B.up, o, sx);
        }
        const float rM = frcp(m1);
        {
            const float hm = frcp(R.rM_prev + rM);
            float s = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) s += a.k.ai[j - 1] * (R.U[sl(PH, NO - 1 + j)] - R.U[sl(PH, NO - j)]);
            R.FX[sl(PH, NO - 1)] = hm * s;
        }
        R.rM_prev = rM;
        S.at(PA, UB, PAD + lane) = R.U[sl(PH, 0)];
        S.at(PA, RM, PAD + lane) = frcp(m0);
        // B
        lds_barrier();
        // C. z-flux at column ix, face (row, row+1)
        {
            float sf = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j)
                sf += a.k.ai[j - 1] * (S.at(PA, UB, PAD + lane + j) - S.at(PA, UB, PAD + lane + 1 - j));
            S.at(PA, FZ, PAD + lane) = sf * frcp(S.at(PA, RM, PAD + lane) + S.at(PA, RM, PAD + lane + 1));
        }
        // D. update of column ix-1
        {
            float X = 0.0f, Z = 0.0f;
#pragma unroll
            for (int k = 1; k <= NO; ++k) {
                X += a.k.aox[k - 1] * (R.FX[sl(PH, k - 2)] - R.FX[sl(PH, -k - 1)]);
                Z += a.k.aoz[k - 1] * (S.at(PB, FZ, PAD + lane + k - 1) - S.at(PB, FZ, PAD + lane - k));
            }
            const float nu = (2.0f * R.U[sl(PH, -1)] - R.po[PB][1]) + R.po[PB][0] * (X + Z);
            if (L.outrow && ix - 1 >= xs && ix - 1 < xe) st(B.uo, o, B.soff(ix - 1), nu);
        }
    }

    static __device__ __forceinline__ void run(const IntArgs<NO> &a, int bid)
    {
        __shared__ Smem S;
        const Geo &g = a.g;
        Lane L;
        L.lane = threadIdx.x;
        for (int i = L.lane; i < LW; i += NT)
            if (i < PAD || i >= NT + PAD)
                for (int q = 0; q < 2; ++q) {
                    S.at(q, UB, i) = 0.0f; S.at(q, RM, i) = 1.0f; S.at(q, FZ, i) = 0.0f;
                }
        // tile id -> tile: full-height tiles walking down z first, then the
        // last partial band of rows as 64-lane segments (as IntKernel2)
        const int nmain = a.nbz * a.nbx;
        int z0, xs, xe, xl, li = L.lane, span = NT;
        bool live = true;
        if (bid < nmain) {
            const int bzid = bid % a.nbz, bxid = bid / a.nbz;
            z0 = g.iz0 + bzid * TZ;
            xs = g.ix0 + bxid * a.lx;
            xe = min(xs + a.lx, g.ix1 + 1);
            xl = xe;
        } else {
            const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
            xs = g.ix0 + ((bid - nmain) * (NT / 64) + w) * a.lx;
            if (xs > g.ix1) { live = false; xs = g.ix1; }
            xe = min(xs + a.lx, g.ix1 + 1);
            xl = xs + a.lx;
            z0 = a.zseg;
            li = L.lane & 63;
            span = 64;
        }
        L.r = z0 - HZ + li;
        const int rc = L.r < 1 ? 1 : (L.r > g.NZ ? g.NZ : L.r);
        L.ofs = (unsigned)(rc - 1) * 4u;
        L.outrow = live && (li >= HZ) && (li < span - HZ) && (L.r <= g.iz1);
        L.rr = L.r <= g.nz + g.nbd;
        L.rb = L.r > g.nz && L.rr;
        const bool bc = a.corr && (xs - WARM + 2 * NO - 1 < 1 || xe + 2 * NO - 1 > g.nx || z0 - HZ + span - 1 > g.nz);
        Regs R;
#pragma unroll
        for (int i = 0; i < P; ++i) { R.U[i] = 0.0f; R.FX[i] = 0.0f; }
        R.rM_prev = 1.0f;
        R.pr[0] = 0.0f; R.pr[1] = 0.0f;
        const int x0 = xs - WARM;
        Rsrc B;
        {
            const int lo = 1 - g.npml;
            B.xb = x0 < lo ? lo : x0;
            B.pitch = g.ld * 4;
            B.u = mk(g, a.ux, B.xb); B.up = mk(g, a.uxp, B.xb); B.uo = mk(g, a.uxo, B.xb);
            B.bf = mk(g, a.bf, B.xb); B.mu = mk(g, a.mu, B.xb); B.ring = mk(g, a.ringx, B.xb);
        }
        {
            const int lo = 1 - g.npml;
            const unsigned o = L.ofs;
            const int cn = x0 + 2 * NO - 1;
            const int cm = xs - NO > lo ? xs - NO : lo, c0 = xs;
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0);
            R.pu[0] = ld(B.u, o, sn);
            if (bc && ring_cell(g, L, cn)) R.pr[0] = ld(B.ring, o, sn);
            R.pl[0] = ld(B.mu, o, sm);
            R.p0[0] = ld(B.mu, o, s0);
            R.pu[1] = 0.0f; R.pl[1] = 1.0f; R.p0[1] = 1.0f;
            R.po[0][0] = R.po[0][1] = R.po[1][0] = R.po[1][1] = 0.0f;
        }
        __syncthreads();
        static_assert(P <= 26, "the step lists below cover ring periods up to 26");
        int base = x0;
        if constexpr (NO <= 4) {
            for (; base + P - 1 <= xl; base += P) {
#define IKS_STEP(PH) if constexpr (P > PH) step<PH>(a, B, S, R, L, bc, base + PH, xs, xe);
                IKS_STEP(0) IKS_STEP(1) IKS_STEP(2) IKS_STEP(3) IKS_STEP(4) IKS_STEP(5) IKS_STEP(6)
                IKS_STEP(7) IKS_STEP(8) IKS_STEP(9) IKS_STEP(10) IKS_STEP(11) IKS_STEP(12) IKS_STEP(13)
#undef IKS_STEP
            }
        }
#define IKS_STEP(PH) if constexpr (P > PH) { if (base + PH <= xl) step<PH>(a, B, S, R, L, bc, base + PH, xs, xe); }
#define IKS_GUARDED_PERIOD                                                                                   \
    IKS_STEP(0) IKS_STEP(1) IKS_STEP(2) IKS_STEP(3) IKS_STEP(4) IKS_STEP(5) IKS_STEP(6) IKS_STEP(7)        \
    IKS_STEP(8) IKS_STEP(9) IKS_STEP(10) IKS_STEP(11) IKS_STEP(12) IKS_STEP(13) IKS_STEP(14) IKS_STEP(15) \
    IKS_STEP(16) IKS_STEP(17) IKS_STEP(18) IKS_STEP(19) IKS_STEP(20) IKS_STEP(21) IKS_STEP(22) IKS_STEP(23) \
    IKS_STEP(24) IKS_STEP(25)
        if constexpr (NO <= 4) {
            if (base <= xl) { IKS_GUARDED_PERIOD }
        } else {
            for (; base <= xl; base += P) { IKS_GUARDED_PERIOD }
        }
#undef IKS_GUARDED_PERIOD
#undef IKS_STEP
    }
};

template <int NO, int NT>
__global__ __launch_bounds__(NT) void k_interior_sh(IntArgs<NO> a)
{
    IntKernelSH<NO, NT>::run(a, xcd_tile(blockIdx.x, gridDim.x));
}

}  // namespace fdfk
