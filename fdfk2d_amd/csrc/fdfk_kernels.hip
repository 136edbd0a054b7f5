// HIP kernels (gfx950) of the FDFK2D time step: the updatedispl role
// (src/FDFK2D.f90:516-522, binary-only in the reference) and the receiver
// recorder loadseis (src/Submain.f90:33-116).
//
// One step n -> n+1 is three launches on the handle's stream:
//   k_prep      background u0 on the hybrid ring at this step,     ring, strips,
//               first-derivative PML memories psi (level n) and,  receivers
//               on recording iterations, the receivers (level n)
//   k_interior2 full-order fused stencil + leapfrog over the       ~97% of cells
//               total-field region below the surface rows
//   k_frame     everything else (fdfk_frame.hip): free-surface     the frame
//               rows, reduced-order edge stencils, split-field
//               PML strips, hybrid TF/SF ring
//
// k_interior2 is the hot kernel: 64-wide wavefronts run along z (coalesced
// rows), each block marches along x keeping x-neighbourhoods in register
// rings (no x halo re-reads) and exchanging z-neighbourhoods through LDS.
// It reads u^n, u^{n-1}, dt^2/rho, lambda, mu once and writes u^{n+1} in
// place of u^{n-1}: 36 algorithmic bytes per cell-update in FP32.
//
// The recorder keeps the reference's FP64 interpolation weights and
// accumulation exactly.  The hybrid TF/SF correction enters as reads of the
// per-step ring field in the view of the output cell's side instead of a
// second stencil pass.
#include "fdfk_internal.h"

namespace fdfk {

// ---------------------------------------------------------------------
// pointwise operators with the reduced-order rule of the reference
// ---------------------------------------------------------------------
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float accf(float s, float c, float d) { return s + c * d; }
// harmonic mean 2/(1/a + 1/b); a zero modulus gives 0 (rcp(0) = inf)
__device__ __forceinline__ float hmean(float a, float b) { return 2.0f * frcp(frcp(a) + frcp(b)); }

// The pointwise operators below (PML memories and their corrections) run
// their stencil loops to the full order NO and give the terms beyond the
// reduced order of the cell a zero coefficient.  Those terms may address
// cells beyond the padded grid: addresses are clamped to the grid (the
// values are finite and multiplied by zero).
__device__ __forceinline__ size_t cellc(const Geo &g, int iz, int ix)
{
    ix = ix < 1 - g.npml ? 1 - g.npml : (ix > g.nx + g.npml ? g.nx + g.npml : ix);
    iz = iz < 1 ? 1 : (iz > g.NZ ? g.NZ : iz);
    return cell(g, iz, ix);
}
// coefficient j of order n from a (NO x NO) table; 0 for j > n (and n < 1)
template <int NO>
__device__ __forceinline__ float coefz(const float *tab, int j, int n)
{
    const int nn = n < 1 ? 1 : n;
    const float c = tab[(j - 1) + NO * (nn - 1)];
    return j <= n ? c : 0.0f;
}

struct FieldRd {
    const float *p;
    const Geo *g;
    __device__ float operator()(int iz, int ix) const { return p[cellc(*g, iz, ix)]; }
};

// background cell offset in the twelve reference arrays (FDFK2D.f90:330-355)
// arr: 0 left, 1 right, 2 bottom; which: 0 inside (reg), 1 outside (pml)
__device__ __forceinline__ long bg_off(const Geo &g, int iz, int ix, int it, int &arr, int &which)
{
    const int nx = g.nx, nz = g.nz, nbd = g.nbd;
    if (iz >= 1 && iz <= nz - nbd && ix >= 1 && ix <= nbd) {
        arr = 0; which = 0;
        return (long)(iz - 1) + (long)(nz - nbd) * ((long)(ix - 1) + (long)nbd * it);
    }
    if (iz >= 1 && iz <= nz - nbd && ix >= nx - nbd + 1 && ix <= nx) {
        arr = 1; which = 0;
        return (long)(iz - 1) + (long)(nz - nbd) * ((long)(ix - (nx - nbd + 1)) + (long)nbd * it);
    }
    if (iz >= nz - nbd + 1 && iz <= nz && ix >= 1 && ix <= nx) {
        arr = 2; which = 0;
        return (long)(iz - (nz - nbd + 1)) + (long)nbd * ((long)(ix - 1) + (long)nx * it);
    }
    if (iz >= 1 && iz <= nz && ix >= 1 - nbd && ix <= 0) {
        arr = 0; which = 1;
        return (long)(iz - 1) + (long)nz * ((long)(ix - (1 - nbd)) + (long)nbd * it);
    }
    if (iz >= 1 && iz <= nz && ix >= nx + 1 && ix <= nx + nbd) {
        arr = 1; which = 1;
        return (long)(iz - 1) + (long)nz * ((long)(ix - (nx + 1)) + (long)nbd * it);
    }
    if (iz >= nz + 1 && iz <= nz + nbd && ix >= 1 - nbd && ix <= nx + nbd) {
        arr = 2; which = 1;
        return (long)(iz - (nz + 1)) + (long)nbd * ((long)(ix - (1 - nbd)) + (long)(nx + 2 * nbd) * it);
    }
    return -1;
}

// side of the hybrid interface: 0 total-field region (1..nx x 1..nz), 1 padding
__device__ __forceinline__ int tf_side(const Geo &g, int iz, int ix)
{
    return (ix >= 1 && ix <= g.nx && iz <= g.nz) ? 0 : 1;
}

template <int NO, class R>
__device__ __forceinline__ float cdx_sum(const Geo &g, const Coefs *cf, const R &f, int iz, int ix)
{
    const int nox = imin(imin(ix + g.npml - 1, g.nx + g.npml - ix), NO);
    float s = 0.0f;
#pragma unroll
    for (int k = 1; k <= NO; ++k) s = accf(s, coefz<NO>(cf->c1f, k, nox), f(iz, ix + k) - f(iz, ix - k));
    return s;
}
template <int NO, class R>
__device__ __forceinline__ float cdz_sum(const Geo &g, const Coefs *cf, const R &f, int iz, int ix)
{
    const int noz = imin(imin(iz - 1, g.nz + g.npml - iz), NO);
    float s = 0.0f;
#pragma unroll
    for (int k = 1; k <= NO; ++k) s = accf(s, coefz<NO>(cf->c1f, k, noz), f(iz + k, ix) - f(iz - k, ix));
    return s;
}

__device__ __forceinline__ int strip_of(const Geo &g, int iz, int ix)
{
    if (iz >= 1 && iz <= g.nz) {
        if (ix <= 0) return S_L;
        if (ix > g.nx) return S_R;
        return S_NONE;
    }
    return S_B;
}
__device__ __forceinline__ size_t strip_off(const Geo &g, int s, int iz, int ix)
{
    if (s == S_L) return (size_t)(ix - (1 - g.npml)) * g.nz + (size_t)(iz - 1);
    if (s == S_R) return (size_t)(ix - (g.nx + 1)) * g.nz + (size_t)(iz - 1);
    return (size_t)(ix - (1 - g.npml)) * g.npml + (size_t)(iz - (g.nz + 1));
}

// background on the hybrid ring at iteration it, scattered into the
// full-grid ring arrays (positions are fixed; values change every step)
__device__ __forceinline__ void ring_prep_body(const Geo &g, const Background *bg, int it, float *ringx,
                                               float *ringz, long bid)
{
    // one thread per ring cell, walking the six background arrays in their
    // own (contiguous) order so the reads of step it coalesce
    long o = bid * blockDim.x + threadIdx.x;
    long n[6];
    ring_sizes(g, n);
    const int nx = g.nx, nz = g.nz, nbd = g.nbd;
    int k = 0;
    while (k < 6 && o >= n[k]) { o -= n[k]; ++k; }
    if (k == 6) return;
    int iz, ix;
    switch (k) {
    case 0: iz = 1 + (int)(o % (nz - nbd)); ix = 1 + (int)(o / (nz - nbd)); break;
    case 1: iz = 1 + (int)(o % (nz - nbd)); ix = nx - nbd + 1 + (int)(o / (nz - nbd)); break;
    case 2: iz = nz - nbd + 1 + (int)(o % nbd); ix = 1 + (int)(o / nbd); break;
    case 3: iz = 1 + (int)(o % nz); ix = 1 - nbd + (int)(o / nz); break;
    case 4: iz = 1 + (int)(o % nz); ix = nx + 1 + (int)(o / nz); break;
    default: iz = nz + 1 + (int)(o % nbd); ix = 1 - nbd + (int)(o / nbd); break;
    }
    const int which = k / 3, arr = k % 3;
    const long src = o + n[k] * (long)(it % bg->period);
    const size_t c = cell(g, iz, ix);
    ringx[c] = bg->a[6 * which + 2 * arr][src];
    ringz[c] = bg->a[6 * which + 2 * arr + 1][src];
}


// ---------------------------------------------------------------------
// PML first-derivative memories (phase A of the step)
// ---------------------------------------------------------------------
template <int NO>
__device__ __forceinline__ void psi_body(const StepArgs &a, long bid)
{
    const Geo &g = a.g;
    long t = bid * blockDim.x + threadIdx.x;
    long nlr = (long)g.npml * g.nz;
    long nb = (long)g.NX * g.npml;
    int s, iz, ix;
    if (t < nlr) { s = S_L; ix = (1 - g.npml) + (int)(t / g.nz); iz = 1 + (int)(t % g.nz); }
    else if (t < 2 * nlr) { t -= nlr; s = S_R; ix = g.nx + 1 + (int)(t / g.nz); iz = 1 + (int)(t % g.nz); }
    else if (t < 2 * nlr + nb) { t -= 2 * nlr; s = S_B; ix = (1 - g.npml) + (int)(t / g.npml); iz = g.nz + 1 + (int)(t % g.npml); }
    else return;
    const size_t o = strip_off(g, s, iz, ix);
    FieldRd rx{a.ux, &g}, rz{a.uz, &g};
    const int jx = ix + g.npml - 1;
    const Strips *st = a.st;
    if (st->dpx[jx] > 0.0f) {
        float b = st->bx[jx], c = 1.0f - b;
        float gx = cdx_sum<NO>(g, a.cf, rx, iz, ix) / g.dx;
        float gz = cdx_sum<NO>(g, a.cf, rz, iz, ix) / g.dx;
        float *p0 = st->psx[s][0], *p1 = st->psx[s][1];
        p0[o] = b * p0[o] + c * gx;
        p1[o] = b * p1[o] + c * gz;
    }
    if (s == S_B && st->dpz[iz - 1] > 0.0f) {
        float b = st->bz[iz - 1], c = 1.0f - b;
        float gx = cdz_sum<NO>(g, a.cf, rx, iz, ix) / g.dz;
        float gz = cdz_sum<NO>(g, a.cf, rz, iz, ix) / g.dz;
        float *p0 = st->psz[s][0], *p1 = st->psz[s][1];
        p0[o] = b * p0[o] + c * gx;
        p1[o] = b * p1[o] + c * gz;
    }
}


// ---------------------------------------------------------------------
// receivers: 7x7 discretised-Dirac interpolation (Submain.f90:3-116)
// ---------------------------------------------------------------------
__device__ double dirac_w(double x)
{
#pragma clang fp contract(off)
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


// One wave per receiver (blockDim.x / 64 receivers per block): lane l < 49
// forms the weight and loads the field of stencil point l, then lane 0
// accumulates the 49 products in the reference's loop order (ix outer, iz
// inner) with FP64 products rounded to FP32 after every term, so the sum is
// the reference's bit for bit.
constexpr int kRecPerBlock = 4;   // receivers per block in the fused prep kernel
__device__ __forceinline__ void record_body(const RecArgs &r, long bid, int istp)
{
#pragma clang fp contract(off)
    __shared__ double wx[kRecPerBlock][64], wz[kRecPerBlock][64];
    const int w = threadIdx.x / 64;
    const int i = (int)bid * (int)(blockDim.x / 64) + w;
    const int l = threadIdx.x % 64;
    if (i >= r.nrcv) return;   // wave-uniform; the exchange below stays inside the wave
    const Geo &g = r.g;
    const float rx = r.rx[i], rz = r.rz[i], x0 = r.x0, z0 = r.z0, dx = g.dx, dz = g.dz;
    int xa, xb, za, zb;
    const int npt = rec_window(g, x0, z0, rx, rz, xa, xb, za, zb);   // 0..49
    const int nzr = zb - za + 1;
    if (l < npt) {
        const int ix = xa + l / nzr, iz = za + l % nzr;
        double x = fabs((double)(x0 + (float)(ix - 1) * dx) - (double)rx) / (double)dx;
        double z = fabs((double)(z0 + (float)(iz - 1) * dz) - (double)rz) / (double)dz;
        double ba = dirac_w(x) * dirac_w(z);
        size_t c = cell(g, iz, ix);
        wx[w][l] = ba * (double)r.ux[c];
        wz[w][l] = ba * (double)r.uz[c];
    }
    // the wave's own LDS writes are complete and visible to its lanes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (l == 0) {
        float sx = 0.0f, sz = 0.0f;
        for (int k = 0; k < npt; ++k) {
            sx = (float)((double)sx + wx[w][k]);
            sz = (float)((double)sz + wz[w][k]);
        }
        r.seisx[(size_t)i * r.nsamp + istp] = sx;
        r.seisz[(size_t)i * r.nsamp + istp] = sz;
    }
}
// The three small per-step kernels in one launch (they are independent:
// the ring gather writes the ring field, the PML memories read u^n and
// write psi, the recorder reads u^n): blocks [0, nring) gather the ring,
// [nring, nring + npsi) update the memories, the rest record receivers.
template <int NO>
__global__ __launch_bounds__(256) void k_prep(StepArgs a, RecArgs r, const Background *bg, int nring, int npsi)
{
    int b = blockIdx.x;
    if (b < nring) { ring_prep_body(a.g, bg, a.it, (float *)a.ringx, (float *)a.ringz, b); return; }
    b -= nring;
    if (b < npsi) { psi_body<NO>(a, b); return; }
    record_body(r, b - npsi, r.istp);
}

#pragma clang fp contract(on)

// word 3 of a gfx9 buffer resource (32-bit raw data, no swizzle)
constexpr int kBufferFlags = 0x00020000;

// Workgroup barrier for LDS exchange only.  __syncthreads() also waits for
// every outstanding global load (gfx9 counts loads and stores in one
// counter), which would drain the one-column-ahead prefetch of the marching
// kernels at every column.  Their lanes never read global data written by
// another lane of the same launch, so waiting for the wave's own LDS
// operations before the barrier is sufficient; the "memory" clobber keeps
// the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// MI355X hands the workgroups of a launch to its 8 XCDs round-robin
// (workgroup p runs on XCD p % 8), and each XCD has its own L2.  This maps
// the hardware id to a tile id so that every XCD owns one contiguous range
// of tile ids: tiles that read the same cache lines (the halo rows shared by
// z-neighbours) then meet in one L2 instead of being fetched from HBM twice.
// A bijection on [0, n) for any n.
constexpr int kXcds = 8;

// Occupancy target of the kernels that run the interior tiles: 3 waves per
// SIMD (at most 168 VGPRs) up to FD order 6; the rings of the higher orders
// need more registers and run at 2 waves per SIMD (orders 8..12) or 1
// (orders 14, 16) rather than spill.
template <int NO>
constexpr int kIntWaves = NO <= 3 ? 3 : (NO <= 6 ? 2 : 1);
__device__ __forceinline__ int xcd_tile(int p, int n)
{
    const int x = p % kXcds, k = p / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------
// interior kernel: full-order fused stencil over the total-field region.
// Lanes run along z (one row each, HZ halo rows on either side of the TZ
// output rows), the block marches along x one column per step.  Per step:
//   A  take the prefetched column ix+2NO-1 into the register ring, prefetch
//      column ix+2NO and the materials of the next step, form the x-flux at
//      the face (ix+NO-1, ix+NO) and stage u(ix+NO), u(ix), 1/M(ix) in LDS;
//   B  one barrier (LDS double-buffered by column parity);
//   C  l Dz uz, m Dz ux at ix+NO (x-mixed ring), z-fluxes and l Dx ux,
//      m Dx uz at ix (to LDS);
//   D  the update of column ix-1, whose z-terms were staged one step before.
// The steady state has no branches: every lane loads (rows clamped into the
// grid; rows beyond the stencil reach of an output are never used) and only
// the stores are predicated.  Addresses are a per-column scalar base plus a
// loop-invariant 32-bit lane offset.
// ---------------------------------------------------------------------
// M16: lam and mu are read as 16-bit residuals against per-row FP32
// profiles, FP16 (M16 = 1) or BF16 (M16 = 2) (half the bytes; the FP32
// arrays hold the same rounded values for the other kernels)
template <int NO, int NT, int M16 = 0>
struct IntKernel2 {
    static constexpr int HZ = 2 * NO - 1;
    static constexpr int TZ = NT - 2 * HZ;      // output rows per block
    static constexpr int P0 = 3 * NO + 1;
    static constexpr int P = P0 + (P0 & 1);     // even ring period: LDS parity is compile-time
    static constexpr int PAD = NO + 1;
    static constexpr int WARM = 4 * NO - 2;
    static constexpr int LW = NT + 2 * PAD;

    // LDS, lane-major: row r of the block holds SW words, ten per column
    // parity (uA, uB, rM, fz, vz for the two components); the odd stride
    // keeps the 64 lanes of a wavefront on distinct banks, and every access
    // of a step is one base register plus an immediate offset
    //   uA: u at column ix+NO   uB: u at column ix   rM: 1/(l+2m), 1/m at ix
    //   fz: z-fluxes at ix      vz: m*Dx uz, l*Dx ux at ix
    enum { UA = 0, UB = 2, RM = 4, FZ = 6, VZ = 8 };
    static constexpr int SW = 21;
    struct Smem {
        float m[LW * SW];
        __device__ __forceinline__ float &at(int par, int q, int c, int i) { return m[i * SW + par * 10 + q + c]; }
    };

    static __device__ __forceinline__ int sl(int ph, int j) { return ((ph + j) % P + P) % P; }

    struct Regs {
        float U[2][P], FX[2][P], VX[2][P];
        float rMa_prev, rMm_prev;
        float pu[2][2];     // [parity][comp] next new u column
        float pl[2][2];     // [parity][l,m] next lead materials (ix+NO)
        float p0[2][2];     // [parity][l,m] next column materials (ix)
        float po[2][3];     // [parity][bf, uxp, uzp] of column ix (used one step later)
        float pr[2][2];     // [parity][comp] u0 of the next new column on the outer ring
    };

    struct Lane {
        int lane, r;
        unsigned ofs;       // clamped row byte offset inside a column
        bool outrow;
        bool rb, rr;        // row on the bottom outer ring / not below it
    };
    // Padding cells are read in the total-field view, u + u0.  u0 is nonzero
    // only on the outer ring (nbd cells wide beside and below the physical
    // region), so only those cells load it; c is step-uniform.
    static __device__ __forceinline__ bool ring_cell(const Geo &g, const Lane &L, int c)
    {
        if (c >= 1 && c <= g.nx) return L.rb;
        return c >= 1 - g.nbd && c <= g.nx + g.nbd && L.rr;
    }
    // Buffer resources on every array, based at the block's first column xb
    // (built once per block); an element is rsrc + scalar column offset
    // (c - xb) * ld * 4 + the lane's 32-bit row offset.  Offsets stay far
    // below 4 GiB because a block spans lx + 6*NO columns.
    struct Rsrc {
        __amdgpu_buffer_rsrc_t ux, uz, uxp, uzp, uxo, uzo, bf, lam, mu, rx, rz, lam16, mu16;
        float plam, pmu;   // M16: the lane row's material profile
        int xb, pitch;   // first column, column pitch in bytes
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
    static __device__ __forceinline__ __amdgpu_buffer_rsrc_t mk16(const Geo &g, const unsigned short *base, int xb)
    {
        const unsigned short *cb = base + (size_t)(xb + g.npml - 1) * (size_t)g.ld;
        return __builtin_amdgcn_make_buffer_rsrc((void *)cb, 0, 0x7fffffff, kBufferFlags);
    }
    // a material value at the lane's row of a column: byte offsets of the
    // 32-bit arrays, halved for the 16-bit copies
    // (M16: the row's profile value plus the scaled 16-bit residual)
    static __device__ __forceinline__ float h16(unsigned short h)
    {
        if constexpr (M16 == 2)
            return __uint_as_float((unsigned)h << 16);
        else
            return (float)__builtin_bit_cast(_Float16, h);
    }
    static __device__ __forceinline__ float ld_lam(const IntArgs<NO> &a, const Rsrc &B, unsigned o, int so)
    {
        if constexpr (M16 != 0)
            return B.plam + h16(__builtin_amdgcn_raw_buffer_load_b16(B.lam16, o >> 1, so >> 1, 0)) * a.sl;
        else
            return ld(B.lam, o, so);
    }
    static __device__ __forceinline__ float ld_mu(const IntArgs<NO> &a, const Rsrc &B, unsigned o, int so)
    {
        if constexpr (M16 != 0)
            return B.pmu + h16(__builtin_amdgcn_raw_buffer_load_b16(B.mu16, o >> 1, so >> 1, 0)) * a.sm;
        else
            return ld(B.mu, o, so);
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
        // Columns whose loads matter: u within the x-reach 2NO-1 of the
        // tile, lead materials within NO (face means, Dx of l Dz uz), column
        // materials and level n-1 on the tile only.  Loads outside are
        // redirected to the nearest such column (a cache hit) instead of
        // fetching data that only feeds unstored updates.
        auto cu = [&](int c) { return colc(c < xs - (2 * NO - 1) ? xs - (2 * NO - 1) : (c > xe + 2 * NO - 2 ? xe + 2 * NO - 2 : c)); };
        auto cl = [&](int c) { return colc(c < xs - NO ? xs - NO : (c > xe + NO - 1 ? xe + NO - 1 : c)); };
        auto ct = [&](int c) { return c < xs ? xs : (c > xe - 1 ? xe - 1 : c); };
        const unsigned o = L.ofs;
        // A. ring update, prefetch of the next step, x-flux, LDS staging.
        // The u0 of the new column is added here, one step after its load,
        // so that its latency stays hidden like every other load's.
        if (bc) {
            const bool rg = ring_cell(g, L, cu(ix + 2 * NO - 1));
            R.U[0][sl(PH, 2 * NO - 1)] = rg ? R.pu[PA][0] + R.pr[PA][0] : R.pu[PA][0];
            R.U[1][sl(PH, 2 * NO - 1)] = rg ? R.pu[PA][1] + R.pr[PA][1] : R.pu[PA][1];
        } else {
            R.U[0][sl(PH, 2 * NO - 1)] = R.pu[PA][0];
            R.U[1][sl(PH, 2 * NO - 1)] = R.pu[PA][1];
        }
        const float l1 = R.pl[PA][0], m1 = R.pl[PA][1];
        const float l0 = R.p0[PA][0], m0 = R.p0[PA][1];
        {
            const int cn = cu(ix + 2 * NO), cm = cl(ix + 1 + NO), c0 = ct(ix + 1), cx = ct(ix);
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0), sx = B.soff(cx);
            R.pu[PB][0] = ld(B.ux, o, sn);
            R.pu[PB][1] = ld(B.uz, o, sn);
            if (bc && ring_cell(g, L, cn)) {
                R.pr[PB][0] = ld(B.rx, o, sn);
                R.pr[PB][1] = ld(B.rz, o, sn);
            }
            R.pl[PB][0] = ld_lam(a, B, o, sm); R.pl[PB][1] = ld_mu(a, B, o, sm);
            R.p0[PB][0] = ld_lam(a, B, o, s0); R.p0[PB][1] = ld_mu(a, B, o, s0);
            R.po[PA][0] = ld(B.bf, o, sx);
            R.po[PA][1] = ld(B.uxp, o, sx);
            R.po[PA][2] = ld(B.uzp, o, sx);
        }
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
        S.at(PA, UA, 0, PAD + lane) = R.U[0][sl(PH, NO)];
        S.at(PA, UA, 1, PAD + lane) = R.U[1][sl(PH, NO)];
        S.at(PA, UB, 0, PAD + lane) = R.U[0][sl(PH, 0)];
        S.at(PA, UB, 1, PAD + lane) = R.U[1][sl(PH, 0)];
        S.at(PA, RM, 0, PAD + lane) = frcp(l0 + 2.0f * m0);
        S.at(PA, RM, 1, PAD + lane) = frcp(m0);
        // B.
        lds_barrier();
        // C. Vx at ix+NO; z-fluxes and Vz at ix
        {
            float dzz = 0.0f, dzx = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                dzz += a.k.bi[j - 1] * (S.at(PA, UA, 1, PAD + lane + j) - S.at(PA, UA, 1, PAD + lane - j));
                dzx += a.k.bi[j - 1] * (S.at(PA, UA, 0, PAD + lane + j) - S.at(PA, UA, 0, PAD + lane - j));
            }
            R.VX[0][sl(PH, NO)] = l1 * dzz;
            R.VX[1][sl(PH, NO)] = m1 * dzx;
        }
        {
            float sfx = 0.0f, sfz = 0.0f;
#pragma unroll
            for (int j = 1; j <= NO; ++j) {
                sfx += a.k.ai[j - 1] * (S.at(PA, UB, 0, PAD + lane + j) - S.at(PA, UB, 0, PAD + lane + 1 - j));
                sfz += a.k.ai[j - 1] * (S.at(PA, UB, 1, PAD + lane + j) - S.at(PA, UB, 1, PAD + lane + 1 - j));
            }
            const float rMa0 = S.at(PA, RM, 0, PAD + lane), rMm0 = S.at(PA, RM, 1, PAD + lane);
            const float fzx = sfx * frcp(rMm0 + S.at(PA, RM, 1, PAD + lane + 1));
            const float fzz = sfz * frcp(rMa0 + S.at(PA, RM, 0, PAD + lane + 1));
            float dxz = 0.0f, dxx = 0.0f;
#pragma unroll
            for (int k = 1; k <= NO; ++k) {
                dxz += a.k.bi[k - 1] * (R.U[1][sl(PH, k)] - R.U[1][sl(PH, -k)]);
                dxx += a.k.bi[k - 1] * (R.U[0][sl(PH, k)] - R.U[0][sl(PH, -k)]);
            }
            S.at(PA, FZ, 0, PAD + lane) = fzx;
            S.at(PA, FZ, 1, PAD + lane) = fzz;
            S.at(PA, VZ, 0, PAD + lane) = m0 * dxz;
            S.at(PA, VZ, 1, PAD + lane) = l0 * dxx;
        }
        // D. update of column ix-1 (its z-terms were staged in the previous step)
        {
            float Xx = 0.0f, Xz = 0.0f, Zx = 0.0f, Zz = 0.0f, Mx = 0.0f, Mz = 0.0f;
#pragma unroll
            for (int k = 1; k <= NO; ++k) {
                Xx += a.k.aox[k - 1] * (R.FX[0][sl(PH, k - 2)] - R.FX[0][sl(PH, -k - 1)]);
                Xz += a.k.aox[k - 1] * (R.FX[1][sl(PH, k - 2)] - R.FX[1][sl(PH, -k - 1)]);
                Zx += a.k.aoz[k - 1] * (S.at(PB, FZ, 0, PAD + lane + k - 1) - S.at(PB, FZ, 0, PAD + lane - k));
                Zz += a.k.aoz[k - 1] * (S.at(PB, FZ, 1, PAD + lane + k - 1) - S.at(PB, FZ, 1, PAD + lane - k));
                Mx += a.k.bo[k - 1] * ((R.VX[0][sl(PH, k - 1)] - R.VX[0][sl(PH, -k - 1)]) +
                                      (S.at(PB, VZ, 0, PAD + lane + k) - S.at(PB, VZ, 0, PAD + lane - k)));
                Mz += a.k.bo[k - 1] * ((R.VX[1][sl(PH, k - 1)] - R.VX[1][sl(PH, -k - 1)]) +
                                      (S.at(PB, VZ, 1, PAD + lane + k) - S.at(PB, VZ, 1, PAD + lane - k)));
            }
            const float bf = R.po[PB][0];
            const float nx_ = (2.0f * R.U[0][sl(PH, -1)] - R.po[PB][1]) + bf * ((Xx + Zx) + Mx);
            const float nz_ = (2.0f * R.U[1][sl(PH, -1)] - R.po[PB][2]) + bf * ((Xz + Zz) + Mz);
            if (L.outrow && ix - 1 >= xs && ix - 1 < xe) {
                st(B.uxo, o, B.soff(ix - 1), nx_);
                st(B.uzo, o, B.soff(ix - 1), nz_);
            }
        }
    }

    static __device__ __forceinline__ void run(const IntArgs<NO> &a, int bid)
    {
        __shared__ Smem S;
        run(a, bid, S);
    }
    static __device__ __forceinline__ void run(const IntArgs<NO> &a, int bid, Smem &S)
    {
        const Geo &g = a.g;
        Lane L;
        L.lane = threadIdx.x;
        for (int i = L.lane; i < LW; i += NT)
            if (i < PAD || i >= NT + PAD)
                for (int q = 0; q < 2; ++q)
                    for (int c = 0; c < 2; ++c) {
                        S.at(q, UA, c, i) = 0.0f; S.at(q, UB, c, i) = 0.0f; S.at(q, RM, c, i) = 1.0f;
                        S.at(q, FZ, c, i) = 0.0f; S.at(q, VZ, c, i) = 0.0f;
                    }
        // tile id -> tile.  Ids below nbz*nbx are full-height tiles;
        // consecutive ids walk down z first, so the z neighbours of one
        // column of tiles share an XCD (xcd_tile).  A last band of rows too
        // short for a full tile (at most 64 - 2HZ rows) is cut into 64-lane
        // segments instead, each wave of a block marching its own columns:
        // a full-height tile there would spend every step on mostly idle
        // lanes.  Lanes of a segment read their LDS neighbours across the
        // segment's edge only from halo lanes, whose results are not stored.
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
            if (xs > g.ix1) { live = false; xs = g.ix1; }   // past the last column: march, store nothing
            xe = min(xs + a.lx, g.ix1 + 1);
            xl = xs + a.lx;   // the same step count for every wave: one barrier per step
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
        // tiles whose loads reach the padding read the ring (wave-uniform)
        const bool bc = a.corr && (xs - WARM + 2 * NO - 1 < 1 || xe + 2 * NO - 1 > g.nx || z0 - HZ + span - 1 > g.nz);
        Regs R;
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int i = 0; i < P; ++i) { R.U[q][i] = 0.0f; R.FX[q][i] = 0.0f; R.VX[q][i] = 0.0f; }
        R.rMa_prev = 1.0f; R.rMm_prev = 1.0f;
#pragma unroll
        for (int q = 0; q < 2; ++q) { R.pr[q][0] = 0.0f; R.pr[q][1] = 0.0f; }
        const int x0 = xs - WARM;
        Rsrc B;
        {
            const int lo = 1 - g.npml;
            B.xb = x0 < lo ? lo : x0;   // every column a step touches is clamped to >= lo
            B.pitch = g.ld * 4;
            B.ux = mk(g, a.ux, B.xb); B.uz = mk(g, a.uz, B.xb);
            B.uxp = mk(g, a.uxp, B.xb); B.uzp = mk(g, a.uzp, B.xb);
            B.uxo = mk(g, a.uxo, B.xb); B.uzo = mk(g, a.uzo, B.xb);
            B.bf = mk(g, a.bf, B.xb); B.lam = mk(g, a.lam, B.xb); B.mu = mk(g, a.mu, B.xb);
            if constexpr (M16 != 0) {
                B.lam16 = mk16(g, a.lam16, B.xb); B.mu16 = mk16(g, a.mu16, B.xb);
                B.plam = a.lamp[rc - 1]; B.pmu = a.mup[rc - 1];
            }
            B.rx = mk(g, a.ringx, B.xb); B.rz = mk(g, a.ringz, B.xb);
        }
        // prologue: the loads the first step consumes
        {
            const int lo = 1 - g.npml;
            const unsigned o = L.ofs;
            const int cn = x0 + 2 * NO - 1;   // = xs - (2NO-1), the first u column the tile reaches
            const int cm = xs - NO > lo ? xs - NO : lo, c0 = xs;   // step x0 feeds no stored update
            const int sn = B.soff(cn), sm = B.soff(cm), s0 = B.soff(c0);
            R.pu[0][0] = ld(B.ux, o, sn); R.pu[0][1] = ld(B.uz, o, sn);
            if (bc && ring_cell(g, L, cn)) {
                R.pr[0][0] = ld(B.rx, o, sn);
                R.pr[0][1] = ld(B.rz, o, sn);
            }
            R.pl[0][0] = ld_lam(a, B, o, sm); R.pl[0][1] = ld_mu(a, B, o, sm);
            R.p0[0][0] = ld_lam(a, B, o, s0); R.p0[0][1] = ld_mu(a, B, o, s0);
#pragma unroll
            for (int q = 0; q < 2; ++q) { R.pu[1][q] = 0.0f; R.pl[1][q] = 1.0f; R.p0[1][q] = 1.0f; }
#pragma unroll
            for (int q = 0; q < 3; ++q) { R.po[0][q] = 0.0f; R.po[1][q] = 0.0f; }
        }
        __syncthreads();
        // whole ring periods without per-step guards (no merges, so the
        // ring registers need no copies), then the guarded remainder.  FD
        // orders above 8 keep only the guarded loop: a second copy of their
        // long ring period would double an already large body.
        static_assert(P <= 26, "the step lists below cover ring periods up to 26");
        int base = x0;
        if constexpr (NO <= 4) {
            for (; base + P - 1 <= xl; base += P) {
#define IK2_STEP(PH) if constexpr (P > PH) step<PH>(a, B, S, R, L, bc, base + PH, xs, xe);
                IK2_STEP(0) IK2_STEP(1) IK2_STEP(2) IK2_STEP(3) IK2_STEP(4) IK2_STEP(5) IK2_STEP(6)
                IK2_STEP(7) IK2_STEP(8) IK2_STEP(9) IK2_STEP(10) IK2_STEP(11) IK2_STEP(12) IK2_STEP(13)
#undef IK2_STEP
            }
        }
#define IK2_STEP(PH) if constexpr (P > PH) { if (base + PH <= xl) step<PH>(a, B, S, R, L, bc, base + PH, xs, xe); }
#define IK2_GUARDED_PERIOD                                                                                   \
    IK2_STEP(0) IK2_STEP(1) IK2_STEP(2) IK2_STEP(3) IK2_STEP(4) IK2_STEP(5) IK2_STEP(6) IK2_STEP(7)        \
    IK2_STEP(8) IK2_STEP(9) IK2_STEP(10) IK2_STEP(11) IK2_STEP(12) IK2_STEP(13) IK2_STEP(14) IK2_STEP(15) \
    IK2_STEP(16) IK2_STEP(17) IK2_STEP(18) IK2_STEP(19) IK2_STEP(20) IK2_STEP(21) IK2_STEP(22) IK2_STEP(23) \
    IK2_STEP(24) IK2_STEP(25)
        if constexpr (NO <= 4) {
            if (base <= xl) { IK2_GUARDED_PERIOD }   // fewer than P steps remain
        } else {
            for (; base <= xl; base += P) { IK2_GUARDED_PERIOD }
        }
#undef IK2_GUARDED_PERIOD
#undef IK2_STEP
    }
};

template <int NO, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kIntWaves<NO>))) void k_interior2(IntArgs<NO> a)
{
    IntKernel2<NO, NT>::run(a, xcd_tile(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------
// layout conversion helpers
// ---------------------------------------------------------------------
// reference (1:3, NZ, NX) -> bf = dt^2/rho, lam, mu in the padded layout
// The damped part of the PML (more than nbd cells into the padding) takes
// the material of the outermost undamped cell along the normal: the
// absorbing layers are normal-invariant, which keeps heterogeneous damped
// corners stable.
__global__ void k_import_material(Geo g, const float *c, float dt2, float *bf, float *lam, float *mu)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * g.NZ;
    if (t >= n) return;
    int j = (int)(t / g.NZ), k = (int)(t % g.NZ);
    int ix = j + 1 - g.npml, iz = k + 1;
    int sx = ix < 1 - g.nbd ? 1 - g.nbd : (ix > g.nx + g.nbd ? g.nx + g.nbd : ix);
    int sz = iz > g.nz + g.nbd ? g.nz + g.nbd : iz;
    long src = (long)(sx + g.npml - 1) * g.NZ + (sz - 1);
    size_t d = (size_t)j * g.ld + k;
    bf[d] = dt2 / c[3 * src + 0];
    lam[d] = c[3 * src + 1];
    mu[d] = c[3 * src + 2];
}

// 16-bit material storage: lam = lamp(row) + h * sl with h the FP16 (bf = 0)
// or BF16 (bf = 1) residual against a per-row FP32 profile (power-of-two
// scale, nearest even), so the rounding falls on the lateral variation
// only; the FP32 arrays receive the same rounded values (the product h * sl
// is exact, the sum rounds once)
__device__ __forceinline__ unsigned short round16(float x, int bf, float &v)
{
    if (bf) {
        unsigned u = __float_as_uint(x);
        u += 0x7FFFu + ((u >> 16) & 1u);   // nearest even (finite values)
        const unsigned short h = (unsigned short)(u >> 16);
        v = __uint_as_float((unsigned)h << 16);
        return h;
    }
    const _Float16 h = (_Float16)x;
    v = (float)h;
    return __builtin_bit_cast(unsigned short, h);
}
__global__ void k_round_material(Geo g, float *lam, float *mu, unsigned short *lam16, unsigned short *mu16,
                                 const float *lamp, const float *mup, float sl, float sm, int bf)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * g.NZ;
    if (t >= n) return;
    const int k = (int)(t % g.NZ);
    const size_t d = (size_t)(t / g.NZ) * g.ld + (size_t)k;
    float vl, vm;
    lam16[d] = round16((lam[d] - lamp[k]) * (1.0f / sl), bf, vl);
    mu16[d] = round16((mu[d] - mup[k]) * (1.0f / sm), bf, vm);
    lam[d] = lamp[k] + vl * sl;
    mu[d] = mup[k] + vm * sm;
}

// reference (1:2, NZ, NX) <-> two padded arrays (level 1 -> cur, level 2 -> prev)
__global__ void k_import_levels(Geo g, const float *ref2, float *cur, float *prev)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * g.NZ;
    if (t >= n) return;
    int j = (int)(t / g.NZ), k = (int)(t % g.NZ);
    size_t d = (size_t)j * g.ld + k;
    cur[d] = ref2[2 * t];
    prev[d] = ref2[2 * t + 1];
}
__global__ void k_export_levels(Geo g, const float *cur, const float *prev, float *ref2)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * g.NZ;
    if (t >= n) return;
    int j = (int)(t / g.NZ), k = (int)(t % g.NZ);
    size_t d = (size_t)j * g.ld + k;
    ref2[2 * t] = cur[d];
    ref2[2 * t + 1] = prev[d];
}
// strip arrays: reference (1:2, cells) <-> cur/prev
__global__ void k_split_levels(long n, const float *ref2, float *cur, float *prev)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    cur[t] = ref2[2 * t];
    prev[t] = ref2[2 * t + 1];
}
__global__ void k_join_levels(long n, const float *cur, const float *prev, float *ref2)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    ref2[2 * t] = cur[t];
    ref2[2 * t + 1] = prev[t];
}

}  // namespace fdfk
