// Internal definitions shared by the HIP kernels and the host API of the
// FDFK2D time-stepping library (gfx950).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace fdfk {

constexpr int kMaxOrder = 8;

// Grid description.  Cell (ix, iz) with ix in 1-npml..nx+npml (reference
// column index) and iz in 1..nz+npml (free surface at iz = 1) lives at
//   (ix + npml - 1) * ld + (iz - 1)
// i.e. z is the contiguous (wavefront-coalesced) axis, like the reference's
// Fortran arrays, with the column pitch ld padded to a 128-byte multiple.
struct Geo {
    int nx, nz, npml, norder, nbd;
    int NX, NZ, ld;
    float dx, dz;
    double dt;
    // interior rectangle handled by the tiled stencil kernel: full-order
    // stencils, no PML, no hybrid correction, no free-surface closure
    int ix0, ix1, iz0, iz1;
};

__host__ __device__ inline size_t cell(const Geo &g, int iz, int ix)
{
    return (size_t)(ix + g.npml - 1) * (size_t)g.ld + (size_t)(iz - 1);
}

// 7x7 window of receiver (rx, rz) for loadseis (src/Submain.f90:61-62,66,
// 85): columns xa..xb, rows za..zb, clipped to the padded grid; returns the
// number of window cells, 0 when the receiver lies more than 3 cells outside
// the grid along either axis (the reference's loops are then empty and it
// records 0).  The cell indices are clamped in floating point before the
// conversion to int, so any coordinate (huge, infinite, NaN) gives a window
// inside the grid or an empty one.
__host__ __device__ inline int rec_window(const Geo &g, float x0, float z0, float rx, float rz, int &xa,
                                          int &xb, int &za, int &zb)
{
    float fx = floorf((rx - x0) / g.dx), fz = floorf((rz - z0) / g.dz);
    const float xlo = (float)(-g.npml - 8), xhi = (float)(g.nx + g.npml + 8);
    const float zlo = -8.0f, zhi = (float)(g.NZ + 8);
    fx = fx > xlo ? (fx < xhi ? fx : xhi) : xlo;   // NaN compares false: -> xlo
    fz = fz > zlo ? (fz < zhi ? fz : zhi) : zlo;
    const int irx = (int)fx + 1, irz = (int)fz + 1;
    xa = irx - 3 > 1 - g.npml ? irx - 3 : 1 - g.npml;
    xb = irx + 3 < g.nx + g.npml ? irx + 3 : g.nx + g.npml;
    za = irz - 3 > 1 ? irz - 3 : 1;
    zb = irz + 3 < g.NZ ? irz + 3 : g.NZ;
    if (xb < xa || zb < za) return 0;
    return (xb - xa + 1) * (zb - za + 1);
}

// Cells per time step of the six hybrid-ring background arrays in the order
// of src/FDFK2D.f90:330-355: left/right/bottom inside (reg), then
// left/right/bottom outside (pml).
__host__ __device__ inline void ring_sizes(const Geo &g, long n[6])
{
    const long nx = g.nx, nz = g.nz, nbd = g.nbd;
    n[0] = n[1] = (nz - nbd) * nbd;
    n[2] = nbd * nx;
    n[3] = n[4] = nz * nbd;
    n[5] = nbd * (nx + 2 * nbd);
}
__host__ __device__ inline long ring_count(const Geo &g)
{
    long n[6];
    ring_sizes(g, n);
    return n[0] + n[1] + n[2] + n[3] + n[4] + n[5];
}

// FD coefficient tables coef(j, i) = c[(j-1) + norder*(i-1)] in FP64 (as the
// reference's fd_coef0/fd_coef1) plus full-order FP32 copies for the
// interior kernel.
struct Coefs {
    double c0[kMaxOrder * kMaxOrder];
    double c1[kMaxOrder * kMaxOrder];
    float c0f[kMaxOrder * kMaxOrder];   // FP32 copies used by the pointwise kernels
    float c1f[kMaxOrder * kMaxOrder];
};

// Padding strips of the PML (left, right, bottom).  Cell offsets inside a
// strip are column-major with z fastest:
//   L: (ix - (1-npml)) * nz + (iz - 1)          ix in 1-npml..0,     iz in 1..nz
//   R: (ix - (nx+1))   * nz + (iz - 1)          ix in nx+1..nx+npml, iz in 1..nz
//   B: (ix - (1-npml)) * npml + (iz - (nz+1))   all ix,              iz in nz+1..nz+npml
// which is exactly the reference's bux1l/bpxlx/... cell order.
enum { S_L = 0, S_R = 1, S_B = 2, S_NONE = -1 };

// Multiaxial damping (Meza-Fajardo & Papageorgiou 2008, BSSA 98:1811): the
// split parts of a strip cell are damped along x by dX = dx(ix) + r dz(iz)
// and along z by dZ = dz(iz) + r dx(ix), dx / dz the classical profiles per
// padded column / row, while the first-derivative memories psi keep the
// classical profiles (psi_x: dx, psi_z: dz).  The classical layer (r = 0)
// lets guided modes of a strong low-velocity surface layer grow in the side
// strips at late time; the cross damping removes that growth.  (Driving the
// memories with dX, dZ as well makes the second-order split system
// unstable.)  FDFK_PML_RATIO overrides the ratio (experiments only).
constexpr float kPmlRatio = 0.1f;

struct Strips {
    // split fields u = u1 + u2 + u3: [strip][comp][part] -> cur / prev
    float *cur[3][2][3];
    float *prev[3][2][3];
    // first-derivative memories psi: [strip][comp] x-direction, z (bottom only)
    float *psx[3][2];
    float *psz[3][2];
    // classical damping profiles per padded column (NX) / row (NZ) and the
    // memory decay factors exp(-d dt) of those profiles
    const float *dpx, *bx;
    const float *dpz, *bz;
    float ratio;   // multiaxial ratio r
};

// damping of strip cell (iz, ix) along x and along z
__host__ __device__ inline float damp_x(const Strips *st, int jx, int jz) { return st->dpx[jx] + st->ratio * st->dpz[jz]; }
__host__ __device__ inline float damp_z(const Strips *st, int jx, int jz) { return st->dpz[jz] + st->ratio * st->dpx[jx]; }

// Twelve background arrays, reference layout and order:
// 0 uxl_reg 1 uzl_reg 2 uxr_reg 3 uzr_reg 4 uxb_reg 5 uzb_reg
// 6 uxl_pml 7 uzl_pml 8 uxr_pml 9 uzr_pml 10 uxb_pml 11 uzb_pml
// the twelve background arrays; step it is stored at slab it % period
// (period nt+1: all steps; 2B: a ring of two blocks of B steps)
struct Background {
    const float *a[12];
    int valid;
    int period;
};

// full-order FP32 coefficients of the fused stencils
template <int NO>
struct ICoef {
    float ai[NO];   // staggered inner  a_j = coef0(j,NO)
    float aox[NO];  // x outer 2*coef0(k,NO)/dx^2
    float aoz[NO];  // z outer 2*coef0(k,NO)/dz^2
    float bi[NO];   // central inner    b_j = coef1(j,NO)
    float bo[NO];   // mixed outer      coef1(k,NO)/(dx*dz)
};

// Strips and Background live in device memory (read through a pointer) so
// that runtime strip / array indices never force the tables into scratch.
struct StepArgs {
    Geo g;
    const Coefs *cf;
    const float *ux, *uz;          // level n
    float *uxp, *uzp;              // level n-1 in
    float *uxo, *uzo;              // level n+1 out (uxp, uzp: in place)
    const float *bf, *lam, *mu;    // dt^2/rho, lambda, mu
    const Strips *st;
    const float *ringx, *ringz;    // background u0 on the hybrid ring at this step (0 elsewhere)
    int ring_valid;
    int it;
    int frame_full;                // full-order frame tiles take their specialised path
};

// frame-kernel work decomposition: up to four bands (columns x0..x1, rows
// z0..z1), each cut into nbx x nbz tiles of tx x tz cells and lying on one
// side of the hybrid interface (view 0: total-field region, 1: padding);
// off[] are block prefix counts
struct ETiles {
    int n;
    int x0[4], x1[4], z0[4], z1[4], tx[4], tz[4], nbx[4], nbz[4], view[4];
    int off[5];
};

struct RecArgs {
    Geo g;
    float x0, z0;
    int nrcv, nsamp, istp;
    const float *rx, *rz;
    const float *ux, *uz;
    float *seisx, *seisz;  // (nsamp, nrcv), time fastest
};

template <int NO>
struct IntArgs {
    Geo g;
    ICoef<NO> k;
    const float *ux, *uz;
    float *uxp, *uzp;          // level n-1 in
    float *uxo, *uzo;          // level n+1 out (uxp, uzp: in place)
    const float *bf, *lam, *mu;
    // FP16 residuals of lam and mu against per-row FP32 profiles (16-bit
    // material storage): value = profile(row) + half * scale, the scales
    // powers of two
    const unsigned short *lam16, *mu16;
    const float *lamp, *mup;
    float sl, sm;
    const float *ringx, *ringz;  // background on the hybrid ring (total-field view of the padding)
    int corr;                    // ring values valid at this step
    int lx;   // output columns per block
    int nbx;  // blocks along x
    int nbz;  // full-height blocks along z
    int nsb;  // blocks of one-wave segments over the last, partial band of rows (0: none)
    int zseg; // first row of that band
};

}  // namespace fdfk
