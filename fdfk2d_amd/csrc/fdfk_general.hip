// ---------------------------------------------------------------------
// General per-cell step (included by fdfk_api.hip after the other kernel
// files).  It serves the absorbing top layer (is_PML_top = .true., izt =
// 1-npml, src/Input.f90:179-183), which the tiled kernels do not: every
// updated cell evaluates the six right-hand-side parts with the pointwise
// operators of src/Submain.f90 (reduced orders by the distance to the
// padded edges, rows counted from izt), so no operator pass needs a row
// above the first one.  Under a top layer the reference's passes are not
// usable as written: diffz2_reg and diffz2_pml leave row 1 of columns 1..nx
// without z-terms (Submain.f90:971,1062-1063) and diffxz_reg builds its
// intermediate from row 1 down (:616).
//
// Layout: fields, materials and the ring fields are [NX][ld] with rows
// zt..nz+npml (cell (iz, ix) at (ix+npml-1)*ld + iz-zt); split parts and
// memories keep the reference's strip arrays, the split parts with their
// two levels interleaved (1:2, cells) and updated in place.
//
// Arithmetic follows the formulation term by term with contraction off and
// FP64 accumulation of every coefficient product (a REAL(4) sum with a
// REAL(8) coefficient), so a run in T = float reproduces the FP32 CPU
// formulation bit for bit and T = double its FP64 build.
//
// Launches per step: k_gen_prep (receivers, hybrid ring values, PML
// memories and the staged operator terms, all from level n), for P-SV
// k_gen_ringst (the staged terms of the masked ring fields), then
// k_gen_step (every cell of level n+1, or, in FP32, the cells around the
// deep interior rectangle that the tiled interior kernel updates in the
// same step; the ring cells first, in waves of their own, with their hybrid
// correction).
// ---------------------------------------------------------------------
namespace fdfk {

template <class T>
struct GenArgs {
    Geo g;                      // NZ rows from zt; ld the column pitch
    int zt;                     // first row: 1 (free surface) or 1-npml (top layer)
    int top;                    // the top strip exists (is_pml_top)
    const double *c0, *c1;      // fd_coef0 / fd_coef1, (norder, norder) column-major
    const T *ux, *uz;           // level n
    const T *uxp, *uzp;         // level n-1
    T *uxo, *uzo;               // level n+1 (may alias uxp, uzp)
    const T *rho, *lam, *mu;    // material planes (damped cells repeat the undamped edge)
    T *bu[4][2][3];             // split parts [strip L R B T][comp][part], (1:2, cells)
    T *bpx[4][2];               // x-memories [strip][comp]
    T *bpz[4][2];               // z-memories (bottom and top strips)
    const T *dpx, *dpz;         // classical damping profiles per column / row
    const T *bx, *bz;           // exp(-d dt) of those profiles
    T ratio;                    // multiaxial ratio
    const T *ringx, *ringz;     // background on the ring cells at this step, 0 elsewhere
    int ring_valid;
    int top_ring;               // the top ring's background is present
    const T *bg[16];            // twelve reference arrays + uxt_reg, uzt_reg, uxt_pml, uzt_pml
    int it;
    // hole: the rectangle columns hx0..hx1 x rows hz0..hz1 that the tiled
    // interior kernel updates in the same step (none when hx0 > hx1)
    int hx0, hx1, hz0, hz1;
    // harmonic means of the moduli at the cell faces, formed once at set-up:
    // [mode] 0 lam+2mu, 1 mu; hx at cell (iz, ix) for the face (ix, ix+1),
    // hz for the face (iz, iz+1) (the same quotients the operators form)
    const T *hx[2], *hz[2];
    // Staged per-cell terms of level n, formed by the prep launch over the
    // cells the step kernel reaches (the box minus the rectangle ex0..ex1 x
    // ez0..ez1, which is the hole shrunk by norder): a face value of the
    // staggered second-derivative operators, face modulus times staggered
    // difference, and the products modulus times first derivative over
    // the spacing that the mixed operators sum.  P-SV: 0 x-face ux
    // [lam+2mu], 1 z-face ux [mu], 2 x-face uz [mu], 3 z-face uz [lam+2mu],
    // 4 mu dz ux / dz, 5 lam dx ux / dx, 6 lam dz uz / dz, 7 mu dx uz / dx,
    // 8..11 weighted memories (gstage); SH: 0 x-face uy [mu], 1 z-face uy
    // [mu], 2..3 weighted memories.
    T *st[12];
    int ex0, ex1, ez0, ez1;
    // P-SV: staged terms of the masked ring field (the other side's values
    // only) over the ring's neighbourhood (gring_region, compact): planes
    // 0..7 of the padding's values (for cells of the total-field region),
    // 8..15 of the total-field region's (for padding cells).  SH evaluates
    // its two masked operators directly (cheaper than staging them).
    T *mq[16];
};

template <class T>
__device__ __forceinline__ size_t gcell(const GenArgs<T> &a, int iz, int ix)
{
    return (size_t)(ix + a.g.npml - 1) * (size_t)a.g.ld + (size_t)(iz - a.zt);
}
__device__ __forceinline__ int gmin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int gmax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ double gco(const double *c, int norder, int j, int n) { return c[(j - 1) + norder * (n - 1)]; }
// REAL(T) accumulation with a REAL(8) coefficient
template <class T>
__device__ __forceinline__ T gacc(T s, double c, T d)
{
#pragma clang fp contract(off)
    return (T)((double)s + c * (double)d);
}
template <class T>
__device__ __forceinline__ T ghmean(T a, T b)
{
#pragma clang fp contract(off)
    return (T)2.0f / ((T)1.0f / a + (T)1.0f / b);
}

// Cell t of the box columns x0..x1 x rows z0..z1 minus the hole columns
// hx0..hx1 x rows hz0..hz1 (inside the box; none when hx0 > hx1 or hz0 >
// hz1), band by band: rows above the hole, rows below it, then the columns
// left and right of it; false past the last one.
__device__ __forceinline__ bool gbox_cell(long t, int x0, int x1, int z0, int z1, int hx0, int hx1, int hz0, int hz1,
                                          int &iz, int &ix)
{
    const long W = x1 - x0 + 1, H = z1 - z0 + 1;
    if (hx0 > hx1 || hz0 > hz1) {
        if (t >= W * H) return false;
        ix = x0 + (int)(t / H); iz = z0 + (int)(t % H);
        return true;
    }
    const long na = hz0 - z0, nb = z1 - hz1, nm = hz1 - hz0 + 1;
    if (t < na * W) { ix = x0 + (int)(t / na); iz = z0 + (int)(t % na); return true; }
    t -= na * W;
    if (t < nb * W) { ix = x0 + (int)(t / nb); iz = hz1 + 1 + (int)(t % nb); return true; }
    t -= nb * W;
    const long nl = (long)(hx0 - x0) * nm;
    if (t < nl) { ix = x0 + (int)(t / nm); iz = hz0 + (int)(t % nm); return true; }
    t -= nl;
    if (t < (long)(x1 - hx1) * nm) { ix = hx1 + 1 + (int)(t / nm); iz = hz0 + (int)(t % nm); return true; }
    return false;
}
// the inverse: the index t of cell (iz, ix) of the box minus the hole
__device__ __forceinline__ long gbox_index(int iz, int ix, int x0, int x1, int z0, int z1, int hx0, int hx1, int hz0,
                                           int hz1)
{
    const long W = x1 - x0 + 1, H = z1 - z0 + 1;
    if (hx0 > hx1 || hz0 > hz1) return (long)(ix - x0) * H + (iz - z0);
    const long na = hz0 - z0, nb = z1 - hz1, nm = hz1 - hz0 + 1;
    if (iz < hz0) return (long)(ix - x0) * na + (iz - z0);
    if (iz > hz1) return na * W + (long)(ix - x0) * nb + (iz - hz1 - 1);
    if (ix < hx0) return (na + nb) * W + (long)(ix - x0) * nm + (iz - hz0);
    return (na + nb) * W + (long)(hx0 - x0) * nm + (long)(ix - hx1 - 1) * nm + (iz - hz0);
}
__host__ __device__ inline long gbox_count(int x0, int x1, int z0, int z1, int hx0, int hx1, int hz0, int hz1)
{
    long n = (long)(x1 - x0 + 1) * (z1 - z0 + 1);
    if (hx0 <= hx1 && hz0 <= hz1) n -= (long)(hx1 - hx0 + 1) * (hz1 - hz0 + 1);
    return n;
}

// The ring cells: the box columns 1-nbd..nx+nbd x rows zr..nz+nbd (zr =
// 1-nbd with the top ring, else 1) minus the cells farther than nbd from
// every edge of the physical region it closes (columns nbd+1..nx-nbd x rows
// zi..nz-nbd, zi = nbd+1 with the top ring, else 1).  b: x0 x1 z0 z1 hx0
// hx1 hz0 hz1.
__host__ __device__ inline void gring_box(const Geo &g, int top_ring, int b[8])
{
    b[0] = 1 - g.nbd; b[1] = g.nx + g.nbd; b[2] = top_ring ? 1 - g.nbd : 1; b[3] = g.nz + g.nbd;
    b[4] = g.nbd + 1; b[5] = g.nx - g.nbd; b[6] = top_ring ? g.nbd + 1 : 1; b[7] = g.nz - g.nbd;
}

// row of "no cell" for a thread of a step launch that has none
constexpr int kNoCell = -(1 << 30);

// Cell t of a step launch: every cell, or with a hole (the rectangle the
// tiled interior kernel updates) the cells around it.
template <class T>
__device__ __forceinline__ bool gstep_cell(const GenArgs<T> &a, long t, int &iz, int &ix)
{
    const Geo &g = a.g;
    return gbox_cell(t, 1 - g.npml, g.nx + g.npml, a.zt, a.zt + g.NZ - 1, a.hx0, a.hx1, a.hz0, a.hz1, iz, ix);
}

// strip of a padded cell: 0 L, 1 R, 2 B, 3 T, -1 none
template <class T>
__device__ __forceinline__ int gstrip(const GenArgs<T> &a, int iz, int ix)
{
    if (iz >= 1 && iz <= a.g.nz) return ix <= 0 ? 0 : (ix > a.g.nx ? 1 : -1);
    return iz > a.g.nz ? 2 : 3;
}
template <class T>
__device__ __forceinline__ size_t gstrip_off(const GenArgs<T> &a, int s, int iz, int ix)
{
    const Geo &g = a.g;
    switch (s) {
    case 0: return (size_t)(iz - 1) + (size_t)g.nz * (size_t)(ix - (1 - g.npml));
    case 1: return (size_t)(iz - 1) + (size_t)g.nz * (size_t)(ix - (g.nx + 1));
    case 2: return (size_t)(iz - (g.nz + 1)) + (size_t)g.npml * (size_t)(ix - (1 - g.npml));
    default: return (size_t)(iz - (1 - g.npml)) + (size_t)g.npml * (size_t)(ix - (1 - g.npml));
    }
}

// Ring cell lookup (the reference's twelve arrays, src/FDFK2D.f90:330-355,
// plus the top ring): array index 0..15 of the x component (z is the next
// one) and the offset of step it, or -1.  inside: the cell is in the
// total-field region.
template <class T>
__device__ __forceinline__ long gring(const GenArgs<T> &a, int iz, int ix, int it, int &arr)
{
    const int nx = a.g.nx, nz = a.g.nz, nbd = a.g.nbd;
    if (iz >= 1 && iz <= nz - nbd && ix >= 1 && ix <= nbd) {
        arr = 0;
        return (long)(iz - 1) + (long)(nz - nbd) * ((long)(ix - 1) + (long)nbd * it);
    }
    if (iz >= 1 && iz <= nz - nbd && ix >= nx - nbd + 1 && ix <= nx) {
        arr = 2;
        return (long)(iz - 1) + (long)(nz - nbd) * ((long)(ix - (nx - nbd + 1)) + (long)nbd * it);
    }
    if (iz >= nz - nbd + 1 && iz <= nz && ix >= 1 && ix <= nx) {
        arr = 4;
        return (long)(iz - (nz - nbd + 1)) + (long)nbd * ((long)(ix - 1) + (long)nx * it);
    }
    if (iz >= 1 && iz <= nz && ix >= 1 - nbd && ix <= 0) {
        arr = 6;
        return (long)(iz - 1) + (long)nz * ((long)(ix - (1 - nbd)) + (long)nbd * it);
    }
    if (iz >= 1 && iz <= nz && ix >= nx + 1 && ix <= nx + nbd) {
        arr = 8;
        return (long)(iz - 1) + (long)nz * ((long)(ix - (nx + 1)) + (long)nbd * it);
    }
    if (iz >= nz + 1 && iz <= nz + nbd && ix >= 1 - nbd && ix <= nx + nbd) {
        arr = 10;
        return (long)(iz - (nz + 1)) + (long)nbd * ((long)(ix - (1 - nbd)) + (long)(nx + 2 * nbd) * it);
    }
    if (a.top_ring) {
        if (iz >= 1 && iz <= nbd && ix >= nbd + 1 && ix <= nx - nbd) {
            arr = 12;
            return (long)(iz - 1) + (long)nbd * ((long)(ix - (nbd + 1)) + (long)(nx - 2 * nbd) * it);
        }
        if (iz >= 1 - nbd && iz <= 0 && ix >= 1 - nbd && ix <= nx + nbd) {
            arr = 14;
            return (long)(iz - (1 - nbd)) + (long)nbd * ((long)(ix - (1 - nbd)) + (long)(nx + 2 * nbd) * it);
        }
    }
    return -1;
}

// field accessors: plain level-n fields, or the ring background masked to
// one side of the hybrid interface (inside = the total-field region)
template <class T>
struct GField {
    const GenArgs<T> *a;
    const T *f;
    __device__ __forceinline__ T operator()(int iz, int ix) const { return f[gcell(*a, iz, ix)]; }
};
template <class T>
struct GMasked {
    const GenArgs<T> *a;
    const T *f;
    bool inside;   // keep the values of the total-field region (else of the padding)
    __device__ __forceinline__ T operator()(int iz, int ix) const
    {
        const bool in = iz >= 1 && iz <= a->g.nz && ix >= 1 && ix <= a->g.nx;
        return in == inside ? f[gcell(*a, iz, ix)] : (T)0.0f;
    }
};

// collocated first derivatives with the reduced-order rule of diff1
// (Submain.f90:192-209), not divided by the spacing
template <class T, class F>
__device__ T gcdx(const GenArgs<T> &a, const F &f, int iz, int ix)
{
    const int norder = a.g.norder;
    const int nox = gmin(gmin(ix + a.g.npml - 1, a.g.nx + a.g.npml - ix), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= nox; ++k) s = gacc(s, gco(a.c1, norder, k, nox), (T)(f(iz, ix + k) - f(iz, ix - k)));
    return s;
}
template <class T, class F>
__device__ T gcdz(const GenArgs<T> &a, const F &f, int iz, int ix)
{
    const int norder = a.g.norder;
    const int noz = gmin(gmin(iz - a.zt, a.g.nz + a.g.npml - iz), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= noz; ++k) s = gacc(s, gco(a.c1, norder, k, noz), (T)(f(iz + k, ix) - f(iz - k, ix)));
    return s;
}

// modulus of the second-derivative terms: lam+2mu (mode 0) or mu (mode 1)
template <class T>
__device__ __forceinline__ T gmat(const GenArgs<T> &a, int mode, size_t c)
{
#pragma clang fp contract(off)
    return mode == 0 ? (T)(a.lam[c] + (T)2.0f * a.mu[c]) : a.mu[c];
}

// diffx2 / diffz2 at one cell (the loop bodies of Submain.f90:815-837,
// :990-1012 with their reduced orders)
template <class T, class F>
__device__ T gx2(const GenArgs<T> &a, int mode, const F &u, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder, nx = a.g.nx, npml = a.g.npml;
    const int nox = gmin(gmin(ix + npml - 1, nx + npml - ix), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= nox; ++k) {
        const int i = (2 * k - 1) / 2;
        const int n1 = gmin(gmin(ix + i + npml - 1, nx + npml - ix - i), norder);
        T s1 = (T)0.0f;
        for (int j = 1; j <= n1; ++j) s1 = gacc(s1, gco(a.c0, norder, j, n1), (T)(u(iz, ix + j + k - 1) - u(iz, ix - j + k)));
        const int n2 = gmin(gmin(ix - i + npml - 1, nx + npml - ix + i), norder);
        T s2 = (T)0.0f;
        for (int j = 1; j <= n2; ++j) s2 = gacc(s2, gco(a.c0, norder, j, n2), (T)(u(iz, ix + j - k) - u(iz, ix - j - k + 1)));
        const T t = a.hx[mode][gcell(a, iz, ix + k - 1)] * s1 - a.hx[mode][gcell(a, iz, ix - k)] * s2;
        s = gacc(s, gco(a.c0, norder, k, nox), t);
    }
    const bool in_reg = iz >= 1 && iz <= a.g.nz && ix >= 1 && ix <= nx;
    const T dxs = (T)a.g.dx;
    const T dx2 = in_reg ? (T)(1.0 / (double)(dxs * dxs)) : (T)1.0f / (dxs * dxs);
    return s * dx2;
}
template <class T, class F>
__device__ T gz2(const GenArgs<T> &a, int mode, const F &u, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder, nz = a.g.nz, npml = a.g.npml, zt = a.zt;
    const int noz = gmin(gmin(iz - zt, nz + npml - iz), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= noz; ++k) {
        const int i = (2 * k - 1) / 2;
        const int n1 = gmin(gmin(iz + i - zt, nz + npml - iz - i), norder);
        T s1 = (T)0.0f;
        for (int j = 1; j <= n1; ++j) s1 = gacc(s1, gco(a.c0, norder, j, n1), (T)(u(iz + j + k - 1, ix) - u(iz - j + k, ix)));
        const int n2 = gmin(gmin(iz - i - zt, nz + npml - iz + i), norder);
        T s2 = (T)0.0f;
        for (int j = 1; j <= n2; ++j) s2 = gacc(s2, gco(a.c0, norder, j, n2), (T)(u(iz + j - k, ix) - u(iz - j - k + 1, ix)));
        const T t = a.hz[mode][gcell(a, iz + k - 1, ix)] * s1 - a.hz[mode][gcell(a, iz - k, ix)] * s2;
        s = gacc(s, gco(a.c0, norder, k, noz), t);
    }
    const T dzs = (T)a.g.dz;
    return s * ((T)1.0f / (dzs * dzs));
}
// ---- staged operators: the sums of gx2 and gz2 and of the mixed
// operators, term by term, with the inner sums read from the staged planes ----
// Order of the staggered difference at the face between p and p+1 as the
// operators of diffx2/diffz2 choose it: the s1 sums of gx2/gz2 take
// gnx(p) / gnz(p) at the face p = i+k-1, the s2 sums gnx(h+1) / gnz(h+1) at
// the face h = i-k.
template <class T>
__device__ __forceinline__ int gnx(const GenArgs<T> &a, int p)
{
    return gmin(gmin(p + a.g.npml - 1, a.g.nx + a.g.npml - p), a.g.norder);
}
template <class T>
__device__ __forceinline__ int gnz(const GenArgs<T> &a, int p)
{
    return gmin(gmin(p - a.zt, a.g.nz + a.g.npml - p), a.g.norder);
}
// staggered difference at the face (p, p+1) along x / z with n terms
template <class T, class F>
__device__ T gsx(const GenArgs<T> &a, const F &u, int iz, int p, int n)
{
    T s = (T)0.0f;
    for (int j = 1; j <= n; ++j) s = gacc(s, gco(a.c0, a.g.norder, j, n), (T)(u(iz, p + j) - u(iz, p - j + 1)));
    return s;
}
template <class T, class F>
__device__ T gsz(const GenArgs<T> &a, const F &u, int p, int ix, int n)
{
    T s = (T)0.0f;
    for (int j = 1; j <= n; ++j) s = gacc(s, gco(a.c0, a.g.norder, j, n), (T)(u(p + j, ix) - u(p - j + 1, ix)));
    return s;
}
// Indexers of staged planes: full planes in cell order, or compact planes
// over a box minus a hole (the ring's neighbourhood)
template <class T>
struct GIdxCell {
    const GenArgs<T> *a;
    __device__ __forceinline__ size_t operator()(int iz, int ix) const { return gcell(*a, iz, ix); }
};
struct GIdxBox {
    int b[8];
    __device__ __forceinline__ long operator()(int iz, int ix) const
    {
        return gbox_index(iz, ix, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]);
    }
};
// gx2 from the staged x-face values F of field u (modulus mode); F at
// idx(cell) holds the face value of the face (cell, next cell)
template <class T, class I, class U>
__device__ T gx2s(const GenArgs<T> &a, int mode, const T *F, const I &idx, const U &u, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder, nx = a.g.nx, npml = a.g.npml;
    const int nox = gmin(gmin(ix + npml - 1, nx + npml - ix), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= nox; ++k) {
        const int h = ix - k, nh = gnx(a, h + 1);
        // the face value holds gnx(h) terms; near the padded edges the s2
        // sum takes another count and is formed here
        const T pb = nh == gnx(a, h) ? F[idx(iz, h)] : a.hx[mode][gcell(a, iz, h)] * gsx(a, u, iz, h, nh);
        s = gacc(s, gco(a.c0, norder, k, nox), (T)(F[idx(iz, ix + k - 1)] - pb));
    }
    const bool in_reg = iz >= 1 && iz <= a.g.nz && ix >= 1 && ix <= nx;
    const T dxs = (T)a.g.dx;
    const T dx2 = in_reg ? (T)(1.0 / (double)(dxs * dxs)) : (T)1.0f / (dxs * dxs);
    return s * dx2;
}
template <class T, class I, class U>
__device__ T gz2s(const GenArgs<T> &a, int mode, const T *F, const I &idx, const U &u, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder, nz = a.g.nz, npml = a.g.npml, zt = a.zt;
    const int noz = gmin(gmin(iz - zt, nz + npml - iz), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= noz; ++k) {
        const int h = iz - k, nh = gnz(a, h + 1);
        const T pb = nh == gnz(a, h) ? F[idx(h, ix)] : a.hz[mode][gcell(a, h, ix)] * gsz(a, u, h, ix, nh);
        s = gacc(s, gco(a.c0, norder, k, noz), (T)(F[idx(iz + k - 1, ix)] - pb));
    }
    const T dzs = (T)a.g.dz;
    return s * ((T)1.0f / (dzs * dzs));
}
// Dx(mat Dz u) (diffzx) and Dz(mat Dx u) (diffxz) at one cell from the staged
// products V = mat Dz u / dz (gzxs) or mat Dx u / dx (gxzs) at the cells of
// the outer sum
template <class T, class I>
__device__ T gzxs(const GenArgs<T> &a, const T *V, const I &idx, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder;
    const int nox = gmin(gmin(ix + a.g.npml - 1, a.g.nx + a.g.npml - ix), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= nox; ++k)
        s = gacc(s, gco(a.c1, norder, k, nox), (T)(V[idx(iz, ix + k)] - V[idx(iz, ix - k)]));
    return s / (T)a.g.dx;
}
template <class T, class I>
__device__ T gxzs(const GenArgs<T> &a, const T *V, const I &idx, int iz, int ix)
{
#pragma clang fp contract(off)
    const int norder = a.g.norder;
    const int noz = gmin(gmin(iz - a.zt, a.g.nz + a.g.npml - iz), norder);
    T s = (T)0.0f;
    for (int k = 1; k <= noz; ++k)
        s = gacc(s, gco(a.c1, norder, k, noz), (T)(V[idx(iz + k, ix)] - V[idx(iz - k, ix)]));
    return s / (T)a.g.dz;
}
// Cell t of the staged region (the box minus the shrunk hole)
template <class T>
__device__ __forceinline__ bool gstage_cell(const GenArgs<T> &a, long t, int &iz, int &ix)
{
    const Geo &g = a.g;
    return gbox_cell(t, 1 - g.npml, g.nx + g.npml, a.zt, a.zt + g.NZ - 1, a.ex0, a.ex1, a.ez0, a.ez1, iz, ix);
}
// the staged terms of one cell (iz, ix) of fields ux, uz into planes P at o;
// d: the undivided first derivatives dz ux, dx ux, dz uz, dx uz
template <class T, class U>
__device__ void gstage_terms(const GenArgs<T> &a, const U &ux, const U &uz, int iz, int ix, T *const *P, size_t o,
                             T d[4])
{
#pragma clang fp contract(off)
    const size_t c = gcell(a, iz, ix);
    const int nfx = gnx(a, ix), nfz = gnz(a, iz);
    P[0][o] = a.hx[0][c] * gsx(a, ux, iz, ix, nfx);
    P[1][o] = a.hz[1][c] * gsz(a, ux, iz, ix, nfz);
    P[2][o] = a.hx[1][c] * gsx(a, uz, iz, ix, nfx);
    P[3][o] = a.hz[0][c] * gsz(a, uz, iz, ix, nfz);
    d[0] = gcdz(a, ux, iz, ix);
    d[1] = gcdx(a, ux, iz, ix);
    d[2] = gcdz(a, uz, iz, ix);
    d[3] = gcdx(a, uz, iz, ix);
    const T dx = (T)a.g.dx, dz = (T)a.g.dz;
    P[4][o] = a.mu[c] * d[0] / dz;    // gzx[mu] of ux
    P[5][o] = a.lam[c] * d[1] / dx;   // gxz[lam] of ux
    P[6][o] = a.lam[c] * d[2] / dz;   // gzx[lam] of uz
    P[7][o] = a.mu[c] * d[3] / dx;    // gxz[mu] of uz
}
template <class T, class U>
__device__ void gstage_terms_sh(const GenArgs<T> &a, const U &u, int iz, int ix, T *const *P, size_t o)
{
#pragma clang fp contract(off)
    const size_t c = gcell(a, iz, ix);
    P[0][o] = a.hx[1][c] * gsx(a, u, iz, ix, gnx(a, ix));
    P[1][o] = a.hz[1][c] * gsz(a, u, iz, ix, gnz(a, iz));
}
// The staged terms of level n at cell t of the staged region, and on a
// strip cell the update of its PML memories (the first derivatives of
// level n, the formulation's phase A); every cell then stores its memories
// weighted by the modulus the memory corrections of the step apply (0 off
// the strips): P-SV planes 8 (lam+2mu) psi_x[ux], 9 mu psi_x[uz], 10 mu
// psi_z[ux], 11 (lam+2mu) psi_z[uz]; SH planes 2 mu psi_x, 3 mu psi_z.
template <class T>
__device__ void gstage(const GenArgs<T> &a, long t)
{
#pragma clang fp contract(off)
    int iz, ix;
    if (!gstage_cell(a, t, iz, ix)) return;
    const size_t c = gcell(a, iz, ix);
    T d[4];
    gstage_terms(a, GField<T>{&a, a.ux}, GField<T>{&a, a.uz}, iz, ix, a.st, c, d);
    const Geo &g = a.g;
    const int s = gstrip(a, iz, ix);
    // off the strips the weighted memories stay 0 from set-up
    if (s >= 0) {
        T w[4] = {(T)0.0f, (T)0.0f, (T)0.0f, (T)0.0f};
        const size_t o = gstrip_off(a, s, iz, ix);
        const int jx = ix + g.npml - 1, jz = iz - a.zt;
        T *px0 = a.bpx[s][0], *px1 = a.bpx[s][1];
        if (a.dpx[jx] > (T)0.0f) {
            const T b = a.bx[jx], cc = (T)1.0f - b;
            px0[o] = b * px0[o] + cc * (d[1] / (T)g.dx);
            px1[o] = b * px1[o] + cc * (d[3] / (T)g.dx);
        }
        const T ma = a.lam[c] + (T)2.0f * a.mu[c], mu = a.mu[c];
        w[0] = ma * px0[o];
        w[1] = mu * px1[o];
        if (T *pz0 = a.bpz[s][0]) {
            T *pz1 = a.bpz[s][1];
            if (a.dpz[jz] > (T)0.0f) {
                const T b = a.bz[jz], cc = (T)1.0f - b;
                pz0[o] = b * pz0[o] + cc * (d[0] / (T)g.dz);
                pz1[o] = b * pz1[o] + cc * (d[2] / (T)g.dz);
            }
            w[2] = mu * pz0[o];
            w[3] = ma * pz1[o];
        }
        for (int k = 0; k < 4; ++k) a.st[8 + k][c] = w[k];
    }
}
template <class T>
__device__ void gstage_sh(const GenArgs<T> &a, long t)
{
#pragma clang fp contract(off)
    int iz, ix;
    if (!gstage_cell(a, t, iz, ix)) return;
    const size_t c = gcell(a, iz, ix);
    const GField<T> fu{&a, a.ux};
    gstage_terms_sh(a, fu, iz, ix, a.st, c);
    const Geo &g = a.g;
    const int s = gstrip(a, iz, ix);
    if (s >= 0) {   // off the strips the weighted memories stay 0 from set-up
        T wx = (T)0.0f, wz = (T)0.0f;
        const size_t o = gstrip_off(a, s, iz, ix);
        const int jx = ix + g.npml - 1, jz = iz - a.zt;
        T *px = a.bpx[s][0];
        if (a.dpx[jx] > (T)0.0f) {
            const T b = a.bx[jx], cc = (T)1.0f - b;
            px[o] = b * px[o] + cc * (gcdx(a, fu, iz, ix) / (T)g.dx);
        }
        wx = a.mu[c] * px[o];
        if (T *pz = a.bpz[s][0]) {
            if (a.dpz[jz] > (T)0.0f) {
                const T b = a.bz[jz], cc = (T)1.0f - b;
                pz[o] = b * pz[o] + cc * (gcdz(a, fu, iz, ix) / (T)g.dz);
            }
            wz = a.mu[c] * pz[o];
        }
        a.st[2][c] = wx;
        a.st[3][c] = wz;
    }
}

// The six RHS parts {Xx, Zx, Mx, Xz, Zz, Mz} of fields ux, uz at one cell
// from their staged terms P; the free-surface closure on row 1 when zt = 1
// (half-cell energy form, traction free)
template <class T, class I, class U>
__device__ void gparts_st(const GenArgs<T> &a, const T *const *P, const I &idx, const U &ux, const U &uz, int iz,
                          int ix, T p[6])
{
#pragma clang fp contract(off)
    if (iz == 1 && a.zt == 1) {
        const size_t ca = gcell(a, 1, ix), cb = gcell(a, 2, ix);
        const T dz = (T)a.g.dz, dx = (T)a.g.dx;
        const T sdz2 = (T)2.0f / (dz * dz);
        const int norder = a.g.norder;
        const int nox = gmin(gmin(ix + a.g.npml - 1, a.g.nx + a.g.npml - ix), norder);
        T sl = (T)0.0f, sm = (T)0.0f;
        for (int k = 1; k <= nox; ++k) {
            const size_t a1 = gcell(a, 1, ix + k), b1 = gcell(a, 2, ix + k), a2 = gcell(a, 1, ix - k), b2 = gcell(a, 2, ix - k);
            const T lp = (T)0.5f * (a.lam[a1] + a.lam[b1]), lm = (T)0.5f * (a.lam[a2] + a.lam[b2]);
            const T mp = (T)0.5f * (a.mu[a1] + a.mu[b1]), mm = (T)0.5f * (a.mu[a2] + a.mu[b2]);
            sl = gacc(sl, gco(a.c1, norder, k, nox), (T)(lp * (uz(2, ix + k) - uz(1, ix + k)) - lm * (uz(2, ix - k) - uz(1, ix - k))));
            sm = gacc(sm, gco(a.c1, norder, k, nox), (T)(mp * (ux(2, ix + k) - ux(1, ix + k)) - mm * (ux(2, ix - k) - ux(1, ix - k))));
        }
        const T lbar = (T)0.5f * (a.lam[ca] + a.lam[cb]), mbar = (T)0.5f * (a.mu[ca] + a.mu[cb]);
        const T dxuz = (gcdx(a, uz, 1, ix) + gcdx(a, uz, 2, ix)) / dx;
        const T dxux = (gcdx(a, ux, 1, ix) + gcdx(a, ux, 2, ix)) / dx;
        p[0] = gx2s(a, 0, P[0], idx, ux, iz, ix);
        p[1] = (a.hz[1][ca] * (ux(2, ix) - ux(1, ix))) * sdz2;
        p[2] = (sl / dx + mbar * dxuz) / dz;
        p[3] = gx2s(a, 1, P[2], idx, uz, iz, ix);
        p[4] = (a.hz[0][ca] * (uz(2, ix) - uz(1, ix))) * sdz2;
        p[5] = (sm / dx + lbar * dxux) / dz;
        return;
    }
    p[0] = gx2s(a, 0, P[0], idx, ux, iz, ix);
    p[1] = gz2s(a, 1, P[1], idx, ux, iz, ix);
    p[2] = gzxs(a, P[6], idx, iz, ix) + gxzs(a, P[7], idx, iz, ix);
    p[3] = gx2s(a, 1, P[2], idx, uz, iz, ix);
    p[4] = gz2s(a, 0, P[3], idx, uz, iz, ix);
    p[5] = gzxs(a, P[4], idx, iz, ix) + gxzs(a, P[5], idx, iz, ix);
}
// gparts_sh from the staged terms
template <class T, class I, class U>
__device__ void gparts_sh_st(const GenArgs<T> &a, const T *const *P, const I &idx, const U &u, int iz, int ix, T &X,
                             T &Z)
{
#pragma clang fp contract(off)
    X = gx2s(a, 1, P[0], idx, u, iz, ix);
    if (iz == 1 && a.zt == 1) {
        const size_t ca = gcell(a, 1, ix);
        const T dz = (T)a.g.dz;
        Z = (a.hz[1][ca] * (u(2, ix) - u(1, ix))) * ((T)2.0f / (dz * dz));
    } else {
        Z = gz2s(a, 1, P[1], idx, u, iz, ix);
    }
}

// ---- per-step preparation: receivers (level n), ring values, memories ----
template <class T>
struct GenRec {
    int nrcv, nsamp, istp;
    T x0, z0;
    const float *rx, *rz;
    T *seisx, *seisz;   // (nsamp, nrcv), time fastest
};

__device__ double gdirac(double x)
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

// loadseis (Submain.f90:3-116) of receiver i, rows from zt: the 7x7
// window's products summed in the reference's loop order
template <class T, int NC = 2>
__device__ void grecord(const GenArgs<T> &a, const GenRec<T> &r, int i)
{
#pragma clang fp contract(off)
    const Geo &g = a.g;
    const T rx = (T)r.rx[i], rz = (T)r.rz[i], dx = (T)g.dx, dz = (T)g.dz;
    // the window, clamped in floating point (any coordinate gives an empty
    // or in-grid window)
    float fx = floorf((float)((rx - r.x0) / dx)), fz = floorf((float)((rz - r.z0) / dz));
    const float xlo = (float)(-g.npml - 8), xhi = (float)(g.nx + g.npml + 8);
    const float zlo = (float)(a.zt - 9), zhi = (float)(g.nz + g.npml + 8);
    fx = fx > xlo ? (fx < xhi ? fx : xhi) : xlo;
    fz = fz > zlo ? (fz < zhi ? fz : zhi) : zlo;
    const int irx = (int)fx + 1, irz = (int)fz + 1;
    T sx = (T)0.0f, sz = (T)0.0f;
    for (int ix = gmax(1 - g.npml, irx - 3); ix <= gmin(g.nx + g.npml, irx + 3); ++ix) {
        const double x = fabs((double)(r.x0 + (T)(ix - 1) * dx) - (double)rx) / (double)dx;
        const double bax = gdirac(x);
        for (int iz = gmax(a.zt, irz - 3); iz <= gmin(g.nz + g.npml, irz + 3); ++iz) {
            const double z = fabs((double)(r.z0 + (T)(iz - 1) * dz) - (double)rz) / (double)dz;
            const double ba = bax * gdirac(z);
            const size_t c = gcell(a, iz, ix);
            sx = (T)((double)sx + ba * (double)a.ux[c]);
            if (NC == 2) sz = (T)((double)sz + ba * (double)a.uz[c]);
        }
    }
    r.seisx[(size_t)i * r.nsamp + r.istp] = sx;
    if (NC == 2) r.seisz[(size_t)i * r.nsamp + r.istp] = sz;
}

// one thread per receiver, per ring cell and per cell of the staged region:
// threads [0, nrec) record, [nrec, nrec + nring) write the ring values of
// the cells of gring_box (the other cells of the ring fields stay zero from
// their allocation), the next nst stage the terms of level n (gstage: a
// strip cell also updates its memories)
template <class T>
__global__ __launch_bounds__(256) void k_gen_prep(GenArgs<T> a, GenRec<T> r, long nrec, long nring, long nst,
                                                   T *ringx, T *ringz)
{
#pragma clang fp contract(off)
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const Geo &g = a.g;
    if (t < nrec) { grecord(a, r, (int)t); return; }
    t -= nrec;
    if (t < nring) {
        int b[8], iz, ix;
        gring_box(g, a.top_ring, b);
        gbox_cell(t, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], iz, ix);
        int arr = 0;
        const long o = a.ring_valid ? gring(a, iz, ix, a.it, arr) : -1;
        const size_t c = gcell(a, iz, ix);
        ringx[c] = o >= 0 ? a.bg[arr][o] : (T)0.0f;
        ringz[c] = o >= 0 ? a.bg[arr + 1][o] : (T)0.0f;
        return;
    }
    t -= nring;
    if (t < nst) gstage(a, t);
}

// one thread per cell of the padded grid: level n+1
template <class T>
__global__ __launch_bounds__(256) void k_gen_step(GenArgs<T> a, long nq)
{
#pragma clang fp contract(off)
    const Geo &g = a.g;
    int iz, ix;
    T q[6];
    const bool ringcell = gstep_job(a, nq, iz, ix);
    if (iz == kNoCell) return;
    if (ringcell) {
        // the hybrid correction: the operators of the ring field masked to
        // the other side of the interface, from its staged terms
        const bool inside = iz >= 1 && iz <= g.nz && ix >= 1 && ix <= g.nx;
        const GMasked<T> wx{&a, a.ringx, !inside}, wz{&a, a.ringz, !inside};
        gparts_st(a, a.mq + (inside ? 0 : 8), gring_region(a), wx, wz, iz, ix, q);
    }
    const size_t gi = gcell(a, iz, ix);
    const int iz_first = a.zt == 1 ? 1 : a.zt + 1;
    if (ix < 2 - g.npml || ix > g.nx + g.npml - 1 || iz < iz_first || iz > g.nz + g.npml - 1) {
        // cells on the padded grid's edge keep their value
        a.uxo[gi] = a.ux[gi];
        a.uzo[gi] = a.uz[gi];
        return;
    }
    T p[6];
    gparts_st(a, a.st, GIdxCell<T>{&a}, GField<T>{&a, a.ux}, GField<T>{&a, a.uz}, iz, ix, p);
    const int s = gstrip(a, iz, ix);
    // PML memory corrections (nonzero only next to damped cells)
    {
        const bool surf = iz == 1 && a.zt == 1;
        const int norder = g.norder;
        const int nox = gmin(gmin(ix + g.npml - 1, g.nx + g.npml - ix), norder);
        const int noz = gmin(gmin(iz - a.zt, g.nz + g.npml - iz), norder);
        T sx0 = (T)0.0f, sx1 = (T)0.0f, sz0 = (T)0.0f, sz1 = (T)0.0f;
        const T *W0 = a.st[8], *W1 = a.st[9], *W2 = a.st[10], *W3 = a.st[11];
        // a sum whose cells are all off the strips is a sum of zeros
        const bool xs = iz < 1 || iz > g.nz || ix - nox < 1 || ix + nox > g.nx;
        const bool zs = ix < 1 || ix > g.nx || iz - noz < 1 || iz + noz > g.nz;
        for (int k = 1; xs && k <= nox; ++k) {
            const size_t ca = gcell(a, iz, ix + k), cb = gcell(a, iz, ix - k);
            sx0 = gacc(sx0, gco(a.c1, norder, k, nox), (T)(W0[ca] - W0[cb]));
            sx1 = gacc(sx1, gco(a.c1, norder, k, nox), (T)(W1[ca] - W1[cb]));
        }
        for (int k = 1; zs && k <= noz; ++k) {
            const size_t ca = gcell(a, iz + k, ix), cb = gcell(a, iz - k, ix);
            sz0 = gacc(sz0, gco(a.c1, norder, k, noz), (T)(W2[ca] - W2[cb]));
            sz1 = gacc(sz1, gco(a.c1, norder, k, noz), (T)(W3[ca] - W3[cb]));
        }
        p[0] = p[0] - sx0 / (T)g.dx;
        p[3] = p[3] - sx1 / (T)g.dx;
        if (!surf) { p[1] = p[1] - sz0 / (T)g.dz; p[4] = p[4] - sz1 / (T)g.dz; }
    }
    // hybrid TF/SF correction: an inside-ring cell adds the operator of the
    // background on the padding, an outside-ring cell subtracts that of the
    // background on the total-field region
    if (ringcell) {
        const bool inside = iz >= 1 && iz <= g.nz && ix >= 1 && ix <= g.nx;
        for (int m = 0; m < 6; ++m) p[m] = inside ? p[m] + q[m] : p[m] - q[m];
    }
    const T dt2 = (T)(g.dt * g.dt);
    const T bf = dt2 / a.rho[gi];
    if (s < 0) {
        const T rx = (p[0] + p[1]) + p[2], rz = (p[3] + p[4]) + p[5];
        a.uxo[gi] = ((T)2.0f * a.ux[gi] - a.uxp[gi]) + bf * rx;
        a.uzo[gi] = ((T)2.0f * a.uz[gi] - a.uzp[gi]) + bf * rz;
        return;
    }
    const size_t o = gstrip_off(a, s, iz, ix);
    const int jx = ix + g.npml - 1, jz = iz - a.zt;
    const T ax = (T)0.5f * (a.dpx[jx] + a.ratio * a.dpz[jz]) * (T)g.dt;
    const T az = (T)0.5f * (a.dpz[jz] + a.ratio * a.dpx[jx]) * (T)g.dt;
    const T c3 = (T)2.0f - (T)4.0f * ax * az;
    for (int comp = 0; comp < 2; ++comp) {
        T *b1 = a.bu[s][comp][0], *b2 = a.bu[s][comp][1], *b3 = a.bu[s][comp][2];
        const T X = p[3 * comp], Z = p[3 * comp + 1], Mm = p[3 * comp + 2];
        const T n1 = ((T)2.0f * b1[2 * o] - ((T)1.0f - ax) * b1[2 * o + 1] + bf * X) / ((T)1.0f + ax);
        const T n2 = ((T)2.0f * b2[2 * o] - ((T)1.0f - az) * b2[2 * o + 1] + bf * Z) / ((T)1.0f + az);
        const T n3 = (c3 * b3[2 * o] - ((T)1.0f - ax - az) * b3[2 * o + 1] + bf * Mm) / ((T)1.0f + ax + az);
        b1[2 * o + 1] = b1[2 * o]; b1[2 * o] = n1;
        b2[2 * o + 1] = b2[2 * o]; b2[2 * o] = n2;
        b3[2 * o + 1] = b3[2 * o]; b3[2 * o] = n3;
        const T tot = (n1 + n2) + n3;
        if (comp == 0) a.uxo[gi] = tot; else a.uzo[gi] = tot;
    }
}

// ---- hybrid-ring correction terms: one thread per ring cell that the step
// kernel updates (the operator parts of the ring background masked to the
// other side of the interface), so that the step kernel's waves, most of
// which hold a few ring cells, read six values instead of evaluating the
// operators on the masked background ----
template <class T>
__device__ __forceinline__ bool gring_job(const GenArgs<T> &a, long t, int &iz, int &ix)
{
    const Geo &g = a.g;
    int b[8];
    gring_box(g, a.top_ring, b);
    if (t >= gbox_count(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7])) return false;
    gbox_cell(t, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], iz, ix);
    // cells the step kernel leaves: the padded grid's edge and the hole
    const int iz_first = a.zt == 1 ? 1 : a.zt + 1;
    if (ix < 2 - g.npml || ix > g.nx + g.npml - 1 || iz < iz_first || iz > g.nz + g.npml - 1) return false;
    return !(ix >= a.hx0 && ix <= a.hx1 && iz >= a.hz0 && iz <= a.hz1);
}
// The ring's neighbourhood: the ring box grown and its hole shrunk by
// norder, within the padded grid; the ring cells' operators read no cell
// outside it
template <class T>
__device__ __forceinline__ GIdxBox gring_region(const GenArgs<T> &a)
{
    const Geo &g = a.g;
    int r[8];
    gring_box(g, a.top_ring, r);
    const int no = g.norder;
    GIdxBox b;
    b.b[0] = gmax(r[0] - no, 1 - g.npml); b.b[1] = gmin(r[1] + no, g.nx + g.npml);
    b.b[2] = gmax(r[2] - no, a.zt); b.b[3] = gmin(r[3] + no, a.zt + g.NZ - 1);
    b.b[4] = r[4] + no; b.b[5] = r[5] - no; b.b[6] = r[6] + no; b.b[7] = r[7] - no;
    return b;
}
// staged terms of both masked ring fields at cell t of the ring's neighbourhood
template <class T>
__global__ __launch_bounds__(256) void k_gen_ringst(GenArgs<T> a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const GIdxBox b = gring_region(a);
    if (t >= gbox_count(b.b[0], b.b[1], b.b[2], b.b[3], b.b[4], b.b[5], b.b[6], b.b[7])) return;
    int iz, ix;
    gbox_cell(t, b.b[0], b.b[1], b.b[2], b.b[3], b.b[4], b.b[5], b.b[6], b.b[7], iz, ix);
    T d[4];
    for (int keep = 0; keep < 2; ++keep)
        gstage_terms(a, GMasked<T>{&a, a.ringx, keep == 1}, GMasked<T>{&a, a.ringz, keep == 1}, iz, ix,
                     a.mq + 8 * keep, (size_t)t, d);
}
// Cell of thread t of a step launch.  With nq > 0 (the ring values are
// valid) the launch starts with the ring cells that the step updates, nq
// threads in the order of gring_box rounded up to whole blocks, so that the
// waves evaluating the hybrid correction hold ring cells only; the cells of
// the band around the interior follow, the ring cells among them left to
// the first range.  Returns true for a ring cell of the first range; iz =
// kNoCell for a thread without a cell.
template <class T>
__device__ __forceinline__ bool gstep_job(const GenArgs<T> &a, long nq, int &iz, int &ix)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nqb = (nq + 255) / 256 * 256;
    iz = kNoCell;
    if (t < nqb) {
        if (t >= nq || !gring_job(a, t, iz, ix)) iz = kNoCell;
        return iz != kNoCell;
    }
    if (!gstep_cell(a, t - nqb, iz, ix)) {
        iz = kNoCell;
        return false;
    }
    if (nq > 0) {
        const Geo &g = a.g;
        const int iz_first = a.zt == 1 ? 1 : a.zt + 1;
        const bool edge = ix < 2 - g.npml || ix > g.nx + g.npml - 1 || iz < iz_first || iz > g.nz + g.npml - 1;
        int arr = 0;
        if (!edge && gring(a, iz, ix, a.it, arr) >= 0) iz = kNoCell;   // updated by the first range
    }
    return false;
}

// ---- SH (transverse) motion: uy in the ux slots (fields, ring values,
// x background arrays, seisx), one split part pair per strip cell (parts 0
// and 1 of component 0), memories of component 0.  rho uy_tt = X + Z with X
// = diffx2[mu] uy, Z = diffz2[mu] uy, the free-surface closure Z = 2/dz^2
// H(mu1, mu2)(uy2 - uy1) on row 1 when zt = 1.
template <class T, class F>
__device__ void gparts_sh(const GenArgs<T> &a, const F &u, int iz, int ix, T &X, T &Z)
{
#pragma clang fp contract(off)
    X = gx2(a, 1, u, iz, ix);
    if (iz == 1 && a.zt == 1) {
        const size_t ca = gcell(a, 1, ix), cb = gcell(a, 2, ix);
        const T dz = (T)a.g.dz;
        Z = (a.hz[1][ca] * (u(2, ix) - u(1, ix))) * ((T)2.0f / (dz * dz));
    } else {
        Z = gz2(a, 1, u, iz, ix);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_gen_prep_sh(GenArgs<T> a, GenRec<T> r, long nrec, long nring, long nst,
                                                      T *ringx)
{
#pragma clang fp contract(off)
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const Geo &g = a.g;
    if (t < nrec) { grecord<T, 1>(a, r, (int)t); return; }
    t -= nrec;
    if (t < nring) {
        int b[8], iz, ix;
        gring_box(g, a.top_ring, b);
        gbox_cell(t, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], iz, ix);
        int arr = 0;
        const long o = a.ring_valid ? gring(a, iz, ix, a.it, arr) : -1;
        ringx[gcell(a, iz, ix)] = o >= 0 ? a.bg[arr][o] : (T)0.0f;
        return;
    }
    t -= nring;
    if (t < nst) gstage_sh(a, t);
}

template <class T>
__global__ __launch_bounds__(256) void k_gen_step_sh(GenArgs<T> a, long nq)
{
#pragma clang fp contract(off)
    const Geo &g = a.g;
    int iz, ix;
    T qx = (T)0.0f, qz = (T)0.0f;
    const bool ringcell = gstep_job(a, nq, iz, ix);
    if (iz == kNoCell) return;
    if (ringcell) {
        const bool inside = iz >= 1 && iz <= g.nz && ix >= 1 && ix <= g.nx;
        gparts_sh(a, GMasked<T>{&a, a.ringx, !inside}, iz, ix, qx, qz);
    }
    const size_t gi = gcell(a, iz, ix);
    const int iz_first = a.zt == 1 ? 1 : a.zt + 1;
    if (ix < 2 - g.npml || ix > g.nx + g.npml - 1 || iz < iz_first || iz > g.nz + g.npml - 1) {
        a.uxo[gi] = a.ux[gi];
        return;
    }
    T X, Z;
    gparts_sh_st(a, a.st, GIdxCell<T>{&a}, GField<T>{&a, a.ux}, iz, ix, X, Z);
    const bool surf = iz == 1 && a.zt == 1;
    {
        const int norder = g.norder;
        const int nox = gmin(gmin(ix + g.npml - 1, g.nx + g.npml - ix), norder);
        const int noz = gmin(gmin(iz - a.zt, g.nz + g.npml - iz), norder);
        T sx = (T)0.0f, sz = (T)0.0f;
        const T *Wx = a.st[2], *Wz = a.st[3];
        // a sum whose cells are all off the strips is a sum of zeros
        const bool xs = iz < 1 || iz > g.nz || ix - nox < 1 || ix + nox > g.nx;
        const bool zs = ix < 1 || ix > g.nx || iz - noz < 1 || iz + noz > g.nz;
        for (int k = 1; xs && k <= nox; ++k)
            sx = gacc(sx, gco(a.c1, norder, k, nox), (T)(Wx[gcell(a, iz, ix + k)] - Wx[gcell(a, iz, ix - k)]));
        for (int k = 1; zs && k <= noz; ++k)
            sz = gacc(sz, gco(a.c1, norder, k, noz), (T)(Wz[gcell(a, iz + k, ix)] - Wz[gcell(a, iz - k, ix)]));
        X = X - sx / (T)g.dx;
        if (!surf) Z = Z - sz / (T)g.dz;
    }
    if (ringcell) {
        const bool inside = iz >= 1 && iz <= g.nz && ix >= 1 && ix <= g.nx;
        if (inside) { X = X + qx; Z = Z + qz; }
        else { X = X - qx; Z = Z - qz; }
    }
    const T bf = (T)(g.dt * g.dt) / a.rho[gi];
    const int s = gstrip(a, iz, ix);
    if (s < 0) {
        a.uxo[gi] = ((T)2.0f * a.ux[gi] - a.uxp[gi]) + bf * (X + Z);
        return;
    }
    const size_t o = gstrip_off(a, s, iz, ix);
    const int jx = ix + g.npml - 1, jz = iz - a.zt;
    const T ax = (T)0.5f * (a.dpx[jx] + a.ratio * a.dpz[jz]) * (T)g.dt;
    const T az = (T)0.5f * (a.dpz[jz] + a.ratio * a.dpx[jx]) * (T)g.dt;
    T *b1 = a.bu[s][0][0], *b2 = a.bu[s][0][1];
    const T n1 = ((T)2.0f * b1[2 * o] - ((T)1.0f - ax) * b1[2 * o + 1] + bf * X) / ((T)1.0f + ax);
    const T n2 = ((T)2.0f * b2[2 * o] - ((T)1.0f - az) * b2[2 * o + 1] + bf * Z) / ((T)1.0f + az);
    b1[2 * o + 1] = b1[2 * o]; b1[2 * o] = n1;
    b2[2 * o + 1] = b2[2 * o]; b2[2 * o] = n2;
    a.uxo[gi] = n1 + n2;
}

// reference c (1:3, rows zt.., columns) -> rho, lam, mu planes; the damped
// cells take the material of the outermost undamped cell along the normal
template <class T>
__global__ void k_gen_import_material(Geo g, int zt, const float *c, T *rho, T *lam, T *mu)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NZ) return;
    const int j = (int)(t / g.NZ), k = (int)(t % g.NZ);
    const int ix = j + 1 - g.npml, iz = zt + k;
    const int sx = ix < 1 - g.nbd ? 1 - g.nbd : (ix > g.nx + g.nbd ? g.nx + g.nbd : ix);
    const int sz = iz > g.nz + g.nbd ? g.nz + g.nbd : (iz < 1 - g.nbd ? 1 - g.nbd : iz);
    const long src = (long)(sx + g.npml - 1) * g.NZ + (sz - zt);
    const size_t d = (size_t)j * g.ld + k;
    rho[d] = (T)c[3 * src + 0];
    lam[d] = (T)c[3 * src + 1];
    mu[d] = (T)c[3 * src + 2];
}

// harmonic means at the x- and z-faces of every cell for the moduli of
// the modes in [m0, m1] (0 lam+2mu, 1 mu); the last column / row has no
// face and holds 0
template <class T>
__global__ void k_gen_faces(Geo g, int zt, const T *lam, const T *mu, int m0, int m1, T *hx0, T *hz0, T *hx1, T *hz1)
{
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NZ) return;
    const int j = (int)(t / g.NZ), k = (int)(t % g.NZ);
    const size_t c = (size_t)j * g.ld + k, cx = c + g.ld, cz = c + 1;
    for (int m = m0; m <= m1; ++m) {
        auto mat = [&](size_t q) { return m == 0 ? (T)(lam[q] + (T)2.0f * mu[q]) : mu[q]; };
        T *hx = m == 0 ? hx0 : hx1, *hz = m == 0 ? hz0 : hz1;
        hx[c] = j + 1 < g.NX ? ghmean(mat(cx), mat(c)) : (T)0.0f;
        hz[c] = k + 1 < g.NZ ? ghmean(mat(cz), mat(c)) : (T)0.0f;
    }
}

// dt^2/rho for the tiled interior kernels of the FP32 general path (the
// update factor the general kernels form per cell)
__global__ void k_gen_bf(long n, float dt2, const float *rho, float *bf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) bf[t] = dt2 / rho[t];
}

// reference (1:2, NZ, NX) <-> two padded arrays (level 1 -> cur, level 2 -> prev)
template <class T, class R>
__global__ void k_gen_levels(Geo g, int to_ref, R *ref2, T *cur, T *prev)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NZ) return;
    const size_t d = (size_t)(t / g.NZ) * g.ld + (size_t)(t % g.NZ);
    if (to_ref) {
        ref2[2 * t] = (R)cur[d];
        ref2[2 * t + 1] = (R)prev[d];
    } else {
        cur[d] = (T)ref2[2 * t];
        prev[d] = (T)ref2[2 * t + 1];
    }
}

// element-type conversion of a contiguous array (widen FP32 inputs for the
// FP64 path, round FP64 results for FP32 getters)
template <class D, class S>
__global__ void k_convert(long n, const S *src, D *dst)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (D)src[t];
}

}  // namespace fdfk
