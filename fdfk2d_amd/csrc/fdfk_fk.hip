// FK background on the GPU in FP64: the comput_background_wavefield role
// (src/FDFK2D.f90:432-434, binary-only libFK3D in the reference).
//
// Plane P or SV wave incident from the half-space of a 1-D layered model
// with a free surface (this project's formulation, z down, e^{iw(px+qz-t)}):
//   k_fk_amp    one thread per frequency: propagate the incident and the two
//               reflected half-space waves up through the layer matrices
//               Q_j = E_j^-1 E_{j+1}, impose zero traction at z = 0, keep the
//               four amplitudes of every layer
//   k_fk_spec   one thread per (frequency, depth): displacement spectrum
//               E_j[0:2] Lambda_j(z - z_j) a_j at each ring depth
//   hipFFT Z2D  frequency -> time per depth (complex frequency w + i eps
//               against wrap-around, undone by e^{eps t})
//   k_fk_fill   every ring point and iteration: the depth trace delayed by
//               p |x - x_entry| (plane-wave translation invariance), 4-point
//               cubic Lagrange interpolation, written in the reference's
//               twelve-array layout (src/FDFK2D.f90:330-355)
// Incident pulse ds exp(-(pi f0 (t - t0))^2), t0 = 1.5/f0, centred at the
// entry column at depth max(ring bottom, half-space top).
#include <hipfft/hipfft.h>

namespace fdfk {

struct zc {
    double r, i;
};
__host__ __device__ inline zc zmk(double r, double i) { return zc{r, i}; }
__host__ __device__ inline zc operator+(zc a, zc b) { return zc{a.r + b.r, a.i + b.i}; }
__host__ __device__ inline zc operator-(zc a, zc b) { return zc{a.r - b.r, a.i - b.i}; }
__host__ __device__ inline zc operator*(zc a, zc b) { return zc{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__host__ __device__ inline zc operator*(double s, zc a) { return zc{s * a.r, s * a.i}; }
__host__ __device__ inline zc zdiv(zc a, zc b)
{
    double d = b.r * b.r + b.i * b.i;
    return zc{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
__host__ __device__ inline zc zexp(zc a)
{
    double e = exp(a.r);
    return zc{e * cos(a.i), e * sin(a.i)};
}
__host__ __device__ inline zc zconj(zc a) { return zc{a.r, -a.i}; }
__host__ __device__ inline double zabs(zc a) { return hypot(a.r, a.i); }
__host__ __device__ inline zc zI(zc a) { return zc{-a.i, a.r}; }  // i*a

constexpr int kMaxLayer = 16;

struct FKModel {
    int nl;
    double zt[kMaxLayer];
    double mu[kMaxLayer];      // rho vs^2 (SH)
    zc ea[kMaxLayer], eb[kMaxLayer];
    zc E[kMaxLayer][4][4];
    zc Q[kMaxLayer][4][4];
    double tref;
};

static zc vslow_h(double v, double p)
{
    double a = 1.0 / (v * v) - p * p;
    return a >= 0.0 ? zmk(std::sqrt(a), 0.0) : zmk(0.0, std::sqrt(-a));
}

// E (rows ux, uz, szz/iw, sxz/iw; columns P down, S down, P up, S up)
static void emat_h(double vp, double vs, double rho, double p, zc ea, zc eb, zc E[4][4])
{
    double g = 1.0 - 2.0 * vs * vs * p * p;
    E[0][0] = zmk(vp * p, 0); E[0][1] = vs * eb; E[0][2] = zmk(vp * p, 0); E[0][3] = vs * eb;
    E[1][0] = vp * ea; E[1][1] = zmk(-vs * p, 0); E[1][2] = -vp * ea; E[1][3] = zmk(vs * p, 0);
    E[2][0] = zmk(rho * vp * g, 0); E[2][1] = (-2.0 * rho * vs * vs * vs * p) * eb;
    E[2][2] = zmk(rho * vp * g, 0); E[2][3] = (-2.0 * rho * vs * vs * vs * p) * eb;
    E[3][0] = (2.0 * rho * vs * vs * vp * p) * ea; E[3][1] = zmk(rho * vs * g, 0);
    E[3][2] = (-2.0 * rho * vs * vs * vp * p) * ea; E[3][3] = zmk(-rho * vs * g, 0);
}

// A X = B, 4x4 complex, B is 4x4 (Gaussian elimination, partial pivoting)
static bool solve4_h(zc A[4][4], zc B[4][4])
{
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (zabs(A[r][c]) > zabs(A[piv][c])) piv = r;
        if (zabs(A[piv][c]) == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < 4; ++k) { std::swap(A[c][k], A[piv][k]); std::swap(B[c][k], B[piv][k]); }
        for (int r = c + 1; r < 4; ++r) {
            zc f = zdiv(A[r][c], A[c][c]);
            for (int k = c; k < 4; ++k) A[r][k] = A[r][k] - f * A[c][k];
            for (int k = 0; k < 4; ++k) B[r][k] = B[r][k] - f * B[c][k];
        }
    }
    for (int c = 3; c >= 0; --c)
        for (int k = 0; k < 4; ++k) {
            zc s = B[c][k];
            for (int j = c + 1; j < 4; ++j) s = s - A[c][j] * B[j][k];
            B[c][k] = zdiv(s, A[c][c]);
        }
    return true;
}

static bool fk_model_setup(FKModel &m, int nl, const float *fk, double p)
{
    if (nl < 1 || nl > kMaxLayer) return false;
    m.nl = nl;
    for (int j = 0; j < nl; ++j) {
        double zt = fk[4 * j], vp = fk[4 * j + 1], vs = fk[4 * j + 2], rho = fk[4 * j + 3];
        if (!(vs > 0.0) || !(vp > vs) || !(rho > 0.0)) return false;
        m.zt[j] = zt;
        m.mu[j] = rho * vs * vs;
        m.ea[j] = vslow_h(vp, p);
        m.eb[j] = vslow_h(vs, p);
        emat_h(vp, vs, rho, p, m.ea[j], m.eb[j], m.E[j]);
    }
    for (int j = 0; j + 1 < nl; ++j) {
        zc A[4][4], B[4][4];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) { A[r][c] = m.E[j][r][c]; B[r][c] = m.E[j + 1][r][c]; }
        if (!solve4_h(A, B)) return false;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) m.Q[j][r][c] = B[r][c];
    }
    return true;
}

struct FKRun {
    int nfft, nk;          // nk = nfft/2 frequencies used (Nyquist dropped)
    double T, eps, dtf;
    double f0, ds;
    int src_is_p;
    int nzq;               // ring depths z0 + (q + zf - 1) dz, q = 0..nzq-1
    int free_surface;      // 1: traction-free top; 0: the top layer extends upward (absorbing top layer)
    int sh;                // SH incidence: one component (uy), stored in the x slots
};

// per frequency: amplitudes of all layers, amps[(k*nl + j)*4 + c]
__global__ void k_fk_amp(const FKModel *mp, FKRun r, zc *amps, zc *work)
{
    const FKModel &m = *mp;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.nk) return;
    const int N = m.nl;
    const double PI = 3.14159265358979323846;
    zc w = zmk(2.0 * PI * k / r.T, r.eps);
    zc x = zmk(w.r / (2.0 * PI * r.f0), w.i / (2.0 * PI * r.f0));
    zc S = (r.ds / (r.f0 * sqrt(PI))) * zexp(zmk(-(x.r * x.r - x.i * x.i), -2.0 * x.r * x.i));
    zc ainc = S * zexp(zI(w * zmk(m.tref, 0.0)));
    // basis amplitudes of each layer: work[((k*N + j)*3 + b)*4 + c]
    zc *ab = work + (size_t)k * N * 12;
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 4; ++c) ab[((N - 1) * 3 + b) * 4 + c] = zmk(0, 0);
    ab[((N - 1) * 3 + 0) * 4 + (r.src_is_p ? 2 : 3)] = zmk(1, 0);
    ab[((N - 1) * 3 + 1) * 4 + 0] = zmk(1, 0);
    ab[((N - 1) * 3 + 2) * 4 + 1] = zmk(1, 0);
    for (int j = N - 2; j >= 0; --j) {
        double h = m.zt[j + 1] - m.zt[j];
        zc iwh_a = zI(w * m.ea[j]), iwh_b = zI(w * m.eb[j]);
        zc inv[4] = {zexp(-h * iwh_a), zexp(-h * iwh_b), zexp(h * iwh_a), zexp(h * iwh_b)};
        for (int b = 0; b < 3; ++b)
            for (int rr = 0; rr < 4; ++rr) {
                zc s = zmk(0, 0);
                for (int c = 0; c < 4; ++c) s = s + m.Q[j][rr][c] * ab[((j + 1) * 3 + b) * 4 + c];
                ab[(j * 3 + b) * 4 + rr] = inv[rr] * s;
            }
    }
    // the two top conditions of each basis vector: zero traction at the
    // surface, or (no free surface) no down-going P and S in the top layer
    zc T[3][2];
    for (int b = 0; b < 3; ++b)
        for (int rr = 0; rr < 2; ++rr) {
            zc s = zmk(0, 0);
            if (r.free_surface)
                for (int c = 0; c < 4; ++c) s = s + m.E[0][2 + rr][c] * ab[(0 * 3 + b) * 4 + c];
            else
                s = ab[(0 * 3 + b) * 4 + rr];
            T[b][rr] = s;
        }
    zc det = T[1][0] * T[2][1] - T[2][0] * T[1][1];
    zc r0 = zmk(0, 0) - T[0][0] * ainc, r1 = zmk(0, 0) - T[0][1] * ainc;
    zc x1 = zdiv(r0 * T[2][1] - T[2][0] * r1, det);
    zc x2 = zdiv(T[1][0] * r1 - r0 * T[1][1], det);
    for (int j = 0; j < N; ++j)
        for (int c = 0; c < 4; ++c)
            amps[((size_t)k * N + j) * 4 + c] =
                ainc * ab[(j * 3 + 0) * 4 + c] + x1 * ab[(j * 3 + 1) * 4 + c] + x2 * ab[(j * 3 + 2) * 4 + c];
}

// SH plane wave, per frequency: b = (uy, syz/(iw)) = E a, a = (S down, S
// up), E = [[1, 1], [mu eta, -mu eta]]; E^-1 b = (u + s/(mu eta), u - s/(mu
// eta)) / 2.  The incident S up-going and the leaving down-going S of the
// half-space propagated up; top: traction free or no down-going wave.
// amps[(k*nl + j)*4 + c], c = 0 down, 1 up.
__global__ void k_fk_amp_sh(const FKModel *mp, FKRun r, zc *amps, zc *work)
{
    const FKModel &m = *mp;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.nk) return;
    const int N = m.nl;
    const double PI = 3.14159265358979323846;
    zc w = zmk(2.0 * PI * k / r.T, r.eps);
    zc x = zmk(w.r / (2.0 * PI * r.f0), w.i / (2.0 * PI * r.f0));
    zc S = (r.ds / (r.f0 * sqrt(PI))) * zexp(zmk(-(x.r * x.r - x.i * x.i), -2.0 * x.r * x.i));
    zc ainc = S * zexp(zI(w * zmk(m.tref, 0.0)));
    zc *ab = work + (size_t)k * N * 4;   // [layer][basis][comp]
    ab[((N - 1) * 2 + 0) * 2 + 0] = zmk(0, 0); ab[((N - 1) * 2 + 0) * 2 + 1] = zmk(1, 0);
    ab[((N - 1) * 2 + 1) * 2 + 0] = zmk(1, 0); ab[((N - 1) * 2 + 1) * 2 + 1] = zmk(0, 0);
    for (int j = N - 2; j >= 0; --j) {
        const double h = m.zt[j + 1] - m.zt[j];
        const zc mek = m.mu[j + 1] * m.eb[j + 1], mej = m.mu[j] * m.eb[j];
        const zc iwh = zI(w * m.eb[j]);
        const zc ph[2] = {zexp(-h * iwh), zexp(h * iwh)};
        for (int b = 0; b < 2; ++b) {
            const zc ad = ab[((j + 1) * 2 + b) * 2 + 0], au = ab[((j + 1) * 2 + b) * 2 + 1];
            const zc u = ad + au, sg = mek * (ad - au);
            const zc q = zdiv(sg, mej);
            ab[(j * 2 + b) * 2 + 0] = ph[0] * (0.5 * (u + q));
            ab[(j * 2 + b) * 2 + 1] = ph[1] * (0.5 * (u - q));
        }
    }
    const zc me0 = m.mu[0] * m.eb[0];
    zc T[2];
    for (int b = 0; b < 2; ++b)
        T[b] = r.free_surface ? me0 * (ab[b * 2 + 0] - ab[b * 2 + 1]) : ab[b * 2 + 0];
    const zc x1 = zdiv(zmk(0, 0) - T[0] * ainc, T[1]);
    for (int j = 0; j < N; ++j)
        for (int c = 0; c < 2; ++c)
            amps[((size_t)k * N + j) * 4 + c] = ainc * ab[(j * 2 + 0) * 2 + c] + x1 * ab[(j * 2 + 1) * 2 + c];
}

// SH displacement spectra at the ring depths (uy in the first component,
// zero in the second)
__global__ void k_fk_spec_sh(const FKModel *mp, FKRun r, const zc *amps, const int *lay, const double *dzl, zc *spec)
{
    const FKModel &m = *mp;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nkp = r.nfft / 2 + 1;
    if (t >= (long)r.nzq * nkp) return;
    int q = (int)(t / nkp), k = (int)(t % nkp);
    zc uy = zmk(0, 0);
    if (k < r.nk) {
        const double PI = 3.14159265358979323846;
        zc w = zmk(2.0 * PI * k / r.T, r.eps);
        int j = lay[q];
        zc iwb = zI(w * m.eb[j]);
        const zc *a = amps + ((size_t)k * m.nl + j) * 4;
        uy = zexp(dzl[q] * iwb) * a[0] + zexp(-dzl[q] * iwb) * a[1];
    }
    spec[(size_t)q * nkp + k] = zconj(uy);
    spec[((size_t)r.nzq + q) * nkp + k] = zmk(0, 0);
}

// displacement spectra at the ring depths, conjugated for the e^{+} inverse
// transform of hipFFT: spec[((comp*nzq + q)*(nfft/2+1)) + k]
__global__ void k_fk_spec(const FKModel *mp, FKRun r, const zc *amps, const int *lay, const double *dzl, zc *spec)
{
    const FKModel &m = *mp;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nkp = r.nfft / 2 + 1;
    if (t >= (long)r.nzq * nkp) return;
    int q = (int)(t / nkp), k = (int)(t % nkp);
    zc ux = zmk(0, 0), uz = zmk(0, 0);
    if (k < r.nk) {
        const double PI = 3.14159265358979323846;
        zc w = zmk(2.0 * PI * k / r.T, r.eps);
        int j = lay[q];
        double d = dzl[q];
        zc iwa = zI(w * m.ea[j]), iwb = zI(w * m.eb[j]);
        zc ph[4] = {zexp(d * iwa), zexp(d * iwb), zexp(-d * iwa), zexp(-d * iwb)};
        const zc *a = amps + ((size_t)k * m.nl + j) * 4;
        for (int c = 0; c < 4; ++c) {
            zc pa = ph[c] * a[c];
            ux = ux + m.E[j][0][c] * pa;
            uz = uz + m.E[j][1][c] * pa;
        }
    }
    spec[(size_t)q * nkp + k] = zconj(ux);
    spec[((size_t)r.nzq + q) * nkp + k] = zconj(uz);
}

// undo the complex-frequency damping and the 1/T of the discrete inverse
__global__ void k_fk_scale(FKRun r, double *g, long ntr)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr * r.nfft) return;
    int n = (int)(t % r.nfft);
    g[t] = g[t] * exp(r.eps * n * r.dtf) / r.T;
}

__device__ inline double fk_interp(const double *g, int nfft, double dtf, double t)
{
    if (t < 0.0) return 0.0;
    double u = t / dtf;
    int i = (int)floor(u);
    double f = u - i;
    if (i + 2 >= nfft / 2) return 0.0;
    double gm1 = i >= 1 ? g[i - 1] : 0.0, g0 = g[i], g1 = g[i + 1], g2 = g[i + 2];
    double wm1 = -f * (f - 1.0) * (f - 2.0) / 6.0, w0 = (f + 1.0) * (f - 1.0) * (f - 2.0) / 2.0;
    double w1 = -(f + 1.0) * f * (f - 2.0) / 2.0, w2 = (f + 1.0) * f * (f - 1.0) / 6.0;
    return wm1 * gm1 + w0 * g0 + w1 * g1 + w2 * g2;
}

struct FillArgs {
    Geo g;
    int t0, t1, period;         // steps [t0, t1), step it stored at slab it % period
    int nfft, nzq;
    double dt, dtf, x0, dx, p, sgn, xe;
    const double *gl, *gr;      // time traces [comp][q][nfft] of the left / right model
    int strip;                  // 0 L reg, 1 R reg, 2 B reg, 3 L pml, 4 R pml, 5 B pml, 6 T reg, 7 T pml
    int rows, cols, row0, col0; // strip shape: rows row0.., columns col0..
    int zf;                     // row of trace depth index 0
};

// ax, az: the destination arrays of the strip, FP32 (the reference's
// arrays) or FP64 (the FP64 storage mode)
template <class R>
__global__ void k_fk_fill(FillArgs f, R *ax, R *az)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long per = (long)f.rows * f.cols;
    if (t >= per * (f.t1 - f.t0)) return;
    int it = f.t0 + (int)(t / per);
    long rem = t % per;
    int ix = f.col0 + (int)(rem / f.rows), iz = f.row0 + (int)(rem % f.rows);
    double x = f.x0 + (double)(ix - 1) * f.dx;
    double tau = f.p * (f.sgn > 0 ? (x - f.xe) : (f.xe - x));
    double tt = it * f.dt - tau;
    int q = iz - f.zf;
    size_t nz2 = (size_t)f.nzq * f.nfft;
    const double *gxl = f.gl + (size_t)q * f.nfft, *gzl = f.gl + nz2 + (size_t)q * f.nfft;
    const double *gxr = f.gr + (size_t)q * f.nfft, *gzr = f.gr + nz2 + (size_t)q * f.nfft;
    double bx, bz;
    bool side = (f.strip == 0 || f.strip == 1 || f.strip == 3 || f.strip == 4);
    if (side) {
        bool left = (f.strip == 0 || f.strip == 3);
        bx = fk_interp(left ? gxl : gxr, f.nfft, f.dtf, tt);
        bz = fk_interp(left ? gzl : gzr, f.nfft, f.dtf, tt);
    } else {
        double w = (double)(ix - 1) / (double)(f.g.nx - 1);
        w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
        double lx = fk_interp(gxl, f.nfft, f.dtf, tt), lz = fk_interp(gzl, f.nfft, f.dtf, tt);
        double rx = fk_interp(gxr, f.nfft, f.dtf, tt), rz = fk_interp(gzr, f.nfft, f.dtf, tt);
        bx = (1.0 - w) * lx + w * rx;
        bz = (1.0 - w) * lz + w * rz;
    }
    size_t o = (size_t)(iz - f.row0) + (size_t)f.rows * ((size_t)(ix - f.col0) + (size_t)f.cols * (it % f.period));
    ax[o] = (R)(f.sgn * bx);
    az[o] = (R)bz;
}

}  // namespace fdfk

namespace {

// Work buffers of one background computation, allocated and freed in the
// order of the solver's stream (no device-wide wait when they are freed,
// so other streams' queued work, such as a running batch, is not waited for)
struct DevBuf {
    void *p = nullptr;
    hipStream_t st = nullptr;
    hipError_t alloc(size_t bytes, hipStream_t stream)
    {
        st = stream;
        return malloc_async_retry(&p, bytes, stream);
    }
    ~DevBuf() { if (p) (void)hipFreeAsync(p, st); }
    template <class T> T *as() { return (T *)p; }
};

// Inverse-FFT plans, kept for the life of the process and shared by shape:
// a caller takes an idle plan of its shape or makes one, and gives it back.
// Plan set-up uploads device tables; reusing plans keeps it out of the path
// of every background computation after the first of a shape.
struct PlanPool {
    struct Entry { int nfft, nkp, batch; hipfftHandle h; size_t work; bool busy; };
    std::mutex mu;
    std::vector<Entry> plans;
};
PlanPool &plan_pool()
{
    static PlanPool *pool = new PlanPool();   // never destroyed: plans outlive static teardown
    return *pool;
}
// A plan of the given shape, made without its own work area: the caller
// supplies `work` bytes of device memory per execution (stream-ordered), so
// idle plans hold only their small tables and a device-wide free never
// follows from dropping one.
static hipfftResult make_plan(int nfft, int nkp, int batch, hipfftHandle &h, size_t &work)
{
    hipfftResult r = hipfftCreate(&h);
    if (r != HIPFFT_SUCCESS) return r;
    int n[1] = {nfft};
    if ((r = hipfftSetAutoAllocation(h, 0)) != HIPFFT_SUCCESS ||
        (r = hipfftMakePlanMany(h, 1, n, nullptr, 1, nkp, nullptr, 1, nfft, HIPFFT_Z2D, batch, &work)) !=
            HIPFFT_SUCCESS)
        (void)hipfftDestroy(h);
    return r;
}
int plan_take(int nfft, int nkp, int batch, hipfftHandle &h, size_t &work)
{
    PlanPool &P = plan_pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        for (auto &e : P.plans)
            if (!e.busy && e.nfft == nfft && e.nkp == nkp && e.batch == batch) {
                e.busy = true;
                h = e.h;
                work = e.work;
                return 0;
            }
    }
    // with the device full, free the kept blocks of destroyed solvers and the
    // idle plans, clear the failed call's error state and try once more
    if (make_plan(nfft, nkp, batch, h, work) != HIPFFT_SUCCESS) {
        (void)hipGetLastError();
        if (!cache_release() || make_plan(nfft, nkp, batch, h, work) != HIPFFT_SUCCESS) {
            (void)hipGetLastError();
            return fail(FDFK_EDEVICE, "hipfftMakePlanMany failed");
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.plans.push_back({nfft, nkp, batch, h, work, true});
    return 0;
}
// Idle plans kept for reuse (plans of many shapes, batches of many grid
// depths, would otherwise accumulate); returned plans beyond this count are
// destroyed, oldest first (a plan is idle only after its stream finished).
constexpr int kPlanKeep = 16;
void plan_give(hipfftHandle h)
{
    PlanPool &P = plan_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto &e : P.plans)
        if (e.h == h) e.busy = false;
    int idle = 0;
    for (auto &e : P.plans) idle += !e.busy;
    for (auto it = P.plans.begin(); idle > kPlanKeep && it != P.plans.end();) {
        if (!it->busy) {
            (void)hipfftDestroy(it->h);
            it = P.plans.erase(it);
            --idle;
        } else {
            ++it;
        }
    }
}

int next_pow2(long n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// traces of one model on the device: g[comp][q][nfft].  One stream
// synchronisation in all (after the scaling): the model, the layer of every
// depth and its depth below the layer top go up in one packed copy from a
// host buffer that lives until then, and the plan and its work area are
// given back only after it.  Under a serving load every synchronisation
// waits behind the other streams' work on the shared hardware queues.
int fk_traces_dev(fdfk_solver *s, FKModel &m, const FKRun &r, const double *zq, double *g_out)
{
    const int nzq = r.nzq;
    const long nkp = r.nfft / 2 + 1;
    hipfftHandle plan;
    size_t wbytes = 0;
    if (int rc = plan_take(r.nfft, (int)nkp, 2 * nzq, plan, wbytes)) return rc;
    const size_t ol = (sizeof(FKModel) + 15) / 16 * 16;
    const size_t od = ol + ((size_t)nzq * sizeof(int) + 15) / 16 * 16;
    std::vector<char> host(od + (size_t)nzq * sizeof(double));
    std::memcpy(host.data(), &m, sizeof(FKModel));
    int *lay = (int *)(host.data() + ol);
    double *dzl = (double *)(host.data() + od);
    for (int q = 0; q < nzq; ++q) {
        int j = 0;
        while (j + 1 < m.nl && zq[q] >= m.zt[j + 1]) ++j;
        lay[q] = j;
        dzl[q] = zq[q] - m.zt[j];
    }
    DevBuf up, amps, work, spec, wa;
    hipError_t e = up.alloc(host.size(), s->stream);
    if (e == hipSuccess) e = amps.alloc((size_t)r.nk * m.nl * 4 * sizeof(zc), s->stream);
    if (e == hipSuccess) e = work.alloc((size_t)r.nk * m.nl * 12 * sizeof(zc), s->stream);
    if (e == hipSuccess) e = spec.alloc((size_t)2 * nzq * nkp * sizeof(zc), s->stream);
    if (e == hipSuccess && wbytes > 0) e = wa.alloc(wbytes, s->stream);
    if (e != hipSuccess) {
        plan_give(plan);
        return fail(FDFK_ENOMEM, std::string("FK work arrays: ") + hipGetErrorString(e));
    }
    e = hipMemcpyAsync(up.p, host.data(), host.size(), hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->stream);
        plan_give(plan);
        return fail(FDFK_EDEVICE, std::string("FK model upload: ") + hipGetErrorString(e));
    }
    const FKModel *dm = (const FKModel *)up.p;
    const int *dlay = (const int *)((char *)up.p + ol);
    const double *ddz = (const double *)((char *)up.p + od);
    long ns = (long)nzq * nkp;
    if (r.sh) {
        hipLaunchKernelGGL(k_fk_amp_sh, dim3((r.nk + 127) / 128), dim3(128), 0, s->stream, dm, r, amps.as<zc>(),
                           work.as<zc>());
        hipLaunchKernelGGL(k_fk_spec_sh, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s->stream, dm, r,
                           amps.as<zc>(), dlay, ddz, spec.as<zc>());
    } else {
        hipLaunchKernelGGL(k_fk_amp, dim3((r.nk + 127) / 128), dim3(128), 0, s->stream, dm, r, amps.as<zc>(),
                           work.as<zc>());
        hipLaunchKernelGGL(k_fk_spec, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s->stream, dm, r,
                           amps.as<zc>(), dlay, ddz, spec.as<zc>());
    }
    hipfftSetStream(plan, s->stream);
    hipfftSetWorkArea(plan, wa.p);
    const hipfftResult fr = hipfftExecZ2D(plan, (hipfftDoubleComplex *)spec.p, (hipfftDoubleReal *)g_out);
    const long tot = 2L * nzq * r.nfft;
    if (fr == HIPFFT_SUCCESS)
        hipLaunchKernelGGL(k_fk_scale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream, r, g_out,
                           2L * nzq);
    e = hipStreamSynchronize(s->stream);
    plan_give(plan);
    if (fr != HIPFFT_SUCCESS) return fail(FDFK_EDEVICE, "hipfftExecZ2D failed");
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(FDFK_EDEVICE, std::string("FK traces: ") + hipGetErrorString(e));
    return 0;
}

// destroys every idle plan; true when one was
bool plan_pool_release_idle()
{
    PlanPool &P = plan_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    bool any = false;
    for (auto it = P.plans.begin(); it != P.plans.end();) {
        if (!it->busy) {
            (void)hipfftDestroy(it->h);
            it = P.plans.erase(it);
            any = true;
        } else {
            ++it;
        }
    }
    return any;
}

}  // namespace

extern "C" int fdfk_compute_background(fdfk_solver *s, const fdfk_fk *fk)
{
    if (!s || !fk || !fk->fk_left || !fk->fk_right || !fk->bazs) return fail(FDFK_EINVAL, "null argument");
    if (!(fk->f0 > 0)) return fail(FDFK_EINVAL, "f0 must be > 0");
    const Geo &g = s->g;
    const int nbd = g.nbd, nx = g.nx, nz = g.nz, nt = s->nt;
    double cs = 0.0;
    for (int i = 0; i < g.NX; ++i) cs += std::cos(fk->bazs[i]);
    double rayp = fk->rayp;
    if (std::fabs(rayp) < 1e-9) rayp = rayp < 0 ? -1e-9 : 1e-9;  // src/FDFK2D.f90:431
    // travel away from the event: along-profile slowness -p cos(baz)
    const double sgn = (rayp < 0.0 ? -1.0 : 1.0) * (cs < 0.0 ? 1.0 : -1.0);
    const double p = std::fabs(rayp);
    const double f0 = fk->f0, t0 = 1.5 / f0, dt = g.dt;
    FKModel ml{}, mr{};
    if (!fk_model_setup(ml, fk->nlayer_left, fk->fk_left, p) || !fk_model_setup(mr, fk->nlayer_right, fk->fk_right, p))
        return fail(FDFK_EINVAL, "invalid FK model (need 1..16 solid layers, vp > vs > 0)");
    // depths of the ring rows zf..nz+nbd; an absorbing top layer adds the
    // top ring (rows 1-nbd..nbd), above the model top the top layer continues
    const bool top = s->zt != 1;
    const int zf = top ? 1 - nbd : 1;
    const int nzq = nz + nbd - zf + 1;
    std::vector<double> zq(nzq);
    for (int q = 0; q < nzq; ++q) zq[q] = (double)s->z0 + (double)(q + zf - 1) * g.dz;
    const double xl = (double)s->x0 + (double)(1 - nbd - 1) * g.dx, xr = (double)s->x0 + (double)(nx + nbd - 1) * g.dx;
    const double xe = sgn > 0 ? xl : xr;
    const double zref = std::max(std::max(zq[nzq - 1], ml.zt[ml.nl - 1]), std::max(zq[nzq - 1], mr.zt[mr.nl - 1]));
    FKRun r{};
    r.dtf = dt / 2.0;
    const double tneed = nt * dt + p * (xr - xl) + 2.0 * t0 + 10.0 / f0;
    r.nfft = next_pow2((long)std::ceil(2.0 * tneed / r.dtf));
    r.nk = r.nfft / 2;
    r.T = r.nfft * r.dtf;
    r.eps = 27.6 / r.T;
    r.f0 = f0; r.ds = fk->ds; r.src_is_p = fk->src_is_p ? 1 : 0; r.nzq = nzq;
    r.free_surface = top ? 0 : 1;
    r.sh = s->sh ? 1 : 0;
    if (r.sh) r.src_is_p = 0;   // an SH solver's incident wave is an S wave
    ml.tref = t0 + (r.src_is_p ? ml.ea[ml.nl - 1].r : ml.eb[ml.nl - 1].r) * (zref - ml.zt[ml.nl - 1]);
    mr.tref = t0 + (r.src_is_p ? mr.ea[mr.nl - 1].r : mr.eb[mr.nl - 1].r) * (zref - mr.zt[mr.nl - 1]);
    const bool same = fk->nlayer_left == fk->nlayer_right &&
                      std::memcmp(fk->fk_left, fk->fk_right, sizeof(float) * 4 * fk->nlayer_left) == 0;
    if (fk->block_steps < 0) return fail(FDFK_EINVAL, "block_steps must be >= 0");
    const int B = fk->block_steps;
    const bool blocked = B > 0 && 2L * B < (long)nt + 1;
    const int period = blocked ? 2 * B : nt + 1;
    int rc = alloc_background(s, period);
    if (rc) return rc;
    // traces: transient when every step is stored at once, kept with the
    // solver when later blocks are synthesised from them
    DevBuf gl, gr;
    const size_t ntr = (size_t)2 * nzq * r.nfft;
    double *tl = nullptr, *tr = nullptr;
    if (blocked) {
        if ((rc = s->dalloc(&tl, ntr))) return rc;
        if (!same && (rc = s->dalloc(&tr, ntr))) return rc;
    } else {
        HIPCHK(gl.alloc(ntr * sizeof(double), s->stream));
        tl = gl.as<double>();
        if (!same) {
            HIPCHK(gr.alloc(ntr * sizeof(double), s->stream));
            tr = gr.as<double>();
        }
    }
    if ((rc = fk_traces_dev(s, ml, r, zq.data(), tl))) return rc;
    if (!same && (rc = fk_traces_dev(s, mr, r, zq.data(), tr))) return rc;
    FillArgs f{};
    f.g = g; f.period = period; f.nfft = r.nfft; f.nzq = nzq; f.dt = dt; f.dtf = r.dtf;
    f.x0 = s->x0; f.dx = g.dx; f.p = p; f.sgn = sgn; f.xe = xe;
    f.gl = tl; f.gr = same ? tl : tr;
    f.zf = zf;
    // strip shapes: (rows, cols, row0, col0) of the twelve arrays, FDFK2D.f90:330-355,
    // then the top ring's inside and outside strips
    const std::array<std::array<int, 4>, 8> shp = {{{nz - nbd, nbd, 1, 1}, {nz - nbd, nbd, 1, nx - nbd + 1},
                                                   {nbd, nx, nz - nbd + 1, 1}, {nz, nbd, 1, 1 - nbd},
                                                   {nz, nbd, 1, nx + 1}, {nbd, nx + 2 * nbd, nz + 1, 1 - nbd},
                                                   {nbd, nx - 2 * nbd, 1, nbd + 1}, {nbd, nx + 2 * nbd, 1 - nbd, 1 - nbd}}};
    // destinations: the reference's FP32 arrays, or the FP64 path's copies
    const bool f64 = s->gen && s->gsize == 8;
    std::array<void *, 16> dst{};
    for (int i = 0; i < 12; ++i) dst[i] = f64 ? s->gbg[i] : (void *)s->bg[i];
    for (int i = 0; i < 4; ++i) dst[12 + i] = f64 ? s->gbg[12 + i] : (void *)s->bg_top[i];
    const int nstrip = top ? 8 : 6;
    auto fill = [f, shp, dst, nstrip, f64](hipStream_t st, int t0, int t1) -> int {
        FillArgs a = f;
        a.t0 = t0; a.t1 = t1;
        for (int k = 0; k < nstrip; ++k) {
            a.strip = k;
            a.rows = shp[k][0]; a.cols = shp[k][1]; a.row0 = shp[k][2]; a.col0 = shp[k][3];
            const long n = (long)a.rows * a.cols * (t1 - t0);
            if (f64)
                hipLaunchKernelGGL(k_fk_fill<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                                   (double *)dst[2 * k], (double *)dst[2 * k + 1]);
            else
                hipLaunchKernelGGL(k_fk_fill<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a,
                                   (float *)dst[2 * k], (float *)dst[2 * k + 1]);
        }
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : fail(FDFK_EDEVICE, std::string("background fill: ") + hipGetErrorString(e));
    };
    // the block the run is in and the next one (all steps when not blocked)
    const int first = blocked ? (s->it / B) * B : 0;
    if ((rc = fill(s->stream, first, std::min(first + period, nt + 1)))) return rc;
    s->bg_block = blocked ? B : 0;
    s->bg_fill = blocked ? std::function<int(hipStream_t, int, int)>(fill) : nullptr;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    s->bg_valid = true;
    s->bg_top_valid = top;
    return 0;
}

extern "C" void comput_background_wavefield_(const int32_t *nx, const int32_t *nz, const int32_t *npml,
    const int32_t *izt, const int32_t *nbd, const int32_t *nt, const int32_t *nlayer_left,
    const int32_t *nlayer_right, const int32_t *src_type, const float *x0, const float *z0,
    const float *dx, const float *dz, const double *dt, const float *ds, const float *f0,
    const double *rayp, const double *bazs, const float *fk_left, const float *fk_right,
    float *uxl_reg, float *uzl_reg, float *uxr_reg, float *uzr_reg, float *uxb_reg,
    float *uzb_reg, float *uxl_pml, float *uzl_pml, float *uxr_pml, float *uzr_pml,
    float *uxb_pml, float *uzb_pml)
{
    // the reference's error contract is print and stop
    auto die = [](const char *m) {
        std::fprintf(stderr, "comput_background_wavefield (fdfk2d_amd): %s\n", m);
        std::fflush(stderr);
        std::exit(1);
    };
    // host checks first: a bad configuration stops before any GPU work
    if (*izt != 1 && *izt != 1 - *npml) die("izt must be 1 (free surface) or 1-npml (is_PML_top)");
    if (*nbd < 1 || *nbd % 2 == 0) die("nbd must be 2*norder-1");
    const int norder = (*nbd + 1) / 2;
    if (validate_geometry(*nx, *nz, *npml, norder, *dx, *dz, *dt, *nt)) die(fdfk_last_error());
    if (!bazs || !fk_left || !fk_right) die("null input array");
    float *const out[12] = {uxl_reg, uzl_reg, uxr_reg, uzr_reg, uxb_reg, uzb_reg,
                            uxl_pml, uzl_pml, uxr_pml, uzr_pml, uxb_pml, uzb_pml};
    for (float *o : out)
        if (!o) die("null output array");
    // a handle with the geometry and the background storage only: no
    // fields, materials or strips are needed to synthesise the background
    auto *s = new fdfk_solver();
    std::unique_ptr<fdfk_solver> own(s);
    s->bg_only = true;
    s->zt = *izt;   // izt < 1: no free surface, and the top ring is computed too
    Geo &g = s->g;
    g.nx = *nx; g.nz = *nz; g.npml = *npml; g.norder = norder; g.nbd = *nbd;
    g.NX = g.nx + 2 * g.npml; g.NZ = g.nz + g.npml;
    g.ld = (g.NZ + 31) / 32 * 32;
    g.dx = *dx; g.dz = *dz; g.dt = *dt;
    s->nt = *nt; s->x0 = *x0; s->z0 = *z0;
    if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess)
        die("no HIP device / stream");
    fdfk_fk fk{};
    fk.nlayer_left = *nlayer_left; fk.nlayer_right = *nlayer_right;
    fk.fk_left = fk_left; fk.fk_right = fk_right; fk.src_is_p = *src_type != 0;
    fk.ds = *ds; fk.f0 = *f0; fk.rayp = *rayp; fk.bazs = bazs;
    if (fdfk_compute_background(s, &fk) || fdfk_get_background(s, out)) die(fdfk_last_error());
    if (s->zt != 1) {
        // the top ring, for updatedispl_ with the same uxl_reg
        std::array<std::vector<float>, 4> t;
        for (int i = 0; i < 4; ++i) t[i].assign((size_t)s->bg_top_n[i], 0.0f);
        float *const t4[4] = {t[0].data(), t[1].data(), t[2].data(), t[3].data()};
        if (fdfk_get_top_background(s, t4)) die(fdfk_last_error());
        std::lock_guard<std::mutex> lk(g_compat_mu);
        g_top_ring[uxl_reg] = std::move(t);
    }
}
