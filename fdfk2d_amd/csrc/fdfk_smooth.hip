// Model smoother on the GPU: utils/gauss_smooth.f90:310-397 (the reference's
// Gaussian or triangular smoothing of an SU model, one trace of nz samples
// per column x).  Included by fdfk_api.hip.
//
// Methods 0 and 1 evaluate the reference's window sum cell by cell in its
// FP32 arithmetic: weights normalised by their FP32 sum in the array's
// memory order (x outer, z inner), then per cell
//     mysum += w * vel, mywgt += w   (x offset outer, z offset inner)
// with a separate rounding after each multiply and add, and mysum / mywgt.
// Cost (2 nwx + 1)(2 nwz + 1) products per cell.
// Method 2 is the Gaussian as two 1-D passes (x, then z) with FP64 sums and
// an FP32 intermediate: the 2-D window exp(-(x/lx)^2/2 - (z/lz)^2/2),
// truncated at ceil(3.6 l / d) cells and normalised by its sum, is the
// product of two 1-D windows truncated and normalised alike, and the mirror
// fold acts per axis.  Cost 2 (2n + 1) FP64 products per cell; same result as
// method 0 up to rounding.
// Both fold indices once at each edge (1-based: 0 -> 1, -1 -> 2, n+1 -> n),
// so a window half-width must be smaller than the model's extent.

namespace fdfk {

__host__ __device__ inline int smooth_fold(int i, int n)
{
    if (i < 0) i = -1 - i;
    if (i >= n) i = 2 * n - 1 - i;
    return i;
}

// one thread per window cell, z fastest (coalesced reads of each shifted row)
__global__ void __launch_bounds__(256) k_smooth_window(const float *__restrict__ v,
    float *__restrict__ out, int nx, int nz, const float *__restrict__ w, int nwx, int nwz,
    float wsum, int ix0, int nxw, int iz0, int nzw)
{
#pragma clang fp contract(off)
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nxw * nzw) return;
    const int i = ix0 + (int)(idx / nzw);
    const int j = iz0 + (int)(idx % nzw);
    const int lz = 2 * nwz + 1;
    float acc = 0.0f;
    for (int i1 = -nwx; i1 <= nwx; ++i1) {
        const float *row = v + (size_t)smooth_fold(i + i1, nx) * nz;
        const float *wc = w + (size_t)(i1 + nwx) * lz + nwz;
        // unrolled so loads run ahead; the sums stay in order
#pragma unroll 8
        for (int j1 = -nwz; j1 <= nwz; ++j1) {
            const float p = wc[j1] * row[smooth_fold(j + j1, nz)];
            acc = acc + p;
        }
    }
    out[(size_t)i * nz + j] = acc / wsum;
}

// 1-D FP64 pass along x (stride nz) or z (stride 1); cells of columns
// ix0..ix1, rows iz0..iz1 (0-based) are written, the others keep out's values
template <bool ALONG_X>
__global__ void __launch_bounds__(256) k_smooth_pass(const float *__restrict__ in,
    float *__restrict__ out, int nx, int nz, const double *__restrict__ w, int n, int ix0, int ix1,
    int iz0, int iz1)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nx * nz) return;
    const int i = (int)(idx / nz);
    const int j = (int)(idx % nz);
    if (i < ix0 || i > ix1 || j < iz0 || j > iz1) return;
    double s = 0.0;
    if (ALONG_X) {
        for (int k = -n; k <= n; ++k)
            s += w[k + n] * (double)in[(size_t)smooth_fold(i + k, nx) * nz + j];
    } else {
        const float *col = in + (size_t)i * nz;
        for (int k = -n; k <= n; ++k)
            s += w[k + n] * (double)col[smooth_fold(j + k, nz)];
    }
    out[idx] = (float)s;
}

// Weight window w[(j + nwx)(2 nwz + 1) + (k + nwz)] (j along x, k along z:
// the memory order of the reference's w(k, j)) and the per-cell weight sum
// in the loop order, gauss_smooth.f90:310-360, 391.
static int smooth_weights(float lambdax, float lambdaz, float dx, float dz, int method, int &nwx,
                          int &nwz, std::vector<float> &w, float &wsum)
{
#pragma clang fp contract(off)
    if (method == 0) {
        nwx = (int)std::ceil(3.6f * lambdax / dx);
        nwz = (int)std::ceil(3.6f * lambdaz / dz);
        const int lz = 2 * nwz + 1;
        w.assign((size_t)(2 * nwx + 1) * lz, 0.0f);
        for (int j = -nwx; j <= nwx; ++j) {
            const float rx = ((float)j * dx) / lambdax;
            for (int k = -nwz; k <= nwz; ++k) {
                const float rz = ((float)k * dz) / lambdaz;
                const float q = rx * rx + rz * rz;         // single precision, as declared
                w[(size_t)(j + nwx) * lz + (k + nwz)] = (float)std::exp(-0.5 * (double)q);
            }
        }
    } else {
        nwx = (int)std::ceil(2.1f * lambdax / dx);
        nwz = (int)std::ceil(2.1f * lambdaz / dz);
        if (nwx != nwz)
            return fail(FDFK_EINVAL, "triangular smoothing needs equal half-widths in x and z "
                                     "(the reference's in-place weights read outside their array otherwise)");
        const int n = nwx, l = 2 * n + 1;
        w.assign((size_t)l * l, 0.0f);
        auto W = [&](int k, int j) -> float & { return w[(size_t)(j + n) * l + (k + n)]; };
        float sigma2 = (float)(1.8 * (double)std::ceil(lambdaz / dz));
        for (int j = -n; j <= n; ++j)
            for (int k = -n; k <= n; ++k) W(k, j) = (float)((double)(nwz - std::abs(k)) / (double)sigma2);
        sigma2 = (float)(1.8 * (double)std::ceil(lambdax / dx));
        // in place, reading (j, k) after its own update when column k came first
        for (int j = -n; j <= n; ++j)
            for (int k = -n; k <= n; ++k)
                W(k, j) = (float)((double)(nwx - std::abs(j)) / (double)sigma2 * (double)W(j, k));
    }
    float total = 0.0f;
    for (float x : w) total = total + x;
    for (float &x : w) x = x / total;
    wsum = 0.0f;
    for (float x : w) wsum = wsum + x;
    return FDFK_OK;
}

// the two normalised 1-D FP64 Gaussian windows of method 2, half-widths
// n = ceil(3.6 l / d)
static int smooth_weights_1d(float lambdax, float lambdaz, float dx, float dz, int nx, int nz, int n[2],
                             std::vector<double> w[2])
{
    n[0] = (int)std::ceil(3.6f * lambdax / dx);
    n[1] = (int)std::ceil(3.6f * lambdaz / dz);
    if (n[0] >= nx || n[1] >= nz)
        return fail(FDFK_EINVAL, "smoothing window wider than the model (the mirror folds once)");
    const double lam[2] = {(double)lambdax, (double)lambdaz}, d[2] = {(double)dx, (double)dz};
    for (int a = 0; a < 2; ++a) {
        w[a].assign(2 * n[a] + 1, 0.0);
        double tot = 0.0;
        for (int k = -n[a]; k <= n[a]; ++k) {
            const double r = (double)k * d[a] / lam[a];
            w[a][k + n[a]] = std::exp(-0.5 * r * r);
            tot += w[a][k + n[a]];
        }
        for (double &x : w[a]) x /= tot;
    }
    return FDFK_OK;
}

static int smooth_check(const float *vel, const float *out, int nx, int nz, float lambdax, float lambdaz,
                        float dx, float dz, int iz0, int iz1, int ix0, int ix1, int method)
{
    if (!vel || !out) return fail(FDFK_EINVAL, "vel and out must be non-null");
    if (nx < 1 || nz < 1) return fail(FDFK_EINVAL, "model must hold at least one cell");
    if (!(lambdax > 0.0f) || !(lambdaz > 0.0f) || !(dx > 0.0f) || !(dz > 0.0f))
        return fail(FDFK_EINVAL, "lambdax, lambdaz, dx and dz must be positive");
    if (iz0 < 1 || iz1 > nz || iz0 > iz1 || ix0 < 1 || ix1 > nx || ix0 > ix1)
        return fail(FDFK_EINVAL, "window must satisfy 1 <= iz0 <= iz1 <= nz and 1 <= ix0 <= ix1 <= nx");
    if (method < 0 || method > 2) return fail(FDFK_EINVAL, "method must be 0, 1 or 2");
    return FDFK_OK;
}

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() { for (void *q : p) if (q) (void)hipFree(q); }
    template <class T> hipError_t get(T **out, size_t n)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 1);
        p.push_back(q);
        *out = (T *)q;
        return e;
    }
};

}  // namespace fdfk

extern "C" int fdfk_smooth_model(const float *vel, float *out, int32_t nx, int32_t nz, float lambdax,
                                 float lambdaz, float dx, float dz, int32_t iz0, int32_t iz1,
                                 int32_t ix0, int32_t ix1, int32_t method)
{
    using namespace fdfk;
    if (int rc = smooth_check(vel, out, nx, nz, lambdax, lambdaz, dx, dz, iz0, iz1, ix0, ix1, method)) return rc;
    const size_t cells = (size_t)nx * nz;
    const int nxw = ix1 - ix0 + 1, nzw = iz1 - iz0 + 1;
    DevBufs b;
    float *dv = nullptr, *dout = nullptr;
    HIPCHK(b.get(&dv, cells));
    HIPCHK(b.get(&dout, cells));
    HIPCHK(hipMemcpy(dv, vel, cells * sizeof(float), hipMemcpyHostToDevice));
    if (method <= 1) {
        int nwx = 0, nwz = 0;
        float wsum = 0.0f;
        std::vector<float> w;
        if (int rc = smooth_weights(lambdax, lambdaz, dx, dz, method, nwx, nwz, w, wsum)) return rc;
        if (nwx >= nx || nwz >= nz)
            return fail(FDFK_EINVAL, "smoothing window wider than the model (the mirror folds once)");
        float *dw = nullptr;
        HIPCHK(b.get(&dw, w.size()));
        HIPCHK(hipMemcpy(dw, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
        const int64_t n = (int64_t)nxw * nzw;
        hipLaunchKernelGGL(k_smooth_window, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dv,
                           dout, nx, nz, dw, nwx, nwz, wsum, ix0 - 1, nxw, iz0 - 1, nzw);
        HIPCHK(hipGetLastError());
    } else {
        int n[2];
        std::vector<double> wv[2];
        if (int rc = smooth_weights_1d(lambdax, lambdaz, dx, dz, nx, nz, n, wv)) return rc;
        double *dw[2] = {nullptr, nullptr};
        for (int a = 0; a < 2; ++a) {
            const std::vector<double> &w = wv[a];
            HIPCHK(b.get(&dw[a], w.size()));
            HIPCHK(hipMemcpy(dw[a], w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        const unsigned nb = (unsigned)((cells + 255) / 256);
        hipLaunchKernelGGL(k_smooth_pass<true>, dim3(nb), dim3(256), 0, 0, dv, dout, nx, nz, dw[0], n[0], 0,
                           nx - 1, 0, nz - 1);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_smooth_pass<false>, dim3(nb), dim3(256), 0, 0, dout, dv, nx, nz, dw[1], n[1], 0,
                           nx - 1, 0, nz - 1);
        HIPCHK(hipGetLastError());
        dout = dv;
    }
    if (out != vel) std::memmove(out, vel, cells * sizeof(float));
    // only the window comes back; the rest of out keeps vel's values
    HIPCHK(hipMemcpy2D(out + (size_t)(ix0 - 1) * nz + (iz0 - 1), (size_t)nz * sizeof(float),
                       dout + (size_t)(ix0 - 1) * nz + (iz0 - 1), (size_t)nz * sizeof(float),
                       (size_t)nzw * sizeof(float), (size_t)nxw, hipMemcpyDeviceToHost));
    return FDFK_OK;
}

// ---------------------------------------------------------------------
// Allocation-free smoother on device buffers.  The scratch memory (weight
// tables, and the intermediate model of the two-pass method) is one
// file-scope device array of kSmoothScratchBytes per GPU, held for the life
// of the process.  In host code the array's name is only a stand-in: its
// device address is taken once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to copies and
// kernels.  Calls on one GPU share the array: calls that overlap on a GPU
// (other streams or threads) race.  Nothing in it must start at zero; every
// call writes what it reads.
// ---------------------------------------------------------------------
namespace fdfk {

constexpr size_t kSmoothScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kSmoothScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (resolved on first use)
static int smooth_scratch(void **ptr)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FDFK_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(FDFK_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return FDFK_OK;
}

}  // namespace fdfk

extern "C" int fdfk_smooth_model_dev(const float *vel, float *out, int32_t nx, int32_t nz, float lambdax,
                                     float lambdaz, float dx, float dz, int32_t iz0, int32_t iz1, int32_t ix0,
                                     int32_t ix1, int32_t method, void *stream)
{
    using namespace fdfk;
    if (int rc = smooth_check(vel, out, nx, nz, lambdax, lambdaz, dx, dz, iz0, iz1, ix0, ix1, method)) return rc;
    if (vel == out) return fail(FDFK_EINVAL, "out must not alias vel (the window reads neighbours of written cells)");
    hipStream_t st = (hipStream_t)stream;
    const size_t cells = (size_t)nx * nz;
    const int nxw = ix1 - ix0 + 1, nzw = iz1 - iz0 + 1;
    auto align = [](size_t b) { return (b + 255) / 256 * 256; };
    // weights (host), then the scratch layout and its limit, before any GPU work
    std::vector<float> wf;
    std::vector<double> wd[2];
    int nwx = 0, nwz = 0, n[2] = {0, 0};
    float wsum = 0.0f;
    size_t need = 0;
    if (method <= 1) {
        if (int rc = smooth_weights(lambdax, lambdaz, dx, dz, method, nwx, nwz, wf, wsum)) return rc;
        if (nwx >= nx || nwz >= nz)
            return fail(FDFK_EINVAL, "smoothing window wider than the model (the mirror folds once)");
        need = align(wf.size() * sizeof(float));
    } else {
        if (int rc = smooth_weights_1d(lambdax, lambdaz, dx, dz, nx, nz, n, wd)) return rc;
        need = align(wd[0].size() * sizeof(double)) + align(wd[1].size() * sizeof(double)) +
               align(cells * sizeof(float));
    }
    if (need > kSmoothScratchBytes)
        return fail(FDFK_EINVAL, "model and window need " + std::to_string(need) + " bytes of scratch; the "
                                 "allocation-free smoother has " + std::to_string(kSmoothScratchBytes) +
                                 " (use fdfk_smooth_model)");
    void *base = nullptr;
    if (int rc = smooth_scratch(&base)) return rc;
    unsigned char *sp = (unsigned char *)base;
    // cells outside the window keep vel's values
    HIPCHK(hipMemcpyAsync(out, vel, cells * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (method <= 1) {
        float *dw = (float *)sp;
        HIPCHK(hipMemcpyAsync(dw, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice, st));
        const int64_t nw = (int64_t)nxw * nzw;
        hipLaunchKernelGGL(k_smooth_window, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, vel, out, nx,
                           nz, dw, nwx, nwz, wsum, ix0 - 1, nxw, iz0 - 1, nzw);
    } else {
        double *dw0 = (double *)sp;
        double *dw1 = (double *)(sp + align(wd[0].size() * sizeof(double)));
        float *mid = (float *)(sp + align(wd[0].size() * sizeof(double)) + align(wd[1].size() * sizeof(double)));
        HIPCHK(hipMemcpyAsync(dw0, wd[0].data(), wd[0].size() * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dw1, wd[1].data(), wd[1].size() * sizeof(double), hipMemcpyHostToDevice, st));
        const unsigned nb = (unsigned)((cells + 255) / 256);
        hipLaunchKernelGGL(k_smooth_pass<true>, dim3(nb), dim3(256), 0, st, vel, mid, nx, nz, dw0, n[0], 0, nx - 1,
                           0, nz - 1);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_smooth_pass<false>, dim3(nb), dim3(256), 0, st, mid, out, nx, nz, dw1, n[1], ix0 - 1,
                           ix1 - 1, iz0 - 1, iz1 - 1);
    }
    HIPCHK(hipGetLastError());
    // the host weight tables are read by the queued copies: wait for them
    HIPCHK(hipStreamSynchronize(st));
    return FDFK_OK;
}
