// Host side of the C ABI declared in include/fdfk.h.
//
// A solver handle owns every array of one simulation on the GPU (fields at
// two time levels, materials, PML strips and memories, the hybrid
// background, receivers and seismograms) and a HIP stream; fdfk_run queues
// the per-iteration kernels of the reference loop (src/FDFK2D.f90:465-532)
// without host synchronisation.
#include "fdfk.h"
#include "fdfk_internal.h"
// single translation unit: the kernels are compiled with the host code
#include "fdfk_kernels.hip"
#include "fdfk_frame.hip"
#include "fdfk_interior2.hip"
#include "fdfk_interior_sh.hip"
#include "fdfk_general.hip"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <array>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace fdfk {

static thread_local std::string g_err;
static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(FDFK_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Host-to-device copy on a solver's own stream, waited for before return
// (sources may be locals).  Creation and background set-up go through the
// solver's stream, never the null stream, so that they add no ordering
// point to other solvers' or a batch's streams running at the same time.
static hipError_t h2d(hipStream_t st, void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

// Device blocks of destroyed solvers, kept for later solvers (per device,
// by size).  hipFree waits for the whole device and holds up other threads'
// launches meanwhile, so freeing a solver while others run (a batch, a
// server's requests) would stall them all; a destroyed solver's blocks go
// here instead, once its stream (ordered after any batch steps that used
// them) is idle.  A request takes the smallest cached block of at least its
// size and at most twice it, else allocates; when an allocation fails the
// device's cached blocks are freed and it is tried once more.
// fdfk_trim_memory frees the cache.
struct BlockCache {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void *>> free;   // device -> size -> block
};
static BlockCache &block_cache()
{
    static BlockCache *c = new BlockCache();   // outlives static teardown
    return *c;
}
}  // namespace fdfk
namespace {
bool plan_pool_release_idle();   // fdfk_fk.hip: idle inverse-FFT plans and their work areas
}
namespace fdfk {

// frees every block kept for the current device and the idle FFT plans;
// false when there was nothing to free
static bool cache_release()
{
    const bool plans = plan_pool_release_idle();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    std::multimap<size_t, void *> blocks;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        blocks.swap(c.free[dev]);
    }
    for (auto &b : blocks) (void)hipFree(b.second);
    return plans || !blocks.empty();
}

// stream-ordered allocation that, when the device is out of memory, frees
// the kept blocks of destroyed solvers and tries once more
static hipError_t malloc_async_retry(void **p, size_t bytes, hipStream_t st)
{
    hipError_t e = hipMallocAsync(p, bytes, st);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        if (cache_release()) e = hipMallocAsync(p, bytes, st);
    }
    return e;
}

static hipError_t block_get(size_t bytes, void **p, size_t *got)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        auto &m = c.free[dev];
        auto it = m.lower_bound(bytes);
        if (it != m.end() && it->first <= 2 * bytes) {
            *p = it->second;
            *got = it->first;
            m.erase(it);
            return hipSuccess;
        }
    }
    *got = bytes;
    e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        // the cache may hold blocks no request fits: give them back and retry once
        (void)hipGetLastError();
        if (cache_release()) e = hipMalloc(p, bytes);
    }
    return e;
}
// The kept blocks of a device are bounded by an eighth of its memory: a
// block that would exceed it (mixes of many sizes reuse few blocks) is
// freed, the largest kept blocks first.  The free waits for the device, as
// any free does; it happens only when the bound is reached.
static void block_put(int dev, void *p, size_t bytes)
{
    static size_t cap = 0;
    if (cap == 0) {
        size_t fr = 0, total = 0;
        cap = (hipMemGetInfo(&fr, &total) == hipSuccess && total > 0) ? total / 8 : (size_t(32) << 30);
    }
    std::vector<void *> drop;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        auto &m = c.free[dev];
        m.emplace(bytes, p);
        size_t held = 0;
        for (auto &b : m) held += b.first;
        while (held > cap && !m.empty()) {
            auto last = std::prev(m.end());
            held -= last->first;
            drop.push_back(last->second);
            m.erase(last);
        }
    }
    for (void *q : drop) (void)hipFree(q);
}

// FD coefficients: closed forms of src/Submain.f90:1111-1198 (fd_coef0:
// staggered first derivative, fd_coef1: collocated first derivative)
static void coef_table(int norder, bool staggered, double *c)
{
    std::memset(c, 0, sizeof(double) * norder * norder);
    for (int i = 1; i <= norder; ++i)
        for (int j = 1; j <= i; ++j) {
            double num = 1.0, lo = 1.0, hi = 1.0;
            auto node = [&](int k) { return staggered ? (double)((2 * k - 1) * (2 * k - 1)) : (double)(k * k); };
            for (int k = 1; k <= j - 1; ++k) { num *= node(k); lo *= node(j) - node(k); }
            for (int k = j + 1; k <= i; ++k) { num *= node(k); hi *= node(k) - node(j); }
            num *= (j % 2 == 1) ? 1.0 : -1.0;
            lo *= staggered ? (double)(2 * j - 1) : 2.0 * (double)j;
            c[(j - 1) + norder * (i - 1)] = num / lo / hi;
        }
}

}  // namespace fdfk

using namespace fdfk;

struct fdfk_solver {
    Geo g{};
    int nt = 0, nstp = 1, nsis = 0, nsamp = 1, nrcv = 0, precision = 0;
    int it = 0, istp = 0, isis = 0;
    float x0 = 0, z0 = 0, vpml = 0;
    double R0 = 1e-3;
    int npower = 2;
    hipStream_t stream = nullptr;
    hipEvent_t ev_order = nullptr;   // orders this stream against a batch's (join, leave)
    // FP32 general path with the tiled interior: the per-cell kernels of a
    // step run on a second stream beside the interior kernel (fork after the
    // prep launch, join before the next one)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int gen_overlap = 2;   // 2: the per-cell kernels first on the solver's stream, 1: the interior first
    std::vector<std::pair<void *, size_t>> allocs;   // device blocks and their sizes
    int device = 0;
    int64_t bytes = 0;
    bool bg_only = false;   // geometry and background only (comput_background_wavefield_)
    // fields
    float *ux[2] = {nullptr, nullptr}, *uz[2] = {nullptr, nullptr};
    int cur = 0;
    float *bf = nullptr, *lam = nullptr, *mu = nullptr;
    Coefs *dcoef = nullptr;
    Coefs hcoef{};
    // strips
    float *strip_fields[3][2][3][2] = {};  // [strip][comp][part][buffer]
    float *psx[3][2] = {}, *psz[3][2] = {};
    float *dpx = nullptr, *dpz = nullptr;   // classical damping profiles per column / row
    float *bx = nullptr, *bz = nullptr;     // exp(-d dt) of those profiles (memory decay)
    float pml_ratio = kPmlRatio;            // multiaxial ratio
    std::vector<float> h_dx, h_dz, h_bx, h_bz;
    long strip_cells[3] = {0, 0, 0};
    Strips *d_strips[2] = {nullptr, nullptr};  // device copies for the two buffer parities
    // background
    float *bg[12] = {};
    long bg_n[12] = {};
    bool bg_valid = false;
    // background storage: step it at slab it % bg_period; with bg_block > 0
    // the arrays hold two blocks and bg_fill(stream, t0, t1) synthesises
    // steps [t0, t1) from the kept FK traces
    int bg_period = 0, bg_block = 0;
    std::function<int(hipStream_t, int, int)> bg_fill;
    Background *d_bg = nullptr;
    float *ringx = nullptr, *ringz = nullptr;
    // receivers
    float *rx = nullptr, *rz = nullptr, *seisx = nullptr, *seisz = nullptr;
    float *stage = nullptr;   // (2, NZ, NX) staging for reference-layout copies
    // interior kernel
    int lx = 45, nt_threads = 256;
    bool fused = true;   // interior and frame in one launch (k_step)
    bool frame_full = true;   // full-order frame tiles specialised (FDFK_FRAME_FULL=0: general path only)
    int lx_model = 0;    // interior tile width of the occupancy model (lx may be tuned)
    // 16-bit material storage (FDFK_FP16, FDFK_BF16): lam, mu of the
    // interior tiles as per-row FP32 profiles plus 16-bit residuals, value =
    // profile + residual * scale (power-of-two scales)
    bool mat16 = false;
    bool matbf = false;                // the residuals in BF16 (FDFK_BF16) instead of FP16
    unsigned short *lam16 = nullptr, *mu16 = nullptr;
    float *lamp = nullptr, *mup = nullptr;   // per-row profiles the residuals are taken against
    float sl = 1.0f, sm = 1.0f;
    // Two-step pairs (FDFK_STEP2=1): the deep rectangle of the interior is
    // stepped twice per launch.  The pair writes out of place, into two
    // spare levels per component that rotate with the two live ones.
    bool step2 = false;
    float *uxs[2] = {nullptr, nullptr}, *uzs[2] = {nullptr, nullptr};
    Geo gdeep{};
    int d_lx = 0, d_nbx = 0, d_nbz = 0, d_tz = 0;
    Geo gband[4]{};
    int band_lx[4] = {0, 0, 0, 0};
    // General per-cell path (fdfk_general.hip): the absorbing top layer
    // (is_pml_top, rows from zt = 1-npml).  g.NZ and g.ld then count the top
    // rows; the background arrays, ring fields, receivers and seismograms
    // are the ones above, the fields, materials and strips these.
    // FDFK_FP64 runs it in double (every array below then holds doubles;
    // gsize is the element size).
    bool gen = false;
    bool sh = false;                   // transverse SH motion (one component, in the ux slots)
    int zt = 1;
    int gsize = 4;
    void *gux[2] = {nullptr, nullptr}, *guz[2] = {nullptr, nullptr};
    void *grho = nullptr, *glam = nullptr, *gmu = nullptr;
    void *ghx[2] = {}, *ghz[2] = {};   // face harmonic means [lam+2mu, mu]
    void *gst[12] = {};                // staged terms of level n (P-SV 12 planes, SH 4)
    void *gmq[16] = {};                // staged terms of the masked ring fields (ring neighbourhood)
    void *gbu[4][2][3] = {};           // split parts (1:2, cells), reference order
    void *gbpx[4][2] = {}, *gbpz[4][2] = {};
    double *gc0 = nullptr, *gc1 = nullptr;
    void *gdamp[4] = {};               // dpx, dpz, bx, bz
    void *gring[2] = {};               // ring fields
    void *gbg[16] = {};                // the twelve arrays and the top ring (FP32: bg[], bg_top[])
    void *gseis[2] = {};               // seismograms (FP32: seisx, seisz)
    double pml_ratio_d = 0.1;          // multiaxial ratio of the FP64 path
    float *bg_top[4] = {};             // top ring: uxt_reg, uzt_reg, uxt_pml, uzt_pml
    long bg_top_n[4] = {};
    bool bg_top_valid = false;         // without it the hybrid contour is open at the top
    // FP32 general path: the deep interior (columns 1..nx, rows 2*norder..nz)
    // runs on the tiled interior kernel (P-SV: k_interior2, SH:
    // k_interior_sh) over a view of the planes that starts at row 1, the
    // general kernels update the cells around it (FDFK_GEN_TILED=0: the
    // general kernels update every cell)
    bool gtiled = false;
    Geo gview{};
    float *gbf = nullptr;              // dt^2/rho of the tiled kernel
    int glx = 45;
    // profiling
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_int, ev_step;
    double int_ms = 0, step_ms = 0;
    int64_t int_n = 0, step_n = 0;

    template <class T>
    int dalloc(T **p, size_t n)
    {
        void *q = nullptr;
        if (n == 0) n = 1;
        size_t got = 0;
        hipError_t e = block_get(n * sizeof(T), &q, &got);
        if (e != hipSuccess) return fail(FDFK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.emplace_back(q, got);
        bytes += (int64_t)(n * sizeof(T));
        *p = (T *)q;
        return 0;
    }
    ~fdfk_solver()
    {
        if (stream) hipStreamSynchronize(stream);
        if (stream2) hipStreamSynchronize(stream2);
        if (ev_order) hipEventDestroy(ev_order);
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        for (auto &e : ev_int) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto &e : ev_step) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto &a : allocs) block_put(device, a.first, a.second);   // stream idle: reusable
        if (stream) hipStreamDestroy(stream);
        if (stream2) hipStreamDestroy(stream2);
    }
};

namespace {

int build_damping(fdfk_solver *s)
{
    const Geo &g = s->g;
    const int nx = g.nx, nz = g.nz, npml = g.npml, nbd = g.nbd;
    const int leff = npml - nbd;
    const double lr = std::log(s->R0);
    const double d0x = leff > 0 ? -(double)(s->npower + 1) * (double)s->vpml * lr / (2.0 * leff * (double)g.dx) : 0.0;
    const double d0z = leff > 0 ? -(double)(s->npower + 1) * (double)s->vpml * lr / (2.0 * leff * (double)g.dz) : 0.0;
    s->h_dx.assign(g.NX, 0.f); s->h_bx.assign(g.NX, 1.f);
    s->h_dz.assign(g.NZ, 0.f); s->h_bz.assign(g.NZ, 1.f);
    auto prof = [&](int q, double d0) {
        if (q <= nbd || leff <= 0) return 0.0;
        double r = (double)(q - nbd) / (double)leff, p = 1.0;
        for (int e = 0; e < s->npower; ++e) p *= r;
        return d0 * p;
    };
    for (int j = 0; j < g.NX; ++j) {
        int ix = j + 1 - npml;
        int q = ix <= 0 ? 1 - ix : (ix > nx ? ix - nx : 0);
        s->h_dx[j] = (float)prof(q, d0x);
        s->h_bx[j] = (float)std::exp(-(double)s->h_dx[j] * g.dt);
    }
    for (int k = 0; k < g.NZ; ++k) {
        int iz = k + s->zt;   // rows from zt: the top strip mirrors the bottom one
        s->h_dz[k] = iz > nz ? (float)prof(iz - nz, d0z) : (iz <= 0 ? (float)prof(1 - iz, d0z) : 0.0f);
        s->h_bz[k] = (float)std::exp(-(double)s->h_dz[k] * g.dt);
    }
    int rc;
    if ((rc = s->dalloc(&s->dpx, g.NX)) || (rc = s->dalloc(&s->dpz, g.NZ)) || (rc = s->dalloc(&s->bx, g.NX)) ||
        (rc = s->dalloc(&s->bz, g.NZ)))
        return rc;
    HIPCHK(h2d(s->stream, s->dpx, s->h_dx.data(), g.NX * 4));
    HIPCHK(h2d(s->stream, s->dpz, s->h_dz.data(), g.NZ * 4));
    HIPCHK(h2d(s->stream, s->bx, s->h_bx.data(), g.NX * 4));
    HIPCHK(h2d(s->stream, s->bz, s->h_bz.data(), g.NZ * 4));
    return 0;
}

int upload_strips(fdfk_solver *s)
{
    for (int par = 0; par < 2; ++par) {
        Strips h{};
        for (int st = 0; st < 3; ++st)
            for (int c = 0; c < 2; ++c) {
                for (int p = 0; p < 3; ++p) {
                    h.cur[st][c][p] = s->strip_fields[st][c][p][par];
                    h.prev[st][c][p] = s->strip_fields[st][c][p][1 - par];
                }
                h.psx[st][c] = s->psx[st][c];
                h.psz[st][c] = s->psz[st][c];
            }
        h.dpx = s->dpx; h.bx = s->bx; h.dpz = s->dpz; h.bz = s->bz; h.ratio = s->pml_ratio;
        if (!s->d_strips[par]) {
            int rc = s->dalloc(&s->d_strips[par], 1);
            if (rc) return rc;
        }
        HIPCHK(h2d(s->stream, s->d_strips[par], &h, sizeof(Strips)));
    }
    return 0;
}

StepArgs step_args(fdfk_solver *s)
{
    StepArgs a{};
    a.g = s->g;
    a.cf = s->dcoef;
    a.ux = s->ux[s->cur];
    a.uz = s->uz[s->cur];
    a.uxp = s->ux[1 - s->cur];
    a.uzp = s->uz[1 - s->cur];
    a.uxo = a.uxp; a.uzo = a.uzp;
    a.bf = s->bf; a.lam = s->lam; a.mu = s->mu;
    a.st = s->d_strips[s->cur];
    a.ringx = s->ringx; a.ringz = s->ringz;
    a.ring_valid = s->bg_valid && s->it <= s->nt;
    a.it = s->it;
    a.frame_full = s->frame_full ? 1 : 0;
    return a;
}

template <int NO>
ICoef<NO> make_icoef(const fdfk_solver *s)
{
    ICoef<NO> k{};
    const Geo &g = s->g;
    const double *c0 = s->hcoef.c0, *c1 = s->hcoef.c1;
    const int n = g.norder;
    for (int j = 1; j <= NO; ++j) {
        double a0 = c0[(j - 1) + n * (NO - 1)], b0 = c1[(j - 1) + n * (NO - 1)];
        k.ai[j - 1] = (float)a0;
        k.aox[j - 1] = (float)(2.0 * a0 / ((double)g.dx * g.dx));
        k.aoz[j - 1] = (float)(2.0 * a0 / ((double)g.dz * g.dz));
        k.bi[j - 1] = (float)b0;
        k.bo[j - 1] = (float)(b0 / ((double)g.dx * g.dz));
    }
    return k;
}

// Calls f(std::integral_constant<int, NO>{}) for NO = norder, 1..kMaxOrder
// (FD orders 2..16, the reference's range: src/Input.f90:119-131).
template <class F>
void with_order(int norder, F &&f)
{
    switch (norder) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
    }
}
static_assert(kMaxOrder == 8, "with_order covers norder 1..8");

// Interior tiling for a column width lx: nbx columns of tiles, nbz bands of
// full-height tiles (TZ rows); a last band of at most 64 - 2HZ rows goes to
// one-wave segments, NT/64 of them per block (nsb blocks from row zseg on).
// A rectangle of fewer rows than that goes to segments entirely when
// thin_segments is set (the thin bands of a two-step pair).
template <int NO, int NT>
void int_tiling(const Geo &g, int lx, int &nbx, int &nbz, int &nsb, int &zseg, bool thin_segments = false)
{
    constexpr int TZ = IntKernel2<NO, NT>::TZ, HZ = IntKernel2<NO, NT>::HZ, W = NT / 64;
    const int ncols = g.ix1 - g.ix0 + 1, nrows = g.iz1 - g.iz0 + 1;
    nbx = (ncols + lx - 1) / lx;
    const int full = nrows / TZ, rem = nrows - full * TZ;
    if (W > 1 && (full > 0 || thin_segments) && rem > 0 && rem <= 64 - 2 * HZ) {
        nbz = full;
        zseg = g.iz0 + full * TZ;
        nsb = (nbx + W - 1) / W;
    } else {
        nbz = (nrows + TZ - 1) / TZ;
        zseg = 0;
        nsb = 0;
    }
}

template <int NO, int NT>
long int_blocks(const Geo &g, int lx)
{
    int nbx, nbz, nsb, zseg;
    int_tiling<NO, NT>(g, lx, nbx, nbz, nsb, zseg);
    return (long)nbx * nbz + nsb;
}

template <int NO, int NT>
IntArgs<NO> interior_args(fdfk_solver *s)
{
    IntArgs<NO> a{};
    const Geo &g = s->g;
    a.g = g;
    a.k = make_icoef<NO>(s);
    a.ux = s->ux[s->cur]; a.uz = s->uz[s->cur];
    a.uxp = s->ux[1 - s->cur]; a.uzp = s->uz[1 - s->cur];
    a.uxo = a.uxp; a.uzo = a.uzp;
    a.bf = s->bf; a.lam = s->lam; a.mu = s->mu;
    a.ringx = s->ringx; a.ringz = s->ringz;
    a.lam16 = s->lam16; a.mu16 = s->mu16; a.lamp = s->lamp; a.mup = s->mup; a.sl = s->sl; a.sm = s->sm;
    a.corr = s->bg_valid && s->it <= s->nt;
    a.lx = s->lx;
    int_tiling<NO, NT>(g, a.lx, a.nbx, a.nbz, a.nsb, a.zseg);
    return a;
}

template <int NO, int NT>
void launch_interior_t(fdfk_solver *s)
{
    IntArgs<NO> a = interior_args<NO, NT>(s);
    hipLaunchKernelGGL((k_interior2<NO, NT>), dim3(a.nbx * a.nbz + a.nsb), dim3(NT), 0, s->stream, a);
}

// Interior block shape.  The stencil blocks all take about the same time,
// so for a block height NT the kernel time is (number of resident rounds) x
// (columns per block + warm-up columns); pick the column width that
// minimises it for the blocks one launch keeps resident (occupancy x CUs).
// The height is 256 lanes (measured fastest on the 8001x2000 solve; 128 and
// 192 can be requested for experiments).
struct IntPlan {
    int lx = 45;
    long cost = -1;
};

ETiles frame_tiles(const Geo &g);
size_t frame_lds_bytes(const ETiles &t, int norder);

template <int NO>
size_t step_lds_bytes(const Geo &g)
{
    return std::max(sizeof(typename IntKernel2<NO, 256>::Smem), frame_lds_bytes(frame_tiles(g), g.norder));
}

template <int NO, int NT>
IntPlan plan_lx(const Geo &g, bool fused, bool m16 = false)
{
    IntPlan best;
    int dev = 0, ncu = 0, bpc = 0;
    hipError_t oe;
    if constexpr (NT == 256)
        oe = fused ? (m16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_step<NO, 1>, NT, step_lds_bytes<NO>(g))
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_step<NO>, NT, step_lds_bytes<NO>(g)))
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_interior2<NO, NT>, NT, 0);
    else
        oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_interior2<NO, NT>, NT, 0);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        oe != hipSuccess || ncu <= 0 || bpc <= 0)
        return best;
    constexpr int WARM = IntKernel2<NO, NT>::WARM;
    const long slots = (long)ncu * bpc;
    const int ncols = g.ix1 - g.ix0 + 1;
    for (int lx = 4; lx <= std::max(4, ncols); ++lx) {
        const long rounds = (int_blocks<NO, NT>(g, lx) + slots - 1) / slots;
        const long cost = rounds * (lx + WARM);
        if (best.cost < 0 || cost < best.cost) { best.cost = cost; best.lx = lx; }
    }
    return best;
}

// Interior tile widths measured by fdfk_autotune, per geometry and launch
// form; plan_interior prefers them to the occupancy model.
struct TunedKey {
    int nx, nz, npml, norder, fused;
    bool operator<(const TunedKey &o) const
    {
        return std::tie(nx, nz, npml, norder, fused) < std::tie(o.nx, o.nz, o.npml, o.norder, o.fused);
    }
};
static std::mutex g_tuned_mu;
static std::map<TunedKey, int> *g_tuned = new std::map<TunedKey, int>();
static TunedKey tuned_key(const fdfk_solver *s)
{
    return TunedKey{s->g.nx, s->g.nz, s->g.npml, s->g.norder, s->fused ? 1 : 0};
}
static bool tuned_lx(const fdfk_solver *s, int *lx)
{
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned->find(tuned_key(s));
    if (it == g_tuned->end()) return false;
    *lx = it->second;
    return true;
}

// block height (threads along z) and column width of the interior launch;
// FDFK_NT / FDFK_LX override them (FDFK_FUSED=0 launches the interior and
// the frame separately)
void plan_interior(fdfk_solver *s)
{
    const Geo &g = s->g;
    const char *ent = std::getenv("FDFK_NT");
    const int want = ent ? std::atoi(ent) : 0;
    const char *ef = std::getenv("FDFK_FUSED");
    const bool fused = !(ef && std::atoi(ef) == 0) && want != 128 && want != 192;
    IntPlan p;
    int nt = 256;
    auto consider = [&](int n, IntPlan q) {
        if (q.cost >= 0 && (p.cost < 0 || q.cost < p.cost)) { p = q; nt = n; }
    };
    if (g.norder == 3 && want == 192) consider(192, plan_lx<3, 192>(g, false));
    else if (g.norder == 3 && want == 128) consider(128, plan_lx<3, 128>(g, false));
    else with_order(g.norder, [&](auto no) { consider(256, plan_lx<decltype(no)::value, 256>(g, fused, s->mat16)); });
    s->nt_threads = nt;
    s->fused = fused && nt == 256;
    s->lx = s->lx_model = p.lx;
    if (s->nt_threads == 256) tuned_lx(s, &s->lx);
    if (const char *e = std::getenv("FDFK_LX")) s->lx = std::max(4, std::atoi(e));
}

void launch_interior(fdfk_solver *s)
{
    if (s->g.norder == 3 && s->nt_threads == 128) launch_interior_t<3, 128>(s);
    else if (s->g.norder == 3 && s->nt_threads == 192) launch_interior_t<3, 192>(s);
    else with_order(s->g.norder, [&](auto no) { launch_interior_t<decltype(no)::value, 256>(s); });
}

// one frame band cut into tiles of at most kFrameThreads cells whose staged
// halo and intermediates fit the kernel's LDS
void add_band(ETiles &t, int x0, int x1, int z0, int z1, int norder, int view)
{
    if (x1 < x0 || z1 < z0) return;
    const int cols = x1 - x0 + 1, rows = z1 - z0 + 1;
    // the thin side of a band is split into equal tiles of at most 16
    int tx, tz;
    if (rows <= cols) {
        const int nz = (rows + 15) / 16;
        tz = (rows + nz - 1) / nz;
        tx = std::max(1, kFrameThreads / tz);
    } else {
        const int nxt = (cols + 15) / 16;
        tx = (cols + nxt - 1) / nxt;
        tz = std::max(1, kFrameThreads / tx);
    }
    while (frame_lds_floats(tx, tz, norder) > kFrameLds) {
        if (tx >= tz) --tx; else --tz;
    }
    int q = t.n++;
    t.x0[q] = x0; t.x1[q] = x1; t.z0[q] = z0; t.z1[q] = z1; t.tx[q] = tx; t.tz[q] = tz; t.view[q] = view;
    t.nbx[q] = (cols + tx - 1) / tx;
    t.nbz[q] = (rows + tz - 1) / tz;
    t.off[q + 1] = t.off[q] + t.nbx[q] * t.nbz[q];
}

// every updated cell outside the interior rectangle, in four bands that
// each lie on one side of the hybrid interface: the surface rows of the
// total-field region, the left and right strips (all rows of the model
// depth), and the bottom strip (all columns)
ETiles frame_tiles(const Geo &g)
{
    const int xa = 2 - g.npml, xb = g.nx + g.npml - 1;
    ETiles t{};
    add_band(t, 1, g.nx, 1, g.iz0 - 1, g.norder, 0);
    add_band(t, xa, 0, 1, g.nz, g.norder, 1);
    add_band(t, g.nx + 1, xb, 1, g.nz, g.norder, 1);
    add_band(t, xa, xb, g.nz + 1, g.nz + g.npml - 1, g.norder, 1);
    return t;
}

// dynamic LDS bytes of a frame launch: its largest tile
size_t frame_lds_bytes(const ETiles &t, int norder)
{
    int m = 0;
    for (int q = 0; q < t.n; ++q) m = std::max(m, frame_lds_floats(t.tx[q], t.tz[q], norder));
    return (size_t)m * sizeof(float);
}

template <int NO>
void launch_step_t(fdfk_solver *s, const StepArgs &a)
{
    IntArgs<NO> ia = interior_args<NO, 256>(s);
    ETiles t = frame_tiles(s->g);
    const int nint = ia.nbx * ia.nbz + ia.nsb;
    if (s->matbf)
        hipLaunchKernelGGL((k_step<NO, 2>), dim3((unsigned)(nint + t.off[t.n])), dim3(256),
                           step_lds_bytes<NO>(s->g), s->stream, ia, a, t, nint);
    else if (s->mat16)
        hipLaunchKernelGGL((k_step<NO, 1>), dim3((unsigned)(nint + t.off[t.n])), dim3(256),
                           step_lds_bytes<NO>(s->g), s->stream, ia, a, t, nint);
    else
        hipLaunchKernelGGL(k_step<NO>, dim3((unsigned)(nint + t.off[t.n])), dim3(256), step_lds_bytes<NO>(s->g),
                           s->stream, ia, a, t, nint);
}

void launch_step(fdfk_solver *s, const StepArgs &a)
{
    with_order(s->g.norder, [&](auto no) { launch_step_t<decltype(no)::value>(s, a); });
}

void launch_frame(fdfk_solver *s, const StepArgs &a)
{
    ETiles t = frame_tiles(s->g);
    if (t.off[t.n] == 0) return;
    dim3 fg((unsigned)t.off[t.n]), fb(kFrameThreads);
    const size_t sh = frame_lds_bytes(t, s->g.norder);
    with_order(s->g.norder, [&](auto no) {
        hipLaunchKernelGGL(k_frame<decltype(no)::value>, fg, fb, sh, s->stream, a, t);
    });
}

// ---------------------------------------------------------------------
// Two-step pairs over the deep interior
// ---------------------------------------------------------------------
template <int NO>
size_t step2_lds_bytes(const Geo &g, bool deep)
{
    size_t b = std::max(sizeof(typename IntKernel2<NO, 256>::Smem), frame_lds_bytes(frame_tiles(g), g.norder));
    if constexpr (NO <= 4)
        if (deep) b = std::max(b, sizeof(typename IntKernel22<NO>::Smem));
    return b;
}

// The deep rectangle: its tiles compute level n+1 on a halo of HZ cells
// around it for their second step, and that halo must consist of regular
// interior cells stepped without the padding's background correction.  So
// it starts HZ rows below the interior's first row (the rows above the
// interior read only physical rows), and ends 2HZ cells inside the
// interior's sides and bottom (the halo's stencils then read no padding).
// Returns false for orders above 4 (the two-engine march holds ring periods
// up to 14) or when no deep rectangle worth a launch remains.
template <int NO>
bool plan_step2_t(fdfk_solver *s)
{
    if constexpr (NO > 4) {
        (void)s;
        return false;
    } else {
        using K = IntKernel22<NO>;
        constexpr int HZ = K::HZ;
        const Geo &g = s->g;
        Geo d = g;
        d.ix0 = g.ix0 + 2 * HZ; d.ix1 = g.ix1 - 2 * HZ;
        d.iz0 = g.iz0 + HZ;     d.iz1 = g.iz1 - 2 * HZ;
        const int ncols = d.ix1 - d.ix0 + 1, nrows = d.iz1 - d.iz0 + 1;
        if (ncols < 4 * HZ || nrows < 2 * HZ) return false;
        int dev = 0, ncu = 0, bpc = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_step2<NO, true>, 256, step2_lds_bytes<NO>(g, true)) !=
                hipSuccess ||
            ncu <= 0 || bpc <= 0)
            return false;
        s->d_nbz = (nrows + K::TZMAX - 1) / K::TZMAX;
        s->d_tz = (nrows + s->d_nbz - 1) / s->d_nbz;
        // column width: fewest (resident rounds) x (march steps of a tile)
        const long slots = (long)ncu * bpc;
        long best = -1;
        for (int lx = 8; lx <= ncols; ++lx) {
            const long nb = (long)s->d_nbz * ((ncols + lx - 1) / lx);
            const long cost = (nb + slots - 1) / slots * (lx + K::WARM + K::LAG2);
            if (best < 0 || cost < best) { best = cost; s->d_lx = lx; }
        }
        if (const char *e = std::getenv("FDFK_LX2")) s->d_lx = std::max(8, std::atoi(e));
        s->d_nbx = (ncols + s->d_lx - 1) / s->d_lx;
        s->gdeep = d;
        // bands: above and below it (whole width), left and right of it
        Geo b[4] = {g, g, g, g};
        b[0].iz1 = d.iz0 - 1;
        b[1].iz0 = d.iz1 + 1;
        b[2].iz0 = b[3].iz0 = d.iz0;
        b[2].iz1 = b[3].iz1 = d.iz1;
        b[2].ix1 = d.ix0 - 1;
        b[3].ix0 = d.ix1 + 1;
        // short marches: the second launch of a pair holds only the bands and
        // the frame, so a band tile's march is on its critical path
        for (int k = 0; k < 4; ++k) {
            s->gband[k] = b[k];
            s->band_lx[k] = k < 2 ? 16 : b[k].ix1 - b[k].ix0 + 1;
        }
        return true;
    }
}

template <int NO>
void launch_step2_t(fdfk_solver *s, const StepArgs &a, bool first)
{
    if constexpr (NO > 4) {
        (void)s; (void)a; (void)first;
    } else {
        static_assert(sizeof(DeepArgs<NO>) + sizeof(Bands<NO>) + sizeof(StepArgs) + sizeof(ETiles) + 16 <= 4096,
                      "the pair launch's arguments exceed the kernel argument limit");
        ETiles t = frame_tiles(s->g);
        Bands<NO> bs{};
        for (int k = 0; k < 4; ++k) {
            IntArgs<NO> b = interior_args<NO, 256>(s);
            b.g = s->gband[k];
            b.lx = s->band_lx[k];
            int_tiling<NO, 256>(b.g, b.lx, b.nbx, b.nbz, b.nsb, b.zseg, true);
            b.uxo = a.uxo; b.uzo = a.uzo;
            bs.b[k] = b;
            bs.off[k + 1] = bs.off[k] + b.nbx * b.nbz + b.nsb;
        }
        DeepArgs<NO> da{};
        int nd = 0;
        if (first) {
            da.g = s->gdeep;
            da.k = make_icoef<NO>(s);
            da.ux = s->ux[s->cur]; da.uz = s->uz[s->cur];
            da.uxp = s->ux[1 - s->cur]; da.uzp = s->uz[1 - s->cur];
            da.o1x = s->uxs[0]; da.o1z = s->uzs[0];
            da.o2x = s->uxs[1]; da.o2z = s->uzs[1];
            da.bf = s->bf; da.lam = s->lam; da.mu = s->mu;
            da.lx = s->d_lx; da.nbx = s->d_nbx; da.nbz = s->d_nbz; da.tz = s->d_tz;
            nd = da.nbx * da.nbz;
            hipLaunchKernelGGL((k_step2<NO, true>), dim3((unsigned)(nd + bs.off[4] + t.off[t.n])), dim3(256),
                               step2_lds_bytes<NO>(s->g, true), s->stream, da, bs, a, t, nd);
        } else {
            hipLaunchKernelGGL((k_step2<NO, false>), dim3((unsigned)(bs.off[4] + t.off[t.n])), dim3(256),
                               step2_lds_bytes<NO>(s->g, false), s->stream, da, bs, a, t, nd);
        }
    }
}

void launch_step2(fdfk_solver *s, const StepArgs &a, bool first)
{
    with_order(s->g.norder, [&](auto no) { launch_step2_t<decltype(no)::value>(s, a, first); });
}

// Blocked background: entering block j >= 1 (iteration jB), synthesise
// block j+1 into the half of the ring that held block j-1, whose last step
// is already queued before this point of the stream.
int bg_refill(fdfk_solver *s, hipStream_t stream)
{
    const int B = s->bg_block;
    if (B <= 0 || !s->bg_fill || s->it < B || s->it % B != 0) return 0;
    const int t0 = s->it + B, t1 = std::min(s->it + 2 * B, s->nt + 1);
    return t0 < t1 ? s->bg_fill(stream, t0, t1) : 0;
}

// One time step.  half = 0: in place (level n+1 over level n-1).
// half = 1, 2: the first and second step of a two-step pair, written into
// the spare level uxs[half-1]; the spare and the replaced level n-1 trade
// places afterwards.  The first launch of a pair also writes level n+2 of
// the deep rectangle into uxs[1], which the second launch completes.
int one_step(fdfk_solver *s, bool record, int half = 0)
{
    const Geo &g = s->g;
    if (int rc = bg_refill(s, s->stream)) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    if (s->prof) {
        hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2); hipEventCreate(&e3);
        hipEventRecord(e0, s->stream);
    }
    RecArgs r{};
    const bool rec = record && s->nrcv > 0 && s->istp <= s->nsis;
    if (rec) {
        r.g = g; r.x0 = s->x0; r.z0 = s->z0; r.nrcv = s->nrcv; r.nsamp = s->nsamp; r.istp = s->istp;
        r.rx = s->rx; r.rz = s->rz; r.ux = s->ux[s->cur]; r.uz = s->uz[s->cur];
        r.seisx = s->seisx; r.seisz = s->seisz;
    }
    StepArgs a = step_args(s);
    if (half) { a.uxo = s->uxs[half - 1]; a.uzo = s->uzs[half - 1]; }
    {
        const int nring = a.ring_valid ? (int)((ring_count(g) + 255) / 256) : 0;
        const int npsi = (int)((2L * g.npml * g.nz + (long)g.NX * g.npml + 255) / 256);
        const int nrec = rec ? (s->nrcv + kRecPerBlock - 1) / kRecPerBlock : 0;
        dim3 pg((unsigned)(nring + npsi + nrec)), pb(256);
        with_order(g.norder, [&](auto no) {
            hipLaunchKernelGGL(k_prep<decltype(no)::value>, pg, pb, 0, s->stream, a, r, s->d_bg, nring, npsi);
        });
    }
    if (s->prof) hipEventRecord(e1, s->stream);
    if (half) {
        launch_step2(s, a, half == 1);
        if (s->prof) hipEventRecord(e2, s->stream);
    } else if (s->fused) {
        launch_step(s, a);
        if (s->prof) hipEventRecord(e2, s->stream);
    } else {
        launch_interior(s);
        if (s->prof) hipEventRecord(e2, s->stream);
        launch_frame(s, a);
    }
    if (s->prof) {
        hipEventRecord(e3, s->stream);
        s->ev_int.emplace_back(e1, e2);
        s->ev_step.emplace_back(e0, e3);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FDFK_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    if (half) {
        std::swap(s->ux[1 - s->cur], s->uxs[half - 1]);
        std::swap(s->uz[1 - s->cur], s->uzs[half - 1]);
    }
    s->cur = 1 - s->cur;
    s->it += 1;
    return 0;
}

// ---------------------------------------------------------------------
// General per-cell path (absorbing top layer)
// ---------------------------------------------------------------------
template <class T>
GenArgs<T> gen_args(fdfk_solver *s)
{
    GenArgs<T> a{};
    a.g = s->g;
    a.zt = s->zt;
    a.hx0 = 1; a.hx1 = 0;   // no hole: the step kernel updates every cell
    a.top = s->zt != 1;
    a.c0 = s->gc0; a.c1 = s->gc1;
    a.ux = (const T *)s->gux[s->cur]; a.uz = (const T *)s->guz[s->cur];
    a.uxp = (const T *)s->gux[1 - s->cur]; a.uzp = (const T *)s->guz[1 - s->cur];
    a.uxo = (T *)s->gux[1 - s->cur]; a.uzo = (T *)s->guz[1 - s->cur];
    a.rho = (const T *)s->grho; a.lam = (const T *)s->glam; a.mu = (const T *)s->gmu;
    for (int m = 0; m < 2; ++m) { a.hx[m] = (const T *)s->ghx[m]; a.hz[m] = (const T *)s->ghz[m]; }
    for (int st = 0; st < 4; ++st)
        for (int c = 0; c < 2; ++c) {
            for (int q = 0; q < 3; ++q) a.bu[st][c][q] = (T *)s->gbu[st][c][q];
            a.bpx[st][c] = (T *)s->gbpx[st][c];
            a.bpz[st][c] = (T *)s->gbpz[st][c];
        }
    a.dpx = (const T *)s->gdamp[0]; a.dpz = (const T *)s->gdamp[1];
    a.bx = (const T *)s->gdamp[2]; a.bz = (const T *)s->gdamp[3];
    a.ratio = sizeof(T) == 8 ? (T)s->pml_ratio_d : (T)s->pml_ratio;
    a.ringx = (const T *)s->gring[0]; a.ringz = (const T *)s->gring[1];
    a.ring_valid = s->bg_valid && s->it <= s->nt;
    a.top_ring = s->bg_top_valid;
    for (int i = 0; i < 16; ++i) a.bg[i] = (const T *)s->gbg[i];
    a.it = s->bg_period > 0 ? s->it % s->bg_period : 0;
    for (int i = 0; i < 12; ++i) a.st[i] = (T *)s->gst[i];
    for (int i = 0; i < 16; ++i) a.mq[i] = (T *)s->gmq[i];
    a.ex0 = 1; a.ex1 = 0; a.ez0 = 1; a.ez1 = 0;   // staged over every cell
    return a;
}

// Tiled interior of an FP32 general-path solver: the kernel arguments over
// the view of the planes from row 1 (offset 1-zt elements into each plane)
template <int NO>
IntArgs<NO> gen_interior_args(fdfk_solver *s)
{
    IntArgs<NO> a{};
    const size_t off = (size_t)(1 - s->zt);
    auto view = [off](void *p) { return p ? (float *)p + off : nullptr; };
    a.g = s->gview;
    a.k = make_icoef<NO>(s);
    a.ux = view(s->gux[s->cur]); a.uz = view(s->guz[s->cur]);
    a.uxp = view(s->gux[1 - s->cur]); a.uzp = view(s->guz[1 - s->cur]);
    a.uxo = a.uxp; a.uzo = a.uzp;
    a.bf = s->gbf + off; a.lam = view(s->glam); a.mu = view(s->gmu);
    a.ringx = view(s->gring[0]); a.ringz = view(s->gring[1]);
    a.corr = s->bg_valid && s->it <= s->nt;
    a.lx = s->glx;
    int_tiling<NO, 256>(a.g, a.lx, a.nbx, a.nbz, a.nsb, a.zseg);
    return a;
}

// its tile width: the occupancy model of plan_lx for the kernel it launches
template <int NO>
int gen_plan_lx(const Geo &v, bool sh)
{
    int dev = 0, ncu = 0, bpc = 0;
    const hipError_t oe = sh ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_interior_sh<NO, 256>, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_interior2<NO, 256>, 256, 0);
    if (oe != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0 || bpc <= 0)
        return 45;
    constexpr int WARM = IntKernel2<NO, 256>::WARM;
    const long slots = (long)ncu * bpc;
    long best = -1;
    int blx = 45;
    for (int lx = 4; lx <= std::max(4, v.ix1 - v.ix0 + 1); ++lx) {
        const long cost = ((int_blocks<NO, 256>(v, lx) + slots - 1) / slots) * (lx + WARM);
        if (best < 0 || cost < best) { best = cost; blx = lx; }
    }
    return blx;
}

// One iteration of the general path: prep (receivers before the update,
// ring values, memories), then the cell update (FP32 with the tiled
// interior: the interior rectangle, then the cells around it); out of place
// into level n-1, then the levels trade places.
template <class T>
int gen_step_t(fdfk_solver *s, bool record)
{
    const Geo &g = s->g;
    if (int rc = bg_refill(s, s->stream)) return rc;
    GenArgs<T> a = gen_args<T>(s);
    GenRec<T> r{};
    const bool rec = record && s->nrcv > 0 && s->istp <= s->nsis;
    if (rec) {
        r.nrcv = s->nrcv; r.nsamp = s->nsamp; r.istp = s->istp;
        r.x0 = (T)s->x0; r.z0 = (T)s->z0;
        r.rx = s->rx; r.rz = s->rz;
        r.seisx = (T *)s->gseis[0]; r.seisz = (T *)s->gseis[1];
    }
    const long nrec = rec ? s->nrcv : 0;
    int rb[8];
    gring_box(g, a.top_ring, rb);
    const long nring = s->gring[0] ? gbox_count(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], rb[6], rb[7]) : 0;
    long nc = (long)g.NX * g.NZ;
    if constexpr (sizeof(T) == 4) {
        if (s->gtiled) {
            const Geo &v = s->gview;
            a.hx0 = v.ix0; a.hx1 = v.ix1; a.hz0 = v.iz0; a.hz1 = v.iz1;
            nc = gbox_count(1 - g.npml, g.nx + g.npml, s->zt, s->zt + g.NZ - 1, a.hx0, a.hx1, a.hz0, a.hz1);
            // the step kernel reaches norder cells into the hole
            a.ex0 = v.ix0 + g.norder; a.ex1 = v.ix1 - g.norder;
            a.ez0 = v.iz0 + g.norder; a.ez1 = v.iz1 - g.norder;
        }
    }
    const long nst = gbox_count(1 - g.npml, g.nx + g.npml, s->zt, s->zt + g.NZ - 1, a.ex0, a.ex1, a.ez0, a.ez1);
    auto interior = [&](hipStream_t st) {
        if constexpr (sizeof(T) == 4) {
            if (!s->gtiled) return;
            with_order(g.norder, [&](auto no) {
                constexpr int NO = decltype(no)::value;
                const IntArgs<NO> ia = gen_interior_args<NO>(s);
                const unsigned nb = (unsigned)(ia.nbx * ia.nbz + ia.nsb);
                if (s->sh)
                    hipLaunchKernelGGL((k_interior_sh<NO, 256>), dim3(nb), dim3(256), 0, st, ia);
                else
                    hipLaunchKernelGGL((k_interior2<NO, 256>), dim3(nb), dim3(256), 0, st, ia);
            });
        }
    };
    // ring cells whose hybrid correction the step launch evaluates first
    const long nq = a.ring_valid ? nring : 0;
    const long nqb = (nq + 255) / 256 * 256;
    // The interior kernel and the per-cell kernels write disjoint cells of
    // level n+1 and read level n and the prep launch's output only, so with
    // a second stream they run side by side: the interior is bound by HBM,
    // the per-cell kernels by cache latency, and their workgroups (no LDS,
    // few registers) fit beside the interior's on the same CUs.  The interior
    // is queued first so that its workgroups take their slots first.
    // mode 0: one stream; 1 (SH's default): the interior first, on the
    // solver's stream, the per-cell kernels on the second; 2 (P-SV's
    // default): the per-cell kernels first, on the solver's stream, the
    // interior on the second
    const int mode = s->stream2 ? s->gen_overlap : 0;
    const bool fork = mode != 0;
    hipStream_t sc = mode == 1 ? s->stream2 : s->stream;
    hipStream_t si = mode == 2 ? s->stream2 : s->stream;
    // the prep launch: receivers, ring values, then the staged terms (one
    // launch: starting the interior after the ring values alone, beside a
    // second launch for the staging, was slower)
    auto prep = [&](long nr, long ng, long ns) {
        const long ntot = nr + ng + ns;
        if (ntot <= 0) return;
        if (s->sh)
            hipLaunchKernelGGL(k_gen_prep_sh<T>, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, s->stream, a, r,
                               nr, ng, ns, (T *)s->gring[0]);
        else
            hipLaunchKernelGGL(k_gen_prep<T>, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, s->stream, a, r,
                               nr, ng, ns, (T *)s->gring[0], (T *)s->gring[1]);
    };
    prep(nrec, nring, nst);
    if (fork) {
        HIPCHK(hipEventRecord(s->ev_fork, s->stream));
        HIPCHK(hipStreamWaitEvent(s->stream2, s->ev_fork, 0));
        if (mode == 1) interior(si);
    }

    long nm = 0;   // cells of the ring's neighbourhood (P-SV's staged masked terms)
    if (nq && !s->sh) {
        const int no = g.norder;
        nm = gbox_count(std::max(rb[0] - no, 1 - g.npml), std::min(rb[1] + no, g.nx + g.npml),
                        std::max(rb[2] - no, s->zt), std::min(rb[3] + no, s->zt + g.NZ - 1), rb[4] + no, rb[5] - no,
                        rb[6] + no, rb[7] - no);
    }
    if (s->sh) {
        if (!fork) interior(si);
        hipLaunchKernelGGL(k_gen_step_sh<T>, dim3((unsigned)((nqb + nc + 255) / 256)), dim3(256), 0, sc, a, nq);
    } else {
        if (nq) hipLaunchKernelGGL(k_gen_ringst<T>, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, sc, a);
        if (!fork) interior(si);
        hipLaunchKernelGGL(k_gen_step<T>, dim3((unsigned)((nqb + nc + 255) / 256)), dim3(256), 0, sc, a, nq);
    }
    // mode 2: the interior is queued after the per-cell kernels, whose
    // workgroups (no LDS, few registers) then take their places first
    if (mode == 2) interior(si);
    if (fork) {
        HIPCHK(hipEventRecord(s->ev_join, s->stream2));
        HIPCHK(hipStreamWaitEvent(s->stream, s->ev_join, 0));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FDFK_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    s->cur = 1 - s->cur;
    s->it += 1;
    return 0;
}
int gen_step(fdfk_solver *s, bool record)
{
    return s->gsize == 8 ? gen_step_t<double>(s, record) : gen_step_t<float>(s, record);
}

// damping profiles of the general path in its own type (the FP64 path keeps
// them unrounded, like the formulation's FP64 build)
template <class T>
int gen_damping(fdfk_solver *s)
{
    const Geo &g = s->g;
    const int nx = g.nx, nz = g.nz, npml = g.npml, nbd = g.nbd;
    const int leff = npml - nbd;
    const double lr = std::log(s->R0);
    const double d0x = leff > 0 ? -(double)(s->npower + 1) * (double)s->vpml * lr / (2.0 * leff * (double)g.dx) : 0.0;
    const double d0z = leff > 0 ? -(double)(s->npower + 1) * (double)s->vpml * lr / (2.0 * leff * (double)g.dz) : 0.0;
    auto prof = [&](int q, double d0) {
        if (q <= nbd || leff <= 0) return 0.0;
        double r = (double)(q - nbd) / (double)leff, p = 1.0;
        for (int e = 0; e < s->npower; ++e) p *= r;
        return d0 * p;
    };
    std::vector<T> dx(g.NX), bx(g.NX), dz(g.NZ), bz(g.NZ);
    for (int j = 0; j < g.NX; ++j) {
        const int ix = j + 1 - npml;
        const int q = ix <= 0 ? 1 - ix : (ix > nx ? ix - nx : 0);
        dx[j] = (T)prof(q, d0x);
        bx[j] = (T)std::exp(-(double)dx[j] * g.dt);
    }
    for (int k = 0; k < g.NZ; ++k) {
        const int iz = k + s->zt;
        dz[k] = iz > nz ? (T)prof(iz - nz, d0z) : (iz <= 0 ? (T)prof(1 - iz, d0z) : (T)0);
        bz[k] = (T)std::exp(-(double)dz[k] * g.dt);
    }
    const std::vector<T> *src[4] = {&dx, &dz, &bx, &bz};
    for (int i = 0; i < 4; ++i) {
        T *d = nullptr;
        if (int rc = s->dalloc(&d, src[i]->size())) return rc;
        HIPCHK(h2d(s->stream, d, src[i]->data(), src[i]->size() * sizeof(T)));
        s->gdamp[i] = d;
    }
    return 0;
}

// geometry of a job: the checks the background alone needs
int validate_geometry(int nx, int nz, int npml, int norder, float dx, float dz, double dt, int nt)
{
    if (norder < 1 || norder > kMaxOrder) return fail(FDFK_EINVAL, "norder must be 1..8 (FD order 2..16)");
    const int nbd = 2 * norder - 1;
    if (npml < nbd) return fail(FDFK_EINVAL, "npml must be >= 2*norder-1");
    if (nx < 2 * nbd + 2) return fail(FDFK_EINVAL, "nx too small for the hybrid ring");
    if (nz < nbd + 2 * norder + 1) return fail(FDFK_EINVAL, "nz too small");
    if (!(dx > 0) || !(dz > 0) || !(dt > 0)) return fail(FDFK_EINVAL, "dx, dz, dt must be > 0");
    if (nt < 0) return fail(FDFK_EINVAL, "nt >= 0 required");
    return 0;
}

int validate(const fdfk_job *j)
{
    if (!j) return fail(FDFK_EINVAL, "null job");
    const bool m16 = j->precision == FDFK_FP16 || j->precision == FDFK_BF16;
    if (j->is_pml_top && m16)
        return fail(FDFK_ENOSYS, "an absorbing top layer (is_pml_top) runs in FP32 or FP64 storage");
    if (j->wave != FDFK_PSV && j->wave != FDFK_SH) return fail(FDFK_EINVAL, "wave must be FDFK_PSV or FDFK_SH");
    if (j->wave == FDFK_SH && m16)
        return fail(FDFK_ENOSYS, "SH solvers run in FP32 or FP64 storage");
    if (int rc = validate_geometry(j->nx, j->nz, j->npml, j->norder, j->dx, j->dz, j->dt, j->nt)) return rc;
    if (j->nstp < 1) return fail(FDFK_EINVAL, "nstp >= 1 required");
    if (j->nrcv < 0 || (j->nrcv > 0 && (!j->rx || !j->rz))) return fail(FDFK_EINVAL, "receiver arrays");
    for (int i = 0; i < j->nrcv; ++i)
        if (!std::isfinite(j->rx[i]) || !std::isfinite(j->rz[i]))
            return fail(FDFK_EINVAL, "receiver coordinates must be finite");
    if (!j->c) return fail(FDFK_EINVAL, "material array c is required");
    if (j->precision != FDFK_FP32 && !m16 && j->precision != FDFK_FP64)
        return fail(FDFK_EINVAL, "storage precision: FDFK_FP32, FDFK_FP16, FDFK_BF16 or FDFK_FP64");
    return 0;
}

}  // namespace

extern "C" {

const char *fdfk_last_error(void) { return g_err.c_str(); }

static int alloc_background(fdfk_solver *s, int period);

int fdfk_autotune(const fdfk_job *job, int32_t steps, int32_t *best_lx, float *best_ms, int32_t *model_lx,
                  float *model_ms)
{
    if (!job || steps < 1) return fail(FDFK_EINVAL, "job required and steps >= 1");
    if (job->is_pml_top || job->precision == FDFK_FP64 || job->wave == FDFK_SH)
        return fail(FDFK_ENOSYS, "the tile autotuner covers the tiled FP32 kernels only");
    // scratch solver of the same geometry, set up for every timed step
    std::vector<int> cand;
    fdfk_job j = *job;
    j.nt = 2 * 9 * (steps + 3) + 64;
    j.nstp = 1;
    fdfk_solver *s = nullptr;
    if (int rc = fdfk_create(&j, &s)) return rc;
    std::unique_ptr<fdfk_solver> own(s);
    // a zero background of two steps: the hybrid-ring work runs as in a real solve
    if (int rc = alloc_background(s, 2)) return rc;
    for (int i = 0; i < 12; ++i) HIPCHK(hipMemsetAsync(s->bg[i], 0, (size_t)s->bg_n[i] * 4, s->stream));
    s->bg_valid = true;
    const int ncols = s->g.ix1 - s->g.ix0 + 1;
    const int m = s->lx_model;
    for (double f : {1.0, 0.5, 0.67, 0.8, 0.9, 1.1, 1.25, 1.5, 2.0}) {
        const int c = std::min(ncols, std::max(4, (int)std::lround(m * f)));
        if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    }
    // warm the GPU to its steady clocks at the model's width, then time every
    // candidate in two passes of opposite order and keep its faster pass
    s->lx = m;
    if (int rc = fdfk_run(s, 50)) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::vector<float> best(cand.size(), 0.0f);
    int rc = 0;
    for (int pass = 0; pass < 2 && rc == 0; ++pass)
        for (size_t q = 0; q < cand.size() && rc == 0; ++q) {
            const size_t k = pass == 0 ? q : cand.size() - 1 - q;
            s->lx = cand[k];
            if ((rc = fdfk_run(s, 3))) break;
            if (hipEventRecord(e0, s->stream) != hipSuccess) { rc = fail(FDFK_EDEVICE, "event"); break; }
            if ((rc = fdfk_run(s, steps))) break;
            float ms = 0.0f;
            if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(FDFK_EDEVICE, "event timing");
                break;
            }
            ms /= (float)steps;
            if (pass == 0 || ms < best[k]) best[k] = ms;
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    // leave the model's width (candidate 0) only for a gain of 1% or more
    size_t bk = 0;
    for (size_t k = 1; k < cand.size(); ++k)
        if (best[k] < best[bk]) bk = k;
    if (best[bk] > 0.99f * best[0]) bk = 0;
    const int blx = cand[bk];
    const float bms = best[bk], mms = best[0];
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        (*g_tuned)[tuned_key(s)] = blx;
    }
    if (best_lx) *best_lx = blx;
    if (best_ms) *best_ms = bms;
    if (model_lx) *model_lx = m;
    if (model_ms) *model_ms = mms;
    return FDFK_OK;
}

int fdfk_trim_memory(void)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::multimap<size_t, void *> blocks;
    {
        BlockCache &c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        blocks.swap(c.free[dev]);
    }
    for (auto &b : blocks) HIPCHK(hipFree(b.second));
    plan_pool_release_idle();
    return FDFK_OK;
}

// A solver of the general path (fdfk_general.hip) in element type T: the
// absorbing top layer (is_pml_top, rows from zt = 1-npml; c holds rows
// zt..nz+npml) and the FP64 storage mode (FDFK_FP64, either surface)
extern "C++" {
template <class T>
static int gen_create(const fdfk_job *job, fdfk_solver **out)
{
    auto *s = new fdfk_solver();
    auto bad = [&](int code) { delete s; return code; };
    int rc = 0;
    s->gen = true;
    s->sh = job->wave == FDFK_SH;
    s->gsize = (int)sizeof(T);
    s->precision = job->precision;
    s->zt = job->is_pml_top ? 1 - job->npml : 1;
    Geo &g = s->g;
    g.nx = job->nx; g.nz = job->nz; g.npml = job->npml; g.norder = job->norder;
    g.nbd = 2 * g.norder - 1;
    g.NX = g.nx + 2 * g.npml; g.NZ = g.nz + g.npml - s->zt + 1;
    g.ld = (g.NZ + 31) / 32 * 32;
    g.dx = job->dx; g.dz = job->dz; g.dt = job->dt;
    s->nt = job->nt; s->nstp = job->nstp;
    s->nsis = job->nt / job->nstp;
    s->nsamp = s->nsis + 1;
    s->nrcv = job->nrcv;
    s->x0 = job->x0; s->z0 = job->z0;
    s->R0 = job->R0; s->npower = job->npower;
    s->fused = false;
    if (const char *e = std::getenv("FDFK_PML_RATIO")) {
        s->pml_ratio = (float)std::atof(e);
        s->pml_ratio_d = std::atof(e);
    }
    const long ncell = (long)g.NX * g.NZ;
    float vpml = job->vpml;
    if (!(vpml > 0)) {
        vpml = 0.0f;
        for (long i = 0; i < ncell; ++i) {
            float v = std::sqrt((job->c[3 * i + 1] + 2.0f * job->c[3 * i + 2]) / job->c[3 * i]);
            if (v > vpml) vpml = v;
        }
    }
    s->vpml = vpml;
    coef_table(g.norder, true, s->hcoef.c0);
    coef_table(g.norder, false, s->hcoef.c1);
    if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess)
        return bad(fail(FDFK_EDEVICE, "hipStreamCreate failed"));
    const size_t nf = (size_t)g.NX * g.ld;
    auto zalloc = [&](void **dst, size_t n) -> int {
        T *p = nullptr;
        if (int r = s->dalloc(&p, n)) return r;
        if (hipMemsetAsync(p, 0, n * sizeof(T), s->stream) != hipSuccess) return fail(FDFK_EDEVICE, "memset");
        *dst = p;
        return 0;
    };
    for (int q = 0; q < 2; ++q)
        if ((rc = zalloc(&s->gux[q], nf)) || (rc = zalloc(&s->guz[q], nf))) return bad(rc);
    if ((rc = zalloc(&s->grho, nf)) || (rc = zalloc(&s->glam, nf)) || (rc = zalloc(&s->gmu, nf)) ||
        (rc = s->dalloc(&s->gc0, (size_t)kMaxOrder * kMaxOrder)) || (rc = s->dalloc(&s->gc1, (size_t)kMaxOrder * kMaxOrder)) ||
        (rc = s->dalloc(&s->stage, (size_t)2 * ncell * (sizeof(T) / 4))))
        return bad(rc);
    if (h2d(s->stream, s->gc0, s->hcoef.c0, sizeof(s->hcoef.c0)) != hipSuccess ||
        h2d(s->stream, s->gc1, s->hcoef.c1, sizeof(s->hcoef.c1)) != hipSuccess)
        return bad(fail(FDFK_EDEVICE, "coef upload"));
    {
        float *dc = nullptr;
        if (malloc_async_retry((void **)&dc, (size_t)3 * ncell * 4, s->stream) != hipSuccess)
            return bad(fail(FDFK_ENOMEM, "staging"));
        if (h2d(s->stream, dc, job->c, (size_t)3 * ncell * 4) != hipSuccess) {
            hipFreeAsync(dc, s->stream);
            return bad(fail(FDFK_EDEVICE, "material upload"));
        }
        hipLaunchKernelGGL(k_gen_import_material<T>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                           s->stream, g, s->zt, dc, (T *)s->grho, (T *)s->glam, (T *)s->gmu);
        hipFreeAsync(dc, s->stream);
    }
    {
        // face harmonic means (SH uses the mu faces only)
        const int m0 = s->sh ? 1 : 0;
        for (int m = m0; m < 2; ++m)
            if ((rc = s->dalloc((T **)&s->ghx[m], nf)) || (rc = s->dalloc((T **)&s->ghz[m], nf))) return bad(rc);
        hipLaunchKernelGGL(k_gen_faces<T>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s->stream, g, s->zt,
                           (const T *)s->glam, (const T *)s->gmu, m0, 1, (T *)s->ghx[0], (T *)s->ghz[0],
                           (T *)s->ghx[1], (T *)s->ghz[1]);
        // staged terms of level n (written by every prep launch before they are read)
        for (int i = 0; i < (s->sh ? 4 : 12); ++i)
            if ((rc = s->dalloc((T **)&s->gst[i], nf))) return bad(rc);
        // the weighted memories are written on strip cells only: 0 elsewhere
        for (int i = s->sh ? 2 : 8; i < (s->sh ? 4 : 12); ++i)
            if (hipMemsetAsync(s->gst[i], 0, nf * sizeof(T), s->stream) != hipSuccess)
                return bad(fail(FDFK_EDEVICE, "memset"));
    }
    if constexpr (sizeof(T) == 4) {
        // the tiled interior: columns 1..nx, rows 2*norder..nz (full-order
        // stencils whose reach stays in rows >= 1), as on the tiled path
        Geo v = g;
        v.NZ = g.nz + g.npml;
        v.ix0 = 1; v.ix1 = g.nx; v.iz0 = 2 * g.norder; v.iz1 = g.nz;
        const char *e = std::getenv("FDFK_GEN_TILED");
        if (!(e && std::atoi(e) == 0) && v.iz1 >= v.iz0) {
            s->gtiled = true;
            s->gview = v;
            // the per-cell kernels beside the interior (FDFK_GEN_OVERLAP=0: in line)
            const char *eo = std::getenv("FDFK_GEN_OVERLAP");
            // default: P-SV the per-cell kernels first, SH the interior
            // first (the faster order of each, measured on the 8001x2000 solve)
            s->gen_overlap = s->sh ? 1 : 2;
            if (eo && (std::atoi(eo) == 1 || std::atoi(eo) == 2)) s->gen_overlap = std::atoi(eo);
            if (!(eo && std::atoi(eo) == 0) &&
                (hipStreamCreate(&s->stream2) != hipSuccess ||
                 hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
                 hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess))
                return bad(fail(FDFK_EDEVICE, "second stream"));
            if ((rc = s->dalloc(&s->gbf, nf))) return bad(rc);
            hipLaunchKernelGGL(k_gen_bf, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s->stream, (long)nf,
                               (float)(g.dt * g.dt), (const float *)s->grho, s->gbf);
            with_order(g.norder, [&](auto no) { s->glx = gen_plan_lx<decltype(no)::value>(v, s->sh); });
            if (const char *el = std::getenv("FDFK_LX")) {
                const int lx = std::atoi(el);
                if (lx >= 4) s->glx = lx;
            }
        }
    }
    // strips L, R, B (and T with a top layer): split parts with two
    // interleaved levels, x-memories, z-memories in the bottom and top strips
    const size_t ncs[4] = {(size_t)g.npml * g.nz, (size_t)g.npml * g.nz, (size_t)g.NX * g.npml, (size_t)g.NX * g.npml};
    for (int st = 0; st < (s->zt != 1 ? 4 : 3); ++st)
        for (int c = 0; c < 2; ++c) {
            for (int q = 0; q < 3; ++q)
                if ((rc = zalloc(&s->gbu[st][c][q], 2 * ncs[st]))) return bad(rc);
            if ((rc = zalloc(&s->gbpx[st][c], ncs[st]))) return bad(rc);
            if (st >= 2 && (rc = zalloc(&s->gbpz[st][c], ncs[st]))) return bad(rc);
        }
    if ((rc = gen_damping<T>(s))) return bad(rc);
    if (s->nrcv > 0) {
        if ((rc = s->dalloc(&s->rx, s->nrcv)) || (rc = s->dalloc(&s->rz, s->nrcv)) ||
            (rc = zalloc(&s->gseis[0], (size_t)s->nrcv * s->nsamp)) ||
            (rc = zalloc(&s->gseis[1], (size_t)s->nrcv * s->nsamp)))
            return bad(rc);
        if (h2d(s->stream, s->rx, job->rx, s->nrcv * 4) != hipSuccess ||
            h2d(s->stream, s->rz, job->rz, s->nrcv * 4) != hipSuccess)
            return bad(fail(FDFK_EDEVICE, "receiver upload"));
        if (sizeof(T) == 4) { s->seisx = (float *)s->gseis[0]; s->seisz = (float *)s->gseis[1]; }
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return bad(fail(FDFK_EDEVICE, std::string("setup: ") + hipGetErrorString(e)));
    *out = s;
    return 0;
}
}  // extern "C++"

int fdfk_create(const fdfk_job *job, fdfk_solver **out)
{
    if (!out) return fail(FDFK_EINVAL, "null output pointer");
    *out = nullptr;
    int rc = validate(job);
    if (rc) return rc;
    if (job->precision == FDFK_FP64) return gen_create<double>(job, out);
    if (job->is_pml_top || job->wave == FDFK_SH) return gen_create<float>(job, out);
    auto *s = new fdfk_solver();
    Geo &g = s->g;
    g.nx = job->nx; g.nz = job->nz; g.npml = job->npml; g.norder = job->norder;
    g.nbd = 2 * g.norder - 1;
    g.NX = g.nx + 2 * g.npml; g.NZ = g.nz + g.npml;
    g.ld = (g.NZ + 31) / 32 * 32;
    g.dx = job->dx; g.dz = job->dz; g.dt = job->dt;
    // interior kernel: the whole total-field region below the surface rows
    // (full order there because npml >= 2*norder-1)
    g.ix0 = 1; g.ix1 = g.nx;
    g.iz0 = 2 * g.norder; g.iz1 = g.nz;
    s->nt = job->nt; s->nstp = job->nstp;
    s->nsis = job->nt / job->nstp;
    s->nsamp = s->nsis + 1;
    s->nrcv = job->nrcv;
    s->x0 = job->x0; s->z0 = job->z0;
    s->R0 = job->R0; s->npower = job->npower;
    s->mat16 = job->precision == FDFK_FP16 || job->precision == FDFK_BF16;
    s->matbf = job->precision == FDFK_BF16;
    if (const char *e = std::getenv("FDFK_PML_RATIO")) s->pml_ratio = (float)std::atof(e);
    if (const char *e = std::getenv("FDFK_FRAME_FULL")) s->frame_full = std::atoi(e) != 0;
    plan_interior(s);
    if (s->mat16 && !s->fused) {
        delete s;
        return fail(FDFK_ENOSYS, "16-bit materials run in the fused step only (FDFK_FUSED=0 is FP32)");
    }
    const long ncell = (long)g.NX * g.NZ;
    // vpml = max Vp (src/FDFK2D.f90:416)
    float vpml = job->vpml;
    if (!(vpml > 0)) {
        vpml = 0.0f;
        for (long i = 0; i < ncell; ++i) {
            float v = std::sqrt((job->c[3 * i + 1] + 2.0f * job->c[3 * i + 2]) / job->c[3 * i]);
            if (v > vpml) vpml = v;
        }
    }
    s->vpml = vpml;
    coef_table(g.norder, true, s->hcoef.c0);
    coef_table(g.norder, false, s->hcoef.c1);
    for (int i = 0; i < kMaxOrder * kMaxOrder; ++i) {
        s->hcoef.c0f[i] = (float)s->hcoef.c0[i];
        s->hcoef.c1f[i] = (float)s->hcoef.c1[i];
    }
    auto bad = [&](int code) { delete s; return code; };
    if (hipGetDevice(&s->device) != hipSuccess || hipStreamCreate(&s->stream) != hipSuccess)
        return bad(fail(FDFK_EDEVICE, "hipStreamCreate failed"));
    const size_t nf = (size_t)g.NX * g.ld;
    if ((rc = s->dalloc(&s->ux[0], nf)) || (rc = s->dalloc(&s->ux[1], nf)) ||
        (rc = s->dalloc(&s->uz[0], nf)) || (rc = s->dalloc(&s->uz[1], nf)) ||
        (rc = s->dalloc(&s->bf, nf)) || (rc = s->dalloc(&s->lam, nf)) || (rc = s->dalloc(&s->mu, nf)) ||
        (rc = s->dalloc(&s->dcoef, 1)) || (rc = s->dalloc(&s->stage, (size_t)2 * ncell)))
        return bad(rc);
    for (int q = 0; q < 2; ++q) {
        if (hipMemsetAsync(s->ux[q], 0, nf * 4, s->stream) || hipMemsetAsync(s->uz[q], 0, nf * 4, s->stream))
            return bad(fail(FDFK_EDEVICE, "memset"));
    }
    if (hipMemsetAsync(s->bf, 0, nf * 4, s->stream) || hipMemsetAsync(s->lam, 0, nf * 4, s->stream) ||
        hipMemsetAsync(s->mu, 0, nf * 4, s->stream))
        return bad(fail(FDFK_EDEVICE, "memset"));
    if (const char *e = std::getenv("FDFK_STEP2")) {
        if (std::atoi(e) != 0 && s->fused && !s->mat16 && s->nt_threads == 256)
            with_order(g.norder, [&](auto no) { s->step2 = plan_step2_t<decltype(no)::value>(s); });
        if (s->step2) {
            // spares start at zero like the live levels: cells no step
            // writes (the outermost padding line) stay zero in every level
            for (int q = 0; q < 2; ++q) {
                if ((rc = s->dalloc(&s->uxs[q], nf)) || (rc = s->dalloc(&s->uzs[q], nf))) return bad(rc);
                if (hipMemsetAsync(s->uxs[q], 0, nf * 4, s->stream) || hipMemsetAsync(s->uzs[q], 0, nf * 4, s->stream))
                    return bad(fail(FDFK_EDEVICE, "memset"));
            }
        }
    }
    if (h2d(s->stream, s->dcoef, &s->hcoef, sizeof(Coefs)) != hipSuccess)
        return bad(fail(FDFK_EDEVICE, "coef upload"));
    // materials: reference layout -> dt^2/rho, lambda, mu
    // staging buffer in the stream's order: freeing it waits for no other stream
    float *dc = nullptr;
    if (malloc_async_retry((void **)&dc, (size_t)3 * ncell * 4, s->stream) != hipSuccess)
        return bad(fail(FDFK_ENOMEM, "staging"));
    if (h2d(s->stream, dc, job->c, (size_t)3 * ncell * 4) != hipSuccess) {
        hipFreeAsync(dc, s->stream);
        return bad(fail(FDFK_EDEVICE, "material upload"));
    }
    hipLaunchKernelGGL(k_import_material, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s->stream, g, dc,
                       (float)(g.dt * g.dt), s->bf, s->lam, s->mu);
    hipFreeAsync(dc, s->stream);
    if (s->mat16) {
        // Residuals against per-row profiles: the profile of a row is the
        // midrange of its values over the undamped columns 1-nbd..nx+nbd
        // (rows below nz+nbd take row nz+nbd's, as their damped cells copy
        // it); one power-of-two scale per field keeps |residual| / scale
        // < 2^15 over the undamped cells (the damped ones are copies).
        const int xa = 1 - g.nbd, xb = g.nx + g.nbd, zb = g.nz + g.nbd;
        auto C = [&](int ix, int iz, int k) { return job->c[3 * ((size_t)(ix + g.npml - 1) * g.NZ + (iz - 1)) + k]; };
        std::vector<float> pl(g.NZ), pm(g.NZ);
        double ml = 0.0, mm = 0.0;
        for (int iz = 1; iz <= g.NZ; ++iz) {
            if (iz > zb) { pl[iz - 1] = pl[zb - 1]; pm[iz - 1] = pm[zb - 1]; continue; }
            float l0 = C(xa, iz, 1), l1 = l0, m0 = C(xa, iz, 2), m1 = m0;
            for (int ix = xa; ix <= xb; ++ix) {
                l0 = std::min(l0, C(ix, iz, 1)); l1 = std::max(l1, C(ix, iz, 1));
                m0 = std::min(m0, C(ix, iz, 2)); m1 = std::max(m1, C(ix, iz, 2));
            }
            pl[iz - 1] = 0.5f * (l0 + l1);
            pm[iz - 1] = 0.5f * (m0 + m1);
            for (int ix = xa; ix <= xb; ++ix) {
                ml = std::max(ml, (double)std::fabs(C(ix, iz, 1) - pl[iz - 1]));
                mm = std::max(mm, (double)std::fabs(C(ix, iz, 2) - pm[iz - 1]));
            }
        }
        auto pow2 = [](double m) {
            int e = 0;
            if (!(m > 0.0)) return 1.0f;
            std::frexp(m, &e);
            return (float)std::ldexp(1.0, e - 15);
        };
        s->sl = pow2(ml);
        s->sm = pow2(mm);
        if ((rc = s->dalloc(&s->lam16, nf)) || (rc = s->dalloc(&s->mu16, nf)) || (rc = s->dalloc(&s->lamp, g.NZ)) ||
            (rc = s->dalloc(&s->mup, g.NZ)))
            return bad(rc);
        if (h2d(s->stream, s->lamp, pl.data(), g.NZ * 4) != hipSuccess ||
            h2d(s->stream, s->mup, pm.data(), g.NZ * 4) != hipSuccess)
            return bad(fail(FDFK_EDEVICE, "material profile upload"));
        hipLaunchKernelGGL(k_round_material, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s->stream, g,
                           s->lam, s->mu, s->lam16, s->mu16, s->lamp, s->mup, s->sl, s->sm, s->matbf ? 1 : 0);
    }
    // PML strips
    s->strip_cells[S_L] = s->strip_cells[S_R] = (long)g.npml * g.nz;
    s->strip_cells[S_B] = (long)g.NX * g.npml;
    for (int st = 0; st < 3; ++st) {
        size_t n = (size_t)s->strip_cells[st];
        for (int c = 0; c < 2; ++c) {
            for (int p = 0; p < 3; ++p)
                for (int b = 0; b < 2; ++b) {
                    if ((rc = s->dalloc(&s->strip_fields[st][c][p][b], n))) return bad(rc);
                    hipMemsetAsync(s->strip_fields[st][c][p][b], 0, n * 4, s->stream);
                }
            if ((rc = s->dalloc(&s->psx[st][c], n))) return bad(rc);
            hipMemsetAsync(s->psx[st][c], 0, n * 4, s->stream);
            if (st == S_B) {
                if ((rc = s->dalloc(&s->psz[st][c], n))) return bad(rc);
                hipMemsetAsync(s->psz[st][c], 0, n * 4, s->stream);
            }
        }
    }
    if ((rc = build_damping(s))) return bad(rc);
    if ((rc = upload_strips(s))) return bad(rc);
    // receivers
    if (s->nrcv > 0) {
        if ((rc = s->dalloc(&s->rx, s->nrcv)) || (rc = s->dalloc(&s->rz, s->nrcv)) ||
            (rc = s->dalloc(&s->seisx, (size_t)s->nrcv * s->nsamp)) ||
            (rc = s->dalloc(&s->seisz, (size_t)s->nrcv * s->nsamp)))
            return bad(rc);
        if (h2d(s->stream, s->rx, job->rx, s->nrcv * 4) != hipSuccess ||
            h2d(s->stream, s->rz, job->rz, s->nrcv * 4) != hipSuccess)
            return bad(fail(FDFK_EDEVICE, "receiver upload"));
        hipMemsetAsync(s->seisx, 0, (size_t)s->nrcv * s->nsamp * 4, s->stream);
        hipMemsetAsync(s->seisz, 0, (size_t)s->nrcv * s->nsamp * 4, s->stream);
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return bad(fail(FDFK_EDEVICE, std::string("setup: ") + hipGetErrorString(e)));
    *out = s;
    return 0;
}

static void bg_sizes(const Geo &g, int nt, long n[12])
{
    long T = nt + 1;
    long reg_lr = (long)(g.nz - g.nbd) * g.nbd * T, reg_b = (long)g.nbd * g.nx * T;
    long pml_lr = (long)g.nz * g.nbd * T, pml_b = (long)g.nbd * (g.nx + 2 * g.nbd) * T;
    long v[12] = {reg_lr, reg_lr, reg_lr, reg_lr, reg_b, reg_b, pml_lr, pml_lr, pml_lr, pml_lr, pml_b, pml_b};
    for (int i = 0; i < 12; ++i) n[i] = v[i];
}

// device storage of the twelve background arrays, their pointer table and
// the two full-grid ring arrays
static int alloc_background(fdfk_solver *s, int period)
{
    if (s->d_bg)
        return s->bg_period == period ? 0
                                      : fail(FDFK_EINVAL, "background storage exists with another block size");
    long n[12];
    bg_sizes(s->g, period - 1, n);
    Background h{};
    for (int i = 0; i < 12; ++i) {
        int rc = s->dalloc(&s->bg[i], (size_t)n[i]);
        if (rc) return rc;
        s->bg_n[i] = n[i];
        h.a[i] = s->bg[i];
    }
    if (s->zt != 1) {
        // top ring: inside rows 1..nbd x cols nbd+1..nx-nbd, outside rows
        // 1-nbd..0 x cols 1-nbd..nx+nbd, step slowest
        const long nin = (long)s->g.nbd * (s->g.nx - 2 * s->g.nbd) * period;
        const long nout = (long)s->g.nbd * (s->g.nx + 2 * s->g.nbd) * period;
        const long nt4[4] = {nin, nin, nout, nout};
        for (int i = 0; i < 4; ++i) {
            int rc = s->dalloc(&s->bg_top[i], (size_t)nt4[i]);
            if (rc) return rc;
            s->bg_top_n[i] = nt4[i];
        }
    }
    if (s->gen) {
        // the general path's background in its own type: FP32 shares the
        // arrays above, FP64 has double copies
        for (int i = 0; i < 16; ++i) {
            float *f = i < 12 ? s->bg[i] : s->bg_top[i - 12];
            const long n = i < 12 ? s->bg_n[i] : s->bg_top_n[i - 12];
            if (!f) continue;
            if (s->gsize == 4) { s->gbg[i] = f; continue; }
            double *d = nullptr;
            if (int rc = s->dalloc(&d, (size_t)n)) return rc;
            s->gbg[i] = d;
        }
    }
    h.valid = 1;
    h.period = period;
    s->bg_period = period;
    const size_t nf = (size_t)s->g.NX * s->g.ld;
    int rc;
    if ((rc = s->dalloc(&s->d_bg, 1))) return rc;
    HIPCHK(h2d(s->stream, s->d_bg, &h, sizeof(Background)));
    if (!s->bg_only) {
        if ((rc = s->dalloc(&s->ringx, nf)) || (rc = s->dalloc(&s->ringz, nf))) return rc;
        HIPCHK(hipMemsetAsync(s->ringx, 0, nf * 4, s->stream));
        HIPCHK(hipMemsetAsync(s->ringz, 0, nf * 4, s->stream));
        s->gring[0] = s->ringx; s->gring[1] = s->ringz;
        if (s->gen && s->gsize == 8) {
            double *rx = nullptr, *rz = nullptr;
            if ((rc = s->dalloc(&rx, nf)) || (rc = s->dalloc(&rz, nf))) return rc;
            HIPCHK(hipMemsetAsync(rx, 0, nf * 8, s->stream));
            HIPCHK(hipMemsetAsync(rz, 0, nf * 8, s->stream));
            s->gring[0] = rx; s->gring[1] = rz;
        }
        if (s->gen && !s->sh && !s->gmq[0]) {
            // P-SV: staged terms of the masked ring fields over the ring's
            // neighbourhood (the largest ring box grown and its hole shrunk
            // by norder, within the grid)
            int b[8];
            gring_box(s->g, s->zt != 1, b);
            const Geo &g = s->g;
            const int no = g.norder;
            const size_t nm = (size_t)gbox_count(std::max(b[0] - no, 1 - g.npml), std::min(b[1] + no, g.nx + g.npml),
                                                 std::max(b[2] - no, s->zt), std::min(b[3] + no, s->zt + g.NZ - 1),
                                                 b[4] + no, b[5] - no, b[6] + no, b[7] - no);
            for (int k = 0; k < 2; ++k)
                for (int i = 0; i < 8; ++i)
                    if ((rc = s->dalloc((char **)&s->gmq[8 * k + i], nm * (size_t)s->gsize))) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_set_background(fdfk_solver *s, const float *const bg[12])
{
    if (!s || !bg) return fail(FDFK_EINVAL, "null argument");
    long n[12];
    bg_sizes(s->g, s->nt, n);
    int rc = alloc_background(s, s->nt + 1);
    if (rc) return rc;
    s->bg_block = 0;
    s->bg_fill = nullptr;
    for (int i = 0; i < 12; ++i) {
        if (!bg[i]) return fail(FDFK_EINVAL, "null background array");
        HIPCHK(hipMemcpyAsync(s->bg[i], bg[i], (size_t)n[i] * 4, hipMemcpyHostToDevice, s->stream));
        if (s->gen && s->gsize == 8)
            hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n[i] + 255) / 256)), dim3(256), 0,
                               s->stream, n[i], s->bg[i], (double *)s->gbg[i]);
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->bg_valid = true;
    return 0;
}

int fdfk_set_background64(fdfk_solver *s, const double *const bg[12])
{
    if (!s || !bg) return fail(FDFK_EINVAL, "null argument");
    if (!(s->gen && s->gsize == 8)) return fail(FDFK_EINVAL, "an FP64 background needs an FDFK_FP64 solver");
    int rc = alloc_background(s, s->nt + 1);
    if (rc) return rc;
    s->bg_block = 0;
    s->bg_fill = nullptr;
    for (int i = 0; i < 12; ++i) {
        if (!bg[i]) return fail(FDFK_EINVAL, "null background array");
        HIPCHK(hipMemcpyAsync(s->gbg[i], bg[i], (size_t)s->bg_n[i] * 8, hipMemcpyHostToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->bg_valid = true;
    return 0;
}

int fdfk_get_background(fdfk_solver *s, float *const bg[12])
{
    if (!s || !bg) return fail(FDFK_EINVAL, "null argument");
    if (!s->bg_valid) return fail(FDFK_EINVAL, "no background on the device");
    if (s->gen && s->gsize == 8) return fail(FDFK_EINVAL, "an FDFK_FP64 solver keeps its background in FP64 only");
    if (s->bg_period != s->nt + 1)
        return fail(FDFK_EINVAL, "the background is synthesised in blocks; its steps are not all stored");
    for (int i = 0; i < 12; ++i)
        HIPCHK(hipMemcpyAsync(bg[i], s->bg[i], (size_t)s->bg_n[i] * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_set_top_background(fdfk_solver *s, const float *const top[4])
{
    if (!s || !top) return fail(FDFK_EINVAL, "null argument");
    if (s->zt == 1) return fail(FDFK_EINVAL, "the top ring exists with an absorbing top layer (is_pml_top) only");
    int rc = alloc_background(s, s->nt + 1);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) {
        if (!top[i]) return fail(FDFK_EINVAL, "null background array");
        HIPCHK(hipMemcpyAsync(s->bg_top[i], top[i], (size_t)s->bg_top_n[i] * 4, hipMemcpyHostToDevice, s->stream));
        if (s->gen && s->gsize == 8)
            hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((s->bg_top_n[i] + 255) / 256)), dim3(256),
                               0, s->stream, s->bg_top_n[i], s->bg_top[i], (double *)s->gbg[12 + i]);
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->bg_top_valid = true;
    return 0;
}

int fdfk_set_top_background64(fdfk_solver *s, const double *const top[4])
{
    if (!s || !top) return fail(FDFK_EINVAL, "null argument");
    if (!(s->gen && s->gsize == 8)) return fail(FDFK_EINVAL, "an FP64 background needs an FDFK_FP64 solver");
    if (s->zt == 1) return fail(FDFK_EINVAL, "the top ring exists with an absorbing top layer (is_pml_top) only");
    int rc = alloc_background(s, s->nt + 1);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i) {
        if (!top[i]) return fail(FDFK_EINVAL, "null background array");
        HIPCHK(hipMemcpyAsync(s->gbg[12 + i], top[i], (size_t)s->bg_top_n[i] * 8, hipMemcpyHostToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->bg_top_valid = true;
    return 0;
}

int fdfk_get_top_background(fdfk_solver *s, float *const top[4])
{
    if (!s || !top) return fail(FDFK_EINVAL, "null argument");
    if (!s->bg_top_valid) return fail(FDFK_EINVAL, "no top ring background on the device");
    if (s->bg_period != s->nt + 1)
        return fail(FDFK_EINVAL, "the background is synthesised in blocks; its steps are not all stored");
    for (int i = 0; i < 4; ++i)
        HIPCHK(hipMemcpyAsync(top[i], s->bg_top[i], (size_t)s->bg_top_n[i] * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_run(fdfk_solver *s, int32_t nsteps)
{
    if (!s) return fail(FDFK_EINVAL, "null solver");
    if (nsteps < 0) return fail(FDFK_EINVAL, "nsteps < 0");
    if (s->gen) {
        for (int k = 0; k < nsteps; ++k) {
            if (s->it > s->nt) return fail(FDFK_ERANGE, "all iterations done");
            const bool rec = (s->isis == s->it) && (s->istp <= s->nsis);
            if (int rc = gen_step(s, rec)) return rc;
            if (rec) { s->istp += 1; s->isis += s->nstp; }
        }
        return 0;
    }
    for (int k = 0; k < nsteps; ++k) {
        if (s->it > s->nt) return fail(FDFK_ERANGE, "all iterations done");
        // a pair never spans two calls: between calls both live levels are whole
        const bool pair = s->step2 && k + 1 < nsteps && s->it + 1 <= s->nt;
        for (int h = 0; h < (pair ? 2 : 1); ++h) {
            bool rec = (s->isis == s->it) && (s->istp <= s->nsis);
            int rc = one_step(s, rec, pair ? h + 1 : 0);
            if (rc) return rc;
            if (rec) { s->istp += 1; s->isis += s->nstp; }
        }
        if (pair) ++k;
    }
    return 0;
}

int fdfk_iteration(const fdfk_solver *s) { return s ? s->it : FDFK_EINVAL; }

int fdfk_query(fdfk_solver *s)
{
    if (!s) return fail(FDFK_EINVAL, "null solver");
    const hipError_t e = hipStreamQuery(s->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(FDFK_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int fdfk_sync(fdfk_solver *s)
{
    if (!s) return fail(FDFK_EINVAL, "null solver");
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_seismograms64(fdfk_solver *s, double *sx, double *sz)
{
    if (!s || !sx || !sz) return fail(FDFK_EINVAL, "null argument");
    if (!(s->gen && s->gsize == 8)) return fail(FDFK_EINVAL, "FP64 seismograms need an FDFK_FP64 solver");
    if (s->nrcv == 0) return 0;
    size_t n = (size_t)s->nrcv * s->nsamp * 8;
    HIPCHK(hipMemcpyAsync(sx, s->gseis[0], n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(sz, s->gseis[1], n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_seismograms(fdfk_solver *s, float *sx, float *sz)
{
    if (!s || !sx || !sz) return fail(FDFK_EINVAL, "null argument");
    if (s->nrcv == 0) return 0;
    if (s->gen && s->gsize == 8) {
        std::vector<double> a((size_t)s->nrcv * s->nsamp), b(a.size());
        if (int rc = fdfk_seismograms64(s, a.data(), b.data())) return rc;
        for (size_t i = 0; i < a.size(); ++i) { sx[i] = (float)a[i]; sz[i] = (float)b[i]; }
        return 0;
    }
    size_t n = (size_t)s->nrcv * s->nsamp * 4;
    HIPCHK(hipMemcpyAsync(sx, s->seisx, n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(sz, s->seisz, n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_get_state(fdfk_solver *s, float *ux, float *uz)
{
    if (!s || !ux || !uz) return fail(FDFK_EINVAL, "null argument");
    const Geo &g = s->g;
    long n = (long)g.NX * g.NZ;
    dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    if (s->gen) {
        for (int c = 0; c < 2; ++c) {
            void *const *f = c == 0 ? s->gux : s->guz;
            if (s->gsize == 8)
                hipLaunchKernelGGL((k_gen_levels<double, float>), gr, bl, 0, s->stream, g, 1, s->stage,
                                   (double *)f[s->cur], (double *)f[1 - s->cur]);
            else
                hipLaunchKernelGGL((k_gen_levels<float, float>), gr, bl, 0, s->stream, g, 1, s->stage,
                                   (float *)f[s->cur], (float *)f[1 - s->cur]);
            HIPCHK(hipMemcpyAsync(c == 0 ? ux : uz, s->stage, (size_t)2 * n * 4, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        return 0;
    }
    hipLaunchKernelGGL(k_export_levels, gr, bl, 0, s->stream, g, s->ux[s->cur], s->ux[1 - s->cur], s->stage);
    HIPCHK(hipMemcpyAsync(ux, s->stage, (size_t)2 * n * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    hipLaunchKernelGGL(k_export_levels, gr, bl, 0, s->stream, g, s->uz[s->cur], s->uz[1 - s->cur], s->stage);
    HIPCHK(hipMemcpyAsync(uz, s->stage, (size_t)2 * n * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_get_state64(fdfk_solver *s, double *ux, double *uz)
{
    if (!s || !ux || !uz) return fail(FDFK_EINVAL, "null argument");
    if (!(s->gen && s->gsize == 8)) return fail(FDFK_EINVAL, "FP64 fields need an FDFK_FP64 solver");
    const Geo &g = s->g;
    long n = (long)g.NX * g.NZ;
    dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    for (int c = 0; c < 2; ++c) {
        void *const *f = c == 0 ? s->gux : s->guz;
        hipLaunchKernelGGL((k_gen_levels<double, double>), gr, bl, 0, s->stream, g, 1, (double *)s->stage,
                           (double *)f[s->cur], (double *)f[1 - s->cur]);
        HIPCHK(hipMemcpyAsync(c == 0 ? ux : uz, s->stage, (size_t)2 * n * 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return 0;
}

int fdfk_set_state(fdfk_solver *s, const float *ux, const float *uz)
{
    if (!s || !ux || !uz) return fail(FDFK_EINVAL, "null argument");
    const Geo &g = s->g;
    long n = (long)g.NX * g.NZ;
    dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    if (s->gen) {
        for (int c = 0; c < 2; ++c) {
            void *const *f = c == 0 ? s->gux : s->guz;
            HIPCHK(hipMemcpyAsync(s->stage, c == 0 ? ux : uz, (size_t)2 * n * 4, hipMemcpyHostToDevice, s->stream));
            if (s->gsize == 8)
                hipLaunchKernelGGL((k_gen_levels<double, float>), gr, bl, 0, s->stream, g, 0, s->stage,
                                   (double *)f[s->cur], (double *)f[1 - s->cur]);
            else
                hipLaunchKernelGGL((k_gen_levels<float, float>), gr, bl, 0, s->stream, g, 0, s->stage,
                                   (float *)f[s->cur], (float *)f[1 - s->cur]);
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        return 0;
    }
    HIPCHK(hipMemcpyAsync(s->stage, ux, (size_t)2 * n * 4, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_import_levels, gr, bl, 0, s->stream, g, s->stage, s->ux[s->cur], s->ux[1 - s->cur]);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpyAsync(s->stage, uz, (size_t)2 * n * 4, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_import_levels, gr, bl, 0, s->stream, g, s->stage, s->uz[s->cur], s->uz[1 - s->cur]);
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int fdfk_profile(fdfk_solver *s, int32_t enable)
{
    if (!s) return fail(FDFK_EINVAL, "null solver");
    s->prof = enable != 0;
    return 0;
}

int fdfk_kernel_stats(fdfk_solver *s, double *int_ms, int64_t *int_n, double *step_ms, int64_t *step_n)
{
    if (!s) return fail(FDFK_EINVAL, "null solver");
    HIPCHK(hipStreamSynchronize(s->stream));
    for (auto &e : s->ev_int) {
        float ms = 0;
        hipEventElapsedTime(&ms, e.first, e.second);
        s->int_ms += ms; s->int_n += 1;
    }
    for (auto &e : s->ev_step) {
        float ms = 0;
        hipEventElapsedTime(&ms, e.first, e.second);
        s->step_ms += ms; s->step_n += 1;
        hipEventDestroy(e.first); hipEventDestroy(e.second);
    }
    for (auto &e : s->ev_int) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    s->ev_int.clear(); s->ev_step.clear();
    if (int_ms) *int_ms = s->int_ms;
    if (int_n) *int_n = s->int_n;
    if (step_ms) *step_ms = s->step_ms;
    if (step_n) *step_n = s->step_n;
    s->int_ms = s->step_ms = 0; s->int_n = s->step_n = 0;
    return 0;
}

int64_t fdfk_device_bytes(const fdfk_solver *s) { return s ? s->bytes : 0; }

void fdfk_destroy(fdfk_solver *s) { delete s; }

// ---------------------------------------------------------------------
// compatibility layer: the reference's external procedures
// ---------------------------------------------------------------------
struct CompatEntry {
    fdfk_solver *s = nullptr;
    int last_it = -2;
    int nx, nz, npml, norder, nt;
    const float *c, *bg0;
};
static std::map<const float *, CompatEntry> g_compat;
static std::mutex g_compat_mu;
// The top ring of an absorbing top layer has no slot in the reference's
// arrays: comput_background_wavefield_ keeps the four arrays it computes
// here, keyed by the caller's uxl_reg array, and updatedispl_ takes them
// from here for the same uxl_reg (without them the hybrid contour stays
// open at the top).
static std::map<const float *, std::array<std::vector<float>, 4>> g_top_ring;

// strips of a general-path solver: the device arrays are the reference's,
// so the exchange with the caller's arrays is a plain copy
static int gen_strips_copy(fdfk_solver *s, int st, float *const parts[2][3], float *const px[2],
                           float *const pz[2], bool to_dev)
{
    const Geo &g = s->g;
    const size_t n = st < 2 ? (size_t)g.npml * g.nz : (size_t)g.NX * g.npml;
    const hipMemcpyKind k = to_dev ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    for (int c = 0; c < 2; ++c) {
        for (int q = 0; q < 3; ++q)
            HIPCHK(hipMemcpyAsync(to_dev ? (void *)s->gbu[st][c][q] : (void *)parts[c][q],
                                  to_dev ? (const void *)parts[c][q] : (const void *)s->gbu[st][c][q], 2 * n * 4, k,
                                  s->stream));
        HIPCHK(hipMemcpyAsync(to_dev ? (void *)s->gbpx[st][c] : (void *)px[c],
                              to_dev ? (const void *)px[c] : (const void *)s->gbpx[st][c], n * 4, k, s->stream));
        if (st >= 2)
            HIPCHK(hipMemcpyAsync(to_dev ? (void *)s->gbpz[st][c] : (void *)pz[c],
                                  to_dev ? (const void *)pz[c] : (const void *)s->gbpz[st][c], n * 4, k, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

static int strip_import(fdfk_solver *s, int st, float *const parts[2][3], float *const psx[2], float *const psz[2])
{
    long n = s->strip_cells[st];
    float *tmp = s->stage;  // 2*NX*NZ floats >= 2*n
    dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    for (int c = 0; c < 2; ++c) {
        for (int p = 0; p < 3; ++p) {
            HIPCHK(hipMemcpyAsync(tmp, parts[c][p], (size_t)2 * n * 4, hipMemcpyHostToDevice, s->stream));
            hipLaunchKernelGGL(k_split_levels, gr, bl, 0, s->stream, n, tmp, s->strip_fields[st][c][p][s->cur],
                               s->strip_fields[st][c][p][1 - s->cur]);
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        HIPCHK(hipMemcpy(s->psx[st][c], psx[c], (size_t)n * 4, hipMemcpyHostToDevice));
        if (st == S_B) HIPCHK(hipMemcpy(s->psz[st][c], psz[c], (size_t)n * 4, hipMemcpyHostToDevice));
    }
    return 0;
}
static int strip_export(fdfk_solver *s, int st, float *const parts[2][3], float *const psx[2], float *const psz[2])
{
    long n = s->strip_cells[st];
    float *tmp = s->stage;
    dim3 gr((unsigned)((n + 255) / 256)), bl(256);
    for (int c = 0; c < 2; ++c) {
        for (int p = 0; p < 3; ++p) {
            hipLaunchKernelGGL(k_join_levels, gr, bl, 0, s->stream, n, s->strip_fields[st][c][p][s->cur],
                               s->strip_fields[st][c][p][1 - s->cur], tmp);
            HIPCHK(hipMemcpyAsync(parts[c][p], tmp, (size_t)2 * n * 4, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        HIPCHK(hipMemcpy(psx[c], s->psx[st][c], (size_t)n * 4, hipMemcpyDeviceToHost));
        if (st == S_B) HIPCHK(hipMemcpy(psz[c], s->psz[st][c], (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    return 0;
}

void updatedispl_(const int32_t *norder, const int32_t *nx, const int32_t *nz, const int32_t *npml,
    const int32_t *izt, const int32_t *is_pml_top, const int32_t *npower, const int32_t *nbd,
    const int32_t *it, const int32_t *nt, const float *dx, const float *dz, const double *dt,
    const float *vpml, const double *R0,
    const float *uxl_reg, const float *uzl_reg, const float *uxr_reg, const float *uzr_reg,
    const float *uxb_reg, const float *uzb_reg, const float *uxl_pml, const float *uzl_pml,
    const float *uxr_pml, const float *uzr_pml, const float *uxb_pml, const float *uzb_pml,
    const float *c, const double *coef0, const double *coef1, float *ux, float *uz,
    float *bpxlx, float *bpzlx, float *bpxrx, float *bpzrx, float *bpxtx, float *bpxtz,
    float *bpztx, float *bpztz, float *bpxbx, float *bpxbz, float *bpzbx, float *bpzbz,
    float *bux1l, float *bux2l, float *bux3l, float *buz1l, float *buz2l, float *buz3l,
    float *bux1r, float *bux2r, float *bux3r, float *buz1r, float *buz2r, float *buz3r,
    float *bux1b, float *bux2b, float *bux3b, float *buz1b, float *buz2b, float *buz3b,
    float *bux1t, float *bux2t, float *bux3t, float *buz1t, float *buz2t, float *buz3t)
{
    (void)coef0; (void)coef1;
    std::lock_guard<std::mutex> lk(g_compat_mu);
    // the reference's error contract is print and stop: a failed step must
    // not return to a driver that would go on writing stale seismograms
    auto die = [](const std::string &m) {
        std::fprintf(stderr, "updatedispl (fdfk2d_amd): %s\n", m.c_str());
        std::fflush(stderr);
        std::exit(1);
    };
    // host checks first: a bad configuration stops before any GPU work
    const bool top = *is_pml_top != 0;
    if (*izt != (top ? 1 - *npml : 1)) die("izt must be 1 (free surface) or 1-npml (is_PML_top)");
    if (*nbd != 2 * *norder - 1) die("nbd must equal 2*norder-1");
    if (validate_geometry(*nx, *nz, *npml, *norder, *dx, *dz, *dt, *nt)) die(fdfk_last_error());
    if (*it < 0 || *it > *nt) die("it outside 0..nt");
    if (!c || !ux || !uz) die("null array");
    auto itr = g_compat.find(ux);
    // a new run (it = 0), a restart (it does not follow the last call) or
    // other arrays: the host arrays are authoritative, so the materials, the
    // background and the state are uploaded again into a new mirror
    const bool fresh = itr == g_compat.end() || itr->second.nx != *nx || itr->second.nz != *nz ||
                       itr->second.npml != *npml || itr->second.norder != *norder || itr->second.nt != *nt ||
                       itr->second.c != c || itr->second.bg0 != uxl_reg || *it == 0 ||
                       itr->second.last_it != *it - 1;
    float *const pl[2][3] = {{bux1l, bux2l, bux3l}, {buz1l, buz2l, buz3l}};
    float *const pr[2][3] = {{bux1r, bux2r, bux3r}, {buz1r, buz2r, buz3r}};
    float *const pb[2][3] = {{bux1b, bux2b, bux3b}, {buz1b, buz2b, buz3b}};
    float *const pt[2][3] = {{bux1t, bux2t, bux3t}, {buz1t, buz2t, buz3t}};
    float *const xl[2] = {bpxlx, bpzlx}, *const xr[2] = {bpxrx, bpzrx}, *const xb[2] = {bpxbx, bpzbx};
    float *const zb[2] = {bpxbz, bpzbz}, *const none[2] = {nullptr, nullptr};
    float *const xt[2] = {bpxtx, bpztx}, *const ztop[2] = {bpxtz, bpztz};
    // the caller's strips in and out: four strips (general path) with a top layer
    auto strips = [&](fdfk_solver *s, bool in) {
        if (s->gen)
            return gen_strips_copy(s, 0, pl, xl, none, in) || gen_strips_copy(s, 1, pr, xr, none, in) ||
                   gen_strips_copy(s, 2, pb, xb, zb, in) || gen_strips_copy(s, 3, pt, xt, ztop, in);
        if (in)
            return strip_import(s, S_L, pl, xl, none) || strip_import(s, S_R, pr, xr, none) ||
                   strip_import(s, S_B, pb, xb, zb);
        return strip_export(s, S_L, pl, xl, none) || strip_export(s, S_R, pr, xr, none) ||
               strip_export(s, S_B, pb, xb, zb);
    };
    if (top) {
        for (float *a : {bux1t, bux2t, bux3t, buz1t, buz2t, buz3t, bpxtx, bpxtz, bpztx, bpztz})
            if (!a) die("null top-strip array");
    }
    if (fresh) {
        if (itr != g_compat.end()) { fdfk_destroy(itr->second.s); g_compat.erase(itr); }
        fdfk_job j{};
        j.nx = *nx; j.nz = *nz; j.npml = *npml; j.norder = *norder; j.npower = *npower;
        j.is_pml_top = top ? 1 : 0;
        j.dx = *dx; j.dz = *dz; j.dt = *dt; j.R0 = *R0; j.vpml = *vpml; j.nt = *nt; j.nstp = 1;
        j.c = c; j.precision = FDFK_FP32;
        fdfk_solver *s = nullptr;
        if (fdfk_create(&j, &s)) die(fdfk_last_error());
        const float *bg[12] = {uxl_reg, uzl_reg, uxr_reg, uzr_reg, uxb_reg, uzb_reg,
                               uxl_pml, uzl_pml, uxr_pml, uzr_pml, uxb_pml, uzb_pml};
        if (fdfk_set_background(s, bg) || fdfk_set_state(s, ux, uz) || strips(s, true)) die(fdfk_last_error());
        if (top) {
            auto tr = g_top_ring.find(uxl_reg);
            if (tr != g_top_ring.end()) {
                const float *t4[4] = {tr->second[0].data(), tr->second[1].data(), tr->second[2].data(),
                                      tr->second[3].data()};
                if (fdfk_set_top_background(s, t4)) die(fdfk_last_error());
            }
        }
        CompatEntry e;
        e.s = s; e.nx = *nx; e.nz = *nz; e.npml = *npml; e.norder = *norder; e.nt = *nt; e.c = c; e.bg0 = uxl_reg;
        itr = g_compat.emplace(ux, e).first;
    }
    CompatEntry &e = itr->second;
    fdfk_solver *s = e.s;
    s->it = *it;
    if (s->gen ? gen_step(s, false) : one_step(s, false)) die(fdfk_last_error());
    if (fdfk_get_state(s, ux, uz) || strips(s, false)) die(fdfk_last_error());
    e.last_it = *it;
}

void fdfk_compat_release(void)
{
    std::lock_guard<std::mutex> lk(g_compat_mu);
    for (auto &kv : g_compat) fdfk_destroy(kv.second.s);
    g_compat.clear();
    g_top_ring.clear();
}

}  // extern "C"

#include "fdfk_fk.hip"
#include "fdfk_batch.hip"
#include "fdfk_smooth.hip"
