// Batched execution: many resident simulations advanced together, two
// launches per step for all of them (kb_prep, kb_step: a tile work-list over
// jobs).  Each job keeps its own HBM state (a fdfk_solver handle); the batch
// only builds, every step, a table of per-job kernel arguments plus prefix
// sums of each job's block counts, uploads it from pinned memory and
// launches the same kernel bodies as the single-job path.  A block finds
// its job by binary search over the prefix sums.
//
// Jobs may have different grids, step counts, receivers and backgrounds;
// they must share the FD order.  Small grids that cannot fill 256 CUs alone
// are packed into one launch, and the per-step launch count is independent
// of the number of jobs.

namespace fdfk {

constexpr int kMaxBatch = 256;
enum { KB_RING = 0, KB_PSI, KB_INT, KB_FRAME, KB_REC, KB_N };

template <int NO>
struct BatchJob {
    StepArgs sa;
    IntArgs<NO> ia;
    ETiles ft;
    RecArgs ra;
    const Background *bg;
    float *ringx, *ringz;
};

struct BatchPrefix {
    int n;
    int pre[KB_N][kMaxBatch + 1];
};
// table layout (host slot and device slot alike): the prefix block, then the
// jobs from this offset, so that the prefix and the n jobs of a step are one
// contiguous copy
constexpr size_t kJobOff = (sizeof(BatchPrefix) + 255) / 256 * 256;

__device__ __forceinline__ int find_job(const int *pre, int n, int bid)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= bid) lo = mid; else hi = mid;
    }
    return lo;
}

// one launch for the per-step preparation of every member: the hybrid ring
// background, the PML memories and (on recording iterations) the receivers,
// in that block order; 256-lane blocks, kRecPerBlock receivers per block
template <int NO>
__global__ __launch_bounds__(256) void kb_prep(const BatchJob<NO> *J, const BatchPrefix *P)
{
    const int n = P->n;
    int b = blockIdx.x;
    const int nr = P->pre[KB_RING][n], np = P->pre[KB_PSI][n];
    if (b < nr) {
        const int j = find_job(P->pre[KB_RING], n, b);
        const BatchJob<NO> &q = J[j];
        ring_prep_body(q.sa.g, q.bg, q.sa.it, q.ringx, q.ringz, b - P->pre[KB_RING][j]);
        return;
    }
    b -= nr;
    if (b < np) {
        const int j = find_job(P->pre[KB_PSI], n, b);
        psi_body<NO>(J[j].sa, b - P->pre[KB_PSI][j]);
        return;
    }
    b -= np;
    const int j = find_job(P->pre[KB_REC], n, b);
    record_body(J[j].ra, b - P->pre[KB_REC][j]);
}

// the interior tiles of every member, then their frame tiles, in one launch
// (the single-job k_step arrangement)
template <int NO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kIntWaves<NO>))) void kb_step(const BatchJob<NO> *J,
                                                                                       const BatchPrefix *P)
{
    extern __shared__ float sm[];
    const int n = P->n;
    const int ni = P->pre[KB_INT][n];
    if ((int)blockIdx.x < ni) {
        const int t = xcd_tile(blockIdx.x, ni);
        const int j = find_job(P->pre[KB_INT], n, t);
        IntKernel2<NO, 256>::run(J[j].ia, t - P->pre[KB_INT][j],
                                 *reinterpret_cast<typename IntKernel2<NO, 256>::Smem *>(sm));
    } else {
        const int b = xcd_tile((int)blockIdx.x - ni, (int)gridDim.x - ni);
        const int j = find_job(P->pre[KB_FRAME], n, b);
        frame_body<NO>(J[j].sa, J[j].ft, b - P->pre[KB_FRAME][j]);
    }
}

}  // namespace fdfk

struct fdfk_batch {
    int norder = 3;
    hipStream_t stream = nullptr;
    std::vector<fdfk_solver *> members;
    static constexpr int R = 4;
    void *dtab[R] = {}, *htab[R] = {};
    hipEvent_t ev[R] = {};
    int slot = 0;
    size_t tabbytes = 0;
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_int, ev_step;
    // Job templates of the current active set: everything a step's job holds
    // except the fields that change with the iteration, rebuilt only when
    // the active set changes (tile width, tilings, frame tiles, coefficients)
    std::vector<fdfk_solver *> tact;
    std::vector<char> tjobs;        // BatchJob<norder> per active member
    std::vector<int> tnb;           // KB_N block counts per active member
    int tlx = 0;
    size_t tsh = 0;
    ~fdfk_batch()
    {
        if (stream) hipStreamSynchronize(stream);
        for (int i = 0; i < R; ++i) {
            if (dtab[i]) hipFree(dtab[i]);
            if (htab[i]) hipHostFree(htab[i]);
            if (ev[i]) hipEventDestroy(ev[i]);
        }
        for (auto &e : ev_int) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto &e : ev_step) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

template <int NO>
size_t batch_table_bytes()
{
    return kJobOff + sizeof(BatchJob<NO>) * kMaxBatch;
}

// fill one job's arguments for the step it is about to take; returns the
// block counts of the kernels
// Interior column width shared by the members of one batched step: the
// launch holds every member's blocks, so the one-round rule of the
// single-job plan (narrow tiles for small grids) would leave the launch
// dominated by warm-up columns.  Minimise rounds x (lx + warm-up) over the
// whole batch instead.
template <int NO>
int batch_lx(const std::vector<fdfk_solver *> &act)
{
    constexpr int NT = 256;
    constexpr int WARM = IntKernel2<NO, NT>::WARM;
    static long slots = 0;
    if (slots == 0) {
        // occupancy at the largest dynamic LDS a batched step can ask for
        const size_t sh = std::max(sizeof(typename IntKernel2<NO, NT>::Smem), (size_t)kFrameLds * sizeof(float));
        int dev = 0, ncu = 0, bpc = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kb_step<NO>, NT, sh) == hipSuccess &&
            ncu > 0 && bpc > 0)
            slots = (long)ncu * bpc;
        else
            slots = 768;
    }
    int best = 64;
    long best_cost = -1;
    for (int lx = 8; lx <= 256; lx += 8) {
        long nb = 0;
        for (auto *s : act) nb += int_blocks<NO, NT>(s->g, lx);
        const long cost = ((nb + slots - 1) / slots) * (lx + WARM);
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = lx; }
    }
    return best;
}

template <int NO>
void fill_job(fdfk_solver *s, BatchJob<NO> &b, int nb[KB_N], bool &record, int lx)
{
    const Geo &g = s->g;
    b.sa = step_args(s);
    b.bg = s->d_bg;
    b.ringx = s->ringx;
    b.ringz = s->ringz;
    b.ia = interior_args<NO, 256>(s);
    b.ia.lx = lx;
    int_tiling<NO, 256>(g, lx, b.ia.nbx, b.ia.nbz, b.ia.nsb, b.ia.zseg);
    b.ft = frame_tiles(g);
    record = (s->isis == s->it) && (s->istp <= s->nsis) && s->nrcv > 0;
    RecArgs &r = b.ra;
    r = RecArgs{};
    r.g = g; r.x0 = s->x0; r.z0 = s->z0; r.nrcv = s->nrcv; r.nsamp = s->nsamp; r.istp = s->istp;
    r.rx = s->rx; r.rz = s->rz; r.ux = s->ux[s->cur]; r.uz = s->uz[s->cur];
    r.seisx = s->seisx; r.seisz = s->seisz;
    long nring = ring_count(g);
    long npsi = 2L * g.npml * g.nz + (long)g.NX * g.npml;
    nb[KB_RING] = b.sa.ring_valid ? (int)((nring + 255) / 256) : 0;
    nb[KB_PSI] = (int)((npsi + 255) / 256);
    nb[KB_INT] = b.ia.nbx * b.ia.nbz + b.ia.nsb;
    nb[KB_FRAME] = b.ft.off[b.ft.n];
    nb[KB_REC] = record ? (s->nrcv + kRecPerBlock - 1) / kRecPerBlock : 0;
}

// the fields of a job template that change from step to step: the level
// pointers (they swap every step), the iteration, the recording index and
// whether this iteration records; nb[KB_REC] follows the recording flag
template <int NO>
bool refresh_job(fdfk_solver *s, BatchJob<NO> &b, int nb[KB_N])
{
    b.sa = step_args(s);
    b.ia.ux = s->ux[s->cur]; b.ia.uz = s->uz[s->cur];
    b.ia.uxp = s->ux[1 - s->cur]; b.ia.uzp = s->uz[1 - s->cur];
    b.ia.corr = s->bg_valid && s->it <= s->nt;
    const bool record = (s->isis == s->it) && (s->istp <= s->nsis) && s->nrcv > 0;
    b.ra.istp = s->istp;
    b.ra.ux = s->ux[s->cur]; b.ra.uz = s->uz[s->cur];
    nb[KB_REC] = record ? (s->nrcv + kRecPerBlock - 1) / kRecPerBlock : 0;
    return record;
}

template <int NO>
int batch_step(fdfk_batch *B)
{
    std::vector<fdfk_solver *> act;
    for (auto *s : B->members)
        if (s->it <= s->nt) act.push_back(s);
    if (act.empty()) return 0;
    constexpr size_t JB = sizeof(BatchJob<NO>);
    if (act != B->tact) {
        B->tlx = batch_lx<NO>(act);
        B->tjobs.assign(JB * act.size(), 0);
        B->tnb.assign((size_t)KB_N * act.size(), 0);
        size_t sh = sizeof(typename IntKernel2<NO, 256>::Smem);
        for (size_t j = 0; j < act.size(); ++j) {
            auto &t = ((BatchJob<NO> *)B->tjobs.data())[j];
            bool r = false;
            fill_job<NO>(act[j], t, &B->tnb[(size_t)KB_N * j], r, B->tlx);
            sh = std::max(sh, frame_lds_bytes(t.ft, NO));
        }
        B->tsh = sh;
        B->tact = act;
    }
    const int slot = B->slot;
    HIPCHK(hipEventSynchronize(B->ev[slot]));   // the table slot is free again
    auto *P = (BatchPrefix *)B->htab[slot];
    auto *J = (BatchJob<NO> *)((char *)B->htab[slot] + kJobOff);
    P->n = (int)act.size();
    for (int k = 0; k < KB_N; ++k) P->pre[k][0] = 0;
    std::vector<char> rec(act.size());
    for (size_t j = 0; j < act.size(); ++j) {
        std::memcpy((void *)&J[j], B->tjobs.data() + JB * j, JB);
        int nb[KB_N];
        std::memcpy(nb, &B->tnb[(size_t)KB_N * j], sizeof(nb));
        rec[j] = refresh_job<NO>(act[j], J[j], nb);
        for (int k = 0; k < KB_N; ++k) P->pre[k][j + 1] = P->pre[k][j] + nb[k];
    }
    // the prefix block and the used jobs: one copy
    auto *dP = (BatchPrefix *)B->dtab[slot];
    auto *dJ = (BatchJob<NO> *)((char *)B->dtab[slot] + kJobOff);
    HIPCHK(hipMemcpyAsync(B->dtab[slot], B->htab[slot], kJobOff + JB * act.size(), hipMemcpyHostToDevice,
                          B->stream));
    const int n = P->n;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    if (B->prof) {
        hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2); hipEventCreate(&e3);
        hipEventRecord(e0, B->stream);
    }
    const int nprep = P->pre[KB_RING][n] + P->pre[KB_PSI][n] + P->pre[KB_REC][n];
    if (nprep) hipLaunchKernelGGL(kb_prep<NO>, dim3(nprep), dim3(256), 0, B->stream, dJ, dP);
    if (B->prof) hipEventRecord(e1, B->stream);
    const size_t sh = B->tsh;
    const int nstep = P->pre[KB_INT][n] + P->pre[KB_FRAME][n];
    if (nstep) hipLaunchKernelGGL(kb_step<NO>, dim3(nstep), dim3(256), sh, B->stream, dJ, dP);
    if (B->prof) hipEventRecord(e2, B->stream);
    if (B->prof) {
        hipEventRecord(e3, B->stream);
        B->ev_int.emplace_back(e1, e2);
        B->ev_step.emplace_back(e0, e3);
    }
    HIPCHK(hipEventRecord(B->ev[slot], B->stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FDFK_EDEVICE, std::string("batch launch: ") + hipGetErrorString(e));
    B->slot = (slot + 1) % fdfk_batch::R;
    for (size_t j = 0; j < act.size(); ++j) {
        fdfk_solver *s = act[j];
        if (rec[j]) { s->istp += 1; s->isis += s->nstp; }
        s->cur = 1 - s->cur;
        s->it += 1;
    }
    return 0;
}

}  // namespace

extern "C" {

int fdfk_batch_create(int32_t norder, fdfk_batch **out)
{
    if (!out) return fail(FDFK_EINVAL, "null output pointer");
    *out = nullptr;
    if (norder < 1 || norder > kMaxOrder) return fail(FDFK_EINVAL, "norder must be 1..8");
    auto *B = new fdfk_batch();
    B->norder = norder;
    with_order(norder, [&](auto no) { B->tabbytes = batch_table_bytes<decltype(no)::value>(); });
    if (hipStreamCreate(&B->stream) != hipSuccess) { delete B; return fail(FDFK_EDEVICE, "hipStreamCreate"); }
    for (int i = 0; i < fdfk_batch::R; ++i) {
        if (hipMalloc(&B->dtab[i], B->tabbytes) != hipSuccess ||
            hipHostMalloc(&B->htab[i], B->tabbytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreate(&B->ev[i]) != hipSuccess) {
            delete B;
            return fail(FDFK_ENOMEM, "batch table allocation");
        }
        hipEventRecord(B->ev[i], B->stream);
    }
    *out = B;
    return 0;
}

int fdfk_batch_add(fdfk_batch *B, fdfk_solver *s)
{
    if (!B || !s) return fail(FDFK_EINVAL, "null argument");
    if (s->g.norder != B->norder) return fail(FDFK_EINVAL, "solver FD order differs from the batch");
    if ((int)B->members.size() >= kMaxBatch) return fail(FDFK_ERANGE, "batch is full");
    for (auto *m : B->members)
        if (m == s) return fail(FDFK_EINVAL, "solver already in the batch");
    // the batch's next steps wait for the solver's queued work (background,
    // state) on the GPU; the host does not wait
    if (!s->ev_order) HIPCHK(hipEventCreateWithFlags(&s->ev_order, hipEventDisableTiming));
    HIPCHK(hipEventRecord(s->ev_order, s->stream));
    HIPCHK(hipStreamWaitEvent(B->stream, s->ev_order, 0));
    B->members.push_back(s);
    B->tact.clear();    // a new solver may sit at the address of a removed one
    return 0;
}

int fdfk_batch_remove(fdfk_batch *B, fdfk_solver *s)
{
    if (!B || !s) return fail(FDFK_EINVAL, "null argument");
    for (size_t i = 0; i < B->members.size(); ++i)
        if (B->members[i] == s) {
            // later work on the solver's stream (read-back, single-job steps,
            // destroy) waits on the GPU for the batch steps already queued
            if (!s->ev_order) HIPCHK(hipEventCreateWithFlags(&s->ev_order, hipEventDisableTiming));
            HIPCHK(hipEventRecord(s->ev_order, B->stream));
            HIPCHK(hipStreamWaitEvent(s->stream, s->ev_order, 0));
            B->members.erase(B->members.begin() + i);
            B->tact.clear();
            return 0;
        }
    return fail(FDFK_EINVAL, "solver not in the batch");
}

int fdfk_batch_run(fdfk_batch *B, int32_t nsteps)
{
    if (!B) return fail(FDFK_EINVAL, "null batch");
    for (int k = 0; k < nsteps; ++k) {
        int rc;
        with_order(B->norder, [&](auto no) { rc = batch_step<decltype(no)::value>(B); });
        if (rc) return rc;
    }
    return 0;
}

int fdfk_batch_active(const fdfk_batch *B)
{
    if (!B) return FDFK_EINVAL;
    int n = 0;
    for (auto *s : B->members) n += s->it <= s->nt;
    return n;
}

int fdfk_batch_sync(fdfk_batch *B)
{
    if (!B) return fail(FDFK_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(B->stream));
    return 0;
}

int fdfk_batch_profile(fdfk_batch *B, int32_t enable)
{
    if (!B) return fail(FDFK_EINVAL, "null batch");
    B->prof = enable != 0;
    return 0;
}

int fdfk_batch_kernel_stats(fdfk_batch *B, double *int_ms, int64_t *int_n, double *step_ms, int64_t *step_n)
{
    if (!B) return fail(FDFK_EINVAL, "null batch");
    HIPCHK(hipStreamSynchronize(B->stream));
    double a = 0, c = 0;
    int64_t na = 0, nc = 0;
    for (auto &e : B->ev_int) {
        float ms = 0;
        hipEventElapsedTime(&ms, e.first, e.second);
        a += ms; ++na;
        hipEventDestroy(e.first); hipEventDestroy(e.second);
    }
    for (auto &e : B->ev_step) {
        float ms = 0;
        hipEventElapsedTime(&ms, e.first, e.second);
        c += ms; ++nc;
        hipEventDestroy(e.first); hipEventDestroy(e.second);
    }
    B->ev_int.clear(); B->ev_step.clear();
    if (int_ms) *int_ms = a;
    if (int_n) *int_n = na;
    if (step_ms) *step_ms = c;
    if (step_n) *step_n = nc;
    return 0;
}

void fdfk_batch_destroy(fdfk_batch *B) { delete B; }

}  // extern "C"
