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

// One member's arguments in a batch table.  A table is built when the set
// of active members changes and then serves every step until the next
// change: the level pointers alternate (one table per step parity), and the
// iteration-dependent values follow from dstep, the steps taken since the
// table was built: iteration sa.it + dstep; a recording at every nstp-th
// step from rec_first on (rec_left of them, -1: none), into sample
// ra.istp + (dstep - rec_first) / nstp.
template <int NO>
struct BatchJob {
    StepArgs sa;
    IntArgs<NO> ia;
    ETiles ft;
    RecArgs ra;
    const Background *bg;
    float *ringx, *ringz;
    int rec_first, nstp, rec_left;
};

struct BatchPrefix {
    int n;
    int pre[KB_N][kMaxBatch + 1];
};
// table layout (host slot and device slot alike): the prefix block, then
// from this offset the n jobs of even steps followed by the n jobs of odd
// steps, so that a table is one contiguous copy
constexpr size_t kJobOff = (sizeof(BatchPrefix) + 255) / 256 * 256;

__device__ __forceinline__ int find_job(const int *pre, int n, int bid)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= bid) lo = mid; else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);   // block-uniform
}

// The step count since the active set's tables were built: launched
// eagerly the kernels take it as an argument (hstep); replayed as a graph
// (FDFK_GRAPH=1) they read it from the device (ctr[0]), which kb_tick
// advances after each step, so that the launches of every step are
// identical and one step can be built as a graph once per active set.  Eager launches
// are the default: on this runtime a replayed three-node graph was not
// faster than the three launches (16 jobs of one active set: 0.220 s
// against 0.212-0.215 s), and the rebuilds at every join and finish cost
// a mix of small requests 12%.
__global__ void kb_tick(int *ctr) { ctr[0] += 1; }
__global__ void kb_set(int *ctr, int v) { ctr[0] = v; }

// one launch for the per-step preparation of every member: the hybrid ring
// background, the PML memories and (on recording iterations) the receivers,
// in that block order; 256-lane blocks, kRecPerBlock receivers per block
template <int NO>
__global__ __launch_bounds__(256) void kb_prep(const BatchJob<NO> *__restrict__ J0, const BatchPrefix *__restrict__ P,
                                                const int *ctr, int hstep)
{
    const int n = P->n;
    const int dstep = ctr ? ctr[0] : hstep;
    const BatchJob<NO> *J = J0 + (dstep & 1) * n;
    int b = blockIdx.x;
    const int nr = P->pre[KB_RING][n], np = P->pre[KB_PSI][n];
    if (b < nr) {
        const int j = find_job(P->pre[KB_RING], n, b);
        const BatchJob<NO> &q = J[j];
        ring_prep_body(q.sa.g, q.bg, q.sa.it + dstep, q.ringx, q.ringz, b - P->pre[KB_RING][j]);
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
    const BatchJob<NO> &q = J[j];
    const int d = dstep - q.rec_first;     // block-uniform: the whole block records or leaves
    if (q.rec_first < 0 || d < 0 || d % q.nstp != 0 || d / q.nstp >= q.rec_left) return;
    record_body(q.ra, b - P->pre[KB_REC][j], q.ra.istp + d / q.nstp);
}

// the interior tiles of every member, then their frame tiles, in one launch
// (the single-job k_step arrangement).  The job table is read-only in these
// kernels and aliases none of the arrays they write (__restrict__): its
// fields are then read through the scalar cache wherever the march needs
// them; without it, every field read after the block's first store was a
// per-lane vector load (+60% time on a one-member batch)
template <int NO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kIntWaves<NO>))) void kb_step(
    const BatchJob<NO> *__restrict__ J0, const BatchPrefix *__restrict__ P, const int *ctr, int hstep)
{
    extern __shared__ float sm[];
    const int n = P->n;
    const BatchJob<NO> *J = J0 + ((ctr ? ctr[0] : hstep) & 1) * n;
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
    // pacing: the host runs at most R steps ahead of the GPU, so that a
    // solver added to the batch joins within a few steps (an event after
    // every fourth step instead was 2 us per step faster on a one-member
    // batch and 3% slower on a mix of small requests)
    static constexpr int R = 4;
    hipEvent_t ev[R] = {};
    int slot = 0;
    // job tables of the current active set, double-buffered: a table is
    // rebuilt (and copied once) only when the active set changes; tev[t]
    // marks the last step that read table t
    static constexpr int TS = 2;
    void *dtab[TS] = {}, *htab[TS] = {};
    hipEvent_t tev[TS] = {};
    int tslot = 0;
    size_t tabbytes = 0;
    std::vector<fdfk_solver *> tact;
    int tn = 0, dstep = 0, nprep = 0, nstepb = 0;
    size_t tsh = 0;
    bool prof = false;
    // device step counter of the current tables, and the step of the
    // current active set as a graph (rebuilt with the tables)
    int *dctr = nullptr;
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // Graphs of earlier active sets whose replays may still be queued: a
    // graph is destroyed only once the step of its last replay is known to
    // be complete (the pacing wait below covers steps up to tot - R)
    struct Retired { hipGraphExec_t exec; hipGraph_t graph; long last; };
    std::vector<Retired> retired;
    long tot = 0;      // batched steps queued so far
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_int, ev_step;
    // Interior tile width of the current tables, and the online tuner
    // (fdfk_batch_autotune): for an active set not seen before, the first
    // batched steps run on candidate widths around the occupancy model's,
    // each timed over kTuneSteps steps after one untimed step, and the set
    // keeps the fastest (the model's unless another is 1% faster).  The
    // width changes the tiling only, not a cell's arithmetic.
    int tlx = 0;
    bool tune = false;
    static constexpr int kTuneSteps = 4;
    std::vector<int> tcand;                  // candidates of the set being tuned
    std::vector<float> tms;                  // their ms per step
    int tk = -1, tc = 0;                     // candidate, steps on its tables
    hipEvent_t te0 = nullptr, te1 = nullptr;
    std::vector<long> tkey;                  // shape of the set being tuned
    std::map<std::vector<long>, int> tuned;  // shape of an active set -> width
    ~fdfk_batch()
    {
        if (stream) hipStreamSynchronize(stream);
        if (te0) hipEventDestroy(te0);
        if (te1) hipEventDestroy(te1);
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        for (auto &r : retired) {
            hipGraphExecDestroy(r.exec);
            hipGraphDestroy(r.graph);
        }
        if (dctr) hipFree(dctr);
        for (int i = 0; i < R; ++i)
            if (ev[i]) hipEventDestroy(ev[i]);
        for (int i = 0; i < TS; ++i) {
            if (dtab[i]) hipFree(dtab[i]);
            if (htab[i]) hipHostFree(htab[i]);
            if (tev[i]) hipEventDestroy(tev[i]);
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
    return kJobOff + 2 * sizeof(BatchJob<NO>) * kMaxBatch;
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
    for (int lx = 4; lx <= 256; lx += (lx < 32 ? 1 : 8)) {
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

// the level pointers of a job for a step whose current level is c
template <int NO>
void set_levels(const fdfk_solver *s, BatchJob<NO> &b, int c)
{
    b.sa.ux = s->ux[c]; b.sa.uz = s->uz[c];
    b.sa.uxp = s->ux[1 - c]; b.sa.uzp = s->uz[1 - c];
    b.sa.uxo = b.sa.uxp; b.sa.uzo = b.sa.uzp;
    b.sa.st = s->d_strips[c];
    b.ia.ux = s->ux[c]; b.ia.uz = s->uz[c];
    b.ia.uxp = s->ux[1 - c]; b.ia.uzp = s->uz[1 - c];
    b.ia.uxo = b.ia.uxp; b.ia.uzo = b.ia.uzp;
    b.ra.ux = s->ux[c]; b.ra.uz = s->uz[c];
}

// build and upload the job tables of a new active set
template <int NO>
int build_tables(fdfk_batch *B, const std::vector<fdfk_solver *> &act, int lx_force = 0)
{
    constexpr size_t JB = sizeof(BatchJob<NO>);
    // last step on the old table (recorded even when add/remove cleared tact:
    // steps in flight may still read it)
    HIPCHK(hipEventRecord(B->tev[B->tslot], B->stream));
    B->tslot = (B->tslot + 1) % fdfk_batch::TS;
    HIPCHK(hipEventSynchronize(B->tev[B->tslot]));   // its previous user is done
    const int n = (int)act.size();
    auto *P = (BatchPrefix *)B->htab[B->tslot];
    auto *J0 = (BatchJob<NO> *)((char *)B->htab[B->tslot] + kJobOff);
    auto *J1 = J0 + n;
    P->n = n;
    for (int k = 0; k < KB_N; ++k) P->pre[k][0] = 0;
    const int lx = lx_force > 0 ? lx_force : batch_lx<NO>(act);
    B->tlx = lx;
    size_t sh = sizeof(typename IntKernel2<NO, 256>::Smem);
    for (int j = 0; j < n; ++j) {
        fdfk_solver *s = act[j];
        int nb[KB_N];
        bool r = false;
        fill_job<NO>(s, J0[j], nb, r, lx);
        const bool rec = s->nrcv > 0 && s->istp <= s->nsis;
        J0[j].rec_first = rec ? s->isis - s->it : -1;
        J0[j].nstp = s->nstp;
        J0[j].rec_left = rec ? s->nsis - s->istp + 1 : 0;
        nb[KB_REC] = rec ? (s->nrcv + kRecPerBlock - 1) / kRecPerBlock : 0;
        J1[j] = J0[j];
        set_levels<NO>(s, J1[j], 1 - s->cur);
        sh = std::max(sh, frame_lds_bytes(J0[j].ft, NO));
        for (int k = 0; k < KB_N; ++k) P->pre[k][j + 1] = P->pre[k][j] + nb[k];
    }
    HIPCHK(hipMemcpyAsync(B->dtab[B->tslot], B->htab[B->tslot], kJobOff + 2 * JB * n, hipMemcpyHostToDevice,
                          B->stream));
    hipLaunchKernelGGL(kb_set, dim3(1), dim3(1), 0, B->stream, B->dctr, 0);
    HIPCHK(hipGetLastError());
    // the graph's step refers to the old tables: build it again; the old
    // graph may still have replays queued, so it is retired, not destroyed
    if (B->gexec) B->retired.push_back({B->gexec, B->graph, B->tot - 1});
    B->gexec = nullptr;
    B->graph = nullptr;
    B->tact = act;
    B->tn = n;
    B->dstep = 0;
    B->tsh = sh;
    B->nprep = P->pre[KB_RING][n] + P->pre[KB_PSI][n] + P->pre[KB_REC][n];
    B->nstepb = P->pre[KB_INT][n] + P->pre[KB_FRAME][n];
    return 0;
}

// shape of an active set for the tuner: the members' grids, sorted
static std::vector<long> set_shape(const std::vector<fdfk_solver *> &act)
{
    std::vector<long> k;
    for (auto *s : act) k.push_back(((long)s->g.nx << 32) | ((long)s->g.nz << 12) | (long)s->g.npml);
    std::sort(k.begin(), k.end());
    return k;
}

// Before a step: new tables for a new active set (its tuned width when
// known, the first candidate when it is to be tuned), and the tuner's
// events and switches between candidates
template <int NO>
int tune_tables(fdfk_batch *B, const std::vector<fdfk_solver *> &act)
{
    if (act != B->tact) {
        B->tk = -1;
        int lx = 0;
        if (B->tune) {
            std::vector<long> key = set_shape(act);
            auto it = B->tuned.find(key);
            if (it != B->tuned.end()) {
                lx = it->second;
            } else {
                const int m = batch_lx<NO>(act);
                B->tcand.clear();
                for (double f : {1.0, 0.5, 0.67, 0.8, 1.25, 1.5, 2.0}) {
                    const int c = std::min(256, std::max(4, (int)std::lround(m * f)));
                    if (std::find(B->tcand.begin(), B->tcand.end(), c) == B->tcand.end()) B->tcand.push_back(c);
                }
                B->tms.assign(B->tcand.size(), 0.0f);
                B->tk = 0;
                B->tc = 0;
                B->tkey = key;
                lx = B->tcand[0];
            }
        }
        return build_tables<NO>(B, act, lx);
    }
    if (B->tk < 0) return 0;
    // one untimed step on the candidate's tables, then kTuneSteps timed
    if (B->tc == 1) HIPCHK(hipEventRecord(B->te0, B->stream));
    if (B->tc == 1 + fdfk_batch::kTuneSteps) {
        HIPCHK(hipEventRecord(B->te1, B->stream));
        float ms = 0.0f;
        HIPCHK(hipEventSynchronize(B->te1));
        HIPCHK(hipEventElapsedTime(&ms, B->te0, B->te1));
        B->tms[B->tk] = ms / fdfk_batch::kTuneSteps;
        B->tc = 0;
        if (++B->tk < (int)B->tcand.size()) return build_tables<NO>(B, act, B->tcand[B->tk]);
        size_t bk = 0;
        for (size_t k = 1; k < B->tcand.size(); ++k)
            if (B->tms[k] < B->tms[bk]) bk = k;
        if (B->tms[bk] > 0.99f * B->tms[0]) bk = 0;
        B->tuned[B->tkey] = B->tcand[bk];
        B->tk = -1;
        return build_tables<NO>(B, act, B->tcand[bk]);
    }
    return 0;
}

template <int NO>
int batch_step(fdfk_batch *B)
{
    std::vector<fdfk_solver *> act;
    for (auto *s : B->members)
        if (s->it <= s->nt) act.push_back(s);
    if (act.empty()) return 0;
    if (int rc = tune_tables<NO>(B, act)) return rc;
    if (B->tk >= 0) B->tc += 1;
    const int slot = B->slot;
    HIPCHK(hipEventSynchronize(B->ev[slot]));   // at most R steps ahead of the GPU: steps < tot - R + 1 are done
    for (size_t i = 0; i < B->retired.size();) {
        const auto r = B->retired[i];
        if (r.last <= B->tot - fdfk_batch::R) {
            HIPCHK(hipGraphExecDestroy(r.exec));
            HIPCHK(hipGraphDestroy(r.graph));
            B->retired.erase(B->retired.begin() + i);
        } else {
            ++i;
        }
    }
    for (fdfk_solver *s : act)                  // members whose background is synthesised in blocks
        if (int rc = bg_refill(s, B->stream)) return rc;
    auto *dP = (const BatchPrefix *)B->dtab[B->tslot];
    auto *dJ = (const BatchJob<NO> *)((char *)B->dtab[B->tslot] + kJobOff);
    // the launches of one step: prep and step, with the step count as an
    // argument (no device counter)
    const int *nctr = nullptr;
    const int hs = B->dstep;
    auto launch = [&](hipStream_t st) {
        if (B->nprep) hipLaunchKernelGGL(kb_prep<NO>, dim3(B->nprep), dim3(256), 0, st, dJ, dP, nctr, hs);
        if (B->nstepb) hipLaunchKernelGGL(kb_step<NO>, dim3(B->nstepb), dim3(256), B->tsh, st, dJ, dP, nctr, hs);
    };
    if (B->prof) {
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
        hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2); hipEventCreate(&e3);
        hipEventRecord(e0, B->stream);
        if (B->nprep) hipLaunchKernelGGL(kb_prep<NO>, dim3(B->nprep), dim3(256), 0, B->stream, dJ, dP, nctr, hs);
        hipEventRecord(e1, B->stream);
        if (B->nstepb)
            hipLaunchKernelGGL(kb_step<NO>, dim3(B->nstepb), dim3(256), B->tsh, B->stream, dJ, dP, nctr, hs);
        hipEventRecord(e2, B->stream);
        // with graph replay the device counter must stay in step for the
        // replays after profiling is switched off
        if (B->use_graph) hipLaunchKernelGGL(kb_tick, dim3(1), dim3(1), 0, B->stream, B->dctr);
        hipEventRecord(e3, B->stream);
        B->ev_int.emplace_back(e1, e2);
        B->ev_step.emplace_back(e0, e3);
    } else if (B->use_graph) {
        if (!B->gexec) {
            // the step of this active set as a graph, built once; every later
            // step of the set replays it (its arguments do not change).  The
            // graph is assembled node by node rather than captured from the
            // stream: a capture in progress makes HIP refuse the
            // stream-ordered allocations other threads make meanwhile for
            // the solvers they prepare.
            const int *ctr = B->dctr;
            int zero = 0;
            void *args[4] = {(void *)&dJ, (void *)&dP, (void *)&ctr, (void *)&zero};
            int *tick = B->dctr;
            void *targs[1] = {(void *)&tick};
            hipGraph_t gr = nullptr;
            HIPCHK(hipGraphCreate(&gr, 0));
            hipGraphNode_t prev = nullptr;
            auto add = [&](const void *fn, unsigned nb, unsigned nt, size_t sh, void **kargs) -> hipError_t {
                hipKernelNodeParams kp{};
                kp.func = const_cast<void *>(fn);
                kp.gridDim = dim3(nb);
                kp.blockDim = dim3(nt);
                kp.sharedMemBytes = (unsigned)sh;
                kp.kernelParams = kargs;
                kp.extra = nullptr;
                hipGraphNode_t n = nullptr;
                hipError_t e = hipGraphAddKernelNode(&n, gr, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
                prev = n;
                return e;
            };
            hipError_t ge = hipSuccess;
            if (B->nprep) ge = add((const void *)kb_prep<NO>, (unsigned)B->nprep, 256, 0, args);
            if (ge == hipSuccess && B->nstepb) ge = add((const void *)kb_step<NO>, (unsigned)B->nstepb, 256, B->tsh, args);
            if (ge == hipSuccess) ge = add((const void *)kb_tick, 1, 1, 0, targs);
            if (ge == hipSuccess) ge = hipGraphInstantiate(&B->gexec, gr, nullptr, nullptr, 0);
            if (ge != hipSuccess) {
                (void)hipGraphDestroy(gr);
                B->gexec = nullptr;
                return fail(FDFK_EDEVICE, std::string("batch step graph: ") + hipGetErrorString(ge));
            }
            B->graph = gr;
        }
        HIPCHK(hipGraphLaunch(B->gexec, B->stream));
    } else {
        launch(B->stream);
    }
    HIPCHK(hipEventRecord(B->ev[slot], B->stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FDFK_EDEVICE, std::string("batch launch: ") + hipGetErrorString(e));
    B->slot = (slot + 1) % fdfk_batch::R;
    B->dstep += 1;
    B->tot += 1;
    // host copies of the members' counters advance as the device tables assume
    for (fdfk_solver *s : act) {
        if ((s->isis == s->it) && (s->istp <= s->nsis) && s->nrcv > 0) { s->istp += 1; s->isis += s->nstp; }
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
        if (hipEventCreateWithFlags(&B->ev[i], hipEventDisableTiming) != hipSuccess) {
            delete B;
            return fail(FDFK_EDEVICE, "batch event");
        }
        hipEventRecord(B->ev[i], B->stream);
    }
    if (const char *e = std::getenv("FDFK_GRAPH")) B->use_graph = std::atoi(e) != 0;
    if (hipMalloc((void **)&B->dctr, sizeof(int)) != hipSuccess) {
        delete B;
        return fail(FDFK_ENOMEM, "batch counter allocation");
    }
    hipLaunchKernelGGL(kb_set, dim3(1), dim3(1), 0, B->stream, B->dctr, 0);
    for (int i = 0; i < fdfk_batch::TS; ++i) {
        if (hipMalloc(&B->dtab[i], B->tabbytes) != hipSuccess ||
            hipHostMalloc(&B->htab[i], B->tabbytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&B->tev[i], hipEventDisableTiming) != hipSuccess) {
            delete B;
            return fail(FDFK_ENOMEM, "batch table allocation");
        }
        hipEventRecord(B->tev[i], B->stream);
    }
    *out = B;
    return 0;
}

int fdfk_batch_add(fdfk_batch *B, fdfk_solver *s)
{
    if (!B || !s) return fail(FDFK_EINVAL, "null argument");
    if (s->g.norder != B->norder) return fail(FDFK_EINVAL, "solver FD order differs from the batch");
    if (s->mat16) return fail(FDFK_ENOSYS, "16-bit-material solvers are not batched");
    if (s->gen) return fail(FDFK_ENOSYS, "solvers with an absorbing top layer are not batched");
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

int fdfk_batch_autotune(fdfk_batch *B, int32_t enable)
{
    if (!B) return fail(FDFK_EINVAL, "null batch");
    if (enable && !B->te0) {
        if (hipEventCreate(&B->te0) != hipSuccess || hipEventCreate(&B->te1) != hipSuccess)
            return fail(FDFK_EDEVICE, "tuner events");
    }
    B->tune = enable != 0;
    if (!B->tune) B->tk = -1;
    return 0;
}

int fdfk_batch_tile_width(const fdfk_batch *B)
{
    if (!B) return FDFK_EINVAL;
    return B->tlx;
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
