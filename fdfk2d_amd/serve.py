"""Continuous-batching request server for FDFK2D simulations on one GPU.

Requests arrive over time (a Poisson stream); each is admitted as soon as
its HBM footprint fits the budget.  A pool of host threads prepares admitted
requests (model, solver handle, FK background computed on the GPU on the
solver's own stream) while the main thread keeps the running batch stepping;
a prepared job joins the batch at the next chunk boundary and advances with
every other admitted job in the same batched launches.  A job leaves the
batch when its last iteration is done; its seismograms are read back and its
memory freed, making room for waiting requests.  Latency is measured from
arrival to seismograms on the host.

The plain path for comparison serves the same requests one at a time with
the single-job API (fdfk_run).
"""
import collections
import concurrent.futures as cf
import math
import os
import time

import numpy as np

from . import configs, geometry, refstate
from .lib import Batch, Solver


# ---------------------------------------------------------------------------
# request mix
# ---------------------------------------------------------------------------
def make_model(nx, nz, npml, dx, seed, perturb=True):
    """3-layer model with a smoothed 5% random perturbation (the large-solve
    recipe) on an (NX, NZ) padded grid."""
    NX, NZ = nx + 2 * npml, nz + npml
    vsl = [(0.0, 1500.0), (2000.0, 3600.0), (45000.0, 4500.0)]
    z = np.arange(NZ) * dx
    vs1 = np.zeros(NZ, np.float32)
    for zt, v in vsl:
        vs1[z >= zt] = v
    vs = np.repeat(vs1[None, :], NX, 0)
    if perturb:
        d = configs.smooth_noise_fft((NX, NZ), 5000.0 / dx, seed)
        vs = (vs * (1.0 + 0.05 * d)).astype(np.float32)
    vp = (np.float32(1.73) * vs).astype(np.float32)
    layers = [(zt, 1.73 * v, v) for zt, v in vsl]
    return vp, vs, configs.fk_layers(layers)


def request_mix(n, seed=5, rate=4.0, nx_range=(1000, 8000), nz_range=(400, 2000),
                steps_range=(3000, 20000), dx_choices=(50.0, 100.0, 200.0), p_range=(0.04, 0.08),
                npml=20, perturb=True):
    """Seeded requests: arrival times (Poisson), log-uniform grid sizes,
    uniform step counts, random P/S and ray parameter."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        f = rng.uniform()
        nx = int(round(math.exp(math.log(nx_range[0]) + f * math.log(nx_range[1] / nx_range[0]))))
        nz = int(round(math.exp(math.log(nz_range[0]) + f * math.log(nz_range[1] / nz_range[0]))))
        dx = float(rng.choice(dx_choices))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        reqs.append(dict(id=i, arrival=float(t[i]), nx=nx, nz=nz, dx=dx, nsteps=steps,
                         src_is_p=bool(rng.integers(0, 2)), p_skm=float(rng.uniform(*p_range)),
                         seed=int(rng.integers(1 << 30)), npml=npml, perturb=perturb))
    return reqs


def build_job(req):
    """Materials, FK model, receivers and dt for one request (dt = 0.9 of the
    reference's stability limit, src/Submain.f90:139-146)."""
    nx, nz, dx, npml = req["nx"], req["nz"], req["dx"], req["npml"]
    vp, vs, fk = make_model(nx, nz, npml, dx, req["seed"], req.get("perturb", True))
    c = refstate.material_law(vp, vs)
    dt = 0.9 * refstate.stable_dt(3, dx, dx, c)
    if "dt_max" in req:
        dt = min(dt, req["dt_max"])
    nrcv = max(2, min(401, nx // 20))
    rx = np.linspace(0.0, (nx - 1) * dx, nrcv).astype(np.float32)
    bazs = geometry.column_bazs(nx, npml, 0.0, dx, configs.LAT_ORG, configs.LON_ORG, configs.AZ_ORG,
                                configs.EVLA, configs.EVLO)
    return dict(c=c, fk=fk, dt=dt, rx=rx, rz=np.zeros_like(rx), bazs=bazs, nt=req["nsteps"] - 1,
                f0=0.5, ds=1e-3, rayp=req["p_skm"] / 1000.0)


# background synthesised in blocks of this many steps (fdfk_fk.block_steps):
# the FP64 traces and two blocks of the twelve arrays stay on the device
# instead of every step
BG_BLOCK = 512


def footprint_bytes(req, block_steps=BG_BLOCK):
    """(resident, transient) device bytes of a request: fields, materials,
    ring, PML and the background stay resident (every step, or with blocks
    two blocks plus the FP64 traces they are synthesised from); the FK work
    arrays (spectra and traces, FP64) exist while the background is
    computed.  dt is estimated at 0.9 of the stability limit for Vp <= 9 km/s."""
    nx, nz, npml = req["nx"], req["nz"], req["npml"]
    nbd = 5
    NX, NZ = nx + 2 * npml, nz + npml
    ld = (NZ + 31) // 32 * 32
    nt = req["nsteps"] - 1
    fields = 9 * NX * ld * 4 + 2 * NX * NZ * 4
    strips = (2 * npml * nz + NX * npml) * 4 * 15
    ring_pts = 2 * (nz - nbd) * nbd + nbd * nx + 2 * nz * nbd + nbd * (nx + 2 * nbd)
    dt = 0.9 * 0.5752 * req["dx"] / 9000.0
    f0 = 0.5
    tneed = nt * dt + req["p_skm"] / 1000.0 * (nx + 2 * nbd) * req["dx"] + 3.0 / f0 + 10.0 / f0
    nfft = 1 << int(math.ceil(math.log2(2.0 * tneed / (dt / 2.0))))
    fk_work = 2 * 2 * (nz + nbd) * nfft * 8 * 2
    if block_steps and 2 * block_steps < nt + 1:
        # two blocks of the arrays and one side's FP64 traces (the request's
        # left and right models are the same)
        bg = 2 * ring_pts * (2 * block_steps) * 4 + 2 * (nz + nbd) * nfft * 8
    else:
        bg = 2 * ring_pts * (nt + 1) * 4
    return fields + strips + bg, fk_work


def _keep(keep_results, rid):
    """keep_results: True (every request), False, or a set of request ids."""
    return keep_results is True or (bool(keep_results) and not isinstance(keep_results, bool) and rid in keep_results)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
# Solo streams (grids of at least solo_cells) are fed in chunks of
# SOLO_CHUNK iterations, at most SOLO_AHEAD chunks queued per stream, so that
# a large request's thousands of steps do not fill the hardware queues ahead
# of the batch and of new requests; while solo streams run, the batch steps
# in chunks of at most BATCH_CHUNK_WITH_SOLO, so that the loop comes back to
# refill them before they drain (a stream is refilled only once idle).
SOLO_AHEAD = 2
SOLO_CHUNK = 500
BATCH_CHUNK_WITH_SOLO = 25
PROGRESS_S = 20.0   # a progress line at least this often on long runs


class Server:
    def __init__(self, hbm_budget_gb=200.0, chunk=100, log=None, workers=None, solo_cells=4_000_000,
                 block_steps=BG_BLOCK):
        self.budget = hbm_budget_gb * 1e9
        self.block_steps = block_steps
        self.peak_used = 0.0     # largest admitted footprint total (bytes)
        self.chunk = chunk
        # grids of at least solo_cells padded cells fill the GPU alone: they
        # step on their own streams (the single-job kernels, whose frame
        # tiles overlap their interior's tail) beside the batch, which takes
        # the grids too small to fill the GPU
        self.solo_cells = solo_cells
        self.solo = {}
        # chunks queued on each solo stream since it was last seen idle: at
        # most SOLO_AHEAD, so that a large request's thousands of steps do
        # not fill the hardware queues ahead of the batch and new requests
        self.solo_queued = {}
        self.workers = workers or max(1, min(8, os.cpu_count() or 1))
        self.batch = Batch(3)
        self.used = 0.0
        self.active = {}
        self.prep_parts = dict(model=0.0, create=0.0, background=0.0)
        self.log = log or (lambda *a: None)

    def _prepare(self, req):
        """Host thread: model, solver and FK background of one request (the
        library calls release the GIL; the background runs on the solver's
        own stream, beside the batch's)."""
        t0 = time.perf_counter()
        job = build_job(req)
        t1 = time.perf_counter()
        s = Solver(req["nx"], req["nz"], req["npml"], 3, req["dx"], req["dx"], job["dt"], job["nt"],
                   job["c"], rx=job["rx"], rz=job["rz"])
        t2 = time.perf_counter()
        try:
            s.compute_background(job["fk"], job["fk"], req["src_is_p"], job["f0"], job["ds"],
                                 job["rayp"], job["bazs"], block_steps=self.block_steps)
        except Exception:
            s.close()
            raise
        t3 = time.perf_counter()
        return s, (t1 - t0, t2 - t1, t3 - t2)

    @staticmethod
    def _finish(s):
        """Host thread: seismograms of a solver that left the batch (its
        stream is ordered after the batch's last steps), then its memory."""
        try:
            return s.seismograms(), time.perf_counter()
        finally:
            s.close()

    def run(self, reqs, keep_results=False):
        """Serve reqs (arrival times relative to the start); returns per
        request latency and the total wall time.  The main thread only
        admits, joins and steps; preparation and read-back run on the pool,
        so the GPU is never drained while requests come and go."""
        pending = collections.deque(sorted(reqs, key=lambda r: r["arrival"]))
        waiting = collections.deque()
        preparing, finishing = {}, {}
        lat, results = {}, {}
        t0 = time.perf_counter()
        try:
            self._loop(pending, waiting, preparing, finishing, lat, results, keep_results, t0)
        except BaseException:
            # a failed request ends the run: leave no solver behind
            for rid, (req, s) in list(self.active.items()):
                self.batch.remove(s)
                s.close()
            self.active.clear()
            for rid, (req, s) in list(self.solo.items()):
                s.close()
            self.solo.clear()
            for f in preparing:
                if not f.cancelled() and f.exception() is None:
                    f.result()[0].close()
            for f in finishing:
                f.exception()               # read-back and close ran to the end
            raise
        total = time.perf_counter() - t0
        return lat, total, results

    def _loop(self, pending, waiting, preparing, finishing, lat, results, keep_results, t0):
        with cf.ThreadPoolExecutor(self.workers) as pool:
            last_log = 0.0
            while pending or waiting or preparing or self.active or self.solo or finishing:
                now = time.perf_counter() - t0
                if now - last_log >= PROGRESS_S:
                    last_log = now
                    self.log("%.0f s: %d served, %d batched, %d solo, %d preparing, %d waiting" % (
                        now, len(lat), len(self.active), len(self.solo), len(preparing), len(waiting) + len(pending)))
                while pending and pending[0]["arrival"] <= now:
                    waiting.append(pending.popleft())
                while waiting:
                    req = waiting[0]
                    res, work = footprint_bytes(req, self.block_steps)
                    # over budget: wait while anything holds device memory
                    # (batched, solo, preparing or finishing requests)
                    if self.used + res + work > self.budget and (self.active or self.solo or preparing or finishing):
                        break
                    waiting.popleft()
                    preparing[pool.submit(self._prepare, req)] = (req, work)
                    self.used += res + work
                    self.peak_used = max(self.peak_used, self.used)
                for f in [f for f in preparing if f.done()]:
                    req, work = preparing.pop(f)
                    s, secs = f.result()
                    for k, v in zip(("model", "create", "background"), secs):
                        self.prep_parts[k] += v
                    self.used -= work                      # FK work arrays are gone
                    if (req["nx"] + 2 * req["npml"]) * (req["nz"] + req["npml"]) >= self.solo_cells:
                        self.solo[req["id"]] = (req, s)
                    else:
                        self.batch.add(s)
                        self.active[req["id"]] = (req, s)
                for f in [f for f in finishing if f.done()]:
                    req = finishing.pop(f)
                    (sx, sz), t_done = f.result()
                    self.used -= footprint_bytes(req, self.block_steps)[0]
                    lat[req["id"]] = t_done - t0 - req["arrival"]
                    if _keep(keep_results, req["id"]):
                        results[req["id"]] = (sx, sz)
                    if len(lat) % 25 == 0:
                        self.log("served %d requests, %.1f s" % (len(lat), time.perf_counter() - t0))
                queued_solo = False
                for rid, (req, s) in list(self.solo.items()):
                    left = s.nt + 1 - s.iteration
                    if left > 0:
                        n = self.solo_queued.get(rid, 0)
                        if n >= SOLO_AHEAD and s.idle():
                            n = 0
                        if n < SOLO_AHEAD:
                            s.run(min(SOLO_CHUNK, left))   # queued on the solver's stream
                            self.solo_queued[rid] = n + 1
                            queued_solo = True
                    if s.iteration > s.nt:
                        del self.solo[rid]
                        self.solo_queued.pop(rid, None)
                        finishing[pool.submit(self._finish, s)] = req
                if not self.active:
                    wait = 0.05 if not pending else max(0.0, min(0.05, pending[0]["arrival"] - now))
                    if self.solo:
                        if not queued_solo:
                            time.sleep(0.0005)   # every solo stream holds SOLO_AHEAD chunks
                        continue
                    if preparing or finishing:
                        cf.wait(list(preparing) + list(finishing), timeout=wait,
                                return_when=cf.FIRST_COMPLETED)
                    elif pending:
                        time.sleep(wait)
                    continue
                self.batch.run(min(self.chunk, BATCH_CHUNK_WITH_SOLO) if self.solo else self.chunk)
                if self.batch.active < len(self.active):
                    for rid, (req, s) in list(self.active.items()):
                        if s.iteration > s.nt:
                            self.batch.remove(s)
                            del self.active[rid]
                            finishing[pool.submit(self._finish, s)] = req

    def close(self):
        self.batch.close()


def serve_plain(reqs, keep_results=False, stats=None, block_steps=BG_BLOCK, log=None):
    """One request at a time with the single-job API, FIFO by arrival, the
    background synthesised in blocks as the server does.  stats, if a dict,
    accumulates the seconds of each phase."""
    lat, results = {}, {}
    st = stats if stats is not None else {}
    for k in ("model", "create", "background", "run"):
        st.setdefault(k, 0.0)
    t0 = time.perf_counter()
    for req in sorted(reqs, key=lambda r: r["arrival"]):
        now = time.perf_counter() - t0
        if req["arrival"] > now:
            time.sleep(req["arrival"] - now)
        ta = time.perf_counter()
        job = build_job(req)
        tb = time.perf_counter()
        s = Solver(req["nx"], req["nz"], req["npml"], 3, req["dx"], req["dx"], job["dt"], job["nt"],
                   job["c"], rx=job["rx"], rz=job["rz"])
        tc = time.perf_counter()
        s.compute_background(job["fk"], job["fk"], req["src_is_p"], job["f0"], job["ds"], job["rayp"],
                             job["bazs"], block_steps=block_steps)
        td = time.perf_counter()
        s.run(job["nt"] + 1)
        sx, sz = s.seismograms()
        s.close()
        te = time.perf_counter()
        for k, v in zip(("model", "create", "background", "run"), (tb - ta, tc - tb, td - tc, te - td)):
            st[k] += v
        lat[req["id"]] = time.perf_counter() - t0 - req["arrival"]
        if _keep(keep_results, req["id"]):
            results[req["id"]] = (sx, sz)
        if log and len(lat) % 25 == 0:
            log("plain path: %d requests, %.1f s" % (len(lat), time.perf_counter() - t0))
    return lat, time.perf_counter() - t0, results


def summarize(lat, total, n):
    v = np.array(sorted(lat.values()))
    return dict(requests=n, wall_s=round(total, 3), solves_per_hour=round(n / total * 3600.0, 1),
                p50_s=round(float(np.percentile(v, 50)), 4), p99_s=round(float(np.percentile(v, 99)), 4))
