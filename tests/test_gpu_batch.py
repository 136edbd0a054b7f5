"""Batched execution and serving on the GPU: every job gives the same
seismograms alone (single-job API) and in the mix (batched launches with
other jobs of different sizes, joining at different times)."""
import numpy as np
import pytest

from fdfk2d_amd import serve
from fdfk2d_amd.lib import Batch, FdfkError, Solver, trim_memory

pytestmark = pytest.mark.gpu


def _reqs():
    return [dict(id=0, arrival=0.0, nx=181, nz=61, dx=200.0, nsteps=240, src_is_p=True, p_skm=0.06, seed=1, npml=20),
            dict(id=1, arrival=0.0, nx=97, nz=45, dx=100.0, nsteps=300, src_is_p=False, p_skm=0.05, seed=2, npml=20),
            dict(id=2, arrival=0.0, nx=240, nz=80, dx=50.0, nsteps=200, src_is_p=True, p_skm=0.075, seed=3, npml=20)]


def _solver(req, nstp=1, block=0):
    job = serve.build_job(req)
    s = Solver(req["nx"], req["nz"], req["npml"], 3, req["dx"], req["dx"], job["dt"], job["nt"], job["c"],
               rx=job["rx"], rz=job["rz"], nstp=nstp)
    s.compute_background(job["fk"], job["fk"], req["src_is_p"], job["f0"], job["ds"], job["rayp"], job["bazs"],
                         block_steps=block)
    return s


@pytest.mark.parametrize("nstp,block", [(1, 0), (3, 0), (1, 32), (3, 32)])
def test_batch_equals_alone_with_staggered_joins(nstp, block):
    """Joins after 37 and 87 batched steps (not multiples of nstp): the
    batch tables are rebuilt at each join and each finish, with the
    recording phase of every member carried across.  With block > 0 the
    members' backgrounds are synthesised in blocks on the batch's stream;
    the reference runs alone store every step."""
    print(f"case: 3 jobs alone vs batched, staggered joins, nstp {nstp}, block {block}", flush=True)
    reqs = _reqs()
    alone = []
    for r in reqs:
        s = _solver(r, nstp)
        s.run(s.nt + 1)
        alone.append(s.seismograms())
        s.close()
    b = Batch(3)
    sol = [_solver(r, nstp, block) for r in reqs]
    b.add(sol[0])
    b.run(37)
    b.add(sol[1])
    b.run(50)
    b.add(sol[2])
    while b.active:
        b.run(25)
    b.sync()
    for s, (ax, az) in zip(sol, alone):
        gx, gz = s.seismograms()
        assert np.abs(az).max() > 0
        assert np.array_equal(gx, ax) and np.array_equal(gz, az)
        s.close()
    b.close()


def test_autotuned_batch_equals_alone():
    """With the online tile-width tuner on, the active sets' first batched
    steps run on several interior tile widths (the tables are rebuilt at
    every candidate) before each set keeps its fastest; a set of a shape
    tuned before takes its width at once.  Every job still gives its
    seismograms alone bit for bit, and the width kept is one of the
    candidates around the occupancy model's."""
    print("case: 3 jobs alone vs batched with the tile-width tuner, joins after 60 and 110 steps", flush=True)
    reqs = _reqs()
    alone = []
    for r in reqs:
        s = _solver(r)
        s.run(s.nt + 1)
        alone.append(s.seismograms())
        s.close()
    b = Batch(3)
    b.autotune(True)
    sol = [_solver(r) for r in reqs]
    b.add(sol[0])
    b.run(60)               # the one-member set: 7 candidates x 5 steps, then its width
    w1 = b.tile_width
    b.add(sol[1])
    b.run(50)
    b.add(sol[2])
    widths = set()
    while b.active:
        b.run(5)
        widths.add(b.tile_width)
    b.sync()
    for s, (ax, az) in zip(sol, alone):
        gx, gz = s.seismograms()
        assert np.abs(az).max() > 0
        assert np.array_equal(gx, ax) and np.array_equal(gz, az)
        s.close()
    b.close()
    assert 4 <= w1 <= 256
    assert len(widths) >= 2      # the three-member set was tuned over several widths


@pytest.mark.parametrize("block", [0, 32])
def test_graph_replayed_steps_equal_eager_launches(block, monkeypatch):
    """A batched step is built as a graph once per active set and replayed
    (FDFK_GRAPH=1); FDFK_GRAPH=0 (the default) launches its kernels one by
    one.  Both give the same seismograms and final wavefields bit for
    bit, with table rebuilds at staggered joins and finishes and (block 32)
    background refills launched eagerly between the replays."""
    print("case: 3 jobs batched, graph replay vs eager launches, block %d" % block, flush=True)

    def batched(graph):
        monkeypatch.setenv("FDFK_GRAPH", "1" if graph else "0")
        b = Batch(3)
        sol = [_solver(r, 1, block) for r in _reqs()]
        b.add(sol[0])
        b.run(41)
        b.add(sol[1])
        b.run(33)
        b.add(sol[2])
        while b.active:
            b.run(19)
        b.sync()
        out = [s.seismograms() + s.get_state() for s in sol]
        for s in sol:
            s.close()
        b.close()
        return out

    eager = batched(False)
    graph = batched(True)
    for g, e in zip(graph, eager):
        assert np.abs(e[1]).max() > 0
        for a, b_ in zip(g, e):
            assert np.array_equal(a, b_)


def test_graph_rebuilds_while_replays_are_queued(monkeypatch):
    """Twelve members of different lengths joining every 7 steps and leaving
    as they finish, steps replayed as graphs (FDFK_GRAPH=1): the active set
    changes about 20 times while the host keeps up to four replays queued
    ahead of the GPU, so the graph of an active set is replaced while
    replays of it are still pending (those graphs are destroyed only after
    their last replay completed).  Every member equals its run alone, bit
    for bit."""
    print("case: 12 members, joins every 7 steps, graph replay", flush=True)
    monkeypatch.setenv("FDFK_GRAPH", "1")
    reqs = []
    for i in range(12):
        reqs.append(dict(id=i, arrival=0.0, nx=61 + 10 * i, nz=40 + 3 * i, dx=200.0, nsteps=40 + 17 * i,
                         src_is_p=(i % 2 == 0), p_skm=0.05 + 0.002 * i, seed=100 + i, npml=20))
    alone = []
    for r in reqs:
        s = _solver(r)
        s.run(s.nt + 1)
        alone.append(s.seismograms())
        s.close()
    b = Batch(3)
    sol = [_solver(r) for r in reqs]
    for s in sol:
        b.add(s)
        b.run(7)
    while b.active:
        b.run(5)
    b.sync()
    for s, (ax, az) in zip(sol, alone):
        gx, gz = s.seismograms()
        assert np.abs(az).max() > 0
        assert np.array_equal(gx, ax) and np.array_equal(gz, az)
        s.close()
    b.close()


def test_server_results_equal_plain_path():
    print("case: serving 6 requests vs plain serial path", flush=True)
    reqs = serve.request_mix(6, seed=7, rate=20.0, nx_range=(120, 400), nz_range=(50, 120),
                             steps_range=(150, 400), dx_choices=(100.0, 200.0))
    srv = serve.Server(hbm_budget_gb=50.0, chunk=40)
    lat, total, res = srv.run(reqs, keep_results=True)
    srv.close()
    trim_memory()           # every solver is closed: the pool's memory goes back
    lat_p, total_p, res_p = serve.serve_plain(reqs, keep_results=True)
    assert sorted(res) == sorted(res_p) == list(range(6))
    for i in res:
        assert np.array_equal(res[i][0], res_p[i][0]) and np.array_equal(res[i][1], res_p[i][1])
    print("served", serve.summarize(lat, total, 6), "plain", serve.summarize(lat_p, total_p, 6), flush=True)


@pytest.mark.parametrize("block", [1, 37, 64])
def test_blocked_background_equals_stored(block):
    """Background synthesised in blocks of `block` steps (a ring of two
    blocks, refilled as the run advances; 37 does not divide the 240
    iterations) against all steps stored: the same seismograms, bit for bit,
    also when the run is advanced in uneven chunks."""
    print(f"case: blocked background, block {block}, 181x61, 240 iterations", flush=True)
    req = _reqs()[0]
    full = _solver(req)
    full.run(full.nt + 1)
    want = full.seismograms()
    full.close()
    s = _solver(req, block=block)
    s.run(50)
    s.run(101)
    s.run(s.nt + 1 - 151)
    got = s.seismograms()
    assert np.abs(want[1]).max() > 0
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    with pytest.raises(FdfkError, match="blocks"):
        s.get_background()
    s.close()


def test_server_failed_request_raises_and_cleans_up():
    """A request that fdfk_create rejects on the host (a PML narrower than the
    hybrid ring: npml 3 < 2*norder - 1 = 5) ends the run with that error;
    no solver stays in the batch, and a new server then serves normally."""
    print("case: serving with one invalid request, then a normal run", flush=True)
    reqs = serve.request_mix(4, seed=9, rate=50.0, nx_range=(120, 300), nz_range=(50, 100),
                             steps_range=(100, 200), dx_choices=(200.0,))
    bad = dict(reqs[2], npml=3)
    srv = serve.Server(hbm_budget_gb=50.0, chunk=40)
    with pytest.raises(FdfkError, match="npml"):
        srv.run(reqs[:2] + [bad] + reqs[3:])
    assert not srv.active and srv.batch.active == 0
    srv.close()
    srv = serve.Server(hbm_budget_gb=50.0, chunk=40)
    lat, total, res = srv.run(reqs, keep_results=True)
    srv.close()
    assert sorted(lat) == [r["id"] for r in reqs] and all(np.isfinite(res[r["id"]][1]).all() for r in reqs)


def test_batch_equals_alone_at_the_batched_workload_shapes():
    """Three jobs drawn as in the batched workload (seed 3: nx uniform in
    2000..8000, nz in 800..2000, dx = dz = 50 m, 40-cell PML, P and SV
    alternating, ray parameter uniform in 0.04..0.08 s/km, layered model
    with its random perturbation), 1200 iterations each (3.6 s: the incident
    pulse, delayed 3 s, is in the grid): batched launches over
    large uneven grids (many interior tiles per member, one tile width for
    the batch) give each job's seismograms and final wavefield of the job
    alone, bit for bit."""
    rng = np.random.default_rng(3)
    reqs = [dict(id=k, arrival=0.0, nx=int(rng.integers(2000, 8001)), nz=int(rng.integers(800, 2001)), dx=50.0,
                 nsteps=1200, src_is_p=(k % 2 == 0), p_skm=float(rng.uniform(0.04, 0.08)), seed=30 + k, npml=40)
            for k in range(3)]
    print("case: batched workload shapes %s, 1200 iterations" % [(r["nx"], r["nz"]) for r in reqs], flush=True)
    alone = []
    for r in reqs:
        s = _solver(r)
        s.run(s.nt + 1)
        alone.append(s.seismograms() + s.get_state())
        s.close()
    b = Batch(3)
    sol = [_solver(r) for r in reqs]
    for s in sol:
        b.add(s)
    while b.active:
        b.run(200)
    b.sync()
    for s, ref in zip(sol, alone):
        got = s.seismograms() + s.get_state()
        print("max |uz| %.3g" % np.abs(ref[3]).max(), flush=True)
        assert np.abs(ref[3]).max() > 1e-6
        for a, e in zip(got, ref):
            assert np.array_equal(a, e)
        s.close()
    b.close()


def test_server_with_solo_requests_equals_plain_path():
    """Requests of at least solo_cells padded cells step on their own
    streams beside the batch (here the threshold is lowered so that the mix
    splits between the two): every result equals the plain path's, bit for
    bit, and both kinds finish."""
    reqs = serve.request_mix(6, seed=7, rate=20.0, nx_range=(120, 400), nz_range=(50, 120),
                             steps_range=(150, 400), dx_choices=(100.0, 200.0))
    cells = sorted((r["nx"] + 2 * r["npml"]) * (r["nz"] + r["npml"]) for r in reqs)
    thr = cells[3]          # the three largest step alone
    print("case: serving 6 requests, solo threshold %d cells, vs plain serial path" % thr, flush=True)
    srv = serve.Server(hbm_budget_gb=50.0, chunk=40, solo_cells=thr)
    lat, total, res = srv.run(reqs, keep_results=True)
    srv.close()
    lat_p, total_p, res_p = serve.serve_plain(reqs, keep_results=True)
    assert sorted(res) == sorted(res_p) == list(range(6))
    for i in res:
        assert np.abs(res_p[i][1]).max() > 0
        assert np.array_equal(res[i][0], res_p[i][0]) and np.array_equal(res[i][1], res_p[i][1])


def test_server_budget_holds_with_solo_requests():
    """Admission by footprint counts the requests stepping on their own
    streams: with a budget that fits one request at a time and every request
    solo, the admitted total never exceeds the budget.  The background is
    synthesised in blocks of 64 steps; results equal the plain path's."""
    reqs = serve.request_mix(4, seed=11, rate=50.0, nx_range=(120, 300), nz_range=(50, 100),
                             steps_range=(300, 500), dx_choices=(100.0, 200.0))
    big = max(sum(serve.footprint_bytes(r, 64)) for r in reqs)
    budget = 1.2 * big
    print("case: serving 4 solo requests, budget %.3g GB (one request), blocks of 64 steps" % (budget / 1e9),
          flush=True)
    srv = serve.Server(hbm_budget_gb=budget / 1e9, chunk=50, solo_cells=1, block_steps=64)
    lat, total, res = srv.run(reqs, keep_results=True)
    srv.close()
    assert srv.peak_used <= srv.budget
    lat_p, total_p, res_p = serve.serve_plain(reqs, keep_results=True)
    for i in res_p:
        assert np.abs(res_p[i][1]).max() > 0
        assert np.array_equal(res[i][0], res_p[i][0]) and np.array_equal(res[i][1], res_p[i][1])


def test_server_serving_workload_requests_at_their_own_sizes():
    """The first four requests of the serving workload itself (seed 5, 4
    arrivals/s: 2338x772 to 6467x1696 cells, 3769 to 17352 steps, dx 50..200 m,
    253 G cell-updates), served with the default split (grids of >= 4 M
    padded cells on their own streams, the rest batched, background in
    blocks) and one at a time: every seismogram identical."""
    reqs = serve.request_mix(4, seed=5, rate=4.0)
    print("case: the serving workload's first 4 requests at their own sizes, server vs plain", flush=True)
    srv = serve.Server(hbm_budget_gb=200.0, chunk=100)
    lat, total, res = srv.run(reqs, keep_results=True)
    srv.close()
    print("server: %.2f s" % total, flush=True)
    lat_p, total_p, res_p = serve.serve_plain(reqs, keep_results=True)
    print("plain: %.2f s" % total_p, flush=True)
    assert sorted(res) == sorted(res_p) == list(range(4))
    for i in res:
        assert np.abs(res_p[i][1]).max() > 0
        assert np.array_equal(res[i][0], res_p[i][0]) and np.array_equal(res[i][1], res_p[i][1])
