#!/usr/bin/env python3
"""Benchmark of the FDFK2D time-stepping hot path on one MI355X.

A step is one iteration of the reference loop (src/FDFK2D.f90:465-532):
receiver recording, PML memories, the fused interior stencil, the boundary
frame (free surface, PML strips, hybrid ring).  The workload is the single
large solve (8001 x 2000 points, dx = dz = 50 m, dt = 0.003 s, 40-cell PML,
FD order 6, 401 receivers, seeded layered + 5% smoothed random model).
value = padded-grid cell-updates per second (NX*NZ per step, src/FDFK2D.f90
loop bounds), inputs resident in HBM.

Prints one JSON line with the roofline of the dominant kernel, k_step (the
interior and the frame of a step in one launch; algorithmic bytes from
step_alg_bytes, duration from HIP events on the solver's stream in a
separate profiled pass after the timed steps) and a CPU baseline: the FP32
CPU oracle (the reference's separate operator passes, OpenMP) timed on a
bounded sample of the 1001 x 400 workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILED_STEPS = 50
# HBM bytes per k_step launch on the 8001x2000 workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by a plain device copy
# (profiles/k_step_8001x2000_hbm_traffic.md)
MEASURED_TRAFFIC_BYTES = 726348914
# the same measurement with FP16 materials (profiles/k_step_fp16_8001x2000_hbm_traffic.md):
# FETCH_SIZE 235573 KiB x 2 + WRITE_SIZE 142402 KiB
MEASURED_TRAFFIC_BYTES_FP16 = 235573 * 2048 + 142402 * 1024


def step_alg_bytes(nx, nz, npml, norder, mat_bytes=12):
    """Algorithmic HBM bytes of one k_step launch and its cell count.

    Total-field cells (the interior, columns 1..nx x rows 2*norder..nz, and
    the surface rows 1..2*norder-1): 36 B each.  PML strip cells (left and
    right strips over rows 1..nz, bottom strip over all updated columns)
    read u^n (8 B) and the materials (12 B) and write u^{n+1} (8 B) like the
    others, but carry the split field instead of u^{n-1}: three parts x two
    components read at two levels and written at one (72 B), plus the
    first-derivative memories, x only in the side strips (8 B), x and z in
    the bottom strip (16 B)."""
    tf = nx * nz
    side = 2 * (npml - 1) * nz
    bottom = (nx + 2 * (npml - 1)) * (npml - 1)
    tfb = 24 + mat_bytes       # 36 B; 32 B with lam and mu read as FP16 (dt^2/rho stays FP32)
    return tfb * tf + (28 + 72 + 8) * side + (28 + 72 + 16) * bottom, tf + side + bottom


def gpu_bench(steps, warmup, workload, bg_block=0, nt=0, autotune=False, precision="fp32"):
    """nt: iterations the solver (and its background) is set up for, at
    least the ones run; bg_block > 0 synthesises the background in blocks of
    that many steps instead of storing all nt+1."""
    from fdfk2d_amd import configs, lib
    nt = max(nt, warmup + steps + PROFILED_STEPS + 1)
    cfg = configs.c2(nt=nt) if workload == "c2" else configs.c1(nt=nt)
    tuned = None
    if autotune:
        tuned = lib.autotune(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"],
                             cfg["c"], rx=cfg["rx"], rz=cfg["rz"])
    s = lib.Solver(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"],
                   nt, cfg["c"], rx=cfg["rx"], rz=cfg["rz"],
                   precision={"fp16": lib.FDFK_FP16, "bf16": lib.FDFK_BF16}.get(precision, lib.FDFK_FP32))
    bg_mode = "fk-gpu" if not bg_block else f"fk-gpu, synthesised in blocks of {bg_block} steps"
    try:
        t_bg = time.perf_counter()
        s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"],
                             cfg["rayp"], cfg["bazs"], block_steps=bg_block)
        s.sync()
        t_bg = time.perf_counter() - t_bg
    except lib.FdfkError:
        bg_mode = "zero"
        s.set_background([np.zeros(n, np.float32) for n in lib.bg_sizes(cfg["nx"], cfg["nz"], 5, nt)])
    device_bytes = s.device_bytes
    s.run(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    t1 = time.perf_counter()
    # kernel durations from a separate pass, so that the timed steps carry
    # no event packets
    s.profile(True)
    s.run(PROFILED_STEPS)
    s.sync()
    st = s.kernel_stats()
    NX, NZ = cfg["NX"], cfg["NZ"]
    alg_bytes, alg_cells = step_alg_bytes(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"],
                                          mat_bytes=8 if precision in ("fp16", "bf16") else 12)
    sx, sz = s.seismograms()
    finite = bool(np.isfinite(sx).all() and np.isfinite(sz).all())
    s.close()
    wall = t1 - t0
    return dict(cfg=cfg, NX=NX, NZ=NZ, wall=wall, stats=st, alg_bytes=alg_bytes, alg_cells=alg_cells,
                bg_mode=bg_mode, finite=finite, nt=nt, device_bytes=device_bytes, tuned=tuned,
                background_s=round(t_bg, 3) if bg_mode != "zero" else None)


def variant_bench(variant, steps=200, warmup=10, tiled=True):
    """The single large solve with the absorbing top layer ("top", P-SV), as
    SH motion ("sh", free surface) or both ("sh-top"), on the general-path
    solver: the deep interior on the tiled interior kernel and the cells
    around it on the per-cell kernels, or (tiled=False) every cell on the
    per-cell kernels.  value = padded cells (rows from izt) per second."""
    from fdfk2d_amd import configs, lib
    top, sh = variant in ("top", "sh-top"), variant.startswith("sh")
    nt = warmup + steps + 1
    cfg = configs.c2(nt=nt)
    c, NX, NZ, npml = cfg["c"], cfg["NX"], cfg["NZ"], cfg["npml"]
    if top:
        # npml rows above the surface repeating row 1 (the damped cells take
        # the material of the outermost undamped one)
        c3 = c.reshape(NX, NZ, 3)
        c = np.ascontiguousarray(np.concatenate([np.repeat(c3[:, :1], npml, axis=1), c3], axis=1)).ravel()
        NZ += npml
    old = os.environ.get("FDFK_GEN_TILED")
    os.environ["FDFK_GEN_TILED"] = "1" if tiled else "0"
    try:
        s = lib.Solver(cfg["nx"], cfg["nz"], npml, cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], nt, c,
                       rx=cfg["rx"], rz=cfg["rz"], pml_top=top, wave=lib.FDFK_SH if sh else lib.FDFK_PSV)
    finally:
        if old is None:
            del os.environ["FDFK_GEN_TILED"]
        else:
            os.environ["FDFK_GEN_TILED"] = old
    try:
        s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"] and not sh, cfg["f0"], cfg["ds"],
                             cfg["rayp"], cfg["bazs"])
        s.run(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.run(steps)
        s.sync()
        wall = time.perf_counter() - t0
        sx = s.seismograms()[0]
        finite = bool(np.isfinite(sx).all()) and float(np.abs(sx).max()) > 0
        dev = s.device_bytes
    finally:
        s.close()
    cells = NX * NZ
    return {"metric": "Gcell-updates/s", "value": round(cells * steps / wall / 1e9, 3), "unit": "Gcell-updates/s",
            "ms_per_step": round(wall / steps * 1e3, 4), "steps": steps, "warmup": warmup, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "single large solve 8001x2000, dx=dz=50 m, dt=0.003 s, 40-cell PML, order 6, "
                                   "401 receivers; " + {"top": "absorbing top layer, P-SV",
                                                        "sh": "SH motion, free surface",
                                                        "sh-top": "SH motion, absorbing top layer"}[variant],
                       "grid_padded": [NX, NZ], "interior": "tiled" if tiled else "per-cell kernels only",
                       "device_gib": round(dev / 2**30, 3), "finite": finite}}


def cpu_baseline(sample_steps):
    from fdfk2d_amd import configs, refstate
    from oracle import pyoracle as po
    cfg = configs.c1(nt=sample_steps)
    st = refstate.make_state(cfg["nx"], cfg["nz"], cfg["npml"], 3, cfg["dx"], cfg["dz"], cfg["dt"],
                             sample_steps, cfg["c"])
    st["bg"] = [np.zeros(n, np.float32) for n in po.bg_shapes(cfg["nx"], cfg["nz"], 5, sample_steps)]
    po.updatedispl(st, 0)  # first-touch / thread start-up outside the timed sample
    t0 = time.perf_counter()
    for it in range(1, sample_steps):
        po.updatedispl(st, it)
    dt = time.perf_counter() - t0
    n = sample_steps - 1
    cells = cfg["NX"] * cfg["NZ"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=round(cells * n / dt / 1e9, 6), unit="Gcell-updates/s", cores=threads,
                kind="port",
                sample=f"CPU oracle, {n} iterations of the 1001x400 (1041x420 padded) workload, "
                       f"{dt:.1f} s, OMP_NUM_THREADS={threads}")


def copy_calibration(mib=512, reps=5):
    """Plain device-to-device copies of a known size (torch's copy kernel):
    the reference point that converts the FETCH_SIZE / WRITE_SIZE counters of
    a profiler run into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "reps": reps}), flush=True)


def serve_bench(n, seed=5, rate=200.0, full=False, log=None):
    """Scaled continuous-batching serving run (the request mix of the serving
    workload with smaller grids and step counts so that both the batched and
    the one-at-a-time plain path finish in about a minute): p50/p99 latency
    and solves per hour for both, plus the equality of their seismograms.
    With full, the first n requests of the serving workload itself (4
    arrivals/s, grids 1000..8000 x 400..2000 log-uniform, 3000..20000 steps,
    dx 50/100/200 m)."""
    from fdfk2d_amd import serve
    if full:
        rate = 4.0
        reqs = serve.request_mix(n, seed=seed, rate=rate)
        shapes = "1000..8000 x 400..2000 log-uniform, 3000..20000 steps, dx 50/100/200 m"
    else:
        reqs = serve.request_mix(n, seed=seed, rate=rate, nx_range=(400, 2000), nz_range=(160, 500),
                                 steps_range=(600, 2500), dx_choices=(100.0, 200.0))
        shapes = "400..2000 x 160..500 log-uniform, 600..2500 steps, dx 100/200 m"
    # untimed pass: rocFFT compiles its kernels for each new FFT length at run
    # time; both timed paths then start with the same warm caches (for long
    # mixes the first 40 requests cover the lengths)
    serve.serve_plain(reqs if n <= 40 else [dict(r, arrival=0.0) for r in reqs[:40]])
    if log:
        log("serving: warm-up done")
    # seismograms kept for the equality check: all of them for short mixes,
    # every tenth request of long ones (a large request's record is 64 MB)
    keep = True if n <= 40 else set(range(0, n, 10))
    srv = serve.Server(hbm_budget_gb=200.0, chunk=100, log=log)
    lat_b, tot_b, res_b = srv.run(reqs, keep_results=keep)
    srv.close()
    if log:
        log("serving: batched path done, %.1f s" % tot_b)
    plain_s = {}
    lat_p, tot_p, res_p = serve.serve_plain(reqs, keep_results=keep, stats=plain_s, log=log)
    worst = 0.0
    for rid in res_p:
        for a_, b_ in zip(res_b[rid], res_p[rid]):
            worst = max(worst, serve.rel_l2(a_, b_))
    b = serve.summarize(lat_b, tot_b, n)
    p = serve.summarize(lat_p, tot_p, n)
    cells = sum((r["nx"] + 2 * r["npml"]) * (r["nz"] + r["npml"]) * r["nsteps"] for r in reqs)
    return {"metric": "solves_per_hour", "value": b["solves_per_hour"], "unit": "solves/h",
            "higher_is_better": True, "data": "synthetic",
            "config": {"workload": f"serving: {n} Poisson arrivals at {rate}/s (seed {seed}), grids "
                                   + shapes + ", HBM budget 200 GB", "cell_updates": cells},
            "batched": dict(b, prep_workers=srv.workers,
                            prep_s={k: round(v, 3) for k, v in srv.prep_parts.items()}),
            "plain": dict(p, phase_s={k: round(v, 3) for k, v in plain_s.items()}),
            "speedup_solves_per_hour": round(b["solves_per_hour"] / p["solves_per_hour"], 3),
            "max_rel_l2_batched_vs_plain": worst, "compared_requests": len(res_p)}


def full_solve_bench(nt=20000, block=512):
    """The whole single large solve, timed end to end on the GPU: FK
    background (synthesised in blocks), all nt+1 iterations with recording,
    seismograms back on the host.  Model set-up on the host is outside."""
    from fdfk2d_amd import configs, lib
    cfg = configs.c2(nt=nt)
    t0 = time.perf_counter()
    s = lib.Solver(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], nt,
                   cfg["c"], rx=cfg["rx"], rz=cfg["rz"])
    s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"], cfg["rayp"],
                         cfg["bazs"], block_steps=block)
    s.sync()
    t1 = time.perf_counter()
    s.run(nt + 1)
    s.sync()
    t2 = time.perf_counter()
    sx, sz = s.seismograms()
    t3 = time.perf_counter()
    gib = s.device_bytes / 2**30
    s.close()
    cells = cfg["NX"] * cfg["NZ"] * (nt + 1)
    # health of the whole record, from a second (untimed) run sampled at
    # thirds of the record: the side PML strips must stay far below the
    # surface record and not grow (a growing absorbing layer shows there)
    NX, NZ, npml, nx, nz = cfg["NX"], cfg["NZ"], cfg["npml"], cfg["nx"], cfg["nz"]
    s = lib.Solver(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], nt,
                   cfg["c"], rx=cfg["rx"], rz=cfg["rz"])
    s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"], cfg["rayp"],
                         cfg["bazs"], block_steps=block)
    sides, finite = [], bool(np.isfinite(sx).all() and np.isfinite(sz).all())
    for k in range(3):
        s.run((nt + 1) // 3 if k < 2 else nt + 1 - 2 * ((nt + 1) // 3))
        ux, uz = s.get_state()
        a = np.maximum(np.abs(ux.reshape(NX, NZ, 2)[:, :, 0]), np.abs(uz.reshape(NX, NZ, 2)[:, :, 0]))
        finite = finite and bool(np.isfinite(a).all())
        sides.append(float(max(a[:npml - 5, :nz].max(), a[npml + nx + 5:, :nz].max())))
    s.close()
    side = sides[-1]
    peak = float(max(np.abs(sx).max(), np.abs(sz).max()))
    late_n = int(round(20.0 / cfg["dt"]))
    late = float(max(np.abs(sx[:, -late_n:]).max(), np.abs(sz[:, -late_n:]).max()))
    return {"metric": "solves_per_hour", "value": round(3600.0 / (t3 - t0), 1), "unit": "solves/h",
            "higher_is_better": True, "data": "synthetic", "dtype": "f32",
            "config": {"workload": f"single large solve 8001x2000, {nt + 1} iterations, 401 receivers, "
                                   f"background in blocks of {block} steps", "device_gib": round(gib, 2)},
            "wall_s": round(t3 - t0, 3), "create_and_background_s": round(t1 - t0, 3),
            "loop_s": round(t2 - t1, 3), "loop_gcells": round(cells / (t2 - t1) / 1e9, 2),
            "seismograms_s": round(t3 - t2, 3),
            "health": {"finite": finite, "surface_peak": peak,
                       "surface_max_last_20s": late,
                       "side_pml_max_at_thirds": sides,
                       "end_s": round((nt + 1) * cfg["dt"], 3),
                       "ok": finite and side < 0.1 * peak and side < 2.0 * max(sides[:2])}}


def compat_bench(steps, warmup=3):
    """The drop-in path on the single large solve: updatedispl_ called per
    iteration with the reference driver's host arrays (src/FDFK2D.f90:516-522),
    which copies both levels of ux, uz and the PML arrays back every call
    because the driver reads them on the host.  The rate includes those PCIe
    copies; it is the price of the unchanged driver, not the kernel's."""
    from fdfk2d_amd import configs, lib, refstate
    nt = warmup + steps + 1
    cfg = configs.c2(nt=nt)
    st = refstate.make_state(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"],
                             nt, cfg["c"])
    st["bg"] = [np.zeros(n, np.float32) for n in lib.bg_sizes(cfg["nx"], cfg["nz"], 5, nt)]
    for it in range(warmup):
        lib.updatedispl(st, it)
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        lib.updatedispl(st, it)
    dt = time.perf_counter() - t0
    lib.lib().fdfk_compat_release()
    cells = cfg["NX"] * cfg["NZ"]
    return {"metric": "Gcell-updates/s", "value": round(cells * steps / dt / 1e9, 3), "unit": "Gcell-updates/s",
            "higher_is_better": True, "data": "synthetic", "dtype": "f32",
            "config": {"workload": "single large solve 8001x2000 through updatedispl_ (host arrays, "
                                   "both levels and the PML arrays copied back every call)",
                       "steps": steps},
            "ms_per_step": round(dt / steps * 1e3, 3)}


def batch_bench(njobs, seed=3, steps=1000):
    """Scaled batched workload: njobs independent solves of mixed sizes
    advanced together (one launch per kernel type per step) against the
    same solves run one after another with the single-job API."""
    from fdfk2d_amd import serve
    from fdfk2d_amd.lib import Batch, Solver
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(njobs):
        nx = int(rng.integers(500, 2001))
        nz = int(rng.integers(200, 501))
        reqs.append(dict(id=i, arrival=0.0, nx=nx, nz=nz, dx=100.0, nsteps=steps, src_is_p=bool(i % 2 == 0),
                         p_skm=float(rng.uniform(0.04, 0.08)), seed=int(rng.integers(1 << 30)), npml=20))

    def make(r):
        job = serve.build_job(r)
        s = Solver(r["nx"], r["nz"], r["npml"], 3, r["dx"], r["dx"], job["dt"], job["nt"], job["c"],
                   rx=job["rx"], rz=job["rz"])
        s.compute_background(job["fk"], job["fk"], r["src_is_p"], job["f0"], job["ds"], job["rayp"], job["bazs"])
        return s

    sol = [make(r) for r in reqs]
    for s in sol:
        s.sync()
    t0 = time.perf_counter()
    for s in sol:
        s.run(s.nt + 1)
    for s in sol:
        s.sync()
    t_seq = time.perf_counter() - t0
    for s in sol:
        s.close()
    def batched(tune):
        sol = [make(r) for r in reqs]
        B = Batch(3)
        B.autotune(tune)
        for s in sol:
            B.add(s)
        t0 = time.perf_counter()
        while B.active:
            B.run(100)
            width = B.tile_width
        B.sync()
        t = time.perf_counter() - t0
        for s in sol:
            s.close()
        B.close()
        return t, width

    t_bat, w_model = batched(False)
    t_tuned, w_tuned = batched(True)     # the tuner's exploration steps are inside the timed region
    cells = sum((r["nx"] + 40) * (r["nz"] + 20) * r["nsteps"] for r in reqs)
    return {"metric": "Gcell-updates/s", "value": round(cells / t_bat / 1e9, 3), "unit": "Gcell-updates/s",
            "higher_is_better": True, "data": "synthetic",
            "config": {"workload": f"batch of {njobs} solves (seed {seed}), nx 500..2000, nz 200..500, "
                                   f"dx 100 m, {steps} steps each"},
            "batched_s": round(t_bat, 3), "sequential_s": round(t_seq, 3),
            "sequential_gcells": round(cells / t_seq / 1e9, 3), "speedup": round(t_seq / t_bat, 3),
            "tile_width_model": w_model,
            "autotuned": {"batched_s": round(t_tuned, 3), "tile_width": w_tuned,
                          "speedup": round(t_seq / t_tuned, 3)}}


def c3_bench(njobs=64, block=512, log=lambda *a: None):
    """The batched workload (fdfk2d_amd.configs.c3_requests: 64 simulations,
    nx 2000..8000, nz 800..2000 at 50 m, 10000 iterations, P/S, p 0.04..0.08
    s/km, all submitted at once): every job one at a time with the
    single-job API, against the continuous-batching server (graph-replayed
    batched steps) with the 200 GB HBM budget; the 64 jobs do not fit at
    once (each keeps about 3 GB of FP64 traces for its background
    synthesis), so they join the batch as memory frees.  Twice: every job in
    the batch, and the server's default split (grids of 4 M cells or more
    step on their own streams beside the batch).  Times include the FK
    backgrounds (in blocks of `block` steps) and the model set-up; every
    eighth job's seismograms must be identical on all paths."""
    from fdfk2d_amd import configs, serve
    from fdfk2d_amd.lib import trim_memory
    reqs = configs.c3_requests(njobs)
    keep = set(range(0, njobs, 8))
    # warm-up: rocFFT compiles a kernel per FFT length at run time
    serve.serve_plain(reqs[:2], block_steps=block)
    lat_p, t_plain, res_p = serve.serve_plain(reqs, keep_results=keep, block_steps=block, log=log)
    log("c3 plain: %.1f s" % t_plain)
    trim_memory()
    out = {}
    for name, solo in (("batched", 10 ** 12), ("batched_with_solo_streams", 4_000_000)):
        srv = serve.Server(hbm_budget_gb=200.0, chunk=100, solo_cells=solo, block_steps=block, log=log)
        lat_b, t_b, res_b = srv.run(reqs, keep_results=keep)
        srv.close()
        trim_memory()
        log("c3 %s: %.1f s" % (name, t_b))
        same = all(np.array_equal(a_, b_) for rid in res_p for a_, b_ in zip(res_b[rid], res_p[rid]))
        out[name] = dict(serve.summarize(lat_b, t_b, njobs), peak_admitted_gb=round(srv.peak_used / 1e9, 1),
                         speedup_solves_per_hour=round(t_plain / t_b, 3), seismograms_identical=same)
    cells = sum((r["nx"] + 2 * r["npml"]) * (r["nz"] + r["npml"]) * r["nsteps"] for r in reqs)
    best = max(out.values(), key=lambda d: d["solves_per_hour"])
    return {"metric": "solves_per_hour", "value": best["solves_per_hour"], "unit": "solves/h",
            "higher_is_better": True, "data": "synthetic", "dtype": "f32",
            "config": {"workload": f"batched: {njobs} simulations (seed 3), nx 2000..8000, nz 800..2000, dx 50 m, "
                                   "10000 iterations, 40-cell PML, P/S alternating, p 0.04..0.08 s/km",
                       "cell_updates": cells, "background": f"FK on the GPU in blocks of {block} steps",
                       "hbm_budget_gb": 200.0},
            "plain": dict(serve.summarize(lat_p, t_plain, njobs), gcells=round(cells / t_plain / 1e9, 2)),
            **out}


def c4_bench(nsteps=8000, rf_every=20, block=512, log=lambda *a: None):
    """The storage-precision workload: the single large solve (8001 x 2000)
    for nsteps iterations in FP64 (the general per-cell kernels in double,
    the reference), FP32 (the tiled kernels) and FP32 with FP16 materials;
    rel-L2 of every receiver trace and of the P receiver functions of every
    rf_every-th receiver against the FP64 run (RF spike picks of the FP64
    run replayed, so that the comparison measures the traces' precision and
    not the discrete decisions of the deconvolution)."""
    from fdfk2d_amd import configs, lib, rf
    cfg = configs.c2(nt=nsteps)
    NX, NZ = cfg["NX"], cfg["NZ"]
    runs = {}
    for name, prec in (("fp64", lib.FDFK_FP64), ("fp32", lib.FDFK_FP32), ("fp16_materials", lib.FDFK_FP16),
                       ("bf16_materials", lib.FDFK_BF16)):
        log(f"c4: {name}, {nsteps + 1} iterations")
        s = lib.Solver(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], nsteps,
                       cfg["c"], rx=cfg["rx"], rz=cfg["rz"], precision=prec)
        try:
            s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"],
                                 cfg["rayp"], cfg["bazs"], block_steps=block)
            s.sync()
            t0 = time.perf_counter()
            s.run(nsteps + 1)
            s.sync()
            t = time.perf_counter() - t0
            sx, sz = s.seismograms()
            dev = s.device_bytes
        finally:
            s.close()
        runs[name] = (np.asarray(sx, np.float64), np.asarray(sz, np.float64), t, dev)
        log(f"c4: {name} {t:.2f} s ({NX * NZ * (nsteps + 1) / t / 1e9:.2f} Gcell/s)")
    rx64, rz64 = runs["fp64"][0], runs["fp64"][1]
    ref_all = np.concatenate([rx64.ravel(), rz64.ravel()])
    # receivers the incident wave has reached within the run (it crosses the
    # profile from one end at the apparent velocity 1/p)
    peak = np.abs(rz64).max()
    sel = [i for i in range(0, rx64.shape[0], rf_every) if np.abs(rz64[i]).max() > 1e-2 * peak]
    dt = cfg["dt"]
    picks, rf64 = {}, {}
    for i in sel:
        rec = []
        rf64[i] = rf.prf(rx64[i], rz64[i], dt, record=rec)[0]
        picks[i] = rec
    out = dict(workload="single large solve 8001x2000, dx=dz=50 m, dt=0.003 s, 40-cell PML, order 6, 401 receivers",
               iterations=nsteps + 1, reference="FDFK_FP64 (general per-cell kernels in double)",
               rf="P-RF (iterative deconvolution, Gaussian 8, FP64 run's spike picks replayed) of "
                  f"{len(sel)} receivers (every {rf_every}th the incident wave reached)", modes={})
    for name, (sx, sz, t, dev) in runs.items():
        m = dict(seconds=round(t, 3), gcell_s=round(NX * NZ * (nsteps + 1) / t / 1e9, 3),
                 device_gib=round(dev / 2 ** 30, 3))
        if name != "fp64":
            err = np.concatenate([sx.ravel(), sz.ravel()]) - ref_all
            m["rel_l2_traces"] = float(np.linalg.norm(err) / np.linalg.norm(ref_all))
            m["rel_l2_x"] = float(np.linalg.norm(sx - rx64) / np.linalg.norm(rx64))
            m["rel_l2_z"] = float(np.linalg.norm(sz - rz64) / np.linalg.norm(rz64))
            num = den = 0.0
            per = {}
            big = max(np.linalg.norm(rf64[i]) for i in sel)
            for i in sel:
                r = rf.prf(sx[i], sz[i], dt, picks=picks[i])[0]
                num += float(np.sum((r - rf64[i]) ** 2))
                den += float(np.sum(rf64[i] ** 2))
                n64 = float(np.linalg.norm(rf64[i]))
                per[str(i + 1)] = dict(rel_l2=float(np.linalg.norm(r - rf64[i]) / max(n64, 1e-300)),
                                       rf_norm_over_largest=n64 / big)
            m["rel_l2_prf"] = float(np.sqrt(num / den))
            m["rel_l2_prf_per_receiver"] = per
        out["modes"][name] = m
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--cpu-steps", type=int, default=201)
    ap.add_argument("--bg-block", type=int, default=0, metavar="B",
                    help="synthesise the background in blocks of B steps (default: store all)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="fp16 / bf16: lam and mu stored as per-row profiles plus FP16 / BF16 residuals with "
                         "power-of-two scales (arithmetic stays FP32)")
    ap.add_argument("--autotune", action="store_true",
                    help="time interior tile widths first (fdfk_autotune) and run with the fastest")
    ap.add_argument("--nt", type=int, default=0,
                    help="iterations the solve is set up for (default: just the ones run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run the copy kernel used to calibrate HBM byte counters")
    ap.add_argument("--serve", type=int, default=0, metavar="N",
                    help="serving measurement with N requests instead of the default line")
    ap.add_argument("--serve-full", action="store_true",
                    help="with --serve N: the first N requests of the serving workload at its own sizes")
    ap.add_argument("--full-solve", action="store_true",
                    help="time the whole 20001-iteration single large solve end to end")
    ap.add_argument("--compat", type=int, default=0, metavar="N",
                    help="N iterations through the drop-in updatedispl_ (PCIe-inclusive rate)")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="batched-execution measurement with N jobs instead of the default line")
    ap.add_argument("--c3", type=int, default=0, metavar="N",
                    help="the batched workload with N of its jobs (64: all), plain vs batched")
    ap.add_argument("--c4", type=int, default=0, metavar="N",
                    help="storage precision on the single large solve: N iterations in FP64, FP32 and FP16 "
                         "materials, rel-L2 of traces and P-RFs against FP64")
    ap.add_argument("--no-solve", action="store_true",
                    help="default line without the whole-solve and serving objects")
    ap.add_argument("--variant", default="", choices=["", "top", "sh", "sh-top"],
                    help="the single large solve with the absorbing top layer, as SH motion, or both "
                         "(general-path solver) instead of the default line")
    ap.add_argument("--per-cell", action="store_true",
                    help="with --variant: every cell on the per-cell kernels (no tiled interior)")
    a = ap.parse_args()
    if a.variant:
        print(json.dumps(variant_bench(a.variant, a.steps, a.warmup, tiled=not a.per_cell)), flush=True)
        return
    if a.copy_calib:
        copy_calibration()
        return
    if a.serve:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        print(json.dumps(serve_bench(a.serve, full=a.serve_full, log=log)), flush=True)
        return
    if a.batch:
        print(json.dumps(batch_bench(a.batch)), flush=True)
        return
    if a.c4:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        print(json.dumps(c4_bench(a.c4, log=log)), flush=True)
        return
    if a.c3:
        log = lambda m: print(m, file=sys.stderr, flush=True)
        print(json.dumps(c3_bench(a.c3, log=log)), flush=True)
        return
    if a.compat:
        print(json.dumps(compat_bench(a.compat)), flush=True)
        return
    if a.full_solve:
        print(json.dumps(full_solve_bench()), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the path would shard as x-slabs with one halo exchange "
              "per step, which is not built, so N > 1 is not run", file=sys.stderr)
        sys.exit(2)
    r = gpu_bench(a.steps, a.warmup, a.workload, a.bg_block, a.nt, a.autotune, a.precision)
    cells = r["NX"] * r["NZ"]
    value = cells * a.steps / r["wall"] / 1e9
    ms = r["wall"] / a.steps * 1e3
    st = r["stats"]
    k_ms = st["interior_ms"] / max(st["interior_n"], 1)
    cfg = r["cfg"]
    # FDFK_FUSED=0 times the interior launch alone (k_interior2): its own
    # algorithmic bytes, the interior tiles' total-field cells at 36 B
    # (32 B with FP16 materials); the measured traffic is the fused launch's
    fused = os.environ.get("FDFK_FUSED", "1") != "0"
    if fused:
        k_name, k_bytes, k_cells = "k_step", r["alg_bytes"], r["alg_cells"]
    else:
        k_cells = cfg["nx"] * (cfg["nz"] - 2 * cfg["norder"] + 1)
        k_name, k_bytes = "k_interior2", (32 if a.precision in ("fp16", "bf16") else 36) * k_cells
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    full_steps = 20000
    out = {
        "metric": "Gcell-updates/s",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single large solve 8001x2000, dx=dz=50 m, dt=0.003 s, 40-cell PML, order 6, "
                        "401 receivers" if a.workload == "c2" else "1001x400, dx=200 m, dt=0.01 s",
            "grid_padded": [r["NX"], r["NZ"]],
            "background": r["bg_mode"],
            "set_up_for_iterations": r["nt"],
            "device_gib": round(r["device_bytes"] / 2**30, 3),
            "background_s": r["background_s"],
            "tile_autotune": r["tuned"],
            "material_storage": ("%s lam, mu residuals on per-row fp32 profiles (power-of-two scales); fp32 dt^2/rho"
                                 % a.precision if a.precision in ("fp16", "bf16") else "fp32"),
            "solves_per_hour_at_20000_steps_loop_only": round(3600.0 / (ms * 1e-3 * full_steps), 3),
            "finite": r["finite"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # the BF16 mode reads the same bytes as the FP16 one
            "traffic": (None if a.workload != "c2" or not fused else
                        MEASURED_TRAFFIC_BYTES_FP16 if a.precision in ("fp16", "bf16") else MEASURED_TRAFFIC_BYTES),
            "traffic_source": "rocprofv3 FETCH_SIZE+WRITE_SIZE per launch, copy-calibrated "
                              + ("(profiles/k_step_fp16_8001x2000_hbm_traffic.md)" if a.precision in ("fp16", "bf16")
                                 else "(profiles/k_step_8001x2000_hbm_traffic.md)"),
            "kernel": k_name,
            "kernel_avg_ms": round(k_ms, 5),
            "alg_bytes_per_launch": k_bytes,
            "cells_per_launch": k_cells,
        },
    }
    if a.workload == "c2" and not a.no_solve and a.precision == "fp32":
        # the rest of the metric, measured rather than projected: one whole
        # solve end to end (its health covers the full 60 s record), and
        # solves/h with p50/p99 latency of continuous batching against the
        # plain one-at-a-time path on a bounded request mix
        fs = full_solve_bench()
        out["full_solve"] = {k: fs[k] for k in ("value", "unit", "wall_s", "create_and_background_s", "loop_s",
                                                "loop_gcells", "health")}
        out["config"]["finite"] = fs["health"]["finite"]
        sv = serve_bench(24)
        sf = serve_bench(12, full=True)
        out["serving"] = {
            name: {"workload": r["config"]["workload"], "solves_per_hour": r["value"],
                   "batched": {k: r["batched"][k] for k in ("wall_s", "solves_per_hour", "p50_s", "p99_s")},
                   "plain": {k: r["plain"][k] for k in ("wall_s", "solves_per_hour", "p50_s", "p99_s")},
                   "speedup_solves_per_hour": r["speedup_solves_per_hour"],
                   "max_rel_l2_batched_vs_plain": r["max_rel_l2_batched_vs_plain"]}
            for name, r in (("scaled_mix_24", sv), ("serving_workload_first_12", sf))}
        # the same solve on the general-path solver: absorbing top layer
        # (P-SV) and SH motion (both top boundaries)
        out["variants"] = {}
        for v in ("top", "sh", "sh-top"):
            vr = variant_bench(v, steps=100, warmup=5)
            out["variants"][v] = {"value": vr["value"], "unit": vr["unit"], "ms_per_step": vr["ms_per_step"],
                                  "grid_padded": vr["config"]["grid_padded"], "finite": vr["config"]["finite"]}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
