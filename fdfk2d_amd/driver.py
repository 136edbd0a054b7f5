"""FDFK2D driver on the GPU path: the command line and loop of
src/FDFK2D.f90:32-588 with the time step and the FK background on the
MI355X (C ABI in include/fdfk.h).

    python -m fdfk2d_amd.driver ./input inpar.dat ./seismograms seis

Reads the unchanged input files, writes <outpath>/<prefix>x.su and z.su
exactly as the reference (radial projection by cos(baz) included), and, when
FD_model.dat asks for them, the wavefield snapshots in ./snapshots/.

With --all-events every event line of a multi-event Source.dat is simulated
(the reference reads only the first, src/Input.f90:142-146), the events
advanced together as batches of resident solvers, and event k is written to
<outpath>/<prefix>_<k>x.su and z.su.
"""
import os
import sys
import time

import numpy as np

from . import geometry, io, refstate
from .lib import Batch, Solver


def write_snapshot(s, it, nx, nz, npml, dz, snapdir="snapshots", izt=1):
    """Snapshot at the top of iteration it (src/FDFK2D.f90:467-480): level 2
    of ux(1:2, :, :) and uz, rows 1..nz of columns 1..nx, one SU trace per
    column, header word 58 = nz and 59 = dz, in ./snapshots/<it>ux.su and
    <it>uz.su (five-digit iteration, I5.5).  Rows start at izt."""
    os.makedirs(snapdir, exist_ok=True)
    ux, uz = s.get_state()
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    for name, u in (("ux", ux), ("uz", uz)):
        lev2 = u.reshape(NX, NZ, 2)[npml: npml + nx, 1 - izt: 1 - izt + nz, 1]
        io.write_su(os.path.join(snapdir, "%05d%s.su" % (it, name)), lev2, int(dz))


def _prepare(job, src, log):
    """Solver with its FK background for one event (src/FDFK2D.f90:311-434)."""
    fd = job["fd"]
    nx, nz, npml, norder = job["nx"], job["nz"], fd["npml"], fd["norder"]
    dt = fd["dt"]
    nt = int(np.rint(fd["tmax"] / dt))                       # FDFK2D.f90:311-315
    tstp = fd["tstp"] if fd["tstp"] >= dt else dt
    nstp = int(np.rint(tstp / dt))
    nsis = nt // nstp
    tstp = np.float32(nstp * dt)
    dts = refstate.stable_dt(norder, fd["dx"], fd["dz"], job["c"])
    if dt > dts:                                             # Submain.f90:149-155
        raise ValueError(f"The dt is too large to stablize the simulation; stable dt = {dts:.7f}")
    bazs = geometry.column_bazs(nx, npml, fd["x0"], fd["dx"], fd["lat_org"], fd["lon_org"],
                                fd["az_org"], src["evla"], src["evlo"])
    if abs(np.cos(bazs.sum() / (nx + 2 * npml + 1))) < 0.5 * np.sqrt(3.0):
        log("WARNING: the profile is more than 30 deg off the back-azimuth (FDFK2D.f90:388-401)")
    rayp = src["rayp"]
    if abs(rayp) < 1e-9:
        rayp = 1e-9 * np.sign(rayp if rayp != 0 else 1.0)
    s = Solver(nx, nz, npml, norder, fd["dx"], fd["dz"], dt, nt, job["c"], x0=fd["x0"], z0=fd["z0"],
               nstp=nstp, rx=job["rx"], rz=job["rz"], R0=1e-3, npower=2,
               vpml=refstate.vp_max(job["c"]), pml_top=fd["pml_top"])
    try:
        s.compute_background(job["fk_left"], job["fk_right"], src["src_is_p"], src["f0"], src["ds"],
                             rayp, bazs)
        s.sync()
    except Exception:
        s.close()
        raise
    return s, dict(nt=nt, nsis=nsis, tstp=float(tstp))


def _finish(job, src, s):
    """Seismograms with the radial projection by cos(baz) (FDFK2D.f90:537-543)."""
    fd = job["fd"]
    sx, sz = s.seismograms()
    br = geometry.backazimuth_grids(fd["lat_org"], fd["lon_org"], fd["az_org"], job["rx"],
                                    src["evla"], src["evlo"]) - fd["az_org"] * np.pi / 180.0
    sx = (sx.astype(np.float64) * np.cos(br)[:, None]).astype(np.float32)
    return sx, sz


def simulate(job, log=print):
    fd, src = job["fd"], job["src"]
    nx, nz, npml = job["nx"], job["nz"], fd["npml"]
    t0 = time.perf_counter()
    s, info = _prepare(job, src, log)
    nt = info["nt"]
    t1 = time.perf_counter()
    log(f"background (FK, GPU FP64): {t1 - t0:.2f} s")
    try:
        if fd.get("outsnap") and fd.get("nstep", 0) > 0:
            done = 0
            for it in range(fd["nstep"], nt + 1, fd["nstep"]):
                s.run(it - done)
                done = it
                write_snapshot(s, it, nx, nz, npml, fd["dz"], izt=job.get("izt", 1))
            s.run(nt + 1 - done)
        else:
            s.run(nt + 1)
        s.sync()
        t2 = time.perf_counter()
        cells = (nx + 2 * npml) * (nz + npml) * (nt + 1)
        log(f"time loop: {nt + 1} iterations in {t2 - t1:.2f} s ({cells / (t2 - t1) / 1e9:.2f} Gcell/s)")
        sx, sz = _finish(job, src, s)
    finally:
        s.close()
    return sx, sz, info


def simulate_events(job, events, hbm_budget_gb=200.0, log=print):
    """Every event of a multi-event Source.dat on the same model, advanced
    together as batches of resident solvers (fdfk_batch).  Events are taken
    in file order, as many per batch as the HBM budget holds; each result is
    bit-identical to the event run alone.  Returns [(sx, sz)], info."""
    fd = job["fd"]
    out, info, k = [], None, 0
    t0 = time.perf_counter()
    if fd["pml_top"]:
        # solvers with an absorbing top layer run one at a time (not batched)
        for ev in events:
            s, info = _prepare(job, ev, log)
            try:
                s.run(info["nt"] + 1)
                s.sync()
                out.append(_finish(job, ev, s))
            finally:
                s.close()
        return out, info
    while k < len(events):
        batch = Batch(fd["norder"])
        members = []
        try:
            while k < len(events):
                s, info = _prepare(job, events[k], log)
                members.append((events[k], s))
                batch.add(s)
                k += 1
                if (len(members) + 1) * s.device_bytes > hbm_budget_gb * 1e9:
                    break
            batch.run(info["nt"] + 1)
            batch.sync()
            for ev, s in members:
                out.append(_finish(job, ev, s))
        finally:
            batch.close()
            for _, s in members:
                s.close()
        log(f"events {k - len(members) + 1}..{k} of {len(events)} done as one batch "
            f"({time.perf_counter() - t0:.2f} s so far)")
    return out, info


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    all_events = "--all-events" in argv
    argv = [a for a in argv if a != "--all-events"]
    if len(argv) < 3:
        print("FDFK2D (MI355X path)\n./driver ./inpath inpar.dat ./outpath seism [--all-events]")
        return 0
    prefix, finpar, outpath = argv[:3]
    fsis = argv[3] if len(argv) > 3 else "seis"
    job = io.read_job(prefix, finpar)
    # *_without_pml.su copies of the models (Input.f90:305-341)
    for key, arr in (("vp_model", job["vp"]), ("vs_model", job["vs"])):
        name = job["names"][key]
        r1 = 1 - job.get("izt", 1)   # the row of iz = 1
        io.write_su(os.path.join(prefix, name[:-3] + "_without_pml.su"),
                    arr[job["fd"]["npml"]: job["fd"]["npml"] + job["nx"], r1: r1 + job["nz"]], 1000)
    os.makedirs(outpath, exist_ok=True)
    if all_events:
        # every line of Source.dat, <fsis>_<k>x.su and z.su for event k (1-based, file order)
        events = io.read_events(os.path.join(prefix, job["names"]["source"]))
        if job["fd"].get("outsnap"):
            print("snapshots are written for single-event runs only")
        results, info = simulate_events(job, events)
        dh = int(np.float32(info["tstp"]) * np.float32(1e6))
        for k, (sx, sz) in enumerate(results, 1):
            io.write_su(os.path.join(outpath, f"{fsis}_{k}x.su"), sx, dh)
            io.write_su(os.path.join(outpath, f"{fsis}_{k}z.su"), sz, dh)
        return 0
    sx, sz, info = simulate(job)
    dh = int(np.float32(info["tstp"]) * np.float32(1e6))
    io.write_su(os.path.join(outpath, fsis + "x.su"), sx, dh)
    io.write_su(os.path.join(outpath, fsis + "z.su"), sz, dh)
    return 0


if __name__ == "__main__":
    sys.exit(main())
