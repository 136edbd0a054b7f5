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

Three components (--3c, optionally --spol DEG).  The reference keeps the
in-plane horizontal motion u_x rotated by the receiver's relative
back-azimuth phi = baz - az_org and projected on the profile, x = u_x cos phi
(src/FDFK2D.f90:537-543).  With --3c the same rotation is applied to the
whole horizontal vector (u_x, u_y) and both of its components are written:

    x = u_x cos phi - u_y sin phi     (<prefix>x.su, the reference's file)
    y = u_x sin phi + u_y cos phi     (<prefix>y.su, 90 deg clockwise of x)

u_y is the transverse (SH) motion.  An incident S wave polarised at psi
degrees from SV toward SH (--spol, default 0: the reference's pure SV) is
simulated as an SV solve of strength ds cos psi and an SH solve (FDFK_SH) of
strength ds sin psi; in the 2-D model the two motions do not couple, and an
incident P wave has no SH part.  With psi = 0 x.su and z.su are the
reference's files bit for bit and y.su is u_x sin phi.
"""
import os
import sys
import time

import numpy as np

from . import geometry, io, refstate
from .lib import FDFK_PSV, FDFK_SH, Batch, Solver


def write_snapshot(s, it, nx, nz, npml, dz, snapdir="snapshots", izt=1):
    """Snapshot at the top of iteration it (src/FDFK2D.f90:467-480): level 2
    of ux(1:2, :, :) and uz, rows 1..nz of columns 1..nx, one SU trace per
    column, header word 58 = nz and 59 = dz, in ./snapshots/<it>ux.su and
    <it>uz.su (five-digit iteration, I5.5).  Rows start at izt.  An SH
    solver writes its one component to <it>uy.su."""
    os.makedirs(snapdir, exist_ok=True)
    ux, uz = s.get_state()
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    fields = (("uy", ux),) if s.wave == FDFK_SH else (("ux", ux), ("uz", uz))
    for name, u in fields:
        lev2 = u.reshape(NX, NZ, 2)[npml: npml + nx, 1 - izt: 1 - izt + nz, 1]
        io.write_su(os.path.join(snapdir, "%05d%s.su" % (it, name)), lev2, int(dz))


def rotate_horizontal(sx, sy, phi):
    """(x, y) = R(phi) (u_x, u_y) per receiver, in FP64, rounded to FP32;
    sy None: no transverse motion, x = u_x cos phi exactly as
    src/FDFK2D.f90:540-542 computes it."""
    c, s = np.cos(phi)[:, None], np.sin(phi)[:, None]
    ux = sx.astype(np.float64)
    if sy is None:
        return (ux * c).astype(np.float32), (ux * s).astype(np.float32)
    uy = sy.astype(np.float64)
    return (ux * c - uy * s).astype(np.float32), (ux * s + uy * c).astype(np.float32)


def split_strength(src, spol_deg):
    """Strengths of the SV and SH solves of one event: (ds_sv, ds_sh); an
    incident P wave has no SH part."""
    if src["src_is_p"] or spol_deg == 0.0:
        return src["ds"], 0.0
    psi = np.deg2rad(spol_deg)
    return src["ds"] * float(np.cos(psi)), src["ds"] * float(np.sin(psi))


def _prepare(job, src, log, wave=FDFK_PSV):
    """Solver with its FK background for one event (src/FDFK2D.f90:311-434);
    wave FDFK_SH: the SH solver of an S event (src["ds"] its strength)."""
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
               vpml=refstate.vp_max(job["c"]), pml_top=fd["pml_top"], wave=wave)
    try:
        s.compute_background(job["fk_left"], job["fk_right"], src["src_is_p"] and wave == FDFK_PSV, src["f0"],
                             src["ds"], rayp, bazs)
        s.sync()
    except Exception:
        s.close()
        raise
    return s, dict(nt=nt, nsis=nsis, tstp=float(tstp))


def _receiver_phi(job, src):
    """Relative back-azimuth of each receiver (src/FDFK2D.f90:537-539)."""
    fd = job["fd"]
    return geometry.backazimuth_grids(fd["lat_org"], fd["lon_org"], fd["az_org"], job["rx"],
                                      src["evla"], src["evlo"]) - fd["az_org"] * np.pi / 180.0


def _finish(job, src, s, sy=None, three=False):
    """Seismograms with the radial projection by cos(baz) (FDFK2D.f90:537-543):
    (x, z); three components: (x, y, z), sy the SH solve's u_y or None."""
    sx, sz = s.seismograms()
    x, y = rotate_horizontal(sx, sy, _receiver_phi(job, src))
    return (x, y, sz) if three else (x, sz)


def _sh_part(job, src, spol, log):
    """u_y seismograms of an event's SH solve, None when it has no SH part."""
    _, ds_sh = split_strength(src, spol)
    if ds_sh == 0.0:
        return None
    s, info = _prepare(job, dict(src, ds=ds_sh), log, wave=FDFK_SH)
    try:
        s.run(info["nt"] + 1)
        s.sync()
        return s.seismograms()[0]
    finally:
        s.close()


def _advance(job, solvers, nt, snap):
    """Queue every iteration of the solvers (each on its own stream, so they
    overlap on the GPU), with the snapshots at every nstep iterations."""
    fd = job["fd"]
    if snap:
        done = 0
        for it in range(fd["nstep"], nt + 1, fd["nstep"]):
            for s in solvers:
                s.run(it - done)
            done = it
            for s in solvers:
                write_snapshot(s, it, job["nx"], job["nz"], fd["npml"], fd["dz"], izt=job.get("izt", 1))
        for s in solvers:
            s.run(nt + 1 - done)
    else:
        for s in solvers:
            s.run(nt + 1)
    for s in solvers:
        s.sync()


def simulate(job, log=print, three=False, spol=0.0):
    """One event: (x, z) seismograms and the run's info; three components
    (three=True): (x, y, z).  spol: polarisation of an incident S wave in
    degrees from SV toward SH (the SH part runs on its own solver)."""
    fd, src = job["fd"], job["src"]
    nx, nz, npml = job["nx"], job["nz"], fd["npml"]
    ds_sv, ds_sh = split_strength(src, spol)
    t0 = time.perf_counter()
    s, info = _prepare(job, dict(src, ds=ds_sv), log)
    s_sh = None
    nt = info["nt"]
    try:
        if ds_sh != 0.0:
            s_sh, _ = _prepare(job, dict(src, ds=ds_sh), log, wave=FDFK_SH)
        t1 = time.perf_counter()
        log(f"background (FK, GPU FP64): {t1 - t0:.2f} s")
        _advance(job, [s] + ([s_sh] if s_sh else []), nt, fd.get("outsnap") and fd.get("nstep", 0) > 0)
        t2 = time.perf_counter()
        cells = (nx + 2 * npml) * (nz + npml - job.get("izt", 1) + 1) * (nt + 1) * (2 if s_sh else 1)
        log(f"time loop: {nt + 1} iterations in {t2 - t1:.2f} s ({cells / (t2 - t1) / 1e9:.2f} Gcell/s)")
        out = _finish(job, src, s, s_sh.seismograms()[0] if s_sh else None, three)
    finally:
        s.close()
        if s_sh is not None:
            s_sh.close()
    return out + (info,)


def simulate_events(job, events, hbm_budget_gb=200.0, log=print, three=False, spol=0.0):
    """Every event of a multi-event Source.dat on the same model, advanced
    together as batches of resident solvers (fdfk_batch).  Events are taken
    in file order, as many per batch as the HBM budget holds; each result is
    bit-identical to the event run alone.  The SH solves of S events
    (spol != 0) run one at a time after their batch.  Returns [(x, z)] or,
    three components, [(x, y, z)], and info."""
    fd = job["fd"]
    out, info, k = [], None, 0
    t0 = time.perf_counter()
    if fd["pml_top"]:
        # solvers with an absorbing top layer run one at a time (not batched)
        for ev in events:
            s, info = _prepare(job, dict(ev, ds=split_strength(ev, spol)[0]), log)
            try:
                s.run(info["nt"] + 1)
                s.sync()
                out.append(_finish(job, ev, s, _sh_part(job, ev, spol, log), three))
            finally:
                s.close()
        return out, info
    while k < len(events):
        batch = Batch(fd["norder"])
        members = []
        try:
            while k < len(events):
                s, info = _prepare(job, dict(events[k], ds=split_strength(events[k], spol)[0]), log)
                members.append((events[k], s))
                batch.add(s)
                k += 1
                if (len(members) + 1) * s.device_bytes > hbm_budget_gb * 1e9:
                    break
            batch.run(info["nt"] + 1)
            batch.sync()
            for ev, s in members:
                out.append(_finish(job, ev, s, _sh_part(job, ev, spol, log), three))
        finally:
            batch.close()
            for _, s in members:
                s.close()
        log(f"events {k - len(members) + 1}..{k} of {len(events)} done as one batch "
            f"({time.perf_counter() - t0:.2f} s so far)")
    return out, info


def _options(argv):
    """--all-events, --3c and --spol DEG (or --spol=DEG) out of argv."""
    opts = dict(all_events=False, three=False, spol=0.0)
    rest, i = [], 0
    while i < len(argv):
        a = argv[i]
        if a == "--all-events":
            opts["all_events"] = True
        elif a == "--3c":
            opts["three"] = True
        elif a.startswith("--spol="):
            opts["spol"] = float(a.split("=", 1)[1])
        elif a == "--spol" and i + 1 < len(argv):
            opts["spol"] = float(argv[i + 1])
            i += 1
        else:
            rest.append(a)
        i += 1
    return opts, rest


def main(argv=None):
    opts, argv = _options(list(sys.argv[1:] if argv is None else argv))
    if len(argv) < 3:
        print("FDFK2D (MI355X path)\n./driver ./inpath inpar.dat ./outpath seism [--all-events] [--3c] [--spol DEG]")
        return 0
    prefix, finpar, outpath = argv[:3]
    fsis = argv[3] if len(argv) > 3 else "seis"
    three, spol = opts["three"], opts["spol"]
    job = io.read_job(prefix, finpar)
    # *_without_pml.su copies of the models (Input.f90:305-341)
    for key, arr in (("vp_model", job["vp"]), ("vs_model", job["vs"])):
        name = job["names"][key]
        r1 = 1 - job.get("izt", 1)   # the row of iz = 1
        io.write_su(os.path.join(prefix, name[:-3] + "_without_pml.su"),
                    arr[job["fd"]["npml"]: job["fd"]["npml"] + job["nx"], r1: r1 + job["nz"]], 1000)
    os.makedirs(outpath, exist_ok=True)
    comps = "xyz" if three else "xz"
    if opts["all_events"]:
        # every line of Source.dat, <fsis>_<k>x.su and z.su for event k (1-based, file order)
        events = io.read_events(os.path.join(prefix, job["names"]["source"]))
        if job["fd"].get("outsnap"):
            print("snapshots are written for single-event runs only")
        results, info = simulate_events(job, events, three=three, spol=spol)
        dh = int(np.float32(info["tstp"]) * np.float32(1e6))
        for k, traces in enumerate(results, 1):
            for c, tr in zip(comps, traces):
                io.write_su(os.path.join(outpath, f"{fsis}_{k}{c}.su"), tr, dh)
        return 0
    *traces, info = simulate(job, three=three, spol=spol)
    dh = int(np.float32(info["tstp"]) * np.float32(1e6))
    for c, tr in zip(comps, traces):
        io.write_su(os.path.join(outpath, fsis + c + ".su"), tr, dh)
    return 0


if __name__ == "__main__":
    sys.exit(main())
