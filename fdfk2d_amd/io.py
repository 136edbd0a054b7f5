"""Input and output formats of the reference driver, unchanged.

Readers follow src/Input.f90 (list-directed reads of the fixed-layout text
files named in inpar.dat, FK models with rho = (vp+980)/2.76, the Vp/Vs
models as SU files holding the padded grid) and writers produce the same
files, including the SU seismograms of src/FDFK2D.f90:547-561 (240-byte
header of 120 int16 with ns at word 58 and dt in microseconds at word 59,
then FP32 samples, native endian, one trace per receiver).
"""
import os
import re

import numpy as np

from . import refstate

SU_HEADER_WORDS = 120


def _tokens(line):
    return [t for t in re.split(r"[\s,]+", line.strip()) if t]


def _num(tok):
    return float(re.sub(r"[dD]", "e", tok))


def _logical(tok):
    t = tok.strip().strip(".").lower()
    return t in ("true", "t")


def _data_lines(path):
    with open(path) as f:
        return f.read().splitlines()


def read_inpar(prefix, finpar):
    """The seven file names of inpar.dat (Input.f90:34-73)."""
    names = [ln.strip() for ln in _data_lines(os.path.join(prefix, finpar)) if ln.strip()]
    keys = ["fd_model", "source", "receiver", "fk_left", "fk_right", "vp_model", "vs_model"]
    if len(names) < 7:
        raise ValueError("inpar file must list seven files")
    return dict(zip(keys, names[:7]))


def read_fd_model(path):
    """FD_model.dat (Input.f90:118-131)."""
    L = _data_lines(path)
    order = int(_num(_tokens(L[1])[0]))
    x0, xn, z0, zn, dx, dz, dt = [_num(t) for t in _tokens(L[3])[:7]]
    t = _tokens(L[5])
    tmax, tstp, npml, pml_top = _num(t[0]), _num(t[1]), int(_num(t[2])), _logical(t[3])
    lat, lon, az = [_num(v) for v in _tokens(L[7])[:3]]
    t = _tokens(L[9])
    outsnap, nstep = _logical(t[0]), int(_num(t[1]))
    norder = order // 2
    if 2 * norder > npml + 1:
        norder = (npml + 1) // 2
    return dict(order=order, norder=norder, x0=x0, xn=xn, z0=z0, zn=zn, dx=dx, dz=dz, dt=dt,
                tmax=tmax, tstp=tstp, npml=npml, pml_top=pml_top, lat_org=lat, lon_org=lon,
                az_org=az, outsnap=outsnap, nstep=nstep)


def read_source(path):
    """Source.dat (Input.f90:139-146); rayp converted from s/deg to s/m."""
    L = _data_lines(path)
    t = _tokens(L[1])
    f0, ds, src_is_p = _num(t[0]), _num(t[1]), _logical(t[2])
    t = _tokens(L[3])
    rayp_sdeg, evla, evlo = _num(t[0]), _num(t[1]), _num(t[2])
    return dict(f0=f0, ds=ds, src_is_p=src_is_p, rayp_sdeg=rayp_sdeg,
                rayp=rayp_sdeg * 180.0 / (np.pi * 6371.0e3), evla=evla, evlo=evlo)


def read_events(path):
    """Every event of a multi-event Source.dat (examples/Altyn/input/Source.dat:4-15):
    the event of line 4 (the only one the reference reads, Input.f90:142-146)
    followed by each later line that starts with three numbers (ray parameter
    in s/deg, latitude, longitude); blank lines and text after '!' are
    skipped.  f0, strength and source type of line 2 apply to all events."""
    L = _data_lines(path)
    first = read_source(path)
    events = []
    for ln in L[3:]:
        t = _tokens(ln.split("!", 1)[0])
        if len(t) < 3:
            continue
        try:
            rayp_sdeg, evla, evlo = _num(t[0]), _num(t[1]), _num(t[2])
        except ValueError:
            continue
        ev = dict(first, rayp_sdeg=rayp_sdeg, rayp=rayp_sdeg * 180.0 / (np.pi * 6371.0e3),
                  evla=evla, evlo=evlo)
        events.append(ev)
    return events


def read_receivers(path):
    """Receiver.dat (Input.f90:153-155, 226-236)."""
    L = _data_lines(path)
    n = int(_num(_tokens(L[1])[0]))
    rx, rz = [], []
    for ln in L[3:3 + n]:
        t = _tokens(ln)
        rx.append(_num(t[0]))
        rz.append(_num(t[1]))
    return np.asarray(rx, np.float32), np.asarray(rz, np.float32)


def read_fk_model(path):
    """FK_model_*.dat -> (nlayer, 4) rows ztop, vp, vs, rho (Input.f90:253-266)."""
    L = _data_lines(path)
    n = int(_num(_tokens(L[1])[0]))
    rows = []
    for ln in L[3:3 + n]:
        vp, vs, zt = [_num(v) for v in _tokens(ln)[:3]]
        vpf = np.float32(vp)
        rows.append([zt, vp, vs, (vpf + np.float32(980.0)) / np.float32(2.760)])
    return np.asarray(rows, np.float32)


def read_su(path, ns=None):
    """SU file -> (ntrace, ns) float32 and the header of the first trace."""
    raw = np.fromfile(path, dtype=np.uint8)
    head = raw[:240].view(np.int16)
    if ns is None:
        ns = int(np.uint16(head[57]))
    rec = 240 + 4 * ns
    ntr = raw.size // rec
    tr = raw[: ntr * rec].reshape(ntr, rec)[:, 240:].copy().view(np.float32)
    return tr.reshape(ntr, ns), head.copy()


def write_su(path, traces, d_header):
    """traces (ntrace, ns); header word 59 = d_header (int16, as the reference)."""
    traces = np.ascontiguousarray(traces, np.float32)
    ntr, ns = traces.shape
    head = np.zeros(SU_HEADER_WORDS, np.int16)
    head[57] = np.int16(ns)
    head[58] = np.int16(d_header)
    with open(path, "wb") as f:
        for i in range(ntr):
            f.write(head.tobytes())
            f.write(traces[i].tobytes())


def read_job(prefix, finpar):
    """Everything get_parameter + read_data set up (src/Input.f90:75-381)."""
    names = read_inpar(prefix, finpar)
    fd = read_fd_model(os.path.join(prefix, names["fd_model"]))
    src = read_source(os.path.join(prefix, names["source"]))
    rx, rz = read_receivers(os.path.join(prefix, names["receiver"]))
    fkl = read_fk_model(os.path.join(prefix, names["fk_left"]))
    fkr = read_fk_model(os.path.join(prefix, names["fk_right"]))
    if np.abs(fkl[-1, 1:] - fkr[-1, 1:]).max() > 1e-6:
        raise ValueError("The bottom half-space must be identical !")
    nx = int(round((fd["xn"] - fd["x0"]) / fd["dx"])) + 1
    nz = int(round((fd["zn"] - fd["z0"]) / fd["dz"])) + 1
    npml = fd["npml"]
    # an absorbing top layer adds npml rows above row 1 (src/Input.f90:179-183)
    izt = 1 - npml if fd["pml_top"] else 1
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    xmin, xmax = fd["x0"], fd["x0"] + (nx - 1) * fd["dx"]
    zmin, zmax = fd["z0"], fd["z0"] + (nz - 1) * fd["dz"]
    bad = (rx < xmin) | (rx > xmax) | (rz < zmin) | (rz > zmax)
    if bad.any():
        i = int(np.argmax(bad))
        raise ValueError(f"receiver {i + 1} at ({rx[i]}, {rz[i]}) locates outside the model")
    vp, _ = read_su(os.path.join(prefix, names["vp_model"]), NZ)
    vs, _ = read_su(os.path.join(prefix, names["vs_model"]), NZ)
    if vp.shape[0] < NX or vs.shape[0] < NX:
        raise ValueError("velocity models must hold the padded grid (nx+2*npml traces)")
    c = refstate.material_law(vp[:NX], vs[:NX])
    return dict(names=names, fd=fd, src=src, rx=rx, rz=rz, fk_left=fkl, fk_right=fkr, nx=nx,
                nz=nz, NX=NX, NZ=NZ, izt=izt, c=c, vp=vp[:NX], vs=vs[:NX])


def write_inputs(prefix, *, order, x0, xn, z0, zn, dx, dz, dt, tmax, tstp, npml, lat_org,
                 lon_org, az_org, f0, ds, src_is_p, rayp_sdeg, evla, evlo, rx, rz, fk_left,
                 fk_right, vp, vs, outsnap=False, nstep=100, more_events=(), pml_top=False):
    """Write a complete input directory in the reference's formats.
    fk_left/right: rows (ztop, vp, vs); vp, vs: padded (NX, NZ) grids (rows
    from izt = 1-npml with pml_top); more_events: further (rayp_sdeg, evla,
    evlo) lines of Source.dat."""
    os.makedirs(prefix, exist_ok=True)
    L = lambda b: ".true." if b else ".false."
    with open(os.path.join(prefix, "inpar.dat"), "w") as f:
        f.write("FD_model.dat\nSource.dat\nReceiver.dat\nFK_model_left.dat\n"
                "FK_model_right.dat\nvp_model.su\nvs_model.su\n")
    with open(os.path.join(prefix, "FD_model.dat"), "w") as f:
        f.write("order of FD(2m>=2)\n")
        f.write(f"  {order}\n")
        f.write("X0(m)        Xn(m)        Z0(m)     Zn(m)       Dx(m)      Dz(m)       Dt(s)\n")
        f.write(f"  {x0!r} {xn!r} {z0!r} {zn!r} {dx!r} {dz!r} {dt!r}\n")
        f.write("tmax(s)      tstep        nPML(number of grid)  is_PML_top\n")
        f.write(f"  {tmax!r} {tstp!r} {npml} {L(pml_top)}\n")
        f.write("latitude of (X0,Z0)       longitude of (X0,Z0)  azimuthe of profile\n")
        f.write(f"  {lat_org!r} {lon_org!r} {az_org!r}\n")
        f.write("output snapshot?(.true./.false.)  nstep\n")
        f.write(f"  {L(outsnap)}  {nstep}\n")
    with open(os.path.join(prefix, "Source.dat"), "w") as f:
        f.write("f0(Hz)   streng of source(m)   source type(P: .true.; S: .false.)\n")
        f.write(f"  {f0!r} {ds!r} {L(src_is_p)}\n")
        f.write("ray parameter(s/deg)   latitude of source(deg)   longtitude of source(deg)\n")
        f.write(f"  {rayp_sdeg!r} {evla!r} {evlo!r}\n")
        if more_events:
            f.write("  \n")
            for k, (p, la, lo) in enumerate(more_events):
                f.write(f"  {p!r}  {la!r}\t{lo!r}   ! event {k + 2}\n")
    with open(os.path.join(prefix, "Receiver.dat"), "w") as f:
        f.write("number of receivers:\n")
        f.write(f"  {len(rx)}\n")
        f.write("rx(m)   rz(m)\n")
        for a, b in zip(rx, rz):
            f.write(f"{float(a)!r}\t{float(b)!r}\n")
    for name, fk in (("FK_model_left.dat", fk_left), ("FK_model_right.dat", fk_right)):
        with open(os.path.join(prefix, name), "w") as f:
            f.write("nlayer\n")
            f.write(f"  {len(fk)}\n")
            f.write("Vp(m/s)   Vs(m/s)   z+(m) <----- top boundary of the layers\n")
            for zt, p, s in fk:
                f.write(f"{float(p)!r}  {float(s)!r}  {float(zt)!r}\n")
    write_su(os.path.join(prefix, "vp_model.su"), vp, int(round(dz * 1e3)) if dz * 1e3 < 32767 else 0)
    write_su(os.path.join(prefix, "vs_model.su"), vs, int(round(dz * 1e3)) if dz * 1e3 < 32767 else 0)
