"""Drop-in integration on the GPU: the reference's own Fortran driver
(src/FDFK2D.f90 + Input.f90, compiled with flang where it lies, linked
against libfdfk.so instead of -lHybrid -lPML -lFK3D) runs an input directory
in the reference's formats; its SU seismograms must equal the package's
driver (handle API) and the CPU oracle pipeline on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from fdfk2d_amd import driver, io
from tests._cases import layered_grid, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "fdfk2d_ref")


SNAP_NSTEP = 100


def make_inputs(prefix, outsnap=True, event=(6.6717, -42.0, 10.0), more_events=(), pml_top=False):
    nx, nz, npml, dx = 101, 50, 14, 200.0
    layers = [(0.0, 6000.0, 3450.0), (6000.0, 8000.0, 4450.0)]
    vp, vs = layered_grid(nx, nz, npml, dx, dx, layers, anomaly=(9000.0, 4000.0, 1500.0, -0.06),
                          izt=1 - npml if pml_top else 1)
    fk = [(0.0, 6000.0, 3450.0), (5900.0, 8000.0, 4450.0)]
    rx = np.arange(0, 20001, 2000, dtype=np.float32)
    io.write_inputs(prefix, order=6, x0=0.0, xn=20000.0, z0=0.0, zn=9800.0, dx=dx, dz=dx, dt=0.01,
                    tmax=2.6, tstp=0.01, npml=npml, lat_org=45.0, lon_org=0.0, az_org=0.0, f0=2.0,
                    ds=1e-3, src_is_p=True, rayp_sdeg=event[0], evla=event[1], evlo=event[2], rx=rx,
                    rz=np.zeros_like(rx), fk_left=fk, fk_right=fk, vp=vp, vs=vs, outsnap=outsnap,
                    nstep=SNAP_NSTEP, more_events=more_events, pml_top=pml_top)


def oracle_pipeline(job):
    """The same run on the CPU oracle: FK, loop with loadseis, projection."""
    from oracle import pyoracle as po
    from fdfk2d_amd import refstate
    fd, src = job["fd"], job["src"]
    nx, nz, npml, norder = job["nx"], job["nz"], fd["npml"], fd["norder"]
    izt = job["izt"]
    nt = int(np.rint(fd["tmax"] / fd["dt"]))
    ix = np.arange(1 - npml, nx + npml + 1)
    coordx = (np.float32(fd["x0"]) + (ix - 1).astype(np.float32) * np.float32(fd["dx"])).astype(np.float32)
    bazs = po.backazimuth_grids(fd["lat_org"], fd["lon_org"], fd["az_org"], coordx, src["evla"],
                                src["evlo"]) - fd["az_org"] * np.pi / 180.0
    bg = po.background(nx, nz, npml, izt, 2 * norder - 1, nt, src["src_is_p"], fd["x0"], fd["z0"],
                       fd["dx"], fd["dz"], fd["dt"], src["ds"], src["f0"], src["rayp"], bazs,
                       job["fk_left"], job["fk_right"])
    st = refstate.make_state(nx, nz, npml, norder, fd["dx"], fd["dz"], fd["dt"], nt, job["c"],
                             pml_top=izt != 1)
    if isinstance(bg, tuple):   # with the top ring of an absorbing top layer
        st["bg"], st["bg_top"] = bg
    else:
        st["bg"] = bg
    sx, sz = [], []
    for it in range(nt + 1):
        row = [po.loadseis(nx, nz, npml, izt, fd["x0"], fd["z0"], fd["dx"], fd["dz"], float(a), float(b),
                           st["ux"], st["uz"]) for a, b in zip(job["rx"], job["rz"])]
        sx.append([r[0] for r in row])
        sz.append([r[1] for r in row])
        po.updatedispl(st, it)
    sx = np.asarray(sx, np.float32).T
    sz = np.asarray(sz, np.float32).T
    br = po.backazimuth_grids(fd["lat_org"], fd["lon_org"], fd["az_org"], job["rx"], src["evla"],
                              src["evlo"]) - fd["az_org"] * np.pi / 180.0
    return (sx.astype(np.float64) * np.cos(br)[:, None]).astype(np.float32), sz


@pytest.mark.parametrize("pml_top", [False, True])
def test_reference_driver_dropin(tmp_path, monkeypatch, pml_top):
    print(f"case: reference Fortran driver linked to libfdfk.so, 101x50, 261 iterations, is_PML_top={pml_top}",
          flush=True)
    if not os.path.exists(REF_DRIVER):
        pytest.skip("oracle/_ref/fdfk2d_ref not built (reference tree absent at build time)")
    prefix = str(tmp_path / "input")
    make_inputs(prefix, pml_top=pml_top)
    out_ref = tmp_path / "seis_ref"
    out_ref.mkdir()
    (tmp_path / "snapshots").mkdir()   # the reference opens ./snapshots/ without creating it
    r = subprocess.run([REF_DRIVER, prefix, "inpar.dat", str(out_ref), "seis"], capture_output=True,
                       text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rx_ref, h = io.read_su(str(out_ref / "seisx.su"))
    rz_ref, _ = io.read_su(str(out_ref / "seisz.su"))
    out_pkg = tmp_path / "seis_pkg"
    pkg_cwd = tmp_path / "pkg"
    pkg_cwd.mkdir()
    monkeypatch.chdir(pkg_cwd)
    assert driver.main([prefix, "inpar.dat", str(out_pkg), "seis"]) == 0
    rx_pkg, h2 = io.read_su(str(out_pkg / "seisx.su"))
    rz_pkg, _ = io.read_su(str(out_pkg / "seisz.su"))
    assert np.array_equal(h, h2) and rx_ref.shape == (11, 261)
    assert np.abs(rz_ref).max() > 1e-4
    assert rel_l2(rx_pkg, rx_ref) < 1e-6 and rel_l2(rz_pkg, rz_ref) < 1e-6
    # snapshots (src/FDFK2D.f90:467-480): the same files, bit for bit
    snaps = sorted(os.listdir(tmp_path / "snapshots"))
    assert snaps == ["00100ux.su", "00100uz.su", "00200ux.su", "00200uz.su"]
    for name in snaps:
        a, ha = io.read_su(str(tmp_path / "snapshots" / name))
        b, hb = io.read_su(str(pkg_cwd / "snapshots" / name))
        assert a.shape == (101, 50) and np.array_equal(ha, hb) and np.array_equal(a, b), name
    job = io.read_job(prefix, "inpar.dat")
    ox, oz = oracle_pipeline(job)
    assert rel_l2(rx_pkg, ox) < 1e-5 and rel_l2(rz_pkg, oz) < 1e-5


def test_all_events_batched_equal_single_runs(tmp_path, monkeypatch):
    """--all-events on a three-event Source.dat: event k's seismograms are
    bit-identical to a single-event run of that event, whether the events
    share one batch or each gets its own (a budget too small for two)."""
    print("case: driver --all-events, 3 events, 101x50, 261 iterations", flush=True)
    events = [(6.6717, -42.0, 10.0), (7.2, -35.0, 5.0), (5.9, -50.0, 20.0)]
    monkeypatch.chdir(tmp_path)
    multi = str(tmp_path / "multi")
    make_inputs(multi, outsnap=False, more_events=events[1:])
    out_all = tmp_path / "out_all"
    assert driver.main([multi, "inpar.dat", str(out_all), "seis", "--all-events"]) == 0
    got = []
    for k, ev in enumerate(events, 1):
        one = str(tmp_path / f"single{k}")
        make_inputs(one, outsnap=False, event=ev)
        out = tmp_path / f"out{k}"
        assert driver.main([one, "inpar.dat", str(out), "seis"]) == 0
        for comp in "xz":
            a, ha = io.read_su(str(out_all / f"seis_{k}{comp}.su"))
            b, hb = io.read_su(str(out / f"seis{comp}.su"))
            assert a.shape == (11, 261) and np.array_equal(ha, hb) and np.array_equal(a, b), (k, comp)
        got.append(io.read_su(str(out_all / f"seis_{k}z.su"))[0])
    assert rel_l2(got[1], got[0]) > 1e-2 and rel_l2(got[2], got[0]) > 1e-2
    job = io.read_job(multi, "inpar.dat")
    ev = io.read_events(os.path.join(multi, "Source.dat"))
    solo, _ = driver.simulate_events(job, ev, hbm_budget_gb=1e-6, log=lambda *a: None)
    assert len(solo) == 3
    for k in range(3):
        assert np.array_equal(solo[k][1], got[k])
