"""The driver's three-component output on the GPU (--3c, --spol): an
incident P event keeps the reference's x.su and z.su bit for bit and adds
y.su; an S event polarised at 30 deg from SV runs an SV solve of ds cos psi
and an SH solve of ds sin psi, rotated by each receiver's relative
back-azimuth."""
import os

import numpy as np
import pytest

from fdfk2d_amd import driver, io
from fdfk2d_amd.lib import FDFK_SH
from tests._cases import layered_grid, rel_l2

pytestmark = pytest.mark.gpu


def _inputs(prefix, src_is_p, outsnap=False, tmax=2.6):
    nx, nz, npml, dx = 101, 50, 14, 200.0
    layers = [(0.0, 6000.0, 3450.0), (6000.0, 8000.0, 4450.0)]
    vp, vs = layered_grid(nx, nz, npml, dx, dx, layers, anomaly=(9000.0, 4000.0, 1500.0, -0.06))
    fk = [(0.0, 6000.0, 3450.0), (5900.0, 8000.0, 4450.0)]
    rx = np.arange(0, 20001, 2000, dtype=np.float32)
    # the event 20 deg off the profile's line (back-azimuth about 200 deg)
    io.write_inputs(prefix, order=6, x0=0.0, xn=20000.0, z0=0.0, zn=9800.0, dx=dx, dz=dx, dt=0.01,
                    tmax=tmax, tstp=0.01, npml=npml, lat_org=45.0, lon_org=0.0, az_org=0.0, f0=2.0,
                    ds=1e-3, src_is_p=src_is_p, rayp_sdeg=6.6717, evla=-40.0, evlo=-30.0, rx=rx,
                    rz=np.zeros_like(rx), fk_left=fk, fk_right=fk, vp=vp, vs=vs, outsnap=outsnap, nstep=100)


def _read(out, comps):
    return [io.read_su(str(out / f"seis{c}.su"))[0] for c in comps]


def test_3c_p_event_keeps_reference_files(tmp_path, monkeypatch):
    print("case: driver --3c, P event, 101x50, 261 iterations", flush=True)
    monkeypatch.chdir(tmp_path)
    prefix = str(tmp_path / "in")
    _inputs(prefix, True)
    assert driver.main([prefix, "inpar.dat", str(tmp_path / "plain"), "seis"]) == 0
    assert driver.main([prefix, "inpar.dat", str(tmp_path / "three"), "seis", "--3c", "--spol", "40"]) == 0
    x0, z0 = _read(tmp_path / "plain", "xz")
    x, y, z = _read(tmp_path / "three", "xyz")
    assert np.array_equal(x, x0) and np.array_equal(z, z0)
    job = io.read_job(prefix, "inpar.dat")
    phi = driver._receiver_phi(job, job["src"])
    assert np.all(np.abs(np.sin(phi)) > 0.2)
    # y = u_x sin(phi), x = u_x cos(phi)
    assert np.abs(y).max() > 0
    assert np.allclose(y, x * np.tan(phi)[:, None], rtol=1e-5, atol=1e-6 * np.abs(y).max())


def test_3c_s_event_split_into_sv_and_sh(tmp_path, monkeypatch):
    print("case: driver --3c --spol 30, S event, 101x50, 501 iterations, snapshots", flush=True)
    monkeypatch.chdir(tmp_path)
    prefix = str(tmp_path / "in")
    # 5 s: the direct S wave reaches the surface after about 3 s
    _inputs(prefix, False, outsnap=True, tmax=5.0)
    assert driver.main([prefix, "inpar.dat", str(tmp_path / "plain"), "seis"]) == 0
    assert driver.main([prefix, "inpar.dat", str(tmp_path / "three"), "seis", "--3c", "--spol", "30"]) == 0
    x0, z0 = _read(tmp_path / "plain", "xz")
    x, y, z = _read(tmp_path / "three", "xyz")
    snaps = sorted(os.listdir(tmp_path / "snapshots"))
    assert snaps == ["%05d%s.su" % (it, c) for it in range(100, 501, 100) for c in ("ux", "uy", "uz")]
    # the SV part scales with cos(psi) (the FK background is linear in ds)
    c = np.cos(np.deg2rad(30.0))
    assert rel_l2(z, c * z0.astype(np.float64)) < 1e-5
    # the horizontal vector: rotate back by -phi
    job = io.read_job(prefix, "inpar.dat")
    phi = driver._receiver_phi(job, job["src"])
    ux, uy = driver.rotate_horizontal(x, y, -phi)
    # in-plane part: the plain run's x = u_x cos(phi) scaled by cos(psi)
    assert rel_l2(ux * np.cos(phi)[:, None], c * x0.astype(np.float64)) < 1e-5
    # SH part: equal to the SH solver run alone with ds sin(psi)
    s, info = driver._prepare(job, dict(job["src"], ds=1e-3 * 0.5), lambda *a: None, wave=FDFK_SH)
    s.run(info["nt"] + 1)
    sy = s.seismograms()[0]
    s.close()
    assert np.abs(sy).max() > 0.1 * np.abs(ux).max()
    assert rel_l2(uy, sy) < 1e-6
