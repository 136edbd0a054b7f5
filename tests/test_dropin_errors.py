"""Error contract of the drop-in entry points (CPU, no GPU needed).

The reference stops on a bad configuration (print and `stop`, e.g.
src/Submain.f90:149-155).  updatedispl_ and comput_background_wavefield_
check their arguments on the host before any GPU work and end the process
with a non-zero status on failure, so the reference's own driver
(src/FDFK2D.f90, built from where it lies and linked against libfdfk.so)
cannot go on writing seismograms from unchanged arrays.  The input here is
one the reference accepts but this library rejects: FD order 18 (norder 9;
the library supports orders 2..16, and the reference clamps norder only to
2 norder <= npml + 1, src/Input.f90:130-131).
"""
import os
import subprocess

import numpy as np
import pytest

from fdfk2d_amd import io
from tests._cases import layered_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "fdfk2d_ref")


def test_reference_driver_stops_on_rejected_order(tmp_path):
    if not os.path.exists(REF_DRIVER):
        pytest.skip("oracle/_ref/fdfk2d_ref not built (reference tree absent at build time)")
    nx, nz, npml, dx = 61, 40, 20, 200.0
    layers = [(0.0, 6000.0, 3450.0), (6000.0, 8000.0, 4450.0)]
    vp, vs = layered_grid(nx, nz, npml, dx, dx, layers)
    fk = [(0.0, 6000.0, 3450.0), (6000.0, 8000.0, 4450.0)]
    rx = np.arange(0, 12001, 2000, dtype=np.float32)
    prefix = str(tmp_path / "input")
    io.write_inputs(prefix, order=18, x0=0.0, xn=12000.0, z0=0.0, zn=7800.0, dx=dx, dz=dx, dt=0.005,
                    tmax=0.5, tstp=0.005, npml=npml, lat_org=45.0, lon_org=0.0, az_org=0.0, f0=2.0,
                    ds=1e-3, src_is_p=True, rayp_sdeg=6.6717, evla=-42.0, evlo=10.0, rx=rx,
                    rz=np.zeros_like(rx), fk_left=fk, fk_right=fk, vp=vp, vs=vs)
    out = tmp_path / "seis"
    out.mkdir()
    # no GPU is visible to the child: the rejection must come before any HIP call
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([REF_DRIVER, prefix, "inpar.dat", str(out), "seis"], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path), env=env)
    print(r.stdout[-800:], r.stderr[-800:])
    assert r.returncode != 0
    assert "norder must be 1..8" in r.stderr
    assert not (out / "seisx.su").exists()
