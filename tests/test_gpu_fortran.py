"""The Fortran binding (include/fdfk.f90, ISO_C_BINDING) on the GPU: a
Fortran program (tests/fortran/handle_api_run.f90, built by AMD flang) runs
a case resident through the handle API; its seismograms must equal, bit for
bit, those of the same case through the Python binding, which checks the
derived types' layout against the C structs and the calling conventions."""
import os
import subprocess

import numpy as np
import pytest

from fdfk2d_amd import lib
from tests._cases import Case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "handle_api_run")


def write_case(path, c, nstp, src_is_p):
    with open(path, "wb") as f:
        np.array([c.nx, c.nz, c.npml, c.norder, c.nt, nstp, len(c.rx), len(c.fk), len(c.fk),
                  1 if src_is_p else 0], np.int32).tofile(f)
        np.array([c.dx, c.dz, 0.0, 0.0, c.ds, c.f0], np.float32).tofile(f)
        np.array([c.dt, c.rayp], np.float64).tofile(f)
        np.ascontiguousarray(c.c, np.float32).tofile(f)
        c.rx.astype(np.float32).tofile(f)
        c.rz.astype(np.float32).tofile(f)
        np.ascontiguousarray(c.fk, np.float32).tofile(f)
        np.ascontiguousarray(c.fk, np.float32).tofile(f)
        np.ascontiguousarray(c.bazs, np.float64).tofile(f)


@pytest.mark.parametrize("nstp,src_is_p", [(1, True), (3, False)])
def test_fortran_handle_api_equals_python(tmp_path, nstp, src_is_p):
    print(f"case: Fortran handle-API driver, 121 x 60, nstp {nstp}, P {src_is_p}", flush=True)
    if not os.path.exists(DRIVER):
        pytest.skip("tests/fortran/handle_api_run not built (AMD flang absent at build time)")
    c = Case(nt=300 if src_is_p else 600, src_is_p=src_is_p)   # the SV wave needs ~5 s to reach the surface
    fin, fout = str(tmp_path / "case.bin"), str(tmp_path / "seis.bin")
    write_case(fin, c, nstp, src_is_p)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    nsamp = c.nt // nstp + 1
    got = np.fromfile(fout, np.float32)
    assert got.size == 2 * len(c.rx) * nsamp
    fx, fz = got[: got.size // 2].reshape(len(c.rx), nsamp), got[got.size // 2:].reshape(len(c.rx), nsamp)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, nstp=nstp, rx=c.rx, rz=c.rz)
    s.compute_background(c.fk, c.fk, src_is_p, c.f0, c.ds, c.rayp, c.bazs)
    s.run(c.nt + 1)
    px, pz = s.seismograms()
    s.close()
    assert np.abs(pz).max() > 1e-5
    assert np.array_equal(fx, px) and np.array_equal(fz, pz)
