"""GPU FK background (FP64 propagators + hipFFT synthesis) against the CPU
oracle's FK, and the complete GPU pipeline (GPU FK + GPU stepping) against
the complete oracle pipeline."""
import numpy as np
import pytest

from fdfk2d_amd import lib
from tests._cases import Case, fk_model, oracle_run, rel_l2

pytestmark = pytest.mark.gpu

LAYERS_L = ((0.0, 6000.0, 3450.0), (5000.0, 7000.0, 4000.0), (8000.0, 8000.0, 4450.0))
LAYERS_R = ((0.0, 5800.0, 3300.0), (6000.0, 7000.0, 4000.0), (8000.0, 8000.0, 4450.0))


def _case(src_is_p, baz, same=True):
    c = Case(nx=101, nz=50, npml=14, nt=260, src_is_p=src_is_p, baz=baz, f0=2.0, nrcv=11,
             layers=LAYERS_L, anomaly=(10000.0, 4000.0, 1500.0, -0.05))
    if not same:
        c.fk_right = fk_model(LAYERS_R)
    else:
        c.fk_right = c.fk
    return c


def _oracle_bg(c):
    from oracle import pyoracle as po
    return po.background(c.nx, c.nz, c.npml, 1, c.nbd, c.nt, c.src_is_p, 0.0, 0.0, c.dx, c.dz,
                         c.dt, c.ds, c.f0, c.rayp, c.bazs, c.fk, c.fk_right)


@pytest.mark.parametrize("src_is_p,baz,same", [(True, np.pi, True), (False, 0.3, True),
                                               (True, 0.2, False)])
def test_gpu_background_matches_oracle(src_is_p, baz, same):
    print(f"case: FK background P={src_is_p} baz={baz} same={same}", flush=True)
    c = _case(src_is_p, baz, same)
    ref = _oracle_bg(c)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c)
    s.compute_background(c.fk, c.fk_right, src_is_p, c.f0, c.ds, c.rayp, c.bazs)
    got = s.get_background()
    s.close()
    scale = max(float(np.abs(a).max()) for a in ref)
    assert scale > 1e-4
    for a, b in zip(got, ref):
        assert np.abs(a.astype(np.float64) - b).max() <= 1e-6 * scale


def test_compat_background_matches_oracle():
    print("case: comput_background_wavefield_", flush=True)
    c = _case(True, np.pi)
    ref = _oracle_bg(c)
    got = lib.comput_background_wavefield(c.nx, c.nz, c.npml, 1, c.nbd, c.nt, True, 0.0, 0.0,
                                          c.dx, c.dz, c.dt, c.ds, c.f0, c.rayp, c.bazs, c.fk, c.fk)
    scale = max(float(np.abs(a).max()) for a in ref)
    for a, b in zip(got, ref):
        assert np.abs(a.astype(np.float64) - b).max() <= 1e-6 * scale


def test_full_gpu_pipeline_matches_oracle():
    print("case: GPU FK + GPU steps vs oracle FK + oracle steps", flush=True)
    c = _case(True, np.pi)
    bg = _oracle_bg(c)
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz)
    s.compute_background(c.fk, c.fk, True, c.f0, c.ds, c.rayp, c.bazs)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert np.abs(oz).max() > 1e-4
    assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5
