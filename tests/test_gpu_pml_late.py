"""Late-time behaviour of the absorbing layer along the free surface.

The layering of the single large solve (2 km of Vs 1.5 km/s sediment over
the crust and mantle) on a small grid, 801 x 200 at 50 m, P incidence,
20001 iterations (60 s): after the wave has left, the field in the side PML
strips must decay.  The classical split-field layer let it grow from about
30 s on with a 40-cell layer (0.037 at 60 s against a physical field of
3e-3; backward guided modes of the low-velocity layer amplified by the
absorbing layer, worse with a wider one).  The multiaxial damping of the
split parts (ratio 0.1) removes that growth; 20-, 40- and 80-cell layers are
checked.
"""
import numpy as np
import pytest

from fdfk2d_amd import configs, lib

pytestmark = pytest.mark.gpu


def layered(nx, nz, npml, dt=0.003, dx=50.0):
    NX, NZ = nx + 2 * npml, nz + npml
    vsl = [(0.0, 1500.0), (2000.0, 3600.0), (45000.0, 4500.0)]
    layers = [(zt, 1.73 * v, v) for zt, v in vsl]
    z = np.arange(NZ) * dx
    vs1 = np.zeros(NZ, np.float32)
    for zt, v in vsl:
        vs1[z >= zt] = v
    vs = np.repeat(vs1[None, :], NX, 0)
    vp = (np.float32(1.73) * vs).astype(np.float32)
    return configs._common(nx, nz, npml, dx, dx, dt, 20000, vp, vs, layers, True, 0.07, 0.5, 41, 50.0 * nx / 41)


def side_pml_peaks(npml):
    nx, nz = 801, 200
    cfg = layered(nx, nz, npml)
    NX, NZ = nx + 2 * npml, nz + npml
    s = lib.Solver(nx, nz, npml, cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], 20000, cfg["c"],
                   rx=cfg["rx"], rz=cfg["rz"])
    s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"], cfg["rayp"],
                         cfg["bazs"], block_steps=512)
    peaks, done = [], 0
    for tsec in (10, 60):
        it = min(20001, int(round(tsec / 0.003)))
        s.run(it - done)
        done = it
        ux, uz = s.get_state()
        a = np.maximum(np.abs(ux.reshape(NX, NZ, 2)[:, :, 0]), np.abs(uz.reshape(NX, NZ, 2)[:, :, 0]))
        side = max(a[:npml - 5, :nz].max(), a[npml + nx + 5:, :nz].max())
        peaks.append((float(a.max()), float(side)))
    s.close()
    print("npml %d: (all, side PML) at 10 s %s, at 60 s %s" % (npml, peaks[0], peaks[1]), flush=True)
    return peaks


@pytest.mark.parametrize("npml", [20, 40, 80])
def test_side_strips_decay_at_60s(npml):
    print("case: 801 x 200 layered, %d-cell PML, 60 s" % npml, flush=True)
    p10, p60 = side_pml_peaks(npml)
    assert p60[1] < 0.01 * p10[0]
