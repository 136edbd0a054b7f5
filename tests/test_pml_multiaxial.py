"""Late-time stability of the absorbing layer on a strong low-velocity
surface layer (CPU oracle).

The layering of the single large solve (2 km of Vs 1.5 km/s sediment over
Vs 3.6 km/s crust, Vp = 1.73 Vs) with a free surface: a pulse in the
sediment next to the left strip launches guided waves into the PML.  With
the classical split-field layer (multiaxial ratio 0) those grow in the side
strips (about x3 every 5 s); with the multiaxial damping of the split parts
(ratio 0.1, the default of the oracle and of the GPU library) the strips
decay.  The growth is a property of the continuous layer (independent of dt,
worse with a wider layer), so the test runs 40- and 80-cell layers.
"""
import numpy as np
import pytest

from fdfk2d_amd import refstate
from oracle import pyoracle as po

DX, DT = 50.0, 0.003


def side_peaks(npml, ratio, tmax=15.0, nx=61, nz=80):
    """Side-strip maxima of |u| (the damped part of the strips, rows of the
    model) at 5 s and at tmax after a Gaussian pulse at x = 1 km, z = 0.5 km."""
    po.set_pml_ratio(ratio, fp64=False)
    try:
        NX, NZ = nx + 2 * npml, nz + npml
        z = np.arange(NZ) * DX
        vs1 = np.where(z >= 2000.0, 3600.0, 1500.0).astype(np.float32)
        vs = np.repeat(vs1[None, :], NX, 0)
        vp = (np.float32(1.73) * vs).astype(np.float32)
        nt = int(round(tmax / DT))
        st = refstate.make_state(nx, nz, npml, 3, DX, DX, DT, nt, refstate.material_law(vp, vs))
        x = (np.arange(NX) - npml) * DX
        X, Z = np.meshgrid(x, z, indexing="ij")
        g = np.exp(-((X - 1000.0) ** 2 + (Z - 500.0) ** 2) / (2 * 300.0 ** 2)).astype(np.float32)
        for k in ("ux", "uz"):
            st[k][0::2] = g.reshape(-1)
            st[k][1::2] = g.reshape(-1)
        peaks = {}
        marks = {int(round(5.0 / DT)): 5.0, nt: tmax}
        for it in range(1, nt + 1):
            po.updatedispl(st, nt + 1)          # no background (it > nt)
            if it in marks:
                ux = st["ux"].reshape(NX, NZ, 2)[:, :, 0]
                uz = st["uz"].reshape(NX, NZ, 2)[:, :, 0]
                a = np.maximum(np.abs(ux), np.abs(uz))
                peaks[marks[it]] = float(max(a[:npml - 5, :nz].max(), a[npml + nx + 5:, :nz].max()))
        return peaks[5.0], peaks[tmax]
    finally:
        po.set_pml_ratio(refstate.PML_RATIO, fp64=False)


def test_default_ratio_matches_library():
    assert po.pml_ratio() == pytest.approx(refstate.PML_RATIO)


@pytest.mark.parametrize("npml", [40, 80])
def test_side_strips_decay(npml):
    p5, p15 = side_peaks(npml, refstate.PML_RATIO)
    print("npml %d: side strips %.2e at 5 s, %.2e at 15 s" % (npml, p5, p15))
    assert np.isfinite(p15) and p15 < 0.15 * p5


def test_classical_layer_grows():
    """The planted mistake: the classical layer (ratio 0) on the same case
    grows, so the decay test above tells the two apart."""
    p5, p15 = side_peaks(40, 0.0)
    print("classical: side strips %.2e at 5 s, %.2e at 15 s" % (p5, p15))
    assert p15 > 2.0 * p5
