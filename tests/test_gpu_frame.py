"""Frame tiles whose faces and derivatives all carry the full FD order take a
specialised path: there the two face fluxes of a face (orders n1 and n2)
are one value and the coefficients are block constants.  The general path
(FDFK_FRAME_FULL=0) evaluates both fluxes from the order tables; the two
must agree bit for bit.  40-cell layers, so that the side and bottom strips
hold full-order tiles as well as tiles at the grid's edges; FD orders 2, 6
and 16; the hybrid background on."""
import numpy as np
import pytest

from fdfk2d_amd import lib
from tests._cases import Case

pytestmark = pytest.mark.gpu


def solve(case, monkeypatch, full):
    monkeypatch.setenv("FDFK_FRAME_FULL", "1" if full else "0")
    s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                   rx=case.rx, rz=case.rz)
    try:
        s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
        s.run(case.nt + 1)
        return s.seismograms() + s.get_state()
    finally:
        s.close()


@pytest.mark.parametrize("norder", [1, 3, 8])
def test_full_order_frame_tiles_equal_general_path(norder, monkeypatch):
    print("case: order %d, 160 x 90, 40-cell PML, 401 iterations" % (2 * norder), flush=True)
    case = Case(nx=160, nz=90, npml=40, norder=norder, nt=400, anomaly=(16000.0, 9000.0, 2000.0, -0.08))
    ref = solve(case, monkeypatch, False)
    got = solve(case, monkeypatch, True)
    assert np.abs(ref[1]).max() > 1e-6 and all(np.isfinite(a).all() for a in ref)
    for name, a, b in zip(("seis x", "seis z", "ux", "uz"), got, ref):
        bad = np.argwhere(a != b)
        assert bad.size == 0, "%s differs at %d places, first %s" % (name, len(bad), bad[:4].tolist())
