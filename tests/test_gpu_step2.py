"""Two time steps per launch over the deep interior (FDFK_STEP2=1) against
one step per launch.  A deep tile's second step reads the level n+1 that
its first stencil engine computed in registers, with the arithmetic of the
single-step kernels in the same order, so wavefields and seismograms must
agree bit for bit.  Covered: FD orders 2, 4, 6, 8; deep rectangles of one
and of several tile rows; narrow and wide tile columns; receivers inside
the deep rectangle and at the surface; pairs mixed with single steps
(fdfk_run calls of odd lengths); the iterations past nt, where the hybrid
background is off."""
import numpy as np
import pytest

from fdfk2d_amd import lib
from tests._cases import Case

pytestmark = pytest.mark.gpu


def make_case(norder, nx, nz, nt):
    c = Case(nx=nx, nz=nz, npml=max(15, 2 * norder + 6), norder=norder, nt=nt,
             anomaly=(nx * 100.0, nz * 100.0, 2000.0, -0.08))
    # surface receivers plus a line inside the deep rectangle
    deep_x = np.linspace(0.3, 0.7, 9) * (nx - 1) * c.dx
    c.rx = np.concatenate([c.rx, deep_x]).astype(np.float32)
    c.rz = np.concatenate([c.rz, np.full(9, 0.5 * (nz - 1) * c.dz)]).astype(np.float32)
    return c


def solve(case, monkeypatch, step2, chunks, lx2=None):
    monkeypatch.setenv("FDFK_STEP2", "1" if step2 else "0")
    if lx2:
        monkeypatch.setenv("FDFK_LX2", str(lx2))
    else:
        monkeypatch.delenv("FDFK_LX2", raising=False)
    s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                   rx=case.rx, rz=case.rz)
    try:
        s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
        nbytes = s.device_bytes
        for n in chunks:
            s.run(n)
        assert s.iteration == sum(chunks)
        return s.seismograms() + s.get_state(), nbytes
    finally:
        s.close()


CASES = [
    # norder, nx, nz, nt, fdfk_run lengths, deep tile width
    (3, 121, 60, 300, [301], None),
    (1, 90, 40, 200, [1, 2, 3, 50, 145], None),
    (2, 140, 70, 240, [120, 121], 17),
    (4, 150, 70, 240, [7, 234], None),
    (3, 700, 300, 900, [450, 451], 37),
]


@pytest.mark.parametrize("norder,nx,nz,nt,chunks,lx2", CASES)
def test_pairs_equal_single_steps(norder, nx, nz, nt, chunks, lx2, monkeypatch):
    print("case: order %d, %d x %d, %d iterations in runs of %s, deep tile width %s"
          % (2 * norder, nx, nz, sum(chunks), chunks, lx2), flush=True)
    case = make_case(norder, nx, nz, nt)
    ref, b0 = solve(case, monkeypatch, False, [sum(chunks)])
    got, b1 = solve(case, monkeypatch, True, chunks, lx2)
    NX, NZ = nx + 2 * case.npml, nz + case.npml
    ld = (NZ + 31) // 32 * 32
    assert b1 - b0 == 4 * NX * ld * 4            # the pair path ran: two spare levels per component
    assert np.abs(ref[1]).max() > 1e-6 and all(np.isfinite(a).all() for a in ref)   # the wave reached the receivers
    names = ("seis x", "seis z", "ux", "uz")
    for name, a, b in zip(names, got, ref):
        assert a.shape == b.shape
        bad = np.argwhere(a != b)
        assert bad.size == 0, "%s differs at %d places, first %s" % (name, len(bad), bad[:4].tolist())


def test_pairs_off_above_order_8(monkeypatch):
    """Orders above 8 (norder > 4) keep single steps: no spare levels."""
    case = make_case(5, 120, 60, 20)
    _, b0 = solve(case, monkeypatch, False, [21])
    _, b1 = solve(case, monkeypatch, True, [21])
    assert b1 == b0
