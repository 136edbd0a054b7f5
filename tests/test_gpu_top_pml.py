"""GPU path of the absorbing top layer (is_PML_top = .true., izt = 1-npml)
against the CPU oracle's formulation of it.

The general per-cell kernels evaluate the oracle's pointwise operators with
contraction off and FP64 accumulation in the same order, so with every cell
on them (FDFK_GEN_TILED=0) the FP32 run equals the oracle bit for bit
(fields, strips and seismograms).  By default the deep interior runs on the
tiled interior kernel (FP32 with FMA contraction and hardware reciprocals),
held to rel-L2 1e-5 against the oracle like the free-surface tiled path.
The GPU FK background (top layer extended upward, no free surface) is held
to 1e-6 of its peak against the oracle's, like the free-surface background.
"""
import numpy as np
import pytest

from fdfk2d_amd import lib
from oracle import pyoracle as po
from tests._cases import Case, fk_model, oracle_run, rel_l2

pytestmark = pytest.mark.gpu


def _solver(case):
    return lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                      rx=case.rx, rz=case.rz, pml_top=True)


CASES = [(3, 121, 60, 15, 300), (1, 90, 40, 12, 200), (5, 110, 50, 16, 220)]


@pytest.mark.parametrize("norder,nx,nz,npml,nt", CASES)
def test_top_layer_equals_oracle_bit_for_bit(monkeypatch, norder, nx, nz, npml, nt):
    print(f"case: top layer, per-cell kernels only, order {2 * norder}, {nx}x{nz}, npml {npml}, {nt} iterations",
          flush=True)
    monkeypatch.setenv("FDFK_GEN_TILED", "0")
    case = Case(nx=nx, nz=nz, npml=npml, norder=norder, nt=nt, nrcv=13, pml_top=True,
                rayp_sdeg=6.0, f0=1.2)
    bg = case.oracle_background()
    st = case.state()
    ox, oz = oracle_run(case, st, bg, nt + 1)
    s = _solver(case)
    s.set_background(bg)
    s.run(nt + 1)
    gx, gz = s.seismograms()
    ux, uz = s.get_state()
    s.close()
    assert np.abs(ox).max() > 0
    assert np.array_equal(gx.view(np.uint32), ox.view(np.uint32))
    assert np.array_equal(gz.view(np.uint32), oz.view(np.uint32))
    assert np.array_equal(ux.view(np.uint32), st["ux"].view(np.uint32))
    assert np.array_equal(uz.view(np.uint32), st["uz"].view(np.uint32))


@pytest.mark.parametrize("norder,nx,nz,npml,nt", CASES + [(6, 100, 45, 14, 200)])
def test_top_layer_tiled_interior_against_oracle(norder, nx, nz, npml, nt):
    print(f"case: top layer, tiled interior, order {2 * norder}, {nx}x{nz}, npml {npml}, {nt} iterations", flush=True)
    case = Case(nx=nx, nz=nz, npml=npml, norder=norder, nt=nt, nrcv=13, pml_top=True, rayp_sdeg=6.0, f0=1.2)
    case.rz = np.linspace(0.0, (nz - 1) * case.dz, 13).astype(np.float32)   # receivers at depth too
    bg = case.oracle_background()
    st = case.state()
    ox, oz = oracle_run(case, st, bg, nt + 1)
    s = _solver(case)
    s.set_background(bg)
    s.run(nt + 1)
    gx, gz = s.seismograms()
    ux, uz = s.get_state()
    s.close()
    e = max(rel_l2(gx, ox), rel_l2(gz, oz), rel_l2(ux, st["ux"]), rel_l2(uz, st["uz"]))
    print(f"tiled top layer vs oracle: rel-L2 {e:.2e}", flush=True)
    assert np.abs(ox).max() > 0 and e < 1e-5


def test_top_layer_fk_background_matches_oracle():
    print("case: GPU FK background with the top ring, 121x60", flush=True)
    case = Case(nt=300, pml_top=True, src_is_p=False, rayp_sdeg=8.0)
    arrs, top = case.oracle_background()
    s = _solver(case)
    s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
    garrs, gtop = s.get_background()
    s.close()
    peak = max(np.abs(a).max() for a in arrs + top)
    assert peak > 0
    for a, b in zip(garrs + gtop, arrs + top):
        assert np.abs(a.astype(np.float64) - b).max() <= 1e-6 * peak


def test_top_layer_pipeline_matches_oracle_and_fk():
    """GPU FK + GPU run vs oracle FK + oracle run; and with the FD model equal
    to the FK model the surface records equal the FK response without a free
    surface (the incident wave leaves through the top)."""
    print("case: top layer pipeline, homogeneous 151x60, 700 iterations", flush=True)
    layers = ((0.0, 6000.0, 3450.0),)
    case = Case(nx=151, nz=60, npml=20, nt=700, layers=layers, anomaly=None, f0=1.0, nrcv=7, pml_top=True)
    bg = case.oracle_background()
    st = case.state()
    ox, oz = oracle_run(case, st, bg, case.nt + 1)
    s = _solver(case)
    s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
    s.run(case.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5
    p = case.rayp
    xe = (1 - case.nbd - 1) * case.dx
    ux, uz = po.fk_traces(case.fk, True, p, case.f0, case.ds, case.dt, case.nt + 400, np.array([0.0]),
                          (case.nz + case.nbd - 1) * case.dz, 1.5 / case.f0, free_surface=False)
    n = np.arange(case.nt + 1)
    for i, rx in enumerate(case.rx):
        sh = p * (rx - xe) / case.dt
        kx = np.interp(n - sh, np.arange(ux.shape[1]), ux[0])
        kz = np.interp(n - sh, np.arange(uz.shape[1]), uz[0])
        assert rel_l2(gx[i], kx) < 0.01 and rel_l2(gz[i], kz) < 0.005


def test_top_layer_refusals():
    case = Case(nt=10, pml_top=True)
    s = _solver(case)
    b = lib.Batch(case.norder)
    with pytest.raises(lib.FdfkError):
        b.add(s)
    b.close()
    s.close()


@pytest.mark.parametrize("wave,pml_top", [("psv", True), ("sh", True), ("sh", False)])
def test_second_stream_equals_one_stream(monkeypatch, wave, pml_top):
    """An FP32 general-path step runs its per-cell kernels and the tiled
    interior kernel on two streams side by side: with FDFK_GEN_OVERLAP=2
    (P-SV's default) the per-cell kernels first on the solver's stream and
    the interior on the second, with FDFK_GEN_OVERLAP=1 (SH's default) the
    interior first, with FDFK_GEN_OVERLAP=0 everything in line on one
    stream.  The schedules write disjoint cells from the same inputs, so
    seismograms and final fields are identical."""
    print(f"case: {wave}, top layer {pml_top}, second stream against one stream, 121x60, 260 iterations", flush=True)
    case = Case(nx=121, nz=60, npml=15, norder=3, nt=260, nrcv=13, pml_top=pml_top, rayp_sdeg=6.0, f0=1.2)
    case.rz = np.linspace(0.0, (case.nz - 1) * case.dz, 13).astype(np.float32)
    bg = case.oracle_background()
    w = lib.FDFK_SH if wave == "sh" else lib.FDFK_PSV
    out = []
    for overlap in (None, "2", "1", "0"):
        if overlap is None:
            monkeypatch.delenv("FDFK_GEN_OVERLAP", raising=False)
        else:
            monkeypatch.setenv("FDFK_GEN_OVERLAP", overlap)
        s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                       rx=case.rx, rz=case.rz, pml_top=pml_top, wave=w)
        s.set_background(bg)
        s.run(case.nt + 1)
        out.append(s.seismograms() + s.get_state())
        s.close()
    assert np.abs(out[0][0]).max() > 0
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
