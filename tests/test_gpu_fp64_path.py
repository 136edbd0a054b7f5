"""The FDFK_FP64 storage mode: every array and all arithmetic in double, the
general per-cell kernels instantiated for double.  It is the FP64 GPU
reference of the FP32 and 16-bit modes at sizes the CPU oracle cannot run.

Pinned against the oracle's FP64 build (oracle/build/liboracle64.so, the
same formulation with every FP32 quantity in double): on the same FP64
background the GPU run equals it bit for bit (seismograms and fields),
with the free surface and with the absorbing top layer.  With its own GPU
FK background (FP64, hipFFT) it lies 4.7e-8 from the oracle's run on the
oracle's FK (two independent FP64 FK implementations, hipFFT and the
oracle's radix-2 transform; held to 1e-6); the FP32 tiled path then lies
~1e-6 from it.
"""
import numpy as np
import pytest

from fdfk2d_amd import lib, refstate
from oracle import pyoracle as po
from tests._cases import Case, oracle_run, rel_l2

pytestmark = pytest.mark.gpu


def _oracle64(case):
    bg = po.background(case.nx, case.nz, case.npml, case.izt, case.nbd, case.nt, case.src_is_p, 0.0, 0.0, case.dx,
                       case.dz, case.dt, case.ds, case.f0, case.rayp, case.bazs, case.fk, case.fk, fp64=True)
    st = po.state64(case.state())
    ox, oz = oracle_run(case, st, bg, case.nt + 1)
    return bg, st, ox, oz


def _solver(case, precision):
    return lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                      rx=case.rx, rz=case.rz, precision=precision, vpml=refstate.vp_max(case.c),
                      pml_top=case.pml_top)


@pytest.mark.parametrize("pml_top,src_is_p,norder", [(False, True, 3), (False, False, 2), (True, True, 3)])
def test_fp64_path_equals_fp64_oracle_bit_for_bit(pml_top, src_is_p, norder):
    print(f"case: FP64 path vs FP64 oracle, 121x60, top={pml_top}, P={src_is_p}, order {2 * norder}", flush=True)
    case = Case(nt=300, pml_top=pml_top, src_is_p=src_is_p, norder=norder, nrcv=13)
    bg, st, ox, oz = _oracle64(case)
    s = _solver(case, lib.FDFK_FP64)
    s.set_background(bg)
    s.run(case.nt + 1)
    gx, gz = s.seismograms()
    ux, uz = s.get_state()
    s.close()
    assert gx.dtype == np.float64 and np.abs(ox).max() > 0
    assert np.array_equal(gx, ox) and np.array_equal(gz, oz)
    assert np.array_equal(ux, st["ux"]) and np.array_equal(uz, st["uz"])


def test_fp64_path_with_gpu_fk_and_fp32_path_against_it():
    print("case: FP64 path with its own FK vs FP64 oracle; FP32 tiled path vs FP64 path, 121x60 P", flush=True)
    case = Case(nt=300, nrcv=13)
    _, _, ox, oz = _oracle64(case)
    s = _solver(case, lib.FDFK_FP64)
    s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
    s.run(case.nt + 1)
    g64 = s.seismograms()
    s.close()
    e_fk = max(rel_l2(g64[0], ox), rel_l2(g64[1], oz))
    print(f"FP64 path on the GPU FK vs FP64 oracle on its FK: rel-L2 {e_fk:.2e}", flush=True)
    assert e_fk < 1e-6
    s = _solver(case, lib.FDFK_FP32)
    s.compute_background(case.fk, case.fk, case.src_is_p, case.f0, case.ds, case.rayp, case.bazs)
    s.run(case.nt + 1)
    g32 = s.seismograms()
    s.close()
    e = max(rel_l2(g32[0], g64[0]), rel_l2(g32[1], g64[1]))
    print(f"FP32 tiled path vs FP64 path: rel-L2 {e:.2e}", flush=True)
    assert e < 1e-5


def test_fp64_refusals():
    case = Case(nt=10)
    s = _solver(case, lib.FDFK_FP64)
    b = lib.Batch(case.norder)
    with pytest.raises(lib.FdfkError):
        b.add(s)
    b.close()
    s.close()
