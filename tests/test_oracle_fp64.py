"""The FP64 build of the oracle (oracle/build/liboracle64.so, -DORC_FP64):
the same sources with every FP32 quantity in double precision, the run the
FP32 results are measured against.

Checks that pin it: on the same inputs its seismograms and background agree
with the FP32 build to FP32 rounding (a dropped term or a wrong sign in
either would be of order one), yet differ from it (it really computes in
double: the FP32 build's errors are above FP64 rounding by many orders),
and a plane wave's background in double precision matches the FP32 one
step by step.  The same for the SH formulation (both top boundaries).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests._cases import Case, oracle_run, rel_l2


def _bg(c, fp64):
    return po.background(c.nx, c.nz, c.npml, 1, c.nbd, c.nt, c.src_is_p, 0.0, 0.0, c.dx, c.dz, c.dt,
                         c.ds, c.f0, c.rayp, c.bazs, c.fk, c.fk, fp64=fp64)


def test_fp64_background_matches_fp32():
    c = Case(nx=61, nz=40, npml=12, nt=200)
    b32, b64 = _bg(c, False), _bg(c, True)
    assert all(a.dtype == np.float64 for a in b64)
    for a, b in zip(b32, b64):
        assert np.abs(b).max() > 0
        assert rel_l2(a, b) < 1e-6
    # the FP64 build computes in double: its arrays are not the FP32
    # build's values widened
    assert any(not np.array_equal(a.astype(np.float64), b) for a, b in zip(b32, b64))


def test_fp64_run_agrees_with_fp32_to_fp32_rounding():
    for src_is_p, nt in ((True, 300), (False, 500)):
        c = Case(nx=81, nz=40, npml=12, nt=nt, src_is_p=src_is_p, nrcv=9, f0=1.5)
        x32, z32 = oracle_run(c, c.state(), _bg(c, False), c.nt + 1)
        st64 = po.state64(c.state())
        x64, z64 = oracle_run(c, st64, _bg(c, True), c.nt + 1)
        assert x64.dtype == np.float64 and st64["ux"].dtype == np.float64
        assert np.abs(z64).max() > 1e-6
        ex, ez = rel_l2(x32, x64), rel_l2(z32, z64)
        print("FP32 oracle vs FP64 oracle, P %s: rel-L2 x %.2e z %.2e" % (src_is_p, ex, ez))
        assert ex < 1e-5 and ez < 1e-5
        assert ex > 1e-9 and ez > 1e-9


@pytest.mark.parametrize("top", [False, True])
def test_fp64_sh_run_agrees_with_fp32_to_fp32_rounding(top):
    from fdfk2d_amd import refstate
    c = Case(nx=81, nz=40, npml=12, nt=300, src_is_p=False, pml_top=top)
    out = {}
    for fp64 in (False, True):
        b = po.background_sh(c.nx, c.nz, c.npml, c.izt, c.nbd, c.nt, 0.0, 0.0, c.dx, c.dz, c.dt, c.ds, c.f0,
                             c.rayp, c.bazs, c.fk, c.fk, fp64=fp64)
        st = po.sh_state(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, pml_top=top,
                         vpml=refstate.vp_max(c.c))
        if fp64:
            st = po.state64(st)
            assert all(x.dtype == np.float64 for pair in st["bu"] for x in pair)
        for it in range(c.nt + 1):
            po.updatedispl_sh(st, it, b)
        out[fp64] = st["uy"]
    assert out[True].dtype == np.float64 and np.abs(out[True]).max() > 1e-6
    e = rel_l2(out[False], out[True])
    assert 1e-9 < e < 1e-5, e
