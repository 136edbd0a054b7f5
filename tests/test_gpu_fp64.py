"""The GPU path against the FP64 oracle (oracle/build/liboracle64.so: the
oracle's algorithm with every FP32 quantity in double precision).

The seismograms of the FP32 GPU path, run alone and in a batch (the served
form), must lie within rel-L2 1e-4 of the FP64 run on the same inputs; the
FP32 oracle itself lies about 1e-6 from it (tests/test_oracle_fp64.py).
The 16-bit material storage (per-row FP32 profiles of lambda and mu plus
FP16 or BF16 residuals) is measured against the same run and held to the
same 1e-4: the rounding falls on the lateral variation only.  Each run computes
its own background: the GPU's FK in FP64 with hipFFT, the oracle's in
double precision.
"""
import numpy as np
import pytest

from fdfk2d_amd import lib, refstate
from oracle import pyoracle as po
from tests._cases import Case, oracle_run, rel_l2

pytestmark = pytest.mark.gpu

# bar of the 16-bit material modes against the FP64 run (measured: FP16
# 8.0e-7 / 3.8e-6, BF16 1.9e-6 / 4.6e-5 on the two cases)
BAR16 = {"fp16": 1e-4, "bf16": 1e-4}

CASES = [dict(src_is_p=True, nt=300), dict(src_is_p=False, nt=600)]


def _case(k):
    return Case(**CASES[k])


@pytest.fixture(scope="module")
def fp64_runs():
    out = []
    for k in range(len(CASES)):
        c = _case(k)
        bg = po.background(c.nx, c.nz, c.npml, 1, c.nbd, c.nt, c.src_is_p, 0.0, 0.0, c.dx, c.dz, c.dt,
                           c.ds, c.f0, c.rayp, c.bazs, c.fk, c.fk, fp64=True)
        out.append(oracle_run(c, po.state64(c.state()), bg, c.nt + 1))
    return out


def _solver(c, precision=lib.FDFK_FP32):
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz,
                   precision=precision, vpml=refstate.vp_max(c.c))
    s.compute_background(c.fk, c.fk, c.src_is_p, c.f0, c.ds, c.rayp, c.bazs)
    return s


def _err(got, ref):
    return max(rel_l2(got[0], ref[0]), rel_l2(got[1], ref[1]))


def test_fp32_alone_and_batched_within_1e4_of_fp64(fp64_runs):
    print("case: 121 x 60, P (300 steps) and SV (600 steps), alone and batched, vs FP64 oracle", flush=True)
    cs = [_case(k) for k in range(len(CASES))]
    alone = []
    for c in cs:
        s = _solver(c)
        s.run(c.nt + 1)
        alone.append(s.seismograms())
        s.close()
    b = lib.Batch(3)
    sol = [_solver(c) for c in cs]
    for s in sol:
        b.add(s)
    while b.active:
        b.run(100)
    b.sync()
    batched = [s.seismograms() for s in sol]
    for s in sol:
        s.close()
    b.close()
    for k, ref in enumerate(fp64_runs):
        assert np.abs(ref[1]).max() > 1e-6
        ea, eb = _err(alone[k], ref), _err(batched[k], ref)
        print("case %d: FP32 alone rel-L2 %.2e, batched %.2e against FP64" % (k, ea, eb), flush=True)
        assert ea < 1e-4 and eb < 1e-4


@pytest.mark.parametrize("kind", ["fp16", "bf16"])
def test_fp16_materials_against_fp64(fp64_runs, kind):
    print("case: 121 x 60, %s materials vs FP64 oracle" % kind, flush=True)
    for k, ref in enumerate(fp64_runs):
        c = _case(k)
        s = _solver(c, lib.FDFK_BF16 if kind == "bf16" else lib.FDFK_FP16)
        s.run(c.nt + 1)
        e = _err(s.seismograms(), ref)
        s.close()
        print("case %d: %s materials rel-L2 %.2e against FP64" % (k, kind, e), flush=True)
        assert e < BAR16[kind]


@pytest.mark.parametrize("norder", [1, 4, 8])
def test_fd_orders_within_1e4_of_fp64(norder):
    """FD orders 2, 8 and 16 (P incidence, 121 x 60, 301 iterations): the FP32
    GPU path against the FP64 oracle."""
    print("case: FD order %d, 121 x 60, P, vs FP64 oracle" % (2 * norder), flush=True)
    c = Case(norder=norder, nt=300, npml=max(15, 2 * (2 * norder - 1)))
    bg = po.background(c.nx, c.nz, c.npml, 1, c.nbd, c.nt, c.src_is_p, 0.0, 0.0, c.dx, c.dz, c.dt,
                       c.ds, c.f0, c.rayp, c.bazs, c.fk, c.fk, fp64=True)
    ref = oracle_run(c, po.state64(c.state()), bg, c.nt + 1)
    s = _solver(c)
    s.run(c.nt + 1)
    e = _err(s.seismograms(), ref)
    s.close()
    assert np.abs(ref[1]).max() > 1e-6
    print("order %d: FP32 rel-L2 %.2e against FP64" % (2 * norder, e), flush=True)
    assert e < 1e-4
