"""Storage precision at the size of the single large solve (8001 x 2000,
dx = dz = 50 m, dt = 0.003 s, 40-cell PML, order 6, 401 receivers): the FP32
tiled path and FP32 with FP16 or BF16 materials against the FDFK_FP64 run of the same
workload (the general per-cell kernels in double, themselves equal to the
FP64 oracle bit for bit on small grids, tests/test_gpu_fp64_path.py).

6001 iterations (18 s: the incident P reaches the surface of the first part
of the profile).  Bar: rel-L2 <= 1e-4 over every trace and over the P
receiver functions of the receivers the wave reached (the FP64 run's
deconvolution spike picks replayed, so the comparison measures the traces'
precision and not the deconvolution's discrete decisions).  Measured with
8001 iterations (bench.py --c4 8000): traces 1.5e-5 for both modes.
"""
import numpy as np
import pytest

from fdfk2d_amd import configs, lib, rf

pytestmark = pytest.mark.gpu

NSTEPS = 6000


def _run(cfg, precision):
    s = lib.Solver(cfg["nx"], cfg["nz"], cfg["npml"], cfg["norder"], cfg["dx"], cfg["dz"], cfg["dt"], NSTEPS,
                   cfg["c"], rx=cfg["rx"], rz=cfg["rz"], precision=precision)
    try:
        s.compute_background(cfg["fk_left"], cfg["fk_right"], cfg["src_is_p"], cfg["f0"], cfg["ds"], cfg["rayp"],
                             cfg["bazs"], block_steps=512)
        s.run(NSTEPS + 1)
        sx, sz = s.seismograms()
    finally:
        s.close()
    return np.asarray(sx, np.float64), np.asarray(sz, np.float64)


def test_storage_modes_against_fp64_at_full_size():
    print(f"case: 8001x2000, {NSTEPS + 1} iterations, FP64 / FP32 / FP16 and BF16 materials", flush=True)
    cfg = configs.c2(nt=NSTEPS)
    x64, z64 = _run(cfg, lib.FDFK_FP64)
    peak = np.abs(z64).max()
    assert peak > 0 and np.isfinite(x64).all()
    sel = [i for i in range(0, x64.shape[0], 10) if np.abs(z64[i]).max() > 1e-2 * peak]
    assert len(sel) >= 5
    picks, rf64 = {}, {}
    for i in sel:
        rec = []
        rf64[i] = rf.prf(x64[i], z64[i], cfg["dt"], record=rec)[0]
        picks[i] = rec
    ref = np.concatenate([x64.ravel(), z64.ravel()])
    for name, prec in (("fp32", lib.FDFK_FP32), ("fp16 materials", lib.FDFK_FP16), ("bf16 materials", lib.FDFK_BF16)):
        x, z = _run(cfg, prec)
        e = np.linalg.norm(np.concatenate([x.ravel(), z.ravel()]) - ref) / np.linalg.norm(ref)
        num = den = 0.0
        for i in sel:
            r = rf.prf(x[i], z[i], cfg["dt"], picks=picks[i])[0]
            num += float(np.sum((r - rf64[i]) ** 2))
            den += float(np.sum(rf64[i] ** 2))
        erf = np.sqrt(num / den)
        print(f"{name}: traces rel-L2 {e:.2e}, P-RF rel-L2 {erf:.2e} ({len(sel)} receivers)", flush=True)
        assert e <= 1e-4 and erf <= 1e-4
