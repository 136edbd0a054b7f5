"""GPU parity against the CPU oracle (same seeded inputs, same formulation).

Tolerances: the receiver interpolation reproduces the oracle's arithmetic
exactly; the stencil kernels evaluate the same operators in FP32 (shared
fluxes in the interior, hardware reciprocals in the harmonic means), so
single steps agree to ~1e-6 (<= 1e-5 max-relative per array) and
seismograms to rel-L2 <= 1e-5.
"""
import os

import numpy as np
import pytest

from fdfk2d_amd import lib, refstate
from oracle import pyoracle as po
from tests._cases import Case, oracle_run, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    return Case(nt=300)


@pytest.fixture(scope="module")
def bg(case):
    return case.oracle_background()


@pytest.fixture(scope="module")
def warm_state(case, bg):
    st = case.state()
    oracle_run(case, st, bg, 180, record=False)
    return st


def _max_rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def _strip_ok(a, b, utot):
    """Split-field strip arrays: 1e-5 of the array's own scale, plus a floor
    of 4 FP32 ulps of the total-field scale.  The first padding row under the
    hybrid interface holds the scattered field (~1e-5 of the total field
    here); the oracle forms it as L[u] - L[u0], a difference of total-field
    sized stencil sums, so it is defined only to the total field's FP32
    resolution."""
    err = np.abs(a.astype(np.float64) - b).max()
    return err <= 1e-5 * np.abs(b).max() + 4.0 * np.finfo(np.float32).eps * utot


def test_single_step_compat_matches_oracle(case, bg, warm_state):
    print("case: single step via updatedispl_, it=180", flush=True)
    so = refstate.copy_state(warm_state)
    sg = refstate.copy_state(warm_state)
    so["bg"] = bg
    sg["bg"] = bg
    po.updatedispl(so, 180)
    lib.updatedispl(sg, 180)
    lib.lib().fdfk_compat_release()
    for k in ("ux", "uz"):
        assert _max_rel(sg[k], so[k]) < 1e-5, k
    utot = max(np.abs(so["ux"]).max(), np.abs(so["uz"]).max())
    for i, (a, b) in enumerate(zip(sg["bu"][:18], so["bu"][:18])):
        assert _strip_ok(a, b, utot), ("bu", i)
    for i in (0, 1, 2, 3, 8, 9, 10, 11):
        assert _max_rel(sg["bp"][i], so["bp"][i]) < 1e-5, ("bp", i)


def test_seismograms_match_oracle(case, bg):
    print("case: 300 iterations, handle API, 31 receivers", flush=True)
    st = case.state()
    ox, oz = oracle_run(case, st, bg, case.nt + 1)
    s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt,
                   case.c, rx=case.rx, rz=case.rz)
    s.set_background(bg)
    s.run(case.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert np.abs(oz).max() > 1e-4
    assert rel_l2(gx, ox) < 1e-5
    assert rel_l2(gz, oz) < 1e-5


def test_compat_sequence_matches_handle(case, bg):
    print("case: 40 compat steps vs handle", flush=True)
    st = case.state()
    st["bg"] = bg
    for it in range(40):
        lib.updatedispl(st, it)
    lib.lib().fdfk_compat_release()
    s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c)
    s.set_background(bg)
    s.run(40)
    ux, uz = s.get_state()
    s.close()
    assert np.array_equal(ux, st["ux"]) and np.array_equal(uz, st["uz"])


def test_fused_step_equals_separate_launches(case, bg, monkeypatch):
    """k_step runs the interior and frame tiles of a step in one launch; with
    FDFK_FUSED=0 the same tile bodies run as k_interior2 then k_frame.  The
    arithmetic is the same, so fields and seismograms agree bit for bit."""
    print("case: 120 steps, fused launch vs separate launches", flush=True)
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("FDFK_FUSED", fused)
        s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt,
                       case.c, rx=case.rx, rz=case.rz)
        s.set_background(bg)
        s.run(120)
        out.append(s.get_state() + s.seismograms())
        s.close()
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_partial_band_segments_match_oracle():
    """nz = 262: the interior rows 6..262 are one band of 246-row tiles plus
    an 11-row band, which the interior cuts into one-wave segments (each
    wave of a block marching its own columns).  Seismograms (900 iterations:
    the incident P reaches the surface from the 52 km deep ring) and final
    fields against the oracle, fused and separate launches."""
    print("case: 121 x 262, 900 iterations, partial-band segments", flush=True)
    c = Case(nz=262, nt=900)
    bg = c.oracle_background()
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    for fused in ("1", "0"):
        os.environ["FDFK_FUSED"] = fused
        try:
            s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz)
            s.set_background(bg)
            s.run(c.nt + 1)
            gx, gz = s.seismograms()
            ux, uz = s.get_state()
            s.close()
        finally:
            os.environ.pop("FDFK_FUSED", None)
        assert np.abs(oz).max() > 1e-4
        assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5, fused
        # both time levels, the reference's ux(1:2, ...) layout
        assert rel_l2(ux, np.ravel(st["ux"])) < 1e-5 and rel_l2(uz, np.ravel(st["uz"])) < 1e-5, fused


@pytest.mark.parametrize("norder", [1, 2, 4, 5, 6, 7, 8])
def test_fd_orders_match_oracle(norder):
    """FD orders 2..16 (the reference's norder = order/2, src/Input.f90:130;
    its examples/test1 runs order 16) on a 121 x 262 grid: after 200
    iterations the incident wave fills the lower part of the model, across
    the last band of interior rows that runs as one-wave segments.  A 20-cell
    PML from norder 5 on (the hybrid ring takes 2*norder-1 cells of it).
    Both time levels against the oracle; fused and separate launches bit for
    bit."""
    print("case: norder %d, 121 x 262, 200 iterations" % norder, flush=True)
    c = Case(norder=norder, nz=262, nt=200, npml=15 if norder <= 4 else 20)
    bg = c.oracle_background()
    st = c.state()
    oracle_run(c, st, bg, c.nt + 1, record=False)
    out = []
    for fused in ("1", "0"):
        os.environ["FDFK_FUSED"] = fused
        try:
            s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz)
            s.set_background(bg)
            s.run(c.nt + 1)
            out.append(s.get_state())
            s.close()
        finally:
            os.environ.pop("FDFK_FUSED", None)
    ux, uz = out[0]
    assert np.abs(ux).max() > 1e-5
    assert rel_l2(ux, np.ravel(st["ux"])) < 1e-5 and rel_l2(uz, np.ravel(st["uz"])) < 1e-5
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("norder", [1, 2, 3, 4, 8])
def test_minimal_grids_match_oracle(norder):
    """The smallest grid fdfk_create accepts for each order: nx = 2*nbd + 2,
    nz = nbd + 2*norder + 1 and npml = nbd (the hybrid ring fills the whole
    PML, no damped cells), nbd = 2*norder - 1.  One tile of every kind,
    clipped at both ends.  Both time levels after 61 iterations against the
    oracle; fused and separate launches bit for bit."""
    nbd = 2 * norder - 1
    nx, nz, npml = 2 * nbd + 2, nbd + 2 * norder + 1, nbd
    print("case: minimal grid norder %d, %d x %d, npml %d" % (norder, nx, nz, npml), flush=True)
    c = Case(norder=norder, nx=nx, nz=nz, npml=npml, nt=60, nrcv=3)
    bg = c.oracle_background()
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    out = []
    for fused in ("1", "0"):
        os.environ["FDFK_FUSED"] = fused
        try:
            s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz)
            s.set_background(bg)
            s.run(c.nt + 1)
            out.append(s.get_state() + s.seismograms())
            s.close()
        finally:
            os.environ.pop("FDFK_FUSED", None)
    ux, uz, gx, gz = out[0]
    assert np.abs(ux).max() > 0 and np.isfinite(ux).all()
    assert rel_l2(ux, np.ravel(st["ux"])) < 1e-5 and rel_l2(uz, np.ravel(st["uz"])) < 1e-5
    assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("norder", [3, 8])
def test_anisotropic_grid_spacing_matches_oracle(norder):
    """dx = 200 m, dz = 120 m: every operator scales its x and z parts by
    their own spacing, which equal spacings would not distinguish.
    Seismograms of 401 iterations and both final time levels against the
    oracle."""
    print("case: dx 200 m, dz 120 m, norder %d, 401 iterations" % norder, flush=True)
    c = Case(norder=norder, npml=20, dx=200.0, dz=120.0, dt=0.008, nt=400)
    bg = c.oracle_background()
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=c.rx, rz=c.rz)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    ux, uz = s.get_state()
    s.close()
    assert np.abs(oz).max() > 1e-4
    assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5
    assert rel_l2(ux, np.ravel(st["ux"])) < 1e-5 and rel_l2(uz, np.ravel(st["uz"])) < 1e-5


def test_receivers_off_grid_at_depth_with_origin():
    """Receivers between grid points, at depth and one left of x0 (inside
    the padding), with a nonzero grid origin (x0, z0): the recorder's
    coordinate arithmetic and its clipping at the padded edge
    (src/Submain.f90:3-116) against the oracle's loadseis, which is pinned
    bit for bit to the reference's."""
    print("case: off-grid receivers, x0 = 730 m, z0 = -40 m", flush=True)
    c = Case(nt=300)
    x0, z0 = 730.0, -40.0
    rx = np.array([x0 - 1700.0, x0 + 10.0, x0 + 1234.5, x0 + 5111.1, x0 + 12345.6, x0 + 23999.0],
                  np.float32)
    rz = np.array([z0, z0 + 55.5, z0 + 333.3, z0 + 1000.0, z0 + 2500.0, z0 + 120.0], np.float32)
    bg = c.oracle_background()
    st = c.state()
    st["bg"] = bg
    from oracle import pyoracle as po
    ox, oz = [], []
    for it in range(c.nt + 1):
        row = [po.loadseis(c.nx, c.nz, c.npml, 1, x0, z0, c.dx, c.dz, float(a), float(b), st["ux"], st["uz"])
               for a, b in zip(rx, rz)]
        ox.append([r[0] for r in row])
        oz.append([r[1] for r in row])
        po.updatedispl(st, it)
    ox, oz = np.asarray(ox, np.float32).T, np.asarray(oz, np.float32).T
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, x0=x0, z0=z0, rx=rx, rz=rz)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert np.abs(oz).max() > 1e-4 and np.abs(oz[0]).max() > 0
    assert rel_l2(gx, ox) < 1e-5 and rel_l2(gz, oz) < 1e-5


def test_receivers_outside_the_grid_record_zero():
    """Receivers more than three cells outside the padded grid along one or
    both axes (beyond nx + npml and above the surface, far left and deep,
    huge coordinates): the reference's loadseis loops are empty and record 0
    (src/Submain.f90:66,85); the recorder's window must be empty too, not
    wrap to a positive count and read outside the fields.  A receiver inside
    the grid records the same as the oracle's."""
    print("case: receivers outside the padded grid, 121x60, 301 iterations", flush=True)
    c = Case(nt=300)
    far = (c.nx + c.npml + 10) * c.dx
    rx = np.array([far, -(c.npml + 10) * c.dx, 1.0e30, 5000.0, far, 3000.0], np.float32)
    rz = np.array([-10 * c.dz, 2000.0, -1.0e30, 0.0, (c.nz + c.npml + 20) * c.dz, 1.0e9], np.float32)
    bg = c.oracle_background()
    from oracle import pyoracle as po
    st = c.state()
    st["bg"] = bg
    ox = []
    for it in range(c.nt + 1):
        ox.append(po.loadseis(c.nx, c.nz, c.npml, 1, 0.0, 0.0, c.dx, c.dz, 5000.0, 0.0, st["ux"], st["uz"]))
        po.updatedispl(st, it)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, rx=rx, rz=rz)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    outside = [0, 1, 2, 4, 5]
    assert not gx[outside].any() and not gz[outside].any()
    oz = np.array([o[1] for o in ox], np.float32)
    assert np.abs(oz).max() > 1e-5 and rel_l2(gz[3], oz) < 1e-5


def test_recording_interval():
    """nstp = 3 (tstp = 3 dt): nt/3 + 1 samples, taken at iterations 0, 3,
    6, ... before the update (src/FDFK2D.f90:311-317, 497-522), equal to the
    oracle's every-iteration record at those iterations."""
    print("case: nstp 3, 301 iterations", flush=True)
    c = Case(nt=300)
    bg = c.oracle_background()
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, nstp=3, rx=c.rx, rz=c.rz)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert gx.shape == (len(c.rx), c.nt // 3 + 1)
    assert rel_l2(gx, ox[:, ::3]) < 1e-5 and rel_l2(gz, oz[:, ::3]) < 1e-5


def test_autotuned_tile_width_same_results(case, bg):
    """fdfk_autotune times interior tile widths for a geometry and later
    solvers of it use the fastest; a cell's arithmetic does not depend on the
    width, so the seismograms and final state are bit for bit those of the
    occupancy model's width."""
    print("case: autotune on 121 x 60, then 301 iterations", flush=True)
    def run():
        s = lib.Solver(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.nt, case.c,
                       rx=case.rx, rz=case.rz)
        s.set_background(bg)
        s.run(case.nt + 1)
        out = s.seismograms() + s.get_state()
        s.close()
        return out
    before = run()
    t = lib.autotune(case.nx, case.nz, case.npml, case.norder, case.dx, case.dz, case.dt, case.c,
                     rx=case.rx, rz=case.rz, steps=5)
    print("autotune:", t, flush=True)
    assert 4 <= t["lx"] <= case.nx + 2 * case.npml and t["ms"] > 0 and t["ms"] <= t["model_ms"]
    after = run()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)


def _bf16(x):
    """float32 -> BF16 (nearest even, finite values) -> float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _round_materials(c, nx, nz, npml, nbd, kind="fp16"):
    """The library's 16-bit material storage restated: per row a profile, the
    midrange (min + max) / 2 in FP32 of the row's values over the undamped
    columns 1-nbd..nx+nbd (rows below nz+nbd take row nz+nbd's); lam and mu
    become profile + FP16 or BF16(residual / scale) * scale, one power-of-two
    scale per field with |residual| / scale < 2^15 over the undamped cells,
    nearest even.  (The damped cells are copies of undamped ones on the
    device.)"""
    NX, NZ = nx + 2 * npml, nz + npml
    flat = np.array(c, np.float32).ravel().copy()
    x0, x1, zb = npml - nbd, npml + nx + nbd, nz + nbd      # 0-based column range [x0, x1), rows [0, zb)
    for k in (1, 2):
        v = flat[k::3].reshape(NX, NZ)
        und = v[x0:x1, :zb]
        prof = np.float32(0.5) * (und.min(axis=0) + und.max(axis=0))
        prof = np.concatenate([prof, np.full(NZ - zb, prof[-1], np.float32)]).astype(np.float32)
        m = float(np.abs(und - prof[None, :zb]).max())
        sc = np.float32(np.ldexp(1.0, np.frexp(m)[1] - 15)) if m > 0 else np.float32(1.0)
        r = ((v - prof[None, :]) * (np.float32(1.0) / sc)).astype(np.float32)
        h = _bf16(r) if kind == "bf16" else np.float16(r).astype(np.float32)
        flat[k::3] = (prof[None, :] + h * sc).astype(np.float32).reshape(-1)
    return flat.reshape(np.shape(c))


@pytest.mark.parametrize("norder,kind", [(3, "fp16"), (8, "fp16"), (3, "bf16"), (8, "bf16")])
def test_fp16_materials_equal_fp32_on_rounded_materials(norder, kind):
    """FDFK_FP16 / FDFK_BF16: the interior tiles read lam and mu as per-row
    FP32 profiles plus FP16 or BF16 residuals; every kernel sees the same
    rounded values, so the run equals, bit for bit, an FP32 run on materials
    rounded the same way on the host.  The rounding itself moves the
    seismograms by about the 16-bit resolution of the lateral variation."""
    print("case: %s materials, norder %d, 121 x 60, 301 iterations" % (kind, norder), flush=True)
    prec = lib.FDFK_BF16 if kind == "bf16" else lib.FDFK_FP16
    c = Case(nt=300, norder=norder, npml=15 if norder <= 4 else 20)
    bg = c.oracle_background()
    vpml = refstate.vp_max(c.c)     # the PML profile from the unrounded model in every run
    def run(mat, precision):
        s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, mat, rx=c.rx, rz=c.rz,
                       precision=precision, vpml=vpml)
        s.set_background(bg)
        s.run(c.nt + 1)
        out = s.seismograms() + s.get_state()
        s.close()
        return out
    h = run(c.c, prec)
    r = run(_round_materials(c.c, c.nx, c.nz, c.npml, c.nbd, kind), lib.FDFK_FP32)
    for a, b in zip(h, r):
        assert np.array_equal(a, b)
    f = run(c.c, lib.FDFK_FP32)
    d = max(rel_l2(h[0], f[0]), rel_l2(h[1], f[1]))
    print("seismogram rel-L2, %s vs FP32 materials: %.3g" % (kind, d), flush=True)
    assert 0 < d < 1e-2


@pytest.mark.parametrize("prec", [lib.FDFK_FP16, lib.FDFK_BF16])
def test_fp16_materials_not_batched_nor_separate(monkeypatch, prec):
    c = Case(nt=20)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, precision=prec)
    b = lib.Batch(3)
    with pytest.raises(lib.FdfkError, match="not batched"):
        b.add(s)
    b.close()
    s.close()
    monkeypatch.setenv("FDFK_FUSED", "0")
    with pytest.raises(lib.FdfkError, match="fused step only"):
        lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, precision=prec)
