"""Seeded sweep of random small configurations against the CPU oracle.

Each seed draws an FD order (2..16), grid shape, grid spacing (dx and dz
apart), PML width, 1-3 layers with a Gaussian anomaly, P or SV incidence,
ray parameter, recording interval and receivers scattered over the model
(between grid points and at depth), with dt at 0.8 of the reference's
stability limit (src/Submain.f90:118-164) and enough iterations (at most
700) for the incident wave to reach the surface.  The seismograms of the GPU path
must match the oracle's record at the same iterations to rel-L2 1e-5, the
bar of the other parity tests.  A second sweep draws, besides, the
absorbing top layer and SH motion for the general-path solver.
"""
import numpy as np
import pytest

from fdfk2d_amd import lib, refstate
from tests._cases import Case, fk_model, layered_grid, oracle_run, rel_l2

pytestmark = pytest.mark.gpu


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    norder = int(rng.integers(1, 9))
    nbd = 2 * norder - 1
    npml = int(rng.integers(max(nbd, 8), max(nbd, 8) + 12))
    nx = int(rng.integers(max(2 * nbd + 2, 40), 150))
    nz = int(rng.integers(max(nbd + 2 * norder + 1, 30), 100))
    dx = float(rng.choice([100.0, 150.0, 200.0]))
    dz = float(rng.choice([100.0, 150.0, 200.0]))
    depth = (nz - 1) * dz
    vs1 = float(rng.uniform(2500.0, 3600.0))
    layers = [(0.0, 1.73 * vs1, vs1)]
    for _ in range(int(rng.integers(0, 3))):
        zt = float(rng.uniform(0.2, 0.9)) * depth
        if zt > layers[-1][0] + 2 * dz:
            vs = layers[-1][2] * float(rng.uniform(1.05, 1.3))
            layers.append((zt, 1.75 * vs, vs))
    anomaly = (float(rng.uniform(0.2, 0.8)) * (nx - 1) * dx, float(rng.uniform(0.2, 0.8)) * depth,
               float(rng.uniform(2.0, 6.0)) * max(dx, dz), float(rng.uniform(-0.1, 0.1)))
    src_is_p = bool(rng.integers(0, 2))
    return dict(norder=norder, nx=nx, nz=nz, npml=npml, dx=dx, dz=dz, layers=layers, anomaly=anomaly,
                src_is_p=src_is_p, rayp_sdeg=float(rng.uniform(4.0, 8.5 if src_is_p else 7.0)),
                nstp=int(rng.choice([1, 1, 2, 3])),
                nrcv=int(rng.integers(3, 21)), f0=float(rng.uniform(0.6, 1.2)), rng=rng)


@pytest.mark.parametrize("seed", range(10))
def test_random_configuration_matches_oracle(seed):
    d = draw(seed)
    rng = d.pop("rng")
    print("case: sweep seed %d: %s" % (seed, {k: v for k, v in d.items() if k not in ("layers", "anomaly")}),
          flush=True)
    vp, vs = layered_grid(d["nx"], d["nz"], d["npml"], d["dx"], d["dz"], d["layers"], d["anomaly"])
    c_mat = refstate.material_law(vp, vs)
    dt = float(np.float32(0.8 * refstate.stable_dt(d["norder"], d["dx"], d["dz"], c_mat)))
    # long enough for the incident wave to cross the model and reach the surface
    vs1 = d["layers"][0][2]
    nt = min(700, int(np.ceil((1.5 / d["f0"] + 2.0 * (d["nz"] - 1) * d["dz"] / vs1 + 1.0) / dt)))
    c = Case(nx=d["nx"], nz=d["nz"], npml=d["npml"], norder=d["norder"], dx=d["dx"], dz=d["dz"], dt=dt,
             nt=nt, layers=tuple(d["layers"]), anomaly=d["anomaly"], src_is_p=d["src_is_p"],
             rayp_sdeg=d["rayp_sdeg"], f0=d["f0"], nrcv=d["nrcv"])
    assert np.array_equal(c.fk, fk_model(d["layers"]))
    c.rx = rng.uniform(0.0, (d["nx"] - 1) * d["dx"], d["nrcv"]).astype(np.float32)
    c.rz = np.where(rng.random(d["nrcv"]) < 0.5, 0.0,
                    rng.uniform(0.0, (d["nz"] - 1) * d["dz"], d["nrcv"])).astype(np.float32)
    bg = c.oracle_background()
    st = c.state()
    ox, oz = oracle_run(c, st, bg, c.nt + 1)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, nstp=d["nstp"], rx=c.rx, rz=c.rz)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    k = d["nstp"]
    assert gx.shape == (d["nrcv"], c.nt // k + 1)
    assert np.abs(oz).max() > 1e-5 and np.isfinite(gx).all()     # the wave has arrived
    assert rel_l2(gx, ox[:, ::k]) < 1e-5 and rel_l2(gz, oz[:, ::k]) < 1e-5


def _sh_record(c, bg8, rx, rz):
    from oracle import pyoracle as po
    st = po.sh_state(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, pml_top=c.pml_top,
                     vpml=refstate.vp_max(c.c))
    tr = []
    for it in range(c.nt + 1):
        tr.append([po.loadseis(c.nx, c.nz, c.npml, c.izt, 0.0, 0.0, c.dx, c.dz, float(a), float(b),
                               st["uy"], st["uy"])[0] for a, b in zip(rx, rz)])
        po.updatedispl_sh(st, it, bg8)
    return np.asarray(tr, np.float32).T


@pytest.mark.parametrize("seed", range(10))
def test_random_general_path_configuration_matches_oracle(seed):
    """The same draw for the general-path solver: the absorbing top layer
    and SH motion (each with probability 1/2, SV-polarised S incidence for
    SH), the deep interior on the tiled kernels and the rest per cell."""
    from oracle import pyoracle as po
    d = draw(100 + seed)
    rng = d.pop("rng")
    top, sh = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
    if sh:
        d["src_is_p"] = False
        d["rayp_sdeg"] = min(d["rayp_sdeg"], 7.0)
    print("case: general-path sweep seed %d: top=%s sh=%s %s" % (
        seed, top, sh, {k: v for k, v in d.items() if k not in ("layers", "anomaly")}), flush=True)
    vp, vs = layered_grid(d["nx"], d["nz"], d["npml"], d["dx"], d["dz"], d["layers"], d["anomaly"])
    dt = float(np.float32(0.8 * refstate.stable_dt(d["norder"], d["dx"], d["dz"], refstate.material_law(vp, vs))))
    vs1 = d["layers"][0][2]
    nt = min(700, int(np.ceil((1.5 / d["f0"] + 2.0 * (d["nz"] - 1) * d["dz"] / vs1 + 1.0) / dt)))
    c = Case(nx=d["nx"], nz=d["nz"], npml=d["npml"], norder=d["norder"], dx=d["dx"], dz=d["dz"], dt=dt,
             nt=nt, layers=tuple(d["layers"]), anomaly=d["anomaly"], src_is_p=d["src_is_p"],
             rayp_sdeg=d["rayp_sdeg"], f0=d["f0"], nrcv=d["nrcv"], pml_top=top)
    c.rx = rng.uniform(0.0, (d["nx"] - 1) * d["dx"], d["nrcv"]).astype(np.float32)
    c.rz = np.where(rng.random(d["nrcv"]) < 0.5, 0.0,
                    rng.uniform(0.0, (d["nz"] - 1) * d["dz"], d["nrcv"])).astype(np.float32)
    s = lib.Solver(c.nx, c.nz, c.npml, c.norder, c.dx, c.dz, c.dt, c.nt, c.c, nstp=d["nstp"], rx=c.rx, rz=c.rz,
                   vpml=refstate.vp_max(c.c), pml_top=top, wave=lib.FDFK_SH if sh else lib.FDFK_PSV)
    k = d["nstp"]
    if sh:
        b = po.background_sh(c.nx, c.nz, c.npml, c.izt, c.nbd, c.nt, 0.0, 0.0, c.dx, c.dz, c.dt, c.ds, c.f0,
                             c.rayp, c.bazs, c.fk, c.fk)
        z = np.zeros_like
        twelve = [b[0], z(b[0]), b[1], z(b[1]), b[2], z(b[2]), b[4], z(b[4]), b[5], z(b[5]), b[6], z(b[6])]
        s.set_background((twelve, [b[3], z(b[3]), b[7], z(b[7])]) if top else twelve)
        oy = _sh_record(c, b, c.rx, c.rz)
        s.run(c.nt + 1)
        gy, _ = s.seismograms()
        s.close()
        assert np.abs(oy).max() > 1e-5 and np.isfinite(gy).all()
        assert rel_l2(gy, oy[:, ::k]) < 1e-5
        return
    bg = c.oracle_background()
    ox, oz = oracle_run(c, c.state(), bg, c.nt + 1)
    s.set_background(bg)
    s.run(c.nt + 1)
    gx, gz = s.seismograms()
    s.close()
    assert np.abs(oz).max() > 1e-5 and np.isfinite(gx).all()
    assert rel_l2(gx, ox[:, ::k]) < 1e-5 and rel_l2(gz, oz[:, ::k]) < 1e-5
