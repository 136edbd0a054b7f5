"""Seeded sweep of random small configurations against the CPU oracle.

Each seed draws an FD order (2..16), grid shape, grid spacing (dx and dz
apart), PML width, 1-3 layers with a Gaussian anomaly, P or SV incidence,
ray parameter, recording interval and receivers scattered over the model
(between grid points and at depth), with dt at 0.8 of the reference's
stability limit (src/Submain.f90:118-164) and enough iterations (at most
700) for the incident wave to reach the surface.  The seismograms of the GPU path
must match the oracle's record at the same iterations to rel-L2 1e-5, the
bar of the other parity tests.
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
