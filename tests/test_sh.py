"""SH (transverse) motion of this project's formulation, pinned on the CPU
oracle (no counterpart in the reference, which keeps the in-plane P-SV
motion only):

* SH FK closed forms: in a half-space the free surface doubles the incident
  wave at any incidence (without a free surface the surface trace is the
  incident wave); over a layer the surface reverberation arrives 2 H eta_1
  after the direct wave with the amplitude ratio of the interface's SH
  reflection coefficient (mu1 eta1 - mu2 eta2) / (mu1 eta1 + mu2 eta2);
* FD == FK: with the FD model equal to the FK model the surface records
  equal the FK surface response (free surface and absorbing top layer) and
  the padding holds only a small scattered field;
* a pulse leaves through the absorbing layers; long runs at 0.97 CFL in a
  cell-wise random model stay bounded.
"""
import numpy as np
import pytest

from fdfk2d_amd import refstate
from oracle import pyoracle as po
from tests._cases import fk_model, layered_grid, rel_l2

P = 6.0 * 180.0 / (np.pi * 6371.0e3)


def test_sh_fk_half_space_free_surface_doubles_incident_wave():
    fk = fk_model(((0.0, 6000.0, 3450.0),))
    dt, f0, ds, t0 = 0.005, 1.0, 1e-3, 1.5
    t = np.arange(1600) * dt
    eta = np.sqrt(1 / 3450.0 ** 2 - P ** 2)
    s = ds * np.exp(-(np.pi * f0 * (t - t0 - eta * 12000.0)) ** 2)
    assert rel_l2(po.fk_traces_sh(fk, P, f0, ds, dt, 1600, np.array([0.0]), 12000.0, t0, True)[0], 2 * s) < 1e-6
    assert rel_l2(po.fk_traces_sh(fk, P, f0, ds, dt, 1600, np.array([0.0]), 12000.0, t0, False)[0], s) < 1e-6


def test_sh_fk_layer_reverberation_delay_and_amplitude():
    H, dt = 10000.0, 0.005
    fk = fk_model(((0.0, 6000.0, 3000.0), (H, 8000.0, 4500.0)))
    m1, m2 = float(fk[0, 3]) * 3000.0 ** 2, float(fk[1, 3]) * 4500.0 ** 2
    e1, e2 = np.sqrt(1 / 3000.0 ** 2 - P ** 2), np.sqrt(1 / 4500.0 ** 2 - P ** 2)
    R12 = (m1 * e1 - m2 * e2) / (m1 * e1 + m2 * e2)
    u = po.fk_traces_sh(fk, P, 0.5, 1e-3, dt, 6000, np.array([0.0]), 30000.0, 3.0, True)[0]
    i1 = int(np.argmax(np.abs(u[:3000])))
    lag = 2 * H * e1 / dt
    lo = int(i1 + lag - 50)
    i2 = lo + int(np.argmax(np.abs(u[lo: lo + 100])))
    assert abs((i2 - i1) * dt - 2 * H * e1) <= dt
    assert abs(u[i2] / u[i1] - R12) < 2e-3


def _fd_vs_fk(top, layers, fkl, nt=700, nx=151, nz=60, npml=20, dx=200.0, dt=0.01, f0=1.0):
    izt = 1 - npml if top else 1
    vp, vs = layered_grid(nx, nz, npml, dx, dx, layers, izt=izt)
    c = refstate.material_law(vp, vs)
    st = po.sh_state(nx, nz, npml, 3, dx, dx, dt, nt, c, pml_top=top)
    nbd = 5
    bazs = np.full(nx + 2 * npml, np.pi)
    fk = fk_model(fkl)
    bg = po.background_sh(nx, nz, npml, izt, nbd, nt, 0.0, 0.0, dx, dx, dt, 1e-3, f0, P, bazs, fk, fk)
    rx = np.linspace(0, (nx - 1) * dx, 7).astype(np.float32)
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    tr = []
    for it in range(nt + 1):
        tr.append([po.loadseis(nx, nz, npml, izt, 0.0, 0.0, dx, dx, float(r), 0.0, st["uy"], st["uy"])[0] for r in rx])
        po.updatedispl_sh(st, it, bg)
    tr = np.asarray(tr).T
    xe = (1 - nbd - 1) * dx
    u = po.fk_traces_sh(fk, P, f0, 1e-3, dt, nt + 400, np.array([0.0]), (nz + nbd - 1) * dx, 1.5 / f0, not top)[0]
    n = np.arange(nt + 1)
    errs = [rel_l2(tr[i], np.interp(n - P * (r - xe) / dt, np.arange(u.size), u)) for i, r in enumerate(rx)]
    uyf = st["uy"].reshape(NX, NZ, 2)[..., 0]
    peak = np.abs(tr).max()
    return max(errs), np.abs(uyf[: npml - nbd]).max() / peak, np.abs(uyf[:, : 1 - izt]).max() / peak if top else 0.0


@pytest.mark.parametrize("top", [False, True])
def test_sh_hybrid_self_consistency_layered(top):
    # the FD material change between rows at 5800 and 6000 m sits at 5900 m
    err, side, topm = _fd_vs_fk(top, ((0.0, 6000.0, 3450.0), (6000.0, 8000.0, 4450.0)),
                                ((0.0, 6000.0, 3450.0), (5900.0, 8000.0, 4450.0)))
    assert err < 0.005 and side < 1e-3 and topm < 5e-3, (err, side, topm)


def _sh_energy(st, c, NX, NZ, dt):
    u = st["uy"].reshape(NX, NZ, 2)
    rho = c.reshape(NX, NZ, 3)[..., 0].astype(np.float64)
    v = (u[..., 0].astype(np.float64) - u[..., 1]) / dt
    return float((rho * v * v).sum())


@pytest.mark.parametrize("top", [False, True])
def test_sh_pml_absorbs_outgoing_pulse(top):
    nx, nz, npml, norder = 101, 60, 20, 3
    izt = 1 - npml if top else 1
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    dx, dt = 200.0, 0.01
    vp, vs = layered_grid(nx, nz, npml, dx, dx, [(0.0, 6000.0, 3450.0)], izt=izt)
    c = refstate.material_law(vp, vs)
    st = po.sh_state(nx, nz, npml, norder, dx, dx, dt, 2000, c, pml_top=top)
    X, Z = np.meshgrid((np.arange(NX) - npml - 50) * dx, (np.arange(NZ) + izt - 1 - 30) * dx, indexing="ij")
    pulse = (1e-3 * np.exp(-(X * X + Z * Z) / (2 * 1200.0 ** 2))).astype(np.float32)
    u = st["uy"].reshape(NX, NZ, 2)
    u[..., 0] = pulse
    u[..., 1] = pulse
    e0 = None
    for it in range(1500):
        po.updatedispl_sh(st, it)
        if it == 5:
            e0 = _sh_energy(st, c, NX, NZ, dt)
    ef = _sh_energy(st, c, NX, NZ, dt)
    assert np.isfinite(ef) and ef < 1e-3 * e0


def test_sh_long_run_stable_at_cfl_limit():
    nx, nz, npml, norder = 61, 40, 12, 3
    NX, NZ = nx + 2 * npml, nz + npml
    dx = 200.0
    rng = np.random.default_rng(13)
    vp = rng.uniform(5500.0, 8000.0, (NX, NZ)).astype(np.float32)
    vs = (vp / 1.75).astype(np.float32)
    vs[:, :6] *= np.float32(0.6)
    c = refstate.material_law(vp, vs)
    rc, dts = po.check_stability(nx, nz, npml, 1, norder, dx, dx, 0.01, po.fd_coefs(norder)[2], c)
    st = po.sh_state(nx, nz, npml, norder, dx, dx, 0.97 * dts, 3000, c)
    st["uy"].reshape(NX, NZ, 2)[npml + 30, 3, :] = 1e-3
    for it in range(4000):
        po.updatedispl_sh(st, it)
    m = float(np.abs(st["uy"]).max())
    assert np.isfinite(m) and m < 1e-3
