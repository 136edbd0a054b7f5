"""Pins of the oracle's FK background (the binary-only libFK3D role) by
closed forms and invariants that do not reuse its propagator algebra:

* traction-free surface: stresses built by finite differences from the
  output traces vanish at z = 0;
* layer over half-space: the conversion delays of the published P receiver
  function (images/PRF.png with examples/test2 inputs) Ps = 4.397 s,
  PpPs = 15.666 s, PsPs = 20.064 s, PsSs = 24.46 s;
* incident amplitude and causality in the half-space;
* plane-wave translation invariance of the twelve ring arrays.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests._cases import fk_model

P_TEST2 = 4.798 * 180.0 / (np.pi * 6371.0e3)   # s/m (Source.dat s/deg -> s/m, Input.f90:146)


def _stress_residual(layers, src_is_p, p, f0=1.0):
    fk = fk_model(layers)
    vp, vs, rho = float(fk[0, 1]), float(fk[0, 2]), float(fk[0, 3])
    lam, mu = rho * (vp * vp - 2 * vs * vs), rho * vs * vs
    h = 2.0
    z = np.array([0.0, h, 2 * h])
    dt = 0.002
    ux, uz = po.fk_traces(fk, src_is_p, p, f0, 1e-3, dt, 8000, z, 60e3, 1.5 / f0)
    dudt_x = np.gradient(ux[0], dt)
    dudt_z = np.gradient(uz[0], dt)
    dzx = (-3 * ux[0] + 4 * ux[1] - ux[2]) / (2 * h)
    dzz = (-3 * uz[0] + 4 * uz[1] - uz[2]) / (2 * h)
    # plane wave u(x, z, t) = g(z, t - p x): d/dx = -p d/dt
    szz = (lam + 2 * mu) * dzz + lam * (-p * dudt_x)
    sxz = mu * (dzx + (-p * dudt_z))
    scale = rho * vp * np.abs(np.hypot(dudt_x, dudt_z)).max()
    return np.abs(szz).max() / scale, np.abs(sxz).max() / scale, np.abs(uz[0]).max()


@pytest.mark.parametrize("src_is_p", [True, False])
def test_surface_is_traction_free(src_is_p):
    layers = [(0.0, 6000.0, 3450.0), (35000.0, 8000.0, 4450.0)]
    rs, rx, amp = _stress_residual(layers, src_is_p, P_TEST2)
    assert amp > 1e-4
    assert rs < 2e-3 and rx < 2e-3


def test_halfspace_incident_amplitude_and_causality():
    fk = fk_model([(0.0, 8000.0, 4500.0)])
    p = P_TEST2
    dt = 0.005
    zref = 100e3
    ux, uz = po.fk_traces(fk, True, p, 1.0, 1e-3, dt, 4000, np.array([zref]), zref, 1.5)
    amp = np.hypot(ux[0], uz[0])
    t = np.arange(ux.shape[1]) * dt
    # incident P reaches z_ref at t0 = 1.5 s with |u| = ds, nothing before
    first = amp[t < 4.0]
    assert abs(first.max() - 1e-3) < 2e-6
    assert abs(t[np.argmax(first)] - 1.5) <= dt
    assert amp[t < 0.2].max() < 1e-9   # Gaussian tail exp(-(pi*1.3)^2) of ds
    # P particle motion is along the up-going ray: ux/uz = -p/eta
    eta = np.sqrt(1 / 8000.0 ** 2 - p * p)
    k = np.argmax(first)
    assert abs(ux[0, k] / uz[0, k] + p / eta) < 1e-6


def _peaks(x, t, n):
    order = np.argsort(-np.abs(x))
    picks = []
    for i in order:
        if all(abs(t[i] - t[j]) > 1.0 for j in picks):
            picks.append(i)
        if len(picks) == n:
            break
    return picks


def test_layer_over_halfspace_conversion_delays():
    """examples/test2: H = 35 km, Vp 6000, Vs 3450 over 8000/4450, p = 4.798 s/deg."""
    fk = fk_model([(0.0, 6000.0, 3450.0), (35000.0, 8000.0, 4450.0)])
    p = P_TEST2
    dt = 0.005
    ux, uz = po.fk_traces(fk, True, p, 2.0, 1e-3, dt, 12000, np.array([0.0]), 35000.0, 1.0)
    t = np.arange(ux.shape[1]) * dt
    tp = t[np.argmax(np.abs(uz[0]))]
    H = 35000.0
    es, ep = np.sqrt(1 / 3450.0 ** 2 - p * p), np.sqrt(1 / 6000.0 ** 2 - p * p)
    expect = {"Ps": H * (es - ep), "PpPs": H * (es + ep), "PsPs": 2 * H * es}
    assert abs(expect["Ps"] - 4.397) < 2e-3 and abs(expect["PpPs"] - 15.666) < 2e-3
    r = ux[0]
    for name, lag in expect.items():
        w = (t > tp + lag - 0.6) & (t < tp + lag + 0.6)
        k = np.argmax(np.abs(r[w]))
        assert abs(t[w][k] - (tp + lag)) < 0.03, name
    # PsPs arrives with opposite polarity to Ps on the radial component
    wps = (t > tp + expect["Ps"] - 0.3) & (t < tp + expect["Ps"] + 0.3)
    wss = (t > tp + expect["PsPs"] - 0.3) & (t < tp + expect["PsPs"] + 0.3)
    ps = r[wps][np.argmax(np.abs(r[wps]))]
    ss = r[wss][np.argmax(np.abs(r[wss]))]
    assert np.sign(ps) != np.sign(ss)


def test_ring_arrays_translation_invariant():
    nx, nz, npml, nbd, nt = 41, 30, 12, 5, 400
    dx = dz = 200.0
    dt = 0.01
    fk = fk_model([(0.0, 6000.0, 3450.0), (4000.0, 8000.0, 4450.0)])
    bazs = np.full(nx + 2 * npml, np.pi)
    arr = po.background(nx, nz, npml, 1, nbd, nt, True, 0.0, 0.0, dx, dz, dt, 1e-3, 2.0,
                        P_TEST2, bazs, fk, fk)
    # bottom outside strip uxb_pml (nbd, nx+2nbd, nt+1): two columns, same row
    b = arr[10].reshape(nt + 1, nx + 2 * nbd, nbd)
    c1, c2 = 3, 3 + 10
    shift = P_TEST2 * 10 * dx / dt          # samples (baz = pi: wave toward +x, c2 lags)
    tr1, tr2 = b[:, c1, 2].astype(np.float64), b[:, c2, 2].astype(np.float64)
    n = np.arange(nt + 1)
    tr1s = np.interp(n - shift, n, tr1)
    assert np.abs(tr2).max() > 1e-5
    assert np.abs(tr2 - tr1s)[n > shift + 2].max() < 2e-3 * np.abs(tr2).max()
