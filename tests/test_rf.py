"""Receiver functions (restatement of utils/makeRFitdecon.m): closed-form
checks and the published P receiver function of the reference
(images/PRF.png with the examples/test2 model: Ps 4.397 s, PpPs 15.666 s,
PsPs 20.064 s negative)."""
import numpy as np

from fdfk2d_amd import rf
from oracle import pyoracle as po
from tests._cases import fk_model


def test_deconvolution_recovers_spike_train():
    dt = 0.01
    n = 4000
    rng = np.random.default_rng(0)
    w = np.convolve(rng.standard_normal(n), np.exp(-np.arange(-50, 51) ** 2 / 200.0), "same")
    spikes = np.zeros(n)
    spikes[0], spikes[300], spikes[1200] = 1.0, 0.4, -0.25
    # u = w * spikes (causal shifts)
    u = np.zeros(n)
    for k, a in ((0, 1.0), (300, 0.4), (1200, -0.25)):
        u[k:] += a * w[: n - k]
    r, t, rms = rf.itdecon(u, w, dt, -2.0, 18.0, 0.0, 8.0, 200, 1e-5)
    for lag, amp in ((0.0, 1.0), (3.0, 0.4), (12.0, -0.25)):
        k = np.argmin(np.abs(t - lag))
        win = slice(k - 5, k + 6)
        pk = r[win][np.argmax(np.abs(r[win]))]
        g = rf._normpdf(0.0, 0.0, 1 / (np.sqrt(2) * 8.0)) * dt   # peak of the Gaussian filter
        assert abs(pk / g - amp) < 0.05 * abs(amp) + 0.01
    assert rms[-1] < 1e-3


def test_prf_lags_of_published_figure():
    """FK surface seismograms of examples/test2 -> P-RF peak lags."""
    fk = fk_model([(0.0, 6000.0, 3450.0), (35000.0, 8000.0, 4450.0)])
    p = 4.798 * 180.0 / (np.pi * 6371.0e3)
    dt = 0.01
    ux, uz = po.fk_traces(fk, True, p, 1.0, 1e-3, dt, 5000, np.array([0.0]), 35000.0, 2.0)
    # the driver's radial projection with cos(baz) = -1 (event south of a
    # north-pointing profile, as in the examples): R = -ux, Z = uz (z down)
    r, t, _ = rf.prf(-ux[0], uz[0], dt)
    H = 35000.0
    es, ep = np.sqrt(1 / 3450.0 ** 2 - p * p), np.sqrt(1 / 6000.0 ** 2 - p * p)
    want = [(0.0, 1), (H * (es - ep), 1), (H * (es + ep), 1), (2 * H * es, -1)]
    assert abs(want[1][0] - 4.397) < 2e-3 and abs(want[3][0] - 20.064) < 2e-3
    for lag, sign in want:
        w = np.abs(t - lag) < 0.5
        k = np.argmax(np.abs(r[w]))
        assert abs(t[w][k] - lag) < 0.05
        assert np.sign(r[w][k]) == sign


def test_replayed_picks_reproduce_the_rf():
    """Replaying the spike picks a run recorded gives that run's RF bit for
    bit (the decisions are the only discrete part of the deconvolution)."""
    rng = np.random.default_rng(4)
    dt, n = 0.05, 600
    t = np.arange(n) * dt
    vert = np.exp(-((t - 5.0) / 0.4) ** 2) + 0.01 * rng.standard_normal(n)
    rad = 0.6 * np.roll(vert, 40) - 0.3 * np.roll(vert, 120) + 0.01 * rng.standard_normal(n)
    picks = []
    r1, t1, m1 = rf.prf(rad, vert, dt, record=picks)
    r2, t2, m2 = rf.prf(rad, vert, dt, picks=picks)
    assert len(picks) == len(m1) and np.array_equal(r1, r2) and np.array_equal(m1, m2)
