"""Receiver functions by iterative time-domain deconvolution.

Restates utils/makeRFitdecon.m:72-228 (Ligorria-Ammon iterative
deconvolution with a Gaussian filter) in numpy, with the MATLAB conventions
kept: conv(..., 'same') is the centred part of the full convolution,
xcorr(R, W, maxlag, 'none') the raw cross-correlation on lags -maxlag..maxlag.
P-RF = radial / vertical with lags -2 .. 0.5*T-2 s (utils/plot_PRF.m:18-24),
S-RF = vertical / radial with lags -6.5 .. 0.6*T-6.5 s (utils/plot_SRF.m:18-23).
"""
import numpy as np


def _conv_same(u, v):
    full = np.convolve(u, v)
    s = len(v) // 2
    return full[s: s + len(u)]


def _xcorr(r, w, maxlag):
    n = len(r)
    full = np.correlate(r, w, "full")          # lag k-(n-1)
    return full[n - 1 - maxlag: n + maxlag], np.arange(-maxlag, maxlag + 1)


def _normpdf(x, mu, sigma):
    return np.exp(-0.5 * ((x - mu) / sigma) ** 2) / (sigma * np.sqrt(2.0 * np.pi))


def itdecon(uin, win, dt, minlag, maxlag, tshift=0.0, f0=8.0, itmax=100, minderr=1e-3, picks=None,
            record=None):
    """makeRFitdecon(UIN, WIN, DT, MINLAG, MAXLAG, TSHIFT, F0, ITMAX, MINDERR).

    Returns (rf, lag_times, rms).  The spike position of each iteration (the
    argmax of the correlation) and the iteration count are discrete
    decisions: traces within rounding of each other can pick different
    spikes where two correlation peaks nearly tie.  picks (a sequence of
    lag indices) replays given decisions instead, one iteration per pick;
    record (a list) receives the decisions taken."""
    uin = np.asarray(uin, np.float64).ravel()
    win = np.asarray(win, np.float64).ravel()
    ntin = uin.size
    if ntin == 0:
        raise ValueError("Zero length numerator")
    endn = max(int(round(maxlag / dt)), 0)
    begn = max(int(round(-minlag / dt)), 0)
    while begn + endn > ntin:
        if begn > 0:
            minlag *= 0.5
            begn = max(int(round(-minlag / dt)), 0)
        if endn > 0:
            maxlag *= 0.5
            endn = max(int(round(maxlag / dt)), 0)
    W = np.zeros(ntin)
    W[begn: ntin - endn] = win[begn: ntin - endn]
    sigma = 1.0 / (np.sqrt(2.0) * f0)
    gauss = dt * _normpdf(dt * np.arange(ntin), dt * round(0.5 * ntin - 0.5), sigma)
    U = _conv_same(uin, gauss)
    W = _conv_same(W, gauss)
    R = U.copy()
    inv_pu = 1.0 / np.sum(U * U)
    inv_pw = 1.0 / np.sum(_conv_same(win, gauss) ** 2)
    lag1, lag2 = -begn, endn
    mx = max(begn, endn)
    P0 = np.zeros(2 * mx + 1)
    it = 0
    sumsq_i = 1.0
    derr = 100.0 * np.sum(U * U) + minderr
    rms = []
    while (abs(derr) > minderr and it < itmax) if picks is None else it < len(picks):
        it += 1
        RW, lags = _xcorr(R, W, mx)
        RW = RW * inv_pw
        if picks is None:
            ok = np.nonzero((lags >= lag1) & (lags <= lag2))[0]
            idx = ok[np.argmax(np.abs(RW[ok]))]
        else:
            idx = int(picks[it - 1])
        if record is not None:
            record.append(int(idx))
        P0[idx] += RW[idx]
        rfp = _conv_same(P0, gauss)
        UP = _conv_same(win, rfp)
        R = U - UP
        sumsq = np.sum(R * R) * inv_pu
        rms.append(sumsq)
        derr = 100.0 * (sumsq_i - sumsq)
        sumsq_i = sumsq
    rfp = _conv_same(P0, gauss)
    i1, i2 = mx + lag1, mx + lag2          # 0-based idx1-1 .. idx2-1
    shift = int(round(tshift / dt))
    if shift == 0:
        rf = rfp[i1: i2 + 1]
    elif shift > 0:
        i1 -= shift
        rf = np.concatenate([np.zeros(-i1), rfp[: i2 + 1]]) if i1 < 0 else rfp[i1: i2 + 1]
    else:
        i1 -= shift
        i2 = min(rfp.size - 1, i2 - shift)
        rf = rfp[i1: i2 + 1]
    t = (np.arange(rf.size) + lag1 - (shift if shift > 0 and mx + lag1 - shift >= 0 else 0)) * dt
    return rf, t, np.asarray(rms)


def prf(radial, vertical, dt, f0=8.0, picks=None, record=None):
    """P receiver function of one station (utils/plot_PRF.m:18-24)."""
    T = (len(radial) - 1) * dt
    return itdecon(radial, vertical, dt, -2.0, 0.5 * T - 2.0, 0.0, f0, 100, 1e-3, picks, record)


def srf(radial, vertical, dt, f0=8.0, picks=None, record=None):
    """S receiver function of one station (utils/plot_SRF.m:18-23)."""
    T = (len(radial) - 1) * dt
    return itdecon(vertical, radial, dt, -6.5, 0.6 * T - 6.5, 0.0, f0, 100, 1e-3, picks, record)
