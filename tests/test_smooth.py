"""The model smoother (fdfk2d_amd.smooth, restating utils/gauss_smooth.f90).

The reference program does not compile with this image's Fortran compiler,
so these tests pin the restatement by what the mathematics fixes: normalised
and symmetric Gaussian weights of the stated shape, constants and linear
ramps preserved, a brute-force double-precision evaluation on a tiny grid
(mirror boundaries at both ends of both axes), the window rule, the
triangular weights against a literal simulation of the reference's two
in-place loops, and the command line's output file.
"""
import math

import numpy as np
import pytest

from fdfk2d_amd import io, smooth


def test_gaussian_weights_shape_and_normalisation():
    lx, lz, dx, dz = 3000.0, 1200.0, 500.0, 200.0
    nwx, nwz, w = smooth.weights(lx, lz, dx, dz, 0)
    assert (nwx, nwz) == (math.ceil(3.6 * lx / dx), math.ceil(3.6 * lz / dz))
    assert w.shape == (2 * nwz + 1, 2 * nwx + 1) and w.dtype == np.float32
    assert abs(float(w.sum(dtype=np.float64)) - 1.0) < 1e-6
    assert np.array_equal(w, w[::-1, :]) and np.array_equal(w, w[:, ::-1])
    j = np.arange(-nwx, nwx + 1)[None, :] * dx
    k = np.arange(-nwz, nwz + 1)[:, None] * dz
    g = np.exp(-0.5 * ((j / lx) ** 2 + (k / lz) ** 2))
    assert np.allclose(w, g / g.sum(), rtol=1e-5, atol=1e-9)


def _brute(v, w, nwx, nwz):
    """Independent double-precision evaluation with explicit mirror folds."""
    nx, nz = v.shape

    def fold(i, n):          # 1-based: 0 -> 1, -1 -> 2, n+1 -> n, n+2 -> n-1
        if i < 1:
            i = 1 - i
        if i > n:
            i = 2 * n - i + 1
        return i
    out = np.zeros((nx, nz))
    for i in range(1, nx + 1):
        for j in range(1, nz + 1):
            s = t = 0.0
            for i1 in range(-nwx, nwx + 1):
                for j1 in range(-nwz, nwz + 1):
                    wt = float(w[j1 + nwz, i1 + nwx])
                    s += wt * float(v[fold(i + i1, nx) - 1, fold(j + j1, nz) - 1])
                    t += wt
            out[i - 1, j - 1] = s / t
    return out


@pytest.mark.parametrize("method", [0, 1])
def test_brute_force_tiny_grid(method):
    rng = np.random.default_rng(7 + method)
    v = rng.uniform(2000.0, 4000.0, (7, 6)).astype(np.float32)
    # Gaussian: half-widths 2 along x and 3 along z (an axis mix-up shows);
    # triangular: 2 and 2 (the only shape the reference defines)
    lx, lz = (400.0, 700.0) if method == 0 else (900.0, 900.0)
    nwx, nwz, w = smooth.weights(lx, lz, 1000.0, 1000.0, method)
    assert (nwx, nwz) == ((2, 3) if method == 0 else (2, 2))
    got = smooth.smooth(v, lx, lz, 1000.0, 1000.0, method=method)
    assert np.allclose(got, _brute(v, w, nwx, nwz), rtol=2e-6)


def test_constants_and_ramps():
    nx, nz = 40, 30
    c = np.full((nx, nz), 3456.5, np.float32)
    out = smooth.smooth(c, 1500.0, 1500.0, 500.0, 500.0)
    assert np.allclose(out, c, rtol=2e-6)
    x = np.arange(nx)[:, None] * 500.0
    z = np.arange(nz)[None, :] * 500.0
    ramp = (3000.0 + 0.02 * x + 0.05 * z).astype(np.float32)
    out = smooth.smooth(ramp, 800.0, 800.0, 500.0, 500.0)
    n = math.ceil(3.6 * 800.0 / 500.0)
    core = (slice(n, nx - n), slice(n, nz - n))
    assert np.allclose(out[core], ramp[core], rtol=2e-6)       # symmetric kernel: exact on a ramp
    assert not np.allclose(out[0], ramp[0], rtol=1e-4)          # the mirror bends it at the edge


def test_window_rule():
    rng = np.random.default_rng(3)
    v = rng.uniform(1.0, 2.0, (30, 25)).astype(np.float32)
    full = smooth.smooth(v, 1000.0, 1000.0, 500.0, 500.0)
    part = smooth.smooth(v, 1000.0, 1000.0, 500.0, 500.0, iz0=5, iz1=20, ix0=3, ix1=12)
    inside = (slice(2, 12), slice(4, 20))
    assert np.array_equal(part[inside], full[inside])
    mask = np.ones(v.shape, bool)
    mask[inside] = False
    assert np.array_equal(part[mask], v[mask])


def test_triangular_weights_match_the_in_place_loops():
    """gauss_smooth.f90:334-359 literally: the first loop fills every column
    with (nwz - |k|)/s2z, the second walks columns left to right and sets
    w(k, j) = (nwx - |j|)/s2x * w(j, k) in place."""
    lam, d = 2400.0, 500.0
    nwx, nwz, w = smooth.weights(lam, lam, d, d, 1)
    n = nwx
    s2z = float(np.float32(1.8 * math.ceil(np.float32(lam) / np.float32(d))))
    s2x = s2z
    W = {}
    for j in range(-n, n + 1):
        for k in range(-n, n + 1):
            W[(k, j)] = np.float32((n - abs(k)) / s2z)
    for j in range(-n, n + 1):
        for k in range(-n, n + 1):
            W[(k, j)] = np.float32((n - abs(j)) / s2x * float(W[(j, k)]))
    lit = np.array([[W[(k, j)] for j in range(-n, n + 1)] for k in range(-n, n + 1)], np.float32)
    tot = np.float32(0.0)
    for j in range(2 * n + 1):                 # sum(w): memory order of w(k, j), k fastest
        for k in range(2 * n + 1):
            tot = np.float32(tot + lit[k, j])
    lit = (lit / tot).astype(np.float32)
    assert np.array_equal(w, lit)
    with pytest.raises(ValueError):
        smooth.weights(2400.0, 1200.0, 500.0, 500.0, 1)      # unequal half-widths


def test_command_line(tmp_path):
    rng = np.random.default_rng(5)
    v = rng.uniform(3000.0, 5000.0, (24, 18)).astype(np.float32)
    path = tmp_path / "vp_model.su"
    io.write_su(str(path), v, 200)
    assert smooth.main([str(path), "600", "400", "250", "200", "0", "0", "3", "20", "0"]) == 0
    out, head = io.read_su(str(tmp_path / "vp_model_smooth.su"))
    assert out.shape == v.shape and int(head[57]) == v.shape[1]
    assert int(head[58]) == int(np.int32(200 * 1000).astype(np.int16))
    want = smooth.smooth(v, 600.0, 400.0, 250.0, 200.0, 1, 18, 3, 20, 0)
    assert np.array_equal(out, want)


def test_separable_form_equals_the_window_sum():
    rng = np.random.default_rng(11)
    v = rng.standard_normal((45, 37)).astype(np.float32)
    direct = smooth.smooth(v, 1100.0, 700.0, 250.0, 200.0)
    sep = smooth.gaussian_separable(v, 1100.0, 700.0, 250.0, 200.0)
    assert np.allclose(sep, direct, rtol=1e-4, atol=1e-5 * np.abs(direct).max())
