"""Host side of the driver's three-component output: the rotation of the
horizontal motion by the relative back-azimuth, the SV/SH split of an
incident S wave and the command-line options."""
import numpy as np
import pytest

from fdfk2d_amd import driver


def test_rotation_is_orthogonal_and_reduces_to_the_reference_projection():
    rng = np.random.default_rng(4)
    sx = rng.standard_normal((9, 50)).astype(np.float32)
    sy = rng.standard_normal((9, 50)).astype(np.float32)
    phi = rng.uniform(-np.pi, np.pi, 9)
    x, y = driver.rotate_horizontal(sx, sy, phi)
    n0 = sx.astype(np.float64) ** 2 + sy.astype(np.float64) ** 2
    assert np.allclose(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2, n0, rtol=1e-6, atol=1e-12)
    # quarter turns: x = -u_y, y = u_x at phi = pi/2; identity at phi = 0
    q = np.full(9, np.pi / 2)
    x, y = driver.rotate_horizontal(sx, sy, q)
    assert np.allclose(x, -sy, atol=1e-6) and np.allclose(y, sx, atol=1e-6)
    x, y = driver.rotate_horizontal(sx, sy, np.zeros(9))
    assert np.array_equal(x, sx) and np.array_equal(y, sy)
    # no transverse motion: the x component is the reference's u_x cos(phi)
    # (src/FDFK2D.f90:540-542, FP64 product rounded to FP32) and the same with
    # an all-zero u_y
    x0, y0 = driver.rotate_horizontal(sx, None, phi)
    x1, y1 = driver.rotate_horizontal(sx, np.zeros_like(sx), phi)
    assert np.array_equal(x0, x1) and np.array_equal(y0, y1)
    assert np.array_equal(x0[:, 0], np.float32(np.float64(sx[:, 0]) * np.cos(phi)))


def test_sv_sh_split():
    p_ev = dict(src_is_p=True, ds=1e-3)
    s_ev = dict(src_is_p=False, ds=1e-3)
    assert driver.split_strength(p_ev, 30.0) == (1e-3, 0.0)
    assert driver.split_strength(s_ev, 0.0) == (1e-3, 0.0)
    sv, sh = driver.split_strength(s_ev, 30.0)
    assert sv == pytest.approx(1e-3 * np.sqrt(3) / 2) and sh == pytest.approx(0.5e-3)
    sv, sh = driver.split_strength(s_ev, -90.0)
    assert abs(sv) < 1e-18 and sh == pytest.approx(-1e-3)


def test_options():
    o, rest = driver._options(["in", "inpar.dat", "out", "seis", "--3c", "--spol", "25", "--all-events"])
    assert rest == ["in", "inpar.dat", "out", "seis"]
    assert o == dict(all_events=True, three=True, spol=25.0)
    o, rest = driver._options(["in", "inpar.dat", "out", "--spol=-10.5"])
    assert rest == ["in", "inpar.dat", "out"] and o["spol"] == -10.5 and not o["three"]
