"""The model smoother on the GPU (fdfk_smooth_model, restating
utils/gauss_smooth.f90:310-397) against the numpy restatement in
fdfk2d_amd.smooth, which the CPU tests pin by brute force and closed forms.

Methods 0 and 1 use the same FP32 operations in the same order as smooth()
(separately rounded products and sums, the weight normalisation in memory
order), so they must agree bit for bit.  Method 2 (two FP64 1-D passes) sums
in another order than scipy's correlate1d: a few FP32 ulps of the model's
scale."""
import numpy as np
import pytest

from fdfk2d_amd import io, smooth
from fdfk2d_amd.lib import FdfkError

pytestmark = pytest.mark.gpu


def ulps_of(a):
    return np.spacing(np.float32(np.abs(a).max()))


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("window", [None, (4, 37, 9, 50)])
def test_window_methods_bit_exact(method, window):
    print(f"case: smooth method {method} window {window}, 61x45", flush=True)
    rng = np.random.default_rng(21 + method)
    v = rng.uniform(2000.0, 4000.0, (61, 45)).astype(np.float32)
    lx, lz = (900.0, 500.0) if method == 0 else (1100.0, 1100.0)
    kw = {} if window is None else dict(iz0=window[0], iz1=window[1], ix0=window[2], ix1=window[3])
    want = smooth.smooth(v, lx, lz, 250.0, 200.0 if method == 0 else 250.0, method=method, **kw)
    got = smooth.smooth_gpu(v, lx, lz, 250.0, 200.0 if method == 0 else 250.0, method=method, **kw)
    assert not np.array_equal(want, v)
    assert np.array_equal(got, want)


def test_separable_method_matches_scipy_form():
    print("case: smooth method 2, 2001x500, half-widths 360 and 90 cells", flush=True)
    rng = np.random.default_rng(8)
    v = rng.standard_normal((2001, 500)).astype(np.float32)
    got = smooth.smooth_gpu(v, 5000.0, 2500.0, 50.0, 100.0, method=2)
    want = smooth.gaussian_separable(v, 5000.0, 2500.0, 50.0, 100.0)
    assert np.abs(got - want).max() <= 4 * ulps_of(want)


def test_separable_method_equals_the_window_sum():
    print("case: smooth method 2 vs method 0, 70x52", flush=True)
    rng = np.random.default_rng(9)
    v = rng.uniform(1.0, 2.0, (70, 52)).astype(np.float32)
    direct = smooth.smooth_gpu(v, 700.0, 400.0, 100.0, 80.0, method=0)
    sep = smooth.smooth_gpu(v, 700.0, 400.0, 100.0, 80.0, method=2)
    assert np.allclose(sep, direct, rtol=2e-5)          # FP32 window sum vs FP64 passes; swapped axes: 1.4e-2
    part = smooth.smooth_gpu(v, 700.0, 400.0, 100.0, 80.0, iz0=3, iz1=40, ix0=10, ix1=61, method=2)
    assert np.array_equal(part[9:61, 2:40], sep[9:61, 2:40])
    mask = np.ones(v.shape, bool)
    mask[9:61, 2:40] = False
    assert np.array_equal(part[mask], v[mask])


def test_errors():
    print("case: smoother argument errors", flush=True)
    v = np.ones((20, 10), np.float32)
    with pytest.raises(FdfkError, match="wider than the model"):
        smooth.smooth_gpu(v, 1000.0, 1000.0, 100.0, 100.0)            # nwz = 36 >= nz
    with pytest.raises(FdfkError, match="equal half-widths"):
        smooth.smooth_gpu(v, 300.0, 150.0, 100.0, 100.0, method=1)
    with pytest.raises(FdfkError, match="window"):
        smooth.smooth_gpu(v, 100.0, 100.0, 100.0, 100.0, iz0=5, iz1=11)
    with pytest.raises(FdfkError, match="method"):
        smooth.smooth_gpu(v, 100.0, 100.0, 100.0, 100.0, method=3)


def test_command_line_gpu_equals_cpu(tmp_path):
    print("case: smoother command line --gpu, 24x18", flush=True)
    rng = np.random.default_rng(5)
    v = rng.uniform(3000.0, 5000.0, (24, 18)).astype(np.float32)
    for sub in ("cpu", "gpu"):
        (tmp_path / sub).mkdir()
        io.write_su(str(tmp_path / sub / "vp.su"), v, 200)
    args = ["600", "400", "250", "200", "0", "0", "3", "20", "0"]
    assert smooth.main([str(tmp_path / "cpu" / "vp.su"), *args]) == 0
    assert smooth.main([str(tmp_path / "gpu" / "vp.su"), *args, "--gpu"]) == 0
    a, ha = io.read_su(str(tmp_path / "cpu" / "vp_smooth.su"))
    b, hb = io.read_su(str(tmp_path / "gpu" / "vp_smooth.su"))
    assert np.array_equal(ha, hb) and np.array_equal(a, b)
