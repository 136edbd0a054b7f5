"""The allocation-free smoother (fdfk_smooth_model_dev) against the
allocating entry point (fdfk_smooth_model) through the C ABI: two calls in
a row per method on the same GPU, the second on a smaller model, each equal
to the allocating entry point's result bit for bit (the same kernels and
arithmetic; only where the scratch memory lives differs).  Oversized
requests fail with a clear error before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from fdfk2d_amd import lib

pytestmark = pytest.mark.gpu


def _model(nx, nz, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float32)
    base = 1500.0 + 40.0 * z[None, :]
    return (base + 300.0 * rng.standard_normal((nx, nz))).astype(np.float32)


def _host(vel, args):
    out = np.empty_like(vel)
    lib._check(lib.lib().fdfk_smooth_model(vel.ctypes.data, out.ctypes.data, *args))
    return out


def _hip():
    """The HIP runtime the library links (device buffers for the test itself;
    torch's own runtime is a different build and is not mixed in)."""
    lib.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


H2D, D2H = 1, 2


def _dalloc(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _dev(vel, args):
    h = _hip()
    n = vel.nbytes
    d_in, d_out, st = _dalloc(h, n), _dalloc(h, n), C.c_void_p()
    assert h.hipStreamCreate(C.byref(st)) == 0
    try:
        assert h.hipMemcpy(d_in, vel.ctypes.data, n, H2D) == 0
        junk = np.full_like(vel, np.nan)
        assert h.hipMemcpy(d_out, junk.ctypes.data, n, H2D) == 0
        lib._check(lib.lib().fdfk_smooth_model_dev(d_in, d_out, *args, st))
        assert h.hipStreamSynchronize(st) == 0
        out = np.empty_like(vel)
        assert h.hipMemcpy(out.ctypes.data, d_out, n, D2H) == 0
        return out
    finally:
        h.hipStreamDestroy(st)
        h.hipFree(d_in)
        h.hipFree(d_out)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_no_alloc_smoother_equals_allocating_entry_point(method):
    for nx, nz, lam, win in ((300, 160, 400.0, (5, 150, 3, 290)), (90, 70, 250.0, (1, 70, 1, 90))):
        print("case: method %d, %d x %d model, lambda %g, window %s" % (method, nx, nz, lam, win), flush=True)
        vel = _model(nx, nz, nx)
        iz0, iz1, ix0, ix1 = win
        args = (nx, nz, C.c_float(lam), C.c_float(lam), C.c_float(50.0), C.c_float(50.0), iz0, iz1, ix0, ix1, method)
        want = _host(vel, args)
        got = _dev(vel, args)
        assert np.array_equal(got, want)
        assert not np.array_equal(want, vel)


def test_no_alloc_smoother_rejects_oversized_models():
    nx, nz = 6000, 3000            # 18 M cells: the method-2 intermediate exceeds the 64 MiB scratch
    print("case: method 2 on %d x %d (over the scratch limit)" % (nx, nz), flush=True)
    h = _hip()
    # full-size buffers, so that even a missing check could not reach outside them
    a, b = _dalloc(h, nx * nz * 4), _dalloc(h, nx * nz * 4)
    try:
        rc = lib.lib().fdfk_smooth_model_dev(a, b, nx, nz, C.c_float(400.0), C.c_float(400.0), C.c_float(50.0),
                                             C.c_float(50.0), 1, nz, 1, nx, 2, None)
    finally:
        h.hipFree(a)
        h.hipFree(b)
    assert rc == -22 and b"scratch" in lib.lib().fdfk_last_error()
