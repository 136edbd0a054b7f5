"""2-D model smoother: restatement of utils/gauss_smooth.f90 (the reference's
tool for Gaussian or triangular smoothing of an SU model, the template of
the smoothed random perturbation in the large-solve workload).

    python -m fdfk2d_amd.smooth model.su [lambdax lambdaz dx dz iz_first iz_end ix_first ix_end method]

An SU model holds one trace per column x, samples along z.  The smoothed
model is written to <name up to the first '.'>_smooth.su.  With --gpu the
window sum runs on the GPU (fdfk_smooth_model in libfdfk.so, the same FP32
arithmetic); method 2 there is the Gaussian as two 1-D FP64 passes.

Differences from the reference, where its behaviour is not defined:
- The model is allocated vel(1:nx,1:nz) but indexed vel(z, x) and filled
  trace by trace (gauss_smooth.f90:72-73, 391, 413), which addresses the
  file's samples correctly only for square models; here the model is
  indexed as stored, one trace per column.
- With exactly one optional argument the reference reads lambdax from
  argument 1, the file name (gauss_smooth.f90:86-87); here it is read from
  argument 2, the documented meaning.
- The triangular weights (gauss_smooth.f90:334-359) are built in place with
  a transposed read w(j,k); that stays inside the weight array only when the
  half-widths in x and z are equal, and only that case is accepted here.
The program itself does not compile with the image's Fortran compiler (it
calls get_command_count, gauss_smooth.f90:42), so the restatement is pinned
by tests of its mathematics rather than against the program's output.
"""
import math
import os
import sys

import numpy as np

from . import io


def weights(lambdax, lambdaz, dx, dz, method=0):
    """Normalised weight window w[k + nwz, j + nwx] (k along z, j along x),
    gauss_smooth.f90:310-360.  Returns (nwx, nwz, w) with w in FP32."""
    f32 = np.float32
    lambdax, lambdaz, dx, dz = f32(lambdax), f32(lambdaz), f32(dx), f32(dz)
    if method == 0:
        nwx = int(math.ceil(f32(3.6) * lambdax / dx))
        nwz = int(math.ceil(f32(3.6) * lambdaz / dz))
        j = np.arange(-nwx, nwx + 1)
        k = np.arange(-nwz, nwz + 1)
        x = (j.astype(f32) * dx).astype(f32)
        z = (k.astype(f32) * dz).astype(f32)
        rx = (x / lambdax).astype(f32)
        rz = (z / lambdaz).astype(f32)
        # single-precision squares and sum, exponential in double, stored in FP32
        s = (rx * rx)[None, :] + (rz * rz)[:, None]
        w = np.exp(-0.5 * s.astype(f32).astype(np.float64)).astype(f32)
    else:
        nwx = int(math.ceil(f32(2.1) * lambdax / dx))
        nwz = int(math.ceil(f32(2.1) * lambdaz / dz))
        if nwx != nwz:
            raise ValueError("triangular smoothing needs equal half-widths in x and z "
                             "(the reference's in-place weights read outside their array otherwise)")
        n = nwx
        s2z = float(f32(1.8 * math.ceil(lambdaz / dz)))
        s2x = float(f32(1.8 * math.ceil(lambdax / dx)))
        a = np.abs(np.arange(-n, n + 1)).astype(np.float64)
        # The first loop sets w(k, j) = (nwz - |k|)/s2z.  The second walks the
        # columns j from left to right and sets w(k, j) = (nwx - |j|)/s2x *
        # w(j, k), reading (j, k) after its own update when k < j (column k
        # came first) and before it otherwise.  Products in double, stored in
        # FP32, as the Fortran expressions round.
        tz = ((n - a) / s2z).astype(f32)
        wjk = ((n - a) / s2x * tz.astype(np.float64)).astype(f32)   # final (j, k) for j > k
        w = np.empty((2 * n + 1, 2 * n + 1), f32)
        for jj in range(2 * n + 1):
            for kk in range(2 * n + 1):
                src = wjk[kk] if kk < jj else tz[jj]
                w[kk, jj] = f32((n - a[jj]) / s2x * float(src))
    # w = w/sum(w): the FP32 sum runs in the array's memory order, x outer, z inner
    total = np.cumsum(w.T.ravel(), dtype=f32)[-1]
    w = (w / total).astype(f32)
    return nwx, nwz, w


def _mirror(i, n):
    """1-based index folded back into 1..n (gauss_smooth.f90:384-387)."""
    i = np.where(i < 1, 1 - i, i)
    return np.where(i > n, 2 * n - i + 1, i)


def smooth(vel, lambdax, lambdaz, dx, dz, iz0=1, iz1=None, ix0=1, ix1=None, method=0):
    """Smoothed copy of vel (nx, nz) inside rows iz0..iz1 and columns
    ix0..ix1 (1-based, inclusive); cells outside the window keep their
    values.  FP32 accumulation in the reference's order (window x outer, z
    inner), gauss_smooth.f90:366-397."""
    vel = np.asarray(vel, np.float32)
    nx, nz = vel.shape
    iz1 = nz if iz1 is None else iz1
    ix1 = nx if ix1 is None else ix1
    nwx, nwz, w = weights(lambdax, lambdaz, dx, dz, method)
    if nwx >= nx or nwz >= nz:
        raise ValueError("smoothing window wider than the model (the mirror folds once)")
    ii = np.arange(ix0, ix1 + 1)
    jj = np.arange(iz0, iz1 + 1)
    acc = np.zeros((ii.size, jj.size), np.float32)
    wsum = np.float32(0.0)
    for i1 in range(-nwx, nwx + 1):
        k1 = _mirror(ii + i1, nx) - 1
        rows = vel[k1]
        for j1 in range(-nwz, nwz + 1):
            k2 = _mirror(jj + j1, nz) - 1
            wij = w[j1 + nwz, i1 + nwx]
            wsum = np.float32(wsum + wij)
            acc = (acc + wij * rows[:, k2]).astype(np.float32)
    out = vel.copy()
    out[ix0 - 1: ix1, iz0 - 1: iz1] = (acc / wsum).astype(np.float32)
    return out


def gaussian_separable(vel, lambdax, lambdaz, dx, dz):
    """The Gaussian method on the whole model as two 1-D passes: the
    reference's window exp(-(x/lx)^2/2 - (z/lz)^2/2), truncated at
    ceil(3.6 l/d) cells and normalised by its sum, is the product of the
    two 1-D windows truncated and normalised alike, and the mirror fold is
    per axis.  Same result as smooth() up to rounding, for models too large
    for the direct window sum (the large-solve perturbation)."""
    from scipy.ndimage import correlate1d
    out = np.asarray(vel, np.float32)
    for axis, lam, d in ((0, lambdax, dx), (1, lambdaz, dz)):
        n = int(math.ceil(np.float32(3.6) * np.float32(lam) / np.float32(d)))
        x = np.arange(-n, n + 1) * float(d)
        w = np.exp(-0.5 * (x / float(lam)) ** 2)
        # scipy's 'reflect' repeats the edge sample: index 0 -> 1, n+1 -> n
        out = correlate1d(out.astype(np.float64), w / w.sum(), axis=axis, mode="reflect").astype(np.float32)
    return out


def smooth_gpu(vel, lambdax, lambdaz, dx, dz, iz0=1, iz1=None, ix0=1, ix1=None, method=0):
    """smooth() on the GPU (fdfk_smooth_model): methods 0 and 1 with the same
    FP32 arithmetic as smooth(), method 2 the Gaussian as two FP64 passes
    (gaussian_separable's arithmetic) inside the window."""
    from .lib import _check, lib
    v = np.ascontiguousarray(vel, np.float32)
    nx, nz = v.shape
    out = np.empty_like(v)
    iz1 = nz if iz1 is None else iz1
    ix1 = nx if ix1 is None else ix1
    _check(lib().fdfk_smooth_model(v.ctypes.data, out.ctypes.data, nx, nz, float(lambdax),
                                   float(lambdaz), float(dx), float(dz), int(iz0), int(iz1),
                                   int(ix0), int(ix1), int(method)))
    return out


def _window(v, n, default):
    """A window bound as the reference clamps it: outside 1..n -> default."""
    return default if (v <= 0 or v > n) else v


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    on_gpu = "--gpu" in argv
    argv = [a for a in argv if a != "--gpu"]
    if len(argv) < 1:
        print("Usage: python -m fdfk2d_amd.smooth filename [lambdax lambdaz dx dz iz_first iz_end "
              "ix_first ix_end method] [--gpu]")
        return 0
    infile = argv[0]
    vel, head = io.read_su(infile)
    nx, nz = vel.shape
    dz = float(head[58])           # the model's sample interval word, read as dz (read_data)
    num = [float(a) for a in argv[1:]]
    lambdax = num[0] if len(num) >= 1 else dz
    lambdaz = num[1] if len(num) >= 2 else lambdax
    dx = num[2] if len(num) >= 3 else dz
    dz = num[3] if len(num) >= 4 else dz
    iz0 = _window(int(num[4]), nz, 1) if len(num) >= 5 else 1
    iz1 = _window(int(num[5]), nz, nz) if len(num) >= 6 else nz
    ix0 = _window(int(num[6]), nx, 1) if len(num) >= 7 else 1
    ix1 = _window(int(num[7]), nx, nx) if len(num) >= 8 else nx
    method = int(num[8]) if len(num) >= 9 else 0
    if method > 1:
        method = 0
    iz0, iz1 = min(iz0, iz1), max(iz0, iz1)
    ix0, ix1 = min(ix0, ix1), max(ix0, ix1)
    run = smooth_gpu if on_gpu else smooth
    out = run(vel, lambdax, lambdaz, dx, dz, iz0, iz1, ix0, ix1, method)
    base = os.path.basename(infile)
    stem = base[: base.index(".")] if "." in base else base
    outfile = os.path.join(os.path.dirname(infile), stem + "_smooth.su")
    # header word 59 = dz * 1e3 stored in 16 bits, as the reference writes it
    io.write_su(outfile, out, np.int32(int(dz * 1e3)).astype(np.int16))
    d = np.abs(vel.astype(np.float64) - out)
    print(f"After gauss_smooth: maximum absolute error = {d.max():.2f} ; maximum relative error = "
          f"{100.0 * np.sqrt(np.mean((d / vel) ** 2)):.2f} %")
    return 0


if __name__ == "__main__":
    sys.exit(main())
