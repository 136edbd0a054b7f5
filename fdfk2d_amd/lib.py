"""ctypes binding of the C ABI in include/fdfk.h (fdfk2d_amd/libfdfk.so).

The library is loaded from the package directory; if it is missing the
import fails loudly -- there is no CPU fallback for the time-stepping path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdfk.so")

FDFK_FP32, FDFK_BF16, FDFK_FP16, FDFK_FP64 = 0, 1, 2, 3
FDFK_PSV, FDFK_SH = 0, 1


class FdfkJob(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("nz", C.c_int32), ("npml", C.c_int32), ("norder", C.c_int32),
        ("is_pml_top", C.c_int32), ("npower", C.c_int32),
        ("x0", C.c_float), ("z0", C.c_float), ("dx", C.c_float), ("dz", C.c_float),
        ("dt", C.c_double), ("R0", C.c_double), ("vpml", C.c_float),
        ("nt", C.c_int32), ("nstp", C.c_int32), ("nrcv", C.c_int32),
        ("rx", C.c_void_p), ("rz", C.c_void_p), ("c", C.c_void_p),
        ("precision", C.c_int32), ("wave", C.c_int32),
    ]


class FdfkFK(C.Structure):
    _fields_ = [
        ("nlayer_left", C.c_int32), ("nlayer_right", C.c_int32),
        ("fk_left", C.c_void_p), ("fk_right", C.c_void_p),
        ("src_is_p", C.c_int32), ("ds", C.c_float), ("f0", C.c_float),
        ("rayp", C.c_double), ("bazs", C.c_void_p), ("block_steps", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m fdfk2d_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fdfk_last_error.restype = C.c_char_p
        L.fdfk_device_bytes.restype = C.c_int64
        L.fdfk_create.argtypes = [C.POINTER(FdfkJob), C.POINTER(C.c_void_p)]
        for name in ("fdfk_run", "fdfk_profile"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int32]
        for name in ("fdfk_sync", "fdfk_iteration", "fdfk_destroy", "fdfk_device_bytes", "fdfk_query"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.fdfk_destroy.restype = None
        L.fdfk_set_background.argtypes = [C.c_void_p, C.c_void_p]
        L.fdfk_get_background.argtypes = [C.c_void_p, C.c_void_p]
        L.fdfk_set_top_background.argtypes = [C.c_void_p, C.c_void_p]
        L.fdfk_get_top_background.argtypes = [C.c_void_p, C.c_void_p]
        for name in ("fdfk_set_background64", "fdfk_set_top_background64"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p]
        for name in ("fdfk_seismograms64", "fdfk_get_state64"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdfk_compute_background.argtypes = [C.c_void_p, C.POINTER(FdfkFK)]
        L.fdfk_seismograms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdfk_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdfk_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.fdfk_kernel_stats.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.fdfk_batch_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
        for name in ("fdfk_batch_add", "fdfk_batch_remove"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p]
        for name in ("fdfk_batch_run", "fdfk_batch_profile", "fdfk_batch_autotune"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int32]
        L.fdfk_batch_tile_width.argtypes = [C.c_void_p]
        for name in ("fdfk_batch_active", "fdfk_batch_sync", "fdfk_batch_destroy"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.fdfk_batch_destroy.restype = None
        L.fdfk_batch_kernel_stats.argtypes = L.fdfk_kernel_stats.argtypes
        L.fdfk_autotune.argtypes = [C.POINTER(FdfkJob), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.fdfk_smooth_model.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int32, C.c_int32]
        L.fdfk_smooth_model_dev.argtypes = L.fdfk_smooth_model.argtypes + [C.c_void_p]
        L.updatedispl_.restype = None
        L.comput_background_wavefield_.restype = None
        L.fdfk_compat_release.restype = None
        _lib = L
    return _lib


class FdfkError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise FdfkError(f"fdfk error {rc}: {lib().fdfk_last_error().decode()}")


def trim_memory():
    """Give device memory that no solver uses back to the driver (fdfk_trim_memory)."""
    _check(lib().fdfk_trim_memory())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def bg_sizes(nx, nz, nbd, nt):
    """Element counts of the twelve background arrays (src/FDFK2D.f90:330-355)."""
    T = nt + 1
    reg_lr, reg_b = (nz - nbd) * nbd * T, nbd * nx * T
    pml_lr, pml_b = nz * nbd * T, nbd * (nx + 2 * nbd) * T
    return [reg_lr] * 4 + [reg_b] * 2 + [pml_lr] * 4 + [pml_b] * 2


def bg_top_sizes(nx, nbd, nt):
    """Element counts of the top ring's uxt_reg, uzt_reg, uxt_pml, uzt_pml
    (an absorbing top layer only)."""
    T = nt + 1
    return [nbd * (nx - 2 * nbd) * T] * 2 + [nbd * (nx + 2 * nbd) * T] * 2


def autotune(nx, nz, npml, norder, dx, dz, dt, c, x0=0.0, z0=0.0, rx=None, rz=None, R0=1e-3, npower=2,
             vpml=0.0, steps=20):
    """Time interior tile widths for this geometry (fdfk_autotune); later
    Solvers of the same geometry use the fastest.  Returns the chosen and the
    occupancy model's width with their ms per step."""
    c = np.ascontiguousarray(c, np.float32)
    rx = None if rx is None else np.ascontiguousarray(rx, np.float32)
    rz = None if rz is None else np.ascontiguousarray(rz, np.float32)
    job = FdfkJob(nx=nx, nz=nz, npml=npml, norder=norder, is_pml_top=0, npower=npower, x0=x0, z0=z0, dx=dx,
                  dz=dz, dt=dt, R0=R0, vpml=vpml, nt=0, nstp=1, nrcv=0 if rx is None else rx.size,
                  rx=_ptr(rx), rz=_ptr(rz), c=_ptr(c), precision=FDFK_FP32)
    blx, bms, mlx, mms = C.c_int32(), C.c_float(), C.c_int32(), C.c_float()
    _check(lib().fdfk_autotune(C.byref(job), int(steps), C.byref(blx), C.byref(bms), C.byref(mlx), C.byref(mms)))
    return dict(lx=blx.value, ms=bms.value, model_lx=mlx.value, model_ms=mms.value)


class Solver:
    """One simulation resident on the GPU (a fdfk_solver handle)."""

    def __init__(self, nx, nz, npml, norder, dx, dz, dt, nt, c, x0=0.0, z0=0.0, nstp=1,
                 rx=None, rz=None, R0=1e-3, npower=2, vpml=0.0, precision=FDFK_FP32, pml_top=False,
                 wave=FDFK_PSV):
        """pml_top: an absorbing top layer of npml rows above row 1 (the
        reference's is_PML_top, izt = 1-npml) instead of the free surface; c
        then holds rows izt..nz+npml.  wave=FDFK_SH: the transverse motion uy
        of an incident SH wave (seismograms() gives it as the first array)."""
        self.wave = wave
        self.nx, self.nz, self.npml, self.norder, self.nt = nx, nz, npml, norder, nt
        self.nbd = 2 * norder - 1
        self.precision = precision
        self.pml_top = bool(pml_top)
        self.izt = 1 - npml if pml_top else 1
        self.NX, self.NZ = nx + 2 * npml, nz + npml - self.izt + 1
        self.nstp = nstp
        self.nsamp = nt // nstp + 1
        self._c = np.ascontiguousarray(c, np.float32)
        assert self._c.size == 3 * self.NX * self.NZ, "c must be (3, NZ, NX) reference layout"
        self._rx = None if rx is None else np.ascontiguousarray(rx, np.float32)
        self._rz = None if rz is None else np.ascontiguousarray(rz, np.float32)
        self.nrcv = 0 if rx is None else self._rx.size
        job = FdfkJob(nx=nx, nz=nz, npml=npml, norder=norder, is_pml_top=int(self.pml_top), npower=npower,
                      x0=x0, z0=z0, dx=dx, dz=dz, dt=dt, R0=R0, vpml=vpml, nt=nt, nstp=nstp,
                      nrcv=self.nrcv, rx=_ptr(self._rx), rz=_ptr(self._rz), c=_ptr(self._c),
                      precision=precision, wave=wave)
        h = C.c_void_p()
        _check(lib().fdfk_create(C.byref(job), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fdfk_destroy(self.h)
            self.h = None

    __del__ = close

    def set_background(self, arrays):
        """The twelve reference arrays, or (twelve, top ring's four) with an
        absorbing top layer.  float64 arrays go to an FDFK_FP64 solver
        unrounded."""
        wide = self.precision == FDFK_FP64 and np.asarray(arrays[0][0] if isinstance(arrays, tuple)
                                                           else arrays[0]).dtype == np.float64
        dt, suffix = (np.float64, "64") if wide else (np.float32, "")
        if isinstance(arrays, tuple) and len(arrays) == 2:
            arrays, top = arrays
            self.set_background(arrays)
            tops = [np.ascontiguousarray(a, dt) for a in top]
            for a, n in zip(tops, bg_top_sizes(self.nx, self.nbd, self.nt)):
                assert a.size == n, (a.size, n)
            tp = (C.c_void_p * 4)(*[a.ctypes.data for a in tops])
            _check(getattr(lib(), "fdfk_set_top_background" + suffix)(self.h, tp))
            return
        if wide:
            arrs = [np.ascontiguousarray(a, np.float64) for a in arrays]
            for a, n in zip(arrs, bg_sizes(self.nx, self.nz, self.nbd, self.nt)):
                assert a.size == n, (a.size, n)
            ptrs = (C.c_void_p * 12)(*[a.ctypes.data for a in arrs])
            _check(lib().fdfk_set_background64(self.h, ptrs))
            return
        arrs = [np.ascontiguousarray(a, np.float32) for a in arrays]
        sizes = bg_sizes(self.nx, self.nz, self.nbd, self.nt)
        for a, n in zip(arrs, sizes):
            assert a.size == n, (a.size, n)
        ptrs = (C.c_void_p * 12)(*[a.ctypes.data for a in arrs])
        _check(lib().fdfk_set_background(self.h, ptrs))

    def compute_background(self, fk_left, fk_right, src_is_p, f0, ds, rayp, bazs, block_steps=0):
        """FK background on the GPU; block_steps > 0 synthesises it in blocks
        of that many steps as the run advances instead of storing all steps."""
        fl = np.ascontiguousarray(fk_left, np.float32)
        fr = np.ascontiguousarray(fk_right, np.float32)
        bz = np.ascontiguousarray(bazs, np.float64)
        fk = FdfkFK(nlayer_left=fl.size // 4, nlayer_right=fr.size // 4, fk_left=_ptr(fl),
                    fk_right=_ptr(fr), src_is_p=1 if src_is_p else 0, ds=ds, f0=f0, rayp=rayp,
                    bazs=_ptr(bz), block_steps=int(block_steps))
        _check(lib().fdfk_compute_background(self.h, C.byref(fk)))

    def get_background(self):
        """The twelve reference arrays; with an absorbing top layer (twelve,
        top ring's four)."""
        arrs = [np.zeros(n, np.float32) for n in bg_sizes(self.nx, self.nz, self.nbd, self.nt)]
        ptrs = (C.c_void_p * 12)(*[a.ctypes.data for a in arrs])
        _check(lib().fdfk_get_background(self.h, ptrs))
        if not self.pml_top:
            return arrs
        tops = [np.zeros(n, np.float32) for n in bg_top_sizes(self.nx, self.nbd, self.nt)]
        tp = (C.c_void_p * 4)(*[a.ctypes.data for a in tops])
        _check(lib().fdfk_get_top_background(self.h, tp))
        return arrs, tops

    def run(self, nsteps):
        _check(lib().fdfk_run(self.h, int(nsteps)))

    def sync(self):
        _check(lib().fdfk_sync(self.h))

    def idle(self):
        """True when every queued iteration has completed (does not wait)."""
        rc = lib().fdfk_query(self.h)
        if rc < 0:
            _check(rc)
        return rc == 1

    @property
    def iteration(self):
        return lib().fdfk_iteration(self.h)

    def seismograms(self):
        """(nrcv, nsamp) traces; float64 from an FDFK_FP64 solver."""
        if self.precision == FDFK_FP64:
            sx = np.zeros((self.nrcv, self.nsamp), np.float64)
            sz = np.zeros((self.nrcv, self.nsamp), np.float64)
            _check(lib().fdfk_seismograms64(self.h, _ptr(sx), _ptr(sz)))
            return sx, sz
        sx = np.zeros((self.nrcv, self.nsamp), np.float32)
        sz = np.zeros((self.nrcv, self.nsamp), np.float32)
        _check(lib().fdfk_seismograms(self.h, _ptr(sx), _ptr(sz)))
        return sx, sz

    def get_state(self):
        """Both levels of ux, uz in the reference layout; float64 from an
        FDFK_FP64 solver."""
        n = 2 * self.NX * self.NZ
        if self.precision == FDFK_FP64:
            ux, uz = np.zeros(n, np.float64), np.zeros(n, np.float64)
            _check(lib().fdfk_get_state64(self.h, _ptr(ux), _ptr(uz)))
            return ux, uz
        ux, uz = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(lib().fdfk_get_state(self.h, _ptr(ux), _ptr(uz)))
        return ux, uz

    def set_state(self, ux, uz):
        ux = np.ascontiguousarray(ux, np.float32)
        uz = np.ascontiguousarray(uz, np.float32)
        _check(lib().fdfk_set_state(self.h, _ptr(ux), _ptr(uz)))

    def profile(self, on=True):
        _check(lib().fdfk_profile(self.h, 1 if on else 0))

    def kernel_stats(self):
        a, b, c, d = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        _check(lib().fdfk_kernel_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(interior_ms=a.value, interior_n=b.value, step_ms=c.value, step_n=d.value)

    @property
    def device_bytes(self):
        return lib().fdfk_device_bytes(self.h)


class Batch:
    """Resident solvers advanced together (fdfk_batch)."""

    def __init__(self, norder=3):
        h = C.c_void_p()
        _check(lib().fdfk_batch_create(int(norder), C.byref(h)))
        self.h = h
        self.members = []

    def add(self, solver):
        _check(lib().fdfk_batch_add(self.h, solver.h))
        self.members.append(solver)

    def remove(self, solver):
        _check(lib().fdfk_batch_remove(self.h, solver.h))
        self.members.remove(solver)

    def run(self, nsteps):
        _check(lib().fdfk_batch_run(self.h, int(nsteps)))

    @property
    def active(self):
        return lib().fdfk_batch_active(self.h)

    def sync(self):
        _check(lib().fdfk_batch_sync(self.h))

    def autotune(self, on=True):
        """Time candidate interior tile widths on the first steps of every
        active set of a new shape and keep the fastest (fdfk_batch_autotune)."""
        _check(lib().fdfk_batch_autotune(self.h, 1 if on else 0))

    @property
    def tile_width(self):
        return lib().fdfk_batch_tile_width(self.h)

    def profile(self, on=True):
        _check(lib().fdfk_batch_profile(self.h, 1 if on else 0))

    def kernel_stats(self):
        a, b, c, d = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        _check(lib().fdfk_batch_kernel_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(interior_ms=a.value, interior_n=b.value, step_ms=c.value, step_n=d.value)

    def close(self):
        if getattr(self, "h", None):
            lib().fdfk_batch_destroy(self.h)
            self.h = None

    __del__ = close


def _i(v):
    return C.byref(C.c_int32(int(v)))


def _f(v):
    return C.byref(C.c_float(float(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def updatedispl(st, it):
    """Drop-in path: call updatedispl_ exactly as src/FDFK2D.f90:516-522 does,
    on a reference-layout state (see fdfk2d_amd.refstate)."""
    bg = st.get("bg")
    bgp = [_ptr(x) for x in bg] if bg is not None else [None] * 12
    lib().updatedispl_(
        _i(st["norder"]), _i(st["nx"]), _i(st["nz"]), _i(st["npml"]), _i(st["izt"]),
        _i(st["pml_top"]), _i(st["npower"]), _i(st["nbd"]), _i(it), _i(st["nt"]),
        _f(st["dx"]), _f(st["dz"]), _d(st["dt"]), _f(st["vpml"]), _d(st["R0"]),
        *bgp, _ptr(st["c"]), _ptr(st["coef0"]), _ptr(st["coef1"]), _ptr(st["ux"]), _ptr(st["uz"]),
        *[_ptr(x) for x in st["bp"]], *[_ptr(x) for x in st["bu"]])


def comput_background_wavefield(nx, nz, npml, izt, nbd, nt, src_is_p, x0, z0, dx, dz, dt, ds,
                                f0, rayp, bazs, fk_left, fk_right):
    """Drop-in path for comput_background_wavefield (src/FDFK2D.f90:432-434)."""
    arrs = [np.zeros(n, np.float32) for n in bg_sizes(nx, nz, nbd, nt)]
    fl = np.ascontiguousarray(fk_left, np.float32)
    fr = np.ascontiguousarray(fk_right, np.float32)
    bz = np.ascontiguousarray(bazs, np.float64)
    lib().comput_background_wavefield_(
        _i(nx), _i(nz), _i(npml), _i(izt), _i(nbd), _i(nt), _i(fl.size // 4), _i(fr.size // 4),
        _i(1 if src_is_p else 0), _f(x0), _f(z0), _f(dx), _f(dz), _d(dt), _f(ds), _f(f0),
        _d(rayp), _ptr(bz), _ptr(fl), _ptr(fr), *[_ptr(a) for a in arrs])
    return arrs
