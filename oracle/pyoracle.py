"""ctypes bindings to the CPU oracle (oracle/build/liboracle.so; its FP64
build oracle/build/liboracle64.so runs on float64 arrays) and to the
reference's own source-visible Fortran routines (oracle/_ref/libfdfk_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
LIB64_PATH = os.path.join(HERE, "build", "liboracle64.so")   # every FP32 quantity in FP64
REF_PATH = os.path.join(HERE, "_ref", "libfdfk_ref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _i(v):
    return C.byref(C.c_int(int(v)))


def _f(v):
    return C.byref(C.c_float(float(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_diffx1.restype = C.c_float
        _lib.orc_diff1.restype = C.c_float
        _lib.orc_check_stability.restype = C.c_int
        _lib.orc_fk_traces.restype = C.c_int
    return _lib


_lib64 = None


def lib64():
    """The FP64 build of the same oracle (-DORC_FP64): arrays and the scalars
    the FP32 build takes as float are double."""
    global _lib64
    if _lib64 is None:
        if not os.path.exists(LIB64_PATH):
            raise RuntimeError("FP64 oracle not built: run `make -C oracle`")
        _lib64 = C.CDLL(LIB64_PATH)
    return _lib64


def _real(fp64):
    """(library, by-reference scalar, by-value scalar type) of one build."""
    if fp64:
        return lib64(), _d, C.c_double
    return lib(), _f, C.c_float


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """Reference routines compiled from src/*.f90 (lazy binding: the LAPACK
    symbols used only by zoeppritz stay unresolved and are never called)."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _ref


def fd_coefs(norder):
    out = []
    for name in ("orc_fd_coef0", "orc_fd_coef1", "orc_fd_coef2"):
        c = np.zeros(norder * norder, np.float64)
        getattr(lib(), name)(C.c_int(norder), _p(c))
        out.append(c)
    return out


OPS = ("diffzx_reg", "diffzx_pml", "diffxz_reg", "diffxz_pml",
       "diffx2_reg", "diffx2_pml", "diffz2_reg", "diffz2_pml")


def op(name, norder, nx, nz, npml, izt, dx, dz, coef, mat, u):
    """Apply one oracle operator to 2-D arrays shaped (NX, NZ) (z fastest)."""
    out = np.zeros_like(u)
    L = lib()
    fn = getattr(L, "orc_" + name)
    ints = [C.c_int(norder), C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(izt)]
    if name.startswith("diffx2"):
        fn(*ints, C.c_float(dx), _p(coef), _p(mat), _p(u), _p(out))
    elif name.startswith("diffz2"):
        fn(*ints, C.c_float(dz), _p(coef), _p(mat), _p(u), _p(out))
    else:
        fn(*ints, C.c_float(dx), C.c_float(dz), _p(coef), _p(mat), _p(u), _p(out))
    return out


def ref_op(name, norder, nx, nz, npml, izt, dx, dz, coef, mat, u):
    """The same operator from the reference's compiled Submain.f90."""
    out = np.zeros_like(u)
    fn = getattr(ref(), name + "_")
    ints = [_i(norder), _i(nx), _i(nz), _i(npml), _i(izt)]
    if name.startswith("diffx2"):
        fn(*ints, _f(dx), _p(coef), _p(mat), _p(u), _p(out))
    elif name.startswith("diffz2"):
        fn(*ints, _f(dz), _p(coef), _p(mat), _p(u), _p(out))
    else:
        fn(*ints, _f(dx), _f(dz), _p(coef), _p(mat), _p(u), _p(out))
    return out


def loadseis(nx, nz, npml, izt, x0, z0, dx, dz, rx, rz, ux, uz):
    """FP64 fields (float64 arrays) run the FP64 build."""
    L, _, R = _real(ux.dtype == np.float64)
    sx, sz = R(), R()
    L.orc_loadseis(C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(izt), R(x0),
                   R(z0), R(dx), R(dz), R(rx),
                   R(rz), _p(ux), _p(uz), C.byref(sx), C.byref(sz))
    return sx.value, sz.value


def ref_loadseis(nx, nz, npml, izt, x0, z0, dx, dz, rx, rz, ux, uz):
    sx, sz = C.c_float(), C.c_float()
    ref().loadseis_(_i(nx), _i(nz), _i(npml), _i(izt), _f(x0), _f(z0), _f(dx), _f(dz),
                    _f(rx), _f(rz), _p(ux), _p(uz), C.byref(sx), C.byref(sz))
    return sx.value, sz.value


def check_stability(nx, nz, npml, izt, norder, dx, dz, dt, coef2, c):
    dts = C.c_float()
    rc = lib().orc_check_stability(C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(izt),
                                   C.c_int(norder), C.c_float(dx), C.c_float(dz),
                                   C.c_double(dt), _p(coef2), _p(c), C.byref(dts))
    return rc, dts.value


def material_law(c):
    lib().orc_material_law(C.c_long(c.size // 3), _p(c))


def backazimuth_grids(lat_org, lon_org, az_org, coordx, evla, evlo):
    coordx = np.ascontiguousarray(coordx, np.float32)
    out = np.zeros(coordx.size, np.float64)
    lib().orc_backazimuth_grids(C.c_double(lat_org), C.c_double(lon_org), C.c_double(az_org),
                                C.c_int(coordx.size), _p(coordx), C.c_double(evla),
                                C.c_double(evlo), _p(out))
    return out


def ref_backazimuth_grids(lat_org, lon_org, az_org, coordx, evla, evlo):
    R = ref()
    for nm, v in (("lat_org", lat_org), ("lon_org", lon_org), ("az_org", az_org)):
        C.c_double.in_dll(R, "_QMfdfk_parE" + nm).value = v
    coordx = np.ascontiguousarray(coordx, np.float32)
    out = np.zeros(coordx.size, np.float64)
    R.compute_backazimuth_grids_(_i(coordx.size), _p(coordx), _d(evla), _d(evlo), _p(out))
    return out


def eval_parts(norder, nx, nz, npml, izt, dx, dz, coef0, coef1, lam, mu, ux, uz, iz, ix):
    parts = np.zeros(6, np.float32)
    lib().orc_eval_parts(C.c_int(norder), C.c_int(nx), C.c_int(nz), C.c_int(npml),
                         C.c_int(izt), C.c_float(dx), C.c_float(dz), _p(coef0), _p(coef1),
                         _p(lam), _p(mu), _p(ux), _p(uz), C.c_int(iz), C.c_int(ix),
                         _p(parts))
    return parts


def damping(nx, nz, npml, nbd, pml_top, npower, dx, dz, vpml, R0):
    izt = 1 - npml if pml_top else 1
    dpx = np.zeros(nx + 2 * npml, np.float32)
    dpz = np.zeros(nz + npml - izt + 1, np.float32)
    lib().orc_damping(C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(nbd), C.c_int(pml_top),
                      C.c_int(npower), C.c_float(dx), C.c_float(dz), C.c_float(vpml),
                      C.c_double(R0), _p(dpx), _p(dpz))
    return dpx, dpz


def updatedispl(st, it):
    """One oracle time step on a state dict produced by fdfk2d_amd-style
    allocation (see tests/_state.py).  Arrays are mutated in place; a state
    of float64 arrays (state64) runs the FP64 build."""
    a = st
    L, fr, _ = _real(a["ux"].dtype == np.float64)
    bg = a.get("bg")
    bgp = [_p(x) for x in bg] if bg is not None else [None] * 12
    top = a.get("bg_top")
    head = [_p(x) for x in top] if (top is not None and a["pml_top"]) else None
    fn = L.orc_updatedispl_ext if head else L.orc_updatedispl
    fn(*(head or []),
        _i(a["norder"]), _i(a["nx"]), _i(a["nz"]), _i(a["npml"]), _i(a["izt"]),
        _i(a["pml_top"]), _i(a["npower"]), _i(a["nbd"]), _i(it), _i(a["nt"]),
        fr(a["dx"]), fr(a["dz"]), _d(a["dt"]), fr(a["vpml"]), _d(a["R0"]),
        *bgp, _p(a["c"]), _p(a["coef0"]), _p(a["coef1"]), _p(a["ux"]), _p(a["uz"]),
        *[_p(x) for x in a["bp"]], *[_p(x) for x in a["bu"]])


def set_pml_ratio(r, fp64=None):
    """Multiaxial ratio of the oracle's PML (both builds unless fp64 names one)."""
    libs = [lib()] if fp64 is False else ([lib64()] if fp64 else [lib()] + ([lib64()] if os.path.exists(LIB64_PATH) else []))
    for L in libs:
        L.orc_set_pml_ratio(C.c_double(r))


def pml_ratio():
    lib().orc_pml_ratio.restype = C.c_double
    return lib().orc_pml_ratio()


def background(nx, nz, npml, izt, nbd, nt, src_is_p, x0, z0, dx, dz, dt, ds, f0, rayp, bazs,
               fk_left, fk_right, fp64=False):
    """Oracle FK background in the reference's twelve-array layout (float64
    arrays from the FP64 build when fp64).  With izt != 1 (an absorbing top
    layer, no free surface) returns (twelve arrays, top ring's four)."""
    L, f, _ = _real(fp64)
    dt_ = np.float64 if fp64 else np.float32
    shapes = bg_shapes(nx, nz, nbd, nt)
    arrs = [np.zeros(s, dt_) for s in shapes]
    fl = np.ascontiguousarray(fk_left, dt_)
    fr = np.ascontiguousarray(fk_right, dt_)
    bz = np.ascontiguousarray(bazs, np.float64)
    top = None
    if izt != 1:   # an absorbing top layer: the top ring as well
        top = [np.zeros(s, dt_) for s in bg_top_shapes(nx, nbd, nt)]
    fn = L.orc_comput_background_wavefield_ext if top else L.orc_comput_background_wavefield
    fn(*([_p(x) for x in top] if top else []),
        _i(nx), _i(nz), _i(npml), _i(izt), _i(nbd), _i(nt), _i(fl.size // 4), _i(fr.size // 4),
        _i(1 if src_is_p else 0), f(x0), f(z0), f(dx), f(dz), _d(dt), f(ds), f(f0),
        _d(rayp), _p(bz), _p(fl), _p(fr), *[_p(x) for x in arrs])
    if top:
        return arrs, top
    return arrs


def state64(st):
    """A copy of an oracle state with every float32 array in float64, for
    the FP64 build (the inputs are the FP32 values, widened exactly)."""
    def wide(v):
        if isinstance(v, np.ndarray):
            return v.astype(np.float64) if v.dtype == np.float32 else v.copy()
        if isinstance(v, list):   # arrays, or lists of them (the SH split parts)
            return [wide(x) for x in v]
        return v
    return {k: wide(v) for k, v in st.items()}


def bg_shapes(nx, nz, nbd, nt):
    """Element counts of uxl_reg..uzb_pml (src/FDFK2D.f90:330-355)."""
    T = nt + 1
    reg_lr = (nz - nbd) * nbd * T
    reg_b = nbd * nx * T
    pml_lr = nz * nbd * T
    pml_b = nbd * (nx + 2 * nbd) * T
    return [reg_lr, reg_lr, reg_lr, reg_lr, reg_b, reg_b,
            pml_lr, pml_lr, pml_lr, pml_lr, pml_b, pml_b]


def bg_top_shapes(nx, nbd, nt):
    """Element counts of the top ring's uxt_reg, uzt_reg, uxt_pml, uzt_pml."""
    T = nt + 1
    return [nbd * (nx - 2 * nbd) * T] * 2 + [nbd * (nx + 2 * nbd) * T] * 2


def fk_traces(fkmodel, src_is_p, p, f0, ds, dt, ntq, z, zref, t0, free_surface=True):
    fk = np.ascontiguousarray(fkmodel, np.float32)
    z = np.ascontiguousarray(z, np.float64)
    ux = np.zeros((z.size, ntq), np.float64)
    uz = np.zeros((z.size, ntq), np.float64)
    rc = lib().orc_fk_traces_fs(C.c_int(fk.size // 4), _p(fk), C.c_int(1 if src_is_p else 0),
                             C.c_int(1 if free_surface else 0),
                             C.c_double(p), C.c_float(f0), C.c_float(ds), C.c_double(dt),
                             C.c_int(ntq), C.c_int(z.size), _p(z), C.c_double(zref),
                             C.c_double(t0), _p(ux), _p(uz))
    if rc:
        raise ValueError("invalid FK model")
    return ux, uz


# ---- SH (transverse) motion ----
def sh_state(nx, nz, npml, norder, dx, dz, dt, nt, c, pml_top=False, npower=2, R0=1e-3, vpml=None):
    """Zeroed SH state for updatedispl_sh: uy (1:2, NZ, NX), memories bp
    (x: L, R, B, T; z: B, T), split parts bu[strip][u1, u2] (1:2, cells)."""
    from fdfk2d_amd import refstate
    izt = 1 - npml if pml_top else 1
    NZ, NX = nz + npml - izt + 1, nx + 2 * npml
    c = np.ascontiguousarray(c, np.float32)
    assert c.size == 3 * NZ * NX
    lr, tb = nz * npml, npml * NX
    cells = [lr, lr, tb, tb]
    return dict(norder=norder, nx=nx, nz=nz, npml=npml, izt=izt, pml_top=int(bool(pml_top)), npower=npower,
                nt=nt, dx=float(dx), dz=float(dz), dt=float(dt),
                vpml=refstate.vp_max(c) if vpml is None else float(vpml), R0=float(R0), c=c,
                coef0=refstate.fd_coef(norder, True), coef1=refstate.fd_coef(norder, False),
                uy=np.zeros(2 * NZ * NX, np.float32),
                bp=[np.zeros(n, np.float32) for n in cells] + [np.zeros(tb, np.float32) for _ in range(2)],
                bu=[[np.zeros(2 * n, np.float32) for _ in range(2)] for n in cells])


def updatedispl_sh(st, it, bg=None):
    """One SH step of the oracle on a sh_state (a state64 of it runs the FP64
    build); bg: the eight ring arrays (uyl_reg, uyr_reg, uyb_reg, uyt_reg,
    uyl_pml, uyr_pml, uyb_pml, uyt_pml; the top two may be None) or None."""
    a = st
    L, _, fv = _real(a["uy"].dtype == np.float64)
    bgp = (C.c_void_p * 8)(*[None if x is None else x.ctypes.data for x in bg]) if bg is not None else None
    bpp = (C.c_void_p * 6)(*[x.ctypes.data for x in a["bp"]])
    bup = ((C.c_void_p * 2) * 4)(*[(C.c_void_p * 2)(*[x.ctypes.data for x in pair]) for pair in a["bu"]])
    L.orc_updatedispl_sh(
        C.c_int(a["norder"]), C.c_int(a["nx"]), C.c_int(a["nz"]), C.c_int(a["npml"]), C.c_int(a["pml_top"]),
        C.c_int(a["npower"]), C.c_int(it), C.c_int(a["nt"]), fv(a["dx"]), fv(a["dz"]),
        C.c_double(a["dt"]), fv(a["vpml"]), C.c_double(a["R0"]), bgp, _p(a["c"]), _p(a["coef0"]),
        _p(a["coef1"]), _p(a["uy"]), bpp, bup)


def background_sh(nx, nz, npml, izt, nbd, nt, x0, z0, dx, dz, dt, ds, f0, rayp, bazs, fk_left, fk_right,
                  fp64=False):
    """SH ring background: eight arrays (the top two None without a top
    layer), float64 from the FP64 build when fp64."""
    L, _, fv = _real(fp64)
    dt_ = np.float64 if fp64 else np.float32
    shapes = bg_shapes(nx, nz, nbd, nt)
    T = nt + 1
    arrs = [np.zeros(shapes[0], dt_), np.zeros(shapes[2], dt_), np.zeros(shapes[4], dt_),
            np.zeros(nbd * (nx - 2 * nbd) * T, dt_) if izt != 1 else None,
            np.zeros(shapes[6], dt_), np.zeros(shapes[8], dt_), np.zeros(shapes[10], dt_),
            np.zeros(nbd * (nx + 2 * nbd) * T, dt_) if izt != 1 else None]
    fl = np.ascontiguousarray(fk_left, dt_)
    fr = np.ascontiguousarray(fk_right, dt_)
    bz = np.ascontiguousarray(bazs, np.float64)
    outp = (C.c_void_p * 8)(*[None if x is None else x.ctypes.data for x in arrs])
    L.orc_background_sh(C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(izt), C.c_int(nbd), C.c_int(nt),
                        fv(x0), fv(z0), fv(dx), fv(dz), C.c_double(dt), fv(ds), fv(f0), C.c_double(rayp), _p(bz),
                        C.c_int(fl.size // 4), _p(fl), C.c_int(fr.size // 4), _p(fr), outp)
    return arrs


def fk_traces_sh(fkmodel, p, f0, ds, dt, ntq, z, zref, t0, free_surface=True):
    fk = np.ascontiguousarray(fkmodel, np.float32)
    z = np.ascontiguousarray(z, np.float64)
    uy = np.zeros((z.size, ntq), np.float64)
    rc = lib().orc_fk_traces_sh(C.c_int(fk.size // 4), _p(fk), C.c_int(1 if free_surface else 0), C.c_double(p),
                                C.c_float(f0), C.c_float(ds), C.c_double(dt), C.c_int(ntq), C.c_int(z.size), _p(z),
                                C.c_double(zref), C.c_double(t0), _p(uy))
    if rc:
        raise ValueError("invalid FK model")
    return uy
