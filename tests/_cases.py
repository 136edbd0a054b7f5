"""Small seeded simulation cases shared by the tests (sizes the CPU oracle
finishes in seconds)."""
import numpy as np

from fdfk2d_amd import refstate


def fk_model(layers):
    """(ztop, vp, vs) rows -> (4, nlayer) FK model with rho = (vp+980)/2.76
    (src/Input.f90:259-265)."""
    out = []
    for zt, vp, vs in layers:
        out.append([zt, vp, vs, (np.float32(vp) + np.float32(980.0)) / np.float32(2.760)])
    return np.asarray(out, np.float32)


def layered_grid(nx, nz, npml, dx, dz, layers, anomaly=None, seed=0, izt=1):
    """vp, vs on the padded grid (NX, NZ) of rows izt..nz+npml: 1-D layers
    (the first one also above its top) plus an optional Gaussian Vs anomaly
    (x_c, z_c, radius, fraction)."""
    NX, NZ = nx + 2 * npml, nz + npml - izt + 1
    x = (np.arange(NX) - npml) * dx
    z = (np.arange(NZ) + izt - 1) * dz
    vp = np.full((NX, NZ), layers[0][1], np.float32)
    vs = np.full((NX, NZ), layers[0][2], np.float32)
    for zt, p, s in layers:
        m = z >= zt
        vp[:, m] = p
        vs[:, m] = s
    if anomaly is not None:
        xc, zc, r, frac = anomaly
        X, Z = np.meshgrid(x, z, indexing="ij")
        g = np.exp(-((X - xc) ** 2 + (Z - zc) ** 2) / (2 * r * r))
        vs = (vs * (1.0 + frac * g)).astype(np.float32)
    return vp, vs


class Case:
    def __init__(self, nx=121, nz=60, npml=15, norder=3, dx=200.0, dz=200.0, dt=0.01, nt=300,
                 layers=((0.0, 6000.0, 3450.0), (8000.0, 8000.0, 4450.0)), anomaly=(12000.0, 5000.0, 2000.0, -0.08),
                 src_is_p=True, rayp_sdeg=6.6717, f0=1.0, ds=1e-3, baz=np.pi, nrcv=31, pml_top=False):
        self.nx, self.nz, self.npml, self.norder = nx, nz, npml, norder
        self.dx, self.dz, self.dt, self.nt = dx, dz, dt, nt
        self.nbd = 2 * norder - 1
        self.pml_top = bool(pml_top)
        self.izt = 1 - npml if pml_top else 1
        self.NX, self.NZ = nx + 2 * npml, nz + npml - self.izt + 1
        vp, vs = layered_grid(nx, nz, npml, dx, dz, layers, anomaly, izt=self.izt)
        self.c = refstate.material_law(vp, vs)
        self.fk = fk_model(layers)
        self.src_is_p = src_is_p
        self.rayp = rayp_sdeg * 180.0 / (np.pi * 6371.0e3)
        self.f0, self.ds = f0, ds
        self.bazs = np.full(self.NX, baz, np.float64)
        self.rx = np.linspace(0.0, (nx - 1) * dx, nrcv).astype(np.float32)
        self.rz = np.zeros(nrcv, np.float32)

    def state(self):
        return refstate.make_state(self.nx, self.nz, self.npml, self.norder, self.dx, self.dz,
                                   self.dt, self.nt, self.c, pml_top=self.pml_top)

    def oracle_background(self):
        from oracle import pyoracle as po
        return po.background(self.nx, self.nz, self.npml, self.izt, self.nbd, self.nt, self.src_is_p,
                             0.0, 0.0, self.dx, self.dz, self.dt, self.ds, self.f0, self.rayp,
                             self.bazs, self.fk, self.fk)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_run(case, st, bg, nsteps, it0=0, record=True):
    """Advance the oracle nsteps from iteration it0, recording every iteration
    before the update (src/FDFK2D.f90:497-522).  A state of float64 arrays
    runs the FP64 build and records in float64."""
    from oracle import pyoracle as po
    if isinstance(bg, tuple):   # the twelve arrays and the top ring (absorbing top layer)
        st["bg"], st["bg_top"] = bg
    else:
        st["bg"] = bg
    sx, sz = [], []
    for it in range(it0, it0 + nsteps):
        if record:
            row = [po.loadseis(case.nx, case.nz, case.npml, getattr(case, "izt", 1), 0.0, 0.0, case.dx, case.dz,
                               float(rx), float(rz), st["ux"], st["uz"])
                   for rx, rz in zip(case.rx, case.rz)]
            sx.append([r[0] for r in row])
            sz.append([r[1] for r in row])
        po.updatedispl(st, it)
    dt = st["ux"].dtype
    return np.asarray(sx, dt).T, np.asarray(sz, dt).T
