"""Reference-layout state of one simulation, as simulate_wavefield allocates
it (src/FDFK2D.f90:253-306, 359-376): the arrays passed to updatedispl.

Also the material law of read_data (src/Input.f90:364-378) and the FD
coefficient tables of src/Submain.f90:1111-1198, restated for the host side.
"""
import numpy as np

# Multiaxial ratio of the absorbing layer's damping (the split parts of a
# strip cell are damped by dX = dx + r dz along x and dZ = dz + r dx along z;
# the library's kPmlRatio, the oracle's ORC_PML_RATIO_DEFAULT).
PML_RATIO = 0.1


def material_law(vp, vs):
    """rho = (vp+980)/2.76, lambda = rho(vp^2-2vs^2), mu = rho vs^2 in FP32,
    returned as the reference's c(1:3, iz, ix) flattened (z fastest).
    vp, vs are (NX, NZ) arrays (x slowest)."""
    vp = np.asarray(vp, np.float32)
    vs = np.asarray(vs, np.float32)
    rho = (vp + np.float32(980.0)) / np.float32(2.760)
    lam = rho * (vp * vp - np.float32(2.0) * vs * vs)
    mu = rho * vs * vs
    return np.stack([rho, lam, mu], axis=-1).astype(np.float32).reshape(-1)


def fd_coef(norder, staggered, second=False):
    """coef(j, i) as a flat Fortran (norder, norder) array: staggered first
    derivative (fd_coef0), collocated first (fd_coef1) or collocated second
    derivative (fd_coef2), src/Submain.f90:1111-1243."""
    c = np.zeros(norder * norder, np.float64)
    for i in range(1, norder + 1):
        for j in range(1, i + 1):
            node = (lambda k: float((2 * k - 1) ** 2)) if staggered else (lambda k: float(k * k))
            num = lo = hi = 1.0
            for k in range(1, j):
                num *= node(k)
                lo *= node(j) - node(k)
            for k in range(j + 1, i + 1):
                num *= node(k)
                hi *= node(k) - node(j)
            num *= 1.0 if j % 2 == 1 else -1.0
            if second:
                lo *= float(j * j)
            else:
                lo *= float(2 * j - 1) if staggered else 2.0 * j
            c[(j - 1) + norder * (i - 1)] = num / lo / hi
    return c


def stable_dt(norder, dx, dz, c):
    """dt_stable of check_stability (src/Submain.f90:139-146)."""
    c2 = fd_coef(norder, False, second=True)
    col = c2[(norder - 1) * norder: norder * norder]
    h = max(dx, dz)
    cx, cz = (h / dx) ** 2, (h / dz) ** 2
    cfl = np.sqrt(2.0 / ((cx + cz) * (np.abs(col).sum() + col.sum())))
    return float(np.float32(cfl) * np.float32(h) / np.float32(vp_max(c)))


def vp_max(c):
    c = np.asarray(c, np.float32).reshape(-1, 3)
    return float(np.sqrt((c[:, 1] + np.float32(2.0) * c[:, 2]) / c[:, 0]).max())


def make_state(nx, nz, npml, norder, dx, dz, dt, nt, c, npower=2, R0=1e-3, vpml=None, pml_top=False):
    """All arrays updatedispl receives, zero-initialised: free surface at row 1
    (izt = 1), or with pml_top an absorbing top layer of npml rows above it
    (izt = 1-npml, src/Input.f90:179-183); c holds rows izt..nz+npml."""
    izt = 1 - npml if pml_top else 1
    NZ, NX = nz + npml - izt + 1, nx + 2 * npml
    c = np.ascontiguousarray(c, np.float32)
    assert c.size == 3 * NZ * NX
    st = dict(norder=norder, nx=nx, nz=nz, npml=npml, izt=izt, pml_top=int(bool(pml_top)), npower=npower,
              nbd=2 * norder - 1, nt=nt, dx=float(dx), dz=float(dz), dt=float(dt),
              vpml=vp_max(c) if vpml is None else float(vpml), R0=float(R0), c=c,
              coef0=fd_coef(norder, True), coef1=fd_coef(norder, False),
              ux=np.zeros(2 * NZ * NX, np.float32), uz=np.zeros(2 * NZ * NX, np.float32))
    lr = nz * npml
    tb = npml * NX
    # bpxlx bpzlx bpxrx bpzrx bpxtx bpxtz bpztx bpztz bpxbx bpxbz bpzbx bpzbz
    st["bp"] = [np.zeros(lr, np.float32) for _ in range(4)] + \
               [np.zeros(tb, np.float32) for _ in range(8)]
    # bux1l..buz3l, ..r, ..b, ..t (call order of src/FDFK2D.f90:520-522)
    st["bu"] = [np.zeros(2 * lr, np.float32) for _ in range(12)] + \
               [np.zeros(2 * tb, np.float32) for _ in range(12)]
    return st


def copy_state(st):
    out = dict(st)
    for k in ("ux", "uz"):
        out[k] = st[k].copy()
    out["bp"] = [a.copy() for a in st["bp"]]
    out["bu"] = [a.copy() for a in st["bu"]]
    return out
