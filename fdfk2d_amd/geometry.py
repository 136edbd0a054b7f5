"""Back-azimuth geometry of the profile (host side).

Restates src/coord_transform.f90:1-61 (compute_backazimuth_grids),
src/define_rotation_matrix.f90:66-87 (the closed-form inverse rotation) and
src/calc_backazimuth.f90:72-267 (Bullen's great-circle back-azimuth, geogr2cart,
cart2geogr) in vectorised numpy.  FP64 throughout like the reference; the
coordinate along the profile is FP32 as in the reference call.
"""
import numpy as np

EARTH_RADIUS = float(np.float32(6371.0e3))  # REAL(4) in calc_backazimuth.f90:7
FLATTENING = 1.0 / 298.257


def backaz(evla, evlo, stla, stlo):
    """Back-azimuth (degrees) from station to event (calc_backazimuth.f90:72-197)."""
    d2r = np.pi / 180.0
    coeff = (1.0 - FLATTENING) ** 2
    stla = np.asarray(stla, np.float64)
    stlo = np.asarray(stlo, np.float64)
    scola = np.pi / 2 - np.arctan(coeff * np.tan(stla * d2r))
    ecola = np.pi / 2 - np.arctan(coeff * np.tan(evla * d2r))
    slon, elon = stlo * d2r, evlo * d2r
    c = np.cos(scola)
    d, e = np.sin(slon), -np.cos(slon)
    g, h, k = -c * e, c * d, -np.sin(scola)
    aa, bb, cc = np.sin(ecola) * np.cos(elon), np.sin(ecola) * np.sin(elon), np.cos(ecola)
    rhs1 = (aa - d) ** 2 + (bb - e) ** 2 + cc * cc - 2.0
    rhs2 = (aa - g) ** 2 + (bb - h) ** 2 + (cc - k) ** 2 - 2.0
    dbaz = np.arctan2(rhs1, rhs2)
    dbaz = np.where(dbaz < 0.0, dbaz + 2 * np.pi, dbaz)
    baz = dbaz * (180.0 / np.pi)
    return np.where(np.abs(baz - 360.0) < 1e-5, 0.0, baz)


def backazimuth_grids(lat_org, lon_org, az_org, coordx, evla, evlo):
    """Back-azimuth (radians) of each profile coordinate x (m) to the event
    (coord_transform.f90:1-61)."""
    d2r = np.pi / 180.0
    az, lat, lon = az_org * d2r, lat_org * d2r, lon_org * d2r
    st, ct, sp, cp, sa, ca = np.sin(lat), np.cos(lat), np.sin(lon), np.cos(lon), np.sin(az), np.cos(az)
    f1 = st * cp
    i11, i21, i31 = -sa * sp + ca * f1, sa * cp + ca * f1, -ca * ct
    colat, lo = (90.0 - lat_org) * d2r, lon_org * d2r
    xo = EARTH_RADIUS * np.sin(colat) * np.cos(lo)
    yo = EARTH_RADIUS * np.sin(colat) * np.sin(lo)
    zo = EARTH_RADIUS * np.cos(colat)
    cx = np.asarray(coordx, np.float32).astype(np.float64)
    x, y, z = i11 * cx + xo, i21 * cx + yo, i31 * cx + zo
    r = np.hypot(x, y)
    stla = np.arctan2(z, r) * (180.0 / np.pi)
    stlo = np.arctan2(y, x) * (180.0 / np.pi)
    tiny = float(np.float32(1e-9))
    stla = np.where(np.abs(stla) < tiny, 0.0, stla)
    stlo = np.where(np.abs(stlo) < tiny, 0.0, stlo)
    pole = r < 1e-6
    stla = np.where(pole, np.where(z > 0, 90.0, -90.0), stla)
    stlo = np.where(pole, 0.0, stlo)
    return backaz(evla, evlo, stla, stlo) * d2r


def column_bazs(nx, npml, x0, dx, lat_org, lon_org, az_org, evla, evlo):
    """bazs of simulate_wavefield (src/FDFK2D.f90:381-387): per padded column,
    relative to the profile azimuth."""
    ix = np.arange(1 - npml, nx + npml + 1)
    coordx = (np.float32(x0) + (ix - 1).astype(np.float32) * np.float32(dx)).astype(np.float32)
    return backazimuth_grids(lat_org, lon_org, az_org, coordx, evla, evlo) - az_org * np.pi / 180.0
