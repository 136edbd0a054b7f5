"""Synthetic inputs of the benchmark workloads (seeded, numpy default_rng).

c1: plain CPU reference run  -- nx=1001, nz=400, dx=dz=200 m, dt=0.01 s,
    3000 iterations, 20-cell PML, 40 km crust (Vp 6.3, Vs 3.6 km/s) over
    mantle (Vp 8.1, Vs 4.5), one -5% Gaussian Vs anomaly (10 km radius,
    centre drawn with seed 1), P incidence p = 0.06 s/km, f0 = 0.5 Hz,
    101 surface receivers every 2 km.
c2: single large solve -- nx=8001, nz=2000, dx=dz=50 m, dt=0.003 s,
    20000 iterations, 40-cell PML; sediment (0-2 km, Vs 1.5), crust (to 45 km,
    Vs 3.6), mantle (Vs 4.5), Vp = 1.73 Vs, plus white noise (seed 2)
    Gaussian-smoothed with a 5 km kernel and scaled to 5% applied to Vp and
    Vs; P incidence p = 0.07 s/km; 401 receivers every 1 km.
Both use FD order 6 (norder 3), rho = (vp+980)/2.76 (src/Input.f90:370),
origin (45N, 0E, azimuth 0) and an event at (42S, 10E) so the profile lies
along the great circle (src/FDFK2D.f90:388 never prompts).
"""
import numpy as np

from . import geometry, refstate

DEG_PER_KM = 180.0 / (np.pi * 6371.0)       # s/km -> s/deg factor (Input.f90:146)
LAT_ORG, LON_ORG, AZ_ORG = 45.0, 0.0, 0.0
EVLA, EVLO = -42.0, 10.0


def rho_law(vp):
    return (np.float32(vp) + np.float32(980.0)) / np.float32(2.760)


def fk_layers(layers):
    return np.asarray([[zt, vp, vs, rho_law(vp)] for zt, vp, vs in layers], np.float32)


def smooth_noise_fft(shape, sigma_cells, seed):
    """White noise smoothed by a Gaussian (std sigma_cells) through the FFT
    (periodic, untruncated), normalised to unit standard deviation: the fast
    form for the many small request models of the serving workload."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    kx = np.fft.fftfreq(shape[0])[:, None]
    kz = np.fft.rfftfreq(shape[1])[None, :]
    filt = np.exp(-2.0 * (np.pi * sigma_cells) ** 2 * (kx * kx + kz * kz))
    s = np.fft.irfft2(np.fft.rfft2(w) * filt, s=shape)
    return (s / s.std()).astype(np.float32)


def smooth_noise(shape, sigma_cells, seed):
    """White noise smoothed as utils/gauss_smooth.f90 smooths a model
    (Gaussian of correlation length sigma_cells, window 3.6 lengths, mirror
    boundaries; fdfk2d_amd.smooth), normalised to unit standard deviation."""
    from . import smooth
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape).astype(np.float32)
    s = smooth.gaussian_separable(w, sigma_cells, sigma_cells, 1.0, 1.0).astype(np.float64)
    return (s / s.std()).astype(np.float32)


def _common(nx, nz, npml, dx, dz, dt, nt, vp, vs, layers, src_is_p, p_skm, f0, nrcv, rx_step):
    NX = nx + 2 * npml
    bazs = geometry.column_bazs(nx, npml, 0.0, dx, LAT_ORG, LON_ORG, AZ_ORG, EVLA, EVLO)
    rx = (np.arange(nrcv) * rx_step).astype(np.float32)
    return dict(nx=nx, nz=nz, npml=npml, norder=3, dx=float(dx), dz=float(dz), dt=float(dt),
                nt=nt, vp=vp, vs=vs, c=refstate.material_law(vp, vs), fk_left=fk_layers(layers),
                fk_right=fk_layers(layers), src_is_p=src_is_p,
                rayp=p_skm / 1000.0, rayp_sdeg=p_skm / DEG_PER_KM, f0=float(f0), ds=1e-3,
                bazs=bazs, rx=rx, rz=np.zeros(nrcv, np.float32), NX=NX, NZ=nz + npml)


def c1(nt=2999):
    nx, nz, npml, dx = 1001, 400, 20, 200.0
    NX, NZ = nx + 2 * npml, nz + npml
    layers = [(0.0, 6300.0, 3600.0), (40000.0, 8100.0, 4500.0)]
    z = np.arange(NZ) * dx
    vp = np.where(z >= 40000.0, 8100.0, 6300.0).astype(np.float32)[None, :].repeat(NX, 0)
    vs = np.where(z >= 40000.0, 4500.0, 3600.0).astype(np.float32)[None, :].repeat(NX, 0)
    rng = np.random.default_rng(1)
    xc, zc = rng.uniform(50e3, 150e3), rng.uniform(10e3, 60e3)
    x = (np.arange(NX) - npml) * dx
    X, Z = np.meshgrid(x, z, indexing="ij")
    vs = (vs * (1.0 - 0.05 * np.exp(-((X - xc) ** 2 + (Z - zc) ** 2) / (2 * 10e3 ** 2)))).astype(np.float32)
    return _common(nx, nz, npml, dx, dx, 0.01, nt, vp, vs, layers, True, 0.06, 0.5, 101, 2000.0)


def c2(nt=19999, perturb=True):
    nx, nz, npml, dx = 8001, 2000, 40, 50.0
    NX, NZ = nx + 2 * npml, nz + npml
    vsl = [(0.0, 1500.0), (2000.0, 3600.0), (45000.0, 4500.0)]
    layers = [(zt, 1.73 * v, v) for zt, v in vsl]
    z = np.arange(NZ) * dx
    vs1 = np.zeros(NZ, np.float32)
    for zt, v in vsl:
        vs1[z >= zt] = v
    vs = np.repeat(vs1[None, :], NX, 0)
    if perturb:
        d = smooth_noise((NX, NZ), 5000.0 / dx, 2)
        vs = (vs * (1.0 + 0.05 * d)).astype(np.float32)
    vp = (np.float32(1.73) * vs).astype(np.float32)
    return _common(nx, nz, npml, dx, dx, 0.003, nt, vp, vs, layers, True, 0.07, 0.5, 401, 1000.0)


def c3_requests(njobs=64, seed=3, nsteps=10000):
    """The batched workload: njobs simulations drawn with seed 3, nx uniform
    in 2000..8000 and nz uniform in 800..2000 points at dx = dz = 50 m, 10000
    iterations each, 40-cell PML, ray parameter uniform in 0.04..0.08 s/km,
    P and S incidence alternating, layered + smoothed 5% random models as
    the single large solve (fdfk2d_amd.serve.make_model, per-job seed).
    dt is the large solve's 0.003 s, or 0.9 of the reference's stability
    limit (src/Submain.f90:139-146) when a perturbed model needs less; the
    request dicts are those of fdfk2d_amd.serve."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(njobs):
        nx = int(rng.integers(2000, 8001))
        nz = int(rng.integers(800, 2001))
        out.append(dict(id=i, arrival=0.0, nx=nx, nz=nz, dx=50.0, nsteps=nsteps, src_is_p=(i % 2 == 0),
                        p_skm=float(rng.uniform(0.04, 0.08)), seed=int(rng.integers(1 << 30)), npml=40,
                        dt_max=0.003))
    return out
