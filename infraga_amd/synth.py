"""Synthetic, seeded inputs for the stratified fan and eigenray workloads.

Base profile (all workloads): z = 0..150 km every 0.1 km (1501 nodes);
temperature piecewise linear through (0, 288 K), (11, 217), (20, 217),
(50, 270), (90, 190), (150, 600); c = sqrt(1.4 * 287 * T) m/s; eastward jet
u = 50 exp(-((z - 60) / 10)^2) m/s; v = 10 exp(-((z - 40) / 8)^2) m/s;
rho = 1.225e-3 exp(-z / 7) g/cm^3.

Perturbed profile (the 162k-ray spherical fan): the base plus seeded
Gaussian-correlated noise, sigma 5 m/s on u and v and 2 K on T, correlation
length 5 km (white PCG64 noise smoothed with exp(-dz^2 / L^2) and rescaled to
the target standard deviation).

Profiles are written as ``zcuvd`` rows (z [km], c [m/s], u [m/s], v [m/s],
rho [g/cm^3]) with 17 significant digits so a text round trip is exact, and
the file ends with a newline, which the reference's line counter expects
(src/util/fileIO.cpp:26-33).
"""
from __future__ import annotations

import numpy as np

T_KNOTS = ((0.0, 288.0), (11.0, 217.0), (20.0, 217.0), (50.0, 270.0), (90.0, 190.0), (150.0, 600.0))


def base_profile(dz: float = 0.1, z_top: float = 150.0):
    """Return (z, T, u, v, rho) of the base profile; u, v in m/s."""
    n = int(round(z_top / dz)) + 1
    z = np.round(np.arange(n) * dz, 10)
    kz, kt = zip(*T_KNOTS)
    T = np.interp(z, kz, kt)
    u = 50.0 * np.exp(-(((z - 60.0) / 10.0) ** 2))
    v = 10.0 * np.exp(-(((z - 40.0) / 8.0) ** 2))
    rho = 1.225e-3 * np.exp(-z / 7.0)
    return z, T, u, v, rho


def _correlated(rng, z, sigma, corr_len):
    w = rng.standard_normal(z.size)
    dz = z[1] - z[0]
    half = int(np.ceil(3.0 * corr_len / dz))
    k = np.exp(-((np.arange(-half, half + 1) * dz) / corr_len) ** 2)
    s = np.convolve(w, k, mode="same")
    s = s - s.mean()
    return s * (sigma / s.std())


def perturbed_profile(seed: int = 20240611, sigma_uv: float = 5.0, sigma_T: float = 2.0,
                      corr_len: float = 5.0):
    """Base profile plus seeded correlated perturbations of u, v and T."""
    z, T, u, v, rho = base_profile()
    rng = np.random.Generator(np.random.PCG64(seed))
    u = u + _correlated(rng, z, sigma_uv, corr_len)
    v = v + _correlated(rng, z, sigma_uv, corr_len)
    T = T + _correlated(rng, z, sigma_T, corr_len)
    return z, T, u, v, rho


def sound_speed(T):
    return np.sqrt(1.4 * 287.0 * T)


def to_engine_units(z, T, u, v, rho):
    """Arrays as the reference loader holds them for a zcuvd file: c, u, v
    divided by 1000.0 (src/atmo/atmo_io.sph.strat.cpp:115-126)."""
    c = sound_speed(T)
    c_txt = np.array([float(f"{x:.17g}") for x in c])
    u_txt = np.array([float(f"{x:.17g}") for x in u])
    v_txt = np.array([float(f"{x:.17g}") for x in v])
    return (np.ascontiguousarray(z, dtype=np.float64), c_txt / 1000.0, u_txt / 1000.0,
            v_txt / 1000.0, np.ascontiguousarray(rho, dtype=np.float64))


def from_zTuvdp(cols):
    """Engine units from zTuvdp columns (z km, T K, u/v m/s, rho g/cm^3,
    p mbar) by the reference loader's conversion (atmo_io.*.strat.cpp):
    c = sqrt(0.1 * 1.4 * p / rho) / 1000 km/s, winds / 1000."""
    d = np.asarray(cols, dtype=np.float64)
    z, u, v, rho, p = d[:, 0], d[:, 2], d[:, 3], d[:, 4], d[:, 5]
    c = np.sqrt(0.1 * 1.4 * p / rho) / 1000.0
    return (np.ascontiguousarray(z), c, u / 1000.0, v / 1000.0, np.ascontiguousarray(rho))


def write_zcuvd(path, z, T, u, v, rho):
    c = sound_speed(T)
    with open(path, "w") as f:
        f.write("# synthetic profile: z [km] c [m/s] u [m/s] v [m/s] rho [g/cm^3]\n")
        for row in zip(z, c, u, v, rho):
            f.write(" ".join(f"{x:.17g}" for x in row) + "\n")


def fan_launches(incl_min, incl_max, incl_step, az_min, az_max, az_step):
    """Explicit launch list (incl, az) in degrees: incl_min + i * step (the
    rank-strided form of main.infraga.sph.mpi.cpp:388-389), azimuth outer."""
    ni = int(np.floor((incl_max - incl_min) / incl_step + 1e-9)) + 1
    na = int(np.floor((az_max - az_min) / az_step + 1e-9)) + 1
    incl = incl_min + np.arange(ni) * incl_step
    az = az_min + np.arange(na) * az_step
    L = np.empty((na * ni, 2))
    L[:, 0] = np.tile(incl, na)
    L[:, 1] = np.repeat(az, ni)
    return L


def serial_loop_launches(incl_min, incl_max, incl_step, az_min, az_max, az_step, theta_grnd=0.1):
    """Launch list exactly as the serial driver's accumulated loops produce it
    (main.infraga.sph.cpp:418-434: phi += step, theta += step)."""
    out = []
    phi = az_min
    while phi <= az_max:
        th0 = max(incl_min, theta_grnd)
        th = th0
        while th <= max(incl_max, theta_grnd):
            out.append((th, phi))
            th += incl_step
        phi += az_step
    return np.array(out, dtype=np.float64).reshape(-1, 2)


def eig_pairs(n_pairs: int = 512, seed: int = 7, rmin: float = 100.0, rmax: float = 800.0):
    """Receiver positions (x, y) [km] for the batched 3D eigenray search:
    range U(rmin, rmax), azimuth U(0, 360) deg clockwise from north."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = rng.uniform(rmin, rmax, n_pairs)
    az = rng.uniform(0.0, 360.0, n_pairs)
    x = r * np.sin(np.radians(az))
    y = r * np.cos(np.radians(az))
    return np.stack([x, y], axis=1)


# ---------------------------------------------------------------------------
# Range-dependent spherical workload: a lat/lon grid of profiles (the base
# profile plus a seeded random field, 500 km horizontal and 5 km vertical
# correlation, peak amplitude 15 m/s on u and v and 5 K on T) and seeded
# fractal terrain (power-law spectrum, scaled to 0..2 km).
# ---------------------------------------------------------------------------
R_EARTH = 6370.0


def _grid_field(rng, lat, lon, z, amp, corr_km, corr_z):
    """Random field on the (lat, lon, z) nodes with Gaussian horizontal
    (great-circle distance, corr_km) and vertical (corr_z) correlation,
    scaled to peak |value| = amp."""
    la, lo = np.meshgrid(np.radians(lat), np.radians(lon), indexing="ij")
    la, lo = la.ravel(), lo.ravel()
    s = np.sin((la[:, None] - la[None, :]) / 2.0) ** 2 + \
        np.cos(la[:, None]) * np.cos(la[None, :]) * np.sin((lo[:, None] - lo[None, :]) / 2.0) ** 2
    d = 2.0 * R_EARTH * np.arcsin(np.sqrt(np.clip(s, 0.0, 1.0)))
    Wh = np.exp(-(d / corr_km) ** 2)
    w = rng.standard_normal((la.size, z.size))
    f = Wh @ w
    dz = z[1] - z[0]
    half = int(np.ceil(3.0 * corr_z / dz))
    k = np.exp(-((np.arange(-half, half + 1) * dz) / corr_z) ** 2)
    f = np.stack([np.convolve(row, k, mode="same") for row in f])
    f = f - f.mean()
    f *= amp / np.max(np.abs(f))
    return f.reshape(lat.size, lon.size, z.size)


def rngdep_grid(seed: int = 3, lat=None, lon=None, dz: float = 0.1, z_top: float = 150.0,
                amp_uv: float = 15.0, amp_T: float = 5.0, corr_km: float = 500.0, corr_z: float = 5.0):
    """Profiles on a lat/lon node grid (default 11 x 11 nodes, 2 deg apart,
    30..50 N, 110..90 W).  Returns (lat, lon, z, T, u, v, rho) with the
    fields shaped [nlat, nlon, nz]; u, v in m/s."""
    lat = np.linspace(30.0, 50.0, 11) if lat is None else np.asarray(lat, dtype=np.float64)
    lon = np.linspace(-110.0, -90.0, 11) if lon is None else np.asarray(lon, dtype=np.float64)
    z, T, u, v, rho = base_profile(dz, z_top)
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (lat.size, lon.size, z.size)
    T3 = T + _grid_field(rng, lat, lon, z, amp_T, corr_km, corr_z)
    u3 = u + _grid_field(rng, lat, lon, z, amp_uv, corr_km, corr_z)
    v3 = v + _grid_field(rng, lat, lon, z, amp_uv, corr_km, corr_z)
    rho3 = np.ascontiguousarray(np.broadcast_to(rho, shape))
    return lat, lon, z, T3, u3, v3, rho3


def rngdep_engine_units(lat, lon, z, T3, u3, v3, rho3):
    """Fields as the reference loader holds them after reading zcuvd files
    written with 17 digits: c, u, v in km/s (text value / 1000.0)."""
    c = np.vectorize(lambda x: float(f"{x:.17g}"))(sound_speed(T3))
    return (np.asarray(lat, dtype=np.float64), np.asarray(lon, dtype=np.float64),
            np.ascontiguousarray(z, dtype=np.float64), c / 1000.0, u3 / 1000.0, v3 / 1000.0,
            np.ascontiguousarray(rho3, dtype=np.float64))


def fractal_topo(seed: int = 5, lat=None, lon=None, amp_km: float = 2.0, beta: float = 2.0):
    """Seeded fractal terrain (spectral synthesis, power ~ |k|^-2 beta) on a
    lat/lon grid (default 401 x 401 nodes every 0.05 deg, 30..50 N,
    110..90 W), scaled to 0..amp_km.  Returns (lat, lon, zg[nlat, nlon])."""
    lat = np.linspace(30.0, 50.0, 401) if lat is None else np.asarray(lat, dtype=np.float64)
    lon = np.linspace(-110.0, -90.0, 401) if lon is None else np.asarray(lon, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = np.fft.fft2(rng.standard_normal((lat.size, lon.size)))
    kx = np.fft.fftfreq(lat.size)[:, None]
    ky = np.fft.fftfreq(lon.size)[None, :]
    k = np.hypot(kx, ky)
    k[0, 0] = 1.0
    f = np.real(np.fft.ifft2(w * k ** (-beta)))
    f = (f - f.min()) / (f.max() - f.min()) * amp_km
    return lat, lon, np.ascontiguousarray(f)
