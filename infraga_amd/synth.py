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
