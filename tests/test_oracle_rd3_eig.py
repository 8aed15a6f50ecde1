"""Pins of the oracle's 3D range-dependent eigenray search (go_rd3_eig_*:
the est_eigenray / find_eigenray of geoac.eigenray.3d.cpp over the
Cartesian range-dependent medium of infraga-3d-rngdep -eig_search /
-eig_direct, main.infraga.3d.rngdep.cpp:754-1052).

- Windless columns c(z) = c0 + g z over flat ground (topo::z = 0): rays are
  circles and a hop covers 2 (c0/g) tan(theta), so the n-bounce eigenray to
  a receiver at range D launches at theta_n = atan(g D / (2 c0 (n + 1)))
  toward it.  The search returns exactly one eigenray per bounce count
  whose theta_n lies in the window and clears the shadow-zone height, at
  theta_n within the 0.1 km acceptance disc; find_eigenray alone converges
  to it from a nearby estimate.  (The 3D find_eigenray reflects once per
  trace, so n >= 1 converges to the one-bounce eigenray, as in the
  stratified 3D search.)
- The Levenberg-Marquardt update over terrain (the is_topo branch,
  geoac.eigenray.3d.cpp:337-348): the derivatives of the ground point with
  respect to the launch angles that it builds from row K's auxiliary
  parameters and the terrain slopes equal finite differences of exact
  ground intercepts of neighbouring rays on a sloped plane, once the
  reference's unnormalised move along nu (a factor c_src / c_grnd on the
  correction) is applied; the flat-ground form (slopes left out) misses
  them by far more.
- Over fractal terrain with a horizontal gradient and a wind, the eigenray
  lands within the tolerance of its receiver and its row is the path
  re-trace's row at the same launch angles.
"""
import math

import numpy as np
import pytest

from ga_oracle_py import Rd3Medium
from infraga_amd import synth

C0, G = 0.30, 0.002


def linear_cols(x=(-800.0, -200.0, 400.0, 800.0), y=(-800.0, 0.0, 800.0)):
    z = np.arange(0.0, 150.0 + 1e-9, 0.5)
    x, y = np.array(x), np.array(y)
    rep = lambda f: np.ascontiguousarray(np.broadcast_to(f, (x.size, y.size, f.size)))
    return x, y, z, rep(C0 + G * z), rep(0.0 * z), rep(0.0 * z), rep(1.2e-3 * np.exp(-z / 7.0))


def admissible(theta_deg, incl_min=0.5, incl_max=45.0):
    apex = (C0 / G) * (1.0 / math.cos(math.radians(theta_deg)) - 1.0)
    return incl_min <= theta_deg <= incl_max and apex > 0.2


@pytest.mark.parametrize("D,az", [(120.0, 30.0), (310.0, 200.0)], ids=["D120", "D310"])
def test_rd3_eig_linear_gradient_closed_form(oracle, D, az):
    m = Rd3Medium(oracle, *linear_cols())
    # circles of radius >= 150 km: RK4 at 0.25 km steps is exact far below the 0.1 km disc
    p = m.params(True, ds_max=0.25)
    p.rng_max = 2000.0
    rcvr = (D * math.sin(math.radians(az)), D * math.cos(math.radians(az)))
    arr, _ = m.eig_search(p, (0.0, 0.0, 0.0), rcvr, bnc_min=0, bnc_max=3, incl_step_max=0.5)
    theta = {n: math.degrees(math.atan(G * D / (2.0 * C0 * (n + 1)))) for n in range(4)}
    swept = [n for n in range(4) if admissible(theta[n])]
    assert sorted(arr["n_bounce"].tolist()) == swept, (arr["n_bounce"], theta)
    for a in arr:
        n = int(a["n_bounce"])
        target = theta[min(n, 1)]
        m_hops = min(n, 1) + 1
        dth = 0.1 / (m_hops * (C0 / G) * 2.0 / math.cos(math.radians(target)) ** 2)
        assert abs(math.radians(a["incl_launch"] - target)) < 2.0 * dth + 2e-6, (n, a["incl_launch"], target)
        daz = (a["az_launch"] - az + 180.0) % 360.0 - 180.0
        assert abs(math.radians(daz)) * D < 0.11
        hop = 2.0 * (C0 / G) * math.tan(math.radians(target))
        assert abs(math.hypot(a["pos0"], a["pos1"]) - (n + 1) * hop) < 0.1 * (n + 1) / m_hops + 0.01


def test_rd3_eig_direct_converges_to_closed_form(oracle):
    m = Rd3Medium(oracle, *linear_cols())
    p = m.params(True, ds_max=0.25)
    D, az = 200.0, 75.0
    rcvr = (D * math.sin(math.radians(az)), D * math.cos(math.radians(az)))
    for n in (0, 2):
        th = math.degrees(math.atan(G * D / (2.0 * C0 * (n + 1))))
        target = math.degrees(math.atan(G * D / (2.0 * C0 * (min(n, 1) + 1))))
        row, _ = m.eig_direct(p, (0.0, 0.0, 0.0), rcvr, th + 0.4, bounces=n)
        assert row is not None
        assert abs(row["incl_launch"] - target) < 0.02 and abs(math.radians(row["az_launch"] - az)) * D < 0.11


def sloped(oracle, a, b, c=0.2, g=G):
    """Windless c(z) = C0 + g z columns over the plane z = c + a x + b y."""
    x = np.linspace(-400.0, 400.0, 5)
    z = np.arange(-10.0, 120.0 + 1e-9, 0.5)
    rep = lambda f: np.ascontiguousarray(np.broadcast_to(f, (x.size, x.size, f.size)))
    m = Rd3Medium(oracle, x, x, z, rep(C0 + g * z), rep(0.0 * z), rep(0.0 * z), rep(1.2e-3 * np.exp(-z / 7.0)))
    tx = np.linspace(-420.0, 420.0, 85)
    TX, TY = np.meshgrid(tx, tx, indexing="ij")
    m.set_topo(tx, tx, c + a * TX + b * TY)
    return m


def intercept(m, p, src, th, ph, a, b, c=0.2):
    """The ray state at the ground: rows k-1, k interpolated to z = c + a x + b y."""
    y0, y1 = m.trace_leg(p, src, th, ph)
    f0 = y0[2] - (c + a * y0[0] + b * y0[1])
    f1 = y1[2] - (c + a * y1[0] + b * y1[1])
    t = f0 / (f0 - f1)
    return y0 + t * (y1 - y0)


CASES = [  # slopes a, b; source height; launch; c gradient
    (0.08, -0.05, 0.2, 14.0, 60.0, G),
    (-0.06, 0.09, 0.2, 22.0, 200.0, G),
    (0.07, 0.04, 6.0, -20.0, 130.0, 0.0),     # homogeneous: straight rays down onto the plane
]


@pytest.mark.parametrize("a,b,z_src,th_deg,az,g", CASES, ids=["lin1", "lin2", "homog"])
def test_rd3_lm_update_matches_finite_differences_on_slope(oracle, a, b, z_src, th_deg, az, g):
    """d(x, y)/d(theta, phi) of the ground point: the auxiliary parameters at
    row K plus the move along the ray to the terrain's tangent plane.  The
    reference moves along nu, not nu / |nu| (geoac.eigenray.3d.cpp:347-348
    after ds_dth of :341-342), so its correction is the exact one times
    |nu| = c_src / c_grnd for a windless medium; the check applies that
    factor (1 in the homogeneous case) to the finite differences."""
    m = sloped(oracle, a, b, g=g)
    p = m.params(True, ds_min=0.01, ds_max=0.01)
    p.rng_max = 2000.0
    src = np.array([0.0, 0.0, z_src])
    th, ph = math.radians(th_deg), math.radians(90.0 - az)
    # the update evaluated on the ray state at the ground (the reference
    # evaluates it at row K, up to one step further on)
    yk = intercept(m, p, src, th, ph, a, b)
    J = m.lm_jac(src, yk)
    h = 2e-4
    fd = []
    for dth, dph in ((h, 0.0), (0.0, h)):
        gp = intercept(m, p, src, th + dth, ph + dph, a, b)[:2]
        gm = intercept(m, p, src, th - dth, ph - dph, a, b)[:2]
        fd.extend((gp - gm) / (2.0 * h))
    fd = np.array(fd)
    X = np.array([yk[6], yk[7], yk[12], yk[13]])
    c_src = C0 + g * z_src
    c_grnd = C0 + g * (0.2 + a * yk[0] + b * yk[1])
    expected = X + (c_src / c_grnd) * (fd - X)
    scale = np.abs(fd).max()
    assert np.abs(J - expected).max() < 1e-6 * scale, (J, expected, fd)
    # the flat-ground form (slope terms left out) misses by far more
    ds_dth = -(c_src / c_grnd) * yk[8] / yk[5]
    ds_dph = -(c_src / c_grnd) * yk[14] / yk[5]
    Jf = np.array([yk[6] + yk[3] * ds_dth, yk[7] + yk[4] * ds_dth, yk[12] + yk[3] * ds_dph, yk[13] + yk[4] * ds_dph])
    assert np.abs(Jf - expected).max() > 1e-3 * scale


def test_rd3_eig_over_terrain_lands_on_receivers(oracle):
    """Linear c(z) with a horizontal gradient in x over fractal terrain: the
    search's eigenrays end within the disc of the receiver and their rows are
    the path re-trace's at the same launch angles."""
    x = np.linspace(-400.0, 400.0, 5)
    z = np.arange(0.0, 120.0 + 1e-9, 0.5)
    X, _, Z = np.meshgrid(x, x, z, indexing="ij")
    zero = 0.0 * Z
    m = Rd3Medium(oracle, x, x, z, C0 + G * Z + 2.0e-5 * X, zero + 0.002 * np.exp(-((Z - 40.0) / 8.0) ** 2), zero,
                  1.2e-3 * np.exp(-Z / 7.0))
    tx = np.linspace(-420.0, 420.0, 57)
    _, _, zg = synth.fractal_topo(9, lat=tx, lon=tx, amp_km=0.8)
    m.set_topo(tx, tx, zg)
    p = m.params(True, ds_max=0.25)
    src = (0.0, 0.0, 0.0)
    rcvr = (110.0, -40.0)
    arr, steps = m.eig_search(p, src, rcvr, incl_min=10.0, incl_max=35.0, bnc_min=0, bnc_max=0, incl_step_max=0.5)
    assert len(arr) == 1 and steps > 0
    for e in arr:
        assert math.hypot(e["pos0"] - rcvr[0], e["pos1"] - rcvr[1]) < 0.1
        q, row = m.eig_path(p, src, e["incl_launch"], e["az_launch"], int(e["n_bounce"]))
        assert len(q) > 10
        # the same launch angles up to the degree round trip of the row
        for f in ("pos0", "pos1", "time", "turn_ht", "incl_arr", "back_az", "tl", "absorption"):
            assert row[f] == pytest.approx(e[f], rel=1e-9, abs=1e-9), f
