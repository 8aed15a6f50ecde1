"""Pins of the oracle's 3D Cartesian range-dependent geometry (go_rd3_*,
infraga-3d-rngdep), whose reference build needs the full source tree.

- A grid of identical columns over flat ground reduces it to the pinned 3D
  stratified oracle: the same arrivals, RK4 step counts and path rows, in
  both bookkeeping modes (initial state, equations, reflection, arrival
  rows, and the refs.dv[3] accumulation the stratified set shares).
- Horizontal gradients: the 6-equation right-hand side equals the ray
  equations built from finite differences of the medium's point functions
  c, u, v at the same point (x and y derivatives in their places).
- The 18-equation auxiliary equations are the linearisation of the 6
  along (X, mu) (finite differences of the right-hand side) on a grid with
  v = 0 (where the refs.dv[3] accumulation vanishes).
- Homogeneous medium over a sloped plane z = a x + b y: after the
  reflection the rows continue as the mirror image about the plane normal
  and TL = 20 log10(1 / s) with s the distance from the image source,
  which pins the reflected auxiliary parameters.
"""
import math

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, PATH_CAUSTICS, PATH_RAYS, Medium, Rd3Medium


def columns(z, c, u, v, rho, x=(-1500.0, -500.0, 500.0, 1500.0), y=(-1500.0, 0.0, 1500.0)):
    x, y = np.array(x), np.array(y)
    rep = lambda f: np.ascontiguousarray(np.broadcast_to(f, (x.size, y.size, f.size)))
    return x, y, z, rep(c), rep(u), rep(v), rep(rho)


def test_identical_columns_equal_3d_stratified(oracle, base_units):
    m3 = Rd3Medium(oracle, *columns(*base_units))
    ms = Medium(oracle, GEOM_3D, *base_units)
    L = [(th, az) for az in (-90.0, 30.0, 200.0) for th in (5.0, 17.0, 33.0)]
    ps = ms.params(True)
    ps.rng_max = 1000.0
    a3, _, s3 = m3.fan3(L, m3.params(True), bounces=1, src=(0.0, 0.0, 0.0))
    a, s = ms.fan(L, ps, bounces=1, src=(0.0, 0.0, 0.0))[:2]
    assert s3 == s and len(a3) == len(a) == 18
    for f in ("pos0", "pos1", "time", "turn_ht", "absorption", "incl_arr", "back_az"):
        np.testing.assert_allclose(a3[f], a[f], rtol=1e-10, atol=1e-10, err_msg=f)
    np.testing.assert_allclose(a3["tl"], a["tl"], atol=1e-8)
    a3w, q3, _ = m3.fan3(L, m3.params(True), bounces=1, write=True, flags=PATH_RAYS | PATH_CAUSTICS,
                         src=(0.0, 0.0, 0.0))
    aw, qw, _ = ms.fan_paths(L, ps, PATH_RAYS | PATH_CAUSTICS, bounces=1, src=(0.0, 0.0, 0.0))
    assert len(q3) == len(qw) > 1000 and (q3["kind"] == 1).sum() > 0
    for f in ("launch_idx", "leg", "step", "kind"):
        assert (q3[f] == qw[f]).all(), f
    for f in ("pos0", "pos1", "z", "time"):
        np.testing.assert_allclose(q3[f], qw[f], rtol=1e-10, atol=1e-10, err_msg=f)
    np.testing.assert_allclose(a3w["time"], aw["time"], rtol=1e-12)


def graded(v_on=True):
    """Fields smooth in z and linear in x, y (exact for the hybrid patches)."""
    x = np.linspace(-300.0, 300.0, 7)
    y = np.linspace(-200.0, 200.0, 5)
    z = np.arange(0.0, 80.01, 0.5)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    c = 0.34 - 0.0008 * Z * np.exp(-Z / 40.0) + 2.0e-5 * X - 1.5e-5 * Y
    u = 0.02 * np.sin(Z / 15.0) + 1.0e-5 * X + 2.0e-5 * Y
    v = (0.01 * np.cos(Z / 20.0) - 1.0e-5 * X) * (1.0 if v_on else 0.0)
    rho = 1.2e-3 * np.exp(-Z / 7.0)
    return x, y, z, c, u, v, rho


def test_rhs_horizontal_gradients(oracle):
    m = Rd3Medium(oracle, *graded())
    p = m.params(False)
    rng = np.random.default_rng(4)
    h = 1e-4
    for _ in range(5):
        P = np.array([rng.uniform(-200, 200), rng.uniform(-150, 150), rng.uniform(10.0, 60.0)])
        nu = rng.normal(size=3)
        nu *= 3.0 / np.linalg.norm(nu)
        f = m.rhs(p, np.concatenate([P, nu]))
        pt = lambda q: m.point(q[2], q[0], q[1])   # (c, rho, u, v, w) at (z, x, y)
        val = pt(P)
        c, u, v = val[0], val[2], val[3]
        grad = np.zeros((3, 3))   # rows c, u, v; columns x, y, z
        for n in range(3):
            e = np.zeros(3)
            e[n] = h
            a, b = pt(P + e), pt(P - e)
            grad[:, n] = [(a[0] - b[0]) / (2 * h), (a[2] - b[2]) / (2 * h), (a[3] - b[3]) / (2 * h)]
        nm = np.linalg.norm(nu)
        cg = c * nu / nm + np.array([u, v, 0.0])
        cgm = np.linalg.norm(cg)
        np.testing.assert_allclose(f[:3], cg / cgm, rtol=0, atol=1e-12)
        dnu = -(nm * grad[0] + nu[0] * grad[1] + nu[1] * grad[2]) / cgm
        np.testing.assert_allclose(f[3:6], dnu, rtol=1e-6, atol=1e-10)


def test_aux_equations_are_the_linearisation(oracle):
    m = Rd3Medium(oracle, *graded(v_on=False))
    p18, p6 = m.params(True), m.params(False)
    rng = np.random.default_rng(11)
    eps = 1e-6
    for _ in range(5):
        y = np.concatenate([[rng.uniform(-200, 200), rng.uniform(-150, 150), rng.uniform(10.0, 60.0)],
                            rng.normal(size=3) * 2.9, rng.normal(size=12)])
        f = m.rhs(p18, y)
        for j, (oX, om) in enumerate(((6, 9), (12, 15))):
            d = np.concatenate([y[oX:oX + 3], y[om:om + 3]])
            fd = (m.rhs(p6, y[:6] + eps * d) - m.rhs(p6, y[:6] - eps * d)) / (2 * eps)
            np.testing.assert_allclose(f[oX:oX + 3], fd[:3], rtol=1e-5, atol=1e-8, err_msg=f"X {j}")
            np.testing.assert_allclose(f[om:om + 3], fd[3:], rtol=1e-5, atol=1e-8, err_msg=f"mu {j}")


@pytest.mark.parametrize("a,b,az", [(0.04, -0.02, 60.0), (-0.03, 0.05, 200.0)])
def test_sloped_plane_image_source(oracle, a, b, az):
    x = np.linspace(-100.0, 100.0, 5)
    z = np.arange(-10.0, 40.01, 0.5)
    cst = lambda val: np.full((x.size, x.size, z.size), val)
    m = Rd3Medium(oracle, x, x, z, cst(0.34), cst(0.0), cst(0.0), cst(1.0e-3))
    tx = np.linspace(-120.0, 120.0, 49)
    TX, TY = np.meshgrid(tx, tx, indexing="ij")
    m.set_topo(tx, tx, a * TX + b * TY)
    p = m.params(True, ds_min=0.01, ds_max=0.01)
    th = -25.0
    arr, q, _ = m.fan3([(th, az)], p, bounces=1, write=True, flags=PATH_RAYS, src=(0.0, 0.0, 4.0))
    assert len(arr) >= 1
    q0, q1 = q[q["leg"] == 0], q[q["leg"] == 1]
    assert len(q0) > 5 and len(q1) > 5
    d_in = np.array([q0["pos0"][4] - q0["pos0"][2], q0["pos1"][4] - q0["pos1"][2], q0["z"][4] - q0["z"][2]])
    d_out = np.array([q1["pos0"][4] - q1["pos0"][2], q1["pos1"][4] - q1["pos1"][2], q1["z"][4] - q1["z"][2]])
    nrm = np.array([-a, -b, 1.0]) / math.sqrt(1 + a * a + b * b)
    mirror = d_in - 2.0 * np.dot(d_in, nrm) * nrm
    np.testing.assert_allclose(d_out / np.linalg.norm(d_out), mirror / np.linalg.norm(mirror), atol=1e-9)
    # unfolded path length = distance from the image of the source in the plane
    S = np.array([0.0, 0.0, 4.0])
    Si = S - 2.0 * np.dot(S, nrm) * nrm
    for row in q1[::3]:
        s = np.linalg.norm(np.array([row["pos0"], row["pos1"], row["z"]]) - Si)
        assert row["tl"] == pytest.approx(20.0 * math.log10(1.0 / s), abs=1e-6), row["step"]


def linear_columns(z_bottom=0.5, c_a=0.33, g=0.001):
    """Windless columns with c = c_a + g (z - z_bottom) from z_bottom up; the
    spline holds c(z_bottom) below the profile (clamped lookups)."""
    z = np.arange(z_bottom, 120.0 + 1e-9, 0.5)
    c = c_a + g * (z - z_bottom)
    rho = 1.2e-3 * np.exp(-z / 7.0)
    return columns(z, c, 0.0 * z, 0.0 * z, rho), c_a, g


def test_ground_at_zero_without_terrain(oracle):
    """Without terrain the Cartesian topo::z is 0.0 whatever z_grnd is
    (atmo_state.3d.rngdep.cpp:45-52): the ground test, the reflection and
    the source clamp use z = 0 while the shadow-zone test still uses z0.
    With c = c_a + g z, no wind, z_grnd = 1.0 and the source at z = 0, a ray
    launched at theta is a circle of radius c_a / (g cos theta) through the
    source and each hop covers 2 (c_a / g) tan(theta) at z = 0; reflecting
    at z0 = 1 km would shorten each hop by 2 / tan(theta) km or more."""
    cols, c_a, g = linear_columns(z_bottom=0.0)
    m = Rd3Medium(oracle, *cols)
    m.set_ground(1.0)
    thetas = (8.0, 15.0, 25.0)
    L = [(th, 90.0) for th in thetas]
    arr, _, _ = m.fan3(L, m.params(False), bounces=2, src=(0.0, 0.0, 0.0))
    assert len(arr) == 3 * len(thetas)
    for a in arr:
        th = math.radians(a["incl_launch"])
        hop = 2.0 * (c_a / g) * math.tan(th)
        rng = math.hypot(a["pos0"], a["pos1"])
        # the arrival is the first RK4 row below z = 0, within a few ds_min of the intercept
        assert abs(rng - (a["n_bounce"] + 1) * hop) < 0.01, (a["incl_launch"], a["n_bounce"], rng, hop)
        apex = (c_a / g) * (1.0 / math.cos(th) - 1.0)
        assert abs(a["turn_ht"] - apex) < 1e-3


def test_no_travel_time_limit(oracle):
    """infraga-3d-rngdep stops rays on region, ground and step limits only
    (main.infraga.3d.rngdep.cpp:494-500): a travel-time limit far below the
    arrival times changes nothing."""
    cols, _, _ = linear_columns()
    m = Rd3Medium(oracle, *cols)
    L = [(th, 90.0) for th in (8.0, 15.0)]
    p = m.params(False)
    a1, _, s1 = m.fan3(L, p, bounces=2, src=(0.0, 0.0, 0.0))
    p.time_max = 1.0e-3          # hours
    a2, _, s2 = m.fan3(L, p, bounces=2, src=(0.0, 0.0, 0.0))
    assert s1 == s2 and len(a1) == len(a2) == 6 and (a2["time"] > 3.6).all()


def test_boundary_layer_vertical_wind_derivatives(oracle, base_units):
    """Cartesian boundary-layer vertical wind (atmo::w / dw and both calc_uvw
    overloads) over a hill, on columns with horizontally uniform winds (where
    the reference's derivative formulas, which leave out the horizontal
    derivatives of u0 and v0, are the exact derivatives): dw is the central
    difference of w, and the second overload's ddw the central difference of
    the first overload's dw, in every direction.  The stratified-over-terrain
    form of the same code is bit-identical to the reference's compiled 3D
    medium (test_oracle_strat_topo3d.py); this pins the range-dependent
    branch, its u-x / v-y pairing and the packed index order (zz, xx, yy, zx,
    zy, xy)."""
    z, c, u, v, rho = base_units
    u = u + 0.012 * np.exp(-z / 5.0)
    v = v - 0.007 * np.exp(-z / 5.0)
    xs = (-600.0, -200.0, 200.0, 600.0)
    m = Rd3Medium(oracle, *columns(z, c, u, v, rho, x=xs, y=xs))
    tx = np.linspace(-600.0, 600.0, 49)
    X, Y = np.meshgrid(tx, tx, indexing="ij")
    # above the profile bottom (z = 0), where the spline clamps its height
    zg = 1.2 * np.exp(-((X - 50.0) ** 2 + (Y + 30.0) ** 2) / 200.0 ** 2) + 0.3 * (1.0 + np.sin(X / 90.0) * np.cos(Y / 130.0))
    m.set_topo(tx, tx, zg)
    m.config(use_blw=True)
    rng = np.random.default_rng(31)
    hs = (1e-6, 1e-3, 1e-3)   # z, x, y steps [km]
    checked, w_max = 0, 0.0
    for _ in range(40):
        x, y = rng.uniform(-400.0, 400.0, 2)
        zz = m.topo(x, y)[0] + rng.uniform(0.02, 0.6)
        P = m.point(zz, x, y)
        c1 = m.calc_uvw(zz, x, y, False)
        c2 = m.calc_uvw(zz, x, y, True)
        w, dw, dw1, ddw = P[4], P[14:17], c1[9:12], c2[24:30]
        w_max = max(w_max, abs(w))
        assert np.allclose(c1[2], w, rtol=1e-13, atol=0) and np.allclose(c2[2], w, rtol=1e-13, atol=0)
        assert np.allclose(dw1, dw, rtol=1e-12, atol=1e-18) and np.allclose(c2[9:12], dw, rtol=1e-12, atol=1e-18)
        fd_w, fd_dw = np.zeros(3), np.zeros((3, 3))
        for n in range(3):
            e = np.zeros(3)
            e[n] = hs[n]
            pp, pm = (zz + e[0], x + e[1], y + e[2]), (zz - e[0], x - e[1], y - e[2])
            fd_w[n] = (m.point(*pp)[4] - m.point(*pm)[4]) / (2.0 * hs[n])
            fd_dw[n] = (m.calc_uvw(*pp, False)[9:12] - m.calc_uvw(*pm, False)[9:12]) / (2.0 * hs[n])
        assert np.allclose(dw, fd_w, rtol=2e-6, atol=1e-6 * np.abs(dw).max()), (dw, fd_w)
        full = lambda q: np.array([[q[0], q[3], q[4]], [q[3], q[1], q[5]], [q[4], q[5], q[2]]])
        D = full(ddw)
        scale = np.abs(D).max()
        assert np.allclose(D, fd_dw, rtol=2e-6, atol=2e-6 * scale), (D, fd_dw)
        checked += 1
    assert checked == 40 and w_max > 1e-6
