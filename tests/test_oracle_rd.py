"""Pins of the range-dependent spherical oracle (oracle/ga_oracle_rd.c).

- The medium (hybrid-spline atmosphere with horizontal gradients, terrain,
  boundary layer over terrain, vertical boundary-layer wind, both calc_uvw
  overloads, Sutherland-Bass) and the propagation region are bit-identical
  to the reference's own compiled atmo_state.sph.rngdep.cpp /
  atmo_io.sph.rngdep.cpp (oracle/_ref/libref_sphrd.so), on files the
  reference's own loader reads.
- On a grid of identical columns over flat ground, the range-dependent
  driver reproduces the stratified spherical oracle (itself pinned by the
  reference's printed answers) bit for bit, apart from two reference
  behaviours: the arrival inclination and back azimuth are written
  "x * 180 / pi" (no leading minus on the inclination) where the stratified
  driver writes "-x * (180 / pi)", and the hybrid spline's d2c/dz2 patch uses
  the unclamped height (interpolation.cpp:1763-1771), so RK4 stages just
  below the ground extrapolate it and TL moves by ~1e-11 relative.
- With horizontal gradients: the ray right-hand side equals Hamilton's
  equations of H = c |nu| + v . nu (plus the reference's stated extra
  2 nu_r v terms), and the auxiliary right-hand side equals the directional
  derivative of the ray right-hand side along (R, mu) (except the stated
  omitted G/|c_g|^2 d|c_g| T term).
- Reflection off sloped terrain: the slowness and the position derivatives
  are mirrored about the terrain normal; across a bounce off terrain sloped
  in latitude the azimuth auxiliary variables match finite differences of
  neighbouring rays to ~1 %, where a plain mirror of mu is off by ~26 %.
- The hybrid spline's horizontal second derivatives come from separate
  bicubic patches of the node centred differences, so the auxiliary
  equations use them as the reference does; the linearisation identity is
  checked on a grid where those patches are exact (quadratic fields).
"""
import math
import os

import numpy as np
import pytest

from ga_oracle_py import GEOM_SPH, Medium, Oracle, RdMedium, RefRd, write_rd_files
from infraga_amd import synth

R0 = 6370.0
HAVE_REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_sphrd.so"))


def small_grid(seed=5, topo=True):
    lat = np.array([30.0, 32.0, 34.5, 36.0])
    lon = np.array([-104.0, -102.0, -100.0, -97.5, -96.0])
    g = synth.rngdep_grid(seed, lat=lat, lon=lon, dz=0.5, z_top=120.0, corr_km=300.0)
    out = {"grid": g, "units": synth.rngdep_engine_units(*g)}
    if topo:
        tl, to, tz = synth.fractal_topo(seed + 1, lat=np.linspace(29.5, 36.5, 29), lon=np.linspace(-104.5, -95.5, 37),
                                        amp_km=1.5)
        out["topo"] = (tl, to, tz)
    return out


@pytest.fixture(scope="module")
def grid_files(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("rdgrid"))
    g = small_grid()
    lat, lon, z, T3, u3, v3, rho3 = g["grid"]
    files = write_rd_files(d, lat, lon, z, synth.sound_speed(T3), u3, v3, rho3, topo=g["topo"])
    return g, files


def rd_medium(oracle, g, topo=True):
    m = RdMedium(oracle, *g["units"])
    if topo and "topo" in g:
        m.set_topo(*g["topo"])
    return m


# ------------------------------------------------ bit-exact vs the reference
@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref_sphrd.so not built (no reference tree)")
@pytest.mark.parametrize("use_blw", [False, True])
def test_medium_bit_exact_vs_reference(oracle, grid_files, use_blw):
    g, (pre, lats, lons, topo) = grid_files
    ref = RefRd()
    ref.load(pre, lats, lons, topo)
    ref.use_blw(use_blw)
    m = rd_medium(oracle, g)
    m.config(use_blw=use_blw)
    assert np.array_equal(m.limits(), ref.limits())
    rng = np.random.default_rng(11)
    for _ in range(300):
        la = math.radians(rng.uniform(30.0, 36.0))
        lo = math.radians(rng.uniform(-104.0, -96.0))
        z = rng.uniform(0.0, 4.0) if rng.random() < 0.6 else rng.uniform(0.0, 120.0)
        r = R0 + z
        assert np.array_equal(m.point(r, la, lo), ref.point(r, la, lo))
        assert np.array_equal(m.calc_uvw(r, la, lo, False), ref.calc_uvw(r, la, lo, False))
        assert np.array_equal(m.calc_uvw(r, la, lo, True), ref.calc_uvw(r, la, lo, True))
        assert np.array_equal(m.topo(la, lo), ref.topo(la, lo))
        assert np.array_equal(m.bl(r, la, lo), ref.bl(r, la, lo))
        assert m.sb_alpha(r, la, lo, 0.1) == ref.sb_alpha(r, la, lo, 0.1)
    if use_blw:
        # the vertical wind is live inside the boundary layer over terrain
        w = [m.point(R0 + 0.05 + m.topo(math.radians(33.0), math.radians(-100.0))[0] - R0,
                     math.radians(33.0), math.radians(-100.0))[4]]
        assert abs(w[0]) > 0.0


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref_sphrd.so not built (no reference tree)")
def test_limits_without_terrain_match_reference(oracle, grid_files, tmp_path):
    g, _ = grid_files
    lat, lon, z, T3, u3, v3, rho3 = g["grid"]
    pre, lats, lons, _ = write_rd_files(str(tmp_path), lat, lon, z, synth.sound_speed(T3), u3, v3, rho3)
    ref = RefRd()
    ref.load(pre, lats, lons, None)
    m = rd_medium(oracle, g, topo=False)
    assert np.array_equal(m.limits(), ref.limits())
    rng = np.random.default_rng(2)
    for _ in range(100):
        la, lo = math.radians(rng.uniform(30.0, 36.0)), math.radians(rng.uniform(-104.0, -96.0))
        r = R0 + rng.uniform(0.0, 6.0)
        assert np.array_equal(m.calc_uvw(r, la, lo, True), ref.calc_uvw(r, la, lo, True))
        assert np.array_equal(m.point(r, la, lo), ref.point(r, la, lo))


# ------------------------------------------- reduction to the stratified path
@pytest.mark.parametrize("calc_amp", [True, False])
def test_uniform_grid_reduces_to_stratified(oracle, pert_units, calc_amp):
    z, c, u, v, rho = pert_units
    lat = np.array([20.0, 30.0, 40.0])
    lon = np.array([-110.0, -100.0, -90.0, -80.0])

    def rep(f):
        return np.ascontiguousarray(np.broadcast_to(f, (3, 4, f.size)))
    m = RdMedium(oracle, lat, lon, z, rep(c), rep(u), rep(v), rep(rho))
    s = Medium(oracle, GEOM_SPH, z, c, u, v, rho)
    pr, ps = m.params(calc_amp), s.params(calc_amp)
    for k in ("lat_min", "lat_max", "lon_min", "lon_max", "alt_max"):
        setattr(ps, k, getattr(pr, k))
    L = np.array([[th, az] for az in (-90.0, 135.0) for th in (3.0, 15.0, 31.0)])
    ar, nr = m.fan(L, pr, bounces=2, src=(30.0, -100.0, 0.0))
    a_s, ns = s.fan(L, ps, bounces=2, src=(30.0, -100.0, 0.0))
    assert nr == ns and len(ar) == len(a_s) > 0
    for f in ("launch_idx", "n_bounce", "pos0", "pos1", "time", "cel", "turn_ht", "absorption"):
        assert np.array_equal(ar[f], a_s[f]), f
    np.testing.assert_allclose(ar["incl_arr"], -a_s["incl_arr"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ar["back_az"], a_s["back_az"], rtol=1e-15, atol=1e-13)
    np.testing.assert_allclose(ar["tl"], a_s["tl"], rtol=1e-10, atol=0)


# --------------------------------------------------------- Hamiltonian form
def _fd(fun, h):
    return (fun(h) - fun(-h)) / (2 * h)


def test_rhs_is_hamiltonian_with_horizontal_gradients(oracle):
    g = small_grid(seed=8, topo=False)
    m = rd_medium(oracle, g, topo=False)
    p = m.params(False)
    for z, lat, lon in ((12.3, 31.1, -101.3), (41.7, 33.7, -98.9), (63.05, 35.2, -103.1), (95.2, 30.6, -96.7)):
        q0 = np.array([R0 + z, math.radians(lat), math.radians(lon)])
        nu0 = np.array([0.3, -0.6, 0.55])
        y = np.concatenate([q0, nu0])
        f = m.rhs(y, p)

        def H(q, P):
            nu = np.array([P[0], P[1] / q[0], P[2] / (q[0] * math.cos(q[1]))])
            pt = m.point(*q)
            return pt[0] * math.sqrt(nu @ nu) + pt[3] * nu[1] + pt[2] * nu[2]
        P0 = np.array([nu0[0], q0[0] * nu0[1], q0[0] * math.cos(q0[1]) * nu0[2]])
        hq = np.array([1e-5, 1e-8, 1e-8])
        hp = np.array([1e-6, 1e-3, 1e-3])
        qdot = np.array([_fd(lambda e: H(q0, P0 + e * np.eye(3)[i]), hp[i]) for i in range(3)])
        pdot = -np.array([_fd(lambda e: H(q0 + e * np.eye(3)[i], P0), hq[i]) for i in range(3)])
        r, t = q0[0], q0[1]
        nudot = np.array([
            pdot[0],
            pdot[1] / r - P0[1] * qdot[0] / r ** 2,
            pdot[2] / (r * math.cos(t)) - P0[2] * (qdot[0] * math.cos(t) - r * math.sin(t) * qdot[1]) / (r * math.cos(t)) ** 2,
        ])
        speed = math.sqrt(qdot[0] ** 2 + (r * qdot[1]) ** 2 + (r * math.cos(t) * qdot[2]) ** 2)
        assert np.allclose(f[:3], qdot / speed, rtol=1e-6, atol=1e-12)
        pt = m.point(*q0)
        extra = np.array([0.0, -(1.0 / r) * 2.0 * nu0[0] * pt[3], -(1.0 / (r * math.cos(t))) * 2.0 * nu0[0] * pt[2] * math.cos(t)])
        assert np.allclose(f[3:6], nudot / speed + extra / speed, rtol=1e-6, atol=1e-11)
        # the horizontal gradients are live: dropping them breaks the match
        assert abs(pt[6]) > 1e-6 and abs(pt[7]) > 1e-6


def quadratic_grid():
    """Uniform 1-degree node grid whose fields are a vertical profile plus
    quadratic polynomials in (lat, lon) times smooth vertical shapes.  On the
    interior cells (two nodes from every edge) the centred node differences
    of such data are exact, so the hybrid spline's horizontal second
    derivatives (bicubic patches of those differences, interpolation.cpp:
    1700-1750) equal the derivatives of its first-derivative patches."""
    lat = 30.0 + np.arange(8.0)
    lon = -104.0 + np.arange(8.0)
    z, T, u, v, rho = synth.base_profile(0.5, 120.0)
    X, Y = np.meshgrid(np.radians(lat - 33.5), np.radians(lon + 100.5), indexing="ij")
    g1 = np.exp(-((z - 40.0) / 25.0) ** 2)
    g2 = np.exp(-((z - 70.0) / 15.0) ** 2)

    def poly(a):
        return (a[0] * X + a[1] * Y + a[2] * X * X + a[3] * X * Y + a[4] * Y * Y)[:, :, None]
    c = synth.sound_speed(T) + poly([300.0, -200.0, 4000.0, -3000.0, 2500.0]) * g1
    uu = u + poly([-500.0, 400.0, 6000.0, 3000.0, -5000.0]) * g2
    vv = v + poly([200.0, 600.0, -4000.0, 5000.0, 3000.0]) * g1
    shape = c.shape
    return (lat, lon, z, c / 1000.0, uu / 1000.0, vv / 1000.0,
            np.ascontiguousarray(np.broadcast_to(rho, shape)))


def test_auxiliary_rhs_is_linearisation_with_horizontal_gradients(oracle):
    """The auxiliary right-hand side equals the directional derivative of the
    ray right-hand side along (R, mu), except the stated omitted
    G/|c_g|^2 d|c_g| T term; checked where the hybrid spline's second
    derivatives are consistent with its first derivatives (quadratic_grid)."""
    m = RdMedium(oracle, *quadratic_grid())
    p6, p18 = m.params(False), m.params(True)
    rng = np.random.default_rng(3)
    for z, lat, lon in ((14.2, 32.3, -101.6), (47.9, 34.4, -99.3), (58.4, 33.6, -100.2)):
        base = np.array([R0 + z, math.radians(lat), math.radians(lon), 0.31, -0.52, 0.61])
        aux = rng.normal(size=12) * np.array([5.0, 1e-3, 1e-3, 0.2, 0.2, 0.2] * 2)
        f = m.rhs(np.concatenate([base, aux]), p18)
        for j in range(2):
            a = aux[6 * j: 6 * j + 6]
            h = 1e-6 / max(1.0, np.max(np.abs(a)))
            d = (m.rhs(base + h * a, p6) - m.rhs(base - h * a, p6)) / (2 * h)

            def cgmag(e):
                yy = base + e * a
                pt = m.point(*yy[:3])
                n = yy[3:6]
                return np.linalg.norm(pt[0] * n / np.linalg.norm(n) + np.array([pt[4], pt[3], pt[2]]))
            dcg = (cgmag(h) - cgmag(-h)) / (2 * h)
            cg = cgmag(0.0)
            r, t = base[0], base[1]
            G = np.array([1.0, 1.0 / r, 1.0 / (r * math.cos(t))])
            ST = -m.rhs(base, p6)[3:6] * cg / G
            pt = m.point(*base[:3])
            nu = base[3:6]
            dc, du, dv = pt[5:8], pt[8:11], pt[11:14]
            S = np.linalg.norm(nu) * dc + nu[1] * dv + nu[2] * du
            expect = d.copy()
            expect[3:6] -= G / cg ** 2 * dcg * (ST - S)
            got = f[6 + 6 * j: 12 + 6 * j]
            scale = np.max(np.abs(expect))
            assert np.max(np.abs(got - expect)) / scale < 1e-6, (j, got, expect)
            # horizontal gradients and curvatures are live at this point
            assert abs(pt[6]) > 1e-5 and abs(pt[7]) > 1e-5


# ------------------------------------------------------ terrain reflection
def _mirror(vec, aa, b):
    n = np.array([1.0, -aa, -b]) / math.sqrt(1.0 + aa * aa + b * b)
    return vec - 2.0 * (vec @ n) * n


def test_reflection_mirrors_about_terrain_normal(oracle):
    g = small_grid(seed=12)
    m = rd_medium(oracle, g)
    p = m.params(True)
    rng = np.random.default_rng(4)
    y0, s = m.set_initial(p, math.radians(10.0), math.radians(30.0), R0 + 1.0, math.radians(33.0), math.radians(-100.0))
    for _ in range(20):
        la, lo = math.radians(rng.uniform(31.0, 35.0)), math.radians(rng.uniform(-103.0, -97.0))
        rg = m.topo(la, lo)[0]
        nu = np.array([-0.6, 0.4, 0.5]) * (1.0 + 0.1 * rng.normal(size=3))
        aux = rng.normal(size=12) * np.array([3.0, 1e-3, 1e-3, 0.2, 0.2, 0.2] * 2)
        rk1 = np.concatenate([[rg, la, lo], nu, aux])          # on the terrain: the intercept is rk1
        rk2 = rk1 - np.concatenate([[-0.01, -1e-6, 1e-6], np.zeros(15)])
        out = m.set_refl(p, s, rk2, rk1)
        t = m.topo(la, lo)
        aa, b = t[1] / rg, t[2] / (rg * math.cos(la))
        assert abs(aa) + abs(b) > 1e-5
        assert out[0] == rg and out[1] == la and out[2] == lo
        np.testing.assert_allclose(out[3:6], _mirror(nu, aa, b), rtol=1e-13, atol=1e-15)
        # position derivatives reflect as tangent vectors (dr, r dlat, r cos(lat) dlon)
        for o in (6, 12):
            R = aux[o - 6: o - 3]
            met = np.array([R[0], rg * R[1], rg * math.cos(la) * R[2]])
            mm = _mirror(met, aa, b)
            np.testing.assert_allclose(out[o: o + 3], [mm[0], mm[1] / rg, mm[2] / (rg * math.cos(la))],
                                       rtol=1e-12, atol=1e-15)


def test_reflection_aux_matches_finite_difference_on_slope(oracle):
    """Azimuth derivatives across a bounce off terrain sloped in latitude
    (windless, stratified sound speed, so the azimuth auxiliary equations
    are exact before the bounce, and the reference's tan(nu_theta) and
    repeated db_dlth terms multiply zeros).  With uniform steps, the
    auxiliary variables after set_refl are derivatives at fixed total arc
    length: finite differences of neighbouring rays at a fixed step after
    the bounce, corrected by the shift of the intercept arc length s_i.
    The reference's jump (dC/dphi, da/dphi and ds0 dnu/ds terms) agrees to
    ~1 % (its position reflection is exact to first order in the slope);
    a plain mirror of mu is off by ~26 %."""
    lat = np.array([30.0, 32.0, 34.5, 36.0])
    lon = np.array([-104.0, -102.0, -100.0, -97.5, -96.0])
    z, T, u, v, rho = synth.base_profile(0.25, 120.0)
    shape = (lat.size, lon.size, z.size)
    zero = np.zeros(shape)
    m = RdMedium(oracle, *synth.rngdep_engine_units(lat, lon, z, np.broadcast_to(T, shape), zero, zero,
                                                    np.broadcast_to(rho, shape)))
    tlat, tlon = np.linspace(29.5, 36.5, 57), np.linspace(-104.5, -95.5, 13)
    m.set_topo(tlat, tlon, np.repeat((1.0 + 0.8 * np.sin(np.radians(tlat - 29.5) * 80.0))[:, None], tlon.size, axis=1))
    ds = 0.01
    p = m.params(True, ds_min=ds, ds_max=ds)
    la0, lo0 = math.radians(33.0), math.radians(-100.0)
    zs = m.topo(la0, lo0)[0] + 2.0
    z_max = m.limits()[6]

    def bounce(phi, nsteps, mirror_mu=False):
        y, src = m.set_initial(p, math.radians(-20.0), phi, zs, la0, lo0)
        rows = [y]
        while not (rows[-1][0] - R0 <= z_max and rows[-1][0] < m.topo(rows[-1][1], rows[-1][2])[0]):
            rows.append(m.rk4_step(rows[-1], p)[0])
        k = len(rows) - 1
        y = m.set_refl(p, src, rows[-3], rows[-2])
        frac = (y[1] - rows[-2][1]) / (rows[-2][1] - rows[-3][1])      # intercept within the last step
        if mirror_mu:
            t = m.topo(y[1], y[2])
            y = y.copy()
            y[15:18] = _mirror(rows[-2][15:18], t[1] / y[0], t[2] / (y[0] * math.cos(y[1])))
        out = [y]
        for _ in range(nsteps):
            out.append(m.rk4_step(out[-1], p)[0])
        return np.array(out), (k - 1 + frac) * ds

    phi0, h, n = math.radians(70.0), 1e-5, 150
    c, _ = bounce(phi0, n + 1)
    plus, s_p = bounce(phi0 + h, n)
    minus, s_m = bounce(phi0 - h, n)
    wrong, _ = bounce(phi0, n, mirror_mu=True)
    t = m.topo(c[0][1], c[0][2])
    assert abs(t[1]) / c[0][0] > 2e-3 and t[2] == 0.0
    for i in (0, n):
        fd = (plus[i] - minus[i]) / (2 * h) - (c[i + 1] - c[i]) / ds * (s_p - s_m) / (2 * h)
        scale = np.abs(fd[:6]) + 1e-3 * np.max(np.abs(fd[:6]))
        assert np.max(np.abs(c[i][12:18] - fd[:6]) / scale) < 0.02
        assert np.max(np.abs(wrong[i][12:18] - fd[:6]) / scale) > 0.1


def test_theta_grnd_follows_terrain_slope(oracle):
    g = small_grid(seed=13)
    m = rd_medium(oracle, g)
    la, lo = math.radians(33.2), math.radians(-99.4)
    zs = m.topo(la, lo)[0] - R0
    t = m.topo(la, lo)
    for az in (0.0, 45.0, 90.0, 200.0):
        # slope of the terrain along the launch azimuth, in degrees, plus 0.1
        slope = (t[2] / (R0 * math.cos(la)) * math.sin(math.radians(az)) + t[1] / R0 * math.cos(math.radians(az)))
        assert m.theta_grnd(la, lo, zs, az) == pytest.approx(math.degrees(math.atan(slope)) + 0.1, rel=1e-12)
    assert m.theta_grnd(la, lo, zs + 1.0, 0.0) == -89.9


def test_rngdep_eigenray_search_reduces_to_stratified(oracle, toyatmo_units):
    """The range-dependent eigenray search (go_rd_eig_search: the spherical
    search over the range-dependent medium) on a grid of identical ToyAtmo
    columns over flat ground finds the stratified search's eigenray with the
    same ray-steps, launch angles and arrival row."""
    z, c, u, v, rho = toyatmo_units
    lat = np.array([20.0, 30.0, 40.0])
    lon = np.array([-110.0, -100.0, -90.0, -80.0])

    def rep(f):
        return np.ascontiguousarray(np.broadcast_to(f, (3, 4, f.size)))
    rd = RdMedium(oracle, lat, lon, z, rep(c), rep(u), rep(v), rep(rho))
    st = Medium(oracle, GEOM_SPH, z, c, u, v, rho)
    pr, ps = rd.params(True), st.params(True)
    for k in ("lat_min", "lat_max", "lon_min", "lon_max", "alt_max"):
        setattr(ps, k, getattr(pr, k))
    kw = dict(incl_min=9.5, incl_max=10.6, bnc_min=0, bnc_max=0)
    src, rcvr = (30.0, -100.0, 0.0), (29.985815, -101.96135)   # the reference's 10 deg ToyAtmo arrival
    a1, s1 = rd.eig_search(pr, src, rcvr, **kw)
    a2, s2 = st.eig_search(ps, src, rcvr, **kw)
    assert s1 == s2 and len(a1) == len(a2) == 1
    for f in ("incl_launch", "az_launch", "pos0", "pos1", "time", "turn_ht", "back_az", "absorption"):
        np.testing.assert_allclose(a1[f], a2[f], rtol=1e-11, atol=1e-11, err_msg=f)
    np.testing.assert_allclose(a1["tl"], a2["tl"], atol=1e-8)
    np.testing.assert_allclose(a1["incl_arr"], a2["incl_arr"], atol=1e-8)
