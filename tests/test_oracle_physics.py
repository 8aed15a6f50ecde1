"""Pins of the oracle restatement by what the mathematics fixes.

The reference's equation-set, RK4 and eigenray sources cannot be built in
this image (they include <fftw3.h>), so these parts of the oracle are pinned
by closed forms, invariants and finite differences instead:

- homogeneous medium: straight rays, t = s / c, Jacobian D = s^2 cos(theta),
  amplitude 1/s (spherical spreading);
- linear sound-speed gradient c = c0 + g z: circular arcs with range
  2 (c0/g) tan(theta0), turning height (c0/g)(1/cos(theta0) - 1) and travel
  time (1/g) ln((1 + sin(theta0)) / (1 - sin(theta0)));
- the ray equations are Hamilton's equations of H = c |nu| + v . nu: the
  oracle's right-hand side equals finite differences of H (3D and spherical,
  with winds);
- the eikonal residual stays at zero along rays; spherical windless rays
  conserve r |nu_horizontal| (Bouguer);
- the auxiliary variables equal finite differences of neighbouring rays with
  respect to the launch angles, before and across a ground reflection;
- specular reflection off flat ground flips the vertical slowness and keeps
  the horizontal ones.
"""
import math

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, GEOM_SPH
from oracle_drive import amp, eikonal, jacobian, make_medium, medium_point, refl, rhs, src_struct, trace_steps


def homogeneous(c=0.34):
    z = np.arange(0.0, 150.01, 1.0)
    return z, np.full_like(z, c), np.zeros_like(z), np.zeros_like(z), np.full_like(z, 1.0e-3)


def linear(c0=0.30, g=0.002):
    z = np.arange(0.0, 150.01, 0.5)
    return z, c0 + g * z, np.zeros_like(z), np.zeros_like(z), np.full_like(z, 1.0e-3)


# ---------------------------------------------------------------- straight rays
@pytest.mark.parametrize("geom", [GEOM_3D, GEOM_SPH])
def test_homogeneous_straight_ray(oracle, geom):
    m = make_medium(oracle, geom, *homogeneous())
    p = m.params(True)
    th, az = 25.0, 60.0
    s = src_struct(m, p, (0.0, 0.0, 5.0) if geom == GEOM_3D else (5.0, 10.0, 20.0), th, az)
    rows, path = trace_steps(m, p, s, 400)
    y = rows[-1]
    S = path[-1]
    if geom == GEOM_3D:
        t, ph = math.radians(th), math.radians(90 - az)
        exp = np.array([S * math.cos(t) * math.cos(ph), S * math.cos(t) * math.sin(ph), 5.0 + S * math.sin(t)])
        assert np.allclose(y[:3], exp, rtol=0, atol=1e-11 * S)
        # Jacobian of a straight-ray fan: D = s^2 cos(theta); amplitude 1/s
        assert abs(jacobian(m, y)) == pytest.approx(S * S * math.cos(t), rel=1e-10)
        assert amp(m, s, y) == pytest.approx(1.0 / S, rel=1e-10)
    else:
        # sph: distance from the source point equals the path length
        r0 = 6370.0
        def xyz(r, la, lo):
            return np.array([r * math.cos(la) * math.cos(lo), r * math.cos(la) * math.sin(lo), r * math.sin(la)])
        d = np.linalg.norm(xyz(*y[:3]) - xyz(r0 + 5.0, math.radians(10), math.radians(20)))
        assert d == pytest.approx(S, rel=1e-10)
        assert amp(m, s, y) == pytest.approx(1.0 / S, rel=1e-8)
    # travel time of a straight leg through a homogeneous medium: t = s / c
    from ctypes import POINTER, c_double
    tt = sum(m.o.lib.go_travel_time_seg(m.h, rows[i].ctypes.data_as(POINTER(c_double)),
                                        rows[i + 1].ctypes.data_as(POINTER(c_double))) for i in range(len(rows) - 1))
    assert tt == pytest.approx(S / 0.34, rel=1e-10)


def test_homogeneous_spherical_spreading_at_ground(oracle):
    """Elevated source, downward ray, flat ground: TL = 20 log10(1 / s)."""
    m = make_medium(oracle, GEOM_3D, *homogeneous())
    p = m.params(True)
    arr, _ = m.fan(np.array([[-30.0, 90.0]]), p, bounces=0, src=(0.0, 0.0, 5.0))
    assert len(arr) == 1
    a = arr[0]
    s_path = math.hypot(a["pos0"], 5.0 - 0.0)        # first sub-ground row lies within one step of z = 0
    assert a["tl"] == pytest.approx(20 * math.log10(1.0 / s_path), abs=0.01)
    assert a["time"] == pytest.approx(10.0 / 0.34, abs=0.01)
    assert a["incl_arr"] == pytest.approx(30.0, abs=1e-6)


# ------------------------------------------------------------ linear gradient
@pytest.mark.parametrize("th0", [8.0, 15.0, 25.0])
def test_linear_gradient_arc(oracle, th0):
    c0, g = 0.30, 0.002
    m = make_medium(oracle, GEOM_3D, *linear(c0, g))
    p = m.params(True)
    arr, _ = m.fan(np.array([[th0, 90.0]]), p, bounces=2, src=(0.0, 0.0, 0.0), turn_ht_min=0.0)
    t0 = math.radians(th0)
    R = 2 * (c0 / g) * math.tan(t0)
    zt = (c0 / g) * (1 / math.cos(t0) - 1)
    T = (1 / g) * math.log((1 + math.sin(t0)) / (1 - math.sin(t0)))
    assert [int(a["n_bounce"]) for a in arr] == [0, 1, 2]
    for a in arr:
        n = a["n_bounce"] + 1
        # arrival = first row below ground: within one boundary-layer step of the intercept
        assert a["pos0"] == pytest.approx(n * R, abs=5e-3 * n)
        assert abs(a["pos1"]) < 1e-9
        assert a["turn_ht"] == pytest.approx(zt, abs=1e-5)
        assert a["time"] == pytest.approx(n * T, abs=2e-2 * n)
        assert a["incl_arr"] == pytest.approx(th0, abs=0.05)
        assert a["back_az"] == pytest.approx(-90.0, abs=1e-9)
        assert a["cel"] == pytest.approx(a["pos0"] / a["time"], rel=1e-12)


# --------------------------------------------------------- Hamiltonian form
def _fd(fun, x, h):
    return (fun(x + h) - fun(x - h)) / (2 * h)


def test_rhs_is_hamiltonian_3d(oracle, pert_units):
    m = make_medium(oracle, GEOM_3D, *pert_units)
    p = m.params(False)
    s = src_struct(m, p, (0.0, 0.0, 0.0), 10.0, 30.0)
    rows, _ = trace_steps(m, p, s, 1)   # sets c0 / nu0 in s
    for z in (12.3, 41.7, 63.05, 95.2):
        y = np.array([10.0, -4.0, z, 0.6, -0.5, 0.45] + [0.0] * 12)
        f = rhs(m, p, s, y)

        def H(x, nu):
            c, u, v, _ = medium_point(m, (x[0], x[1], x[2]))
            return c * math.sqrt(nu @ nu) + u * nu[0] + v * nu[1]
        X, NU = y[:3].copy(), y[3:6].copy()
        dHdnu = np.array([_fd(lambda e: H(X, NU + e * np.eye(3)[i]), 0.0, 1e-6) for i in range(3)])
        dHdx = np.array([_fd(lambda e: H(X + e * np.eye(3)[i], NU), 0.0, 1e-5) for i in range(3)])
        cg = np.linalg.norm(dHdnu)
        assert np.allclose(f[:3], dHdnu / cg, atol=1e-8)
        assert np.allclose(f[3:6], -dHdx / cg, atol=1e-7)


def test_rhs_is_hamiltonian_sph(oracle, pert_units):
    """Canonical momenta p = (nu_r, r nu_t, r cos(t) nu_p): the spherical
    geometry terms (G, T) must reproduce Hamilton's equations."""
    m = make_medium(oracle, GEOM_SPH, *pert_units)
    p = m.params(False)
    s = src_struct(m, p, (0.0, 30.0, 0.0), 10.0, 30.0)
    trace_steps(m, p, s, 1)
    for z, lat in ((12.3, 0.4), (41.7, -0.7), (63.05, 1.1), (95.2, 0.05)):
        q0 = np.array([6370.0 + z, lat, 0.3])
        nu0 = np.array([0.3, -0.6, 0.55])
        y = np.concatenate([q0, nu0, np.zeros(12)])
        f = rhs(m, p, s, y)

        def nu_of(q, P):
            return np.array([P[0], P[1] / q[0], P[2] / (q[0] * math.cos(q[1]))])

        def H(q, P):
            nu = nu_of(q, P)
            c, u, v, _ = medium_point(m, tuple(q))
            return c * math.sqrt(nu @ nu) + v * nu[1] + u * nu[2]
        P0 = np.array([nu0[0], q0[0] * nu0[1], q0[0] * math.cos(q0[1]) * nu0[2]])
        hq = np.array([1e-5, 1e-7, 1e-7])
        hp = np.array([1e-6, 1e-3, 1e-3])
        qdot = np.array([_fd(lambda e: H(q0, P0 + e * np.eye(3)[i]), 0.0, hp[i]) for i in range(3)])
        pdot = -np.array([_fd(lambda e: H(q0 + e * np.eye(3)[i], P0), 0.0, hq[i]) for i in range(3)])
        r, t = q0[0], q0[1]
        nudot = np.array([
            pdot[0],
            pdot[1] / r - P0[1] * qdot[0] / r ** 2,
            pdot[2] / (r * math.cos(t)) - P0[2] * (qdot[0] * math.cos(t) - r * math.sin(t) * qdot[1]) / (r * math.cos(t)) ** 2,
        ])
        speed = math.sqrt(qdot[0] ** 2 + (r * qdot[1]) ** 2 + (r * math.cos(t) * qdot[2]) ** 2)
        assert np.allclose(f[:3], qdot / speed, rtol=1e-6, atol=1e-12)
        # The reference's T1, T2 terms (geoac.eqset.sph.strat.cpp:254-256) carry
        # 2 nu_r v and 2 nu_r u cos(lat) beyond Hamilton's equations (w = 0);
        # the radial equation is exactly Hamiltonian.
        _, u, v, _ = medium_point(m, tuple(q0))
        extra = np.array([0.0, -(1.0 / r) * 2.0 * nu0[0] * v, -(1.0 / (r * math.cos(t))) * 2.0 * nu0[0] * u * math.cos(t)])
        assert np.allclose(f[3:6], nudot / speed + extra / speed, rtol=1e-6, atol=1e-11)
        # windless: exactly Hamiltonian
        assert abs(extra[1]) > 1e-9 and abs(extra[2]) > 1e-9


# ----------------------------------------------------------------- invariants
@pytest.mark.parametrize("geom,windy,tol", [(GEOM_3D, True, 1e-6), (GEOM_SPH, False, 1e-6)])
def test_eikonal_residual_along_ray(oracle, pert_units, geom, windy, tol):
    # sph only windless: with winds the reference's lat/lon equations are not
    # Hamiltonian (extra 2 nu_r v terms) and the eikonal drifts at ~1e-4.
    z, c, u, v, rho = pert_units
    if not windy:
        u, v = 0 * u, 0 * v
    m = make_medium(oracle, geom, z, c, u, v, rho)
    p = m.params(True)
    s = src_struct(m, p, (0.0, 30.0, 0.0) if geom == GEOM_SPH else (0.0, 0.0, 0.0), 12.0, 75.0)
    rows, _ = trace_steps(m, p, s, 3000)
    res = [abs(eikonal(m, s, rows[k])) for k in range(0, 3001, 250)]
    assert max(res) < tol


def test_bouguer_invariant_sph(oracle, base_units):
    z, c, u, v, rho = base_units
    m = make_medium(oracle, GEOM_SPH, z, c, 0 * u, 0 * v, rho)
    p = m.params(True)
    s = src_struct(m, p, (0.0, 30.0, 0.0), 15.0, 40.0)
    rows, _ = trace_steps(m, p, s, 4000)
    inv = rows[:, 0] * np.hypot(rows[:, 4], rows[:, 5])
    assert np.max(np.abs(inv - inv[0])) / inv[0] < 1e-10


@pytest.mark.parametrize("geom", [GEOM_3D, GEOM_SPH])
def test_flat_ground_specular_reflection(oracle, pert_units, geom):
    m = make_medium(oracle, geom, *pert_units)
    p = m.params(True)
    s = src_struct(m, p, (0.0, 30.0, 0.0) if geom == GEOM_SPH else (0.0, 0.0, 0.0), 5.0, 75.0)
    rows, _ = trace_steps(m, p, s, 3)
    y0 = refl(m, p, s, rows[1], rows[2])
    iv = 0 if geom == GEOM_SPH else 2
    hv = [1, 2] if geom == GEOM_SPH else [0, 1]
    # slowness: vertical component flips, horizontal ones are unchanged
    assert y0[3 + iv] == pytest.approx(-(rows[2][3 + iv] + (rows[2][3 + iv] - rows[1][3 + iv])
                                          / (rows[2][iv] - rows[1][iv]) * ((6370.0 if geom == GEOM_SPH else 0.0) - rows[2][iv])), rel=1e-12)
    for h in hv:
        assert np.sign(y0[3 + h]) == np.sign(rows[2][3 + h])
    assert y0[iv] == (6370.0 if geom == GEOM_SPH else 0.0)


# --------------------------------------------------- auxiliary equations (FD)
def _aux_fd(oracle, units, geom, nsteps, th=20.0, az=45.0, d=1e-6):
    m = make_medium(oracle, geom, *units)
    p = m.params(True)
    pos = (5.0, 30.0, 0.0) if geom == GEOM_SPH else (0.0, 0.0, 5.0)
    rows, _ = trace_steps(m, p, src_struct(m, p, pos, th, az), nsteps)
    ddeg = math.degrees(d)
    rp, _ = trace_steps(m, p, src_struct(m, p, pos, th + ddeg, az), nsteps)
    rm, _ = trace_steps(m, p, src_struct(m, p, pos, th - ddeg, az), nsteps)
    # azimuth: phi = pi/2 - az, so d/dphi = -d/daz
    ap, _ = trace_steps(m, p, src_struct(m, p, pos, th, az - ddeg), nsteps)
    am, _ = trace_steps(m, p, src_struct(m, p, pos, th, az + ddeg), nsteps)
    return rows[-1], (rp[-1] - rm[-1]) / (2 * d), (ap[-1] - am[-1]) / (2 * d)


def test_auxiliary_equals_finite_difference_3d(oracle, pert_units):
    units = list(pert_units)
    units[3] = 0.0 * units[3]   # v = 0: the 3D dv[3] accumulation quirk is inert
    y, dth, dph = _aux_fd(oracle, units, GEOM_3D, 1500)
    for aux, fd in ((y[6:9], dth[:3]), (y[9:12], dth[3:6]), (y[12:15], dph[:3]), (y[15:18], dph[3:6])):
        scale = np.max(np.abs(fd)) + 1e-300
        assert np.max(np.abs(aux - fd)) / scale < 1e-7, (aux, fd)


def test_auxiliary_equals_finite_difference_sph_azimuth(oracle, pert_units):
    # windless: the azimuth derivative of |c_g| vanishes, so the term the
    # reference omits from its mu equations is zero and FD must match
    z, c, u, v, rho = pert_units
    y, dth, dph = _aux_fd(oracle, (z, c, 0 * u, 0 * v, rho), GEOM_SPH, 1500)
    for aux, fd in ((y[12:15], dph[:3]), (y[15:18], dph[3:6])):
        scale = np.max(np.abs(fd)) + 1e-300
        assert np.max(np.abs(aux - fd)) / scale < 1e-6, (aux, fd)


@pytest.mark.parametrize("geom", [GEOM_3D, GEOM_SPH])
def test_auxiliary_rhs_is_linearisation(oracle, pert_units, geom):
    """Local pin of the auxiliary right-hand side: it equals the directional
    derivative of the (Hamiltonian-pinned) ray right-hand side along (R, mu),
    except for two stated reference deviations:
      - sph mu equations omit + G/|c_g|^2 (d|c_g|) T
        (geoac.eqset.sph.strat.cpp:410-417 multiplies d|c_g| by the gradient
        terms only);
      - 3D adds the azimuth term into dv[3] (geoac.eqset.3d.strat.cpp:298)."""
    m = make_medium(oracle, geom, *pert_units)
    p6, p18 = m.params(False), m.params(True)
    s = src_struct(m, p18, (0.0, 30.0, 0.0) if geom == GEOM_SPH else (0.0, 0.0, 0.0), 10.0, 30.0)
    trace_steps(m, p18, s, 1)
    rng = np.random.default_rng(3)
    for z in (14.2, 47.9, 88.4):
        if geom == GEOM_SPH:
            base = np.array([6370.0 + z, 0.45, -0.2, 0.31, -0.52, 0.61])
        else:
            base = np.array([3.0, -7.0, z, 0.31, -0.52, 0.61])
        aux = rng.normal(size=12) * np.array([5.0, 1e-3, 1e-3, 0.2, 0.2, 0.2] * 2 if geom == GEOM_SPH else [5.0, 5.0, 5.0, 0.2, 0.2, 0.2] * 2)
        y = np.concatenate([base, aux])
        f = rhs(m, p18, s, y)
        for j in range(2):
            a = aux[6 * j: 6 * j + 6]
            h = 1e-6 / max(1.0, np.max(np.abs(a)))
            f6 = lambda e: rhs(m, p6, s, np.concatenate([base + e * a, np.zeros(12)]))[:6]
            d = (f6(h) - f6(-h)) / (2 * h)
            c, u, v, _ = medium_point(m, tuple(base[:3]))
            nu = base[3:6]

            def cgmag(e):
                yy = base + e * a
                cc, uu, vv, _ = medium_point(m, tuple(yy[:3]))
                n = yy[3:6]
                w = np.array([0.0, vv, uu]) if geom == GEOM_SPH else np.array([uu, vv, 0.0])
                return np.linalg.norm(cc * n / np.linalg.norm(n) + w)
            dcg = (cgmag(h) - cgmag(-h)) / (2 * h)
            cg = cgmag(0.0)
            expect = d.copy()
            if geom == GEOM_SPH:
                r, t = base[0], base[1]
                G = np.array([1.0, 1.0 / r, 1.0 / (r * math.cos(t))])
                f_ray = rhs(m, p6, s, np.concatenate([base, np.zeros(12)]))
                # S + T from the ray equations: f_nu = -G/|cg| (S + T)
                ST = -f_ray[3:6] * cg / G
                S = np.zeros(3)
                dwind = (medium_point(m, (r + 1e-5, t, 0))[1:3], medium_point(m, (r - 1e-5, t, 0))[1:3])
                du_dr = (dwind[0][0] - dwind[1][0]) / 2e-5
                dv_dr = (dwind[0][1] - dwind[1][1]) / 2e-5
                dc_dr = (medium_point(m, (r + 1e-5, t, 0))[0] - medium_point(m, (r - 1e-5, t, 0))[0]) / 2e-5
                S[0] = np.linalg.norm(nu) * dc_dr + nu[1] * dv_dr + nu[2] * du_dr
                T = ST - S
                expect[3:6] -= G / cg ** 2 * dcg * T
            else:
                zz = base[2]
                dv_dz = (medium_point(m, (0, 0, zz + 1e-5))[2] - medium_point(m, (0, 0, zz - 1e-5))[2]) / 2e-5
                if True:
                    # dv[3] also holds X_ph,z dv/dz and dv[4] stays zero: the
                    # theta system gains, the phi system lacks X_ph,z dv/dz in dc_g,y
                    dcgy = (1.0 if j == 0 else -1.0) * aux[6 + 2] * dv_dz
                    dcgm = (c * nu[1] / np.linalg.norm(nu) + v) * dcgy / cg
                    expect[0:3] += np.array([0.0, dcgy / cg, 0.0]) - np.array(
                        [c * nu[0] / np.linalg.norm(nu) + u, c * nu[1] / np.linalg.norm(nu) + v, c * nu[2] / np.linalg.norm(nu)]) / cg ** 2 * dcgm
                    S = np.linalg.norm(nu) * ((medium_point(m, (0, 0, zz + 1e-5))[0] - medium_point(m, (0, 0, zz - 1e-5))[0]) / 2e-5) \
                        + nu[0] * ((medium_point(m, (0, 0, zz + 1e-5))[1] - medium_point(m, (0, 0, zz - 1e-5))[1]) / 2e-5) + nu[1] * dv_dz
                    expect[5] += dcgm / cg ** 2 * S
            got = f[6 + 6 * j: 12 + 6 * j]
            scale = np.max(np.abs(expect)) + 1e-300
            assert np.max(np.abs(got - expect)) / scale < 1e-5, (geom, j, got, expect)


@pytest.mark.parametrize("th0", [-5.0, -12.0, -20.0])
def test_auxiliary_across_reflection_3d(oracle, th0):
    """Arrival-position derivative after one bounce from the auxiliary
    variables (R plus the ds0 ground-crossing term) vs finite differences of
    the exact ground intercepts, uniform steps, linear sound speed.  The
    reflection's mu jump term (-2 dnu_z/ds ds0) matters here: without it the
    two differ by 30 % to 1100 %; with it they agree to within the sampling
    noise of the finite difference (about 2 %)."""
    z, c, u, v, rho = linear(0.30, 0.002)
    m = make_medium(oracle, GEOM_3D, z, c, u, v, rho)
    p = m.params(True)
    p.ds_min = p.ds_max
    az0, d = 60.0, 1e-4
    import ctypes as C

    def run(th):
        s = src_struct(m, p, (0.0, 0.0, 3.0), th, az0)
        rows, _ = trace_steps(m, p, s, 0)
        prev2, prev1, legs = None, rows[0], 0
        for _ in range(200000):
            y1 = np.zeros(18)
            m.o.lib.go_rk4_step(m.h, C.byref(p), C.byref(s), prev1.ctypes.data_as(C.POINTER(C.c_double)),
                                y1.ctypes.data_as(C.POINTER(C.c_double)))
            if y1[2] < 0.0:
                if legs == 1:
                    f = prev1[2] / (prev1[2] - y1[2])
                    return prev1 + f * (y1 - prev1)
                prev2, prev1, legs = None, refl(m, p, s, prev2, prev1), 1
                continue
            prev2, prev1 = prev1, y1
        raise AssertionError("ray did not return")
    yi = run(th0)
    ip, im = run(th0 + math.degrees(d)), run(th0 - math.degrees(d))
    fd = (ip[:2] - im[:2]) / (2 * d)
    cc, uu, vv, _ = medium_point(m, tuple(yi[:3]))
    nu = yi[3:6]
    cg = cc * nu / np.linalg.norm(nu) + np.array([uu, vv, 0.0])
    cgm = np.linalg.norm(cg)
    ds0 = -yi[8] / (cg[2] / cgm)
    aux = np.array([yi[6] + cg[0] / cgm * ds0, yi[7] + cg[1] / cgm * ds0])
    assert np.linalg.norm(aux - fd) / np.linalg.norm(fd) < 0.03, (aux, fd)
