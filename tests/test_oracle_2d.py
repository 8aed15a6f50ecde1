"""Pins of the oracle's 2D effective-sound-speed geometry (infraga-2d,
geoac.eqset.2d.cpp), which the reference's build cannot run here (its
equation set includes <fftw3.h>).

- Linear sound speed c0 + g z, windless, constant step: every sampled row
  lies on the circular arc of the closed form and carries its amplitude
  1 / sqrt(D) with D = r (dr/ds dz/dtheta - dz/ds dr/dtheta) from the arc's
  analytic derivatives (so the auxiliary R, Z equations are right).
- Effective sound speed: a constant wind u0 along the azimuth (az 90 deg,
  phi = 0) or v0 (az 0 deg) shifts the arc to c_eff = c0 + u0 (+ v0).
- Windless, smooth profile: the 2D system equals the pinned 3D system in
  the vertical plane (rows, travel times, TL).
- Homogeneous medium: straight rays; the perpendicular drift of a leg is
  v_perp / (c + v_perp) per km of path for a cross wind; an elevated
  reflection altitude mirrors the ray (arrival at 2 h / tan(theta), bounce
  count unchanged); sloped terrain z = a r mirrors about the slope.
"""
import math

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, PATH_CAUSTICS, PATH_RAYS, Medium, Medium2D


def column(c, u=0.0, v=0.0, zt=150.0, dz=0.5, rho=1.0e-3):
    z = np.arange(0.0, zt + 1e-9, dz)
    cc = c(z) if callable(c) else np.full_like(z, c)
    return z, cc, np.full_like(z, u), np.full_like(z, v), np.full_like(z, rho)


def arc(c0, g, th, s):
    """Point at arc length s of the ray launched at th from (0, 0) in c = c0 + g z."""
    rh = c0 / (g * math.cos(th))
    return np.array([rh * math.sin(th) - rh * math.sin(th - s / rh), -rh * math.cos(th) + rh * math.cos(th - s / rh)])


def arc_dth(c0, g, th, s):
    """Analytic derivative of arc() with respect to the launch angle at fixed s."""
    rh = c0 / (g * math.cos(th))
    drh = rh * math.tan(th)
    psi = th - s / rh
    dpsi = 1.0 + s * math.tan(th) / rh
    dx = drh * (math.sin(th) - math.sin(psi)) + rh * (math.cos(th) - math.cos(psi) * dpsi)
    dz = -drh * math.cos(th) + rh * math.sin(th) + drh * math.cos(psi) - rh * math.sin(psi) * dpsi
    return np.array([dx, dz])


@pytest.mark.parametrize("wind", [(0.0, 0.0, 90.0), (0.015, 0.0, 90.0), (0.0, 0.01, 0.0), (0.015, 0.0, -90.0)],
                         ids=["windless", "u_along", "v_along", "u_against"])
def test_linear_gradient_arcs_and_amplitude(oracle, wind):
    c0, g, ds = 0.30, 0.002, 0.02
    u0, v0, az = wind
    phi = math.pi / 2 - math.radians(az)
    ce0 = c0 + u0 * math.cos(phi) + v0 * math.sin(phi)
    m = Medium2D(oracle, *column(lambda z: c0 + g * z, u0, v0), ds_min=ds, ds_max=ds)
    for th_deg in (10.0, 25.0):
        th = math.radians(th_deg)
        arr, pts, _ = m.fan([(th_deg, az)], bounces=0)
        q = pts[(pts["kind"] == 0) & (pts["leg"] == 0)]
        assert len(q) > 50
        for row in q[::7]:
            s = row["step"] * ds
            p = arc(ce0, g, th, s)
            assert abs(row["pos0"] - p[0]) < 1e-9 and abs(row["z"] - p[1]) < 1e-9, (th_deg, row["step"])
            dth = arc_dth(ce0, g, th, s)
            t = th - s / (ce0 / (g * math.cos(th)))
            D = p[0] * (math.cos(t) * dth[1] - math.sin(t) * dth[0])
            # amp = sqrt(rho c cos(th) / (rho0 c_eff0 D)) with the 2D driver's
            # atmo::c (not c_eff) in the Jacobian's dr/ds, dz/ds
            cz = c0 + g * row["z"]
            D_ref = D * cz / (cz + u0 * math.cos(phi) + v0 * math.sin(phi))
            amp = math.sqrt(abs(cz * math.cos(th) / (ce0 * D_ref)))
            assert row["tl"] == pytest.approx(20 * math.log10(amp), abs=1e-6), (th_deg, row["step"])


def test_windless_2d_equals_3d_in_the_vertical_plane(oracle):
    c0, g = 0.31, 0.0015
    col = column(lambda z: c0 + g * z)
    m2 = Medium2D(oracle, *col)
    m3 = Medium(oracle, GEOM_3D, *col)
    p3 = m3.params(True)
    L = [(th, 90.0) for th in (6.0, 15.0, 27.0)]
    a2, q2, s2 = m2.fan(L, bounces=1, flags=PATH_CAUSTICS)
    a3, q3, s3 = m3.fan_paths(L, p3, PATH_RAYS | PATH_CAUSTICS, bounces=1, src=(0.0, 0.0, 0.0))
    assert s2 == s3 and len(a2) == len(a3) == 6 and len(q2) == len(q3)
    assert (q2["step"] == q3["step"]).all() and (q2["leg"] == q3["leg"]).all() and (q2["kind"] == q3["kind"]).all()
    # same trajectories up to the rounding of two algebraically equal right-hand sides
    np.testing.assert_allclose(q2["pos0"], q3["pos0"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(q2["z"], q3["z"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(q2["time"], q3["time"], rtol=1e-9)
    np.testing.assert_allclose(q2["tl"], q3["tl"], atol=1e-6)
    np.testing.assert_allclose(a2["pos0"], a3["pos0"], rtol=1e-9)
    np.testing.assert_allclose(a2["time"], a3["time"], rtol=1e-9)
    np.testing.assert_allclose(a2["turn_ht"], a3["turn_ht"], rtol=1e-9)


def test_cross_wind_drift(oracle):
    c, v0, ds, h = 0.34, 0.012, 0.01, 5.003
    m = Medium2D(oracle, *column(c, 0.0, v0), ds_min=ds, ds_max=ds)
    arr, _, _ = m.fan([(-30.0, 90.0)], bounces=0, src_alt=h)
    assert len(arr) == 1
    f = v0 / (c + v0)              # v_perp / c_eff / sqrt(1 + 2 v/c + v^2/c^2)
    s_leg = arr["pos0"][0] / math.cos(math.radians(30.0))     # path length to the first sub-ground row
    assert arr["pos1"][0] == pytest.approx(f * s_leg, rel=1e-9)
    assert h / math.sin(math.radians(30.0)) <= s_leg < h / math.sin(math.radians(30.0)) + 2 * ds


def test_elevated_reflection_mirror(oracle):
    c, h = 0.34, 5.0
    m = Medium2D(oracle, *column(c), refl_alt=h, ds_min=0.01, ds_max=0.01)
    arr, pts, _ = m.fan([(30.0, 90.0)], bounces=1)
    assert list(arr["n_bounce"]) == [0, 1]
    r1 = 2 * h / math.tan(math.radians(30.0))
    assert arr["pos0"][0] == pytest.approx(r1, abs=0.02)
    assert arr["pos0"][1] == pytest.approx(2 * r1, abs=0.04)
    assert arr["turn_ht"][0] == pytest.approx(h, abs=0.01)
    # three legs (up, down after the elevated reflection, up after the ground)
    assert sorted(set(pts["leg"])) == [0, 1]


def test_sloped_terrain_mirror(oracle):
    c, a = 0.34, 0.05
    r = np.linspace(-10.0, 400.0, 411)
    m = Medium2D(oracle, *column(c), topo=(r, a * r), ds_min=0.01, ds_max=0.01)
    arr, pts, _ = m.fan([(-20.0, 90.0)], bounces=1, src_alt=3.0)
    assert len(arr) >= 1 and arr["n_bounce"][0] == 0
    q = pts[(pts["kind"] == 0) & (pts["leg"] == 1)]
    assert len(q) > 10
    ang = math.degrees(math.atan2(q["z"][8] - q["z"][3], q["pos0"][8] - q["pos0"][3]))
    assert ang == pytest.approx(2 * math.degrees(math.atan(a)) + 20.0, abs=1e-6)
