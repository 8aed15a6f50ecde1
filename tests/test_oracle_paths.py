"""Pins of the oracle's write_rays / write_caustics bookkeeping
(go_fan_paths, go_rd_fan_paths, go_eig_path).

- The reference's own probe of its 90-ray ToyAtmo fan (inclinations
  0.5..45 step 0.5, azimuth -90, 2 bounces, 800 km, source 30 N 100 W) took
  1,399,246 RK4 steps and wrote 93,397 rows to raypaths.dat: the oracle
  gives the same step count, row count and 233 arrivals.
- Homogeneous medium (straight rays, pinned elsewhere): every sampled row m
  carries time s_m / c and TL 20 log10(1 / s_m), with s_m the arc length of
  the RK4 rows traced independently; the arrival carries s_{k-1} / c, i.e.
  the segment into the sub-ground row k is left out, while write_rays=false
  includes it (s_k / c).  No caustics (D = s^2 cos(theta) > 0).
- Caustic records sit exactly where the Jacobian determinant of consecutive
  RK4 rows of the same leg changes sign.
- The range-dependent driver over identical columns reduces to the
  stratified one (same rows, arrivals and caustics).
"""
import math

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, GEOM_SPH, PATH_CAUSTICS, PATH_RAYS, Medium
from oracle_drive import jacobian, make_medium, src_struct, trace_steps


def homogeneous(c=0.34):
    z = np.arange(0.0, 150.01, 1.0)
    return z, np.full_like(z, c), np.zeros_like(z), np.zeros_like(z), np.full_like(z, 1.0e-3)


def reference_fan90():
    """run_prop's accumulated inclination loop (main.infraga.sph.cpp:434)."""
    L, th = [], 0.5
    while th <= 45.0:
        L.append((th, -90.0))
        th += 0.5
    return np.array(L)


def test_raypath_rows_match_reference_probe(oracle, toyatmo_units):
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    p = m.params(True)
    p.rng_max = 800.0
    L = reference_fan90()
    assert len(L) == 90
    arr, pts, steps = m.fan_paths(L, p, PATH_RAYS, bounces=2, src=(30.0, -100.0, 0.0))
    assert steps == 1399246
    assert len(pts) == 93397
    assert len(arr) == 233
    assert (pts["kind"] == 0).all()
    # rows of each ray come in (leg, step) order; first row of every leg is m = 1
    for r in (0, 45, 89):
        q = pts[pts["launch_idx"] == r]
        key = q["leg"] * 10**7 + q["step"]
        assert (np.diff(key) > 0).all()
        for leg in np.unique(q["leg"]):
            st = q["step"][q["leg"] == leg]
            assert st[0] == 1 and (st[1:] % 15 == 0).all()


@pytest.mark.parametrize("geom", [GEOM_3D, GEOM_SPH])
def test_homogeneous_path_rows_and_arrival_closed_form(oracle, geom):
    c = 0.34
    m = make_medium(oracle, geom, *homogeneous(c))
    p = m.params(True)
    th, az = -30.0, 90.0
    if geom == GEOM_3D:
        src_fan, src_tr = (0.0, 0.0, 5.0), (0.0, 0.0, 5.0)
    else:
        src_fan, src_tr = (10.0, 20.0, 5.0), (5.0, 10.0, 20.0)
    arr_w, pts, _ = m.fan_paths([(th, az)], p, PATH_RAYS | PATH_CAUSTICS, bounces=0, src=src_fan)
    arr_l, _ = m.fan([(th, az)], p, bounces=0, src=src_fan)
    # independent trace of the same ray: rows and arc lengths
    s = src_struct(m, p, src_tr, th, az)
    rows, path = trace_steps(m, p, s, 400)
    z = rows[:, 2] if geom == GEOM_3D else rows[:, 0] - 6370.0
    k = int(np.argmax(z < 0.0))
    assert k > 20 and (z[:k] >= 0.0).all()
    # (spherical segment lengths from coordinate differences: O(ds^3 / r^2) off the arc)
    assert arr_w["time"][0] == pytest.approx(path[k - 1] / c, rel=1e-10)
    assert arr_l["time"][0] == pytest.approx(path[k] / c, rel=1e-10)
    assert arr_l["time"][0] - arr_w["time"][0] == pytest.approx((path[k] - path[k - 1]) / c, abs=1e-9)
    ms = [mm for mm in range(1, k) if mm == 1 or mm % 15 == 0]
    assert list(pts["step"]) == ms
    assert (pts["kind"] == 0).all() and (pts["leg"] == 0).all()
    for q, mm in zip(pts, ms):
        assert q["time"] == pytest.approx(path[mm] / c, rel=1e-10)
        assert q["tl"] == pytest.approx(20.0 * math.log10(1.0 / path[mm]), abs=1e-9)
        assert q["z"] == pytest.approx(z[mm], abs=1e-12)
        assert q["absorption"] <= 0.0


def test_caustics_at_jacobian_sign_changes(oracle, toyatmo_units):
    """Caustic rows of a ToyAtmo ray are the rows m of its first leg where
    D(row m-1) * D(row m) < 0, from an independent trace of the leg."""
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    p = m.params(True)
    src = (30.0, -100.0, 0.0)
    found = 0
    for th in (8.0, 12.0, 20.0, 33.0):
        arr, pts, _ = m.fan_paths([(th, -90.0)], p, PATH_CAUSTICS, bounces=0, src=src)
        s = src_struct(m, p, (0.0, 30.0, -100.0), th, -90.0)
        rows, _ = trace_steps(m, p, s, 60000)
        z = rows[:, 0] - 6370.0
        k = int(np.argmax(z[1:] < 0.0)) + 1
        D = np.array([jacobian(m, rows[i]) for i in range(1, k)])
        expect = [i + 1 for i in range(1, len(D)) if D[i] * D[i - 1] < 0.0]
        got = list(pts["step"][pts["kind"] == 1])
        assert got == expect, th
        found += len(got)
    assert found > 0


def test_eig_path_rows_match_fan_rows(oracle, toyatmo_units):
    """An eigenray's path file rows (go_eig_path) are the write_rays rows of
    the same launch traced through the same reflections (spherical)."""
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    p = m.params(True)
    src = (30.0, -100.0, 0.0)
    incl, az = 4.1385204, -84.970504
    pts_e, row = m.eig_path(p, src, incl, az, 1)
    _, pts_f, _ = m.fan_paths([(incl, az)], p, PATH_RAYS, bounces=1, src=src)
    q = pts_f[pts_f["leg"] <= 1]
    assert len(pts_e) == len(q)
    for f in ("leg", "step"):
        assert (pts_e[f] == q[f]).all()
    for f in ("pos0", "pos1", "z", "tl", "absorption", "time"):
        assert np.allclose(pts_e[f], q[f], rtol=1e-12, atol=1e-12), f
    assert row["n_bounce"] == 1
    assert row["time"] == pytest.approx(pts_e["time"][-1], rel=1e-3)


def test_rngdep_paths_reduce_to_stratified(oracle, pert_units):
    """Range-dependent write-mode driver over a grid of identical columns =
    the stratified write-mode driver (rows, caustics, arrivals)."""
    from ga_oracle_py import RdMedium
    z, c, u, v, rho = pert_units
    lat = np.array([20.0, 30.0, 40.0])
    lon = np.array([-110.0, -100.0, -90.0, -80.0])

    def rep(f):
        return np.ascontiguousarray(np.broadcast_to(f, (3, 4, f.size)))
    rd = RdMedium(oracle, lat, lon, z, rep(c), rep(u), rep(v), rep(rho))
    st = Medium(oracle, GEOM_SPH, z, c, u, v, rho)
    prd, pst = rd.params(True), st.params(True)
    for k in ("lat_min", "lat_max", "lon_min", "lon_max", "alt_max"):
        setattr(pst, k, getattr(prd, k))
    L = np.array([[th, az] for az in (-90.0, 135.0) for th in (3.0, 15.0, 31.0)])
    src = (30.0, -100.0, 0.0)
    a1, p1, s1 = rd.fan_paths(L, prd, PATH_RAYS | PATH_CAUSTICS, bounces=2, src=src)
    a2, p2, s2 = st.fan_paths(L, pst, PATH_RAYS | PATH_CAUSTICS, bounces=2, src=src)
    assert s1 == s2 and len(p1) == len(p2) > 0 and len(a1) == len(a2) > 0
    for f in ("launch_idx", "leg", "step", "kind"):
        assert np.array_equal(p1[f], p2[f]), f
    for f in ("pos0", "pos1", "z", "time", "absorption"):
        assert np.array_equal(p1[f], p2[f]), f
    np.testing.assert_allclose(p1["tl"], p2["tl"], rtol=1e-10, atol=1e-10)
    for f in ("launch_idx", "n_bounce", "pos0", "pos1", "time", "turn_ht", "absorption"):
        assert np.array_equal(a1[f], a2[f]), f
