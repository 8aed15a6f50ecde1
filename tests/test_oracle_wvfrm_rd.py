"""Pins of the oracle's waveform over the range-dependent media
(ga_wvfrm.c go_wv_rd_med: wnl_wvfrm of geoac.eqset.{sph,3d}.rngdep.cpp and
run_wnl_wvfrm of main.infraga.{sph,3d}.rngdep.cpp).

Over identical columns and flat ground the range-dependent media return the
stratified ones' values (pinned in tests/test_oracle_rd.py and
tests/test_oracle_rd3.py), so the waveform must be the pinned stratified
waveform (tests/test_oracle_wvfrm.py): the same evolution steps and the
pressure within 1e-9 of its peak, for the spherical and the Cartesian grid,
with and without the stratified semantics (which only change the smallest
evolution step, 1e-6 km for the range-dependent sets against 1e-7 / 1e-5).
"""
import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, GEOM_SPH, Medium, Rd3Medium, RdMedium, WvParams, wv_build, wv_run


def _rep(f, n0, n1):
    return np.ascontiguousarray(np.broadcast_to(f, (n0, n1, f.size)))


@pytest.mark.parametrize("strat", [False, True])
def test_sph_identical_columns_equal_stratified_waveform(oracle, base_units, strat):
    z, c, u, v, rho = base_units
    lat, lon = np.array([20.0, 30.0, 40.0]), np.array([-10.0, 0.0, 10.0, 20.0])
    m = RdMedium(oracle, lat, lon, z, _rep(c, 3, 4), _rep(u, 3, 4), _rep(v, 3, 4), _rep(rho, 3, 4))
    if strat:
        oracle.lib.go_rd_h_set_strat(m.h, 1)
    s = Medium(oracle, GEOM_SPH, *base_units)
    pr, ps = m.params(True), s.params(True)
    for k in ("lat_min", "lat_max", "lon_min", "lon_max", "alt_max"):
        setattr(ps, k, getattr(pr, k))
    wv = WvParams.default(len=1024)
    t, p, _ = wv_build(oracle, "impulse", p0=500.0, t0=0.1, alpha=4.0, length=1024)
    for incl, az in ((12.0, 90.0), (25.0, 90.0)):
        r = wv_run(m, pr, wv, (30.0, 0.0, 0.0), incl, az, 1, t, p)
        q = wv_run(s, ps, wv, (30.0, 0.0, 0.0), incl, az, 1, t, p)
        assert r["ok"] and q["ok"] and r["steps"] == q["steps"] > 100
        np.testing.assert_allclose(r["t_out"], q["t_out"], rtol=1e-12)
        np.testing.assert_allclose(r["t_init"], q["t_init"], rtol=1e-12)
        assert np.abs(r["p_out"] - q["p_out"]).max() <= 1e-9 * np.abs(q["p_out"]).max()


@pytest.mark.parametrize("strat", [False, True])
def test_cart_identical_columns_equal_stratified_waveform(oracle, base_units, strat):
    z, c, u, v, rho = base_units
    x = np.linspace(-600.0, 600.0, 5)
    y = np.linspace(-600.0, 600.0, 4)
    m = Rd3Medium(oracle, x, y, z, _rep(c, 5, 4), _rep(u, 5, 4), _rep(v, 5, 4), _rep(rho, 5, 4))
    if strat:
        oracle.lib.go_rd_h_set_strat(m.h, 1)
    s = Medium(oracle, GEOM_3D, *base_units)
    pr, ps = m.params(True), s.params(True)
    wv = WvParams.default(len=1024)
    t, p, _ = wv_build(oracle, "Uwave", p0=300.0, t0=0.2, alpha=2.0, length=1024)
    for incl, az in ((6.0, 90.0), (18.0, 60.0)):
        r = wv_run(m, pr, wv, (0.0, 0.0, 0.0), incl, az, 1, t, p)
        q = wv_run(s, ps, wv, (0.0, 0.0, 0.0), incl, az, 1, t, p)
        assert r["ok"] and q["ok"] and r["steps"] == q["steps"] > 100
        np.testing.assert_allclose(r["t_out"], q["t_out"], rtol=1e-10)
        assert np.abs(r["p_out"] - q["p_out"]).max() <= 1e-8 * np.abs(q["p_out"]).max()
