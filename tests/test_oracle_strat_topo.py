"""Pins of the stratified-over-terrain oracle (infraga-sph with topo_file).

The oracle runs a stratified profile over terrain as its range-dependent
restatement over identical columns spanning the terrain grid, with the
stratified sources' boundary-layer top (z_max = terrain max + 0.01, then
+ 2 km: atmo_state.sph.strat.cpp:42-54) and the stratified driver's arrival
inclination sign (main.infraga.sph.cpp:523).  The equation set, reflection,
checks and eigenray sources of the two spherical binaries are the same code
(geoac.eqset.sph.{strat,rngdep}.cpp differ only in whitespace there;
geoac.eigenray.sph.cpp is shared), so what remains to pin is the medium:
- against the reference's own atmo_state.sph.strat.cpp / atmo_io.sph.strat.cpp
  compiled here (oracle/_ref/libref_sph.so, set_region with a terrain file),
  the propagation region, c, u, v, w, both calc_uvw overloads and the
  terrain with its derivatives are bit-identical at random points inside
  and above the boundary layer, with and without the boundary-layer
  vertical wind;
- the stratified driver's arrival inclination is the negative of the
  range-dependent driver's on the same rays, with everything else equal.
"""
import math
import os

import numpy as np
import pytest

from ga_oracle_py import GEOM_SPH, RdMedium, Ref
from infraga_amd import synth

R0 = 6370.0
HAVE_REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_sph.so"))


def scene(seed=3):
    z, c, u, v, rho = synth.to_engine_units(*synth.base_profile())
    tl, tn, zg = synth.fractal_topo(seed, lat=np.linspace(30.0, 36.0, 25), lon=np.linspace(-104.0, -97.0, 29),
                                    amp_km=1.5)
    return (z, c, u, v, rho), (tl, tn, zg)


def write_files(d, prof, topo):
    z, c, u, v, rho = prof
    met = os.path.join(d, "strat.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    tl, tn, zg = topo
    tp = os.path.join(d, "topo.dat")
    with open(tp, "w") as f:
        for i, la in enumerate(tl):
            f.write("".join("%.17g %.17g %.17g\n" % (la, lo, zg[i, j]) for j, lo in enumerate(tn)))
    return met, tp


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref_sph.so not built (no reference tree)")
@pytest.mark.parametrize("use_blw", [False, True])
def test_medium_matches_reference_strat_over_terrain(oracle, tmp_path, use_blw):
    prof, topo = scene()
    met, tp = write_files(str(tmp_path), prof, topo)
    ref = Ref(GEOM_SPH)
    ref.load_topo(met, tp)
    ref.lib.ref_use_blw(int(use_blw))
    # the oracle reads the same doubles the reference parsed from the text
    z, c, u, v, rho = prof
    m = RdMedium.strat_over_terrain(oracle, z, (c * 1e3) / 1000.0, (u * 1e3) / 1000.0, (v * 1e3) / 1000.0, rho,
                                    *topo)
    m.config(use_blw=use_blw)
    try:
        lim, rlim = m.limits(), ref.limits()
        assert np.array_equal(lim, rlim), (lim, rlim)
        assert abs(lim[7] - (topo[2].max() + 2.01)) < 1e-12
        rng = np.random.default_rng(17)
        for _ in range(400):
            la = math.radians(rng.uniform(30.0, 36.0))
            lo = math.radians(rng.uniform(-104.0, -97.0))
            zt = m.topo(la, lo)[0] - R0
            z_ = zt + rng.uniform(0.0, 3.0) if rng.random() < 0.7 else rng.uniform(0.0, 120.0)
            r = R0 + z_
            got = np.concatenate([[m.point(r, la, lo)[0], m.point(r, la, lo)[4]],
                                  m.calc_uvw(r, la, lo, False), m.calc_uvw(r, la, lo, True), m.topo(la, lo)[:6]])
            exp = np.concatenate([[ref.lib.ref_c(r, la, lo), ref.lib.ref_w(r, la, lo)],
                                  ref.calc_uvw(r, la, lo, False), ref.calc_uvw(r, la, lo, True), ref.topo(la, lo)])
            # identical columns reduce the hybrid splines to the 1D splines
            # exactly: bit-identical values
            assert np.array_equal(got, exp), (z_, np.flatnonzero(got != exp), got[got != exp], exp[got != exp])
    finally:
        ref.lib.ref_use_blw(0)
        ref.lib.ref_clear()


def test_arrival_inclination_sign(oracle):
    """Same rays through the range-dependent and the stratified semantics
    over the same columns and terrain: the two differ only in the
    boundary-layer top (+2.0 against +2.01 km above the terrain maximum) and
    the arrival inclination's sign.  On rows whose travel time agrees (the
    ray's near-ground legs did not see the 0.01 km band of different wind
    scaling) the arrival inclination is negated."""
    prof, topo = scene()
    tl, tn, zg = topo
    z, c, u, v, rho = prof
    strat = RdMedium.strat_over_terrain(oracle, z, c, u, v, rho, tl, tn, zg)
    col = [np.broadcast_to(f, (4, 4, len(z))).copy() for f in (c, u, v, rho)]
    rd = RdMedium(oracle, np.linspace(tl[0], tl[-1], 4), np.linspace(tn[0], tn[-1], 4), z, *col)
    rd.set_topo(tl, tn, zg)
    ls, lr = strat.limits(), rd.limits()
    assert ls[6] == lr[6] and abs(ls[7] - lr[7] - 0.01) < 1e-12
    L = np.array([[th, az] for az in (0.0, 120.0) for th in (4.0, 12.0, 25.0)])
    a_s, _ = strat.fan(L, strat.params(True), bounces=2, src=None)
    a_r, _ = rd.fan(L, rd.params(True), bounces=2, src=None)
    assert len(a_s) > 0
    same = [i for i in range(min(len(a_s), len(a_r)))
            if a_s["launch_idx"][i] == a_r["launch_idx"][i] and a_s["n_bounce"][i] == a_r["n_bounce"][i]
            and abs(a_s["time"][i] - a_r["time"][i]) < 1e-9 * a_r["time"][i]]
    assert same
    for i in same:
        assert abs(a_s["incl_arr"][i] + a_r["incl_arr"][i]) < 1e-9
