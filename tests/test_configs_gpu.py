"""GPU parity at the exact parameters of the benchmark configurations, on
strided subsets of their launch lists (the oracle finishes them in seconds).

- C1: base profile, 3D, source at the origin on the ground, 2 bounces,
  800 km, 18 equations; every 30th ray of the 90 x 72 launch list.
- C2: perturbed profile (seeded), spherical, source 30 N 0 E on the ground,
  5 bounces, 1000 km, 18 equations; every 1620th ray of the 450 x 360 list.
- C3: the range-dependent grid (11 x 11 profiles x 1501 levels over 20 x 20
  degrees) with the 401 x 401 fractal terrain, source at the region centre,
  4 bounces, 18 equations; every 9000th ray of the 200 x 360 list.
Bar: identical (launch, bounce) keys and per-fan ray-step counts; range,
turning height and travel time within 1e-6 relative; TL and absorption
within 0.01 dB.
"""
import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, prop_params
from parity import compare

pytestmark = pytest.mark.gpu


def _check(arr, st, ref, steps, sph, src):
    fails, summ = compare(arr, ref, sph, src=src)
    print(f"\n{len(arr)} arrivals (oracle {len(ref)}), {st.ray_steps} ray-steps (oracle {steps}); worst {summ['worst']}")
    assert len(ref) > 0
    assert not fails, (fails, summ)
    assert st.ray_steps == steps


def test_c1_3d_stratified_exact_parameters(engine, oracle, base_units):
    from ga_oracle_py import GEOM_3D, Medium
    L = synth.fan_launches(0.5, 45.0, 0.5, 0.0, 355.0, 5.0)[::30]
    aid = engine.load_atmo(*base_units)
    try:
        p = prop_params(GA_GEOM_3D, calc_amp=1, bounces=2, src=(0.0, 0.0, 0.0), rng_max=800.0)
        arr, st = engine.fan(aid, p, L)
    finally:
        engine.release_atmo(aid)
    m = Medium(oracle, GEOM_3D, *base_units)
    op = m.params(True)
    op.rng_max = 800.0
    ref, steps = m.fan(L, op, bounces=2, src=(0.0, 0.0, 0.0))
    _check(arr, st, ref, steps, False, (0.0, 0.0))


def test_c2_sph_fan_exact_parameters(engine, oracle, pert_units):
    from ga_oracle_py import GEOM_SPH, Medium
    L = synth.fan_launches(0.1, 45.0, 0.1, 0.0, 359.0, 1.0)
    assert len(L) == 162000
    L = L[::1620]
    aid = engine.load_atmo(*pert_units)
    try:
        p = prop_params(GA_GEOM_SPH, calc_amp=1, bounces=5, src=(30.0, 0.0, 0.0), rng_max=1000.0)
        arr, st = engine.fan(aid, p, L)
    finally:
        engine.release_atmo(aid)
    m = Medium(oracle, GEOM_SPH, *pert_units)
    ref, steps = m.fan(L, m.params(True), bounces=5, src=(30.0, 0.0, 0.0))
    _check(arr, st, ref, steps, True, (30.0, 0.0))


def test_c3_rngdep_terrain_exact_parameters(engine, oracle):
    from ga_oracle_py import RdMedium
    units = synth.rngdep_engine_units(*synth.rngdep_grid())
    topo = synth.fractal_topo()
    L = synth.fan_launches(0.2, 40.0, 0.2, 0.0, 359.0, 1.0)
    assert len(L) == 72000
    L = L[::9000]
    rid = engine.rngdep_load(*units)
    try:
        engine.topo_load(rid, *topo)
        arr, st = engine.fan(rid, prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=1, bounces=4), L)
    finally:
        engine.rngdep_release(rid)
    m = RdMedium(oracle, *units)
    m.set_topo(*topo)
    ref, steps = m.fan(L, m.params(True), bounces=4, src=None)
    lat0, lon0 = m.centre()
    _check(arr, st, ref, steps, True, (lat0, lon0))
