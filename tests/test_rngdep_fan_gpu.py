"""GPU parity of the range-dependent spherical fan (GA_GEOM_SPH_RNGDEP).

- against the oracle's restatement of infraga-sph-rngdep -prop on a 6 x 6
  node grid with 0..2 km fractal terrain (fixture tests/golden/rngdep_fan.npz,
  written by tests/golden/make_rngdep_golden.py): 18 and 6 equations and the
  boundary-layer vertical wind; arrival keys exact, range, time, turning
  height within 1e-6 relative, TL and absorption within 0.01 dB, arrival
  inclination and back azimuth within 1e-6 degree;
- identical columns over flat ground reproduce the stratified spherical GPU
  fan (the two equation sets coincide there, up to the driver's
  inclination sign);
- the file loader (the reference's node files, prefix{n}.met, terrain rows)
  gives the same fan as the array loader;
- argument errors.
"""
import os

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, EngineError, prop_params
from parity import compare

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rngdep_fan.npz")


def workload():
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    units = synth.rngdep_engine_units(*grid)
    topo = synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81), amp_km=2.0)
    launches = np.array([[th, az] for az in (-90.0, 0.0, 45.0, 135.0, 270.0) for th in (3.0, 8.0, 12.0, 18.0)])
    return grid, units, topo, launches


@pytest.fixture(scope="module")
def rd(engine):
    grid, units, topo, launches = workload()
    rid = engine.rngdep_load(*units)
    engine.topo_load(rid, *topo)
    yield rid, launches
    engine.rngdep_release(rid)


@pytest.mark.parametrize("case,amp,blw", [("amp18", 1, 0), ("amp6", 0, 0), ("blw18", 1, 1)])
def test_fan_matches_oracle(engine, rd, case, amp, blw):
    rid, launches = rd
    ref = np.load(GOLDEN)[case]
    p = prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=amp, bounces=3, topo_use_blw=blw)
    arr, st = engine.fan(rid, p, launches)
    print(f"{case}: {len(arr)} arrivals (oracle {len(ref)}), {st.ray_steps} ray-steps "
          f"(oracle {int(np.load(GOLDEN)[case + '_steps'])}), {st.kernel_ms:.1f} ms")
    src = (40.0, -100.0)
    fails, summ = compare(arr, ref, True, src=src)
    assert not fails, (fails, summ)
    assert st.ray_steps == int(np.load(GOLDEN)[case + "_steps"])
    np.testing.assert_allclose(arr["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(arr["back_az"], ref["back_az"], rtol=0, atol=1e-6)
    assert (ref["n_bounce"] > 0).sum() >= 10   # terrain reflections are exercised


def test_uniform_grid_matches_stratified_gpu(engine, pert_units):
    z, c, u, v, rho = pert_units
    lat = np.array([20.0, 30.0, 40.0])
    lon = np.array([-110.0, -100.0, -90.0, -80.0])

    def rep(f):
        return np.ascontiguousarray(np.broadcast_to(f, (3, 4, f.size)))
    rid = engine.rngdep_load(lat, lon, z, rep(c), rep(u), rep(v), rep(rho))
    aid = engine.load_atmo(z, c, u, v, rho)
    try:
        L = np.array([[th, az] for az in (-90.0, 135.0) for th in (3.0, 15.0, 31.0)])
        pr = prop_params(GA_GEOM_SPH_RNGDEP, bounces=2, src=(30.0, -100.0, 0.0))
        buffer = 1.0e-3
        ps = prop_params(GA_GEOM_SPH, bounces=2, src=(30.0, -100.0, 0.0),
                         bounds=(20.0 + buffer, 40.0 - buffer, -110.0 + buffer, -80.0 - buffer))
        a, sa = engine.fan(rid, pr, L)
        b, sb = engine.fan(aid, ps, L)
        assert sa.ray_steps == sb.ray_steps and len(a) == len(b) > 0
        fails, summ = compare(a, b, True, src=(30.0, -100.0), rel=1e-10, db=1e-8)
        assert not fails, (fails, summ)
        np.testing.assert_allclose(a["incl_arr"], -b["incl_arr"], rtol=0, atol=1e-10)
    finally:
        engine.rngdep_release(rid)
        engine.release_atmo(aid)


def test_file_loader_matches_arrays(engine, rd, tmp_path):
    from ga_oracle_py import write_rd_files
    rid, launches = rd
    grid, units, topo, _ = workload()
    lat, lon, z, T3, u3, v3, rho3 = grid
    pre, lats, lons, tp = write_rd_files(str(tmp_path), lat, lon, z, synth.sound_speed(T3), u3, v3, rho3, topo=topo)
    fid = engine.rngdep_load_files(pre, lats, lons, tp, fmt="zcuvd")
    try:
        p = prop_params(GA_GEOM_SPH_RNGDEP, bounces=3)
        a, sa = engine.fan(rid, p, launches[:8])
        b, sb = engine.fan(fid, p, launches[:8])
        assert sa.ray_steps == sb.ray_steps
        for f in ("pos0", "pos1", "time", "tl", "absorption", "turn_ht"):
            assert np.array_equal(a[f], b[f]), f
    finally:
        engine.rngdep_release(fid)


def test_arguments(engine, rd, tmp_path):
    rid, launches = rd
    with pytest.raises(EngineError):
        engine.fan(rid + 1000, prop_params(GA_GEOM_SPH_RNGDEP), launches[:2])
    with pytest.raises(EngineError):
        engine.rngdep_load_files(str(tmp_path / "none"), str(tmp_path / "a.dat"), str(tmp_path / "b.dat"))
    arr, st = engine.fan(rid, prop_params(GA_GEOM_SPH_RNGDEP), np.zeros((0, 2)))
    assert len(arr) == 0 and st.ray_steps == 0
