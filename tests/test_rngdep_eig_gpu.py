"""GPU parity of the range-dependent eigenray search (ga_eig_search /
ga_eig_direct with GA_GEOM_SPH_RNGDEP, infraga-sph-rngdep -eig_search).

- Identical ToyAtmo columns over flat ground: the same eigenrays as the GPU
  stratified search (the reference's own ToyAtmo eigenrays).
- Terrain grid of tests/golden/rdeig.npz (written by make_rdeig_golden.py
  from the oracle's go_rd_eig_search): the same eigenray count per
  (receiver, bounces), launch angles within 1e-4 deg, and every GPU row and
  its eigenray-N.dat rows against the oracle's accepted-eigenray trace at the
  GPU's own launch angles (go_rd_eig_path) at the arrival bar.
- The front end writes the arrivals and eigenray files.
"""
import os
import subprocess

import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, eig_params, prop_params
from parity import compare_paths

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "rdeig.npz")


def test_identical_columns_equal_stratified_search(engine, toyatmo_units):
    z, c, u, v, rho = toyatmo_units
    lat = np.array([20.0, 30.0, 40.0])
    lon = np.array([-110.0, -100.0, -90.0, -80.0])

    def rep(f):
        return np.ascontiguousarray(np.broadcast_to(f, (3, 4, f.size)))
    rid = engine.rngdep_load(lat, lon, z, rep(c), rep(u), rep(v), rep(rho))
    aid = engine.load_atmo(*toyatmo_units)
    src = (30.0, -100.0, 0.0)
    e = eig_params(bnc_min=0, bnc_max=1)
    a_rd, _, st_rd = engine.eig_search(rid, prop_params(GA_GEOM_SPH_RNGDEP, src=src), e, [(30.25, -104.25)])
    a_st, _, st_st = engine.eig_search(aid, prop_params(GA_GEOM_SPH, src=src, bounds=(20.0, 40.0, -110.0, -80.0)), e,
                                       [(30.25, -104.25)])
    engine.rngdep_release(rid)
    engine.release_atmo(aid)
    print(f"\nrngdep {len(a_rd)} eigenrays {st_rd.ray_steps} steps; stratified {len(a_st)} {st_st.ray_steps}")
    assert len(a_rd) == len(a_st) == 3
    assert ["%.8g" % x for x in a_rd["incl"]] == ["34.559248", "4.1385204", "30.078055"]
    for f in ("incl", "az", "pos0", "pos1", "time", "turn_ht"):
        np.testing.assert_allclose(a_rd[f], a_st[f], rtol=1e-9, err_msg=f)
    np.testing.assert_allclose(a_rd["tl"], a_st["tl"], atol=1e-6)


@pytest.fixture(scope="module")
def terrain_grid(engine):
    from test_rngdep_fan_gpu import workload
    _, units, topo, _ = workload()
    rid = engine.rngdep_load(*units)
    engine.topo_load(rid, *topo)
    yield rid, units, topo
    engine.rngdep_release(rid)


def test_terrain_search_matches_oracle(engine, oracle, terrain_grid):
    from ga_oracle_py import RdMedium
    rid, units, topo = terrain_grid
    g = np.load(GOLDEN)
    ref = g["eigenrays"]
    e = eig_params(incl_min=0.5, incl_max=float(g["incl_max"]), bnc_min=0, bnc_max=1)
    p = prop_params(GA_GEOM_SPH_RNGDEP, src=(float("nan"), float("nan"), 0.0))
    eig, pts, lat, st = engine.eig_search_paths(rid, p, e, g["receivers"])
    print(f"\n{len(eig)} eigenrays (oracle {len(ref)}), {st.ray_steps} ray-steps, {st.kernel_ms:.0f} ms")
    key = lambda a: [(int(r["launch_idx"]), int(r["n_bounce"])) for r in a]
    assert key(eig) == key(ref)
    np.testing.assert_allclose(eig["incl"], ref["incl_launch"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(eig["az"], ref["az_launch"], rtol=0, atol=1e-4)
    m = RdMedium(oracle, *units)
    m.set_topo(*topo)
    la, lo = m.centre()
    op = m.params(True)
    for i, a in enumerate(eig):
        rp, row = m.eig_path(op, (la, lo, 0.0), a["incl"], a["az"], int(a["n_bounce"]))
        for f in ("time", "turn_ht"):
            assert abs(a[f] - row[f]) <= 1e-6 * abs(row[f]), (i, f, a[f], row[f])
        d = np.hypot(a["pos0"] - row["pos0"], (a["pos1"] - row["pos1"]) * np.cos(np.radians(row["pos0"])))
        assert d * 111.0 <= 1e-6 * 300.0, (i, "pos")
        for f in ("tl", "absorption"):
            assert abs(a[f] - row[f]) <= 0.01, (i, f, a[f], row[f])
        assert abs(a["incl_arr"] - row["incl_arr"]) <= 1e-6, i
        q = pts[pts["launch_idx"] == i]
        rp["launch_idx"] = i
        pf, ps = compare_paths(q, rp, True, src=(la, lo, 0.0))
        assert not pf, (i, pf, ps)


def test_terrain_direct_matches_search(engine, terrain_grid):
    """-eig_direct from an estimate near a found eigenray converges to it."""
    rid, _, _ = terrain_grid
    g = np.load(GOLDEN)
    ref = g["eigenrays"][0]
    p = prop_params(GA_GEOM_SPH_RNGDEP, src=(float("nan"), float("nan"), 0.0))
    row, pts, st = engine.eig_direct(rid, p, eig_params(), g["receivers"][int(ref["launch_idx"])],
                                     ref["incl_launch"] + 0.05, ref["az_launch"], int(ref["n_bounce"]))
    assert row is not None and len(pts) > 10
    assert abs(row["incl"] - ref["incl_launch"]) < 0.01 and abs(row["az"] - ref["az_launch"]) < 0.01


def test_cli_rngdep_eig_search(engine, tmp_path, terrain_grid):
    from infraga_amd import synth
    from ga_oracle_py import write_rd_files
    from test_rngdep_fan_gpu import workload
    grid, _, topo, _ = workload()
    lat, lon, z, T3, u3, v3, rho3 = grid
    pre, lats, lons, tp = write_rd_files(str(tmp_path), lat, lon, z, synth.sound_speed(T3), u3, v3, rho3, topo=topo)
    g = np.load(GOLDEN)
    rc = g["receivers"][1]
    exe = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin", "infraga-sph-rngdep")
    r = subprocess.run([exe, "-eig_search", pre, lats, lons, f"topo_file={tp}", "prof_format=zcuvd",
                        f"rcvr_lat={rc[0]}", f"rcvr_lon={rc[1]}", "bnc_max=1", f"incl_max={float(g['incl_max'])}"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in open(pre + ".arrivals.dat") if l.strip() and not l.startswith("#")]
    ref = g["eigenrays"][g["eigenrays"]["launch_idx"] == 1]
    assert len(rows) == len(ref)
    for i, row in enumerate(rows):
        assert abs(float(row[0]) - ref["incl_launch"][i]) < 1e-4
        assert os.path.exists(pre + f".eigenray-{i}.dat")
