"""GPU parity of the 3D stratified profile over terrain (infraga-3d with
topo_file=, ga_atmo_load_strat_topo_xy / ga_atmo_load_file_topo_xy).

Against the oracle's 3D stratified-over-terrain restatement (fixture
tests/golden/strat_topo3d.npz, tests/golden/make_strat_topo3d_golden.py; the
oracle's medium is pinned bit-exact against the reference's compiled 3D
stratified medium with terrain, and its upper-triangle contraction against
the dropped terms, in tests/test_oracle_strat_topo3d.py):
- fans with 18 and 6 equations: arrival keys and ray-step counts exact;
  range, time, turning height within 1e-6 relative; TL and absorption
  within 0.01 dB; arrival inclination and back azimuth within 1e-6 degree;
- the write_rays / write_caustics bookkeeping: path-row keys exact, rows at
  the same bar (caustic rows floored at the terrain, as the stratified
  driver writes them);
- the eigenray search to two receivers: the oracle's eigenrays (keys exact,
  launch angles within 1e-4 degree);
- the front end: infraga-3d -prop, -eig_search and -eig_direct with
  topo_file= on the reference's file formats.
"""
import os
import subprocess

import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_3D_RNGDEP, GA_PATH_CAUSTICS, GA_PATH_RAYS, eig_params, prop_params
from parity import compare, compare_paths

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "strat_topo3d.npz")
BIN = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin")


def scene():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_strat_topo3d_golden as mk
    return mk.scene()


@pytest.fixture(scope="module")
def st3(engine):
    prof, topo, launches = scene()
    rid = engine.load_strat_topo(*prof, *topo, xy=True)
    yield rid, launches
    engine.rngdep_release(rid)


@pytest.mark.parametrize("case,amp", [("amp18", 1), ("amp6", 0)])
def test_fan_matches_oracle(engine, st3, case, amp):
    rid, launches = st3
    g = np.load(GOLDEN)
    ref = g[case]
    p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=amp, bounces=int(g["bounces"]))
    arr, st = engine.fan(rid, p, launches)
    print(f"\n{case}: {len(arr)} arrivals (oracle {len(ref)}), {st.ray_steps} ray-steps (oracle {int(g[case + '_steps'])})")
    assert len(ref) > 5
    fails, summ = compare(arr, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)
    assert st.ray_steps == int(g[case + "_steps"])
    np.testing.assert_allclose(arr["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(arr["back_az"], ref["back_az"], rtol=0, atol=1e-6)


def test_paths_match_oracle(engine, st3):
    rid, launches = st3
    g = np.load(GOLDEN)
    ref, rpts = g["paths18"], g["paths18_pts"]
    p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=1, bounces=int(g["bounces"]))
    arr, pts, st = engine.fan_paths(rid, p, launches, GA_PATH_RAYS | GA_PATH_CAUSTICS)
    assert (rpts["kind"] == 1).sum() > 0
    fails, summ = compare(arr, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)
    pf, ps = compare_paths(pts, rpts, False, src=(0.0, 0.0, 0.0))
    assert not pf, (pf, ps)


def test_eig_search_matches_oracle(engine, st3):
    rid, _ = st3
    g = np.load(GOLDEN)
    ref = g["eigenrays"]
    e = eig_params(incl_min=0.5, incl_max=float(g["incl_max"]), bnc_min=0, bnc_max=1)
    p = prop_params(GA_GEOM_3D_RNGDEP, src=(0.0, 0.0, 0.0))
    eig, lat, st = engine.eig_search(rid, p, e, g["receivers"])
    print(f"\n{len(eig)} eigenrays (oracle {len(ref)}), {st.ray_steps} ray-steps")
    key = lambda a: [(int(r["launch_idx"]), int(r["n_bounce"])) for r in a]
    assert key(eig) == key(ref) and len(ref) > 0
    np.testing.assert_allclose(eig["incl"], ref["incl_launch"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(eig["az"], ref["az_launch"], rtol=0, atol=1e-4)
    for f in ("time", "turn_ht"):
        np.testing.assert_allclose(eig[f], ref[f], rtol=1e-6)
    np.testing.assert_allclose(eig["tl"], ref["tl"], atol=0.01)


def _write_files(d):
    prof, topo, _ = scene()
    z, c, u, v, rho = prof
    met = os.path.join(d, "strat.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    tx, ty, zg = topo
    tp = os.path.join(d, "topo.dat")
    with open(tp, "w") as f:
        for i, x in enumerate(tx):
            f.write("".join("%.17g %.17g %.17g\n" % (x, y, zg[i, j]) for j, y in enumerate(ty)))
    return met, tp


def test_cli_prop_eig_search_and_direct_with_topo_file(engine, oracle, tmp_path):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_strat_topo3d_golden as mk
    met, tp = _write_files(str(tmp_path))
    exe = os.path.join(BIN, "infraga-3d")
    r = subprocess.run([exe, "-prop", met, f"topo_file={tp}", "prof_format=zcuvd", "bounces=3", "write_rays=false",
                        "incl_min=9", "incl_max=14", "incl_step=5", "azimuth=100"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(os.path.join(str(tmp_path), "strat.arrivals.dat")).read().splitlines()
    assert any(l.startswith("#\ttopo file:") for l in lines)
    rows = [[float(t) for t in l.split("\t")] for l in lines if l.strip() and not l.startswith("#")]
    m = mk.medium()
    ref, _, _ = m.fan3([(9.0, 100.0), (14.0, 100.0)], m.params(True), bounces=3)
    assert len(rows) == len(ref) > 0
    for row, a in zip(rows, ref):
        exp = [a["incl_launch"], a["az_launch"], a["n_bounce"], a["pos0"], a["pos1"], a["time"], a["cel"],
               a["turn_ht"], a["incl_arr"], a["back_az"], a["tl"], a["absorption"]]
        np.testing.assert_allclose(row, exp, rtol=1e-5, atol=1e-4)
    g = np.load(GOLDEN)
    rc = g["receivers"][0]
    r = subprocess.run([exe, "-eig_search", met, f"topo_file={tp}", "prof_format=zcuvd", f"rcvr_x={rc[0]}",
                        f"rcvr_y={rc[1]}", "bnc_max=1", f"incl_max={float(g['incl_max'])}"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in open(os.path.join(str(tmp_path), "strat.arrivals.dat"))
            if l.strip() and not l.startswith("#")]
    ref = g["eigenrays"][g["eigenrays"]["launch_idx"] == 0]
    assert len(rows) == len(ref) > 0
    for i, row in enumerate(rows):
        assert abs(float(row[0]) - ref["incl_launch"][i]) < 1e-4
    a0 = ref[0]
    r = subprocess.run([exe, "-eig_direct", met, f"topo_file={tp}", "prof_format=zcuvd", f"rcvr_x={rc[0]}",
                        f"rcvr_y={rc[1]}", f"incl_est={a0['incl_launch'] + 0.05}", f"az_est={a0['az_launch']}",
                        "output_id=direct"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Eigenray identified" in r.stdout
