"""GPU parity of the stratified profile over terrain (infraga-sph with
topo_file=, ga_atmo_load_strat_topo / ga_atmo_load_file_topo).

Against the oracle's stratified-over-terrain restatement (fixture
tests/golden/strat_topo.npz, tests/golden/make_strat_topo_golden.py; the
oracle's medium is pinned bit-exact against the reference's compiled
stratified medium with terrain in tests/test_oracle_strat_topo.py):
- fans with 18 and 6 equations and with the boundary-layer vertical wind:
  arrival keys and ray-step counts exact; range, time, turning height
  within 1e-6 relative; TL and absorption within 0.01 dB; arrival
  inclination (the stratified driver's sign) and back azimuth within 1e-6
  degree;
- the eigenray search to two receivers: the oracle's eigenrays (keys
  exact, launch angles within 1e-4 degree);
- the front end: infraga-sph -prop and -eig_search with topo_file= on the
  reference's file formats write the same rows, with the stratified
  driver's header ("topo file:", turn_ht_min).
"""
import os
import subprocess

import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_SPH_RNGDEP, eig_params, prop_params
from parity import compare

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "strat_topo.npz")
BIN = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin")


def scene():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_strat_topo_golden as mk
    return mk.scene()


@pytest.fixture(scope="module")
def st_atmo(engine):
    prof, topo, launches = scene()
    rid = engine.load_strat_topo(*prof, *topo)
    yield rid, launches
    engine.rngdep_release(rid)


@pytest.mark.parametrize("case,amp,blw", [("amp18", 1, 0), ("amp6", 0, 0), ("blw18", 1, 1)])
def test_fan_matches_oracle(engine, st_atmo, case, amp, blw):
    rid, launches = st_atmo
    g = np.load(GOLDEN)
    ref = g[case]
    p = prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=amp, bounces=int(g["bounces"]), topo_use_blw=blw)
    arr, st = engine.fan(rid, p, launches)
    print(f"\n{case}: {len(arr)} arrivals (oracle {len(ref)}), {st.ray_steps} ray-steps (oracle {int(g[case + '_steps'])})")
    assert len(ref) > 5
    fails, summ = compare(arr, ref, True, src=(40.0, -100.0))
    assert not fails, (fails, summ)
    assert st.ray_steps == int(g[case + "_steps"])
    np.testing.assert_allclose(arr["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(arr["back_az"], ref["back_az"], rtol=0, atol=1e-6)


def test_eig_search_matches_oracle(engine, st_atmo):
    rid, _ = st_atmo
    g = np.load(GOLDEN)
    ref = g["eigenrays"]
    e = eig_params(incl_min=0.5, incl_max=float(g["incl_max"]), bnc_min=0, bnc_max=1)
    p = prop_params(GA_GEOM_SPH_RNGDEP, src=(float("nan"), float("nan"), 0.0))
    eig, lat, st = engine.eig_search(rid, p, e, g["receivers"])
    print(f"\n{len(eig)} eigenrays (oracle {len(ref)}), {st.ray_steps} ray-steps")
    key = lambda a: sorted((int(r["launch_idx"]), int(r["n_bounce"])) for r in a)
    assert key(eig) == key(ref) and len(ref) > 0
    order = lambda a: np.lexsort((a["incl"] if "incl" in a.dtype.names else a["incl_launch"], a["n_bounce"], a["launch_idx"]))
    a, b = eig[order(eig)], ref[np.lexsort((ref["incl_launch"], ref["n_bounce"], ref["launch_idx"]))]
    np.testing.assert_allclose(a["incl"], b["incl_launch"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(a["az"], b["az_launch"], rtol=0, atol=1e-4)


def _write_files(d):
    prof, topo, _ = scene()
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


def test_cli_prop_with_topo_file(engine, tmp_path):
    met, tp = _write_files(str(tmp_path))
    g = np.load(GOLDEN)
    ref = g["amp18"]
    exe = os.path.join(BIN, "infraga-sph")
    r = subprocess.run([exe, "-prop", met, f"topo_file={tp}", "prof_format=zcuvd", "bounces=3", "write_rays=false",
                        "incl_min=4", "incl_max=22", "incl_step=5", "az_min=0", "az_max=300", "az_step=60"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    text = open(os.path.join(tmp_path, "strat.arrivals.dat")).read()
    assert "#\ttopo file:" + tp in text and "#\tturn_ht_min: 0.2" in text and "ground elevation" not in text
    rows = np.array([[float(x) for x in l.split("\t")] for l in text.splitlines() if l.strip() and not l.startswith("#")])
    # the reference's accumulated loops from the terrain-slope floor: compare
    # the rays the fixture traced (inclinations 4, 9, 14; azimuths 0, 60, ...)
    for a in ref[ref["az_launch"] % 60 == 0]:
        if a["incl_launch"] not in (4.0, 9.0, 14.0):
            continue
        hit = rows[(np.abs(rows[:, 0] - a["incl_launch"]) < 1e-9) & (np.abs(rows[:, 1] - a["az_launch"]) < 1e-9)
                   & (rows[:, 2] == a["n_bounce"])]
        assert len(hit) == 1, (a["incl_launch"], a["az_launch"], a["n_bounce"])
        h = hit[0]
        # 8 significant digits in the text; 1e-5 deg is ~1 m, inside the
        # 1e-6 relative range bar at these ranges
        assert abs(h[3] - a["pos0"]) < 1e-5 and abs(h[4] - a["pos1"]) < 1e-5
        assert abs(h[5] - a["time"]) <= 1e-6 * a["time"] and abs(h[8] - a["incl_arr"]) < 1e-5
        assert abs(h[10] - a["tl"]) < 0.01 and abs(h[11] - a["absorption"]) < 0.01


def test_cli_eig_search_with_topo_file(engine, tmp_path):
    met, tp = _write_files(str(tmp_path))
    g = np.load(GOLDEN)
    rc = g["receivers"][0]
    exe = os.path.join(BIN, "infraga-sph")
    r = subprocess.run([exe, "-eig_search", met, f"topo_file={tp}", "prof_format=zcuvd", f"rcvr_lat={rc[0]}",
                        f"rcvr_lon={rc[1]}", "bnc_max=1", f"incl_max={float(g['incl_max'])}"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    text = open(os.path.join(tmp_path, "strat.arrivals.dat")).read()
    assert "#\ttopo file:" + tp in text
    rows = [l.split("\t") for l in text.splitlines() if l.strip() and not l.startswith("#")]
    ref = g["eigenrays"][g["eigenrays"]["launch_idx"] == 0]
    assert len(rows) == len(ref) > 0
    got = sorted((int(r_[2]), float(r_[0])) for r_ in rows)
    exp = sorted((int(a["n_bounce"]), float(a["incl_launch"])) for a in ref)
    for (nb, th), (nb2, th2) in zip(got, exp):
        assert nb == nb2 and abs(th - th2) < 1e-4
