"""GPU parity of the 3D range-dependent eigenray search (ga_eig_search /
ga_eig_direct with GA_GEOM_3D_RNGDEP, infraga-3d-rngdep -eig_search /
-eig_direct).

- Windless linear c(z) columns over flat ground: exactly one eigenray per
  admissible bounce count at the closed-form theta_n = atan(g D / (2 c0
  (n + 1))) within the acceptance disc (the pin of the oracle's search,
  applied to the GPU's).
- Terrain grid of tests/golden/rd3eig.npz (written by make_rd3eig_golden.py
  from the oracle's go_rd3_eig_search): the same eigenray count per
  (receiver, bounces), launch angles within 1e-4 deg, and every GPU row and
  its eigenray-N.dat rows against the oracle's accepted-eigenray trace at the
  GPU's own launch angles (go_rd3_eig_path) at the arrival bar.
- -eig_direct from an estimate near a found eigenray converges to it.
- The front end writes the arrivals and eigenray files in the 3D layout.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_3D_RNGDEP, eig_params, prop_params
from parity import compare_paths

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "rd3eig.npz")
C0, G = 0.30, 0.002


def test_linear_gradient_closed_form(engine):
    x = np.array([-800.0, -200.0, 400.0, 800.0])
    y = np.array([-800.0, 0.0, 800.0])
    z = np.arange(0.0, 150.0 + 1e-9, 0.5)
    rep = lambda f: np.ascontiguousarray(np.broadcast_to(f, (x.size, y.size, f.size)))
    rid = engine.rngdep_load_xy(x, y, z, rep(C0 + G * z), rep(0.0 * z), rep(0.0 * z), rep(1.2e-3 * np.exp(-z / 7.0)))
    D, az = 310.0, 200.0
    rcvr = (D * math.sin(math.radians(az)), D * math.cos(math.radians(az)))
    p = prop_params(GA_GEOM_3D_RNGDEP, src=(0.0, 0.0, 0.0), rng_max=2000.0)
    arr, _, st = engine.eig_search(rid, p, eig_params(bnc_min=0, bnc_max=3, incl_step_max=0.5), [rcvr])
    engine.rngdep_release(rid)
    theta = {n: math.degrees(math.atan(G * D / (2.0 * C0 * (n + 1)))) for n in range(4)}
    apex = lambda t: (C0 / G) * (1.0 / math.cos(math.radians(t)) - 1.0)
    swept = [n for n in range(4) if 0.5 <= theta[n] <= 45.0 and apex(theta[n]) > 0.2]
    print(f"\n{len(arr)} eigenrays {arr['n_bounce'].tolist()}, {st.ray_steps} ray-steps")
    assert sorted(arr["n_bounce"].tolist()) == swept
    for a in arr:
        n = int(a["n_bounce"])
        target = theta[min(n, 1)]     # the 3D find_eigenray reflects once per trace
        m_hops = min(n, 1) + 1
        dth = 0.1 / (m_hops * (C0 / G) * 2.0 / math.cos(math.radians(target)) ** 2)
        assert abs(math.radians(a["incl"] - target)) < 2.0 * dth + 2e-6, (n, a["incl"], target)
        daz = (a["az"] - az + 180.0) % 360.0 - 180.0
        assert abs(math.radians(daz)) * D < 0.11


@pytest.fixture(scope="module")
def terrain3d(engine):
    from test_rngdep3d_gpu import workload
    units, topo, _ = workload()
    rid = engine.rngdep_load_xy(*units)
    engine.topo_load(rid, *topo)
    yield rid, units, topo
    engine.rngdep_release(rid)


def test_terrain_search_matches_oracle(engine, oracle, terrain3d):
    from ga_oracle_py import Rd3Medium
    rid, units, topo = terrain3d
    g = np.load(GOLDEN)
    ref = g["eigenrays"]
    e = eig_params(incl_min=0.5, incl_max=float(g["incl_max"]), bnc_min=0, bnc_max=1)
    p = prop_params(GA_GEOM_3D_RNGDEP, src=(0.0, 0.0, 0.0))
    eig, pts, lat, st = engine.eig_search_paths(rid, p, e, g["receivers"])
    print(f"\n{len(eig)} eigenrays (oracle {len(ref)}), {st.ray_steps} ray-steps, {st.kernel_ms:.0f} ms")
    key = lambda a: [(int(r["launch_idx"]), int(r["n_bounce"])) for r in a]
    assert key(eig) == key(ref)
    np.testing.assert_allclose(eig["incl"], ref["incl_launch"], rtol=0, atol=1e-4)
    np.testing.assert_allclose(eig["az"], ref["az_launch"], rtol=0, atol=1e-4)
    m = Rd3Medium(oracle, *units)
    m.set_topo(*topo)
    op = m.params(True)
    for i, a in enumerate(eig):
        rp, row = m.eig_path(op, (0.0, 0.0, 0.0), a["incl"], a["az"], int(a["n_bounce"]))
        for f in ("time", "turn_ht"):
            assert abs(a[f] - row[f]) <= 1e-6 * abs(row[f]), (i, f, a[f], row[f])
        assert math.hypot(a["pos0"] - row["pos0"], a["pos1"] - row["pos1"]) <= 1e-6 * 300.0, (i, "pos")
        for f in ("tl", "absorption"):
            assert abs(a[f] - row[f]) <= 0.01, (i, f, a[f], row[f])
        assert abs(a["incl_arr"] - row["incl_arr"]) <= 1e-6, i
        assert abs(a["back_az"] - row["back_az"]) <= 1e-6, i
        assert math.hypot(a["pos0"] - g["receivers"][int(a["launch_idx"])][0],
                          a["pos1"] - g["receivers"][int(a["launch_idx"])][1]) < 0.1
        q = pts[pts["launch_idx"] == i]
        rp["launch_idx"] = i
        pf, ps = compare_paths(q, rp, False, src=(0.0, 0.0, 0.0))
        assert not pf, (i, pf, ps)


def test_terrain_direct_matches_search(engine, terrain3d):
    """-eig_direct from an estimate near a found eigenray converges to it."""
    rid, _, _ = terrain3d
    g = np.load(GOLDEN)
    ref = g["eigenrays"][0]
    p = prop_params(GA_GEOM_3D_RNGDEP, src=(0.0, 0.0, 0.0))
    row, pts, st = engine.eig_direct(rid, p, eig_params(), g["receivers"][int(ref["launch_idx"])],
                                     ref["incl_launch"] + 0.05, ref["az_launch"], int(ref["n_bounce"]))
    assert row is not None and len(pts) > 10
    assert abs(row["incl"] - ref["incl_launch"]) < 0.01 and abs(row["az"] - ref["az_launch"]) < 0.01


def test_cli_3d_rngdep_eig_search_and_direct(engine, tmp_path, terrain3d):
    from ga_oracle_py import write_rd_files
    from test_rngdep3d_gpu import workload
    units, topo, _ = workload()
    x, y, z, c, u, v, rho = units
    pre, xs, ys, tp = write_rd_files(str(tmp_path), x, y, z, c * 1000.0, u * 1000.0, v * 1000.0, rho, topo=topo)
    g = np.load(GOLDEN)
    rc = g["receivers"][1]
    exe = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin", "infraga-3d-rngdep")
    r = subprocess.run([exe, "-eig_search", pre, xs, ys, f"topo_file={tp}", "prof_format=zcuvd", f"rcvr_x={rc[0]}",
                        f"rcvr_y={rc[1]}", "src_x=0.0", "src_y=0.0", "bnc_max=1", f"incl_max={float(g['incl_max'])}"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(pre + ".arrivals.dat").read().splitlines()
    assert lines[0] == "# infraga-3d-rngdep -eig_search summary:"
    assert any(l.startswith("# \tprofile grid files: ") for l in lines)
    assert any(l.startswith("# incl [deg]\taz [deg]\tn_b\tx_0 [km]") for l in lines)
    rows = [l.split("\t") for l in lines if l.strip() and not l.startswith("#")]
    ref = g["eigenrays"][g["eigenrays"]["launch_idx"] == 1]
    assert len(rows) == len(ref)
    for i, row in enumerate(rows):
        assert abs(float(row[0]) - ref["incl_launch"][i]) < 1e-4
        hdr = open(pre + f".eigenray-{i}.dat").readline()
        assert hdr.startswith("# x [km]\ty [km]\tz [km]")
    a0 = ref[0]
    r = subprocess.run([exe, "-eig_direct", pre, xs, ys, f"topo_file={tp}", "prof_format=zcuvd", f"rcvr_x={rc[0]}",
                        f"rcvr_y={rc[1]}", "src_x=0.0", "src_y=0.0", f"incl_est={a0['incl_launch'] + 0.05}",
                        f"az_est={a0['az_launch']}", "output_id=direct"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Eigenray identified" in r.stdout
    assert os.path.exists(os.path.join(str(tmp_path), "direct.eigenray-0.dat"))
