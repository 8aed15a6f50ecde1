"""GPU parity of the boundary-layer vertical wind (topo_use_BLw=true) in the
3D Cartesian geometries, against the oracle's fans over the same media with
use_blw on:
- range-dependent (GA_GEOM_3D_RNGDEP over ga_rngdep_load_xy, infraga-3d-rngdep):
  the 6 x 6 node grid with horizontally varying c, u, v, rho and fractal
  terrain of tests/test_rngdep3d_gpu.py;
- stratified over terrain (ga_atmo_load_strat_topo_xy, infraga-3d with
  topo_file=): the profile with near-ground winds over the fractal terrain of
  tests/golden/make_strat_topo3d_golden.py.
The oracle's Cartesian vertical wind is pinned in its stratified form
bit-exact against the reference's compiled 3D medium
(tests/test_oracle_strat_topo3d.py) and in its range-dependent form by
central differences of w and dw (tests/test_oracle_rd3.py).
Bar, as for the fans without it: arrival keys and RK4 step counts exact;
range, time and turning height within 1e-6 relative; TL and absorption
within 0.01 dB; arrival inclination and back azimuth within 1e-6 degree;
path-row keys exact with rows at the same bar.  Each case also checks that
the vertical wind changed the fan (travel times differ from the fan without
it).
"""
import os
import sys

import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_3D_RNGDEP, GA_PATH_CAUSTICS, GA_PATH_RAYS, prop_params
from parity import compare, compare_paths

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rngdep_case():
    from test_rngdep3d_gpu import workload
    return workload()


def strat_case():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_strat_topo3d_golden as mk
    prof, topo, launches = mk.scene()
    return prof, topo, launches[::2]


@pytest.fixture(scope="module")
def media(engine, oracle):
    from ga_oracle_py import Rd3Medium
    units, topo, l_rd = rngdep_case()
    rid = engine.rngdep_load_xy(*units)
    engine.topo_load(rid, *topo)
    m_rd = Rd3Medium(oracle, *units)
    m_rd.set_topo(*topo)
    m_rd.config(use_blw=True)
    prof, ttopo, l_st = strat_case()
    sid = engine.load_strat_topo(*prof, *ttopo, xy=True)
    m_st = Rd3Medium.strat_over_terrain(oracle, *prof, *ttopo)
    m_st.config(use_blw=True)
    yield {"rngdep": (rid, m_rd, l_rd), "strat": (sid, m_st, l_st)}
    engine.rngdep_release(rid)
    engine.rngdep_release(sid)


CASES = [("rngdep", 1, 0), ("rngdep", 0, 0), ("rngdep", 1, GA_PATH_RAYS | GA_PATH_CAUSTICS),
         ("strat", 1, 0), ("strat", 0, 0), ("strat", 1, GA_PATH_RAYS | GA_PATH_CAUSTICS)]


@pytest.mark.parametrize("medium,amp,flags", CASES,
                         ids=[f"{m}-{'amp18' if a else 'amp6'}{'-paths' if f else ''}" for m, a, f in CASES])
def test_fan_blw_matches_oracle(engine, media, medium, amp, flags):
    aid, m, launches = media[medium]
    bounces = 2
    ref, rpts, steps = m.fan3(launches, m.params(bool(amp)), bounces=bounces, write=bool(flags), flags=flags)
    p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=amp, bounces=bounces, topo_use_blw=1)
    if flags:
        arr, pts, st = engine.fan_paths(aid, p, launches, flags)
    else:
        arr, st = engine.fan(aid, p, launches)
    print(f"\n{medium} amp={amp} flags={flags}: {len(arr)} arrivals (oracle {len(ref)}), "
          f"{st.ray_steps} ray-steps (oracle {steps})", flush=True)
    assert len(ref) > 3 and st.ray_steps == steps
    fails, summ = compare(arr, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)
    np.testing.assert_allclose(arr["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(arr["back_az"], ref["back_az"], rtol=0, atol=1e-6)
    if flags:
        pf, ps = compare_paths(pts, rpts, False, src=(0.0, 0.0, 0.0))
        assert not pf, (pf, ps)
    # the vertical wind is on: the same fan without it has other travel times
    p0 = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=amp, bounces=bounces, topo_use_blw=0)
    arr0, _ = engine.fan(aid, p0, launches)
    n = min(len(arr0), len(arr))
    assert len(arr0) != len(arr) or np.abs(arr0["time"][:n] - arr["time"][:n]).max() > 1e-9


def test_cli_3d_rngdep_prop_blw(tmp_path, oracle):
    """infraga-3d-rngdep -prop with topo_use_BLw=true over the reference's
    node-file layout: the arrivals file equals the oracle's rows with the
    vertical wind to the printed digits (source 3 km up, above the terrain,
    so the inclinations start at incl_min; amplitudes on, the binary's
    default, so TL is compared with the oracle's 18-equation fan)."""
    import subprocess
    from ga_oracle_py import Rd3Medium, write_rd_files
    units, topo, _ = rngdep_case()
    x, y, z, c, u, v, rho = units
    pre, xs, ys, tp = write_rd_files(str(tmp_path), x, y, z, c * 1000.0, u * 1000.0, v * 1000.0, rho, topo=topo)
    exe = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin", "infraga-3d-rngdep")
    r = subprocess.run([exe, "-prop", pre, xs, ys, f"topo_file={tp}", "prof_format=zcuvd", "write_rays=false",
                        "incl_min=4", "incl_max=18", "incl_step=7", "azimuth=-90", "bounces=2",
                        "src_alt=3.0", "topo_use_BLw=true"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [[float(t) for t in l.split("\t")] for l in open(pre + ".arrivals.dat") if l.strip() and not l.startswith("#")]
    assert len(rows) > 0
    m = Rd3Medium(oracle, *units)
    m.set_topo(*topo)
    m.config(use_blw=True)
    L = sorted({(row[0], row[1]) for row in rows})
    ref, _, _ = m.fan3(L, m.params(True), bounces=2, src=(float("nan"), float("nan"), 3.0))
    assert len(ref) == len(rows)
    for row, a in zip(rows, ref):
        exp = [a["incl_launch"], a["az_launch"], a["n_bounce"], a["pos0"], a["pos1"], a["time"], a["cel"],
               a["turn_ht"], a["incl_arr"], a["back_az"], a["tl"], a["absorption"]]
        np.testing.assert_allclose(row, exp, rtol=1e-5, atol=1e-5)
