"""GPU parity of the 3D Cartesian range-dependent geometry
(GA_GEOM_3D_RNGDEP, infraga-3d-rngdep -prop) against the oracle's
go_rd3_fan, whose outputs for this workload are the committed fixture
tests/golden/rd3.npz (make_rd3_golden.py).

Workload: a 6 x 6 node grid over +-500 km with horizontally varying c, u,
v, rho (seeded), terrain 0-2 km on an 81 x 81 grid, source at the region
centre, 18 and 6 equations, write_rays=false sums and the write-mode
bookkeeping with ray-path and caustic rows.  Bar: arrival keys and RK4
step counts exact, positions, time, turning height within 1e-6 relative,
TL and absorption within 0.01 dB; path-row keys exact, values at the same
bar (compare_paths).
"""
import os

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D_RNGDEP, GA_PATH_CAUSTICS, GA_PATH_RAYS, prop_params
from parity import compare, compare_paths

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "rd3.npz")


def workload():
    """(units, topo, launches): node x, y [km], z, c, u, v, rho in engine
    units; terrain (x, y, height); (inclination, azimuth) launches."""
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    _, _, z, c, u, v, rho = synth.rngdep_engine_units(*grid)
    x = np.linspace(-500.0, 500.0, 6)
    units = (x, x.copy(), z, c, u, v, rho)
    t = synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81), amp_km=2.0)
    tx = np.linspace(-520.0, 520.0, 81)
    topo = (tx, tx.copy(), t[2])
    launches = np.array([[th, az] for az in (-90.0, 45.0, 160.0) for th in (4.0, 12.0, 19.0)])
    return units, topo, launches


CASES = [("amp18", 1, 0), ("amp6", 0, 0), ("paths18", 1, GA_PATH_RAYS | GA_PATH_CAUSTICS)]


@pytest.fixture(scope="module")
def rd3(engine):
    units, topo, launches = workload()
    rid = engine.rngdep_load_xy(*units)
    engine.topo_load(rid, *topo)
    yield rid, launches
    engine.rngdep_release(rid)


@pytest.mark.gpu
@pytest.mark.parametrize("case,amp,flags", CASES, ids=[c[0] for c in CASES])
def test_fan_matches_oracle(engine, rd3, case, amp, flags):
    rid, launches = rd3
    g = np.load(GOLDEN)
    ref, steps = g[case], int(g[case + "_steps"])
    p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=amp, bounces=2)
    if flags:
        arr, pts, st = engine.fan_paths(rid, p, launches, flags)
    else:
        arr, st = engine.fan(rid, p, launches)
    print(f"\n{case}: {len(arr)} arrivals (oracle {len(ref)}), {st.ray_steps} ray-steps (oracle {steps}), "
          f"{st.kernel_ms:.1f} ms")
    assert len(ref) > 0 and st.ray_steps == steps
    fails, summ = compare(arr, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)
    np.testing.assert_allclose(arr["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(arr["back_az"], ref["back_az"], rtol=0, atol=1e-6)
    if flags:
        rp = g[case + "_pts"]
        pf, ps = compare_paths(pts, rp, False, src=(0.0, 0.0, 0.0))
        print(ps)
        assert not pf, (pf, ps)


@pytest.mark.gpu
def test_geometry_and_grid_must_agree(engine, rd3):
    """A Cartesian grid with the spherical geometry is an error."""
    from infraga_amd.engine import GA_GEOM_SPH_RNGDEP, EngineError
    rid, launches = rd3
    with pytest.raises(EngineError):
        engine.fan(rid, prop_params(GA_GEOM_SPH_RNGDEP), launches[:1])


@pytest.mark.gpu
def test_cli_3d_rngdep_prop(tmp_path, oracle):
    """infraga-3d-rngdep -prop over the reference's node-file layout: the
    arrivals file equals the oracle's rows at the same launch angles to the
    printed digits (source 3 km up, above the terrain, so the inclinations
    start at incl_min).  calc_amp= is given, which the reference's 3D binary
    reads as false (TL column 0)."""
    import subprocess
    from ga_oracle_py import Rd3Medium, write_rd_files
    units, topo, _ = workload()
    x, y, z, c, u, v, rho = units
    pre, xs, ys, tp = write_rd_files(str(tmp_path), x, y, z, c * 1000.0, u * 1000.0, v * 1000.0, rho, topo=topo)
    exe = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin", "infraga-3d-rngdep")
    r = subprocess.run([exe, "-prop", pre, xs, ys, f"topo_file={tp}", "prof_format=zcuvd", "write_rays=false",
                        "incl_min=6", "incl_max=20", "incl_step=7", "azimuth=45", "bounces=2", "calc_amp=true",
                        "src_alt=3.0"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [[float(t) for t in l.split("\t")] for l in open(pre + ".arrivals.dat") if l.strip() and not l.startswith("#")]
    assert len(rows) > 0
    m = Rd3Medium(oracle, *units)
    m.set_topo(*topo)
    L = sorted({(row[0], row[1]) for row in rows})
    ref, _, _ = m.fan3(L, m.params(False), bounces=2, src=(float("nan"), float("nan"), 3.0))
    assert len(ref) == len(rows)
    for row, a in zip(rows, ref):
        exp = [a["incl_launch"], a["az_launch"], a["n_bounce"], a["pos0"], a["pos1"], a["time"], a["cel"],
               a["turn_ht"], a["incl_arr"], a["back_az"], 0.0, a["absorption"]]
        np.testing.assert_allclose(row, exp, rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("z_bottom", [0.0, 0.5])
def test_ground_at_zero_with_z_grnd(engine, oracle, z_bottom):
    """Without terrain the 3D range-dependent ground is z = 0 whatever
    z_grnd is (atmo_state.3d.rngdep.cpp:45-52); z_grnd = 1.0 moves only the
    shadow-zone test and the boundary-layer top.  Linear c from z_bottom
    (held below the profile), GPU against the oracle at the parity bar."""
    from ga_oracle_py import Rd3Medium
    from test_oracle_rd3 import linear_columns
    cols, _, _ = linear_columns(z_bottom=z_bottom)
    L = np.array([[th, az] for az in (90.0, 200.0) for th in (8.0, 15.0, 25.0)])
    rid = engine.rngdep_load_xy(*cols)
    try:
        for amp in (1, 0):
            p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=amp, bounces=2, z_grnd=1.0, src=(0.0, 0.0, 0.0))
            arr, st = engine.fan(rid, p, L)
            m = Rd3Medium(oracle, *cols)
            m.set_ground(1.0)
            ref, _, steps = m.fan3(L, m.params(bool(amp)), bounces=2, src=(0.0, 0.0, 0.0))
            assert len(ref) == 18 and st.ray_steps == steps
            fails, summ = compare(arr, ref, False, src=(0.0, 0.0))
            assert not fails, (fails, summ)
    finally:
        engine.rngdep_release(rid)


@pytest.mark.gpu
def test_no_travel_time_limit(engine):
    """The 3D range-dependent geometry ignores time_max (main.infraga.3d.rngdep.cpp:494-500)."""
    from test_oracle_rd3 import linear_columns
    cols, _, _ = linear_columns(z_bottom=0.0)
    L = np.array([[th, 90.0] for th in (8.0, 15.0)])
    rid = engine.rngdep_load_xy(*cols)
    try:
        a1, s1 = engine.fan(rid, prop_params(GA_GEOM_3D_RNGDEP, calc_amp=0, bounces=2, src=(0.0, 0.0, 0.0)), L)
        a2, s2 = engine.fan(rid, prop_params(GA_GEOM_3D_RNGDEP, calc_amp=0, bounces=2, src=(0.0, 0.0, 0.0),
                                             time_max=1.0e-3), L)
        assert len(a1) == len(a2) == 6 and s1.ray_steps == s2.ray_steps and (a2["time"] > 3.6).all()
    finally:
        engine.rngdep_release(rid)


@pytest.mark.gpu
def test_cli_3d_rngdep_caustics_force_amplitudes(tmp_path):
    """write_caustics=true turns amplitudes on (main.infraga.3d.rngdep.cpp:310)
    even though the 3D binary reads calc_amp=... as false; caustic rows are
    x, y, z, 0.0, time (:467-471) and the arrivals header is the 3D
    range-dependent one (:346-368)."""
    import subprocess
    from ga_oracle_py import write_rd_files
    units, topo, _ = workload()
    x, y, z, c, u, v, rho = units
    pre, xs, ys, tp = write_rd_files(str(tmp_path), x, y, z, c * 1000.0, u * 1000.0, v * 1000.0, rho, topo=topo)
    exe = os.path.join(os.path.dirname(HERE), "infraga_amd", "bin", "infraga-3d-rngdep")
    r = subprocess.run([exe, "-prop", pre, xs, ys, f"topo_file={tp}", "prof_format=zcuvd", "write_rays=false",
                        "write_caustics=true", "calc_amp=false", "incl_min=4", "incl_max=30", "incl_step=2",
                        "azimuth=45", "bounces=1", "src_alt=3.0"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    head = [l for l in open(pre + ".arrivals.dat") if l.startswith("#")]
    assert any("profile prefix: " in l for l in head) and any("profile grid files: " in l for l in head)
    assert any("arrival incl [deg]" in l and "back az [deg]" in l for l in head)
    rows = [[float(t) for t in l.split("\t")] for l in open(pre + ".arrivals.dat") if l.strip() and not l.startswith("#")]
    assert len(rows) > 0 and any(row[10] != 0.0 for row in rows)     # TL computed
    caus = [l.split("\t") for l in open(pre + ".caustics-0.dat") if l.strip() and not l.startswith("#")]
    assert len(caus) > 0 and all(len(cr) == 5 and float(cr[3]) == 0.0 for cr in caus)
