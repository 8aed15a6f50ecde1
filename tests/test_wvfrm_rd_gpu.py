"""GPU parity of the weakly non-linear waveforms over range-dependent media
(ga_wnl_wvfrm with GA_GEOM_SPH_RNGDEP / GA_GEOM_3D_RNGDEP: the
range-dependent traces with every row kept, the absorption evaluated in the
hybrid-spline medium) against the oracle's range-dependent waveform
(go_h_rd_wnl_run, pinned by reduction to the stratified waveform over
identical columns in tests/test_oracle_wvfrm_rd.py), on the terrain grids of
the range-dependent fan tests:
- status and evolution steps exact, wvfrm_init / wvfrm_out times 1e-9
  relative, the evolved pressure within 1e-6 of the output peak, arrival
  position and time 1e-9 relative, for the spherical grid (with a ray that
  leaves the region: status 0), the Cartesian grid and a stratified profile
  over terrain (the stratified semantics, infraga-sph topo_file=).
"""
import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_3D_RNGDEP, GA_GEOM_SPH_RNGDEP, prop_params, wvfrm_params

pytestmark = pytest.mark.gpu


def _src_wave(oracle, n=1024):
    from ga_oracle_py import wv_build
    t, p, _ = wv_build(oracle, "impulse", p0=500.0, t0=0.1, alpha=4.0, length=n)
    return t, p


def _check(g, i, r):
    assert g["status"][i] == (1 if r["ok"] else 0), (g["status"][i], r["ok"])
    assert g["steps"][i] == r["steps"], (g["steps"][i], r["steps"])
    np.testing.assert_allclose(g["t_init"][i], r["t_init"], rtol=1e-9)
    np.testing.assert_allclose(g["t_out"][i], r["t_out"], rtol=1e-9)
    peak = np.abs(r["p_out"]).max()
    err = np.abs(g["p_out"][i] - r["p_out"]).max()
    assert err <= 1e-6 * peak, (err, peak)
    if r["ok"]:
        a = g["arrivals"][i]
        np.testing.assert_allclose([a["pos0"], a["pos1"], a["time"]], r["arrival"][:3], rtol=1e-9, atol=1e-9)
    return err / peak


def _run_case(engine, oracle, rid, m, geom, src, launches, n=1024):
    from ga_oracle_py import WvParams, wv_run
    t, ps = _src_wave(oracle, n)
    p = prop_params(geom, bounces=1, src=src)
    g = engine.wnl_wvfrm(rid, p, wvfrm_params(n), launches, t, ps)
    worst = 0.0
    for i, (incl, az) in enumerate(launches):
        r = wv_run(m, m.params(True), WvParams.default(len=n), src, incl, az, 1, t, ps)
        worst = max(worst, _check(g, i, r))
    st = g["stats"]
    print(f"\n{len(launches)} waveforms, status {list(g['status'])}, {st.wv_steps} evolution steps, "
          f"trace+set-up {st.trace_ms:.1f} ms, evolve {st.evolve_ms:.1f} ms, worst |dp|/peak {worst:.2e}")
    return g


def test_sph_rngdep_terrain(engine, oracle):
    from ga_oracle_py import RdMedium
    from test_rngdep_fan_gpu import workload
    _, units, topo, _ = workload()
    rid = engine.rngdep_load(*units)
    engine.topo_load(rid, *topo)
    m = RdMedium(oracle, *units)
    m.set_topo(*topo)
    la, lo = m.centre()
    try:
        g = _run_case(engine, oracle, rid, m, GA_GEOM_SPH_RNGDEP, (la, lo, 0.0),
                      [(8.0, -90.0), (12.0, 0.0), (18.0, 45.0), (12.0, 135.0), (3.0, -90.0)])
        assert list(g["status"]) == [1, 1, 1, 1, 0]
    finally:
        engine.rngdep_release(rid)


def test_cart_rngdep_terrain(engine, oracle):
    from ga_oracle_py import Rd3Medium
    from test_rngdep3d_gpu import workload
    units, topo, _ = workload()
    rid = engine.rngdep_load_xy(*units)
    engine.topo_load(rid, *topo)
    m = Rd3Medium(oracle, *units)
    m.set_topo(*topo)
    try:
        g = _run_case(engine, oracle, rid, m, GA_GEOM_3D_RNGDEP, (0.0, 0.0, 0.0), [(12.0, 45.0), (4.0, -90.0)])
        assert list(g["status"]) == [1, 0]
    finally:
        engine.rngdep_release(rid)


def test_stratified_profile_over_terrain(engine, oracle):
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_strat_topo_golden as mk
    from ga_oracle_py import RdMedium
    prof, topo, _ = mk.scene()
    rid = engine.load_strat_topo(*prof, *topo)
    m = mk.medium()
    la, lo = m.centre()
    try:
        g = _run_case(engine, oracle, rid, m, GA_GEOM_SPH_RNGDEP, (la, lo, 0.0), [(10.0, 60.0), (14.0, -100.0)])
        assert list(g["status"]) == [1, 1]
    finally:
        engine.rngdep_release(rid)


def _rows(path):
    return np.array([[float(x) for x in l.split("\t")] for l in open(path) if l.strip() and not l.startswith("#")])


def _cli(binary, args, cwd):
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "infraga_amd", "bin", binary)
    return subprocess.run([exe, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


def _abi_at_centre(engine, rid, geom, launch, t, ps, n):
    """The ABI's waveform from the region's centre on the ground, where the
    front ends put the source by default."""
    info = engine.rngdep_info(rid)
    la, lo = (info[1] + info[0]) / 2.0, (info[3] + info[2]) / 2.0
    _, tp = engine.rngdep_eval(rid, [(la, lo, 0.0)])
    zg = max(0.0, float(tp[0, 0]))
    deg = 1.0 if geom == GA_GEOM_3D_RNGDEP else 180.0 / np.pi
    p = prop_params(geom, bounces=1, src=(la * deg, lo * deg, zg))
    return engine.wnl_wvfrm(rid, p, wvfrm_params(n), [launch], t, ps)


def test_cli_sph_rngdep_wnl_wvfrm(engine, oracle, tmp_path):
    """infraga-sph-rngdep -wnl_wvfrm on the reference's node-file layout with
    terrain: source at the region's centre, wvfrm_out.dat equal to the ABI's
    waveform for that source at the file's 8 digits, the writer's header
    (including its extra "ground evevation" line)."""
    import os
    from ga_oracle_py import write_rd_files
    from infraga_amd import synth
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    topo = synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81), amp_km=2.0)
    lat, lon, z, T3, u3, v3, rho3 = grid
    pre, lats, lons, tp = write_rd_files(str(tmp_path), lat, lon, z, synth.sound_speed(T3), u3, v3, rho3, topo=topo)
    r = _cli("infraga-sph-rngdep", ["-wnl_wvfrm", pre, lats, lons, "topo_file=" + tp, "prof_format=zcuvd",
                                    "inclination=12.0", "azimuth=0.0", "bounces=1", "wvfrm_len=1024",
                                    "wvfrm_p0=500.0"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    t, ps = _src_wave(oracle)
    rid = engine.rngdep_load_files(pre, lats, lons, tp, fmt="zcuvd")
    try:
        g = _abi_at_centre(engine, rid, GA_GEOM_SPH_RNGDEP, (12.0, 0.0), t, ps, 1024)
    finally:
        engine.rngdep_release(rid)
    assert g["status"][0] == 1
    text = open(pre + ".wvfrm_out.dat").read()
    assert text.startswith("# infraga-sph-rngdep -wnl_wvfrm summary:\n")
    assert "#\ttopo file:" + tp + "\n" in text and "#\tground evevation: " in text
    out = _rows(pre + ".wvfrm_out.dat")
    np.testing.assert_allclose(out[:, 0], g["t_out"][0], rtol=1e-7)
    peak = np.abs(g["p_out"][0]).max()
    np.testing.assert_allclose(out[:, 1], g["p_out"][0], rtol=1e-7, atol=1e-7 * peak)
    assert os.path.exists(pre + ".wvfrm_init.dat")


def test_cli_sph_topo_file_wnl_wvfrm(engine, oracle, tmp_path):
    """infraga-sph -wnl_wvfrm with topo_file=: the stratified profile over
    terrain, source at the terrain's centre, equal to the ABI's waveform."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_strat_topo_golden as mk
    prof, topo, _ = mk.scene()
    z, c, u, v, rho = prof
    met = os.path.join(str(tmp_path), "strat.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    tl, tn, zg = topo
    tp = os.path.join(str(tmp_path), "topo.dat")
    with open(tp, "w") as f:
        for i, a in enumerate(tl):
            f.write("".join("%.17g %.17g %.17g\n" % (a, b, zg[i, j]) for j, b in enumerate(tn)))
    r = _cli("infraga-sph", ["-wnl_wvfrm", met, "topo_file=" + tp, "prof_format=zcuvd", "inclination=10.0",
                             "azimuth=60.0", "bounces=1", "wvfrm_len=1024", "wvfrm_p0=500.0"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    t, ps = _src_wave(oracle)
    rid = engine.load_file_topo(met, tp, fmt="zcuvd")
    try:
        g = _abi_at_centre(engine, rid, GA_GEOM_SPH_RNGDEP, (10.0, 60.0), t, ps, 1024)
    finally:
        engine.rngdep_release(rid)
    assert g["status"][0] == 1
    text = open(os.path.join(str(tmp_path), "strat.wvfrm_out.dat")).read()
    assert "#\ttopo file:" + tp + "\n" in text
    out = _rows(os.path.join(str(tmp_path), "strat.wvfrm_out.dat"))
    peak = np.abs(g["p_out"][0]).max()
    np.testing.assert_allclose(out[:, 1], g["p_out"][0], rtol=1e-7, atol=1e-7 * peak)
