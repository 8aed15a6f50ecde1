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
