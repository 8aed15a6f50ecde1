"""GPU parity of the 2D weakly non-linear waveform (ga_wnl_wvfrm with
GA_GEOM_2D: the 2D trace with every row kept and every leg counted, the 2D
beta and absorption forms) against the oracle's 2D waveform (go2_wnl_run,
pinned by equality with the 3D waveform in a homogeneous medium,
tests/test_oracle_wvfrm_2d.py), on the seeded perturbed profile of the 2D fan
tests: along, against and across the jet; an elevated reflection altitude
(whose legs count toward bounces); a terrain profile.  Bar: status and
evolution steps exact, times 1e-9 relative, the evolved pressure within 1e-6
of the output peak, arrival range and time 1e-9 relative.
"""
import numpy as np
import pytest

from infraga_amd.engine import GA_GEOM_2D, prop_params, wvfrm_params

pytestmark = pytest.mark.gpu

CASES = [("flat", None, False), ("refl60", 60.0, False), ("hill", None, True)]


@pytest.mark.parametrize("name,refl,topo", CASES, ids=[c[0] for c in CASES])
def test_2d_waveforms_match_oracle(engine, oracle, pert_units, name, refl, topo):
    from ga_oracle_py import Medium2D, WvParams, wv_build, wv_run_2d
    from test_2d_gpu import hill
    t, ps, _ = wv_build(oracle, "impulse", p0=500.0, t0=0.1, alpha=4.0, length=1024)
    L = [(8.0, 90.0), (14.0, -90.0), (20.0, 30.0), (30.0, 90.0)]
    aid = engine.load_atmo(*pert_units)
    kw = {} if refl is None else {"refl_alt": refl}
    if topo:
        engine.topo_load_2d(aid, *hill())
    try:
        g = engine.wnl_wvfrm(aid, prop_params(GA_GEOM_2D, bounces=1, **kw), wvfrm_params(1024), L, t, ps)
    finally:
        engine.release_atmo(aid)
    m = Medium2D(oracle, *pert_units, refl_alt=refl if refl is not None else float("nan"),
                 topo=hill() if topo else None)
    worst = 0.0
    for i, (incl, az) in enumerate(L):
        r = wv_run_2d(m, WvParams.default(len=1024), incl, az, 1, t, ps)
        assert g["status"][i] == (1 if r["ok"] else 0)
        assert g["steps"][i] == r["steps"], (i, g["steps"][i], r["steps"])
        np.testing.assert_allclose(g["t_init"][i], r["t_init"], rtol=1e-9)
        np.testing.assert_allclose(g["t_out"][i], r["t_out"], rtol=1e-9)
        peak = np.abs(r["p_out"]).max()
        err = np.abs(g["p_out"][i] - r["p_out"]).max()
        assert err <= 1e-6 * peak, (i, err, peak)
        worst = max(worst, err / peak)
        a = g["arrivals"][i]
        np.testing.assert_allclose([a["pos0"], a["time"]], r["arrival"][[0, 2]], rtol=1e-9)
    st = g["stats"]
    print(f"\n{name}: status {list(g['status'])}, {st.wv_steps} evolution steps, worst |dp|/peak {worst:.2e}")
    assert (g["status"] == 1).all()
