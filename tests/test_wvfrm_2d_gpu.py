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


def test_cli_2d_wnl_wvfrm(engine, oracle, pert_units, tmp_path):
    """infraga-2d -wnl_wvfrm: wvfrm_out.dat equal to the oracle's 2D waveform
    at the file's 8 digits, the writer's header; write_ray refused."""
    import os
    import subprocess
    from ga_oracle_py import Medium2D, WvParams, wv_build, wv_run_2d
    z, c, u, v, rho = pert_units
    met = os.path.join(str(tmp_path), "pert.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    units = (z, (c * 1e3) / 1000.0, (u * 1e3) / 1000.0, (v * 1e3) / 1000.0, rho)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "infraga_amd", "bin", "infraga-2d")
    args = [exe, "-wnl_wvfrm", met, "prof_format=zcuvd", "inclination=14.0", "azimuth=-90.0", "bounces=1",
            "wvfrm_len=1024", "wvfrm_p0=500.0"]
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    t, ps, _ = wv_build(oracle, "impulse", p0=500.0, t0=0.1, alpha=4.0, length=1024)
    ref = wv_run_2d(Medium2D(oracle, *units), WvParams.default(len=1024), 14.0, -90.0, 1, t, ps)
    text = open(os.path.join(str(tmp_path), "pert.wvfrm_out.dat")).read()
    assert text.startswith("# infraga-2d -wnl_wvfrm summary:\n") and "#\tsource location: 0, 0\n" in text
    out = np.array([[float(x) for x in l.split("\t")] for l in text.splitlines() if l.strip() and not l.startswith("#")])
    np.testing.assert_allclose(out[:, 0], ref["t_out"], rtol=1e-7)
    peak = np.abs(ref["p_out"]).max()
    np.testing.assert_allclose(out[:, 1], ref["p_out"], rtol=1e-7, atol=1e-7 * peak)
    r = subprocess.run(args + ["write_ray=true"], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 2
