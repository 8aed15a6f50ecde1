"""GPU parity of the weakly non-linear waveforms (ga_wnl_wvfrm: the traces
with every row kept, the reference point and Burgers coefficients from those
rows, wv_evolve_kernel's LDS transforms) against the oracle's run_wnl_wvfrm
restatement (oracle/ga_wvfrm.c, pinned in tests/test_oracle_wvfrm.py by
numpy's transforms, linear-propagation closed forms and the Earnshaw
solution) on the same launches, sources and settings:
- status (returned / left the region / no reference point) and the number of
  evolution steps exact;
- wvfrm_init times and wvfrm_out times within 1e-9 relative, the source
  pressure exact, the evolved pressure within 1e-6 of the output peak;
- the arrival summary: position and time 1e-9 relative, turning height,
  inclination and back azimuth 1e-6, TL and absorption 0.01 dB;
for both stratified geometries, nonlinear impulses that steepen, an N-wave,
lengths 1024 to 8192, two and three legs.  A batch of
launches returns exactly what the launches return one at a time.
"""
import math

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, EngineError, prop_params, wvfrm_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def base(engine, base_units):
    aid = engine.load_atmo(*base_units)
    yield aid
    engine.release_atmo(aid)


def oracle_run(oracle, units, geom, src, incl, az, bounces, t, ps, length, ref=1.0, **kw):
    from ga_oracle_py import Medium, WvParams, wv_run
    m = Medium(oracle, geom, *units)
    p = m.params(True, **kw)
    return wv_run(m, p, WvParams.default(len=length), src, incl, az, bounces, t, ps, freq=0.1, wvfrm_ref=ref)


def source(oracle, option, length, p0=500.0, t0=0.1):
    from ga_oracle_py import wv_build
    t, ps, _ = wv_build(oracle, option, p0=p0, t0=t0, alpha=4.0, length=length)
    return t, ps


def check_one(g, i, r, sph):
    assert g["status"][i] == (1 if r["ok"] else 0)
    assert g["steps"][i] == r["steps"], (g["steps"][i], r["steps"])
    np.testing.assert_allclose(g["t_init"][i], r["t_init"], rtol=1e-9)
    assert np.array_equal(g["p_init"][i], r["p_init"])
    np.testing.assert_allclose(g["t_out"][i], r["t_out"], rtol=1e-9)
    peak = np.abs(r["p_out"]).max()
    err = np.abs(g["p_out"][i] - r["p_out"]).max()
    assert err <= 1e-6 * peak, (err, peak)
    if r["ok"]:
        a, o = g["arrivals"][i], r["arrival"]
        np.testing.assert_allclose([a["pos0"], a["pos1"], a["time"], a["cel"]], o[:4], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose([a["turn_ht"], a["incl_arr"], a["back_az"]], o[4:7], rtol=0, atol=1e-6)
        np.testing.assert_allclose([a["tl"], a["absorption"]], o[7:9], rtol=0, atol=0.01)
    return err / peak


CASES = [
    # name, geometry, source, launches, bounces, length, waveform
    ("sph-impulse", GA_GEOM_SPH, (30.0, 0.0, 0.0), [(6.0, 90.0), (12.0, 90.0), (25.0, 90.0), (35.0, 90.0)], 1,
     1024, "impulse"),
    ("3d-impulse", GA_GEOM_3D, (0.0, 0.0, 0.0), [(6.0, 90.0), (18.0, 90.0), (25.0, 90.0), (35.0, 60.0)], 1,
     1024, "impulse"),
    ("sph-nwave-8192", GA_GEOM_SPH, (30.0, 0.0, 0.0), [(12.0, 45.0)], 1, 8192, "Nwave"),
    ("3d-3legs", GA_GEOM_3D, (0.0, 0.0, 0.0), [(6.0, 60.0)], 2, 2048, "Uwave"),
]


@pytest.mark.parametrize("name,geom,src,launches,bounces,length,wv", CASES, ids=[c[0] for c in CASES])
def test_waveforms_match_oracle(engine, oracle, base, base_units, name, geom, src, launches, bounces, length, wv):
    t, ps = source(oracle, wv, length)
    p = prop_params(geom, bounces=bounces, src=src)
    g = engine.wnl_wvfrm(base, p, wvfrm_params(length), launches, t, ps)
    st = g["stats"]
    worst = 0.0
    for i, (incl, az) in enumerate(launches):
        r = oracle_run(oracle, base_units, geom, src, incl, az, bounces, t, ps, length)
        worst = max(worst, check_one(g, i, r, geom == GA_GEOM_SPH))
    print(f"\n{name}: {len(launches)} waveforms, {st.wv_steps} evolution steps, {st.ray_steps} ray-steps, "
          f"trace {st.trace_ms:.1f} ms, evolve {st.evolve_ms:.1f} ms, worst |dp|/peak {worst:.2e}")
    assert (g["status"] == 1).all()


def test_batch_equals_single(engine, oracle, base):
    t, ps = source(oracle, "impulse", 1024)
    p = prop_params(GA_GEOM_SPH, bounces=1, src=(30.0, 0.0, 0.0))
    L = [(8.0, 90.0), (16.0, 0.0), (30.0, -120.0)]
    b = engine.wnl_wvfrm(base, p, wvfrm_params(1024), L, t, ps)
    for i, l in enumerate(L):
        s = engine.wnl_wvfrm(base, p, wvfrm_params(1024), [l], t, ps)
        for k in ("t_init", "p_init", "t_out", "p_out"):
            assert np.array_equal(b[k][i], s[k][0]), k
        assert b["steps"][i] == s["steps"][0] and b["status"][i] == s["status"][0]


def test_status_codes(engine, oracle, base, base_units):
    t, ps = source(oracle, "impulse", 1024)
    # the ray leaves the region (range limit) before its ground return:
    # evolved over the traced leg, status 0, as run_wnl_wvfrm reports it
    p = prop_params(GA_GEOM_SPH, bounces=1, src=(30.0, 0.0, 0.0), rng_max=150.0)
    g = engine.wnl_wvfrm(base, p, wvfrm_params(1024), [(12.0, 90.0)], t, ps)
    r = oracle_run(oracle, base_units, GA_GEOM_SPH, (30.0, 0.0, 0.0), 12.0, 90.0, 1, t, ps, 1024, rng_max=150.0)
    assert not r["ok"] and g["status"][0] == 0
    check_one(g, 0, r, True)
    # the reference point beyond the whole ray: no waveform
    p = prop_params(GA_GEOM_SPH, bounces=0, src=(30.0, 0.0, 0.0))
    g = engine.wnl_wvfrm(base, p, wvfrm_params(1024, ref=1.0e5), [(12.0, 90.0)], t, ps)
    assert g["status"][0] == -1 and np.isnan(g["p_out"][0]).all()


def test_argument_errors(engine, oracle, base):
    t, ps = source(oracle, "impulse", 1024)
    p = prop_params(GA_GEOM_SPH, bounces=0, src=(30.0, 0.0, 0.0))
    with pytest.raises(EngineError):
        engine.wnl_wvfrm(base, p, wvfrm_params(1000), [(12.0, 90.0)], t[:1000], ps[:1000])
    with pytest.raises(EngineError):
        engine.wnl_wvfrm(base, p, wvfrm_params(1024, ds=0.0), [(12.0, 90.0)], t, ps)
    with pytest.raises(EngineError):
        engine.wnl_wvfrm(base, prop_params(GA_GEOM_SPH_RNGDEP), wvfrm_params(1024), [(12.0, 90.0)], t, ps)


def _write_met(d, units):
    import os
    z, c, u, v, rho = units
    met = os.path.join(d, "base.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    # the loader's values: km/s from the file's m/s
    return met, (z, (c * 1e3) / 1000.0, (u * 1e3) / 1000.0, (v * 1e3) / 1000.0, rho)


def _rows(path):
    return np.array([[float(x) for x in l.split("\t")] for l in open(path) if l.strip() and not l.startswith("#")])


@pytest.mark.parametrize("binary,geom,src,extra", [
    ("infraga-sph", GA_GEOM_SPH, (30.0, 0.0, 0.0), []),
    ("infraga-3d", GA_GEOM_3D, (0.0, 0.0, 0.0), ["src_x=0.0", "src_y=0.0"]),
])
def test_cli_wnl_wvfrm(oracle, base_units, tmp_path, binary, geom, src, extra):
    """infraga-sph / infraga-3d -wnl_wvfrm: the built-in impulse of
    wvfrm_init.dat equals the oracle's waveforms.cpp restatement, and
    wvfrm_out.dat the oracle's run_wnl_wvfrm, at the files' 8 significant
    digits; header lines as the reference writes them; write_ray=true gives
    the ray's raypaths.dat."""
    import os
    import subprocess
    met, units = _write_met(str(tmp_path), base_units)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "infraga_amd", "bin", binary)
    args = [exe, "-wnl_wvfrm", met, "prof_format=zcuvd", "inclination=12.0", "azimuth=90.0", "bounces=1",
            "wvfrm_len=1024", "wvfrm_p0=500.0", "write_ray=true", *extra]
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Arrival summary:" in r.stdout
    t, ps = source(oracle, "impulse", 1024)
    ref = oracle_run(oracle, units, geom, src, 12.0, 90.0, 1, t, ps, 1024)
    init = _rows(os.path.join(str(tmp_path), "base.wvfrm_init.dat"))
    np.testing.assert_allclose(init[:, 0], ref["t_init"], rtol=1e-7, atol=1e-6)
    np.testing.assert_allclose(init[:, 1], ref["p_init"], rtol=1e-7, atol=1e-12)
    out_text = open(os.path.join(str(tmp_path), "base.wvfrm_out.dat")).read()
    assert out_text.startswith(f"# {binary} -wnl_wvfrm summary:\n")
    assert "#\tazimuth: 0\n" in out_text             # geoac::phi in degrees, as the reference prints it
    assert "#\twaveform option: impulse\n" in out_text and "#\twaveform p0: 500\n" in out_text
    out = _rows(os.path.join(str(tmp_path), "base.wvfrm_out.dat"))
    np.testing.assert_allclose(out[:, 0], ref["t_out"], rtol=1e-7)
    peak = np.abs(ref["p_out"]).max()
    np.testing.assert_allclose(out[:, 1], ref["p_out"], rtol=1e-7, atol=1e-7 * peak)
    rp = _rows(os.path.join(str(tmp_path), "base.raypaths.dat"))
    assert len(rp) > 10 and rp.shape[1] == 6


def test_cli_wnl_wvfrm_nwave_from_file(oracle, base_units, tmp_path):
    """wvfrm_file= (two columns, '#' comments) through the -w alias: the
    file's samples are the source waveform."""
    import os
    import subprocess
    met, units = _write_met(str(tmp_path), base_units)
    t, ps = source(oracle, "Nwave", 2048, p0=200.0, t0=0.5)
    wf = os.path.join(str(tmp_path), "src.dat")
    with open(wf, "w") as f:
        f.write("# t p\n")
        for a, b in zip(t, ps):
            f.write("%.17g %.17g\n" % (a, b))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "infraga_amd", "bin",
                       "infraga-sph")
    r = subprocess.run([exe, "-w", met, "prof_format=zcuvd", "inclination=8.0", "azimuth=45.0", "bounces=1",
                        "wvfrm_file=" + wf, "output_id=nw"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    ref = oracle_run(oracle, units, GA_GEOM_SPH, (30.0, 0.0, 0.0), 8.0, 45.0, 1, t, ps, 2048)
    out = _rows(os.path.join(str(tmp_path), "nw.wvfrm_out.dat"))
    peak = np.abs(ref["p_out"]).max()
    np.testing.assert_allclose(out[:, 1], ref["p_out"], rtol=1e-7, atol=1e-7 * peak)
    assert "#\twaveform source file: " + wf + "\n" in open(os.path.join(str(tmp_path), "nw.wvfrm_out.dat")).read()


@pytest.mark.gpu
def test_out_step_dumps_match_oracle(engine, oracle, base, base_units, tmp_path):
    """wvfrm_out_step snapshots (ga_wnl_wvfrm_dump) against the oracle's on
    the same waveform: the same labels in the same order, pressure and
    spectrum within 1e-6 of their peaks; the run's outputs equal those of
    ga_wnl_wvfrm.  Then infraga-sph -wnl_wvfrm wvfrm_out_step= writes one
    sc_waveform / sc_spectrum file per label with those values."""
    import os
    import subprocess
    from ga_oracle_py import GEOM_SPH, Medium, WvParams, wv_run
    length, out_step = 1024, 50.0
    t, ps = source(oracle, "impulse", length)
    p = prop_params(GA_GEOM_SPH, bounces=1, src=(30.0, 0.0, 0.0))
    plain = engine.wnl_wvfrm(base, p, wvfrm_params(length), [(8.0, 90.0)], t, ps)
    g = engine.wnl_wvfrm(base, p, wvfrm_params(length), [(8.0, 90.0)], t, ps, out_step=out_step, dump_cap=200)
    for k in ("t_out", "p_out", "steps"):
        assert np.array_equal(g[k], plain[k]), k
    m = Medium(oracle, GEOM_SPH, *base_units)
    r = wv_run(m, m.params(True), WvParams.default(len=length), (30.0, 0.0, 0.0), 8.0, 90.0, 1, t, ps,
               freq=0.1, wvfrm_ref=1.0, dump=(out_step, 200))
    print(f"\n{r['dump_count']} dumps, labels {[d[0] for d in r['dumps']][:6]}", flush=True)
    assert g["dump_count"][0] == r["dump_count"] > 5
    for (lg, ug, Ug), (lo, uo, Uo) in zip(g["dumps"][0], r["dumps"]):
        assert lg == lo
        assert np.abs(ug - uo).max() <= 1e-6 * np.abs(uo).max()
        assert np.abs(Ug - Uo).max() <= 1e-6 * np.abs(Uo).max()
    met, _ = _write_met(str(tmp_path), base_units)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "infraga_amd", "bin", "infraga-sph")
    cli = subprocess.run([exe, "-wnl_wvfrm", met, "prof_format=zcuvd", "inclination=8.0", "azimuth=90.0", "bounces=1",
                          "src_lat=30.0", "src_lon=0.0", f"wvfrm_len={length}", "wvfrm_p0=500.0",
                          f"wvfrm_out_step={out_step:g}"],
                         capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert cli.returncode == 0, cli.stderr
    labels = sorted({lab for lab, _, _ in g["dumps"][0]})
    assert len(list(tmp_path.glob("sc_waveform-*.dat"))) == len(labels)
    assert len(list(tmp_path.glob("sc_spectrum-*.dat"))) == len(labels)
    lab, ug, Ug = g["dumps"][0][-1]
    w = np.loadtxt(tmp_path / ("sc_waveform-%04g.dat" % lab))
    sp = np.loadtxt(tmp_path / ("sc_spectrum-%04g.dat" % lab))
    np.testing.assert_allclose(w[:, 0], t, rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(w[:, 1], ug, rtol=0, atol=1e-5 * np.abs(ug).max())
    np.testing.assert_allclose(sp[:, 1] + 1j * sp[:, 2], Ug, rtol=0, atol=1e-5 * np.abs(Ug).max())
