"""GPU parity of the write_rays / write_caustics output (ga_fan_run_paths,
ga_eig_search_paths) and the front ends' path files.

- Against the reference's own probe of its 90-ray ToyAtmo fan (azimuth -90,
  inclinations 0.5..45 step 0.5, 2 bounces, 800 km): 1,399,246 ray-steps
  and 93,397 raypaths.dat rows, 233 arrivals.
- Against the oracle's write-mode fan on the same launch doubles (3D and
  spherical, 18 and 6 equations, with caustics): arrival keys exact and the
  arrival bar (range, time, turning height 1e-6 relative; TL, absorption
  0.01 dB); path-row keys (ray, leg, step, kind) exact; row position within
  1e-6 of the range from the source, altitude within 1e-6 km per km of
  range, time 1e-6 relative, TL and absorption 0.01 dB.
- Range-dependent fan over terrain against the oracle fixture
  tests/golden/paths_rngdep.npz (make_paths_golden.py).
- Eigenray path rows against the oracle's go_eig_path at the GPU's launch
  angles (ToyAtmo, the reference's three eigenrays).
"""
import os
import subprocess

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import (GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, GA_PATH_CAUSTICS, GA_PATH_RAYS,
                                EngineError, eig_params, prop_params)
from parity import compare, compare_paths

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def reference_fan90():
    L, th = [], 0.5
    while th <= 45.0:
        L.append((th, -90.0))
        th += 0.5
    return np.array(L)


def test_toyatmo_fan_matches_reference_probe(engine, toyatmo_units):
    aid = engine.load_atmo(*toyatmo_units)
    p = prop_params(GA_GEOM_SPH, bounces=2, src=(30.0, -100.0, 0.0), rng_max=800.0)
    arr, pts, st = engine.fan_paths(aid, p, reference_fan90(), GA_PATH_RAYS)
    print(f"\n90-ray ToyAtmo fan: {st.ray_steps} ray-steps, {len(pts)} path rows, {len(arr)} arrivals, "
          f"{st.kernel_ms:.1f} ms")
    assert st.ray_steps == 1399246
    assert len(pts) == 93397
    assert len(arr) == 233
    engine.release_atmo(aid)


CASES = [
    # geometry, calc_amp, incl, az, bounces, src, rng_max
    ("sph18", GA_GEOM_SPH, True, (1.0, 40.0, 3.0), (-90.0, 45.0, 135.0), 2, (30.0, -100.0, 0.0), 800.0),
    ("sph6", GA_GEOM_SPH, False, (1.0, 41.0, 5.0), (10.0, 190.0, 180.0), 2, (30.0, 0.0, 0.0), 1000.0),
    ("3d18", GA_GEOM_3D, True, (1.0, 40.0, 3.0), (0.0, 300.0, 150.0), 2, (0.0, 0.0, 0.0), 800.0),
    ("3d6", GA_GEOM_3D, False, (1.0, 41.0, 5.0), (90.0, 270.0, 180.0), 2, (0.0, 0.0, 0.0), 800.0),
]


@pytest.mark.parametrize("name,geom,calc_amp,incl,az,bounces,src,rng", CASES, ids=[c[0] for c in CASES])
def test_fan_paths_parity(engine, oracle, pert_units, name, geom, calc_amp, incl, az, bounces, src, rng):
    from ga_oracle_py import PATH_CAUSTICS, PATH_RAYS, Medium
    L = synth.fan_launches(incl[0], incl[1], incl[2], az[0], az[1], az[2])
    aid = engine.load_atmo(*pert_units)
    flags = GA_PATH_RAYS | (GA_PATH_CAUSTICS if calc_amp else 0)
    p = prop_params(geom, calc_amp=int(calc_amp), bounces=bounces, src=src, rng_max=rng)
    gpu, gpts, st = engine.fan_paths(aid, p, L, flags)
    m = Medium(oracle, geom, *pert_units)
    op = m.params(calc_amp)
    op.rng_max = rng
    ref, rpts, steps = m.fan_paths(L, op, PATH_RAYS | (PATH_CAUSTICS if calc_amp else 0), bounces=bounces, src=src)
    print(f"\n{name}: rays={len(L)} arrivals gpu={len(gpu)} oracle={len(ref)} rows gpu={len(gpts)} "
          f"oracle={len(rpts)} caustics={int((rpts['kind'] == 1).sum())} steps gpu={st.ray_steps} oracle={steps}")
    fails, summ = compare(gpu, ref, geom == GA_GEOM_SPH, src=src[:2])
    assert not fails, (fails, summ)
    pf, ps = compare_paths(gpts, rpts, geom == GA_GEOM_SPH, src=src)
    print(ps)
    assert not pf, (pf, ps)
    assert len(rpts) > 0 and (not calc_amp or (rpts["kind"] == 1).any())
    engine.release_atmo(aid)


def test_paths_leave_write_false_arrivals_alone(engine, pert_units):
    """The write-mode arrivals differ from write_rays=false ones only by the
    segment into the sub-ground row of each leg: same keys and positions,
    earlier times by at most a few tenths of a second per leg."""
    L = synth.fan_launches(2.0, 30.0, 4.0, 0.0, 180.0, 90.0)
    aid = engine.load_atmo(*pert_units)
    p = prop_params(GA_GEOM_SPH, bounces=2, src=(30.0, 0.0, 0.0))
    a0, _ = engine.fan(aid, p, L)
    a1, pts, _ = engine.fan_paths(aid, p, L, GA_PATH_RAYS)
    assert len(a0) == len(a1) > 0
    assert (a0["launch_idx"] == a1["launch_idx"]).all() and (a0["n_bounce"] == a1["n_bounce"]).all()
    assert (a0["pos0"] == a1["pos0"]).all() and (a0["pos1"] == a1["pos1"]).all()
    dt = a0["time"] - a1["time"]
    assert (dt > 0).all() and (dt < 0.2 * (a0["n_bounce"] + 1)).all()
    engine.release_atmo(aid)


def test_path_capacity_errors(engine, pert_units):
    """Too small a capacity reports the count needed (GA_E_CAP) and a retry
    with that capacity returns the same rows."""
    import ctypes as C
    from infraga_amd.engine import GA_E_CAP, PATH_DTYPE, ARRIVAL_DTYPE, Arrival, PathPt, RunStats
    L = np.ascontiguousarray(synth.fan_launches(5.0, 25.0, 10.0, 0.0, 0.0, 1.0))
    aid = engine.load_atmo(*pert_units)
    p = prop_params(GA_GEOM_SPH, bounces=1)
    out = np.zeros(8, dtype=ARRIVAL_DTYPE)
    pts = np.zeros(10, dtype=PATH_DTYPE)
    n_out, n_pts, st = C.c_int(0), C.c_int64(0), RunStats()
    _dp = C.POINTER(C.c_double)
    rc = engine.L.ga_fan_run_paths(engine.h, aid, C.byref(p), L.ctypes.data_as(_dp), len(L), GA_PATH_RAYS,
                                   out.ctypes.data_as(C.POINTER(Arrival)), 8, C.byref(n_out),
                                   pts.ctypes.data_as(C.POINTER(PathPt)), 10, C.byref(n_pts), C.byref(st))
    assert rc == GA_E_CAP and n_pts.value > 10
    a2, p2, _ = engine.fan_paths(aid, p, L, GA_PATH_RAYS, pts_cap=int(n_pts.value))
    assert len(p2) == n_pts.value
    with pytest.raises(EngineError):
        engine.fan_paths(aid, prop_params(GA_GEOM_SPH, calc_amp=0, bounces=1), L, GA_PATH_CAUSTICS)
    engine.release_atmo(aid)


def test_rngdep_fan_paths_match_oracle(engine):
    from test_rngdep_fan_gpu import workload
    g = np.load(os.path.join(ROOT, "tests", "golden", "paths_rngdep.npz"))
    _, units, topo, _ = workload()
    rid = engine.rngdep_load(*units)
    engine.topo_load(rid, *topo)
    p = prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=1, bounces=2)
    arr, pts, st = engine.fan_paths(rid, p, g["launches"], GA_PATH_RAYS | GA_PATH_CAUSTICS)
    print(f"\nrngdep: {len(arr)} arrivals, {len(pts)} rows, {st.ray_steps} ray-steps (oracle {int(g['steps'])})")
    src = (40.0, -100.0)
    fails, summ = compare(arr, g["arrivals"], True, src=src)
    assert not fails, (fails, summ)
    pf, ps = compare_paths(pts, g["paths"], True, src=src + (0.0,))
    print(ps)
    assert not pf, (pf, ps)
    assert st.ray_steps == int(g["steps"])
    engine.rngdep_release(rid)


def test_eigenray_paths_match_oracle(engine, oracle, toyatmo_units):
    from ga_oracle_py import GEOM_SPH, Medium
    aid = engine.load_atmo(*toyatmo_units)
    src = (30.0, -100.0, 0.0)
    p = prop_params(GA_GEOM_SPH, src=src)
    e = eig_params(bnc_min=0, bnc_max=1)
    eig, pts, lat, st = engine.eig_search_paths(aid, p, e, [(30.25, -104.25)])
    assert len(eig) == 3
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    op = m.params(True)
    for i, a in enumerate(eig):
        q = pts[pts["launch_idx"] == i]
        rp, row = m.eig_path(op, src, a["incl"], a["az"], int(a["n_bounce"]))
        rp["launch_idx"] = i
        pf, ps = compare_paths(q, rp, True, src=src)
        print(i, len(q), ps)
        assert not pf, (i, pf, ps)
        assert abs(row["time"] - a["time"]) <= 1e-6 * row["time"]
    engine.release_atmo(aid)


def test_cli_writes_path_files(tmp_path, oracle, toyatmo_units):
    """infraga-sph -prop with write_rays and write_caustics: raypaths.dat has
    one block per ray (rows at precision 8, blank line after each ray),
    caustics-N.dat one file per bounce count, and the rows equal the
    oracle's write-mode rows to the printed precision."""
    from ga_oracle_py import GEOM_SPH, PATH_CAUSTICS, PATH_RAYS, Medium
    met = tmp_path / "toy.met"
    np.savetxt(met, np.load(os.path.join(ROOT, "tests", "golden", "toyatmo.npz"))["zTuvdp"])
    exe = os.path.join(ROOT, "infraga_amd", "bin", "infraga-sph")
    args = [exe, "-prop", str(met), "src_lat=30.0", "src_lon=-100.0", "azimuth=-90.0", "incl_min=4",
            "incl_max=36", "incl_step=8", "bounces=1", "write_caustics=true", "max_rng=800"]
    subprocess.run(args, check=True, cwd=tmp_path, capture_output=True, timeout=120)
    txt = (tmp_path / "toy.raypaths.dat").read_text().split("\n")
    assert txt[0].startswith("# lat [deg]")
    blocks = "\n".join(txt[1:]).split("\n\n")
    rows = [l for l in txt[1:] if l.strip()]
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    op = m.params(True)
    op.rng_max = 800.0
    L = np.array([[th, -90.0] for th in (4.0, 12.0, 20.0, 28.0, 36.0)])
    _, rp, _ = m.fan_paths(L, op, PATH_RAYS | PATH_CAUSTICS, bounces=1, src=(30.0, -100.0, 0.0))
    r_rows = rp[rp["kind"] == 0]
    assert len(rows) == len(r_rows)
    assert len([b for b in blocks if b.strip()]) == 5
    got = np.array([[float(x) for x in l.split("\t")] for l in rows])
    exp = np.stack([r_rows[f] for f in ("pos0", "pos1", "z", "tl", "absorption", "time")], axis=1)
    np.testing.assert_allclose(got, exp, rtol=2e-7, atol=1e-6)
    for b in (0, 1):
        c = (tmp_path / f"toy.caustics-{b}.dat").read_text().split("\n")
        assert c[0] == "# lat [deg]\tlon [deg]\tz [km]\ttime [s]"
        n_c = len([l for l in c[1:] if l.strip()])
        assert n_c == int(((rp["kind"] == 1) & (rp["leg"] == b)).sum())
