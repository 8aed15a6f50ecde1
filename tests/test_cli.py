"""The drop-in front ends infraga-sph / infraga-3d.

CPU: argument handling that ends before any GPU use (usage, unknown keys,
unsupported options).  GPU: the reference's own printed answers for its
example profile ToyAtmo.met, read back from the .arrivals.dat text the front
ends write (-prop single rays, -eig_search).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "infraga_amd", "bin")


def run(name, *args, cwd=None):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    return subprocess.run([exe, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


def test_usage_without_arguments():
    r = run("infraga-sph")
    assert r.returncode == 0 and "-prop" in r.stdout and "-eig_search" in r.stdout


@pytest.mark.parametrize("args", [
    ("-prop", "x.met", "no_such_key=1"),
    ("-prop", "x.met", "write_topo=true"),
    ("-eig_search", "x.met", "no_such_key=1"),
    ("-back_proj", "x.met"),
])
def test_rejected_arguments(args):
    r = run("infraga-3d", *args)
    assert r.returncode == 2, (r.stdout, r.stderr)


@pytest.mark.parametrize("name", ["infraga-accel-sph", "infraga-accel-3d", "infraga-accel-sph-rngdep"])
def test_accel_ranks_above_zero_exit_quietly(name, tmp_path):
    """Under mpirun every rank starts the binary; ranks above 0 leave at once
    (before any GPU or file work), rank 0 does the whole job."""
    env = dict(os.environ, PMI_RANK="3", PMI_SIZE="8")
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    r = subprocess.run([exe, "-prop", "missing.met"], capture_output=True, text=True, cwd=tmp_path, env=env,
                       timeout=60)
    assert r.returncode == 0 and r.stdout == "" and r.stderr == ""
    assert list(tmp_path.iterdir()) == []
    r = subprocess.run([exe, "-prop", "missing.met", "no_such_key=1"], capture_output=True, text=True, cwd=tmp_path,
                       env=dict(os.environ, OMPI_COMM_WORLD_RANK="0", OMPI_COMM_WORLD_SIZE="2"), timeout=60)
    assert r.returncode == 2


@pytest.fixture
def toy_met(tmp_path, toyatmo_cols):
    path = tmp_path / "ToyAtmo.met"
    np.savetxt(path, toyatmo_cols, fmt="%.17g")
    return str(path)


@pytest.mark.gpu
def test_prop_known_answer_text(toy_met, tmp_path):
    r = run("infraga-sph", "-prop", toy_met, "src_lat=30.0", "src_lon=-100.0", "inclination=10.0",
            "azimuth=-90.0", "bounces=1", "write_rays=false", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in open(tmp_path / "ToyAtmo.arrivals.dat") if l.strip() and not l.startswith("#")]
    assert rows[0][3:6] == ["29.985815", "-101.96135", "659.74062"]


@pytest.mark.gpu
def test_eig_search_known_answer_text(toy_met, tmp_path):
    r = run("infraga-sph", "-eig_search", toy_met, "src_lat=30.0", "src_lon=-100.0", "rcvr_lat=30.25",
            "rcvr_lon=-104.25", "bnc_max=1", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in open(tmp_path / "ToyAtmo.arrivals.dat") if l.strip() and not l.startswith("#")]
    assert [x[0] for x in rows] == ["34.559248", "4.1385204", "30.078055"]
    assert [x[2] for x in rows] == ["0", "1", "1"]
    for i in range(3):
        txt = open(tmp_path / f"ToyAtmo.eigenray-{i}.dat").read().split("\n")
        assert txt[0] == "# lat [deg]\tlon [deg]\tz [km]\ttrans. coeff. [dB]\tabsorption [dB]\ttime [s]"
        assert len([l for l in txt[1:] if l.strip()]) > 100


@pytest.mark.gpu
def test_eig_direct_text(toy_met, tmp_path):
    """-eig_direct from an estimate near the reference's 1-bounce eigenray
    converges to it (launch angles as -eig_search prints them, within the
    acceptance tolerance) and writes eigenray-0.dat only."""
    r = run("infraga-sph", "-eig_direct", toy_met, "src_lat=30.0", "src_lon=-100.0", "rcvr_lat=30.25",
            "rcvr_lon=-104.25", "incl_est=4.0", "bounces=1", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "Eigenray identified" in r.stdout
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["ToyAtmo.eigenray-0.dat", "ToyAtmo.met"]
    rows = [l for l in open(tmp_path / "ToyAtmo.eigenray-0.dat") if l.strip() and not l.startswith("#")]
    assert len(rows) > 100


@pytest.mark.gpu
def test_accel_launch_list_is_the_mpi_drivers(toy_met, tmp_path):
    """infraga-accel-sph as rank 0 of 8: inclinations 0.1..45 step 0.1 give
    450 launches per azimuth (the MPI driver's list; the serial loop gives
    449), with the MPI driver's theta + step * j values."""
    env = dict(os.environ, PMI_RANK="0", PMI_SIZE="8")
    exe = os.path.join(BIN, "infraga-accel-sph")
    r = subprocess.run([exe, "-prop", toy_met, "incl_min=0.1", "incl_max=45", "incl_step=0.1", "azimuth=-90",
                        "bounces=0", "max_rng=60", "write_rays=false", "src_lat=30", "src_lon=-100"],
                       capture_output=True, text=True, cwd=tmp_path, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "450 rays" in r.stdout, r.stdout
    r = run("infraga-sph", "-prop", toy_met, "incl_min=0.1", "incl_max=45", "incl_step=0.1", "azimuth=-90",
            "bounces=0", "max_rng=60", "write_rays=false", "src_lat=30", "src_lon=-100", cwd=str(tmp_path))
    assert r.returncode == 0 and "449 rays" in r.stdout, r.stdout


@pytest.mark.parametrize("args", [
    ("-prop", "prefix", "lats.dat"),
    ("-prop", "prefix", "lats.dat", "lons.dat", "no_such_key=1"),
    ("-prop", "prefix", "lats.dat", "lons.dat", "write_topo=true"),
    ("-eig_search", "prefix", "lats.dat"),
    ("-eig_search", "prefix", "lats.dat", "lons.dat", "no_such_key=1"),
    ("-eig_direct", "prefix", "lats.dat"),
    ("-wnl_wvfrm", "prefix", "lats.dat"),
])
def test_rngdep_rejected_arguments(args):
    r = run("infraga-sph-rngdep", *args)
    assert r.returncode == 2, (r.stdout, r.stderr)


@pytest.mark.parametrize("args", [
    ("-wnl_wvfrm", "x.met", "no_such_key=1"),
])
def test_wnl_wvfrm_rejected_arguments(args):
    """Unknown keys are refused before any GPU work."""
    r = run("infraga-sph", *args)
    assert r.returncode == 2, (r.stdout, r.stderr)


@pytest.mark.gpu
def test_rngdep_prop_text(engine, tmp_path):
    """infraga-sph-rngdep -prop on grid files with terrain: header, the
    reference's launch loops (each azimuth starts at the terrain slope along
    it plus 0.1 degree, main.infraga.sph.rngdep.cpp:415-428) and rows equal
    to the engine's fan over the same launches."""
    from ga_oracle_py import write_rd_files
    from infraga_amd import synth
    from infraga_amd.engine import GA_GEOM_SPH_RNGDEP, prop_params
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    topo = synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81), amp_km=2.0)
    lat, lon, z, T3, u3, v3, rho3 = grid
    pre, lats, lons, tp = write_rd_files(str(tmp_path), lat, lon, z, synth.sound_speed(T3), u3, v3, rho3, topo=topo)
    r = run("infraga-sph-rngdep", "-prop", pre, lats, lons, "topo_file=" + tp, "prof_format=zcuvd", "incl_min=0.0",
            "incl_max=14.0", "incl_step=2.0", "az_min=-60.0", "az_max=60.0", "az_step=60.0", "bounces=2",
            "write_rays=false", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    text = open(pre + ".arrivals.dat").read()
    assert text.startswith("# infraga-sph-rngdep -prop summary:\n")
    assert "#\ttopo file:" + tp + "\n" in text
    rows = np.array([[float(x) for x in l.split("\t")] for l in text.splitlines() if l.strip() and not l.startswith("#")])
    assert len(rows) > 0
    rid = engine.rngdep_load_files(pre, lats, lons, tp, fmt="zcuvd")
    try:
        info = engine.rngdep_info(rid)
        la = (info[1] + info[0]) / 2.0
        lo = (info[3] + info[2]) / 2.0
        _, t = engine.rngdep_eval(rid, [(la, lo, 0.0)])
        # first launch of each azimuth: atan(terrain slope along it) + 0.1 degree
        for az in (-60.0, 0.0, 60.0):
            slope = (t[0, 2] / (6370.0 * np.cos(la)) * np.sin(np.radians(az)) + t[0, 1] / 6370.0 * np.cos(np.radians(az)))
            th_g = np.degrees(np.arctan(slope)) + 0.1
            mine = rows[rows[:, 1] == az]
            if len(mine):
                assert mine[:, 0].min() >= max(0.0, th_g) - 1e-6   # rows carry 8 significant digits
        launches = np.unique(rows[:, :2], axis=0)
        launches = launches[np.lexsort((launches[:, 0], launches[:, 1]))]
        p = prop_params(GA_GEOM_SPH_RNGDEP, bounces=2)
        arr, _ = engine.fan(rid, p, launches)
        got = {(round(a, 9), round(b, 9), int(c)): row for a, b, c, *row in rows}
        for A in arr:
            k = (round(launches[A["launch_idx"]][0], 9), round(launches[A["launch_idx"]][1], 9), int(A["n_bounce"]))
            if k not in got:
                continue
            row = got[k]
            assert abs(row[0] - A["pos0"]) < 1e-6 and abs(row[1] - A["pos1"]) < 1e-5
            assert abs(row[2] - A["time"]) / A["time"] < 1e-5
    finally:
        engine.rngdep_release(rid)
