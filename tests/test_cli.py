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
    ("-prop", "x.met", "write_caustics=true"),
    ("-eig_search", "x.met", "topo_file=t.dat"),
    ("-back_proj", "x.met"),
])
def test_rejected_arguments(args):
    r = run("infraga-3d", *args)
    assert r.returncode == 2, (r.stdout, r.stderr)


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
