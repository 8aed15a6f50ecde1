"""Oracle fixtures for the range-dependent spherical fan (terrain).

Grid: 6 x 6 profile nodes over 30..50 N, 110..90 W (synth.rngdep_grid,
seed 3, 0.5 km levels to 120 km) and an 81 x 81 fractal terrain
(synth.fractal_topo, seed 5, 0..2 km); source at the region centre on the
ground; 20 launches (azimuths -90, 0, 45, 135, 270 deg x inclinations 3, 8,
12, 18 deg); 3 bounces.  Cases: 18 equations, 6 equations, and 18 equations
with the boundary-layer vertical wind.  Runs go_rd_fan (the oracle's
restatement of run_prop, main.infraga.sph.rngdep.cpp:412-554) and writes
tests/golden/rngdep_fan.npz.  About a minute of CPU per case:

    python tests/golden/make_rngdep_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASES = (("amp18", True, False), ("amp6", False, False), ("blw18", True, True))
BOUNCES = 3


def workload():
    from infraga_amd import synth
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    units = synth.rngdep_engine_units(*grid)
    topo = synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81), amp_km=2.0)
    launches = np.array([[th, az] for az in (-90.0, 0.0, 45.0, 135.0, 270.0) for th in (3.0, 8.0, 12.0, 18.0)])
    return units, topo, launches


def one(case):
    from ga_oracle_py import Oracle, RdMedium
    name, amp, blw = case
    units, topo, launches = workload()
    m = RdMedium(Oracle(), *units)
    m.set_topo(*topo)
    m.config(use_blw=blw)
    arr, steps = m.fan(launches, m.params(amp), bounces=BOUNCES, src=None)
    return name, arr, steps


def main():
    from multiprocessing import Pool
    with Pool(len(CASES)) as pool:
        res = pool.map(one, CASES)
    out = {}
    for name, arr, steps in res:
        out[name] = arr
        out[name + "_steps"] = np.int64(steps)
        print(f"{name}: {len(arr)} arrivals, {steps} ray-steps", flush=True)
    np.savez_compressed(os.path.join(HERE, "rngdep_fan.npz"), **out)


if __name__ == "__main__":
    main()
