"""Oracle fixture for the range-dependent write_rays / write_caustics fan.

Same grid, terrain and source as make_rngdep_golden.py (6 x 6 profile nodes
over 30..50 N, 110..90 W, 81 x 81 fractal terrain, source at the region
centre on the ground); 4 launches (azimuths -90, 135 deg x inclinations
8, 18 deg), 2 bounces, 18 equations.  Runs go_rd_fan_paths (the oracle's
restatement of run_prop with write_rays and write_caustics,
main.infraga.sph.rngdep.cpp:451-490) and writes tests/golden/paths_rngdep.npz
(arrivals, path rows, ray-steps).  About a minute of CPU:

    python tests/golden/make_paths_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

BOUNCES = 2
LAUNCHES = np.array([[th, az] for az in (-90.0, 135.0) for th in (8.0, 18.0)])


def main():
    from ga_oracle_py import PATH_CAUSTICS, PATH_RAYS, Oracle, RdMedium
    from make_rngdep_golden import workload
    units, topo, _ = workload()
    m = RdMedium(Oracle(), *units)
    m.set_topo(*topo)
    arr, pts, steps = m.fan_paths(LAUNCHES, m.params(True), PATH_RAYS | PATH_CAUSTICS, bounces=BOUNCES, src=None)
    print(f"{len(arr)} arrivals, {len(pts)} path rows ({int((pts['kind'] == 1).sum())} caustics), {steps} ray-steps")
    np.savez_compressed(os.path.join(HERE, "paths_rngdep.npz"), launches=LAUNCHES, arrivals=arr, paths=pts,
                        steps=np.int64(steps))


if __name__ == "__main__":
    main()
