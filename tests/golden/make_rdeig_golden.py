"""Oracle fixture for the range-dependent eigenray search with terrain
(infraga-sph-rngdep -eig_search).

Same grid, terrain and source as make_rngdep_golden.py (6 x 6 profile nodes
over 30..50 N, 110..90 W, 81 x 81 fractal terrain, source at the region
centre on the ground); receivers at 38.35 N 97.90 W and 41.77 N 97.53 W;
inclinations up to 12 deg (the search starts at the driver's terrain
inclination, in degrees read as radians); bounces 0..1; 0.1 km tolerance.
Runs go_rd_eig_search (the spherical search of geoac.eigenray.sph.cpp over
the oracle's range-dependent medium) and writes tests/golden/rdeig.npz.  A
few minutes of CPU:

    python tests/golden/make_rdeig_golden.py
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

RECEIVERS = np.array([[38.35, -97.90], [41.77, -97.53]])
INCL_MAX = 12.0


def one(i):
    from ga_oracle_py import Oracle, RdMedium
    from make_rngdep_golden import workload
    units, topo, _ = workload()
    m = RdMedium(Oracle(), *units)
    m.set_topo(*topo)
    la, lo = m.centre()
    arr, steps = m.eig_search(m.params(True), (la, lo, 0.0), RECEIVERS[i], incl_min=0.5, incl_max=INCL_MAX,
                              bnc_min=0, bnc_max=1)
    arr["launch_idx"] = i
    return arr, steps


def main():
    from multiprocessing import Pool
    with Pool(len(RECEIVERS)) as pool:
        res = pool.map(one, range(len(RECEIVERS)))
    arrs = np.concatenate([a for a, _ in res])
    for i, (a, s) in enumerate(res):
        print(f"receiver {i}: {len(a)} eigenrays, {s} ray-steps", flush=True)
    np.savez_compressed(os.path.join(HERE, "rdeig.npz"), receivers=RECEIVERS, incl_max=INCL_MAX, eigenrays=arrs,
                        steps=np.array([s for _, s in res]))


if __name__ == "__main__":
    main()
