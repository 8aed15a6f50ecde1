"""Oracle fixture for the 3D range-dependent eigenray search with terrain
(infraga-3d-rngdep -eig_search).

Same grid, terrain and source as tests/test_rngdep3d_gpu.py (6 x 6 profile
nodes over +-500 km, 81 x 81 fractal terrain 0-2 km, source at the region
centre on the ground); the receivers below; inclinations 0.5-20 deg;
bounces 0..1; 0.1 km tolerance.  Runs go_rd3_eig_search (the 3D search of
geoac.eigenray.3d.cpp over the oracle's Cartesian range-dependent medium)
and writes tests/golden/rd3eig.npz.  A few minutes of CPU:

    python tests/golden/make_rd3eig_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

RECEIVERS = np.array([[255.0, 20.0], [195.0, 205.0], [-15.0, -279.0]])
INCL_MAX = 20.0


def one(i):
    from ga_oracle_py import Oracle, Rd3Medium
    from test_rngdep3d_gpu import workload
    units, topo, _ = workload()
    m = Rd3Medium(Oracle(), *units)
    m.set_topo(*topo)
    arr, steps = m.eig_search(m.params(True), (0.0, 0.0, 0.0), RECEIVERS[i], incl_min=0.5, incl_max=INCL_MAX,
                              bnc_min=0, bnc_max=1)
    arr["launch_idx"] = i
    return arr, steps


def main():
    from multiprocessing import Pool
    with Pool(len(RECEIVERS)) as pool:
        res = pool.map(one, range(len(RECEIVERS)))
    arrs = np.concatenate([a for a, _ in res])
    for i, (a, s) in enumerate(res):
        print(f"receiver {i}: {len(a)} eigenrays {a['n_bounce'].tolist()}, {s} ray-steps", flush=True)
    np.savez_compressed(os.path.join(HERE, "rd3eig.npz"), receivers=RECEIVERS, incl_max=INCL_MAX, eigenrays=arrs,
                        steps=np.array([s for _, s in res]))


if __name__ == "__main__":
    main()
