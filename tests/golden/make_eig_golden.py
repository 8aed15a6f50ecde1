"""Oracle eigenray fixtures for the batched 3D search (512 seeded pairs).

Runs the oracle's serial search (go_eig_search_3d: est_eigenray +
find_eigenray per receiver and bounce count) on the base stratified profile,
source at the origin on the ground, bounces 0..3, and writes
tests/golden/eig3d_base.npz with every eigenray found and the per-pair work
counters.  Takes about 35 s of CPU per pair; run with a process pool:

    python tests/golden/make_eig_golden.py [n_pairs] [n_procs]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

BNC_MAX = 3


def one(i):
    from ga_oracle_py import GEOM_3D, Medium, Oracle
    from infraga_amd import synth
    units = synth.to_engine_units(*synth.base_profile())
    o = Oracle()
    m = Medium(o, GEOM_3D, *units)
    p = m.params(True)
    R = synth.eig_pairs()
    t0 = time.time()
    arr, steps = m.eig_search(p, (0.0, 0.0, 0.0), R[i], bnc_min=0, bnc_max=BNC_MAX)
    c = o.eig_counts()
    return i, arr, steps, time.time() - t0, [c[k] for k in o.EIG_COUNT_NAMES]


def main():
    from multiprocessing import Pool
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    rows, steps, secs, counts = [], np.zeros(n, np.int64), np.zeros(n), np.zeros((n, 8), np.int64)
    with Pool(procs) as pool:
        for i, arr, st, dt, c in pool.imap_unordered(one, range(n)):
            arr = arr.copy()
            arr["launch_idx"] = i
            rows.append(arr)
            steps[i], secs[i], counts[i] = st, dt, c
            print(f"pair {i}: {len(arr)} eigenrays, {st} steps, {dt:.1f} s", flush=True)
    allr = np.concatenate(rows) if rows else np.zeros(0)
    order = np.lexsort((np.arange(len(allr)), allr["n_bounce"], allr["launch_idx"])) if len(allr) else []
    allr = allr[order]
    np.savez_compressed(os.path.join(HERE, "eig3d_base.npz"), eigenrays=allr, steps=steps, seconds=secs,
                        counts=counts, bnc_max=BNC_MAX, n_pairs=n)


if __name__ == "__main__":
    main()
