"""Oracle eigenray fixtures for the batched 3D search (512 seeded pairs).

Runs the oracle's serial search (go_eig_search_3d: est_eigenray +
find_eigenray per receiver and bounce count) on the base stratified profile,
source at the origin on the ground, bounces 0..3, and writes
tests/golden/eig3d_base.npz with every eigenray found and the per-pair work
counters.  Takes about 35 s of CPU per pair; run with a process pool:

    python tests/golden/make_eig_golden.py [n_pairs] [n_procs] [--fma]

--fma repeats the search with the same oracle source compiled with x86 FMA
contraction (gcc -mfma -ffp-contract=fast: about one rounding per operation
changes) and writes eig3d_base_fma.npz.  The two fixtures together mark the
receivers whose eigenrays the reference's search itself only fixes up to its
arrival tolerance: where the two builds disagree, rounding-level
differences already move the result.
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
LIB = None


def build_fma(out_dir):
    import subprocess
    src = [os.path.join(ROOT, "oracle", f) for f in ("ga_oracle.c", "ga_interp.c", "ga_oracle_rd.c")]
    lib = os.path.join(out_dir, "libga_oracle_fma.so")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-mfma", "-ffp-contract=fast", "-std=c99",
                           "-I" + os.path.join(ROOT, "oracle"), "-shared", "-o", lib] + src + ["-lm"])
    return lib


def init(lib):
    global LIB
    LIB = lib


def one(i):
    from ga_oracle_py import GEOM_3D, Medium, Oracle
    from infraga_amd import synth
    units = synth.to_engine_units(*synth.base_profile())
    o = Oracle(LIB)
    m = Medium(o, GEOM_3D, *units)
    p = m.params(True)
    R = synth.eig_pairs()
    t0 = time.time()
    arr, steps = m.eig_search(p, (0.0, 0.0, 0.0), R[i], bnc_min=0, bnc_max=BNC_MAX)
    c = o.eig_counts()
    return i, arr, steps, time.time() - t0, [c[k] for k in o.EIG_COUNT_NAMES]


def main():
    import tempfile
    from multiprocessing import Pool
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    fma = "--fma" in sys.argv
    n = int(args[0]) if len(args) > 0 else 512
    procs = int(args[1]) if len(args) > 1 else 8
    lib = build_fma(tempfile.mkdtemp()) if fma else None
    rows, steps, secs, counts = [], np.zeros(n, np.int64), np.zeros(n), np.zeros((n, 8), np.int64)
    with Pool(procs, initializer=init, initargs=(lib,)) as pool:
        for i, arr, st, dt, c in pool.imap_unordered(one, range(n)):
            arr = arr.copy()
            arr["launch_idx"] = i
            rows.append(arr)
            steps[i], secs[i], counts[i] = st, dt, c
            print(f"pair {i}: {len(arr)} eigenrays, {st} steps, {dt:.1f} s", flush=True)
    allr = np.concatenate(rows) if rows else np.zeros(0)
    order = np.lexsort((np.arange(len(allr)), allr["n_bounce"], allr["launch_idx"])) if len(allr) else []
    allr = allr[order]
    name = "eig3d_base_fma.npz" if fma else "eig3d_base.npz"
    np.savez_compressed(os.path.join(HERE, name), eigenrays=allr, steps=steps, seconds=secs,
                        counts=counts, bnc_max=BNC_MAX, n_pairs=n)


if __name__ == "__main__":
    main()
