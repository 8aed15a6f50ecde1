"""Conditioning of the 3D eigenray iteration on the 512-pair batch.

For every starting point of find_eigenray in the batch
(tests/golden/eig3d_lm_starts.txt) the oracle's find_eigenray
(geoac.eigenray.3d.cpp:146-366 restated, go_eig_find_3d_traj) runs from the
start and from starts perturbed by relative 1e-13 and 1e-11 in theta and
phi, recording every iterate.  FP64 implementations of the traces differ
at 1e-12 to 1e-9 relative, so an iteration is not determined at FP64
precision (unstable) when a perturbation of that size
- changes whether it accepts, or
- separates the two runs' iterates by more than 1e-6 rad, or changes the
  number of iterations.
The second case covers iterations that accept from every probed start but
not from starts a few 1e-13 away (receiver 510 direct: the start moved by
2.5e-13 in phi wanders from the second iterate on and fails after 25
iterations, the probed ones accept after 5), and iterations that wander and
accept by chance (receiver 502: the start moved by 3.5e-14 accepts in 4
iterations).
The eigenray count of a (receiver, bounces) group with an unstable start may
legitimately differ between implementations, including two builds of the
oracle itself.  Writes tests/golden/eig3d_sensitivity.npz:

    start: (receiver, bounces, seq, theta, phi) of each start
    accepted[i, j]: outcome under perturbation j (PERTURB below)
    incl[i, j], az[i, j]: accepted launch angles [deg] (NaN if not accepted)
    max_sep[i, j]: largest separation of the iterates from the unperturbed run [rad]
    unstable[i]: the criterion above

    python tests/golden/make_eig_sensitivity.py [n_procs]
    python tests/golden/make_eig_sensitivity.py --reclassify   (re-apply
        the criterion to the stored probe results)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# (relative change of theta, relative change of phi)
PERTURB = [(0.0, 0.0), (1e-13, 1e-13), (-1e-13, 0.0), (0.0, -1e-13), (1e-11, -1e-11)]


def classify(accepted, max_sep):
    """unstable[i] for the criterion of the module docstring."""
    accepted = np.asarray(accepted, dtype=bool)
    changes = accepted.any(axis=1) & ~accepted.all(axis=1)
    return changes | (np.asarray(max_sep).max(axis=1) > 1e-6)


def load_starts():
    rows = []
    for line in open(os.path.join(HERE, "eig3d_lm_starts.txt")):
        if line.startswith("#") or not line.strip():
            continue
        t = line.split()
        rows.append((int(t[0]), int(t[1]), int(t[2]), float(t[3]), float(t[4])))
    return rows


def probe(start):
    from ga_oracle_py import GEOM_3D, Medium, Oracle
    from infraga_amd import synth
    ri, nb, seq, th, ph = start
    m = Medium(Oracle(), GEOM_3D, *synth.to_engine_units(*synth.base_profile()))
    p = m.params(True)
    R = synth.eig_pairs()
    out, base = [], None
    for dth, dph in PERTURB:
        row, traj = m.eig_find_traj(p, (0.0, 0.0, 0.0), R[ri], th * (1.0 + dth), ph * (1.0 + dph), nb)
        if base is None:
            base = traj
        n = min(len(base), len(traj))
        sep = float(np.abs(base[:n, :2] - traj[:n, :2]).sum(axis=1).max()) if n else 0.0
        if len(base) != len(traj):
            sep = max(sep, 1.0)
        out.append((row is not None, np.nan if row is None else row["incl_launch"],
                    np.nan if row is None else row["az_launch"], sep))
    return out


def reclassify():
    path = os.path.join(HERE, "eig3d_sensitivity.npz")
    d = dict(np.load(path))
    d["unstable"] = classify(d["accepted"], d["max_sep"])
    np.savez(path, **d)
    print(f"{len(d['start'])} starts, {int(d['unstable'].sum())} not determined at FP64 precision")


def main():
    import multiprocessing as mp
    if sys.argv[1:] == ["--reclassify"]:
        return reclassify()
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()
    starts = load_starts()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(probe, starts, chunksize=1)
    acc = np.array([[r[0] for r in x] for x in res])
    incl = np.array([[r[1] for r in x] for x in res])
    az = np.array([[r[2] for r in x] for x in res])
    sep = np.array([[r[3] for r in x] for x in res])
    unstable = classify(acc, sep)
    np.savez(os.path.join(HERE, "eig3d_sensitivity.npz"), start=np.array(starts, dtype=np.float64), accepted=acc,
             incl=incl, az=az, max_sep=sep, unstable=unstable, perturb=np.array(PERTURB))
    print(f"{len(starts)} starts, {int(unstable.sum())} not determined at FP64 precision")


if __name__ == "__main__":
    main()
