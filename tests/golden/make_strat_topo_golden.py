"""Oracle fixture for the stratified profile over terrain (infraga-sph -prop
and -eig_search with topo_file=).

The base profile over a 41 x 41 fractal terrain (seed 3, 30..50 N,
110..90 W, up to 2 km) flattened towards the centre by 1 - exp(-d^2 / 8)
(d in degrees from 40 N 100 W, so the source sits on level ground and the
eigenray search starts at the driver's floor), run by the oracle's stratified-over-terrain
restatement (RdMedium.strat_over_terrain: identical columns with the
stratified semantics, pinned bit-exact against the reference's compiled
stratified medium with terrain in tests/test_oracle_strat_topo.py).
Source at the terrain's centre on the ground.
- fans: 20 rays (azimuths 0, 60, 135, 250, 300 x inclinations 4, 9, 14,
  22 deg), 3 bounces: 18 equations, 6 equations, 18 equations with the
  boundary-layer vertical wind;
- eigenray searches to two receivers, inclinations 0.5..12 deg, bounces
  0..1, 0.1 km tolerance.
Writes tests/golden/strat_topo.npz (a few minutes of CPU):

    python tests/golden/make_strat_topo_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

RECEIVERS = np.array([[39.071, -96.864], [43.101, -95.645]])
INCL_MAX = 25.0
BOUNCES = 3
CASES = {"amp18": (True, False), "amp6": (False, False), "blw18": (True, True)}


def scene():
    from infraga_amd import synth
    prof = synth.to_engine_units(*synth.base_profile())
    tl, tn, zg = synth.fractal_topo(3, lat=np.linspace(30.0, 50.0, 41), lon=np.linspace(-110.0, -90.0, 41), amp_km=2.0)
    d2 = (tl[:, None] - 40.0) ** 2 + (tn[None, :] + 100.0) ** 2
    topo = (tl, tn, zg * (1.0 - np.exp(-d2 / 8.0)))
    launches = np.array([[th, az] for az in (0.0, 60.0, 135.0, 250.0, 300.0) for th in (4.0, 9.0, 14.0, 22.0)])
    return prof, topo, launches


def medium():
    from ga_oracle_py import Oracle, RdMedium
    prof, topo, _ = scene()
    return RdMedium.strat_over_terrain(Oracle(), *prof, *topo)


def fan(case):
    amp, blw = CASES[case]
    _, _, L = scene()
    m = medium()
    m.config(use_blw=blw)
    return m.fan(L, m.params(amp), bounces=BOUNCES, src=None)


def eig(i):
    m = medium()
    la, lo = m.centre()
    arr, steps = m.eig_search(m.params(True), (la, lo, 0.0), RECEIVERS[i], incl_min=0.5, incl_max=INCL_MAX,
                              bnc_min=0, bnc_max=1)
    arr["launch_idx"] = i
    return arr, steps


def job(j):
    return fan(j) if isinstance(j, str) else eig(j)


def main():
    from multiprocessing import get_context
    jobs = list(CASES) + list(range(len(RECEIVERS)))
    with get_context("spawn").Pool(len(jobs)) as pool:
        res = pool.map(job, jobs, chunksize=1)
    out = {"receivers": RECEIVERS, "incl_max": INCL_MAX, "bounces": BOUNCES}
    for j, (a, s) in zip(jobs, res):
        print(f"{j}: {len(a)} rows, {s} ray-steps", flush=True)
        if isinstance(j, str):
            out[j] = a
            out[j + "_steps"] = s
    out["eigenrays"] = np.concatenate([a for j, (a, _) in zip(jobs, res) if not isinstance(j, str)])
    np.savez_compressed(os.path.join(HERE, "strat_topo.npz"), **out)


if __name__ == "__main__":
    main()
