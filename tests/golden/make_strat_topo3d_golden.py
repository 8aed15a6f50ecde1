"""Oracle fixture for the 3D stratified profile over terrain (infraga-3d
-prop and -eig_search with topo_file=).

The base profile with 12 / -7 m/s winds at the ground (decaying over 5 km,
so that the boundary-layer scaling over sloped terrain has second
derivatives for the upper-triangle contraction to drop) over a 41 x 41
fractal terrain (seed 3) on x, y in -500..500 km, up to 2 km, flattened
towards the centre by 1 - exp(-d^2 / 8000) (d in km from the centre, so the
source sits on level ground), run by the oracle's 3D stratified-over-terrain
restatement (Rd3Medium.strat_over_terrain: identical columns with the
stratified 3D semantics, its medium pinned bit-exact against the
reference's compiled 3D stratified medium with terrain in
tests/test_oracle_strat_topo3d.py).  Source at the terrain's centre on the
ground.
- fans: 20 rays (azimuths 0, 60, 100, 250, 300 x inclinations 4, 9, 14,
  22 deg), 3 bounces: 18 equations, 6 equations, and 18 equations with the
  write_rays / write_caustics bookkeeping;
- eigenray searches to the receivers below, inclinations 0.5..20 deg,
  bounces 0..1, 0.1 km tolerance.
Writes tests/golden/strat_topo3d.npz (a few minutes of CPU):

    python tests/golden/make_strat_topo3d_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

RECEIVERS = np.array([[257.0, 2.0], [239.0, 139.0]])
INCL_MAX = 20.0
BOUNCES = 3


def scene():
    from infraga_amd import synth
    z, c, u, v, rho = synth.to_engine_units(*synth.base_profile())
    u = u + 0.012 * np.exp(-z / 5.0)
    v = v - 0.007 * np.exp(-z / 5.0)
    _, _, zg = synth.fractal_topo(3, lat=np.linspace(30.0, 50.0, 41), lon=np.linspace(-110.0, -90.0, 41), amp_km=2.0)
    tx = np.linspace(-500.0, 500.0, 41)
    d2 = tx[:, None] ** 2 + tx[None, :] ** 2
    topo = (tx, tx.copy(), zg * (1.0 - np.exp(-d2 / 8000.0)))
    launches = np.array([[th, az] for az in (0.0, 60.0, 100.0, 250.0, 300.0) for th in (4.0, 9.0, 14.0, 22.0)])
    return (z, c, u, v, rho), topo, launches


def medium():
    from ga_oracle_py import Oracle, Rd3Medium
    prof, topo, _ = scene()
    return Rd3Medium.strat_over_terrain(Oracle(), *prof, *topo)


def job(i):
    from ga_oracle_py import PATH_CAUSTICS, PATH_RAYS
    m = medium()
    _, _, L = scene()
    if i == 0:
        a, _, s = m.fan3(L, m.params(True), bounces=BOUNCES)
        return "amp18", a, s, None
    if i == 1:
        a, _, s = m.fan3(L, m.params(False), bounces=BOUNCES)
        return "amp6", a, s, None
    if i == 2:
        a, q, s = m.fan3(L, m.params(True), bounces=BOUNCES, write=True, flags=PATH_RAYS | PATH_CAUSTICS)
        return "paths18", a, s, q
    r = i - 3
    arr, s = m.eig_search(m.params(True), (0.0, 0.0, 0.0), RECEIVERS[r], incl_min=0.5, incl_max=INCL_MAX,
                          bnc_min=0, bnc_max=1)
    arr["launch_idx"] = r
    return f"eig{r}", arr, s, None


def main():
    from multiprocessing import Pool
    with Pool(3 + len(RECEIVERS)) as pool:
        res = pool.map(job, range(3 + len(RECEIVERS)))
    out = {"receivers": RECEIVERS, "incl_max": INCL_MAX, "bounces": BOUNCES}
    eig = []
    for name, a, s, q in res:
        print(name, len(a), "rows", s, "ray-steps", flush=True)
        if name.startswith("eig"):
            eig.append(a)
            continue
        out[name], out[name + "_steps"] = a, np.int64(s)
        if q is not None:
            out[name + "_pts"] = q
    out["eigenrays"] = np.concatenate(eig)
    np.savez_compressed(os.path.join(HERE, "strat_topo3d.npz"), **out)


if __name__ == "__main__":
    main()
