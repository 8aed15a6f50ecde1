"""Write golden vectors from the reference's own compiled sources.

Runs the reference's spline, medium, Sutherland-Bass and globe code, built
from /root/reference by oracle/Makefile into oracle/_ref/, on the synthetic
profiles and stores inputs and outputs as small .npz fixtures next to this
script.  The fixtures let the oracle be checked where the reference tree is
absent.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from ga_oracle_py import GEOM_3D, GEOM_SPH, Ref  # noqa: E402
from infraga_amd import synth  # noqa: E402


def sample_points(geom, rng):
    zs = np.concatenate([
        np.linspace(-0.5, 150.5, 301),              # incl. below / above the table (clamped)
        rng.uniform(0.0, 150.0, 200),
        np.arange(0.0, 150.01, 7.3),                 # some exactly on nodes
        np.array([0.0, 0.05, 0.1, 1.99, 2.0, 2.01, 149.9, 150.0]),
    ])
    if geom == GEOM_SPH:
        lat = rng.uniform(-1.2, 1.2, zs.size)
        lon = rng.uniform(-3.1, 3.1, zs.size)
        return np.stack([6370.0 + zs, lat, lon], axis=1)
    x = rng.uniform(-500, 500, zs.size)
    y = rng.uniform(-500, 500, zs.size)
    return np.stack([x, y, zs], axis=1)


def main():
    rng = np.random.default_rng(11)
    profiles = {"base": synth.base_profile(), "pert": synth.perturbed_profile()}
    freqs = np.array([0.01, 0.1, 1.0, 5.0])
    for pname, prof in profiles.items():
        with tempfile.TemporaryDirectory() as td:
            met = os.path.join(td, f"{pname}.met")
            synth.write_zcuvd(met, *prof)
            for geom, gname in ((GEOM_SPH, "sph"), (GEOM_3D, "3d")):
                ref = Ref(geom)
                ref.load(met, "zcuvd")
                L = ref.lib
                tables = np.stack([np.stack(ref.tables(w)) for w in range(4)])   # (4, 3, n)
                pts = sample_points(geom, rng)
                med = np.zeros((len(pts), 7 + 9))
                uvw1 = np.zeros((len(pts), 12))
                uvw2 = np.zeros((len(pts), 30))
                sb = np.zeros((len(pts), len(freqs)))
                buf = np.zeros(30)
                for i, (a, b, c) in enumerate(pts):
                    med[i, :4] = [L.ref_c(a, b, c), L.ref_u(a, b, c), L.ref_v(a, b, c), L.ref_rho(a, b, c)]
                    med[i, 4:7] = [L.ref_dc(a, b, c, 0 if geom == GEOM_SPH else 2), 0.0, 0.0]
                    med[i, 7:10] = [L.ref_du(a, b, c, n) for n in range(3)]
                    med[i, 10:13] = [L.ref_dv(a, b, c, n) for n in range(3)]
                    med[i, 13:16] = [L.ref_dc(a, b, c, n) for n in range(3)]
                    L.ref_calc_uvw(a, b, c, 0, buf.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
                    uvw1[i] = buf[:12]
                    L.ref_calc_uvw(a, b, c, 1, buf.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
                    uvw2[i] = buf[:30]
                    for j, f in enumerate(freqs):
                        sb[i, j] = L.ref_sb_alpha(a, b, c, f)
                extra = {}
                if geom == GEOM_SPH:
                    g = rng.uniform(-1.4, 1.4, (100, 4))
                    extra["gc_in"] = g
                    extra["gc_out"] = np.array([L.ref_gc_dist(*row) for row in g])
                    extra["bearing_out"] = np.array([L.ref_bearing(*row) for row in g])
                np.savez_compressed(os.path.join(HERE, f"ref_{pname}_{gname}.npz"), tables=tables, pts=pts, med=med,
                                    uvw1=uvw1, uvw2=uvw2, freqs=freqs, sb=sb,
                                    z_bndlyr=L.ref_z_bndlyr(), z0=L.ref_z0(), alt_max=L.ref_alt_max(), **extra)
                print("wrote", pname, gname, len(pts))


if __name__ == "__main__":
    main()
