"""Copies the reference's example atmosphere examples/ToyAtmo.met (zTuvdp
columns: z [km], T [K], u, v [m/s], rho [g/cm^3], p [mbar]) into
tests/golden/toyatmo.npz, so that tests and GPU runs need no reference tree.

The known answers checked against it (tests/test_oracle_kat.py) are the
reference's own printed outputs for this profile, recorded from a build of
the reference sources (setprecision(8) text):
  infraga-sph -prop, src 30 N 100 W, one ray incl 10, az -90, bounces 1:
      first ground arrival  29.985815  -101.96135  659.74062 s
  one ray incl 4.1392114, az -84.970504, bounces 1, 1-bounce arrival:
      30.249254  -104.24987  1412.4563 s  TL -39.828658  abs -0.1609091
  infraga-sph -eig_search, src 30 N 100 W, rcvr 30.25 N 104.25 W, bnc_max 1:
      eigenray inclinations 34.559248, 4.1385204, 30.078055
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples/ToyAtmo.met"

if __name__ == "__main__":
    d = np.loadtxt(SRC)
    np.savez_compressed(os.path.join(HERE, "toyatmo.npz"), zTuvdp=d)
    print(d.shape)
