"""Writes tests/golden/rd3.npz: the oracle's go_rd3_fan (3D Cartesian
range-dependent run_prop) on the workload of tests/test_rngdep3d_gpu.py.

    PYTHONPATH=. python tests/golden/make_rd3_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

from ga_oracle_py import PATH_CAUSTICS, PATH_RAYS, Oracle, Rd3Medium  # noqa: E402
from test_rngdep3d_gpu import workload  # noqa: E402


def main():
    units, topo, launches = workload()
    m = Rd3Medium(Oracle(), *units)
    m.set_topo(*topo)
    out = {}
    for case, amp, write in (("amp18", True, False), ("amp6", False, False), ("paths18", True, True)):
        a, q, s = m.fan3(launches, m.params(amp), bounces=2, write=write,
                         flags=(PATH_RAYS | PATH_CAUSTICS) if write else 0)
        out[case], out[case + "_steps"] = a, np.int64(s)
        if write:
            out[case + "_pts"] = q
        print(case, len(a), "arrivals", s, "ray-steps", len(q), "path rows")
    np.savez_compressed(os.path.join(HERE, "rd3.npz"), **out)


if __name__ == "__main__":
    main()
