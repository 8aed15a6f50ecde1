"""Known answers of the reference itself, checked against the oracle.

The values are the reference's own printed outputs (setprecision(8)) for
its example profile ToyAtmo.met, recorded from a build of the reference
sources: two single-ray -prop runs and one -eig_search run of infraga-sph
(source 30 N, 100 W on the ground).  The oracle must print the same digits.
"""
import pytest

from ga_oracle_py import GEOM_SPH, Medium

SRC = (30.0, -100.0, 0.0)


def printed(x):
    """The value as the reference prints it (8 significant digits)."""
    return float("%.8g" % x)


@pytest.fixture(scope="module")
def toy(oracle, toyatmo_units):
    return Medium(oracle, GEOM_SPH, *toyatmo_units)


FAN_KATS = [
    # (incl, az), bounce row, printed columns
    ((10.0, -90.0), 0, dict(pos0=29.985815, pos1=-101.96135, time=659.74062)),
    ((4.1392114, -84.970504), 1, dict(pos0=30.249254, pos1=-104.24987, time=1412.4563, tl=-39.828658,
                                      absorption=-0.1609091)),
]


@pytest.mark.parametrize("launch,bounce,cols", FAN_KATS, ids=["incl10_az-90_b0", "incl4.139_az-84.97_b1"])
def test_single_ray_known_answers(toy, launch, bounce, cols):
    p = toy.params(True)
    arr, _ = toy.fan([launch], p, bounces=1, src=SRC)
    row = arr[arr["n_bounce"] == bounce]
    assert len(row) == 1
    for k, v in cols.items():
        assert printed(row[k][0]) == v, (k, row[k][0], v)


def test_eigenray_search_known_answers(toy):
    """infraga-sph -eig_search rcvr 30.25 N 104.25 W, bnc_max 1: three
    eigenrays (one direct, two 1-bounce) at these launch inclinations."""
    p = toy.params(True)
    arr, _ = toy.eig_search(p, SRC, (30.25, -104.25), bnc_min=0, bnc_max=1)
    assert [printed(x) for x in arr["incl_launch"]] == [34.559248, 4.1385204, 30.078055]
    assert list(arr["n_bounce"]) == [0, 1, 1]
    # each arrival lies within the 0.1 km acceptance distance of the receiver
    from ga_oracle_py import Oracle
    import math
    for r in arr:
        d = Oracle().lib.go_gc_dist(6370.0, math.radians(r["pos0"]), math.radians(r["pos1"]),
                                    math.radians(30.25), math.radians(-104.25))
        assert d < 0.1


def test_eigenray_row_function_reproduces_search_rows(oracle, base_units):
    """go_eig_row_3d (the accepted-eigenray trace as a function of the launch
    angles) gives back the rows the oracle's 3D search wrote into the
    committed 512-pair fixture, from their printed launch angles: every
    third row, to 1e-9 relative (TL, absorption to 1e-6 dB)."""
    import os
    import numpy as np
    from ga_oracle_py import GEOM_3D
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "eig3d_base.npz"))["eigenrays"]
    m = Medium(oracle, GEOM_3D, *base_units)
    p = m.params(True)
    for r in g[::3]:
        row = m.eig_row(p, (0.0, 0.0, 0.0), r["incl_launch"], r["az_launch"], int(r["n_bounce"]))
        for f in ("pos0", "pos1", "time", "turn_ht", "incl_arr", "back_az"):
            assert abs(row[f] - r[f]) <= 1e-9 * max(abs(r[f]), 1.0), (int(r["launch_idx"]), f, row[f], r[f])
        for f in ("tl", "absorption"):
            assert abs(row[f] - r[f]) <= 1e-6, (int(r["launch_idx"]), f, row[f], r[f])
