"""GPU evaluation of the range-dependent atmosphere (3D hybrid spline) and
of the topography (2D multi-stage spline) against the oracle's restatement
of src/util/interpolation.cpp (itself bit-exact against the reference's
compiled interpolation.cpp, tests/test_oracle_interp.py).

The device topography uses a tensor Hermite form that is the same function
as the reference's per-call natural spline (the natural-spline solve is
linear in the data); agreement is to rounding: 1e-10 of each component's
scale over the sample.
"""
import ctypes as C
import os

import numpy as np
import pytest

import test_oracle_interp as TI

pytestmark = pytest.mark.gpu
_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _rel(got, ref):
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-300)
    return (np.abs(got - ref).max(axis=0) / scale).max()


@pytest.mark.parametrize("even", [True, False])
def test_hybrid_and_topo_match_oracle(engine, even):
    """Evenly spaced axes (segments from the spacing) and uneven ones
    (binary search)."""
    O = TI._L(os.path.join(TI.ORACLE, "libga_oracle.so"))
    TI._bind(O, "go_")
    O["go_h3d_all"].argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
    rng = np.random.Generator(np.random.PCG64(21))
    if even:
        lat = np.linspace(25.0, 35.0, 6)
        lon = np.linspace(-5.0, 5.0, 5)
        z = np.linspace(0.0, 30.0, 121)
    else:
        lat = np.array([25.0, 26.5, 29.0, 31.0, 33.2, 35.0])
        lon = np.array([-5.0, -3.0, 0.5, 1.0, 5.0])
        z = np.concatenate([np.linspace(0.0, 10.0, 81), np.linspace(10.5, 30.0, 40)])
    fields = []
    for k in range(4):
        base = [0.34 - 0.001 * z, 0.02 * np.sin(z / 7.0), 0.01 * np.cos(z / 5.0), 1.2e-3 * np.exp(-z / 7.0)][k]
        pert = 0.02 * np.abs(base).max() * rng.normal(size=(6, 5, 121)).cumsum(axis=2) / 11.0
        fields.append(np.ascontiguousarray(base[None, None, :] + pert))
    rid = engine.rngdep_load(lat, lon, z, *fields)
    tlat = np.linspace(24.0, 36.0, 41)
    tlon = np.linspace(-6.0, 6.0, 37)
    if not even:
        tlat = np.sort(np.concatenate([tlat[::2], tlat[1::2] + 0.07]))
        tlon = np.sort(np.concatenate([tlon[::2], tlon[1::2] - 0.05]))
    zg = np.ascontiguousarray(np.abs(rng.normal(size=(41, 37)).cumsum(axis=0).cumsum(axis=1)) * 0.01)
    engine.topo_load(rid, tlat, tlon, zg)
    try:
        n = 400
        pts = np.stack([np.radians(rng.uniform(24.0, 36.0, n)), np.radians(rng.uniform(-6.0, 6.0, n)),
                        rng.uniform(-1.0, 31.0, n)], axis=1)
        A, T = engine.rngdep_eval(rid, pts)
        x, y = np.ascontiguousarray(np.radians(lat)), np.ascontiguousarray(np.radians(lon))
        for k in range(4):
            h = O["go_h3d_new"](_p(x), len(x), _p(y), len(y), _p(z), len(z), _p(fields[k]))
            ref = np.zeros((n, 10))
            for i, (a, b, c) in enumerate(pts):
                fo, df, ddf = np.zeros(1), np.zeros(3), np.zeros(6)
                O["go_h3d_all"](h, a, b, c, _p(fo), _p(df), _p(ddf))
                ref[i] = np.concatenate([fo, df, ddf])
            O["go_h3d_delete"](h)
            assert _rel(A[:, k, :], ref) < 1e-10, (k, _rel(A[:, k, :], ref))
        tx, ty = np.ascontiguousarray(np.radians(tlat)), np.ascontiguousarray(np.radians(tlon))
        h = O["go_s2d_new"](_p(tx), len(tx), _p(ty), len(ty), _p(zg))
        ref = np.zeros((n, 10))
        for i, (a, b, _) in enumerate(pts):
            a = min(max(a, tx[0]), tx[-1])
            b = min(max(b, ty[0]), ty[-1])
            ref[i] = [O["go_s2d_f"](h, a, b), O["go_s2d_df"](h, a, b, 0), O["go_s2d_df"](h, a, b, 1),
                      O["go_s2d_ddf"](h, a, b, 0, 0), O["go_s2d_ddf"](h, a, b, 0, 1), O["go_s2d_ddf"](h, a, b, 1, 1),
                      O["go_s2d_dddf"](h, a, b, 0, 0, 0), O["go_s2d_dddf"](h, a, b, 0, 0, 1),
                      O["go_s2d_dddf"](h, a, b, 0, 1, 1), O["go_s2d_dddf"](h, a, b, 1, 1, 1)]
        O["go_s2d_delete"](h)
        assert _rel(T, ref) < 1e-10, _rel(T, ref)
    finally:
        engine.rngdep_release(rid)


def test_rngdep_arguments(engine):
    with pytest.raises(Exception):
        engine.rngdep_load([1.0, 0.0], [0.0, 1.0], [0.0, 1.0, 2.0], *[np.zeros((2, 2, 3))] * 4)
    with pytest.raises(Exception):
        engine.rngdep_eval(987654, np.zeros((1, 3)))
