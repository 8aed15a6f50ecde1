"""The oracle's 2D multi-stage spline (topography) and 3D hybrid spline
(range-dependent atmosphere) against the reference's own compiled
src/util/interpolation.cpp (oracle/_ref/libref_interp.so, built from the
reference sources where they are present), and against what the
interpolants must reproduce exactly: linear data, and node values.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _bind(lib, pre):
    lib[pre + "s2d_new"].restype = C.c_void_p
    lib[pre + "s2d_new"].argtypes = [_dp, C.c_int, _dp, C.c_int, _dp]
    lib[pre + "h3d_new"].restype = C.c_void_p
    lib[pre + "h3d_new"].argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
    for n, extra in (("s2d_f", []), ("s2d_df", [C.c_int]), ("s2d_ddf", [C.c_int] * 2), ("s2d_dddf", [C.c_int] * 3)):
        lib[pre + n].restype = C.c_double
        lib[pre + n].argtypes = [C.c_void_p, C.c_double, C.c_double] + extra
    for n, extra in (("h3d_f", []), ("h3d_df", [C.c_int]), ("h3d_ddf", [C.c_int] * 2)):
        lib[pre + n].restype = C.c_double
        lib[pre + n].argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double] + extra
    lib[pre + "s2d_delete"].argtypes = [C.c_void_p]
    lib[pre + "h3d_delete"].argtypes = [C.c_void_p]


class _L:
    def __init__(self, path):
        self.lib = C.CDLL(path)

    def __getitem__(self, k):
        return getattr(self.lib, k)


@pytest.fixture(scope="module")
def orc():
    L = _L(os.path.join(ORACLE, "libga_oracle.so"))
    _bind(L, "go_")
    L["go_h3d_all"].argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
    return L


def _ref():
    path = os.path.join(ORACLE, "_ref", "libref_interp.so")
    if not os.path.exists(path):
        pytest.skip("reference interpolation.cpp not built here (oracle/_ref)")
    L = _L(path)
    _bind(L, "ref_")
    L["ref_h3d_all"].argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp]
    return L


def _grid2d(seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.cumsum(rng.uniform(0.03, 0.07, 33)) + 0.2     # uneven lat nodes [rad-like]
    y = np.linspace(-0.4, 0.5, 29)
    f = rng.normal(0.0, 0.5, (33, 29)).cumsum(axis=0).cumsum(axis=1) * 0.05 + 1.0
    return x, y, np.ascontiguousarray(f)


def _grid3d(seed=4):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.linspace(0.1, 0.5, 6)
    y = np.cumsum(rng.uniform(0.05, 0.12, 5)) - 0.3
    z = np.cumsum(rng.uniform(0.08, 0.3, 48))
    f = (np.sin(3 * x)[:, None, None] + np.cos(2 * y)[None, :, None] + 0.1 * z[None, None, :]
         + 0.05 * rng.normal(size=(6, 5, 48)))
    return x, y, z, np.ascontiguousarray(f)


def _pts2d(x, y, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.stack([rng.uniform(x[0], x[-1], n), rng.uniform(y[0], y[-1], n)], axis=1)


def _pts3d(x, y, z, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.stack([rng.uniform(x[0], x[-1], n), rng.uniform(y[0], y[-1], n), rng.uniform(z[0], z[-1], n)], axis=1)


S2D_CALLS = [("s2d_f", ()), ("s2d_df", (0,)), ("s2d_df", (1,)), ("s2d_ddf", (0, 0)), ("s2d_ddf", (0, 1)),
             ("s2d_ddf", (1, 1)), ("s2d_dddf", (0, 0, 0)), ("s2d_dddf", (0, 0, 1)), ("s2d_dddf", (0, 1, 1)),
             ("s2d_dddf", (1, 1, 1))]
H3D_CALLS = [("h3d_f", ())] + [("h3d_df", (n,)) for n in range(3)] + \
            [("h3d_ddf", (a, b)) for a in range(3) for b in range(3)]


def test_s2d_bit_exact_vs_reference(orc):
    ref = _ref()
    x, y, f = _grid2d()
    ho = orc["go_s2d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(f))
    hr = ref["ref_s2d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(f))
    try:
        for (px, py) in _pts2d(x, y, 60, 1):
            for name, args in S2D_CALLS:
                a = orc["go_" + name](ho, px, py, *args)
                b = ref["ref_" + name](hr, px, py, *args)
                assert a == b, (name, args, px, py, a, b)
    finally:
        orc["go_s2d_delete"](ho)
        ref["ref_s2d_delete"](hr)


def test_h3d_bit_exact_vs_reference(orc):
    ref = _ref()
    x, y, z, f = _grid3d()
    ho = orc["go_h3d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(z), len(z), _ptr(f))
    hr = ref["ref_h3d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(z), len(z), _ptr(f))
    try:
        for (px, py, pz) in _pts3d(x, y, z, 40, 2):
            for name, args in H3D_CALLS:
                a = orc["go_" + name](ho, px, py, pz, *args)
                b = ref["ref_" + name](hr, px, py, pz, *args)
                assert a == b, (name, args, a, b)
            of, odf, oddf = np.zeros(1), np.zeros(3), np.zeros(6)
            orc["go_h3d_all"](ho, px, py, pz, _ptr(of), _ptr(odf), _ptr(oddf))
            rall = np.zeros(10)
            ref["ref_h3d_all"](hr, px, py, pz, _ptr(rall))
            assert np.array_equal(np.concatenate([of, odf, oddf]), rall)
    finally:
        orc["go_h3d_delete"](ho)
        ref["ref_h3d_delete"](hr)


def test_interpolants_reproduce_linear_data(orc):
    """Natural splines and centred differences are exact on linear data."""
    x, y, _ = _grid2d()
    f = np.ascontiguousarray(2.0 + 0.7 * x[:, None] - 1.3 * y[None, :])
    h = orc["go_s2d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(f))
    try:
        for (px, py) in _pts2d(x, y, 20, 5):
            assert abs(orc["go_s2d_f"](h, px, py) - (2.0 + 0.7 * px - 1.3 * py)) < 1e-12
            assert abs(orc["go_s2d_df"](h, px, py, 0) - 0.7) < 1e-10
            assert abs(orc["go_s2d_df"](h, px, py, 1) + 1.3) < 1e-10
            assert abs(orc["go_s2d_ddf"](h, px, py, 0, 1)) < 1e-8
    finally:
        orc["go_s2d_delete"](h)
    x, y, z, _ = _grid3d()
    f = np.ascontiguousarray(1.0 + 0.5 * x[:, None, None] + 0.25 * y[None, :, None] - 0.125 * z[None, None, :]
                             + 0.0 * np.zeros((len(x), len(y), len(z))))
    h = orc["go_h3d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(z), len(z), _ptr(f))
    try:
        for (px, py, pz) in _pts3d(x, y, z, 20, 6):
            fo, df, ddf = np.zeros(1), np.zeros(3), np.zeros(6)
            orc["go_h3d_all"](h, px, py, pz, _ptr(fo), _ptr(df), _ptr(ddf))
            assert abs(fo[0] - (1.0 + 0.5 * px + 0.25 * py - 0.125 * pz)) < 1e-12
            assert np.allclose(df, [0.5, 0.25, -0.125], atol=1e-10)
            assert np.allclose(ddf, 0.0, atol=1e-8)
    finally:
        orc["go_h3d_delete"](h)


def test_interpolants_pass_through_nodes(orc):
    x, y, f = _grid2d()
    h = orc["go_s2d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(f))
    try:
        for i in (0, 5, 17, 32):
            for j in (0, 3, 28):
                assert abs(orc["go_s2d_f"](h, x[i], y[j]) - f[i, j]) < 1e-12
    finally:
        orc["go_s2d_delete"](h)
    x, y, z, f = _grid3d()
    h = orc["go_h3d_new"](_ptr(x), len(x), _ptr(y), len(y), _ptr(z), len(z), _ptr(f))
    try:
        for (i, j, k) in ((0, 0, 0), (2, 3, 10), (5, 4, 47), (3, 1, 30)):
            assert abs(orc["go_h3d_f"](h, x[i], y[j], z[k]) - f[i, j, k]) < 1e-12
    finally:
        orc["go_h3d_delete"](h)
