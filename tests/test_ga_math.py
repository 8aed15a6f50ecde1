"""The engine's exp / exp10 (infraga_amd/csrc/ga_math.hpp), compiled for the
host from the same source, against long double references: within 2 ulp
of the double-rounded reference over the argument ranges the medium and
absorption formulas use, exact special values, saturation and NaN."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cmath>
#include "ga_math.hpp"
extern "C" void run(const double *x, int n, double *e, double *e10, long double *re, long double *re10) {
    for (int i = 0; i < n; ++i) {
        e[i] = ga::ga_exp(x[i]);
        e10[i] = ga::ga_exp10(x[i]);
        re[i] = expl((long double)x[i]);
        re10[i] = powl(10.0L, (long double)x[i]);
    }
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("gamath")
    (d / "t.cpp").write_text(SRC)
    so = d / "t.so"
    subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I",
                           os.path.join(ROOT, "infraga_amd", "csrc"), "-o", str(so), str(d / "t.cpp")])
    L = C.CDLL(str(so))
    L.run.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _eval(L, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    e, e10 = np.empty(n), np.empty(n)
    re, re10 = np.empty(n, dtype=np.longdouble), np.empty(n, dtype=np.longdouble)
    L.run(x.ctypes.data, n, e.ctypes.data, e10.ctypes.data, re.ctypes.data, re10.ctypes.data)
    return e, e10, re, re10


def _ulps(got, ref):
    ref_d = ref.astype(np.float64)
    sp = np.spacing(np.abs(ref_d))
    return np.abs(got.astype(np.longdouble) - ref) / sp.astype(np.longdouble)


def test_exp_and_exp10_within_two_ulp(lib):
    rng = np.random.default_rng(11)
    # exp: boundary-layer factors exp(-10 h), set_ds exp(-5 h), Sutherland-Bass
    # exp(a Tr) and exp(-Theta / T); exp10: composition polynomials (-60 .. 3)
    xe = np.concatenate([rng.uniform(-700.0, 700.0, 20000), rng.uniform(-20.0, 20.0, 20000),
                         rng.uniform(-1e-3, 1e-3, 2000), [0.0, -0.0, 1.0, -1.0, math.log(2.0), 0.5]])
    e, _, re, _ = _eval(lib, xe)
    assert _ulps(e, re).max() <= 2.0
    x10 = np.concatenate([rng.uniform(-300.0, 300.0, 20000), rng.uniform(-60.0, 3.0, 20000), [0.0, 1.0, 2.0, -3.0]])
    _, e10, _, re10 = _eval(lib, x10)
    assert _ulps(e10, re10).max() <= 2.0
    e, e10, _, _ = _eval(lib, np.array([0.0, 1.0, 2.0]))
    assert e[0] == 1.0 and e10[0] == 1.0 and e10[1] == 10.0 and e10[2] == 100.0


def test_saturation_and_nan(lib):
    e, e10, _, _ = _eval(lib, np.array([-1500.0, -5000.0, 750.0, 5000.0, np.nan, -np.inf, np.inf]))
    assert e[0] == 0.0 and e[1] == 0.0 and np.isinf(e[2]) and np.isinf(e[3]) and np.isnan(e[4])
    assert e[5] == 0.0 and np.isinf(e[6])
    assert e10[1] == 0.0 and np.isinf(e10[3]) and np.isnan(e10[4])
    # subnormal range: still a fraction of an ulp of the subnormal spacing is not required,
    # only the right order of magnitude
    e, _, re, _ = _eval(lib, np.array([-740.0, -720.0]))
    assert np.allclose(e, re.astype(np.float64), rtol=1e-3, atol=0)
