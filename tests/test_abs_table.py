"""The absorption table (ga_device.hpp abs_fill_record / abs_alpha,
ga_abs_host.hpp abs_layout), compiled for the host from the engine's own
source: over the synthetic profiles, a profile whose nodes put the
composition switch heights inside pieces and a non-uniform profile, the
table's absorption equals direct Sutherland-Bass evaluation (sb_alpha)
within 1e-12 relative at random heights, at the switch heights and just
above them, for both geometries and three frequencies; every piece passes
the build's own accuracy check; z equal to a switch height takes the lower
formula, as the reference's "z > h" tests do."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

SRC = r'''
#define GA_HOST_BUILD 1
#define GA_HD inline
#include "ga_device.hpp"
#include "ga_abs_host.hpp"
using namespace ga;
static void medium(Medium &M, const double *tab, int n) {
    M = Medium{};
    M.tab = tab; M.n = n; M.zlo = tab[0]; M.zhi = tab[(n - 1) * NODE]; M.tweak = 1.0;
    M.sbc[0] = std::pow(10.0, -0.67887); M.sbc[1] = std::pow(10.0, -0.10744); M.sbc[2] = std::pow(10.0, -3.3979);
}
extern "C" int build(const double *tab, int n, int sph, double freq, double *rec, int cap, int *per, double *err) {
    std::vector<double> z(n), geo;
    for (int i = 0; i < n; ++i) z[i] = tab[i * NODE];
    if (!abs_layout(z, geo, *per)) return -1;
    const int nrec = (int)(geo.size() / ABS_GEO);
    if (nrec > cap) return -2;
    Medium M; medium(M, tab, n);
    const AbsBasis B = abs_basis();
    for (int k = 0; k < nrec; ++k) {
        const double *g = &geo[(size_t)k * ABS_GEO];
        if (sph) abs_fill_record<true>(M, B, g[0], g[1], g[2], g[3], g[4], (int)g[5], freq, rec + (size_t)k * ABS_REC, err + k);
        else abs_fill_record<false>(M, B, g[0], g[1], g[2], g[3], g[4], (int)g[5], freq, rec + (size_t)k * ABS_REC, err + k);
    }
    return nrec;
}
extern "C" void eval(const double *tab, int n, int sph, double freq, const double *rec, int per, const double *zs, int m,
                     double *direct, double *table) {
    Medium M; medium(M, tab, n);
    M.abs = rec; M.abs_per = per;
    int cur = 0;
    for (int i = 0; i < m; ++i) {
        const double z = zs[i];
        const SegPt s = seg_point(M, clampz(M, z), cur);
        const double c = herm_f(s, 1), rho = herm_f(s, 7);
        direct[i] = sph ? sb_alpha<true>(z, c, rho, freq, 1.0, M.sbc) : sb_alpha<false>(z, c, rho, freq, 1.0, M.sbc);
        table[i] = abs_alpha(M, s, cur, z, c, rho, freq);
    }
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("abstab")
    (d / "t.cpp").write_text(SRC)
    so = d / "t.so"
    subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", "-I",
                           os.path.join(ROOT, "infraga_amd", "csrc"), "-o", str(so), str(d / "t.cpp")])
    L = C.CDLL(str(so))
    L.build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                       C.c_void_p, C.c_void_p]
    return L


def _tab(units):
    from scipy.interpolate import CubicSpline
    z = units[0]
    T = np.zeros((len(z), 10))
    T[:, 0] = z
    for j, f in ((1, units[1]), (3, units[2]), (5, units[3]), (7, units[4])):
        T[:, j] = f
        T[:, j + 1] = CubicSpline(z, f, bc_type="natural")(z, 1)
    return np.ascontiguousarray(T)


def _profiles():
    from infraga_amd import synth
    base = synth.to_engine_units(*synth.base_profile())
    z, c, u, v, rho = base
    rng = np.random.default_rng(5)
    zn = np.cumsum(np.concatenate([[0.0], rng.uniform(0.05, 0.9, 400)]))
    zn = zn[zn < z[-1]]
    return {"perturbed": synth.to_engine_units(*synth.perturbed_profile()), "base": base,
            "shifted": (z + 0.013, c, u, v, rho),
            "nonuniform": (zn,) + tuple(np.interp(zn, z, f) for f in (c, u, v, rho))}


@pytest.mark.parametrize("name", ["perturbed", "base", "shifted", "nonuniform"])
def test_table_equals_direct(lib, name):
    T = _tab(_profiles()[name])
    n = len(T)
    cap = 400000
    rng = np.random.default_rng(7)
    sw = np.array([30.0, 76.0, 80.0, 90.0, 95.0])
    zs = np.concatenate([rng.uniform(T[0, 0], T[-1, 0], 20000), sw, sw + 1e-12, sw - 1e-12, [T[0, 0], T[-1, 0]]])
    for sph in (1, 0):
        for freq in (0.005, 0.1, 1.0):
            rec, err, per = np.zeros(cap * 22), np.zeros(cap), C.c_int(0)
            nrec = lib.build(T.ctypes.data, n, sph, freq, rec.ctypes.data, cap, C.byref(per), err.ctypes.data)
            assert nrec > 0
            assert err[:nrec].max() < 1e-12, (name, sph, freq, err[:nrec].max())
            d, t = np.zeros(len(zs)), np.zeros(len(zs))
            lib.eval(T.ctypes.data, n, sph, freq, rec.ctypes.data, per.value, zs.ctypes.data, len(zs), d.ctypes.data,
                     t.ctypes.data)
            rel = np.abs(t - d) / np.abs(d)
            assert rel.max() < 1e-12, (name, sph, freq, rel.max(), zs[np.argmax(rel)])


def test_switch_heights_take_the_lower_formula(lib):
    """At z equal to a switch height the table and sb_alpha agree (lower
    formula); the two formulas differ there by far more than the bar, so a
    table that took the upper one fails."""
    T = _tab(_profiles()["base"])
    n = len(T)
    rec, err, per = np.zeros(100000 * 22), np.zeros(100000), C.c_int(0)
    nrec = lib.build(T.ctypes.data, n, 1, 0.1, rec.ctypes.data, 100000, C.byref(per), err.ctypes.data)
    assert nrec > 0
    zs = np.array([30.0, 30.0 + 1e-9])
    d, t = np.zeros(2), np.zeros(2)
    lib.eval(T.ctypes.data, n, 1, 0.1, rec.ctypes.data, per.value, zs.ctypes.data, 2, d.ctypes.data, t.ctypes.data)
    assert abs(d[1] - d[0]) / d[0] > 1e-3          # the formula switches at 30 km
    assert np.allclose(t, d, rtol=1e-12, atol=0)
