"""The oracle's spline, medium, absorption and globe functions against the
reference's own compiled code: golden vectors in tests/golden/ (written by
tests/golden/make_golden.py from oracle/_ref) and, where that build is
present, a live comparison on fresh points.  Bar: bit-exact, since both
sides run the same binary64 operations through the same libm."""
import ctypes as C
import os

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, GEOM_SPH, Medium, Ref
from infraga_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
dp = C.POINTER(C.c_double)


def _units(name):
    prof = synth.base_profile() if name == "base" else synth.perturbed_profile()
    return synth.to_engine_units(*prof)


@pytest.mark.parametrize("pname", ["base", "pert"])
@pytest.mark.parametrize("geom,gname", [(GEOM_SPH, "sph"), (GEOM_3D, "3d")])
def test_oracle_matches_reference_golden(oracle, pname, geom, gname):
    g = np.load(os.path.join(HERE, "golden", f"ref_{pname}_{gname}.npz"))
    m = Medium(oracle, geom, *_units(pname))
    L = oracle.lib
    n = g["tables"].shape[2]
    for w in range(4):
        x, f, s = np.zeros(n), np.zeros(n), np.zeros(n)
        L.go_h_tables(m.h, w, x.ctypes.data_as(dp), f.ctypes.data_as(dp), s.ctypes.data_as(dp))
        assert np.array_equal(x, g["tables"][w, 0])
        assert np.array_equal(f, g["tables"][w, 1])
        assert np.array_equal(s, g["tables"][w, 2]), f"spline slopes differ (field {w})"
    for i, (a, b, c) in enumerate(g["pts"]):
        got = [L.go_c(m.h, a, b, c), L.go_u(m.h, a, b, c), L.go_v(m.h, a, b, c), L.go_rho(m.h, a, b, c)]
        assert got == list(g["med"][i, :4]), (i, got, g["med"][i, :4])
        assert [L.go_du(m.h, a, b, c, k) for k in range(3)] == list(g["med"][i, 7:10])
        assert [L.go_dv(m.h, a, b, c, k) for k in range(3)] == list(g["med"][i, 10:13])
        assert [L.go_dc(m.h, a, b, c, k) for k in range(3)] == list(g["med"][i, 13:16])
        o = [np.zeros(3) for _ in range(3)]
        u, v, w = C.c_double(), C.c_double(), C.c_double()
        L.go_calc_uvw.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), dp, dp, dp, dp, dp, dp]
        L.go_calc_uvw(m.h, a, b, c, C.byref(u), C.byref(v), C.byref(w), *[x.ctypes.data_as(dp) for x in o], None, None, None)
        assert [u.value, v.value, w.value] + list(np.concatenate(o)) == list(g["uvw1"][i]), ("calc_uvw", i)
        o2 = [np.zeros(6) for _ in range(3)]
        L.go_calc_uvw(m.h, a, b, c, C.byref(u), C.byref(v), C.byref(w), *[x.ctypes.data_as(dp) for x in o],
                      *[x.ctypes.data_as(dp) for x in o2])
        assert [u.value, v.value, w.value] + list(np.concatenate(o)) + list(np.concatenate(o2)) == list(g["uvw2"][i]), ("calc_uvw2", i)
        for j, fr in enumerate(g["freqs"]):
            assert L.go_sb_alpha(m.h, a, b, c, fr) == g["sb"][i, j], ("SB_alpha", i, fr)
    if geom == GEOM_SPH:
        for row, ref in zip(g["gc_in"], g["gc_out"]):
            assert L.go_gc_dist(6370.0, *row) == ref


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_sph.so")),
                    reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("geom", [GEOM_SPH, GEOM_3D])
def test_oracle_matches_reference_live(oracle, tmp_path, geom):
    prof = synth.perturbed_profile(seed=99)
    met = str(tmp_path / "p.met")
    synth.write_zcuvd(met, *prof)
    ref = Ref(geom)
    ref.load(met, "zcuvd")
    m = Medium(oracle, geom, *synth.to_engine_units(*prof))
    rng = np.random.default_rng(5)
    for _ in range(300):
        z = rng.uniform(-1, 151)
        p = (6370.0 + z, rng.uniform(-1, 1), rng.uniform(-3, 3)) if geom == GEOM_SPH else (rng.uniform(-9, 9), rng.uniform(-9, 9), z)
        assert oracle.lib.go_c(m.h, *p) == ref.lib.ref_c(*p)
        assert oracle.lib.go_u(m.h, *p) == ref.lib.ref_u(*p)
        assert oracle.lib.go_v(m.h, *p) == ref.lib.ref_v(*p)
        assert oracle.lib.go_sb_alpha(m.h, *p, 0.2) == ref.lib.ref_sb_alpha(*p, 0.2)
