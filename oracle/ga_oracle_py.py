"""ctypes bindings for the test-only oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

- ``Oracle`` wraps libga_oracle.so (the C restatement in ga_oracle.c).
- ``Ref`` wraps _ref/libref_{sph,3d}.so, the reference's own spline, medium,
  absorption and globe sources compiled from /root/reference (only present
  where that tree was available at build time).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEOM_3D, GEOM_SPH = 0, 1

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Params(C.Structure):
    _fields_ = [("calc_amp", C.c_int), ("eq_cnt", C.c_int),
                ("s_max", C.c_double), ("ds_min", C.c_double), ("ds_max", C.c_double),
                ("alt_max", C.c_double), ("rng_max", C.c_double), ("time_max", C.c_double),
                ("x_min", C.c_double), ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double),
                ("lat_min", C.c_double), ("lat_max", C.c_double), ("lon_min", C.c_double), ("lon_max", C.c_double)]


class Prop(C.Structure):
    _fields_ = [("bounces", C.c_int), ("freq", C.c_double), ("turn_ht_min", C.c_double),
                ("src0", C.c_double), ("src1", C.c_double), ("src_alt", C.c_double)]


class Arrival(C.Structure):
    _fields_ = [("launch_idx", C.c_int), ("n_bounce", C.c_int),
                ("incl_launch", C.c_double), ("az_launch", C.c_double),
                ("pos0", C.c_double), ("pos1", C.c_double),
                ("time", C.c_double), ("cel", C.c_double), ("turn_ht", C.c_double),
                ("incl_arr", C.c_double), ("back_az", C.c_double), ("tl", C.c_double),
                ("absorption", C.c_double)]


ARRIVAL_FIELDS = [f for f, _ in Arrival._fields_]


def arrivals_to_array(arr, n):
    """Structured numpy array from a ctypes Arrival buffer."""
    dt = np.dtype([("launch_idx", "i4"), ("n_bounce", "i4")] + [(f, "f8") for f in ARRIVAL_FIELDS[2:]])
    raw = np.frombuffer(C.string_at(C.addressof(arr), C.sizeof(Arrival) * n), dtype=dt)
    return raw.copy()


class Oracle:
    def __init__(self, path=None):
        path = path or os.path.join(HERE, "libga_oracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        L.go_h_atmo.restype = C.c_void_p
        L.go_h_atmo.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.go_h_free.argtypes = [C.c_void_p]
        L.go_h_set_ground.argtypes = [C.c_void_p, C.c_double]
        L.go_h_set_tweak.argtypes = [C.c_void_p, C.c_double]
        L.go_h_set_r0.argtypes = [C.c_void_p, C.c_double]
        L.go_h_params.argtypes = [C.c_void_p, C.c_int, C.POINTER(Params)]
        L.go_h_tables.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.go_h_fan.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(Prop), _dp, C.c_int,
                               C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
        L.go_h_fan.restype = C.c_int
        for name in ("go_c", "go_rho", "go_u", "go_v"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
            getattr(L, name).restype = C.c_double
        for name in ("go_dc", "go_du", "go_dv"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int]
            getattr(L, name).restype = C.c_double
        L.go_sb_alpha.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
        L.go_sb_alpha.restype = C.c_double
        L.go_gc_dist.argtypes = [C.c_double] * 5
        L.go_gc_dist.restype = C.c_double
        L.go_calc_uvw.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                  _dp, _dp, _dp, _dp, _dp, _dp]
        L.go_spline_all.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        L.go_set_initial.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp]
        L.go_rhs.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp, _dp]
        L.go_rk4_step.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp, _dp]
        L.go_rk4_step.restype = C.c_double
        L.go_eikonal_resid.argtypes = [C.c_void_p, C.c_void_p, _dp]
        L.go_eikonal_resid.restype = C.c_double
        L.go_jacobian.argtypes = [C.c_void_p, _dp]
        L.go_jacobian.restype = C.c_double
        L.go_amp.argtypes = [C.c_void_p, C.c_void_p, _dp]
        L.go_amp.restype = C.c_double
        L.go_travel_time_seg.argtypes = [C.c_void_p, _dp, _dp]
        L.go_travel_time_seg.restype = C.c_double
        L.go_set_refl.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp, _dp, _dp]
        L.go_eig_search_3d.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double,
                                       C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.POINTER(Arrival), C.c_int,
                                       C.POINTER(C.c_longlong)]
        L.go_eig_search_3d.restype = C.c_int
        L.go_eig_search_sph.argtypes = L.go_eig_search_3d.argtypes
        L.go_eig_search_sph.restype = C.c_int
        L.go_eig_counts.argtypes = [C.POINTER(C.c_longlong)]

    EIG_COUNT_NAMES = ("est_calls", "sweep_traces", "sweep_traces_corrected", "sweep_traces_adaptive",
                       "lm_traces", "final_traces", "finds", "found")

    def eig_counts(self):
        """Work counters of the last eigenray search."""
        buf = (C.c_longlong * len(self.EIG_COUNT_NAMES))()
        self.lib.go_eig_counts(buf)
        return dict(zip(self.EIG_COUNT_NAMES, list(buf)))


class Src(C.Structure):
    _fields_ = [("src", C.c_double * 3), ("c0", C.c_double), ("nu0", C.c_double),
                ("theta", C.c_double), ("phi", C.c_double)]


class Medium:
    """One oracle atmosphere (engine units: km, km/s, g/cm^3)."""

    def __init__(self, oracle: Oracle, geom, z, c, u, v, rho):
        self.o = oracle
        self.geom = geom
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (z, c, u, v, rho)]
        self.h = oracle.lib.go_h_atmo(geom, *[_ptr(a) for a in self._keep], len(self._keep[0]))
        if not self.h:
            raise ValueError("oracle rejected the atmosphere")

    def __del__(self):
        try:
            self.o.lib.go_h_free(self.h)
        except Exception:
            pass

    def params(self, calc_amp=True, **kw):
        p = Params()
        self.o.lib.go_h_params(self.h, int(calc_amp), C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def fan(self, launches, params, bounces=2, freq=0.1, turn_ht_min=0.2, src=(30.0, 0.0, 0.0)):
        L = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        pr = Prop(bounces, freq, turn_ht_min, src[0], src[1], src[2])
        cap = max(1, L.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out = C.c_int(0)
        steps = C.c_longlong(0)
        rc = self.o.lib.go_h_fan(self.h, C.byref(params), C.byref(pr), _ptr(L), L.shape[0], out, cap,
                                 C.byref(n_out), C.byref(steps))
        if rc != 0:
            raise RuntimeError("oracle fan overflow")
        return arrivals_to_array(out, n_out.value), steps.value

    def eig_search(self, params, src, rcvr, incl_min=0.5, incl_max=45.0, bnc_min=0, bnc_max=0,
                   iterations=25, damping=1.0e-3, tolerance=0.1, az_dev_lim=2.0, incl_step_max=0.1,
                   incl_step_min=0.001, freq=0.1):
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        cap = 64
        out = (Arrival * cap)()
        steps = C.c_longlong(0)
        fn = self.o.lib.go_eig_search_sph if self.geom == GEOM_SPH else self.o.lib.go_eig_search_3d
        n = fn(self.h, C.byref(params), _ptr(s), _ptr(r), incl_min, incl_max,
               bnc_min, bnc_max, iterations, damping, tolerance, az_dev_lim,
               incl_step_max, incl_step_min, freq, out, cap, C.byref(steps))
        return arrivals_to_array(out, min(n, cap)), steps.value


class Ref:
    """The reference's own compiled medium/spline/absorption sources."""

    def __init__(self, geom):
        name = "libref_sph.so" if geom == GEOM_SPH else "libref_3d.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        self.geom = geom
        L.ref_load.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_load.restype = C.c_int
        L.ref_len.restype = C.c_int
        L.ref_spline_tables.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ref_spline_eval.argtypes = [C.c_int, C.c_double, _dp]
        for name in ("ref_c", "ref_rho", "ref_u", "ref_v"):
            getattr(L, name).argtypes = [C.c_double] * 3
            getattr(L, name).restype = C.c_double
        for name in ("ref_dc", "ref_du", "ref_dv"):
            getattr(L, name).argtypes = [C.c_double] * 3 + [C.c_int]
            getattr(L, name).restype = C.c_double
        L.ref_calc_uvw.argtypes = [C.c_double] * 3 + [C.c_int, _dp]
        L.ref_sb_alpha.argtypes = [C.c_double] * 4
        L.ref_sb_alpha.restype = C.c_double
        for name in ("ref_z_bndlyr", "ref_z0", "ref_alt_max"):
            getattr(L, name).restype = C.c_double
        if geom == GEOM_SPH:
            L.ref_gc_dist.argtypes = [C.c_double] * 4
            L.ref_gc_dist.restype = C.c_double
            L.ref_bearing.argtypes = [C.c_double] * 4
            L.ref_bearing.restype = C.c_double
            L.ref_r0.restype = C.c_double

    def load(self, path, fmt="zcuvd"):
        if self.lib.ref_load(path.encode(), fmt.encode()) != 1:
            raise RuntimeError(f"reference set_region failed on {path}")
        n = self.lib.ref_len()
        return n

    def tables(self, which):
        n = self.lib.ref_len()
        x, f, s = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.ref_spline_tables(which, _ptr(x), _ptr(f), _ptr(s))
        return x, f, s
