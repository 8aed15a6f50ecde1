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


class PathPt(C.Structure):
    _fields_ = [("launch_idx", C.c_int), ("leg", C.c_int), ("step", C.c_int), ("kind", C.c_int),
                ("pos0", C.c_double), ("pos1", C.c_double), ("z", C.c_double),
                ("tl", C.c_double), ("absorption", C.c_double), ("time", C.c_double)]


PATH_DTYPE = np.dtype([("launch_idx", "i4"), ("leg", "i4"), ("step", "i4"), ("kind", "i4"),
                       ("pos0", "f8"), ("pos1", "f8"), ("z", "f8"), ("tl", "f8"), ("absorption", "f8"),
                       ("time", "f8")])
PATH_RAYS, PATH_CAUSTICS = 1, 2


def paths_to_array(buf, n):
    """Structured numpy array from a ctypes PathPt buffer."""
    return np.frombuffer(C.string_at(C.addressof(buf), C.sizeof(PathPt) * n), dtype=PATH_DTYPE).copy()


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
        L.go_h_fan_paths.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(Prop), _dp, C.c_int, C.c_int,
                                     C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(PathPt),
                                     C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.go_h_fan_paths.restype = C.c_int
        L.go_eig_path.argtypes = [C.c_void_p, C.POINTER(Params), _dp, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.POINTER(PathPt), C.c_longlong, C.POINTER(Arrival)]
        L.go_eig_path.restype = C.c_longlong
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
        L.go_eig_find_3d.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Arrival),
                                     C.POINTER(C.c_longlong)]
        L.go_eig_find_3d.restype = C.c_int
        L.go_eig_find_3d_traj.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double, C.c_int,
                                          C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Arrival), _dp,
                                          C.c_int, C.POINTER(C.c_int)]
        L.go_eig_find_3d_traj.restype = C.c_int
        L.go_eig_row_3d.argtypes = [C.c_void_p, C.POINTER(Params), _dp, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.POINTER(Arrival)]
        L.go_eig_row_3d.restype = C.c_int
        L.go_eig_counts.argtypes = [C.POINTER(C.c_longlong)]
        L.go_eig_direct.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Arrival),
                                    C.POINTER(C.c_longlong)]
        L.go_eig_direct.restype = C.c_int
        L.go2_h_new.restype = C.c_void_p
        L.go2_h_new.argtypes = [C.c_void_p]
        L.go2_h_free.argtypes = [C.c_void_p]
        L.go2_h_config.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 6
        L.go2_set_topo.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.go2_set_ground.argtypes = [C.c_void_p, C.c_double]
        L.go2_fan.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                              C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(PathPt), C.c_longlong,
                              C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.go2_fan.restype = C.c_int

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

    def fan_paths(self, launches, params, flags=PATH_RAYS, bounces=2, freq=0.1, turn_ht_min=0.2,
                  src=(30.0, 0.0, 0.0)):
        """-prop with write_rays / write_caustics bookkeeping (go_fan_paths):
        (arrivals, path rows, ray-steps)."""
        L = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        pr = Prop(bounces, freq, turn_ht_min, src[0], src[1], src[2])
        cap = max(1, L.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out = C.c_int(0)
        steps = C.c_longlong(0)
        n_pts = C.c_longlong(0)
        pcap = 1
        while True:
            pts = (PathPt * pcap)()
            rc = self.o.lib.go_h_fan_paths(self.h, C.byref(params), C.byref(pr), _ptr(L), L.shape[0], int(flags),
                                           out, cap, C.byref(n_out), pts, pcap, C.byref(n_pts), C.byref(steps))
            if n_pts.value <= pcap:
                break
            pcap = n_pts.value
        if rc != 0:
            raise RuntimeError("oracle fan overflow")
        return arrivals_to_array(out, n_out.value), paths_to_array(pts, n_pts.value), steps.value

    def eig_path(self, params, src, incl_deg, az_deg, bounces, freq=0.1):
        """The accepted eigenray's path rows and arrival row (go_eig_path)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        A = Arrival()
        n = self.o.lib.go_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg), int(bounces),
                                   freq, None, 0, C.byref(A))
        pts = (PathPt * max(1, n))()
        n = self.o.lib.go_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg), int(bounces),
                                   freq, pts, n, C.byref(A))
        return paths_to_array(pts, n), arrivals_to_array((Arrival * 1)(A), 1)[0]

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


    def eig_direct(self, params, src, rcvr, incl_est, az_est=float("nan"), bounces=0, iterations=25,
                   damping=1.0e-3, tolerance=0.1, freq=0.1):
        """-eig_direct (go_eig_direct): (row or None, ray-steps)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        A = Arrival()
        steps = C.c_longlong(0)
        ok = self.o.lib.go_eig_direct(self.h, C.byref(params), _ptr(s), _ptr(r), float(incl_est), float(az_est),
                                      int(bounces), int(iterations), damping, tolerance, freq, C.byref(A),
                                      C.byref(steps))
        return (arrivals_to_array((Arrival * 1)(A), 1)[0] if ok else None), steps.value

    def eig_find(self, params, src, rcvr, th_rad, ph_rad, bounces, iterations=25, damping=1.0e-3, tolerance=0.1,
                 freq=0.1):
        """3D: find_eigenray alone from (theta, phi) in radians (go_eig_find_3d):
        (row or None, ray-steps)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        A = Arrival()
        steps = C.c_longlong(0)
        ok = self.o.lib.go_eig_find_3d(self.h, C.byref(params), _ptr(s), _ptr(r), float(th_rad), float(ph_rad),
                                       int(bounces), int(iterations), damping, tolerance, freq, C.byref(A),
                                       C.byref(steps))
        return (arrivals_to_array((Arrival * 1)(A), 1)[0] if ok else None), steps.value

    def eig_find_traj(self, params, src, rcvr, th_rad, ph_rad, bounces, iterations=25, damping=1.0e-3,
                      tolerance=0.1, freq=0.1):
        """eig_find that also returns the iterates: (row or None, array of (theta, phi, dr))."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        A = Arrival()
        traj = np.zeros((iterations + 1, 3))
        n = C.c_int(0)
        ok = self.o.lib.go_eig_find_3d_traj(self.h, C.byref(params), _ptr(s), _ptr(r), float(th_rad), float(ph_rad),
                                            int(bounces), int(iterations), damping, tolerance, freq, C.byref(A),
                                            _ptr(traj), traj.shape[0], C.byref(n))
        return (arrivals_to_array((Arrival * 1)(A), 1)[0] if ok else None), traj[: n.value].copy()

    def eig_row(self, params, src, incl_deg, az_deg, bounces, freq=0.1):
        """3D: the row find_eigenray writes for an accepted eigenray with these
        launch angles (go_eig_row_3d)."""
        if self.geom == GEOM_SPH:
            raise NotImplementedError("eig_row: 3D only")
        s = np.ascontiguousarray(src, dtype=np.float64)
        A = Arrival()
        self.o.lib.go_eig_row_3d(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg), int(bounces),
                                 freq, C.byref(A))
        return arrivals_to_array((Arrival * 1)(A), 1)[0]


class Medium2D:
    """The oracle's infraga-2d geometry over a stratified profile (engine
    units).  Ground defaults to the profile bottom, as set_limits does."""

    def __init__(self, oracle: Oracle, z, c, u, v, rho, calc_amp=True, refl_alt=float("nan"),
                 alt_max=float("nan"), rng_max=float("nan"), s_max=float("nan"), ds_min=float("nan"),
                 ds_max=float("nan"), z_grnd=None, topo=None):
        self.o = oracle
        self.m = Medium(oracle, GEOM_3D, z, c, u, v, rho)
        L = oracle.lib
        self.h = L.go2_h_new(self.m.h)
        if topo is not None:
            r, zt = (np.ascontiguousarray(a, dtype=np.float64) for a in topo)
            self._topo = (r, zt)
            L.go2_set_topo(self.h, _ptr(r), _ptr(zt), len(r))
        if z_grnd is not None:
            L.go2_set_ground(self.h, float(z_grnd))
        L.go2_h_config(self.h, int(calc_amp), refl_alt, alt_max, rng_max, s_max, ds_min, ds_max)

    def __del__(self):
        try:
            self.o.lib.go2_h_free(self.h)
        except Exception:
            pass

    def fan(self, launches, bounces=2, src_alt=0.0, freq=0.1, flags=0):
        """(arrivals, path rows, ray-steps); arrivals pos0 = r, pos1 = perpendicular drift."""
        Lr = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        cap = max(1, Lr.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out = C.c_int(0)
        steps = C.c_longlong(0)
        n_pts = C.c_longlong(0)
        pcap = 1
        while True:
            pts = (PathPt * pcap)()
            rc = self.o.lib.go2_fan(self.h, _ptr(Lr), Lr.shape[0], bounces, src_alt, freq, int(flags), out, cap,
                                    C.byref(n_out), pts, pcap, C.byref(n_pts), C.byref(steps))
            if n_pts.value <= pcap and n_out.value <= cap:
                break
            pcap = max(pcap, n_pts.value)
            if n_out.value > cap:
                cap = n_out.value
                out = (Arrival * cap)()
        return arrivals_to_array(out, n_out.value), paths_to_array(pts, n_pts.value), steps.value


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
        L.ref_limits.argtypes = [_dp]
        L.ref_load_topo.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.ref_load_topo.restype = C.c_int
        L.ref_use_blw.argtypes = [C.c_int]
        L.ref_z_max.restype = C.c_double
        L.ref_topo.argtypes = [C.c_double, C.c_double, _dp]
        L.ref_w.argtypes = [C.c_double] * 3
        L.ref_w.restype = C.c_double

    def load_topo(self, path, topo_path, fmt="zcuvd"):
        """set_region(atmo, topo, format): the stratified profile over terrain."""
        if self.lib.ref_load_topo(path.encode(), topo_path.encode(), fmt.encode()) != 1:
            raise RuntimeError(f"reference set_region failed on {path}, {topo_path}")
        return self.lib.ref_len()

    def limits(self):
        out = np.zeros(8)
        self.lib.ref_limits(_ptr(out))
        return out

    def topo(self, a, b):
        out = np.zeros(6)
        self.lib.ref_topo(a, b, _ptr(out))
        return out

    def calc_uvw(self, a, b, c, second):
        out = np.zeros(30 if second else 12)
        self.lib.ref_calc_uvw(a, b, c, int(second), _ptr(out))
        return out

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


# ---------------------------------------------------------------------------
# Range-dependent spherical medium (ga_oracle_rd.c) and the reference's own
# compiled range-dependent atmosphere (_ref/libref_sphrd.so).
# ---------------------------------------------------------------------------
def _rd_bind(L):
    if getattr(L, "_rd_bound", False):
        return
    L.go_rd_new.restype = C.c_void_p
    L.go_rd_new.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]
    L.go_rd_delete.argtypes = [C.c_void_p]
    L.go_rd_set_topo.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp]
    L.go_rd_set_topo.restype = C.c_int
    L.go_rd_set_ground.argtypes = [C.c_void_p, C.c_double]
    L.go_rd_h_config.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.go_rd_h_limits.argtypes = [C.c_void_p, _dp]
    L.go_rd_h_set_strat.argtypes = [C.c_void_p, C.c_int]
    L.go_rd_h_set_upper.argtypes = [C.c_void_p, C.c_int]
    L.go_rd_params.argtypes = [C.c_void_p, C.c_int, C.POINTER(Params)]
    L.go_rd_centre.argtypes = [C.c_void_p, _dp, _dp]
    for name in ("go_rd_c", "go_rd_rho", "go_rd_u", "go_rd_v", "go_rd_w", "go_rd_topo_z", "go_rd_bl_sc"):
        f = getattr(L, name)
        f.restype = C.c_double
        f.argtypes = [C.c_void_p] + [C.c_double] * (2 if name == "go_rd_topo_z" else 3)
    for name in ("go_rd_dc", "go_rd_du", "go_rd_dv", "go_rd_dw", "go_rd_bl_dsc"):
        f = getattr(L, name)
        f.restype = C.c_double
        f.argtypes = [C.c_void_p] + [C.c_double] * 3 + [C.c_int]
    L.go_rd_bl_ddsc.restype = C.c_double
    L.go_rd_bl_ddsc.argtypes = [C.c_void_p] + [C.c_double] * 3 + [C.c_int, C.c_int]
    L.go_rd_topo_dz.restype = C.c_double
    L.go_rd_topo_dz.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
    L.go_rd_topo_ddz.restype = C.c_double
    L.go_rd_topo_ddz.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int]
    L.go_rd_topo_dddz.restype = C.c_double
    L.go_rd_topo_dddz.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
    L.go_rd_calc_uvw.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                 _dp, _dp, _dp, _dp, _dp, _dp]
    L.go_rd_sb_alpha.restype = C.c_double
    L.go_rd_sb_alpha.argtypes = [C.c_void_p] + [C.c_double] * 4
    L.go_rd_set_initial.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp]
    L.go_rd_rhs.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp]
    L.go_rd_rk4_step.restype = C.c_double
    L.go_rd_rk4_step.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp]
    L.go_rd_set_refl.argtypes = [C.c_void_p, C.POINTER(Params), C.c_void_p, _dp, _dp, _dp]
    for name in ("go_rd_jacobian",):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [C.c_void_p, _dp]
    for name in ("go_rd_amp", "go_rd_eikonal_resid"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, _dp]
    L.go_rd_eikonal_deriv_resid.restype = C.c_double
    L.go_rd_eikonal_deriv_resid.argtypes = [C.c_void_p, _dp]
    L.go_rd_travel_time_seg.restype = C.c_double
    L.go_rd_travel_time_seg.argtypes = [C.c_void_p, _dp, _dp]
    L.go_rd_atten_seg.restype = C.c_double
    L.go_rd_atten_seg.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    L.go_rd_theta_grnd.restype = C.c_double
    L.go_rd_theta_grnd.argtypes = [C.c_void_p] + [C.c_double] * 4
    L.go_rd_fan.restype = C.c_int
    L.go_rd_fan.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(Prop), _dp, C.c_int,
                            C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.go_rd_fan_paths.restype = C.c_int
    L.go_rd_eig_search.restype = C.c_int
    L.go_rd_eig_search.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.POINTER(Arrival), C.c_int, C.POINTER(C.c_longlong)]
    L.go_rd_eig_theta_grnd.restype = C.c_double
    L.go_rd_eig_theta_grnd.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
    L.go_rd_eig_path.restype = C.c_longlong
    L.go_rd_eig_path.argtypes = [C.c_void_p, C.POINTER(Params), _dp, C.c_double, C.c_double, C.c_int, C.c_double,
                                 C.POINTER(PathPt), C.c_longlong, C.POINTER(Arrival)]
    L.go_rd_fan_paths.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(Prop), _dp, C.c_int, C.c_int,
                                  C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(PathPt), C.c_longlong,
                                  C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.go_rd_new_xy.restype = C.c_void_p
    L.go_rd_new_xy.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp]
    L.go_rd3_fan.restype = C.c_int
    L.go_rd3_fan.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(Prop), _dp, C.c_int, C.c_int, C.c_int,
                             C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(PathPt), C.c_longlong,
                             C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.go_rd3_rhs.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp]
    L.go_rd3_eig_search.restype = C.c_int
    L.go_rd3_eig_search.argtypes = L.go_rd_eig_search.argtypes
    L.go_rd3_eig_direct.restype = C.c_int
    L.go_rd3_eig_direct.argtypes = [C.c_void_p, C.POINTER(Params), _dp, _dp, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(Arrival),
                                    C.POINTER(C.c_longlong)]
    L.go_rd3_eig_path.restype = C.c_longlong
    L.go_rd3_eig_path.argtypes = L.go_rd_eig_path.argtypes
    L.go_rd3_lm_jac.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.go_rd3_trace_leg.restype = C.c_int
    L.go_rd3_trace_leg.argtypes = [C.c_void_p, C.POINTER(Params), _dp, C.c_double, C.c_double, _dp, _dp]
    L._rd_bound = True


class RdMedium:
    """One range-dependent spherical atmosphere of the oracle.

    Fields are [nlat, nlon, nz] arrays in km/s (c, u, v) and g/cm^3 (rho);
    node coordinates in degrees, z in km."""

    def __init__(self, oracle: Oracle, lat_deg, lon_deg, z, c, u, v, rho):
        self.o = oracle
        L = oracle.lib
        _rd_bind(L)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (lat_deg, lon_deg, z, c, u, v, rho)]
        la, lo, zz, cc, uu, vv, rr = self._keep
        nlat, nlon, nz = la.size, lo.size, zz.size
        for f in (cc, uu, vv, rr):
            if f.size != nlat * nlon * nz:
                raise ValueError("field shape must be [nlat, nlon, nz]")
        self.h = L.go_rd_new(_ptr(la), nlat, _ptr(lo), nlon, _ptr(zz), nz, _ptr(cc), _ptr(uu), _ptr(vv), _ptr(rr))
        if not self.h:
            raise ValueError("oracle rejected the range-dependent atmosphere")

    def __del__(self):
        try:
            self.o.lib.go_rd_delete(self.h)
        except Exception:
            pass

    def set_topo(self, lat_deg, lon_deg, zg):
        la = np.ascontiguousarray(lat_deg, dtype=np.float64)
        lo = np.ascontiguousarray(lon_deg, dtype=np.float64)
        zz = np.ascontiguousarray(zg, dtype=np.float64)
        self._topo_keep = (la, lo, zz)
        if self.o.lib.go_rd_set_topo(self.h, _ptr(la), la.size, _ptr(lo), lo.size, _ptr(zz)) != 0:
            raise ValueError("oracle rejected the topography")

    def set_ground(self, z0):
        self.o.lib.go_rd_set_ground(self.h, z0)

    def config(self, use_blw=False, tweak_abs=1.0):
        self.o.lib.go_rd_h_config(self.h, int(use_blw), tweak_abs)

    def set_strat(self, on=True):
        """The stratified binaries' semantics over identical columns
        (go_rd_h_set_strat)."""
        self.o.lib.go_rd_h_set_strat(self.h, int(on))

    def set_upper(self, on=True):
        """Only the stratified 3D set's upper-triangle contraction (go_rd_h_set_upper)."""
        self.o.lib.go_rd_h_set_upper(self.h, int(on))

    @classmethod
    def strat_over_terrain(cls, oracle: "Oracle", z, c, u, v, rho, tlat, tlon, zg, n=4):
        """infraga-sph with topo_file: the stratified profile (engine units)
        as n x n identical columns spanning the terrain grid's extent, the
        terrain, and the stratified semantics."""
        tlat = np.asarray(tlat, dtype=np.float64)
        tlon = np.asarray(tlon, dtype=np.float64)
        lat = np.linspace(tlat[0], tlat[-1], n)
        lon = np.linspace(tlon[0], tlon[-1], n)
        col = [np.broadcast_to(np.asarray(f, dtype=np.float64), (n, n, len(z))).copy() for f in (c, u, v, rho)]
        m = cls(oracle, lat, lon, z, *col)
        m.set_strat(True)
        m.set_topo(tlat, tlon, zg)
        return m

    def limits(self):
        out = np.zeros(8)
        self.o.lib.go_rd_h_limits(self.h, _ptr(out))
        return out

    def centre(self):
        a, b = C.c_double(), C.c_double()
        self.o.lib.go_rd_centre(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def params(self, calc_amp=True, **kw):
        p = Params()
        self.o.lib.go_rd_params(self.h, int(calc_amp), C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def point(self, r, t, p):
        L, h = self.o.lib, self.h
        out = [L.go_rd_c(h, r, t, p), L.go_rd_rho(h, r, t, p), L.go_rd_u(h, r, t, p),
               L.go_rd_v(h, r, t, p), L.go_rd_w(h, r, t, p)]
        for fn in (L.go_rd_dc, L.go_rd_du, L.go_rd_dv, L.go_rd_dw):
            out += [fn(h, r, t, p, n) for n in range(3)]
        return np.array(out)

    def calc_uvw(self, r, t, p, second):
        u, v, w = C.c_double(), C.c_double(), C.c_double()
        d = [np.zeros(3) for _ in range(3)]
        dd = [np.zeros(6) for _ in range(3)]
        ddp = [_ptr(x) for x in dd] if second else [None, None, None]
        self.o.lib.go_rd_calc_uvw(self.h, r, t, p, C.byref(u), C.byref(v), C.byref(w),
                                  _ptr(d[0]), _ptr(d[1]), _ptr(d[2]), *ddp)
        out = [u.value, v.value, w.value] + list(d[0]) + list(d[1]) + list(d[2])
        if second:
            out += list(dd[0]) + list(dd[1]) + list(dd[2])
        return np.array(out)

    def topo(self, lat, lon):
        L, h = self.o.lib, self.h
        return np.array([L.go_rd_topo_z(h, lat, lon), L.go_rd_topo_dz(h, lat, lon, 0), L.go_rd_topo_dz(h, lat, lon, 1),
                         L.go_rd_topo_ddz(h, lat, lon, 0, 0), L.go_rd_topo_ddz(h, lat, lon, 0, 1),
                         L.go_rd_topo_ddz(h, lat, lon, 1, 1), L.go_rd_topo_dddz(h, lat, lon, 0, 0, 0),
                         L.go_rd_topo_dddz(h, lat, lon, 0, 0, 1), L.go_rd_topo_dddz(h, lat, lon, 0, 1, 1),
                         L.go_rd_topo_dddz(h, lat, lon, 1, 1, 1)])

    def bl(self, r, t, p):
        L, h = self.o.lib, self.h
        out = [L.go_rd_bl_sc(h, r, t, p)] + [L.go_rd_bl_dsc(h, r, t, p, n) for n in range(3)]
        out += [L.go_rd_bl_ddsc(h, r, t, p, n, m) for n in range(3) for m in range(n, 3)]
        return np.array(out)

    def sb_alpha(self, r, t, p, freq):
        return self.o.lib.go_rd_sb_alpha(self.h, r, t, p, freq)

    def rhs(self, y, params):
        y = np.ascontiguousarray(y, dtype=np.float64)
        f = np.zeros(18)
        self.o.lib.go_rd_rhs(self.h, C.byref(params), _ptr(y), _ptr(f))
        return f[:params.eq_cnt]

    def rk4_step(self, y, params):
        y = np.zeros(18) + np.pad(np.asarray(y, dtype=np.float64), (0, 18 - len(y)))
        y1 = np.zeros(18)
        ds = self.o.lib.go_rd_rk4_step(self.h, C.byref(params), _ptr(y), _ptr(y1))
        return y1[:params.eq_cnt], ds

    def set_initial(self, params, theta, phi, src_r, src_lat, src_lon):
        s = Src((C.c_double * 3)(src_r, src_lat, src_lon), 0.0, 0.0, theta, phi)
        y0 = np.zeros(18)
        self.o.lib.go_rd_set_initial(self.h, C.byref(params), C.byref(s), _ptr(y0))
        return y0[:params.eq_cnt], s

    def set_refl(self, params, src, rk2, rk1):
        a = np.ascontiguousarray(np.pad(rk2, (0, 18 - len(rk2))), dtype=np.float64)
        b = np.ascontiguousarray(np.pad(rk1, (0, 18 - len(rk1))), dtype=np.float64)
        y0 = np.zeros(18)
        self.o.lib.go_rd_set_refl(self.h, C.byref(params), C.byref(src), _ptr(a), _ptr(b), _ptr(y0))
        return y0[:params.eq_cnt]

    def eikonal_resid(self, src, y):
        y = np.ascontiguousarray(np.pad(y, (0, 18 - len(y))), dtype=np.float64)
        return self.o.lib.go_rd_eikonal_resid(self.h, C.byref(src), _ptr(y))

    def amp(self, src, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        return self.o.lib.go_rd_amp(self.h, C.byref(src), _ptr(y))

    def aux_resid(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        return self.o.lib.go_rd_eikonal_deriv_resid(self.h, _ptr(y))

    def trace(self, params, src, nsteps, y0):
        """nsteps RK4 steps from y0 (no checks); returns rows [nsteps+1, neq]."""
        rows = [np.asarray(y0, dtype=np.float64)]
        for _ in range(nsteps):
            y1, _ = self.rk4_step(rows[-1], params)
            rows.append(y1)
        return np.array(rows)

    def theta_grnd(self, lat_src, lon_src, z_src, az_deg):
        return self.o.lib.go_rd_theta_grnd(self.h, lat_src, lon_src, z_src, az_deg)

    def fan(self, launches, params, bounces=2, freq=0.1, turn_ht_min=0.2, src=None):
        """src = (lat_deg, lon_deg, alt_km); None = the region centre at the ground."""
        if src is None:
            la, lo = self.centre()
            src = (la, lo, 0.0)
        Lr = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        pr = Prop(bounces, freq, turn_ht_min, src[0], src[1], src[2])
        cap = max(1, Lr.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out = C.c_int(0)
        steps = C.c_longlong(0)
        rc = self.o.lib.go_rd_fan(self.h, C.byref(params), C.byref(pr), _ptr(Lr), Lr.shape[0], out, cap,
                                  C.byref(n_out), C.byref(steps))
        if rc != 0:
            raise RuntimeError("oracle fan overflow")
        return arrivals_to_array(out, n_out.value), steps.value

    def eig_search(self, params, src, rcvr, incl_min=0.5, incl_max=45.0, bnc_min=0, bnc_max=0, iterations=25,
                   damping=1.0e-3, tolerance=0.1, az_dev_lim=2.0, incl_step_max=0.1, incl_step_min=0.001, freq=0.1):
        """infraga-sph-rngdep -eig_search (go_rd_eig_search): (eigenrays, ray-steps)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        cap = 64
        out = (Arrival * cap)()
        steps = C.c_longlong(0)
        n = self.o.lib.go_rd_eig_search(self.h, C.byref(params), _ptr(s), _ptr(r), incl_min, incl_max, bnc_min,
                                        bnc_max, iterations, damping, tolerance, az_dev_lim, incl_step_max,
                                        incl_step_min, freq, out, cap, C.byref(steps))
        return arrivals_to_array(out, min(n, cap)), steps.value

    def eig_path(self, params, src, incl_deg, az_deg, bounces, freq=0.1):
        """The accepted eigenray's path rows and row (go_rd_eig_path)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        A = Arrival()
        n = self.o.lib.go_rd_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg), int(bounces),
                                      freq, None, 0, C.byref(A))
        pts = (PathPt * max(1, n))()
        n = self.o.lib.go_rd_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg), int(bounces),
                                      freq, pts, n, C.byref(A))
        return paths_to_array(pts, n), arrivals_to_array((Arrival * 1)(A), 1)[0]

    def fan_paths(self, launches, params, flags=PATH_RAYS, bounces=2, freq=0.1, turn_ht_min=0.2, src=None):
        """write_rays / write_caustics run_prop (go_rd_fan_paths): (arrivals, path rows, ray-steps)."""
        if src is None:
            la, lo = self.centre()
            src = (la, lo, 0.0)
        Lr = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        pr = Prop(bounces, freq, turn_ht_min, src[0], src[1], src[2])
        cap = max(1, Lr.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out = C.c_int(0)
        steps = C.c_longlong(0)
        n_pts = C.c_longlong(0)
        pcap = 1
        while True:
            pts = (PathPt * pcap)()
            rc = self.o.lib.go_rd_fan_paths(self.h, C.byref(params), C.byref(pr), _ptr(Lr), Lr.shape[0], int(flags),
                                            out, cap, C.byref(n_out), pts, pcap, C.byref(n_pts), C.byref(steps))
            if n_pts.value <= pcap:
                break
            pcap = n_pts.value
        if rc != 0:
            raise RuntimeError("oracle fan overflow")
        return arrivals_to_array(out, n_out.value), paths_to_array(pts, n_pts.value), steps.value


def write_rd_files(d, lat_deg, lon_deg, z, c_ms, u_ms, v_ms, rho, topo=None, prefix="prof"):
    """Write a range-dependent grid in the reference's input layout: node
    files (one degree value per line), prefix{i*nlon+j}.met in format zcuvd
    (z [km], c, u, v [m/s], rho [g/cm^3]) and optionally a topography file of
    "lat lon z" rows, longitude fastest.  Values are written with 17
    significant digits so the text parses back to the same doubles.
    Returns (prefix_path, lat_path, lon_path, topo_path or None)."""
    os.makedirs(d, exist_ok=True)
    lat_p, lon_p = os.path.join(d, "lats.dat"), os.path.join(d, "lons.dat")
    with open(lat_p, "w") as f:
        f.write("".join("%.17g\n" % x for x in lat_deg))
    with open(lon_p, "w") as f:
        f.write("".join("%.17g\n" % x for x in lon_deg))
    pre = os.path.join(d, prefix)
    nlat, nlon = len(lat_deg), len(lon_deg)
    for i in range(nlat):
        for j in range(nlon):
            rows = np.column_stack([z, c_ms[i, j], u_ms[i, j], v_ms[i, j], rho[i, j]])
            with open("%s%d.met" % (pre, i * nlon + j), "w") as f:
                f.write("".join(" ".join("%.17g" % x for x in r) + "\n" for r in rows))
    topo_p = None
    if topo is not None:
        tlat, tlon, tz = topo
        topo_p = os.path.join(d, "topo.dat")
        with open(topo_p, "w") as f:
            for i, la in enumerate(tlat):
                f.write("".join("%.17g %.17g %.17g\n" % (la, lo, tz[i, j]) for j, lo in enumerate(tlon)))
    return pre, lat_p, lon_p, topo_p


class RefRd:
    """The reference's own compiled range-dependent spherical atmosphere."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libref_sphrd.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.ref_rd_load.argtypes = [C.c_char_p] * 5
        L.ref_rd_load.restype = C.c_int
        L.ref_rd_use_blw.argtypes = [C.c_int]
        L.ref_rd_set_tweak.argtypes = [C.c_double]
        L.ref_rd_limits.argtypes = [_dp]
        L.ref_rd_point.argtypes = [C.c_double] * 3 + [_dp]
        L.ref_rd_calc_uvw.argtypes = [C.c_double] * 3 + [C.c_int, _dp]
        L.ref_rd_topo.argtypes = [C.c_double] * 2 + [_dp]
        L.ref_rd_bl.argtypes = [C.c_double] * 3 + [_dp]
        L.ref_rd_sb_alpha.argtypes = [C.c_double] * 4
        L.ref_rd_sb_alpha.restype = C.c_double
        self.loaded = False

    def load(self, prefix, lats, lons, topo=None, fmt="zcuvd"):
        if self.loaded:
            self.lib.ref_rd_clear()
        rc = self.lib.ref_rd_load(prefix.encode(), lats.encode(), lons.encode(),
                                  (topo or "").encode(), fmt.encode())
        if rc != 1:
            raise RuntimeError("reference set_region failed")
        self.loaded = True

    def use_blw(self, on):
        self.lib.ref_rd_use_blw(int(on))

    def limits(self):
        out = np.zeros(8)
        self.lib.ref_rd_limits(_ptr(out))
        return out

    def point(self, r, t, p):
        out = np.zeros(17)
        self.lib.ref_rd_point(r, t, p, _ptr(out))
        return out

    def calc_uvw(self, r, t, p, second):
        out = np.zeros(30)
        self.lib.ref_rd_calc_uvw(r, t, p, int(second), _ptr(out))
        return out[:30 if second else 12]

    def topo(self, lat, lon):
        out = np.zeros(10)
        self.lib.ref_rd_topo(lat, lon, _ptr(out))
        return out

    def bl(self, r, t, p):
        out = np.zeros(10)
        self.lib.ref_rd_bl(r, t, p, _ptr(out))
        return out

    def sb_alpha(self, r, t, p, freq):
        return self.lib.ref_rd_sb_alpha(r, t, p, freq)


class Rd3Medium(RdMedium):
    """One 3D Cartesian range-dependent atmosphere of the oracle
    (infraga-3d-rngdep): node coordinates x, y in km, fields [nx, ny, nz];
    a topography set on it takes x, y in km."""

    def __init__(self, oracle: Oracle, x_km, y_km, z, c, u, v, rho):
        self.o = oracle
        L = oracle.lib
        _rd_bind(L)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x_km, y_km, z, c, u, v, rho)]
        xx, yy, zz, cc, uu, vv, rr = self._keep
        if any(f.size != xx.size * yy.size * zz.size for f in (cc, uu, vv, rr)):
            raise ValueError("field shape must be [nx, ny, nz]")
        self.h = L.go_rd_new_xy(_ptr(xx), xx.size, _ptr(yy), yy.size, _ptr(zz), zz.size, _ptr(cc), _ptr(uu), _ptr(vv),
                                _ptr(rr))
        if not self.h:
            raise ValueError("oracle rejected the range-dependent atmosphere")

    def rhs(self, params, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        f = np.zeros(18)
        self.o.lib.go_rd3_rhs(self.h, C.byref(params), _ptr(y), _ptr(f))
        return f[:params.eq_cnt]

    def fan3(self, launches, params, bounces=2, write=False, flags=0, freq=0.1, turn_ht_min=0.2,
             src=(float("nan"), float("nan"), 0.0)):
        """run_prop (go_rd3_fan): (arrivals, path rows, ray-steps)."""
        Lr = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        pr = Prop(bounces, freq, turn_ht_min, src[0], src[1], src[2])
        cap = max(1, Lr.shape[0] * (bounces + 1))
        out = (Arrival * cap)()
        n_out, steps, n_pts = C.c_int(0), C.c_longlong(0), C.c_longlong(0)
        pcap = 1
        while True:
            pts = (PathPt * pcap)()
            self.o.lib.go_rd3_fan(self.h, C.byref(params), C.byref(pr), _ptr(Lr), Lr.shape[0], int(write), int(flags),
                                  out, cap, C.byref(n_out), pts, pcap, C.byref(n_pts), C.byref(steps))
            if n_pts.value <= pcap:
                break
            pcap = n_pts.value
        return arrivals_to_array(out, n_out.value), paths_to_array(pts, n_pts.value), steps.value

    def eig_search(self, params, src, rcvr, incl_min=0.5, incl_max=45.0, bnc_min=0, bnc_max=0, iterations=25,
                   damping=1.0e-3, tolerance=0.1, az_dev_lim=2.0, incl_step_max=0.1, incl_step_min=0.001, freq=0.1):
        """infraga-3d-rngdep -eig_search (go_rd3_eig_search): (eigenrays, ray-steps); src = (x, y, z) km."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        cap = 64
        out = (Arrival * cap)()
        steps = C.c_longlong(0)
        n = self.o.lib.go_rd3_eig_search(self.h, C.byref(params), _ptr(s), _ptr(r), incl_min, incl_max, bnc_min,
                                         bnc_max, iterations, damping, tolerance, az_dev_lim, incl_step_max,
                                         incl_step_min, freq, out, cap, C.byref(steps))
        return arrivals_to_array(out, min(n, cap)), steps.value

    def eig_direct(self, params, src, rcvr, incl_est, az_est=float("nan"), bounces=0, iterations=25,
                   damping=1.0e-3, tolerance=0.1, freq=0.1):
        """infraga-3d-rngdep -eig_direct (go_rd3_eig_direct): (row or None, ray-steps)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        r = np.ascontiguousarray(rcvr, dtype=np.float64)
        A = Arrival()
        steps = C.c_longlong(0)
        ok = self.o.lib.go_rd3_eig_direct(self.h, C.byref(params), _ptr(s), _ptr(r), float(incl_est), float(az_est),
                                          int(bounces), int(iterations), damping, tolerance, freq, C.byref(A),
                                          C.byref(steps))
        row = arrivals_to_array((Arrival * 1)(A), 1)[0] if ok else None
        return row, steps.value

    def eig_path(self, params, src, incl_deg, az_deg, bounces, freq=0.1):
        """The accepted eigenray's path rows and row (go_rd3_eig_path)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        A = Arrival()
        n = self.o.lib.go_rd3_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg),
                                       int(bounces), freq, None, 0, C.byref(A))
        pts = (PathPt * max(1, n))()
        n = self.o.lib.go_rd3_eig_path(self.h, C.byref(params), _ptr(s), float(incl_deg), float(az_deg),
                                       int(bounces), freq, pts, n, C.byref(A))
        return paths_to_array(pts, n), arrivals_to_array((Arrival * 1)(A), 1)[0]

    def lm_jac(self, src, yk):
        """find_eigenray's d(x, y)/d(theta, phi) at row K: (dx/dth, dy/dth, dx/dph, dy/dph)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        y = np.ascontiguousarray(yk, dtype=np.float64)
        J = np.zeros(4)
        self.o.lib.go_rd3_lm_jac(self.h, _ptr(s), _ptr(y), _ptr(J))
        return J

    def trace_leg(self, params, src, th, ph):
        """Rows k-1 and k of the first leg (18 equations) from (th, ph) [rad]; None on a break."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        a, b = np.zeros(18), np.zeros(18)
        k = self.o.lib.go_rd3_trace_leg(self.h, C.byref(params), _ptr(s), float(th), float(ph), _ptr(a), _ptr(b))
        return None if k < 0 else (a, b)


# ---- weakly non-linear waveforms (ga_wvfrm.c) --------------------------------
class WvParams(C.Structure):
    """go_wv_params: length, geoac::ds_wvfrm and the wvfrm:: filter settings
    (reference defaults: 2^13 points, ds 0.5, filters (0.5, 4, 0.5) and (0.75, 2, 1))."""
    _fields_ = [("len", C.c_int), ("ds", C.c_double), ("filt1_g", C.c_double), ("filt1_n1", C.c_double),
                ("filt1_n2", C.c_double), ("filt2_g", C.c_double), ("filt2_n1", C.c_double),
                ("filt2_n2", C.c_double)]

    @classmethod
    def default(cls, **kw):
        w = cls(8192, 0.5, 0.5, 4.0, 0.5, 0.75, 2.0, 1.0)
        for k, v in kw.items():
            setattr(w, k, v)
        return w


def _wv_bind(L):
    if getattr(L, "_wv_bound", False):
        return
    L.go_fft.argtypes = [_dp, _dp, C.c_int, C.c_int]
    L.go_fft.restype = C.c_int
    L.go_rfft.argtypes = [_dp, C.c_int, _dp, _dp]
    L.go_irfft.argtypes = [_dp, _dp, C.c_int, _dp]
    L.go_wvfrm_build.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int,
                                 _dp, _dp]
    L.go_h_wnl_run.restype = C.c_longlong
    L.go_h_wnl_run.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(WvParams), _dp, C.c_double, C.c_double,
                               C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               C.POINTER(C.c_int)]
    L.go_h_rd_wnl_run.restype = C.c_longlong
    L.go_h_rd_wnl_run.argtypes = L.go_h_wnl_run.argtypes
    L.go2_h_wnl_run.restype = C.c_longlong
    L.go2_h_wnl_run.argtypes = [C.c_void_p, C.POINTER(WvParams), C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]
    L._wv_bound = True


def wv_fft(oracle, x, inverse=False):
    """go_fft of a complex series (length a power of two)."""
    _wv_bind(oracle.lib)
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    if oracle.lib.go_fft(_ptr(re), _ptr(im), re.size, int(inverse)) != 0:
        raise ValueError("length must be a power of two")
    return re + 1j * im


def wv_rfft(oracle, x):
    _wv_bind(oracle.lib)
    x = np.ascontiguousarray(x, dtype=np.float64)
    re, im = np.zeros(x.size // 2 + 1), np.zeros(x.size // 2 + 1)
    oracle.lib.go_rfft(_ptr(x), x.size, _ptr(re), _ptr(im))
    return re + 1j * im


def wv_irfft(oracle, X, n):
    _wv_bind(oracle.lib)
    re = np.ascontiguousarray(np.real(X), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(X), dtype=np.float64)
    out = np.zeros(n)
    oracle.lib.go_irfft(_ptr(re), _ptr(im), n, _ptr(out))
    return out


WV_OPTIONS = {"impulse": 0, "Uwave": 1, "Nwave": 2}


def wv_build(oracle, option="impulse", p0=10.0, t0=0.1, alpha=4.0, length=8192, n_rise=10, n_period=40):
    """A built-in source waveform (waveforms.cpp build_wvfrm): (t, p, alpha as raised)."""
    _wv_bind(oracle.lib)
    t, p = np.zeros(length), np.zeros(length)
    a = C.c_double(alpha)
    oracle.lib.go_wvfrm_build(WV_OPTIONS.get(option, 3), p0, t0, C.byref(a), length, n_rise, n_period, _ptr(t),
                              _ptr(p))
    return t, p, a.value


def wv_run(medium, params, wv, src, incl, az, bounces, t_src, p_src, freq=0.1, wvfrm_ref=1.0, dump=None):
    """run_wnl_wvfrm through a stratified Medium or a range-dependent
    RdMedium / Rd3Medium (src: lat, lon, alt or x, y, z): dict with t_init, p_init,
    t_out, p_out, arrival (pos0, pos1, time, cel, turn_ht, incl_arr, back_az,
    tl, absorption), ok (ray returned to the ground), steps.  dump =
    (out_step, cap): also "dumps", the wvfrm_out_step snapshots (label,
    u[len], U[len/2 + 1] complex), and "dump_count"."""
    L = medium.o.lib
    _wv_bind(L)
    if dump is not None:
        L.go_wv_dump_set.argtypes = [C.c_double, _dp, C.c_int]
        L.go_wv_dump_count.restype = C.c_int
        width = 1 + wv.len + 2 * (wv.len // 2 + 1)
        dbuf = np.zeros((int(dump[1]), width))
        L.go_wv_dump_set(float(dump[0]), _ptr(dbuf), int(dump[1]))
        try:
            out = wv_run(medium, params, wv, src, incl, az, bounces, t_src, p_src, freq, wvfrm_ref)
            cnt = L.go_wv_dump_count()
        finally:
            L.go_wv_dump_set(1.0e10, None, 0)
        n = wv.len
        out["dump_count"] = cnt
        out["dumps"] = [(S[0], S[1:1 + n].copy(), S[1 + n::2] + 1j * S[2 + n::2]) for S in dbuf[:min(cnt, len(dbuf))]]
        return out
    n = wv.len
    ts = np.ascontiguousarray(t_src, dtype=np.float64)
    ps = np.ascontiguousarray(p_src, dtype=np.float64)
    s = np.ascontiguousarray(src, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("t_init", "p_init", "t_out", "p_out")}
    arr = np.zeros(9)
    ok = C.c_int(0)
    run = L.go_h_rd_wnl_run if isinstance(medium, RdMedium) else L.go_h_wnl_run
    steps = run(medium.h, C.byref(params), C.byref(wv), _ptr(s), float(incl), float(az), int(bounces),
                           float(freq), float(wvfrm_ref), _ptr(ts), _ptr(ps), _ptr(out["t_init"]), _ptr(out["p_init"]),
                           _ptr(out["t_out"]), _ptr(out["p_out"]), _ptr(arr), C.byref(ok))
    out.update(arrival=arr, ok=bool(ok.value), steps=int(steps))
    return out


def wv_run_2d(medium, wv, incl, az, bounces, t_src, p_src, src_alt=0.0, freq=0.1, wvfrm_ref=1.0):
    """infraga-2d run_wnl_wvfrm through a Medium2D: dict with t_init,
    p_init, t_out, p_out, arrival (range, -, time, cel, turn_ht, -, -, tl,
    absorption), ok, steps."""
    L = medium.o.lib
    _wv_bind(L)
    n = wv.len
    ts = np.ascontiguousarray(t_src, dtype=np.float64)
    ps = np.ascontiguousarray(p_src, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("t_init", "p_init", "t_out", "p_out")}
    arr = np.zeros(9)
    ok = C.c_int(0)
    steps = L.go2_h_wnl_run(medium.h, C.byref(wv), float(src_alt), float(incl), float(az), int(bounces), float(freq),
                            float(wvfrm_ref), _ptr(ts), _ptr(ps), _ptr(out["t_init"]), _ptr(out["p_init"]),
                            _ptr(out["t_out"]), _ptr(out["p_out"]), _ptr(arr), C.byref(ok))
    out.update(arrival=arr, ok=bool(ok.value), steps=int(steps))
    return out
