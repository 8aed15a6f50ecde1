"""ctypes binding of libinfraga_gpu.so (the C ABI declared in include/infraga_gpu.h).

This is host plumbing for tests, the bench and Python callers; the product
compute path is the HIP library itself.  Loading fails loudly when the
library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libinfraga_gpu.so")

GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, GA_GEOM_2D, GA_GEOM_3D_RNGDEP = 0, 1, 2, 3, 4
GA_OK, GA_E_ARG, GA_E_HIP, GA_E_NOMEM, GA_E_CAP, GA_E_NODEV, GA_E_ATMO, GA_E_IO = 0, -1, -2, -3, -4, -5, -6, -7

GA_REQ_FAN, GA_REQ_EIG = 0, 1
GA_SERVE_MERGE, GA_SERVE_SERIAL = 1, 2
GA_PENDING = 1

EXPORTED = ["ga_prop_params_default", "ga_eig_params_default", "ga_ctx_create", "ga_ctx_destroy",
            "ga_last_error", "ga_atmo_load_strat", "ga_atmo_load_file", "ga_atmo_release", "ga_fan_run",
            "ga_fan_run_device", "ga_eig_search", "ga_server_create", "ga_server_submit", "ga_server_poll",
            "ga_server_stats", "ga_server_destroy", "ga_rngdep_load", "ga_rngdep_load_xy", "ga_rngdep_load_files_xy",
            "ga_topo_load", "ga_rngdep_eval",
            "ga_rngdep_release", "ga_rngdep_load_files", "ga_rngdep_info", "ga_fan_run_paths",
            "ga_eig_search_paths", "ga_eig_direct", "ga_topo_load_2d", "ga_topo_eval_2d",
            "ga_fan_run_device_noalloc", "ga_fan_autotune", "ga_atmo_load_strat_topo", "ga_atmo_load_file_topo",
            "ga_atmo_load_strat_topo_xy", "ga_atmo_load_file_topo_xy", "ga_wnl_wvfrm",
            "ga_wnl_wvfrm_dump"]

GA_PATH_RAYS, GA_PATH_CAUSTICS = 1, 2
GA_PT_RAY, GA_PT_CAUSTIC = 0, 1

_dp = C.POINTER(C.c_double)


class PropParams(C.Structure):
    _fields_ = [("geometry", C.c_int), ("calc_amp", C.c_int), ("bounces", C.c_int),
                ("freq", C.c_double), ("abs_coeff", C.c_double), ("turn_ht_min", C.c_double),
                ("src", C.c_double * 3), ("z_grnd", C.c_double), ("s_max", C.c_double),
                ("ds_min", C.c_double), ("ds_max", C.c_double), ("alt_max", C.c_double),
                ("rng_max", C.c_double), ("time_max", C.c_double), ("bounds", C.c_double * 4),
                ("r_earth", C.c_double), ("topo_use_blw", C.c_int32), ("refl_alt", C.c_double)]


class EigParams(C.Structure):
    _fields_ = [("incl_min", C.c_double), ("incl_max", C.c_double), ("bnc_min", C.c_int),
                ("bnc_max", C.c_int), ("iterations", C.c_int), ("damping", C.c_double),
                ("tolerance", C.c_double), ("az_dev_lim", C.c_double), ("incl_step_min", C.c_double),
                ("incl_step_max", C.c_double)]


class Arrival(C.Structure):
    _fields_ = [("launch_idx", C.c_int32), ("n_bounce", C.c_int32), ("incl", C.c_double), ("az", C.c_double),
                ("pos0", C.c_double), ("pos1", C.c_double), ("time", C.c_double), ("cel", C.c_double),
                ("turn_ht", C.c_double), ("incl_arr", C.c_double), ("back_az", C.c_double),
                ("tl", C.c_double), ("absorption", C.c_double)]


class PathPt(C.Structure):
    _fields_ = [("launch_idx", C.c_int32), ("leg", C.c_int32), ("step", C.c_int32), ("kind", C.c_int32),
                ("pos0", C.c_double), ("pos1", C.c_double), ("z", C.c_double), ("tl", C.c_double),
                ("absorption", C.c_double), ("time", C.c_double)]


PATH_DTYPE = np.dtype([(f, "i4" if t is C.c_int32 else "f8") for f, t in PathPt._fields_])


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("atmo_id", C.c_int32), ("prop", PropParams), ("eig", EigParams),
                ("data", _dp), ("n", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_rows", C.c_int32), ("queue_ms", C.c_double),
                ("total_ms", C.c_double), ("ray_steps", C.c_int64)]


class FanTuning(C.Structure):
    _fields_ = [("slice_steps", C.c_int32), ("slices_per_graph", C.c_int32), ("pass_ms", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("ray_steps", C.c_int64), ("kernel_ms", C.c_double), ("wall_ms", C.c_double),
                ("n_rays", C.c_int32), ("n_launches", C.c_int32)]


class WvfrmParams(C.Structure):
    """-wnl_wvfrm settings (waveforms.cpp:20-35 and run_wnl_wvfrm defaults)."""
    _fields_ = [("len", C.c_int32), ("reserved", C.c_int32), ("ds", C.c_double), ("ref", C.c_double),
                ("filt1_g", C.c_double), ("filt1_n1", C.c_double), ("filt1_n2", C.c_double),
                ("filt2_g", C.c_double), ("filt2_n1", C.c_double), ("filt2_n2", C.c_double)]


def wvfrm_params(length=8192, ds=0.5, ref=1.0, filt1=(0.5, 4.0, 0.5), filt2=(0.75, 2.0, 1.0)) -> WvfrmParams:
    return WvfrmParams(int(length), 0, float(ds), float(ref), *map(float, filt1), *map(float, filt2))


class WvfrmStats(C.Structure):
    _fields_ = [("ray_steps", C.c_int64), ("wv_steps", C.c_int64), ("trace_ms", C.c_double),
                ("evolve_ms", C.c_double), ("wall_ms", C.c_double), ("n_waveforms", C.c_int32),
                ("reserved", C.c_int32)]


ARRIVAL_DTYPE = np.dtype([("launch_idx", "i4"), ("n_bounce", "i4"), ("incl", "f8"), ("az", "f8"),
                          ("pos0", "f8"), ("pos1", "f8"), ("time", "f8"), ("cel", "f8"), ("turn_ht", "f8"),
                          ("incl_arr", "f8"), ("back_az", "f8"), ("tl", "f8"), ("absorption", "f8")])
assert ARRIVAL_DTYPE.itemsize == C.sizeof(Arrival)


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"infraga_gpu error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str | None = None):
    """Load libinfraga_gpu.so and declare its signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(p)
    L.ga_prop_params_default.argtypes = [C.POINTER(PropParams), C.c_int]
    L.ga_eig_params_default.argtypes = [C.POINTER(EigParams)]
    L.ga_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    L.ga_ctx_destroy.argtypes = [C.c_void_p]
    L.ga_last_error.argtypes = [C.c_void_p]
    L.ga_last_error.restype = C.c_char_p
    L.ga_atmo_load_strat.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int)]
    L.ga_atmo_load_file.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    L.ga_atmo_load_strat_topo.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                          _dp, C.POINTER(C.c_int)]
    L.ga_atmo_load_file_topo.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int,
                                         C.POINTER(C.c_int)]
    L.ga_atmo_load_strat_topo_xy.argtypes = L.ga_atmo_load_strat_topo.argtypes
    L.ga_atmo_load_file_topo_xy.argtypes = L.ga_atmo_load_file_topo.argtypes
    L.ga_atmo_release.argtypes = [C.c_void_p, C.c_int]
    L.ga_fan_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), _dp, C.c_int, C.POINTER(Arrival),
                             C.c_int, C.POINTER(C.c_int), C.POINTER(RunStats)]
    L.ga_fan_run_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ga_fan_run_device_noalloc.argtypes = L.ga_fan_run_device.argtypes
    L.ga_fan_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.c_void_p, C.c_int, C.c_void_p,
                                  C.POINTER(FanTuning)]
    L.ga_eig_search.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.POINTER(EigParams), _dp, C.c_int,
                                C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), _dp, C.POINTER(RunStats)]
    L.ga_fan_run_paths.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), _dp, C.c_int, C.c_int,
                                   C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), C.POINTER(PathPt), C.c_int64,
                                   C.POINTER(C.c_int64), C.POINTER(RunStats)]
    L.ga_eig_search_paths.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.POINTER(EigParams), _dp,
                                      C.c_int, C.POINTER(Arrival), C.c_int, C.POINTER(C.c_int), _dp,
                                      C.POINTER(PathPt), C.c_int64, C.POINTER(C.c_int64), C.POINTER(RunStats)]
    L.ga_topo_load_2d.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
    L.ga_topo_eval_2d.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
    L.ga_eig_direct.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.POINTER(EigParams), _dp, C.c_double,
                                C.c_double, C.c_int, C.POINTER(Arrival), C.POINTER(C.c_int), C.POINTER(PathPt),
                                C.c_int64, C.POINTER(C.c_int64), C.POINTER(RunStats)]
    L.ga_rngdep_load.argtypes = [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp,
                                 C.POINTER(C.c_int)]
    L.ga_topo_load.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
    L.ga_rngdep_eval.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _dp]
    L.ga_rngdep_info.argtypes = [C.c_void_p, C.c_int, _dp]
    L.ga_rngdep_load_files.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                       C.c_int, C.POINTER(C.c_int)]
    L.ga_rngdep_release.argtypes = [C.c_void_p, C.c_int]
    L.ga_wnl_wvfrm.argtypes = [C.c_void_p, C.c_int, C.POINTER(PropParams), C.POINTER(WvfrmParams), _dp, C.c_int,
                               _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(Arrival), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int64), C.POINTER(WvfrmStats)]
    L.ga_wnl_wvfrm_dump.argtypes = L.ga_wnl_wvfrm.argtypes + [C.c_double, C.c_int, _dp, C.POINTER(C.c_int32)]
    L.ga_server_create.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_void_p)]
    L.ga_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.ga_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(Result), C.POINTER(Arrival), C.c_int]
    L.ga_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.ga_server_destroy.argtypes = [C.c_void_p]
    for f in EXPORTED:
        getattr(L, f)
    if path is None:
        _lib = L
    return L


def prop_params(geometry, **kw) -> PropParams:
    """Reference defaults (main.infraga.sph.cpp:208-221) with overrides."""
    L = load_library()
    p = PropParams()
    L.ga_prop_params_default(C.byref(p), geometry)
    for k, v in kw.items():
        if k in ("src", "bounds"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def eig_params(**kw) -> EigParams:
    L = load_library()
    e = EigParams()
    L.ga_eig_params_default(C.byref(e))
    for k, v in kw.items():
        setattr(e, k, v)
    return e


class Engine:
    """One device context (ga_ctx) with resident atmospheres."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.ga_ctx_create(C.byref(h), device)
        if rc != GA_OK:
            raise EngineError(rc, "ga_ctx_create failed (no usable GPU?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.ga_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GA_OK:
            raise EngineError(rc, self.L.ga_last_error(self.h).decode())

    def load_atmo(self, z, c, u, v, rho) -> int:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (z, c, u, v, rho)]
        aid = C.c_int(0)
        self._check(self.L.ga_atmo_load_strat(self.h, *[a.ctypes.data_as(_dp) for a in arrs], len(arrs[0]),
                                              C.byref(aid)))
        return aid.value

    def load_atmo_file(self, path, fmt="zTuvdp", reverse_winds=False) -> int:
        aid = C.c_int(0)
        self._check(self.L.ga_atmo_load_file(self.h, path.encode(), fmt.encode(), int(reverse_winds), C.byref(aid)))
        return aid.value

    def release_atmo(self, aid):
        self._check(self.L.ga_atmo_release(self.h, aid))

    def load_strat_topo(self, z, c, u, v, rho, lat_deg, lon_deg, zg, xy=False) -> int:
        """A stratified profile (engine units) over terrain zg[nlat][nlon] [km]:
        a range-dependent id to run with GA_GEOM_SPH_RNGDEP (ga_atmo_load_strat_topo);
        xy: terrain nodes x, y in km, for GA_GEOM_3D_RNGDEP (ga_atmo_load_strat_topo_xy)."""
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (z, c, u, v, rho, lat_deg, lon_deg, zg)]
        rid = C.c_int(0)
        fn = self.L.ga_atmo_load_strat_topo_xy if xy else self.L.ga_atmo_load_strat_topo
        self._check(fn(self.h, *[q.ctypes.data_as(_dp) for q in a[:5]], len(a[0]), a[5].ctypes.data_as(_dp),
                       len(a[5]), a[6].ctypes.data_as(_dp), len(a[6]), a[7].ctypes.data_as(_dp), C.byref(rid)))
        return rid.value

    def load_file_topo(self, path, topo_path, fmt="zTuvdp", reverse_winds=False, xy=False) -> int:
        rid = C.c_int(0)
        fn = self.L.ga_atmo_load_file_topo_xy if xy else self.L.ga_atmo_load_file_topo
        self._check(fn(self.h, path.encode(), topo_path.encode(), fmt.encode(), int(reverse_winds), C.byref(rid)))
        return rid.value

    def fan(self, aid, params: PropParams, launches):
        """Trace a fan; returns (arrivals structured array, RunStats)."""
        Lc = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        n = Lc.shape[0]
        cap = max(1, n * (params.bounces + 1))
        out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
        n_out = C.c_int(0)
        st = RunStats()
        self._check(self.L.ga_fan_run(self.h, aid, C.byref(params), Lc.ctypes.data_as(_dp), n,
                                      out.ctypes.data_as(C.POINTER(Arrival)), cap, C.byref(n_out), C.byref(st)))
        return out[: n_out.value].copy(), st

    def fan_paths(self, aid, params: PropParams, launches, flags=GA_PATH_RAYS, pts_cap=1 << 20):
        """Trace a fan with the write_rays / write_caustics bookkeeping;
        returns (arrivals, path rows, RunStats).  A too-small pts_cap is
        raised to the count the call reports and the fan traced again."""
        Lc = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        n = Lc.shape[0]
        cap = max(1, n * (params.bounces + 1))
        out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
        n_out = C.c_int(0)
        n_pts = C.c_int64(0)
        st = RunStats()
        while True:
            pts = np.zeros(max(1, pts_cap), dtype=PATH_DTYPE)
            rc = self.L.ga_fan_run_paths(self.h, aid, C.byref(params), Lc.ctypes.data_as(_dp), n, int(flags),
                                         out.ctypes.data_as(C.POINTER(Arrival)), cap, C.byref(n_out),
                                         pts.ctypes.data_as(C.POINTER(PathPt)), pts_cap, C.byref(n_pts),
                                         C.byref(st))
            if rc == GA_E_CAP and n_pts.value > pts_cap:
                pts_cap = n_pts.value
                continue
            break
        self._check(rc)
        return out[: n_out.value].copy(), pts[: n_pts.value].copy(), st

    def wnl_wvfrm(self, aid, params: PropParams, w: WvfrmParams, launches, t_src, p_src, out_step=None,
                  dump_cap=64):
        """-wnl_wvfrm for each (incl, az) launch with the source waveform
        (t_src, p_src); returns a dict of t_init, p_init, t_out, p_out
        ([n, len] each), arrivals, status, steps and the WvfrmStats.  With
        out_step (wvfrm_out_step=, km) it also returns "dumps": per launch a
        list of (label, u[len], U[len/2 + 1] complex) for the first dump_cap
        snapshots, and "dump_count" (ga_wnl_wvfrm_dump)."""
        Lc = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        n = Lc.shape[0]
        ln = int(w.len)
        t = np.ascontiguousarray(t_src, dtype=np.float64)
        pz = np.ascontiguousarray(p_src, dtype=np.float64)
        if t.shape != (ln,) or pz.shape != (ln,):
            raise ValueError(f"source waveform must have {ln} samples")
        out = {k: np.zeros((n, ln)) for k in ("t_init", "p_init", "t_out", "p_out")}
        arr = np.zeros(max(n, 1), dtype=ARRIVAL_DTYPE)
        status = np.zeros(max(n, 1), dtype=np.int32)
        steps = np.zeros(max(n, 1), dtype=np.int64)
        st = WvfrmStats()
        dp = lambda a: a.ctypes.data_as(_dp)
        args = [self.h, aid, C.byref(params), C.byref(w), dp(Lc), n, dp(t), dp(pz),
                dp(out["t_init"]), dp(out["p_init"]), dp(out["t_out"]), dp(out["p_out"]),
                arr.ctypes.data_as(C.POINTER(Arrival)), status.ctypes.data_as(C.POINTER(C.c_int32)),
                steps.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(st)]
        if out_step is None:
            self._check(self.L.ga_wnl_wvfrm(*args))
        else:
            width = 1 + ln + 2 * (ln // 2 + 1)
            snaps = np.zeros((max(n, 1), int(dump_cap), width))
            cnt = np.zeros(max(n, 1), dtype=np.int32)
            self._check(self.L.ga_wnl_wvfrm_dump(*args, float(out_step), int(dump_cap), dp(snaps),
                                                 cnt.ctypes.data_as(C.POINTER(C.c_int32))))
            out["dump_count"] = cnt[:n].copy()
            out["dumps"] = [[(S[0], S[1:1 + ln].copy(), S[1 + ln::2] + 1j * S[2 + ln::2])
                             for S in snaps[i, :min(int(cnt[i]), int(dump_cap))]] for i in range(n)]
        out.update(arrivals=arr[:n].copy(), status=status[:n].copy(), steps=steps[:n].copy(), stats=st)
        return out

    def fan_device(self, aid, params: PropParams, d_launches: int, n: int, d_out: int, d_counts: int,
                   d_steps: int, stream: int = 0):
        """Enqueue a fan on device-resident buffers (raw device pointers)."""
        self._check(self.L.ga_fan_run_device(self.h, aid, C.byref(params), C.c_void_p(d_launches), n,
                                             C.c_void_p(d_out), C.c_void_p(d_counts), C.c_void_p(d_steps),
                                             C.c_void_p(stream) if stream else None))

    def fan_autotune(self, aid, params: PropParams, d_launches: int, n: int, stream: int = 0) -> FanTuning:
        """Pick the slice scheduler's settings on this HBM-resident launch list."""
        best = FanTuning()
        self._check(self.L.ga_fan_autotune(self.h, aid, C.byref(params), C.c_void_p(d_launches), n,
                                           C.c_void_p(stream) if stream else None, C.byref(best)))
        return best

    def fan_device_noalloc(self, aid, params: PropParams, d_launches: int, n: int, d_out: int, d_counts: int,
                           d_steps: int, stream: int = 0):
        """fan_device through the entry point that allocates no device memory."""
        self._check(self.L.ga_fan_run_device_noalloc(self.h, aid, C.byref(params), C.c_void_p(d_launches), n,
                                                     C.c_void_p(d_out), C.c_void_p(d_counts), C.c_void_p(d_steps),
                                                     C.c_void_p(stream) if stream else None))

    def topo_load_2d(self, aid, r_km, z_km):
        """Terrain profile z(r) for GA_GEOM_2D runs on a stratified atmosphere."""
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (r_km, z_km)]
        self._check(self.L.ga_topo_load_2d(self.h, aid, a[0].ctypes.data_as(_dp), a[1].ctypes.data_as(_dp),
                                           len(a[0])))

    def rngdep_load(self, lat_deg, lon_deg, z_km, c, u, v, rho) -> int:
        """Range-dependent atmosphere: node lat/lon [deg], z [km], fields [nlat][nlon][nz]."""
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (lat_deg, lon_deg, z_km, c, u, v, rho)]
        rid = C.c_int(0)
        self._check(self.L.ga_rngdep_load(self.h, a[0].ctypes.data_as(_dp), len(a[0]), a[1].ctypes.data_as(_dp),
                                          len(a[1]), a[2].ctypes.data_as(_dp), len(a[2]),
                                          *[q.ctypes.data_as(_dp) for q in a[3:]], C.byref(rid)))
        self._keep_rd = a
        return rid.value

    def rngdep_load_xy(self, x_km, y_km, z_km, c, u, v, rho) -> int:
        """3D Cartesian range-dependent atmosphere (GA_GEOM_3D_RNGDEP): node x/y, z [km], fields [nx][ny][nz]."""
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (x_km, y_km, z_km, c, u, v, rho)]
        rid = C.c_int(0)
        self._check(self.L.ga_rngdep_load_xy(self.h, a[0].ctypes.data_as(_dp), len(a[0]), a[1].ctypes.data_as(_dp),
                                             len(a[1]), a[2].ctypes.data_as(_dp), len(a[2]),
                                             *[q.ctypes.data_as(_dp) for q in a[3:]], C.byref(rid)))
        self._keep_rd = a
        return rid.value

    def topo_load(self, rid, lat_deg, lon_deg, zg_km):
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (lat_deg, lon_deg, zg_km)]
        self._check(self.L.ga_topo_load(self.h, rid, a[0].ctypes.data_as(_dp), len(a[0]), a[1].ctypes.data_as(_dp),
                                        len(a[1]), a[2].ctypes.data_as(_dp)))

    def rngdep_eval(self, rid, pts):
        """pts: (lat, lon [rad], z [km]) rows -> (atmo [n, 4, 10], topo [n, 10])."""
        P = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        n = P.shape[0]
        A = np.zeros((n, 4, 10))
        T = np.zeros((n, 10))
        self._check(self.L.ga_rngdep_eval(self.h, rid, P.ctypes.data_as(_dp), n, A.ctypes.data_as(_dp),
                                          T.ctypes.data_as(_dp)))
        return A, T

    def rngdep_release(self, rid):
        self._check(self.L.ga_rngdep_release(self.h, rid))

    def rngdep_info(self, rid):
        """(lat_min, lat_max, lon_min, lon_max [rad], alt_max, z0, z_max, boundary-layer top [km])."""
        out = np.zeros(8)
        self._check(self.L.ga_rngdep_info(self.h, rid, out.ctypes.data_as(_dp)))
        return out

    def rngdep_load_files(self, prefix, lat_path, lon_path, topo_path=None, fmt="zTuvdp", reverse_winds=False) -> int:
        """The reference's range-dependent file layout (node files, prefix{n}.met, terrain rows)."""
        rid = C.c_int(0)
        self._check(self.L.ga_rngdep_load_files(self.h, prefix.encode(), lat_path.encode(), lon_path.encode(),
                                                topo_path.encode() if topo_path else None, fmt.encode(),
                                                int(reverse_winds), C.byref(rid)))
        return rid.value

    def eig_search(self, aid, params: PropParams, eig: EigParams, rcvr_xy):
        R = np.ascontiguousarray(rcvr_xy, dtype=np.float64).reshape(-1, 2)
        n = R.shape[0]
        cap = max(16, 8 * n * (eig.bnc_max - eig.bnc_min + 1))
        out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
        lat = np.zeros(n)
        n_out = C.c_int(0)
        st = RunStats()
        rc = self.L.ga_eig_search(self.h, aid, C.byref(params), C.byref(eig), R.ctypes.data_as(_dp), n,
                                  out.ctypes.data_as(C.POINTER(Arrival)), cap, C.byref(n_out),
                                  lat.ctypes.data_as(_dp), C.byref(st))
        if rc == GA_E_CAP:
            cap = n_out.value
            out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
            rc = self.L.ga_eig_search(self.h, aid, C.byref(params), C.byref(eig), R.ctypes.data_as(_dp), n,
                                      out.ctypes.data_as(C.POINTER(Arrival)), cap, C.byref(n_out),
                                      lat.ctypes.data_as(_dp), C.byref(st))
        self._check(rc)
        return out[: n_out.value].copy(), lat, st


    def eig_search_paths(self, aid, params: PropParams, eig: EigParams, rcvr_xy, pts_cap=1 << 18):
        """eig_search plus the eigenray-N.dat rows of every eigenray found
        (launch_idx of a row = index of its eigenray in the result)."""
        R = np.ascontiguousarray(rcvr_xy, dtype=np.float64).reshape(-1, 2)
        n = R.shape[0]
        cap = max(16, 8 * n * (eig.bnc_max - eig.bnc_min + 1))
        lat = np.zeros(n)
        n_out = C.c_int(0)
        n_pts = C.c_int64(0)
        st = RunStats()
        while True:
            out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
            pts = np.zeros(max(1, pts_cap), dtype=PATH_DTYPE)
            rc = self.L.ga_eig_search_paths(self.h, aid, C.byref(params), C.byref(eig), R.ctypes.data_as(_dp), n,
                                            out.ctypes.data_as(C.POINTER(Arrival)), cap, C.byref(n_out),
                                            lat.ctypes.data_as(_dp), pts.ctypes.data_as(C.POINTER(PathPt)),
                                            pts_cap, C.byref(n_pts), C.byref(st))
            if rc == GA_E_CAP and (n_out.value > cap or n_pts.value > pts_cap):
                cap = max(cap, n_out.value)
                pts_cap = max(pts_cap, n_pts.value)
                continue
            break
        self._check(rc)
        return out[: n_out.value].copy(), pts[: n_pts.value].copy(), lat, st


    def eig_direct(self, aid, params: PropParams, eig: EigParams, rcvr_xy, incl_est, az_est=math.nan,
                   bounces=0, pts_cap=1 << 16):
        """-eig_direct: (eigenray row or None, its path rows, RunStats)."""
        R = np.ascontiguousarray(rcvr_xy, dtype=np.float64).reshape(2)
        out = np.zeros(1, dtype=ARRIVAL_DTYPE)
        found = C.c_int(0)
        n_pts = C.c_int64(0)
        st = RunStats()
        while True:
            pts = np.zeros(max(1, pts_cap), dtype=PATH_DTYPE)
            rc = self.L.ga_eig_direct(self.h, aid, C.byref(params), C.byref(eig), R.ctypes.data_as(_dp),
                                      float(incl_est), float(az_est), int(bounces),
                                      out.ctypes.data_as(C.POINTER(Arrival)), C.byref(found),
                                      pts.ctypes.data_as(C.POINTER(PathPt)), pts_cap, C.byref(n_pts), C.byref(st))
            if rc == GA_E_CAP and n_pts.value > pts_cap:
                pts_cap = n_pts.value
                continue
            break
        self._check(rc)
        return (out[0].copy() if found.value else None), pts[: n_pts.value].copy(), st


class Server:
    """Request server on an Engine's context (ga_server_*): submit fans and
    eigenray queries, poll their results.  While it exists, use only the
    server on that engine."""

    def __init__(self, engine: Engine, slots: int = 4, budget_bytes: float = 16e9, merge: bool = True,
                 serial: bool = False):
        self.eng = engine
        self.L = engine.L
        h = C.c_void_p()
        flags = (GA_SERVE_MERGE if merge else 0) | (GA_SERVE_SERIAL if serial else 0)
        engine._check(self.L.ga_server_create(engine.h, slots, budget_bytes, flags, C.byref(h)))
        self.h = h

    def submit_fan(self, aid, params: PropParams, launches) -> int:
        Lc = np.ascontiguousarray(launches, dtype=np.float64).reshape(-1, 2)
        r = Request(GA_REQ_FAN, aid, params, eig_params(), Lc.ctypes.data_as(_dp), Lc.shape[0])
        return self._submit(r)

    def submit_eig(self, aid, params: PropParams, eig: EigParams, receivers) -> int:
        R = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 2)
        r = Request(GA_REQ_EIG, aid, params, eig, R.ctypes.data_as(_dp), R.shape[0])
        return self._submit(r)

    def _submit(self, r: Request) -> int:
        t = C.c_uint64(0)
        rc = self.L.ga_server_submit(self.h, C.byref(r), C.byref(t))
        if rc != GA_OK:
            raise EngineError(rc, "ga_server_submit rejected the request")
        return t.value

    def poll(self, ticket: int, timeout_ms: int = -1):
        """(rows, Result) once finished, None while pending."""
        res = Result()
        cap = 1024
        while True:
            out = np.zeros(cap, dtype=ARRIVAL_DTYPE)
            rc = self.L.ga_server_poll(self.h, ticket, timeout_ms, C.byref(res),
                                       out.ctypes.data_as(C.POINTER(Arrival)), cap)
            if rc == GA_PENDING:
                return None
            if rc == GA_E_CAP:
                cap = res.n_rows
                continue
            if rc != GA_OK:
                raise EngineError(rc, self.L.ga_last_error(self.eng.h).decode())
            return out[: res.n_rows].copy(), res

    def stats(self):
        st, la, ms = C.c_int64(0), C.c_int32(0), C.c_double(0.0)
        self.L.ga_server_stats(self.h, C.byref(st), C.byref(la), C.byref(ms))
        return st.value, la.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.ga_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nan():
    return math.nan
