"""Step-level driving of the oracle from Python (test infrastructure)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ga_oracle_py import GEOM_SPH, Medium, Src

_dp = C.POINTER(C.c_double)


def make_medium(oracle, geom, z, c, u, v, rho):
    return Medium(oracle, geom, z, c, u, v, rho)


def src_struct(m, params, pos, theta_deg, az_deg):
    """pos: sph (alt km, lat deg, lon deg); 3D (x, y, z)."""
    s = Src()
    if m.geom == GEOM_SPH:
        s.src[0] = 6370.0 + pos[0]
        s.src[1] = math.radians(pos[1])
        s.src[2] = math.radians(pos[2])
    else:
        s.src[0], s.src[1], s.src[2] = pos
    s.theta = theta_deg * (math.pi / 180.0)
    s.phi = math.pi / 2.0 - az_deg * (math.pi / 180.0)
    return s


def trace_steps(m, params, s, nsteps, stop_at_ground=True):
    """Integrate nsteps RK4 steps from set_initial; returns rows (n+1, 18) and s."""
    L = m.o.lib
    rows = np.zeros((nsteps + 1, 18))
    y0 = np.zeros(18)
    L.go_set_initial(m.h, C.byref(params), C.byref(s), y0.ctypes.data_as(_dp))
    rows[0] = y0
    path = [0.0]
    y = y0.copy()
    for k in range(nsteps):
        y1 = np.zeros(18)
        ds = L.go_rk4_step(m.h, C.byref(params), C.byref(s), y.ctypes.data_as(_dp), y1.ctypes.data_as(_dp))
        rows[k + 1] = y1
        path.append(path[-1] + ds)
        y = y1
    return rows, np.array(path)


def rhs(m, params, s, y):
    f = np.zeros(18)
    yy = np.ascontiguousarray(y, dtype=np.float64)
    m.o.lib.go_rhs(m.h, C.byref(params), C.byref(s), yy.ctypes.data_as(_dp), f.ctypes.data_as(_dp))
    return f


def eikonal(m, s, y):
    yy = np.ascontiguousarray(y, dtype=np.float64)
    return m.o.lib.go_eikonal_resid(m.h, C.byref(s), yy.ctypes.data_as(_dp))


def jacobian(m, y):
    yy = np.ascontiguousarray(y, dtype=np.float64)
    return m.o.lib.go_jacobian(m.h, yy.ctypes.data_as(_dp))


def amp(m, s, y):
    yy = np.ascontiguousarray(y, dtype=np.float64)
    return m.o.lib.go_amp(m.h, C.byref(s), yy.ctypes.data_as(_dp))


def medium_point(m, p):
    L = m.o.lib
    return (L.go_c(m.h, *p), L.go_u(m.h, *p), L.go_v(m.h, *p), L.go_rho(m.h, *p))


def refl(m, params, s, rk2, rk1):
    y0 = np.zeros(18)
    a = np.ascontiguousarray(rk2, dtype=np.float64)
    b = np.ascontiguousarray(rk1, dtype=np.float64)
    m.o.lib.go_set_refl(m.h, C.byref(params), C.byref(s), a.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                        y0.ctypes.data_as(_dp))
    return y0
