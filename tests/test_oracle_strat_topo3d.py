"""Pins of the 3D stratified-over-terrain oracle (infraga-3d with topo_file).

The oracle runs a stratified profile over terrain as its 3D Cartesian
range-dependent restatement over identical columns spanning the terrain
grid, with the stratified semantics (go_rd_h_set_strat).  What differs from
the range-dependent 3D binary, and what the checks fix:
- the medium: against the reference's own atmo_state.3d.strat.cpp /
  atmo_io.3d.strat.cpp compiled here (oracle/_ref/libref_3d.so, set_region
  with a terrain file), the propagation region (x, y limits, profile bottom
  z0, the boundary-layer top at terrain max + 0.1 + 2 km), c, both calc_uvw
  overloads and the terrain with its derivatives are bit-identical at random
  points inside and above the boundary layer;
- the equation set: geoac.eqset.3d.strat.cpp:305-310 contracts the
  auxiliary equations' second-derivative tensors over the upper triangle
  only (m = n..2, quirk Q2).  On the same medium, the 18-equation
  right-hand sides with and without that contraction differ by exactly the
  dropped terms, built from the pinned calc_uvw second derivatives:
  -(nu_x sum_{m<n} X_m u_nm + nu_y sum_{m<n} X_m v_nm) / |c_g| in the mu
  equations, nothing elsewhere.
"""
import math
import os

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, Rd3Medium, Ref
from infraga_amd import synth

HAVE_REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_3d.so"))


def scene(seed=4):
    z, c, u, v, rho = synth.to_engine_units(*synth.base_profile())
    # winds at the ground, which the boundary layer scales over the terrain
    u = u + 0.012 * np.exp(-z / 5.0)
    v = v - 0.007 * np.exp(-z / 5.0)
    _, _, zg = synth.fractal_topo(seed, lat=np.linspace(30.0, 36.0, 25), lon=np.linspace(-104.0, -97.0, 29),
                                  amp_km=1.5)
    tx = np.linspace(-300.0, 300.0, 25)
    ty = np.linspace(-350.0, 350.0, 29)
    return (z, c, u, v, rho), (tx, ty, zg)


def write_files(d, prof, topo):
    z, c, u, v, rho = prof
    met = os.path.join(d, "strat.met")
    with open(met, "w") as f:
        for row in zip(z, c * 1e3, u * 1e3, v * 1e3, rho):
            f.write(" ".join("%.17g" % x for x in row) + "\n")
    tx, ty, zg = topo
    tp = os.path.join(d, "topo.dat")
    with open(tp, "w") as f:
        for i, x in enumerate(tx):
            f.write("".join("%.17g %.17g %.17g\n" % (x, y, zg[i, j]) for j, y in enumerate(ty)))
    return met, tp


def strat_medium(oracle, prof, topo):
    z, c, u, v, rho = prof
    # the oracle reads the same doubles the reference parsed from the text
    return Rd3Medium.strat_over_terrain(oracle, z, (c * 1e3) / 1000.0, (u * 1e3) / 1000.0, (v * 1e3) / 1000.0, rho,
                                        *topo)


# oracle calc_uvw derivative order (z, x, y), packed (zz, xx, yy, zx, zy, xy);
# the reference's (x, y, z), packed (xx, yy, zz, xy, xz, yz)
D1 = [1, 2, 0]
D2 = [1, 2, 0, 5, 3, 4]


def as_reference_order(o, second):
    out = list(o[:3])
    for f in range(3):
        out += [o[3 + 3 * f + i] for i in D1]
    if second:
        for f in range(3):
            out += [o[12 + 6 * f + i] for i in D2]
    return np.array(out)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref_3d.so not built (no reference tree)")
@pytest.mark.parametrize("use_blw", [False, True])
def test_medium_matches_reference_3d_strat_over_terrain(oracle, tmp_path, use_blw):
    prof, topo = scene()
    met, tp = write_files(str(tmp_path), prof, topo)
    ref = Ref(GEOM_3D)
    ref.load_topo(met, tp)
    ref.lib.ref_use_blw(int(use_blw))
    m = strat_medium(oracle, prof, topo)
    m.config(use_blw=use_blw)
    try:
        lim, rlim = m.limits(), ref.limits()
        assert np.array_equal(lim, rlim), (lim, rlim)
        assert abs(lim[6] - (topo[2].max() + 0.1)) < 1e-12 and abs(lim[7] - lim[6] - 2.0) < 1e-12
        rng = np.random.default_rng(23)
        w_seen = 0
        for _ in range(400):
            x = rng.uniform(-300.0, 300.0)
            y = rng.uniform(-350.0, 350.0)
            zt = m.topo(x, y)[0]
            z_ = zt + rng.uniform(0.0, 3.0) if rng.random() < 0.7 else rng.uniform(0.0, 120.0)
            got = np.concatenate([[m.point(z_, x, y)[0], m.point(z_, x, y)[4]],
                                  as_reference_order(m.calc_uvw(z_, x, y, False), False),
                                  as_reference_order(m.calc_uvw(z_, x, y, True), True), m.topo(x, y)[:6]])
            exp = np.concatenate([[ref.lib.ref_c(x, y, z_), ref.lib.ref_w(x, y, z_)],
                                  ref.calc_uvw(x, y, z_, False), ref.calc_uvw(x, y, z_, True), ref.topo(x, y)])
            assert np.array_equal(got, exp), (z_, np.flatnonzero(got != exp), got[got != exp], exp[got != exp])
            if use_blw and z_ - zt < 2.0:
                w_seen += abs(got[1]) > 1e-6
        if use_blw:
            assert w_seen > 10   # the vertical wind is on inside the boundary layer
    finally:
        ref.lib.ref_use_blw(0)
        ref.lib.ref_clear()


def test_upper_triangle_contraction(oracle):
    prof, topo = scene()
    m = strat_medium(oracle, prof, topo)
    p = m.params(True)
    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(20):
        x, y = rng.uniform(-250.0, 250.0), rng.uniform(-300.0, 300.0)
        zz = m.topo(x, y)[0] + rng.uniform(0.03, 1.5)
        nu = rng.normal(size=3)
        nu *= 3.0 / np.linalg.norm(nu)
        aux = rng.normal(size=12)
        Y = np.concatenate([[x, y, zz], nu, aux])
        fs = m.rhs(p, Y)
        m.set_upper(False)
        ff = m.rhs(p, Y)
        m.set_upper(True)
        o = m.calc_uvw(zz, x, y, True)
        cc = m.point(zz, x, y)[0]
        # second derivatives in Cartesian index order (x, y, z)
        pk = lambda f: np.array(as_reference_order(o, True)[12 + 6 * f:18 + 6 * f])
        full_mat = lambda q: np.array([[q[0], q[3], q[4]], [q[3], q[1], q[5]], [q[4], q[5], q[2]]])
        Du, Dv = full_mat(pk(0)), full_mat(pk(1))
        cg = np.array([cc * nu[0] / np.linalg.norm(nu) + o[0], cc * nu[1] / np.linalg.norm(nu) + o[1],
                       cc * nu[2] / np.linalg.norm(nu)])
        cgm = np.linalg.norm(cg)
        exp = np.zeros(18)
        for j, oX in ((0, 6), (1, 12)):
            X = Y[oX:oX + 3]
            for n in range(3):
                du = sum(X[mm] * Du[n, mm] for mm in range(n))
                dv = sum(X[mm] * Dv[n, mm] for mm in range(n))
                exp[(9 if j == 0 else 15) + n] = -(nu[0] * du + nu[1] * dv) / cgm
        d = ff - fs
        assert np.abs(d[:9]).max() == 0.0 and np.abs(d[12:15]).max() == 0.0
        np.testing.assert_allclose(d, exp, rtol=1e-8, atol=1e-12 * np.abs(ff).max())
        checked += np.abs(exp).max() > 1e-9
    assert checked >= 10     # the dropped terms are not zero near sloped ground
