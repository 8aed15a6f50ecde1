"""Pin of the oracle's 2D waveform (ga_wvfrm.c go2_wnl_run: wnl_wvfrm of
geoac.eqset.2d.cpp and run_wnl_wvfrm of main.infraga.2d.cpp).

In a homogeneous windless medium the 2D effective-sound-speed geometry is
the 3D one restricted to the vertical plane of the launch: the same straight
rays, the Jacobian s^2 cos(theta) in both, the 2D beta
beta0 / c0 sqrt(D0 rho0 c0 / (D rho c)) equal to the 3D form (whose
ray-direction factors are 1 there), the same absorption.  So the 2D waveform
must be the pinned 3D waveform (tests/test_oracle_wvfrm.py): the same steps,
times and pressure to rounding, linear and strongly nonlinear, on a ray that
leaves through the top and on one that reflects off the ground first.  The
2D forms that no homogeneous medium reaches — the absorption scaled by the
effective sound speed of each leg's last evolved row, elevated reflections
counted as legs — are parity unpinned.
"""
import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, Medium, Medium2D, WvParams, wv_run, wv_run_2d
from test_oracle_wvfrm import gaussian, homogeneous


@pytest.mark.parametrize("p0", [1e-9, 2000.0])
@pytest.mark.parametrize("incl,bounces,alt", [(40.0, 0, 0.0), (-30.0, 1, 20.0)])
def test_2d_equals_3d_in_homogeneous_medium(oracle, p0, incl, bounces, alt):
    z, c, u, v, rho = homogeneous(1.0)
    m2 = Medium2D(oracle, z, c, u, v, rho)
    m3 = Medium(oracle, GEOM_3D, z, c, u, v, rho)
    wv = WvParams.default(len=1024)
    t, ps = gaussian(1024, 0.02, 0.5, p0)
    r2 = wv_run_2d(m2, wv, incl, 90.0, bounces, t, ps, src_alt=alt)
    r3 = wv_run(m3, m3.params(True), wv, (0.0, 0.0, alt), incl, 90.0, bounces, t, ps)
    assert r2["steps"] == r3["steps"] > 100
    np.testing.assert_allclose(r2["t_out"], r3["t_out"], rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(r2["t_init"], r3["t_init"], rtol=1e-14, atol=1e-12)
    assert np.abs(r2["p_out"] - r3["p_out"]).max() <= 1e-12 * np.abs(r3["p_out"]).max()
