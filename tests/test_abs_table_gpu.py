"""The absorption table on the GPU: a fan with the table (the default)
against the same fan with direct Sutherland-Bass evaluation
(GA_ABS_TABLE=0), spherical and 3D.  Absorption does not feed back into the
ray, so arrival keys, ray-step counts and every other column are
bit-identical; absorption agrees within 1e-10 relative (the table's own
bound is 1e-12 per step), and differs in the last bits for most arrivals,
which shows the table path ran."""
import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, prop_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geom,src", [(GA_GEOM_SPH, (30.0, 0.0, 0.0)), (GA_GEOM_3D, (0.0, 0.0, 0.0))], ids=["sph", "3d"])
def test_table_equals_direct(engine, pert_units, monkeypatch, geom, src):
    L = synth.fan_launches(0.5, 44.5, 1.0, 0.0, 350.0, 10.0)
    aid = engine.load_atmo(*pert_units)
    try:
        p = prop_params(geom, calc_amp=1, bounces=4, src=src, rng_max=900.0)
        tab, st_t = engine.fan(aid, p, L)
        monkeypatch.setenv("GA_ABS_TABLE", "0")
        direct, st_d = engine.fan(aid, p, L)
    finally:
        monkeypatch.delenv("GA_ABS_TABLE", raising=False)
        engine.release_atmo(aid)
    print(f"\n{len(L)} rays, {len(tab)} arrivals, {st_t.ray_steps} ray-steps; kernels {st_t.kernel_ms:.1f} ms "
          f"(table) vs {st_d.kernel_ms:.1f} ms (direct)")
    assert len(tab) == len(direct) > 100 and st_t.ray_steps == st_d.ray_steps
    for f in tab.dtype.names:
        if f != "absorption":
            assert np.array_equal(tab[f], direct[f]), f
    a, b = tab["absorption"], direct["absorption"]
    assert np.all(np.abs(a - b) <= 1e-10 * np.abs(b) + 1e-12), np.max(np.abs(a - b) / np.abs(b))
    assert np.count_nonzero(a != b) > len(a) // 2
