"""ga_fan_run_device_noalloc against ga_fan_run_device, through the C ABI.

The no-allocation entry point runs the same kernels on a fixed device
array, so its arrivals, arrival counts and ray-step counts must equal those
of ga_fan_run_device bit for bit.  Two calls in a row, the second smaller,
check that what the array keeps from the first call does not leak into the
second; a fan beyond the array's limit is refused before anything runs.
Device buffers come from the HIP runtime the engine links (hip_mem).
"""
import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import (ARRIVAL_DTYPE, GA_E_ARG, GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, Engine,
                                EngineError, prop_params)

pytestmark = pytest.mark.gpu

NOALLOC_SLOTS = 157285   # arrival slots the fixed array holds (n x (bounces + 1))


def _run(eng, hip, fn, aid, p, L):
    n = len(L)
    slots = p.bounces + 1
    d_l = hip.upload(np.asarray(L, dtype=np.float64).reshape(-1, 2)) if n else hip.alloc(16)
    d_out = hip.alloc(max(n, 1) * slots * ARRIVAL_DTYPE.itemsize, fill=0)
    d_cnt = hip.alloc(max(n, 1) * 4, fill=0xFF)       # -1: never written
    d_st = hip.alloc(max(n, 1) * 4, fill=0xFF)
    fn(aid, p, d_l, n, d_out, d_cnt, d_st)
    hip.sync()
    cnt = hip.download(d_cnt, np.int32, n)
    st = hip.download(d_st, np.int32, n)
    raw = hip.download(d_out, ARRIVAL_DTYPE, n * slots)
    rows = np.concatenate([raw[r * slots: r * slots + cnt[r]] for r in range(n)]) if n else raw[:0]
    return cnt, st, rows


def _same(a, b):
    assert np.array_equal(a[0], b[0]), "arrival counts differ"
    assert np.array_equal(a[1], b[1]), "ray-step counts differ"
    assert a[2].tobytes() == b[2].tobytes(), "arrival rows differ"


@pytest.mark.parametrize("geom,calc_amp", [(GA_GEOM_SPH, 1), (GA_GEOM_3D, 0)], ids=["sph18", "3d6"])
def test_noalloc_equals_device_entry(pert_units, geom, calc_amp):
    from hip_mem import Hip
    eng = Engine(0)
    hip = Hip()
    try:
        aid = eng.load_atmo(*pert_units)
        src = (30.0, -100.0, 0.0) if geom == GA_GEOM_SPH else (0.0, 0.0, 0.0)
        p = prop_params(geom, calc_amp=calc_amp, bounces=2, src=src, rng_max=700.0)
        big = synth.fan_launches(1.0, 40.0, 3.0, 0.0, 350.0, 10.0)       # 13 x 36 rays
        small = synth.fan_launches(2.0, 30.0, 4.0, 45.0, 225.0, 90.0)    # 7 x 3 rays
        assert len(small) < len(big)
        for L in (big, small):
            ref = _run(eng, hip, eng.fan_device, aid, p, L)
            got = _run(eng, hip, eng.fan_device_noalloc, aid, p, L)
            print(f"\nrays {len(L)}: arrivals {len(got[2])}, ray-steps {int(got[1].sum())}")
            assert len(ref[2]) > 0
            assert (got[0] >= 0).all() and (got[1] > 0).all()
            _same(got, ref)
    finally:
        hip.free_all()
        eng.close()


def test_noalloc_refuses_beyond_its_limit(pert_units):
    from hip_mem import Hip
    eng = Engine(0)
    hip = Hip()
    try:
        aid = eng.load_atmo(*pert_units)
        p = prop_params(GA_GEOM_SPH, bounces=4)
        n = NOALLOC_SLOTS // 5 + 1
        d_l = hip.alloc(n * 16, fill=0)
        d_out = hip.alloc(n * 5 * ARRIVAL_DTYPE.itemsize, fill=0)
        d_cnt = hip.alloc(n * 4, fill=0xFF)
        d_st = hip.alloc(n * 4, fill=0xFF)
        with pytest.raises(EngineError) as ei:
            eng.fan_device_noalloc(aid, p, d_l, n, d_out, d_cnt, d_st)
        assert ei.value.code == GA_E_ARG
        hip.sync()
        assert (hip.download(d_cnt, np.int32, n) == -1).all(), "nothing may run on a refused call"
        q = prop_params(GA_GEOM_SPH_RNGDEP)
        with pytest.raises(EngineError):
            eng.fan_device_noalloc(aid, q, d_l, 1, d_out, d_cnt, d_st)
    finally:
        hip.free_all()
        eng.close()
