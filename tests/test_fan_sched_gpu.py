"""The rotating slice scheduler of plain fans against the persistent ray
queue, through the C ABI.

Both run the same per-ray arithmetic (lane_loop / slice_trace), so arrival
rows, arrival counts and ray-step counts must agree bit for bit whatever
the slice length, the number of lanes (rays rotating through fewer lanes
than rays, many slices ending in the middle of legs and at reflections)
and however often the captured graph of slices is replayed.  The graph
replays are interleaved with eager work of the persistent path and each
replay is checked against it.
"""
import os

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, Engine, prop_params

pytestmark = pytest.mark.gpu


def _fan(eng, hip, aid, p, L, env):
    """ga_fan_run_device with the scheduler settings in env: per-ray arrival
    counts, per-ray ray-steps and the arrival rows in ray order."""
    from infraga_amd.engine import ARRIVAL_DTYPE
    n, slots = len(L), p.bounces + 1
    d_l = hip.upload(np.asarray(L, dtype=np.float64))
    d_out = hip.alloc(n * slots * ARRIVAL_DTYPE.itemsize, fill=0)
    d_cnt = hip.alloc(n * 4, fill=0xFF)
    d_st = hip.alloc(n * 4, fill=0xFF)
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        eng.fan_device(aid, p, d_l, n, d_out, d_cnt, d_st)
        hip.sync()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    cnt = hip.download(d_cnt, np.int32, n)
    st = hip.download(d_st, np.int32, n)
    raw = hip.download(d_out, ARRIVAL_DTYPE, n * slots)
    assert (cnt >= 0).all() and (st > 0).all(), "every ray must be finished"
    return cnt, st, np.concatenate([raw[r * slots: r * slots + cnt[r]] for r in range(n)])


def _same(a, b, what):
    assert np.array_equal(a[0], b[0]), f"{what}: arrival counts differ"
    assert np.array_equal(a[1], b[1]), f"{what}: ray-step counts differ"
    assert a[2].tobytes() == b[2].tobytes(), f"{what}: arrival rows differ"


CASES = [
    ("sph18", GA_GEOM_SPH, 1, (30.0, -100.0, 0.0), 3),
    ("3d6", GA_GEOM_3D, 0, (0.0, 0.0, 0.0), 2),
]


@pytest.mark.parametrize("name,geom,amp,src,bounces", CASES, ids=[c[0] for c in CASES])
def test_slices_equal_persistent_queue(pert_units, name, geom, amp, src, bounces):
    from hip_mem import Hip
    eng = Engine(0)
    hip = Hip()
    try:
        aid = eng.load_atmo(*pert_units)
        p = prop_params(geom, calc_amp=amp, bounces=bounces, src=src, rng_max=900.0)
        L = synth.fan_launches(0.5, 44.5, 0.5, 0.0, 345.0, 15.0)        # 89 x 24 = 2136 rays
        ref = _fan(eng, hip, aid, p, L, {"GA_FAN_SLICE": 0})
        assert len(ref[2]) > 1000
        for slice_steps, grid in ((7, 2), (256, 1), (4096, 3), (256, 0)):
            got = _fan(eng, hip, aid, p, L, {"GA_FAN_SLICE": slice_steps, "GA_FAN_GRID": grid})
            _same(got, ref, f"slice {slice_steps}, grid {grid}")
    finally:
        hip.free_all()
        eng.close()


def test_graph_replays_equal_eager(pert_units):
    """The cached graph of a slice batch is replayed by repeated calls with
    the same buffers; eager persistent-queue runs go between the replays."""
    from hip_mem import Hip
    from infraga_amd.engine import ARRIVAL_DTYPE
    eng = Engine(0)
    hip = Hip()
    try:
        aid = eng.load_atmo(*pert_units)
        p = prop_params(GA_GEOM_SPH, calc_amp=1, bounces=2, src=(30.0, -100.0, 0.0), rng_max=700.0)
        L = synth.fan_launches(1.0, 40.0, 1.0, 0.0, 340.0, 20.0)
        n, slots = len(L), p.bounces + 1
        d_l = hip.upload(np.asarray(L, dtype=np.float64))
        d_out = hip.alloc(n * slots * ARRIVAL_DTYPE.itemsize, fill=0)
        d_cnt = hip.alloc(n * 4, fill=0xFF)
        d_st = hip.alloc(n * 4, fill=0xFF)

        def read():
            hip.sync()
            cnt = hip.download(d_cnt, np.int32, n)
            st = hip.download(d_st, np.int32, n)
            raw = hip.download(d_out, ARRIVAL_DTYPE, n * slots)
            return cnt, st, np.concatenate([raw[r * slots: r * slots + cnt[r]] for r in range(n)])

        os.environ["GA_FAN_GRAPH"] = "4"
        os.environ["GA_FAN_SLICE"] = "64"
        try:
            # the first call builds the atmosphere's absorption table, which
            # the no-allocation entry point then uses as well
            eng.fan_device(aid, p, d_l, n, d_out, d_cnt, d_st)
            hip.sync()
            eng.fan_device_noalloc(aid, p, d_l, n, d_out, d_cnt, d_st)
            eager = read()
            assert len(eager[2]) > 100
            for rep in range(4):
                hip.L.hipMemset(d_cnt, 0xFF, n * 4)
                eng.fan_device(aid, p, d_l, n, d_out, d_cnt, d_st)
                _same(read(), eager, f"replay {rep}")
                eng.fan_device_noalloc(aid, p, d_l, n, d_out, d_cnt, d_st)
                _same(read(), eager, f"eager after replay {rep}")
        finally:
            os.environ.pop("GA_FAN_GRAPH", None)
            os.environ.pop("GA_FAN_SLICE", None)
    finally:
        hip.free_all()
        eng.close()


def test_autotune_keeps_results(pert_units):
    """ga_fan_autotune picks candidate settings and later calls with them
    give the persistent queue's results bit for bit."""
    from hip_mem import Hip
    eng = Engine(0)
    hip = Hip()
    try:
        aid = eng.load_atmo(*pert_units)
        p = prop_params(GA_GEOM_SPH, calc_amp=1, bounces=2, src=(30.0, -100.0, 0.0), rng_max=700.0)
        L = synth.fan_launches(1.0, 40.0, 1.0, 0.0, 340.0, 20.0)
        ref = _fan(eng, hip, aid, p, L, {"GA_FAN_SLICE": 0})
        d_l = hip.upload(np.asarray(L, dtype=np.float64))
        best = eng.fan_autotune(aid, p, d_l, len(L))
        assert best.slice_steps in (32, 64, 128, 256, 512) and best.slices_per_graph in (8, 32)
        assert best.pass_ms > 0.0
        _same(_fan(eng, hip, aid, p, L, {}), ref, "after autotune")
    finally:
        hip.free_all()
        eng.close()
