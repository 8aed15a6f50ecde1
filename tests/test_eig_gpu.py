"""GPU parity of the batched eigenray search (ga_eig_search).

- spherical: the reference's own -eig_search answer for its example profile
  (three eigenrays, printed inclinations) and the oracle's full rows;
- 3D: the 512 seeded source-receiver pairs of the batched workload (base
  profile, bounces 0..3) against the oracle's serial search, stored in
  tests/golden/eig3d_base.npz by tests/golden/make_eig_golden.py;
- a receiver's eigenrays are the same alone and inside a batch;
- argument errors.
Bar: identical eigenray counts per (receiver, bounces); launch angles,
arrival position, travel time and turning height within 1e-6 relative;
TL and absorption within 0.01 dB.
"""
import os

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, EngineError, eig_params, prop_params
from parity import compare_eig

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "eig3d_base.npz")


def printed(x):
    return float("%.8g" % x)


def test_sph_eigenrays_known_answers(engine, oracle, toyatmo_units):
    from ga_oracle_py import GEOM_SPH, Medium
    aid = engine.load_atmo(*toyatmo_units)
    try:
        src, rcvr = (30.0, -100.0, 0.0), (30.25, -104.25)
        arr, lat, st = engine.eig_search(aid, prop_params(GA_GEOM_SPH, src=src), eig_params(bnc_min=0, bnc_max=1),
                                         [rcvr])
    finally:
        engine.release_atmo(aid)
    assert [printed(x) for x in arr["incl"]] == [34.559248, 4.1385204, 30.078055]
    assert list(arr["n_bounce"]) == [0, 1, 1]
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    ref, _ = m.eig_search(m.params(True), src, rcvr, bnc_min=0, bnc_max=1)
    ref = ref.copy()
    ref["launch_idx"] = 0
    fails, summ = compare_eig(arr, ref, True, src=src[:2])
    assert not fails, (fails, summ)
    assert lat.shape == (1,) and lat[0] > 0.0


@pytest.fixture(scope="module")
def base3d(engine, base_units):
    aid = engine.load_atmo(*base_units)
    yield aid
    engine.release_atmo(aid)


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="oracle eigenray fixture not generated")
def test_3d_batch_matches_oracle(engine, base3d):
    g = np.load(GOLDEN)
    n, bmax = int(g["n_pairs"]), int(g["bnc_max"])
    R = synth.eig_pairs()[:n]
    arr, lat, st = engine.eig_search(base3d, prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0)),
                                     eig_params(bnc_min=0, bnc_max=bmax), R)
    ref = g["eigenrays"]
    print(f"{n} pairs: {len(arr)} eigenrays (oracle {len(ref)}), {st.ray_steps} ray-steps, "
          f"{st.kernel_ms:.0f} ms kernels, p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms")
    fails, summ = compare_eig(arr, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)


def test_receiver_alone_equals_in_batch(engine, base3d):
    R = synth.eig_pairs()[:8]
    p = prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0))
    e = eig_params(bnc_min=0, bnc_max=2)
    batch, _, _ = engine.eig_search(base3d, p, e, R)
    for i in (1, 5):
        alone, _, _ = engine.eig_search(base3d, p, e, R[i:i + 1])
        mine = batch[batch["launch_idx"] == i]
        assert len(alone) == len(mine) and len(mine) > 0
        for f in ("incl", "az", "pos0", "pos1", "time", "turn_ht", "tl", "absorption"):
            assert np.array_equal(alone[f], mine[f]), f


def test_no_receivers(engine, base3d):
    arr, lat, st = engine.eig_search(base3d, prop_params(GA_GEOM_3D), eig_params(), np.zeros((0, 2)))
    assert len(arr) == 0 and st.ray_steps == 0


def test_argument_errors(engine, base3d):
    with pytest.raises(EngineError):
        engine.eig_search(base3d, prop_params(GA_GEOM_3D), eig_params(bnc_min=2, bnc_max=1), [(100.0, 0.0)])
    with pytest.raises(EngineError):
        engine.eig_search(base3d + 1000, prop_params(GA_GEOM_3D), eig_params(), [(100.0, 0.0)])
