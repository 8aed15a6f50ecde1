"""GPU parity of the batched eigenray search (ga_eig_search).

- spherical: the reference's own -eig_search answer for its example profile
  (three eigenrays, printed inclinations) and the oracle's full rows;
- 3D: the 512 seeded source-receiver pairs of the batched workload (base
  profile, bounces 0..3) against the oracle's serial search, stored in
  tests/golden/eig3d_base.npz by tests/golden/make_eig_golden.py.  The
  Levenberg-Marquardt iteration is chaotic near some starts: a 1e-13
  relative change of the start changes whether it accepts
  (tests/golden/make_eig_sensitivity.py classifies every start of the batch),
  and the same oracle compiled with FMA contraction (eig3d_base_fma.npz)
  changes the eigenray count of 5 groups, all of them holding such a start.
  So every GPU row is checked tightly against the oracle's accepted-eigenray
  trace at the GPU's own launch angles (go_eig_row_3d), and the eigenray
  count of every (receiver, bounces) group whose starts are all determined
  at FP64 precision must equal the serial search's;
- a receiver's eigenrays are the same alone and inside a batch;
- argument errors.
Bar (rows): arrival position, travel time and turning height within 1e-6
relative; TL and absorption within 0.01 dB; for the spherical known answer
also identical counts and launch angles within 1e-6 relative.
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
GOLDEN_FMA = os.path.join(HERE, "golden", "eig3d_base_fma.npz")
SENS = os.path.join(HERE, "golden", "eig3d_sensitivity.npz")


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


def _groups(arr):
    out = {}
    for r in arr:
        out.setdefault((int(r["launch_idx"]), int(r["n_bounce"])), []).append(r)
    return out


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="oracle eigenray fixture not generated")
def test_3d_batch_matches_oracle(engine, base3d, oracle, base_units):
    from ga_oracle_py import GEOM_3D, Medium
    g = np.load(GOLDEN)
    g_fma = np.load(GOLDEN_FMA)
    n, bmax = int(g["n_pairs"]), int(g["bnc_max"])
    R = synth.eig_pairs()[:n]
    arr, lat, st = engine.eig_search(base3d, prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0)),
                                     eig_params(bnc_min=0, bnc_max=bmax), R)
    ref, ref_fma = g["eigenrays"], g_fma["eigenrays"]
    print(f"{n} pairs: {len(arr)} eigenrays (oracle {len(ref)}, FMA build {len(ref_fma)}), {st.ray_steps} "
          f"ray-steps, {st.kernel_ms:.0f} ms kernels, p50 {np.percentile(lat, 50):.0f} ms, "
          f"p99 {np.percentile(lat, 99):.0f} ms")

    # 1. rows: the oracle's accepted-eigenray trace at the GPU's launch angles
    m = Medium(oracle, GEOM_3D, *base_units)
    p = m.params(True)
    for r in arr:
        o = m.eig_row(p, (0.0, 0.0, 0.0), r["incl"], r["az"], int(r["n_bounce"]))
        rng = max(np.hypot(o["pos0"], o["pos1"]), 1.0)
        key = (int(r["launch_idx"]), int(r["n_bounce"]))
        assert np.hypot(r["pos0"] - o["pos0"], r["pos1"] - o["pos1"]) <= 1e-6 * rng, (key, "pos")
        for f in ("time", "turn_ht"):
            assert abs(r[f] - o[f]) <= 1e-6 * abs(o[f]), (key, f, r[f], o[f])
        for f in ("tl", "absorption"):
            assert abs(r[f] - o[f]) <= 0.01, (key, f, r[f], o[f])
        for f in ("incl_arr", "back_az"):
            assert abs(r[f] - o[f]) <= 1e-6, (key, f, r[f], o[f])

    # 2. acceptance: within the 0.1 km tolerance of the receiver, unless the
    #    oracle's rows of that group leave it too (the reference's re-trace
    #    after an accepted search can end elsewhere; both builds share it)
    G, O, F = _groups(arr), _groups(ref), _groups(ref_fma)

    def miss(rows, i):
        return max(np.hypot(x["pos0"] - R[i, 0], x["pos1"] - R[i, 1]) for x in rows)
    for k, rows in G.items():
        if miss(rows, k[0]) > 0.1:
            assert k in O and miss(O[k], k[0]) > 0.1, (k, miss(rows, k[0]))

    # 3. search: eigenray counts per (receiver, bounces) identical to the
    #    serial search's in every group whose starts are all determined at
    #    FP64 precision (eig3d_sensitivity.npz); groups holding a start that
    #    a 1e-13 relative change moves may differ, as the FMA build of the
    #    oracle itself does.  Where the counts agree, launch angles within
    #    0.1 degree (the LM stops anywhere inside the tolerance disc)
    sens = np.load(SENS)
    ung = {(int(s[0]), int(s[1])) for s, u in zip(sens["start"], sens["unstable"]) if u}
    keys = set(G) | set(O) | set(F)
    n_ref_flips = sum(1 for k in keys if len(O.get(k, [])) != len(F.get(k, [])))
    flips = sorted(k for k in keys if len(G.get(k, [])) != len(O.get(k, [])))
    hard = [k for k in flips if k not in ung]
    print(f"count differences: GPU vs oracle {len(flips)} {flips} (all in the {len(ung)} groups with a start not "
          f"determined at FP64 precision: {not hard}), oracle vs FMA build {n_ref_flips}")
    assert not hard, hard
    for k in set(G) & set(O):
        if len(G[k]) != len(O[k]):
            continue
        for a, b in zip(sorted(G[k], key=lambda x: x["incl"]), sorted(O[k], key=lambda x: x["incl_launch"])):
            assert abs(a["incl"] - b["incl_launch"]) <= 0.1 and abs(a["az"] - b["az_launch"]) <= 0.1, k


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


DIRECT = [
    # (incl_est, az_est, bounces): ToyAtmo, source 30 N 100 W, receiver 30.25 N 104.25 W
    (4.0, float("nan"), 1),
    (34.0, float("nan"), 0),
    (30.5, -84.9, 1),
    (10.0, float("nan"), 0),    # does not converge within 25 iterations (oracle alike)
]


@pytest.mark.parametrize("incl,az,nb", DIRECT, ids=["b1_4deg", "b0_34deg", "b1_30.5deg_az", "b0_10deg_none"])
def test_eig_direct_matches_oracle(engine, oracle, toyatmo_units, incl, az, nb):
    """-eig_direct (ga_eig_direct) against the oracle's go_eig_direct from the
    same estimate: same outcome; launch angles within 1e-6 relative; the row
    and its eigenray-0.dat rows against the oracle's accepted trace at the
    GPU's launch angles (1e-6 relative, 0.01 dB)."""
    from ga_oracle_py import GEOM_SPH, Medium
    from parity import compare_paths
    src, rcvr = (30.0, -100.0, 0.0), (30.25, -104.25)
    aid = engine.load_atmo(*toyatmo_units)
    row, pts, st = engine.eig_direct(aid, prop_params(GA_GEOM_SPH, src=src), eig_params(), rcvr, incl, az, nb)
    m = Medium(oracle, GEOM_SPH, *toyatmo_units)
    p = m.params(True)
    ref, _ = m.eig_direct(p, src, rcvr, incl, az, nb)
    engine.release_atmo(aid)
    assert (row is None) == (ref is None)
    if ref is None:
        assert len(pts) == 0
        return
    assert abs(row["incl"] - ref["incl_launch"]) <= 1e-6 * abs(ref["incl_launch"])
    assert abs(row["az"] - ref["az_launch"]) <= 1e-6 * 360.0
    rp, rrow = m.eig_path(p, src, row["incl"], row["az"], nb)
    for f in ("time", "turn_ht"):
        assert abs(row[f] - rrow[f]) <= 1e-6 * abs(rrow[f]), f
    for f in ("tl", "absorption"):
        assert abs(row[f] - rrow[f]) <= 0.01, f
    pf, ps = compare_paths(pts, rp, True, src=src)
    assert not pf, (pf, ps)
