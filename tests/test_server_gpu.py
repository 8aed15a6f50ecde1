"""The request server (ga_server_*): continuous batching of fans and
eigenray queries over resident atmospheres.

Each request's result in a mixed stream must equal the same request run
alone through the direct entry points (ga_fan_run, ga_eig_search), to the
parity bar: identical arrival / eigenray keys, range, time and turning
height within 1e-6 relative, TL and absorption within 0.01 dB.
"""
import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP, Server, eig_params, prop_params
from parity import compare, compare_eig

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def atmos(engine, base_units, pert_units):
    ids = {"3d": engine.load_atmo(*base_units), "sph": engine.load_atmo(*pert_units)}
    grid = synth.rngdep_grid(3, lat=np.linspace(30, 50, 6), lon=np.linspace(-110, -90, 6), dz=0.5, z_top=120.0)
    rid = engine.rngdep_load(*synth.rngdep_engine_units(*grid))
    engine.topo_load(rid, *synth.fractal_topo(5, lat=np.linspace(30, 50, 81), lon=np.linspace(-110, -90, 81)))
    yield dict(ids, rd=rid)
    for a in ids.values():
        engine.release_atmo(a)
    engine.rngdep_release(rid)


def _requests(atmos):
    R = synth.eig_pairs()
    fan3 = synth.fan_launches(0.5, 40.0, 1.5, 0.0, 300.0, 60.0)
    fans = synth.fan_launches(0.5, 40.0, 2.5, -90.0, 45.0, 135.0)
    p3 = prop_params(GA_GEOM_3D, bounces=2, src=(0.0, 0.0, 0.0), rng_max=800.0)
    ps = prop_params(GA_GEOM_SPH, bounces=2, src=(30.0, -100.0, 0.0), rng_max=800.0)
    pe = prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0))
    pr = prop_params(GA_GEOM_SPH_RNGDEP, bounces=2, src=(40.0, -100.0, 0.0))
    e = eig_params(bnc_min=0, bnc_max=1)
    return [
        ("fan", atmos["3d"], p3, None, fan3),
        ("eig", atmos["3d"], pe, e, R[[1, 5]]),
        ("fan", atmos["sph"], ps, None, fans),
        ("fan", atmos["rd"], pr, None, fans[::3]),    # range-dependent atmosphere over terrain
        ("eig", atmos["3d"], pe, e, R[[0, 7]]),      # joins the running search of the same source
        ("fan", atmos["3d"], p3, None, fan3[::2]),
    ]


def _alone(engine, kind, aid, p, e, data):
    if kind == "fan":
        arr, _ = engine.fan(aid, p, data)
        return arr
    arr, _, _ = engine.eig_search(aid, p, e, data)
    return arr


@pytest.mark.parametrize("merge,serial,slots", [(True, False, 4), (False, True, 1)], ids=["batched", "serial"])
def test_mixed_stream_equals_alone(engine, atmos, merge, serial, slots):
    reqs = _requests(atmos)
    alone = [_alone(engine, *r) for r in reqs]
    srv = Server(engine, slots=slots, merge=merge, serial=serial)
    try:
        tickets = []
        for kind, aid, p, e, data in reqs:
            tickets.append(srv.submit_fan(aid, p, data) if kind == "fan" else srv.submit_eig(aid, p, e, data))
        results = [srv.poll(t, timeout_ms=120000) for t in tickets]
        steps, launches, kms = srv.stats()
    finally:
        srv.close()
    assert steps > 0 and launches > 0
    for (kind, aid, p, e, data), ref, got in zip(reqs, alone, results):
        assert got is not None, "request did not finish"
        rows, res = got
        assert res.status == 0 and res.total_ms >= res.queue_ms >= 0.0
        sph = p.geometry in (GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP)
        if kind == "fan":
            fails, summ = compare(rows, ref, sph, src=(p.src[0], p.src[1]))
        else:
            fails, summ = compare_eig(rows, ref, sph, src=(p.src[0], p.src[1]))
        assert not fails, (kind, fails, summ)
        assert len(rows) == len(ref) and len(rows) > 0


def test_server_reports_request_errors(engine, atmos):
    srv = Server(engine, slots=2)
    try:
        t = srv.submit_fan(atmos["3d"] + 1000, prop_params(GA_GEOM_3D), [(10.0, 0.0)])
        with pytest.raises(Exception):
            srv.poll(t, timeout_ms=60000)
        t2 = srv.submit_eig(atmos["3d"], prop_params(GA_GEOM_3D), eig_params(), np.zeros((0, 2)))
        rows, res = srv.poll(t2, timeout_ms=60000)
        assert res.status == 0 and len(rows) == 0
    finally:
        srv.close()


def test_admission_by_footprint_queues_and_results_equal_alone(engine, atmos):
    """A budget below any two requests' footprints admits one request at a
    time: each later request is admitted only after the one before it has
    finished (submission-relative times from the results), and every result
    still equals the same request run alone.  With the default budget the
    same requests overlap (some admission precedes an earlier completion)."""
    import time
    reqs = _requests(atmos)
    alone = [_alone(engine, *r) for r in reqs]

    def run(budget):
        srv = Server(engine, slots=4, merge=True, budget_bytes=budget)
        try:
            t_sub, tickets = [], []
            t0 = time.perf_counter()
            for kind, aid, p, e, data in reqs:
                t_sub.append((time.perf_counter() - t0) * 1e3)
                tickets.append(srv.submit_fan(aid, p, data) if kind == "fan" else srv.submit_eig(aid, p, e, data))
            out = [srv.poll(t, timeout_ms=120000) for t in tickets]
        finally:
            srv.close()
        adm = [ts + r.queue_ms for ts, (_, r) in zip(t_sub, out)]
        done = [ts + r.total_ms for ts, (_, r) in zip(t_sub, out)]
        return out, adm, done

    out, adm, done = run(1.0e6)             # below every request's footprint
    for i in range(1, len(reqs)):
        assert adm[i] >= max(done[:i]) - 1.0, (i, adm[i], max(done[:i]))
    for (kind, aid, p, e, data), ref, (rows, res) in zip(reqs, alone, out):
        assert res.status == 0
        sph = p.geometry in (GA_GEOM_SPH, GA_GEOM_SPH_RNGDEP)
        fails, summ = (compare if kind == "fan" else compare_eig)(rows, ref, sph, src=(p.src[0], p.src[1]))
        assert not fails, (kind, fails, summ)
        assert len(rows) == len(ref) and len(rows) > 0
    _, adm2, done2 = run(16e9)
    assert any(adm2[i] < max(done2[:i]) - 1.0 for i in range(1, len(reqs)))
