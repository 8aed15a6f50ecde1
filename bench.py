"""Benchmark: ray-steps/s of the 162k-ray spherical fan (FP64) on one MI355X.

Workload (one "step" = one full pass of the fan through the engine):
  perturbed stratified profile (base + seeded 5 m/s / 2 K Gaussian noise,
  5 km correlation, 1501 levels 0-150 km), spherical geometry, source at
  30 N 0 E on the ground, inclinations 0.1..45.0 step 0.1 x azimuths 0..359
  step 1 (162,000 rays), 5 bounces, 1000 km range, 18 equations (TL on),
  0.1 Hz Sutherland-Bass absorption.  Inputs are resident in HBM before the
  timed region; the pass enqueues reset, trace and arrival kernels on torch's
  current stream, which the HIP events bracket.

A second leg times the batched eigenray search: 512 seeded source-receiver
pairs (range U(100, 800) km, azimuth U(0, 360) deg) on the base stratified
profile in the 3D geometry, source at the origin on the ground, bounces 0..3,
0.1 km tolerance, one call; it reports per-receiver completion latency
(p50/p99) beside the batch time, and the oracle's serial search on one pair
as its CPU baseline.

A third leg replays a served request stream: 16 resident atmospheres (8 in
the 3D geometry, 8 spherical, each a seeded perturbation of the base
profile), Poisson arrivals at a fixed rate, a seeded mix of fans (1k-200k
rays, log-uniform) and eigenray queries (1-64 receivers, bounces 0..3).
The same trace runs through the request server twice: continuous batching
(4 launch slots, eigenray queries merged into running searches) and one
request at a time; both report p50/p99 request latency and ray-steps/s.

A fourth leg times the range-dependent spherical fan over terrain: 11 x 11
profiles (2 degree nodes over 30..50 N, 110..90 W; the base profile plus a
seeded random field, +-15 m/s u and v, +-5 K T, 500 km correlation; 1501
levels to 150 km), 401 x 401 fractal terrain (0.05 degree, 0..2 km), source
at the region centre on the ground, inclinations 0.2..40.0 step 0.2 x
azimuths 0..359 step 1 (72,000 rays), 4 bounces, 18 equations; the oracle on
a few strided rays is its CPU baseline.

Prints one JSON line.  --gpus N with N > 1 exits: this build runs one GPU.
"""
from __future__ import annotations

import argparse
import math
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_STEP = 5690.9          # algorithmic FP64 flops per ray-step (oracle/opcount.cpp, 400-ray sample)
# HBM bytes of one pass (all slice_kernel dispatches) on this workload, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, each scaled by what the counter
# reads for a 256 MiB device copy (FETCH_SIZE reads half the copy's bytes:
# factor 2.00; WRITE_SIZE factor 1.00): 22.3 GB read + 20.0 GB written.  The
# ray records' round trip through the ring is 232 B each way per ray and
# slice, 12.9 GB each way over the pass's 5.56e7 ray-slices.
TRAFFIC_BYTES = 10909465.1875 * 2.00 * 1024 + 19569994.84375 * 1.00 * 1024
# FP64 flops the hardware executes per ray-step on this workload (rocprofv3
# SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 over one pass, FMA = 2, x 64 lanes / ray-steps).
# Below F_STEP: the engine evaluates the smooth Sutherland-Bass terms from a
# per-frequency table of piecewise polynomials instead of the formula.
F_STEP_EXEC = 3305.2
TRANS_STEP = 167.8       # transcendental calls per ray-step (same count, not included in F_STEP)
PEAK_FP64 = 78.6e12      # MI355X FP64 vector / matrix peak
METRIC = "ray-steps/s on 162k-ray sph fan (FP64)"


def workload():
    from infraga_amd import synth
    units = synth.to_engine_units(*synth.perturbed_profile())
    L = synth.fan_launches(0.1, 45.0, 0.1, 0.0, 359.0, 1.0)
    return units, L


def host_cpu():
    """Cores this process may use (capped at the 16-core share of a one-GPU
    box), nproc, CPU model and sockets of the host."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model, sockets = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"cores": max(1, min(avail, 16)), "nproc": os.cpu_count(), "affinity": avail, "cpu_model": model,
            "sockets": max(1, len(sockets))}


def _pool_map(fn, items, cores):
    """fn over items on `cores` worker processes (spawned: this process holds
    the GPU); returns the results and the wall time of the compute phase
    (first worker start to last worker end)."""
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(cores) as pool:
        res = pool.map(fn, items, chunksize=1)
    t_start = min(r["t0"] for r in res)
    t_end = max(r["t1"] for r in res)
    return res, t_end - t_start


def _fan_worker(job):
    """Oracle restatement of the spherical fan on a chunk of the launch list."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import GEOM_SPH, Medium, Oracle
    units, idx, L = job
    m = Medium(Oracle(), GEOM_SPH, *units)
    p = m.params(True)
    t0 = time.time()
    arr, steps = m.fan(L, p, bounces=5, src=(30.0, 0.0, 0.0))
    t1 = time.time()
    arr = arr.copy()
    arr["launch_idx"] = np.asarray(idx)[arr["launch_idx"]]
    return {"arr": arr, "steps": steps, "t0": t0, "t1": t1}


def cpu_baseline(units, L, cores, per_core=480):
    """The oracle restatement on all host cores (one process per core) over
    an evenly strided sample of the same launch list (per_core rays per
    core, about 12 s of CPU each), beside the single-core rate.  Returns the
    baseline and the oracle's arrivals of the sampled rays."""
    n = per_core * cores
    stride = max(1, len(L) // n)
    idx = np.arange(0, len(L), stride)[:n]
    chunks = [(units, idx[k::cores], L[idx[k::cores]]) for k in range(cores)]
    res, wall = _pool_map(_fan_worker, chunks, cores)
    steps = sum(r["steps"] for r in res)
    per_core_rates = [r["steps"] / (r["t1"] - r["t0"]) for r in res]
    arr = np.concatenate([r["arr"] for r in res])
    base = {"value": steps / wall, "unit": "ray-steps/s", "cores": cores, "kind": "port",
            "per_core": float(np.mean(per_core_rates)),
            "sample": f"{len(idx)} rays (every {stride}th of the 162k launch list) on {cores} processes, "
                      f"{steps} ray-steps in {wall:.1f} s wall; oracle restatement, one thread per process"}
    return base, idx, arr


def sample_parity(gpu_rows, ref_rows, geom_sph, src):
    """GPU rows of the sampled rays against the oracle's at the parity bar
    (keys exact; range, time, turning height 1e-6 relative; TL, absorption
    0.01 dB)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import compare
    fails, summ = compare(gpu_rows, ref_rows, geom_sph, src=src)
    worst = {k: float(v) for k, v in summ["worst"].items()}
    return {"ok": not fails, "arrivals_gpu": int(len(gpu_rows)), "arrivals_oracle": int(len(ref_rows)),
            "key_mismatches": len(fails.get("missing", [])) + len(fails.get("extra", [])), "worst": worst,
            "failures": {k: str(v)[:200] for k, v in fails.items()}}


def _eig_worker(i):
    """The oracle's serial search for pair i (one process)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import GEOM_3D, Medium, Oracle
    from infraga_amd import synth
    m = Medium(Oracle(), GEOM_3D, *synth.to_engine_units(*synth.base_profile()))
    R = synth.eig_pairs()
    t0 = time.time()
    arr, steps = m.eig_search(m.params(True), (0.0, 0.0, 0.0), R[i], bnc_min=0, bnc_max=3)
    t1 = time.time()
    arr = arr.copy()
    arr["launch_idx"] = i
    return {"arr": arr, "steps": steps, "t0": t0, "t1": t1}


def eig_cpu_baseline(cores):
    """The oracle's serial search on all host cores, one pair per process
    (pairs 0, 32, 64, ... of the batch)."""
    pairs = list(range(0, 512, max(1, 512 // cores)))[:cores]
    res, wall = _pool_map(_eig_worker, pairs, cores)
    per_pair = [r["t1"] - r["t0"] for r in res]
    arr = np.concatenate([r["arr"] for r in res])
    return {"value": len(pairs) / wall, "unit": "pairs/s", "cores": cores, "kind": "port",
            "s_per_pair_mean": float(np.mean(per_pair)), "s_per_pair_max": float(np.max(per_pair)),
            "eigenrays": int(len(arr)),
            "sample": f"pairs {pairs[0]}..{pairs[-1]} step {pairs[1] - pairs[0] if len(pairs) > 1 else 0} "
                      f"({len(pairs)} pairs, bounces 0..3) on {cores} processes, {wall:.1f} s wall; "
                      f"oracle restatement"}, pairs, arr


def eig_leg(eng, with_cpu, cores=1):
    from infraga_amd import synth
    from infraga_amd.engine import GA_GEOM_3D, eig_params, prop_params
    aid = eng.load_atmo(*synth.to_engine_units(*synth.base_profile()))
    p = prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0))
    e = eig_params(bnc_min=0, bnc_max=3, tolerance=0.1)
    R = synth.eig_pairs()
    eng.eig_search(aid, p, e, R[:4])                 # warm-up (kernel set-up, buffers)
    t0 = time.perf_counter()
    arr, lat, st = eng.eig_search(aid, p, e, R)
    wall = time.perf_counter() - t0
    eng.release_atmo(aid)
    out = {"workload": "512 pairs, 100-800 km, az 0-360, base profile 3D, bounces 0..3, tolerance 0.1 km",
           "wall_ms": wall * 1e3, "kernel_ms": st.kernel_ms, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()), "eigenrays": int(len(arr)),
           "ray_steps": int(st.ray_steps), "traces": int(st.n_rays), "kernel_launches": int(st.n_launches),
           "ray_steps_per_s": st.ray_steps / wall}
    if with_cpu:
        base, pairs, ref = eig_cpu_baseline(cores)
        out["cpu_baseline"] = base
        # per (pair, bounces) eigenray counts of the GPU batch against the
        # oracle's serial search on the sampled pairs
        sel = arr[np.isin(arr["launch_idx"], pairs)]
        cg, co = {}, {}
        for r in sel:
            k = (int(r["launch_idx"]), int(r["n_bounce"]))
            cg[k] = cg.get(k, 0) + 1
        for r in ref:
            k = (int(r["launch_idx"]), int(r["n_bounce"]))
            co[k] = co.get(k, 0) + 1
        keys = set(cg) | set(co)
        out["sample_parity"] = {"pairs": len(pairs), "groups": len(keys),
                                "groups_count_equal": sum(1 for k in keys if cg.get(k, 0) == co.get(k, 0)),
                                "count_differences": sorted([list(k) for k in keys if cg.get(k, 0) != co.get(k, 0)])}
    return out


def rngdep_workload():
    from infraga_amd import synth
    units = synth.rngdep_engine_units(*synth.rngdep_grid())
    topo = synth.fractal_topo()
    L = synth.fan_launches(0.2, 40.0, 0.2, 0.0, 359.0, 1.0)
    return units, topo, L


def _rd_worker(job):
    """Oracle range-dependent fan (spherical or 3D) on a chunk of rays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import Oracle, Rd3Medium, RdMedium
    cart, units, topo, idx, L = job
    m = (Rd3Medium if cart else RdMedium)(Oracle(), *units)
    m.set_topo(*topo)
    p = m.params(True)
    t0 = time.time()
    if cart:
        arr, _, steps = m.fan3(L, p, bounces=4)
    else:
        arr, steps = m.fan(L, p, bounces=4, src=None)
    t1 = time.time()
    arr = arr.copy()
    arr["launch_idx"] = np.asarray(idx)[arr["launch_idx"]]
    return {"arr": arr, "steps": steps, "t0": t0, "t1": t1}


def rd_cpu_baseline(cart, units, topo, L, cores, per_core=6):
    """The oracle's range-dependent fan on all host cores over evenly strided
    rays of the same launch list (per_core rays per core, about 10 s each)."""
    n = per_core * cores
    stride = max(1, len(L) // n)
    idx = np.arange(0, len(L), stride)[:n]
    chunks = [(cart, units, topo, idx[k::cores], L[idx[k::cores]]) for k in range(cores)]
    res, wall = _pool_map(_rd_worker, chunks, cores)
    steps = sum(r["steps"] for r in res)
    arr = np.concatenate([r["arr"] for r in res])
    per_core_rates = [r["steps"] / (r["t1"] - r["t0"]) for r in res]
    return {"value": steps / wall, "unit": "ray-steps/s", "cores": cores, "kind": "port",
            "per_core": float(np.mean(per_core_rates)),
            "sample": f"{len(idx)} rays (every {stride}th launch) on {cores} processes, {steps} ray-steps in "
                      f"{wall:.1f} s wall; oracle restatement"}, idx, arr


def rngdep_leg(eng, with_cpu, cores=1):
    from infraga_amd.engine import GA_GEOM_SPH_RNGDEP, prop_params
    units, topo, L = rngdep_workload()
    rid = eng.rngdep_load(*units)
    eng.topo_load(rid, *topo)
    p = prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=1, bounces=4)
    eng.fan(rid, p, L[::997])                        # warm-up (buffers, kernel load)
    t0 = time.perf_counter()
    arr, st = eng.fan(rid, p, L)
    wall = time.perf_counter() - t0
    eng.rngdep_release(rid)
    out = {"workload": "sph range-dependent: 11 x 11 profiles x 1501 levels, 401 x 401 terrain 0-2 km, "
                       "200 incl x 360 az = 72000 rays, 4 bounces, 18 eq",
           "rays": int(len(L)), "ray_steps": int(st.ray_steps), "arrivals": int(len(arr)),
           "kernel_ms": st.kernel_ms, "wall_ms": wall * 1e3, "ray_steps_per_s": st.ray_steps / wall,
           "reflections": int((arr["n_bounce"] > 0).sum())}
    if with_cpu:
        base, idx, ref = rd_cpu_baseline(False, units, topo, L, cores)
        out["cpu_baseline"] = base
        lat0 = (units[0][0] + units[0][-1]) / 2.0
        lon0 = (units[1][0] + units[1][-1]) / 2.0
        out["sample_parity"] = sample_parity(arr[np.isin(arr["launch_idx"], idx)], ref, True, (lat0, lon0))
    return out


def rngdep3d_leg(eng, with_cpu, cores=1):
    """3D Cartesian range-dependent fan (infraga-3d-rngdep): the grid and
    terrain of the spherical range-dependent leg with node coordinates read
    as km over +-800 km."""
    from infraga_amd import synth
    from infraga_amd.engine import GA_GEOM_3D_RNGDEP, prop_params
    _, _, z, c, u, v, rho = synth.rngdep_engine_units(*synth.rngdep_grid())
    x = np.linspace(-800.0, 800.0, c.shape[0])
    units = (x, np.linspace(-800.0, 800.0, c.shape[1]), z, c, u, v, rho)
    t = synth.fractal_topo()
    topo = (np.linspace(-820.0, 820.0, len(t[0])), np.linspace(-820.0, 820.0, len(t[1])), t[2])
    L = synth.fan_launches(0.2, 40.0, 0.2, 0.0, 359.0, 1.0)
    rid = eng.rngdep_load_xy(*units)
    eng.topo_load(rid, *topo)
    p = prop_params(GA_GEOM_3D_RNGDEP, calc_amp=1, bounces=4)
    eng.fan(rid, p, L[::997])                        # warm-up (buffers, kernel load)
    t0 = time.perf_counter()
    arr, st = eng.fan(rid, p, L)
    wall = time.perf_counter() - t0
    eng.rngdep_release(rid)
    out = {"workload": f"3D range-dependent: {len(x)} x {len(units[1])} profiles x {len(z)} levels over +-800 km, "
                       f"{len(topo[0])} x {len(topo[1])} terrain 0-2 km, 200 incl x 360 az = {len(L)} rays, 4 bounces, 18 eq",
           "rays": int(len(L)), "ray_steps": int(st.ray_steps), "arrivals": int(len(arr)),
           "kernel_ms": st.kernel_ms, "wall_ms": wall * 1e3, "ray_steps_per_s": st.ray_steps / wall}
    if with_cpu:
        base, idx, ref = rd_cpu_baseline(True, units, topo, L, cores)
        out["cpu_baseline"] = base
        out["sample_parity"] = sample_parity(arr[np.isin(arr["launch_idx"], idx)], ref, False, (0.0, 0.0))
    return out


def serve_trace(n_req=24, qps=1.5, seed=11):
    """Seeded request trace: (arrival_s, kind, atmosphere index, size)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / qps, n_req))
    out = []
    for i in range(n_req):
        atm = int(rng.integers(0, 16))
        if rng.random() < 0.5:
            out.append((float(t[i]), "fan", atm, int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))))
        else:
            out.append((float(t[i]), "eig", atm, int(np.clip(np.exp(rng.uniform(0.0, np.log(64.0))), 1, 64))))
    return out


def serve_leg(eng, trace):
    from infraga_amd import synth
    from infraga_amd.engine import GA_GEOM_3D, GA_GEOM_SPH, Server, eig_params, prop_params
    atm = [eng.load_atmo(*synth.to_engine_units(*synth.perturbed_profile(seed=1000 + k))) for k in range(16)]
    geom = [GA_GEOM_3D] * 8 + [GA_GEOM_SPH] * 8
    rng = np.random.Generator(np.random.PCG64(5))
    reqs = []
    for (t, kind, a, size) in trace:
        g = geom[a]
        if kind == "fan":
            n_az = max(1, int(round(np.sqrt(size / 2.0))))
            n_in = max(1, size // n_az)
            L = synth.fan_launches(0.5, 0.5 + 44.5 * (n_in - 1) / max(n_in - 1, 1), 44.5 / max(n_in - 1, 1),
                                   0.0, 360.0 * (n_az - 1) / n_az, 360.0 / n_az)
            p = (prop_params(GA_GEOM_3D, bounces=2, src=(0.0, 0.0, 0.0), rng_max=800.0) if g == GA_GEOM_3D
                 else prop_params(GA_GEOM_SPH, bounces=5, src=(30.0, 0.0, 0.0), rng_max=1000.0))
            reqs.append((t, "fan", atm[a], p, None, L))
        else:
            r = rng.uniform(100.0, 800.0, size)
            az = rng.uniform(0.0, 360.0, size)
            if g == GA_GEOM_3D:
                R = np.stack([r * np.sin(np.radians(az)), r * np.cos(np.radians(az))], axis=1)
                p = prop_params(GA_GEOM_3D, src=(0.0, 0.0, 0.0))
            else:
                d = r / 6370.0
                la0, lo0 = np.radians(30.0), 0.0
                la = np.arcsin(np.sin(la0) * np.cos(d) + np.cos(la0) * np.sin(d) * np.cos(np.radians(az)))
                lo = lo0 + np.arctan2(np.sin(np.radians(az)) * np.sin(d) * np.cos(la0),
                                      np.cos(d) - np.sin(la0) * np.sin(la))
                R = np.degrees(np.stack([la, lo], axis=1))
                p = prop_params(GA_GEOM_SPH, src=(30.0, 0.0, 0.0))
            reqs.append((t, "eig", atm[a], p, eig_params(bnc_min=0, bnc_max=3), R))
    out = {"workload": f"{len(trace)} requests, Poisson {len(trace) / trace[-1][0]:.2f}/s over 16 resident "
                       "atmospheres (8 3D, 8 sph); fans 1k-200k rays, eigenray queries 1-64 receivers"}
    for mode, kw in (("batched", dict(slots=4, merge=True, serial=False)),
                     ("one_request_at_a_time", dict(slots=1, merge=False, serial=True))):
        srv = Server(eng, **kw)
        t0 = time.perf_counter()
        tickets = []
        for (t, kind, a, p, e, data) in reqs:
            dt = t - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append((kind, srv.submit_fan(a, p, data) if kind == "fan" else srv.submit_eig(a, p, e, data)))
        lat = {"fan": [], "eig": []}
        for kind, tk in tickets:
            rows, res = srv.poll(tk, timeout_ms=-1)
            lat[kind].append(res.total_ms)
        wall = time.perf_counter() - t0
        steps, launches, kms = srv.stats()
        srv.close()
        allv = lat["fan"] + lat["eig"]
        out[mode] = {"wall_s": wall, "ray_steps": int(steps), "ray_steps_per_s": steps / wall,
                     "p50_ms": float(np.percentile(allv, 50)), "p99_ms": float(np.percentile(allv, 99)),
                     "eig_p50_ms": float(np.percentile(lat["eig"], 50)) if lat["eig"] else None,
                     "eig_p99_ms": float(np.percentile(lat["eig"], 99)) if lat["eig"] else None,
                     "fan_p50_ms": float(np.percentile(lat["fan"], 50)) if lat["fan"] else None,
                     "fan_p99_ms": float(np.percentile(lat["fan"], 99)) if lat["fan"] else None,
                     "kernel_launches": int(launches)}
    for a in atm:
        eng.release_atmo(a)
    return out


WV_LEN = 8192


def wvfrm_workload():
    """The reference's own -wnl_wvfrm example (ToyAtmo.met, source 30 N
    100 W, inclination 4.1314876, impulse p0 = 500 Pa, t0 = 0.1 s, 8192
    samples, 1 bounce) swept over 256 azimuths within +-40 degrees of its
    azimuth -84.969455: one waveform per CU."""
    from infraga_amd import synth, waveforms
    cols = np.load(os.path.join(ROOT, "tests", "golden", "toyatmo.npz"))["zTuvdp"]
    units = synth.from_zTuvdp(cols)
    az = -84.969455 + np.linspace(-40.0, 40.0, 256)
    L = np.stack([np.full(len(az), 4.1314876), az], axis=1)
    t, ps = waveforms.build("impulse", p0=500.0, t0=0.1, alpha=4.0, length=WV_LEN)
    return units, L, t, ps


def _wv_worker(job):
    """The oracle's run_wnl_wvfrm for one launch (one process)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import GEOM_SPH, Medium, Oracle, WvParams, wv_run
    units, i, incl, az, t, ps = job
    m = Medium(Oracle(), GEOM_SPH, *units)
    t0 = time.time()
    r = wv_run(m, m.params(True), WvParams.default(len=WV_LEN), (30.0, -100.0, 0.0), incl, az, 1, t, ps)
    t1 = time.time()
    return {"i": i, "r": r, "t0": t0, "t1": t1}


def wvfrm_leg(eng, with_cpu, cores=1):
    """Weakly non-linear waveforms (ga_wnl_wvfrm) on the example's sweep:
    trace and evolution times, evolution steps per second, the
    evolution kernel's LDS roofline; the oracle on all host cores over one
    waveform per core, with the GPU's waveforms checked against it."""
    from infraga_amd.engine import GA_GEOM_SPH, prop_params, wvfrm_params
    units, L, t, ps = wvfrm_workload()
    aid = eng.load_atmo(*units)
    p = prop_params(GA_GEOM_SPH, bounces=1, src=(30.0, -100.0, 0.0))
    w = wvfrm_params(WV_LEN)
    eng.wnl_wvfrm(aid, p, w, L[:2], t, ps)           # warm-up (kernel load, buffers)
    t0 = time.perf_counter()
    g = eng.wnl_wvfrm(aid, p, w, L, t, ps)
    wall = time.perf_counter() - t0
    eng.release_atmo(aid)
    st = g["stats"]
    lg = int(round(math.log2(WV_LEN)))
    # LDS bytes per evolution step: three transforms of len complex doubles
    # in passes of three fused radix-2 stages, each pass reading and writing
    # every element (32 len bytes), plus the step's other LDS passes
    # (u and its neighbours, the two scatters, bin reads, the window pass:
    # 112 len bytes)
    lds_step = 3 * 32 * WV_LEN * ((lg + 2) // 3) + 112 * WV_LEN
    ach = lds_step * st.wv_steps / (st.evolve_ms * 1e-3) / 1e9
    peak = 256 * 128 * 2.4e9 / 1e9                   # 256 CUs x 128 B/clk x 2.4 GHz
    out = {"workload": f"{len(L)} waveforms of {WV_LEN} samples: ToyAtmo, source 30N 100W, inclination 4.1314876, "
                       "azimuth -84.969455 +- 40 deg, impulse p0 500 Pa t0 0.1 s, 1 bounce",
           "waveforms": int(len(L)), "returned": int((g["status"] == 1).sum()), "wv_steps": int(st.wv_steps),
           "ray_steps": int(st.ray_steps), "trace_ms": st.trace_ms, "evolve_ms": st.evolve_ms, "wall_ms": wall * 1e3,
           "waveforms_per_s": len(L) / wall, "wv_steps_per_s": st.wv_steps / (st.evolve_ms * 1e-3),
           "roofline": {"kernel": "wv_evolve_kernel", "bound": "lds", "achieved": ach, "peak": peak, "unit": "GB/s",
                        "frac": ach / peak, "traffic": None,
                        "note": f"{lds_step} LDS bytes per step (3 transforms x {(lg + 2) // 3} passes x 32 B x len "
                                f"+ 112 B x len); peak 128 B/clk per CU"}}
    if with_cpu:
        n = max(1, min(cores, len(L)))
        pick = list(range(0, len(L), max(1, len(L) // n)))[:n]
        jobs = [(units, i, float(L[i, 0]), float(L[i, 1]), t, ps) for i in pick]
        res, cwall = _pool_map(_wv_worker, jobs, n)
        steps = sum(r["r"]["steps"] for r in res)
        out["cpu_baseline"] = {"value": steps / cwall, "unit": "wv-steps/s", "cores": n, "kind": "port",
                               "sample": f"{len(pick)} waveforms (every {len(L) // n}th azimuth) on {n} processes, "
                                         f"{steps} steps in {cwall:.1f} s wall"}
        worst, same = 0.0, 0
        for r in res:
            i, o = r["i"], r["r"]
            peak_p = np.abs(o["p_out"]).max()
            worst = max(worst, float(np.abs(g["p_out"][i] - o["p_out"]).max() / peak_p))
            same += int(g["steps"][i] == o["steps"] and g["status"][i] == (1 if o["ok"] else 0))
        out["sample_parity"] = {"waveforms": len(res), "steps_and_status_equal": same,
                                "max_rel_pressure_diff": worst, "ok": bool(same == len(res) and worst <= 1e-6)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eig", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-rngdep", action="store_true")
    ap.add_argument("--no-wvfrm", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the scheduler autotuner (library defaults)")
    ap.add_argument("--one-pass", action="store_true",
                    help="one untimed fan pass (for hardware-counter runs); no JSON line")
    ap.add_argument("--leg", choices=("fan", "rngdep"), default="fan",
                    help="with --one-pass: the spherical fan or the range-dependent fan")
    ap.add_argument("--calibrate", action="store_true",
                    help="with --one-pass: copy a 256 MiB buffer first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "detail": "rays are independent; a multi-GPU run would shard the launch list "
                                    "(see DESIGN.md), which this build does not launch"}))
        sys.exit(2)

    import torch
    from infraga_amd.engine import ARRIVAL_DTYPE, GA_GEOM_SPH, Engine, prop_params

    units, L = workload()
    n = len(L)
    bounces = 5
    eng = Engine(0)
    aid = eng.load_atmo(*units)
    p = prop_params(GA_GEOM_SPH, calc_amp=1, bounces=bounces, src=(30.0, 0.0, 0.0), rng_max=1000.0)
    dev = torch.device("cuda:0")
    d_l = torch.from_numpy(np.ascontiguousarray(L)).to(dev)
    d_out = torch.empty(n * (bounces + 1) * ARRIVAL_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d_cnt = torch.empty(n, dtype=torch.int32, device=dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def one_pass():
        eng.fan_device(aid, p, d_l.data_ptr(), n, d_out.data_ptr(), d_cnt.data_ptr(), d_st.data_ptr(), stream)

    if args.one_pass and args.leg == "rngdep":
        from infraga_amd.engine import GA_GEOM_SPH_RNGDEP
        units_r, topo_r, L_r = rngdep_workload()
        rid = eng.rngdep_load(*units_r)
        eng.topo_load(rid, *topo_r)
        p_r = prop_params(GA_GEOM_SPH_RNGDEP, calc_amp=1, bounces=4)
        if args.calibrate:
            a = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda:0")
            b = torch.empty_like(a)
            torch.cuda.synchronize()
            b.copy_(a)
            torch.cuda.synchronize()
        _, st = eng.fan(rid, p_r, L_r)
        print(f"one range-dependent pass: {st.ray_steps} ray-steps, {st.kernel_ms:.0f} ms", flush=True)
        eng.close()
        return

    if args.one_pass:
        if args.calibrate:
            a = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device=dev)
            b = torch.empty_like(a)
            torch.cuda.synchronize()
            b.copy_(a)
        torch.cuda.synchronize()
        one_pass()
        torch.cuda.synchronize()
        print(f"one pass: {int(d_st.to(torch.int64).sum().item())} ray-steps", flush=True)
        eng.close()
        return

    # the scheduler's steps per slice and slices per graph, picked by the
    # library's autotuner on this launch list before the timed region
    tuning = eng.fan_autotune(aid, p, d_l.data_ptr(), n, stream) if not args.no_tune else None
    for _ in range(args.warmup):
        one_pass()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        one_pass()
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pass_ms = [e0.elapsed_time(e1) for e0, e1 in evs]
    ms_per_step = wall * 1e3 / args.steps
    steps_per_pass = int(d_st.to(torch.int64).sum().item())
    arrivals = int(d_cnt.to(torch.int64).sum().item())
    value = steps_per_pass / (ms_per_step / 1e3)
    kern_ms = sum(pass_ms) / len(pass_ms)
    achieved = F_STEP * steps_per_pass / (kern_ms / 1e3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed stratified profile)",
        "config": {"workload": "sph fan 450 incl x 360 az = 162000 rays, 5 bounces, 1000 km, 18 eq",
                   "rays": n, "ray_steps_per_pass": steps_per_pass, "arrivals_per_pass": arrivals,
                   "scheduler": ({"slice_steps": tuning.slice_steps, "slices_per_graph": tuning.slices_per_graph,
                                  "tuned_pass_ms": tuning.pass_ms} if tuning else "defaults")},
        "roofline": {"bound": "mfma", "pipe": "fp64 (VALU; MI355X FP64 vector = matrix peak)",
                     "bound_note": "compute-bound on the FP64 VALU pipe (no MFMA instructions); 'mfma' is the "
                                   "compute roof of the hbm|mfma enumeration, priced at the FP64 peak",
                     "per": "one pass = all slice_kernel launches of the fan (ring_init, slice batches, "
                            "arrival_kernel)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": TRAFFIC_BYTES,
                     "kernel_ms_events": kern_ms, "flops_per_step": F_STEP,
                     "transcendentals_per_step": TRANS_STEP,
                     "executed_flops_per_step": F_STEP_EXEC,
                     "frac_executed": F_STEP_EXEC * steps_per_pass / (kern_ms / 1e3) / PEAK_FP64,
                     "flops_note": "achieved/frac count the reference algorithm's FP64 work per ray-step "
                                   "(op-counted oracle); frac_executed counts the FP64 work the hardware runs "
                                   "(rocprof), smaller because Sutherland-Bass comes from a table"},
    }
    host = host_cpu()
    ok = True
    if not args.no_cpu:
        base, idx, ref = cpu_baseline(units, L, host["cores"])
        base.update({k: host[k] for k in ("nproc", "affinity", "cpu_model", "sockets")})
        base["gpu_over_cpu"] = value / base["value"]
        line["cpu_baseline"] = base
        raw = d_out.cpu().numpy().view(ARRIVAL_DTYPE)
        cnt = d_cnt.cpu().numpy()
        nb = bounces + 1
        gpu_rows = np.concatenate([raw[i * nb: i * nb + cnt[i]] for i in idx])
        line["sample_parity"] = sample_parity(gpu_rows, ref, True, (30.0, 0.0))
        ok = ok and line["sample_parity"]["ok"]
    if not args.no_eig:
        line["eigenray"] = eig_leg(eng, not args.no_cpu, host["cores"])
    if not args.no_serve:
        line["served"] = serve_leg(eng, serve_trace())
    if not args.no_rngdep:
        line["rngdep"] = rngdep_leg(eng, not args.no_cpu, host["cores"])
        line["rngdep3d"] = rngdep3d_leg(eng, not args.no_cpu, host["cores"])
        for leg in ("rngdep", "rngdep3d"):
            if "sample_parity" in line[leg]:
                ok = ok and line[leg]["sample_parity"]["ok"]
    if not args.no_wvfrm:
        line["wvfrm"] = wvfrm_leg(eng, not args.no_cpu, host["cores"])
        if "sample_parity" in line["wvfrm"]:
            ok = ok and line["wvfrm"]["sample_parity"]["ok"]
    eng.close()
    print(json.dumps(line))
    if not ok:
        # the full-size passes disagree with the oracle on their sampled rays
        print("bench: sample parity failed (see sample_parity in the line)", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
