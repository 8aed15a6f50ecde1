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

Prints one JSON line.  --gpus N with N > 1 exits: this build runs one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_STEP = 5690.9          # algorithmic FP64 flops per ray-step (oracle/opcount.cpp, 400-ray sample)
# HBM bytes per fan_kernel dispatch on this workload, from rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE passes scaled by a 256 MiB copy (factors 2.00 and
# 1.00): 65.7 MB read + 657.5 MB written (profiles/fan_kernel_sph162k_pmc.md)
TRAFFIC_BYTES = 65.7e6 + 657.5e6
TRANS_STEP = 167.8       # transcendental calls per ray-step (same count, not included in F_STEP)
PEAK_FP64 = 78.6e12      # MI355X FP64 vector / matrix peak
METRIC = "ray-steps/s on 162k-ray sph fan (FP64)"


def workload():
    from infraga_amd import synth
    units = synth.to_engine_units(*synth.perturbed_profile())
    L = synth.fan_launches(0.1, 45.0, 0.1, 0.0, 359.0, 1.0)
    return units, L


def cpu_baseline(units, L, target_s=12.0):
    """The oracle restatement (one host core) on a bounded, evenly strided
    sample of the same launch list."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import GEOM_SPH, Medium, Oracle
    m = Medium(Oracle(), GEOM_SPH, *units)
    p = m.params(True)
    n = 24
    total_steps, total_t, used = 0, 0.0, 0
    idx = list(range(0, len(L), max(1, len(L) // 480)))
    while used < len(idx) and total_t < target_s:
        sel = L[idx[used: used + n]]
        t0 = time.perf_counter()
        _, steps = m.fan(sel, p, bounces=5, src=(30.0, 0.0, 0.0))
        total_t += time.perf_counter() - t0
        total_steps += steps
        used += len(sel)
    return {"value": total_steps / total_t, "unit": "ray-steps/s", "cores": 1, "kind": "port",
            "sample": f"{used} rays (every {max(1, len(L) // 480)}th of the 162k launch list), "
                      f"{total_steps} ray-steps in {total_t:.1f} s, oracle restatement, 1 thread"}


def eig_cpu_baseline(pair=6):
    """The oracle's serial search (one host core) for one pair of the batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from ga_oracle_py import GEOM_3D, Medium, Oracle
    from infraga_amd import synth
    m = Medium(Oracle(), GEOM_3D, *synth.to_engine_units(*synth.base_profile()))
    R = synth.eig_pairs()
    t0 = time.perf_counter()
    arr, steps = m.eig_search(m.params(True), (0.0, 0.0, 0.0), R[pair], bnc_min=0, bnc_max=3)
    dt = time.perf_counter() - t0
    return {"value": dt, "unit": "s per receiver", "cores": 1, "kind": "port",
            "sample": f"pair {pair} ({float(np.hypot(*R[pair])):.0f} km), bounces 0..3: {len(arr)} eigenrays, "
                      f"{steps} ray-steps in {dt:.1f} s, oracle restatement, 1 thread"}


def eig_leg(eng, with_cpu):
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
        out["cpu_baseline"] = eig_cpu_baseline()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eig", action="store_true")
    ap.add_argument("--one-pass", action="store_true",
                    help="one untimed fan pass (for hardware-counter runs); no JSON line")
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
                   "rays": n, "ray_steps_per_pass": steps_per_pass, "arrivals_per_pass": arrivals},
        "roofline": {"bound": "mfma", "pipe": "fp64 (VALU; MI355X FP64 vector = matrix peak)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": TRAFFIC_BYTES,
                     "kernel_ms_events": kern_ms, "flops_per_step": F_STEP,
                     "transcendentals_per_step": TRANS_STEP},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(units, L)
    if not args.no_eig:
        line["eigenray"] = eig_leg(eng, not args.no_cpu)
    eng.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
