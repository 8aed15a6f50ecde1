"""Arrival-set comparison at the stated parity bar.

Arrivals are keyed by (launch index, bounce).  The key sets must be equal;
range, turning height and travel time must agree within 1e-6 relative; TL
and absorption within 0.01 dB.  Positions are compared through the range
from the source (great-circle for the spherical geometry) and directly to
1e-6 relative of that range.
"""
from __future__ import annotations

import numpy as np

R0 = 6370.0


def gc_range(lat_deg, lon_deg, lat0_deg, lon0_deg, r0=R0):
    la, lo, la0, lo0 = map(np.radians, (lat_deg, lon_deg, lat0_deg, lon0_deg))
    h = np.sin((la - la0) / 2) ** 2 + np.cos(la) * np.cos(la0) * np.sin((lo - lo0) / 2) ** 2
    return 2 * r0 * np.arcsin(np.sqrt(h))


def keyed(arr):
    return {(int(a["launch_idx"]), int(a["n_bounce"])): a for a in arr}


def keyed_eig(arr):
    """Eigenrays keyed by (receiver, bounces, order found within those)."""
    out, seen = {}, {}
    for a in arr:
        k = (int(a["launch_idx"]), int(a["n_bounce"]))
        j = seen.get(k, 0)
        seen[k] = j + 1
        out[k + (j,)] = a
    return out


def _incl(a):
    return a["incl"] if "incl" in a.dtype.names else a["incl_launch"]


def _az(a):
    return a["az"] if "az" in a.dtype.names else a["az_launch"]


def compare_eig(gpu, ref, geom_sph, src=(0.0, 0.0), rel=1e-6, db=0.01):
    """compare() for eigenray sets; launch angles must also agree to rel."""
    fails, summ = compare(gpu, ref, geom_sph, src, rel, db, key=keyed_eig)
    g, r = keyed_eig(gpu), keyed_eig(ref)
    worst = 0.0
    bad = []
    for k in sorted(set(g) & set(r)):
        e = max(abs(_incl(g[k]) - _incl(r[k])) / max(abs(_incl(r[k])), 1e-12),
                abs(_az(g[k]) - _az(r[k])) / 360.0)
        worst = max(worst, e)
        if e > rel:
            bad.append((k, e))
    summ["worst"]["launch"] = worst
    if bad:
        fails["launch"] = bad[:10]
    return fails, summ


def compare(gpu, ref, geom_sph, src=(0.0, 0.0), rel=1e-6, db=0.01, near=None, key=keyed):
    """Return a dict of failures; empty means parity holds.

    near: optional set of launch indices reported separately (threshold cases)."""
    g, r = key(gpu), key(ref)
    fails = {}
    missing = sorted(set(r) - set(g))
    extra = sorted(set(g) - set(r))
    if missing:
        fails["missing"] = missing[:20]
    if extra:
        fails["extra"] = extra[:20]
    worst = {"range": 0.0, "pos": 0.0, "time": 0.0, "turn_ht": 0.0, "tl": 0.0, "absorption": 0.0}
    bad = []
    for k in sorted(set(g) & set(r)):
        a, b = g[k], r[k]
        if geom_sph:
            ra = gc_range(a["pos0"], a["pos1"], *src)
            rb = gc_range(b["pos0"], b["pos1"], *src)
            dpos = R0 * np.radians(np.hypot(a["pos0"] - b["pos0"], (a["pos1"] - b["pos1"]) * np.cos(np.radians(b["pos0"]))))
        else:
            ra = np.hypot(a["pos0"] - src[0], a["pos1"] - src[1])
            rb = np.hypot(b["pos0"] - src[0], b["pos1"] - src[1])
            dpos = np.hypot(a["pos0"] - b["pos0"], a["pos1"] - b["pos1"])
        e = {
            "range": abs(ra - rb) / max(abs(rb), 1e-12),
            "pos": dpos / max(abs(rb), 1e-12),
            "time": abs(a["time"] - b["time"]) / max(abs(b["time"]), 1e-12),
            "turn_ht": abs(a["turn_ht"] - b["turn_ht"]) / max(abs(b["turn_ht"]), 1e-12),
            "tl": abs(a["tl"] - b["tl"]),
            "absorption": abs(a["absorption"] - b["absorption"]),
        }
        for kk, v in e.items():
            worst[kk] = max(worst[kk], v)
        if (e["range"] > rel or e["pos"] > rel or e["time"] > rel or e["turn_ht"] > rel
                or e["tl"] > db or e["absorption"] > db):
            bad.append((k, e))
    if bad:
        fails["values"] = bad[:10]
    fails_summary = {"worst": worst, "n_common": len(set(g) & set(r))}
    return fails, fails_summary


def compare_paths(gpu, ref, geom_sph, src=(0.0, 0.0, 0.0), rel=1e-6, db=0.01):
    """Path-row sets (ga_path_pt / go_path_pt): keys (ray, leg, step, kind)
    must be equal and in the same order; the row position within rel of the
    range from the source (plus 1 km), the altitude within rel of the same
    scale, time within rel, TL and absorption within db."""
    fails = {}
    kg = np.stack([gpu[f] for f in ("launch_idx", "leg", "step", "kind")], axis=1) if len(gpu) else np.zeros((0, 4))
    kr = np.stack([ref[f] for f in ("launch_idx", "leg", "step", "kind")], axis=1) if len(ref) else np.zeros((0, 4))
    if kg.shape != kr.shape or not (kg == kr).all():
        sg, sr = set(map(tuple, kg.tolist())), set(map(tuple, kr.tolist()))
        fails["keys"] = {"n_gpu": len(kg), "n_ref": len(kr), "missing": sorted(sr - sg)[:10],
                         "extra": sorted(sg - sr)[:10]}
        return fails, {"n": 0}
    if geom_sph:
        rng = gc_range(ref["pos0"], ref["pos1"], src[0], src[1])
        dpos = R0 * np.radians(np.hypot(gpu["pos0"] - ref["pos0"],
                                        (gpu["pos1"] - ref["pos1"]) * np.cos(np.radians(ref["pos0"]))))
    else:
        rng = np.hypot(ref["pos0"] - src[0], ref["pos1"] - src[1])
        dpos = np.hypot(gpu["pos0"] - ref["pos0"], gpu["pos1"] - ref["pos1"])
    scale = rng + 1.0
    e = {
        "pos": dpos / scale,
        "z": np.abs(gpu["z"] - ref["z"]) / scale,
        "time": np.abs(gpu["time"] - ref["time"]) / np.maximum(np.abs(ref["time"]), 1e-9),
        "tl": np.abs(gpu["tl"] - ref["tl"]),
        "absorption": np.abs(gpu["absorption"] - ref["absorption"]),
    }
    worst = {k: float(v.max()) if len(v) else 0.0 for k, v in e.items()}
    bad = (e["pos"] > rel) | (e["z"] > rel) | (e["time"] > rel) | (e["tl"] > db) | (e["absorption"] > db)
    if bad.any():
        idx = np.nonzero(bad)[0][:10]
        fails["values"] = [(tuple(kg[i].tolist()), {k: float(v[i]) for k, v in e.items()}) for i in idx]
        fails["n_bad"] = int(bad.sum())
    return fails, {"worst": worst, "n": len(kg)}
