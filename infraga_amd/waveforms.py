"""Source waveforms of -wnl_wvfrm (src/util/waveforms.cpp:37-145, the
reference's wvfrm::build_wvfrm / load_wvfrm), for callers of
Engine.wnl_wvfrm.

build(option, p0, t0, alpha, length, n_rise, n_period) returns (t, p):
- "impulse": p0 x^a (1 - x / (1 + a)) e^-x, x = t / t0, normalised at its
  peak x_pk = (1 + a) - sqrt(1 + a), zero before t = 0 (a raised to 0.01);
- "Uwave" / "u-wave" / "U-wave": the impulse's central difference with
  eps = 0.01 s, scaled by 1.126 times its slope at 0 (a raised to 2);
- "Nwave" / "n-wave" / "N-wave": -p0 (t - t0/2) / t0 on |t - t0/2| <= t0;
- any other option: the impulse with the shorter record length.
The record spans t0 len max(x_pk / n_rise, 2 pi' sqrt(1 + a) / n_period)
seconds (min for an unrecognised option; pi' = 3.14159 as the reference
writes it), (t0 / n_period) len for the N-wave, sampled at
t_n = -T / 2 + T n / len.
load(path) reads a two-column "t p" file (lines starting with '#'
skipped), as load_wvfrm does.
"""
from __future__ import annotations

import math

import numpy as np


def _impulse(t, t0, a):
    if t < 0.0:
        return 0.0
    x = t / t0
    x_pk = (a + 1.0) - math.sqrt(a + 1.0)
    f = math.pow(x, a) * (1.0 - x / (1.0 + a)) * math.exp(-x)
    return f / (math.pow(x_pk, a) * (1.0 - x_pk / (1.0 + a)) * math.exp(-x_pk))


def build(option="impulse", p0=10.0, t0=0.1, alpha=4.0, length=8192, n_rise=10, n_period=40):
    """A built-in source waveform: (t, p) arrays of `length` samples."""
    imp = option.startswith("impulse")
    uw = option.startswith(("Uwave", "u-wave", "U-wave"))
    nw = option.startswith(("Nwave", "n-wave", "N-wave"))
    if nw:
        duration = (t0 / n_period) * length
    else:
        alpha = max(alpha, 2.0 if uw else 0.01)
        a = alpha
        r = ((1.0 + a) - math.sqrt(1.0 + a)) / n_rise
        q = 2.0 * 3.14159 * math.sqrt(1 + a) / n_period
        duration = t0 * length * (max(r, q) if (imp or uw) else min(r, q))
    t = np.empty(length)
    p = np.empty(length)
    eps = 1.0e-2
    for n in range(length):
        tn = -duration / 2.0 + duration * n / length
        t[n] = tn
        if nw:
            p[n] = -p0 * ((tn - t0 / 2.0) / t0) if abs(tn - t0 / 2.0) <= t0 else -p0 * 0.0
        elif uw:
            v = p0 * (_impulse(tn + eps, t0, alpha) - _impulse(tn - eps, t0, alpha)) / (2.0 * eps)
            p[n] = v / (1.126 * (_impulse(eps, t0, alpha) - _impulse(-eps, t0, alpha)) / (2.0 * eps))
        else:
            p[n] = p0 * _impulse(tn, t0, alpha)
    return t, p


def load(path):
    """A waveform file: (t, p) of its rows not starting with '#'."""
    rows = [l.split() for l in open(path) if not l.startswith("#") and l.strip()]
    a = np.array([[float(r[0]), float(r[1])] for r in rows])
    return a[:, 0].copy(), a[:, 1].copy()
