"""Pins of the oracle's weakly non-linear waveform code (ga_wvfrm.c:
waveforms.cpp, wnl_wvfrm of geoac.eqset.{sph,3d}.strat.cpp and
run_wnl_wvfrm of main.infraga.{sph,3d}.cpp), whose reference build needs
FFTW (absent here).

- The transform: go_fft / go_rfft / go_irfft equal numpy.fft's (FFTW's sign
  convention, unnormalised) to rounding, lengths 2..8192.
- The built-in waveforms: the impulse peaks at p0 at x_pk = (1 + a) -
  sqrt(1 + a) and vanishes before t = 0; the U-wave is the impulse's
  central difference scaled to p0 / 1.126 of its slope at 0; the N-wave is
  the ramp -p0 (t - t0/2) / t0 on |t - t0/2| <= t0; record durations as the
  formulas give.
- Linear propagation (p0 so small that the nonlinear terms vanish) in a
  homogeneous medium: every evolution step is 1 km (the step cap), the
  spectrum is multiplied by filter2 and by 10^(-alpha(f) ds / 10) per step,
  the time series carries the Tukey window of the last step, and the output
  is scaled by sqrt(D0 / D_k) = S0 / S_k of spherical spreading; checked
  against numpy transforms of the source waveform with and without
  absorption.
- Nonlinear propagation (Burgers, no absorption, flat filters) in a
  homogeneous medium: du/ds = beta(s) u du/dt with beta(s) = beta0 S0 / S
  (spherical spreading, beta0 = 1.2 p0 / (rho c^3) in the reference's units)
  has the implicit Earnshaw solution u(t) = u0(t + sigma u), sigma the
  integral of beta over the evolution (beta0 S0 ln(S / S0)); a Gaussian pulse up to
  half its shock distance matches it to 1e-3 of the peak, and the sign of the
  steepening (the compression moves earlier) is fixed.
The caustic phase shift and the beta_caustic branch are parity unpinned
(no closed form here); they follow geoac.eqset.sph.strat.cpp:1058-1068.
"""
import math

import numpy as np
import pytest

from ga_oracle_py import GEOM_3D, Medium, WvParams, wv_build, wv_fft, wv_irfft, wv_rfft, wv_run


@pytest.mark.parametrize("n", [2, 8, 64, 1024, 8192])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    scale = np.abs(np.fft.fft(x)).max()
    assert np.abs(wv_fft(oracle, x) - np.fft.fft(x)).max() < 1e-13 * scale * math.log2(n + 1)
    assert np.abs(wv_fft(oracle, x, True) - n * np.fft.ifft(x)).max() < 1e-13 * scale * math.log2(n + 1)
    xr = rng.normal(size=n)
    R = np.fft.rfft(xr)
    assert np.abs(wv_rfft(oracle, xr) - R).max() < 1e-13 * np.abs(R).max() * math.log2(n + 1)
    # c2r: unnormalised, the imaginary parts of bins 0 and n/2 ignored (FFTW)
    R2 = R.copy()
    R2[0] += 5j
    R2[-1] += 3j
    assert np.abs(wv_irfft(oracle, R2, n) - n * xr).max() < 1e-12 * n * np.abs(xr).max()


def test_builtin_waveforms(oracle):
    p0, t0, a = 10.0, 0.1, 4.0
    t, p, a_out = wv_build(oracle, "impulse", p0=p0, t0=t0, alpha=a, length=4096)
    assert a_out == a
    x_pk = (1 + a) - math.sqrt(1 + a)
    dur = t0 * 4096 * max(x_pk / 10, 2.0 * 3.14159 * math.sqrt(1 + a) / 40)
    assert t[0] == pytest.approx(-dur / 2.0) and t[1] - t[0] == pytest.approx(dur / 4096)
    assert (p[t < 0] == 0).all() and p.max() <= p0 * (1 + 1e-12)
    f = lambda x: x ** a * (1 - x / (1 + a)) * math.exp(-x)
    j = int(np.searchsorted(t, 0.3))
    assert p[j] == pytest.approx(p0 * f(t[j] / t0) / f(x_pk), rel=1e-12)
    # the peak sits at t0 x_pk: sampled values next to it straddle p0
    i = np.argmax(p)
    assert abs(t[i] - t0 * x_pk) <= (t[1] - t[0])
    tu, pu, au = wv_build(oracle, "Uwave", p0=p0, t0=t0, alpha=1.0, length=4096)
    assert au == 2.0                                # the U-wave raises alpha to 2
    imp = lambda tt: 0.0 if tt < 0 else f2(tt / t0)
    x2 = 3 - math.sqrt(3)
    f2 = lambda x: x ** 2 * (1 - x / 3) * math.exp(-x) / (x2 ** 2 * (1 - x2 / 3) * math.exp(-x2))
    eps = 1e-2
    j = len(tu) // 2 + 40
    exp_u = p0 * (imp(tu[j] + eps) - imp(tu[j] - eps)) / (2 * eps) / (1.126 * (imp(eps) - imp(-eps)) / (2 * eps))
    assert pu[j] == pytest.approx(exp_u, rel=1e-12)
    tn, pn, _ = wv_build(oracle, "Nwave", p0=p0, t0=1.0, length=4096)
    assert tn[1] - tn[0] == pytest.approx(1.0 / 40)
    inside = np.abs(tn - 0.5) <= 1.0
    np.testing.assert_allclose(pn[inside], -p0 * (tn[inside] - 0.5), rtol=1e-12, atol=1e-12)
    assert (pn[~inside] == 0).all()


def homogeneous(tweak):
    z = np.arange(0.0, 120.0 + 1e-9, 0.5)
    return z, 0.34 + 0.0 * z, 0.0 * z, 0.0 * z, 1.2e-3 + 0.0 * z


def gaussian(length, dt, tau, p0):
    t = (np.arange(length) - length // 2) * dt
    return t, p0 * np.exp(-(t / tau) ** 2)


def run_homog(oracle, p0, tweak, wv, incl=40.0, ref=1.0, freq=0.1):
    m = Medium(oracle, GEOM_3D, *homogeneous(tweak))
    oracle.lib.go_h_set_tweak(m.h, tweak)
    p = m.params(True)
    t, ps = gaussian(wv.len, 0.02, 0.5, p0)
    r = wv_run(m, p, wv, (0.0, 0.0, 0.0), incl, 90.0, 0, t, ps, freq=freq, wvfrm_ref=ref)
    return m, t, ps, r


def tukey(n, tp=0.8):
    w = np.ones(n)
    m = np.arange(n)
    lo = m < tp * (n - 1) / 2.0
    hi = m > (n - 1) * (1.0 - tp / 2.0)
    w[lo] = 0.5 * (1 + np.cos(np.pi * (2.0 * m[lo] / (tp * (n - 1)) - 1.0)))
    w[hi] = 0.5 * (1 + np.cos(np.pi * (2.0 * m[hi] / (tp * (n - 1)) - 2.0 / tp + 1.0)))
    return w


def path_lengths(r, t_src, c=0.34):
    """S0 (source to the reference row) from the wvfrm_init times, S_{k-1}
    from the arrival time, the evolution's extent s_tot = S_{k-1} - S0 and
    the output row's distance S_k = S_{k-1} + ds_max (straight rays at c)."""
    S0 = c * (r["t_init"][0] - t_src[0])
    Sk1 = c * r["arrival"][2]
    return S0, Sk1, Sk1 - S0, Sk1 + 0.05


def test_linear_propagation_closed_form(oracle):
    wv = WvParams.default(len=4096)
    m, t, ps, r = run_homog(oracle, 1e-9, 0.0, wv)
    assert r["steps"] > 0 and not r["ok"]          # the ray leaves through the ceiling
    S0, Sk1, s_tot, Sk = path_lengths(r, t)
    n_steps = math.ceil(s_tot)                      # 1 km steps (the cap) until s >= s_tot
    assert r["steps"] == n_steps
    n = wv.len
    fl = n // 2 + 1
    k = np.arange(fl)
    filt2 = (1.0 + (k / (wv.filt2_g * fl)) ** wv.filt2_n1) ** (-wv.filt2_n2)
    U = np.fft.rfft(ps / ps.max()) * filt2 ** n_steps
    u = np.fft.irfft(U, n) * tukey(n)
    scale = S0 / Sk                                 # sqrt(D0 / D_k), D = s^2 cos(theta)
    np.testing.assert_allclose(r["p_out"], ps.max() * scale * u, rtol=0, atol=2e-5 * ps.max() * scale)
    np.testing.assert_allclose(r["t_out"] - t, r["arrival"][2], rtol=0, atol=1e-9)


def test_linear_absorption_per_step(oracle):
    """With absorption: each step multiplies by 10^(-alpha(z, f) ds / 10)
    at the step's start; straight up the ray (incl 90) z = S, so the product
    over steps is computed from go_sb_alpha at the steps' heights."""
    wv = WvParams.default(len=4096)
    m, t, ps, r = run_homog(oracle, 1e-9, 1.0, wv, incl=89.999)
    S0, Sk1, s_tot, Sk = path_lengths(r, t)
    n_steps = math.ceil(s_tot)
    assert r["steps"] == n_steps
    n = wv.len
    fl = n // 2 + 1
    k = np.arange(fl)
    filt2 = (1.0 + (k / (wv.filt2_g * fl)) ** wv.filt2_n1) ** (-wv.filt2_n2)
    freq = k / (t[1] - t[0]) / n
    sel = slice(0, 400)                             # the pulse's band (|U| > 1e-12 of its peak)
    att = np.zeros(fl)
    sin_i = math.sin(math.radians(89.999))
    for j in range(n_steps):
        z = (S0 + j * 1.0) * sin_i                  # the ray point at the step's start
        a = np.array([oracle.lib.go_sb_alpha(m.h, 0.0, 0.0, z, f) for f in freq[sel]])
        att[sel] += -a * 1.0 / 10.0
    U = np.zeros(fl, dtype=complex)
    U[sel] = np.fft.rfft(ps / ps.max())[sel] * 10.0 ** att[sel] * filt2[sel] ** n_steps
    u = np.fft.irfft(U, n) * tukey(n)
    scale = S0 / Sk
    np.testing.assert_allclose(r["p_out"], ps.max() * scale * u, rtol=0, atol=2e-5 * ps.max() * scale)


def earnshaw(t, f, sigma):
    """u = f(t + sigma u) for the source shape f, solved per sample by
    bisection on the characteristic (pre-shock: a unique root)."""
    lo, hi = np.full_like(t, -1.5), np.full_like(t, 1.5)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        up = mid - f(t + sigma * mid) > 0
        hi = np.where(up, mid, hi)
        lo = np.where(up, lo, mid)
    return 0.5 * (lo + hi)


def test_nonlinear_steepening_is_earnshaw(oracle):
    wv = WvParams.default(len=4096, filt1_g=1e12, filt2_g=1e12)
    p0, c, rho = 230.0, 0.34, 1.2e-3
    m, t, ps, r = run_homog(oracle, p0, 0.0, wv, incl=40.0, ref=30.0)
    S0, Sk1, s_tot, Sk = path_lengths(r, t)
    beta0 = 1.2 * p0 / ((rho * 1.0e3) * (c * 1.0e3) ** 2) * (1.0 * c) / (c * c)
    # the steps follow the evolving spectrum (ds = ds_wvfrm / (pi Uf_max beta) < 1 km
    # here); their trapezoid sum of beta is the integral to well below 1e-4
    assert r["steps"] > s_tot
    sigma = beta0 * S0 * math.log((S0 + s_tot) / S0)
    shape = lambda tt: np.exp(-(tt / 0.5) ** 2)     # the source pulse over its peak
    grad = math.sqrt(2.0 / math.e) / 0.5
    assert 0.3 < sigma * grad < 0.7                 # well before the shock (sigma |u0'| = 1)
    u = r["p_out"] / (p0 * S0 / Sk)
    exp = earnshaw(t, shape, sigma)
    err = np.abs(u - exp).max()
    assert err < 1e-3, err
    # the compression arrives earlier: the peak moves to negative t
    assert t[np.argmax(u)] < -0.5 * sigma
    assert np.abs(u - earnshaw(t, shape, -sigma)).max() > 100 * err


@pytest.mark.parametrize("option,alpha", [("impulse", 4.0), ("Uwave", 1.0), ("Nwave", 4.0), ("other", 3.0)])
def test_package_waveforms_match_oracle(oracle, option, alpha):
    """infraga_amd.waveforms.build (the package's builder for callers of
    Engine.wnl_wvfrm) equals the oracle's waveforms.cpp restatement to
    rounding, for the three built-in shapes and an unrecognised option."""
    from infraga_amd import waveforms
    t, p = waveforms.build(option, p0=300.0, t0=0.2, alpha=alpha, length=2048, n_rise=8, n_period=30)
    to, po, _ = wv_build(oracle, option, p0=300.0, t0=0.2, alpha=alpha, length=2048, n_rise=8, n_period=30)
    np.testing.assert_allclose(t, to, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(p, po, rtol=1e-13, atol=1e-12)


def test_out_step_dumps(oracle, base_units):
    """wvfrm_out_step snapshots (the sc_spectrum / sc_waveform dumps of
    wnl_wvfrm): labels are multiples of out_step in increasing order, one per
    crossing, the first one out_step ceil(wvfrm_ref / out_step) (the first
    leg starts at s_prev = wvfrm_ref); each snapshot is the state at the top
    of a step after the first, so its pressure is the windowed inverse
    transform of its spectrum (u = IFFT(U) w_tukey, the end of the previous
    step); and the dumps do not change the run."""
    from ga_oracle_py import GEOM_SPH, Medium, WvParams, wv_run
    from infraga_amd import waveforms
    m = Medium(oracle, GEOM_SPH, *base_units)
    p = m.params(True)
    length = 1024
    t, ps = waveforms.build("impulse", 500.0, 0.1, 4.0, length, 10, 40)
    wv = WvParams.default(len=length)
    plain = wv_run(m, p, wv, (30.0, 0.0, 0.0), 8.0, 90.0, 1, t, ps)
    out_step = 50.0
    r = wv_run(m, p, wv, (30.0, 0.0, 0.0), 8.0, 90.0, 1, t, ps, dump=(out_step, 200))
    assert np.array_equal(r["p_out"], plain["p_out"]) and r["steps"] == plain["steps"]
    d = r["dumps"]
    assert r["dump_count"] == len(d) > 5
    labels = np.array([x[0] for x in d])
    assert np.all(np.diff(labels) > 0)
    assert np.allclose(labels / out_step, np.round(labels / out_step), rtol=0, atol=1e-9)
    lab0, u0, U0 = d[0]
    assert lab0 == out_step * math.ceil(1.0 / out_step)        # s = 0, s_prev = wvfrm_ref = 1 km
    tp = 0.8
    mm = np.arange(length)
    tk = np.where(mm < tp * (length - 1) / 2.0, 0.5 * (1.0 + np.cos(np.pi * (2.0 * mm / (tp * (length - 1)) - 1.0))),
                  np.where(mm > (length - 1) * (1.0 - tp / 2.0),
                           0.5 * (1.0 + np.cos(np.pi * (2.0 * mm / (tp * (length - 1)) - 2.0 / tp + 1.0))), 1.0))
    for lab, u, U in d:
        assert np.allclose(u, np.fft.irfft(U, length) * tk, rtol=0, atol=1e-9 * np.abs(u).max())
