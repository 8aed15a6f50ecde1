"""GPU parity of the 2D effective-sound-speed geometry (GA_GEOM_2D,
infraga-2d -prop) against the oracle's go2_fan on the same launch doubles.

Seeded perturbed profile with winds; azimuths along, against and across
the jet; 8 equations with caustics and 4 without; an elevated reflection
altitude; a terrain profile.  Bar: arrival keys exact, range, drift, time,
turning height within 1e-6 relative, TL and absorption within 0.01 dB;
path-row keys exact and values at the same bar (compare_paths).
"""
import os
import subprocess

import numpy as np
import pytest

from infraga_amd import synth
from infraga_amd.engine import GA_GEOM_2D, GA_PATH_CAUSTICS, GA_PATH_RAYS, prop_params
from parity import compare, compare_paths

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def launches():
    return np.array([[th, az] for az in (90.0, -90.0, 30.0, 200.0) for th in np.arange(2.0, 41.0, 3.0)])


def hill():
    r = np.linspace(-50.0, 1500.0, 621)
    return r, 0.8 * np.exp(-((r - 120.0) / 25.0) ** 2) + 0.3 * np.exp(-((r - 400.0) / 60.0) ** 2)


CASES = [
    # name, calc_amp, bounces, refl_alt, terrain
    ("amp8", True, 2, None, False),
    ("noamp4", False, 2, None, False),
    ("refl8", True, 2, 60.0, False),
    ("topo8", True, 2, None, True),
]


@pytest.mark.parametrize("name,amp,bounces,refl,topo", CASES, ids=[c[0] for c in CASES])
def test_2d_fan_matches_oracle(engine, oracle, pert_units, name, amp, bounces, refl, topo):
    from ga_oracle_py import PATH_CAUSTICS, Medium2D
    L = launches()
    aid = engine.load_atmo(*pert_units)
    kw = {}
    if refl is not None:
        kw["refl_alt"] = refl
    if topo:
        engine.topo_load_2d(aid, *hill())
    p = prop_params(GA_GEOM_2D, calc_amp=int(amp), bounces=bounces, **kw)
    flags = GA_PATH_RAYS | (GA_PATH_CAUSTICS if amp else 0)
    gpu, gpts, st = engine.fan_paths(aid, p, L, flags)
    m = Medium2D(oracle, *pert_units, calc_amp=amp, refl_alt=refl if refl is not None else float("nan"),
                 topo=hill() if topo else None)
    ref, rpts, steps = m.fan(L, bounces=bounces, flags=PATH_CAUSTICS if amp else 0)
    engine.release_atmo(aid)
    print(f"\n{name}: rays={len(L)} arrivals gpu={len(gpu)} oracle={len(ref)} rows gpu={len(gpts)} "
          f"oracle={len(rpts)} caustics={int((rpts['kind'] == 1).sum())} steps gpu={st.ray_steps} oracle={steps} "
          f"{st.kernel_ms:.1f} ms")
    assert len(ref) > 0 and st.ray_steps == steps
    fails, summ = compare(gpu, ref, False, src=(0.0, 0.0))
    assert not fails, (fails, summ)
    np.testing.assert_allclose(gpu["incl_arr"], ref["incl_arr"], rtol=0, atol=1e-6)
    pf, ps = compare_paths(gpts, rpts, False, src=(0.0, 0.0, 0.0))
    print(ps)
    assert not pf, (pf, ps)


def test_cli_2d_prop_text(tmp_path, oracle, toyatmo_units):
    """infraga-2d -prop on ToyAtmo: the arrivals and raypaths files (stream
    default precision) equal the oracle's rows to the printed digits."""
    from ga_oracle_py import Medium2D
    met = tmp_path / "toy.met"
    np.savetxt(met, np.load(os.path.join(ROOT, "tests", "golden", "toyatmo.npz"))["zTuvdp"], fmt="%.17g")
    exe = os.path.join(ROOT, "infraga_amd", "bin", "infraga-2d")
    r = subprocess.run([exe, "-prop", str(met), "incl_min=5", "incl_max=35", "incl_step=10", "azimuth=-90",
                        "bounces=1"], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split("\t") for l in open(tmp_path / "toy.arrivals.dat") if l.strip() and not l.startswith("#")]
    m = Medium2D(oracle, *toyatmo_units)
    ref, rp, _ = m.fan([(th, -90.0) for th in (5.0, 15.0, 25.0, 35.0)], bounces=1)
    assert len(rows) == len(ref) > 0
    for row, a in zip(rows, ref):
        vals = [float(x) for x in row]
        exp = [a["incl_launch"], a["az_launch"], a["n_bounce"], a["pos0"], a["time"], a["cel"], a["turn_ht"],
               a["incl_arr"], a["tl"], a["absorption"], a["pos1"]]
        np.testing.assert_allclose(vals, exp, rtol=1e-5, atol=1e-5)
    txt = (tmp_path / "toy.raypaths.dat").read_text().split("\n")
    assert txt[0] == "# r [km]\tz [km]\ttrans. coeff. [dB]\tabsorption [dB]\ttime [s]"
    assert len([l for l in txt[1:] if l.strip()]) == len(rp)
