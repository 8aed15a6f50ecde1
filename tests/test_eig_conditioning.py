"""Conditioning of the 3D eigenray iteration on the 512-pair batch (CPU).

tests/golden/eig3d_sensitivity.npz (tests/golden/make_eig_sensitivity.py)
classifies every start of find_eigenray (geoac.eigenray.3d.cpp:146-366) in
the batch as determined or not determined at FP64 precision, by running the
oracle's iteration from the start and from starts moved by 1e-13 and 1e-11
relative.  These tests check that
- the stored probe results are what the oracle computes now (a sample of
  starts, including every start of the receivers named below);
- the criterion flags the iterations known to be chaotic: receiver 510
  direct accepts from every probed start but its iterates separate at the
  second iterate;
- the oracle compiled with and without FMA contraction (eig3d_base.npz,
  eig3d_base_fma.npz) disagree on the eigenray count only in groups that hold
  a start the criterion flags.  The GPU test of the batch holds the engine to
  the same rule: identical counts in every other group.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)

import make_eig_sensitivity as sens  # noqa: E402


@pytest.fixture(scope="module")
def table():
    return np.load(os.path.join(GOLD, "eig3d_sensitivity.npz"))


def _key(s):
    return int(s[0]), int(s[1])


def unstable_groups(table):
    return {_key(s) for s, u in zip(table["start"], table["unstable"]) if u}


def _counts(path):
    c = {}
    for r in np.load(path)["eigenrays"]:
        k = (int(r["launch_idx"]), int(r["n_bounce"]))
        c[k] = c.get(k, 0) + 1
    return c


def test_table_matches_starts_and_criterion(table):
    starts = sens.load_starts()
    assert len(starts) == len(table["start"]) > 400
    assert np.array_equal(np.array(starts, dtype=np.float64), table["start"])
    assert np.array_equal(sens.classify(table["accepted"], table["max_sep"]), table["unstable"])
    assert 0 < table["unstable"].sum() < len(starts)


def test_probes_reproduce(oracle, table):
    starts = sens.load_starts()
    pick = [i for i, s in enumerate(starts) if _key(s) in ((510, 0), (16, 2))] + list(range(0, len(starts), 76))
    for i in pick:
        got = sens.probe(starts[i])
        acc = np.array([g[0] for g in got])
        sep = np.array([g[3] for g in got])
        assert np.array_equal(acc, table["accepted"][i]), (starts[i], acc)
        assert np.allclose(sep, table["max_sep"][i], rtol=1e-6, atol=1e-15), (starts[i], sep)
        assert bool(sens.classify(acc[None], sep[None])[0]) == bool(table["unstable"][i])


def test_chaotic_iteration_is_flagged(table):
    rows = [i for i, s in enumerate(table["start"]) if _key(s) == (510, 0) and int(s[2]) == 0]
    assert len(rows) == 1
    i = rows[0]
    assert table["accepted"][i].all() and table["max_sep"][i].max() > 1e-6 and table["unstable"][i]
    stable = [i for i in range(len(table["start"])) if not table["unstable"][i]]
    assert max(table["max_sep"][i].max() for i in stable) <= 1e-6


def test_fma_build_differs_only_in_unstable_groups(table):
    plain = _counts(os.path.join(GOLD, "eig3d_base.npz"))
    fma = _counts(os.path.join(GOLD, "eig3d_base_fma.npz"))
    flips = sorted(k for k in set(plain) | set(fma) if plain.get(k, 0) != fma.get(k, 0))
    ung = unstable_groups(table)
    assert flips, "the two builds agree everywhere: the fixture no longer shows the sensitivity"
    assert all(k in ung for k in flips), [k for k in flips if k not in ung]
