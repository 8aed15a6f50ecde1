import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    from ga_oracle_py import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def base_units():
    from infraga_amd import synth
    return synth.to_engine_units(*synth.base_profile())


@pytest.fixture(scope="session")
def pert_units():
    from infraga_amd import synth
    return synth.to_engine_units(*synth.perturbed_profile())


@pytest.fixture(scope="session")
def toyatmo_cols():
    """The reference's examples/ToyAtmo.met (zTuvdp), from the committed fixture."""
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "toyatmo.npz"))["zTuvdp"]


@pytest.fixture(scope="session")
def toyatmo_units(toyatmo_cols):
    from infraga_amd import synth
    return synth.from_zTuvdp(toyatmo_cols)


@pytest.fixture(scope="session")
def engine():
    from infraga_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()
