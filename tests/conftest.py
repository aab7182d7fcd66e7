"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "spi-active_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libspi_hip.so")


@pytest.fixture(scope="session")
def model():
    from spi_amd.model import load_model
    return load_model()


@pytest.fixture(scope="session")
def sim():
    from spi_amd.params import SimParams
    return SimParams()


@pytest.fixture(scope="session")
def theta0(model):
    from spi_amd.params import nominal_params
    return nominal_params(model)


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from spi_amd import _lib
    _lib.lib()
    return torch.device("cuda:0")


def random_state(rng, sim, height=0.30, spread=0.3, vel=0.5):
    """A 37-state near the standing pose with random orientation tilt and velocities."""
    s = np.zeros(37)
    s[0:2] = rng.normal(0, 0.1, 2)
    s[2] = height + rng.normal(0, 0.01)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0, 0.2)
    s[3:6] = axis * np.sin(ang / 2)
    s[6] = np.cos(ang / 2)
    s[7:10] = rng.normal(0, vel, 3)
    s[10:13] = rng.normal(0, vel, 3)
    s[13:25] = np.asarray(sim.q0) + rng.normal(0, spread, 12)
    s[25:37] = rng.normal(0, 2 * vel, 12)
    return s
