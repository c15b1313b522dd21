"""Shared fixtures.  Tests marked ``gpu`` need an MI355X (gfx950) and the
built libdosehip.so; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libdosehip.so")


@pytest.fixture(scope="session")
def proton_model():
    from dosecuda_amd.beam_models import ProtonBeamModel
    return ProtonBeamModel()


@pytest.fixture(scope="session")
def photon_model():
    from dosecuda_amd.beam_models import PhotonBeamModel
    return PhotonBeamModel()


@pytest.fixture(scope="session")
def gpu_ok():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def rel_err(a, b):
    """(max |a-b|, rms |a-b|) relative to max |b| (Dmax of the oracle)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    dmax = float(np.max(np.abs(b))) or 1.0
    d = np.abs(a - b)
    return float(d.max()) / dmax, float(np.sqrt(np.mean(d * d))) / dmax
