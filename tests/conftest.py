import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def orc():
    """The oracle (test infrastructure), built with g++ if its library is missing."""
    from oracle import oracle as o
    if not o.LIB.exists():
        o.build()
    return o


@pytest.fixture(scope="session")
def synth_lib():
    from reconflow_amd import _native
    from reconflow_amd.build import build_synth
    if not (_native.BUILD / "librfsynth.so").exists():
        build_synth()
    return _native.synth()


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(1234)
