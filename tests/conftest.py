"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU.

The oracle (oracle/) is the checker here: it is loaded by tests only."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgpl_hip.so on the device)")


@pytest.fixture(scope="session")
def kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def c1_design():
    """The small workload (2,000 cells, 64x64 bins), initially placed, Nesterov placer built."""
    import gpl_amd as G
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


@pytest.fixture(scope="session")
def c3_design():
    """The 1M-cell workload (1024 x 1024 bins), initially placed, Nesterov placer built (GPU tests
    only: about ten seconds of host work)."""
    import gpl_amd as G
    d = G.Design.from_config("c3")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()
