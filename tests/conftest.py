import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def capi():
    from mvkit_amd import capi as c
    c.lib()
    return c


@pytest.fixture(scope="session")
def scene_plain():
    from mvkit_amd import synth
    return synth.workload_plain_640()


@pytest.fixture(scope="session")
def scene_rot():
    from mvkit_amd import synth
    return synth.workload_rot_2048()


@pytest.fixture(scope="session")
def scene_small():
    from mvkit_amd import synth
    return synth.small_rot()


# Tolerances of the parity contract.
POS_TOL = 0.02                      # px
ANGLE_TOL = math.radians(0.01)      # rad
SCORE_TOL = 1e-4
