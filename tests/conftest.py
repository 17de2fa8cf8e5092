import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def solver_huber():
    import cuvslam_amd as ca
    return ca.SchurComplementBundlerGpu(loss=ca.LOSS_HUBER)


@pytest.fixture(scope="session")
def solver_student():
    import cuvslam_amd as ca
    return ca.SchurComplementBundlerGpu(loss=ca.LOSS_STUDENT_T)
