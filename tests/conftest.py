import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cornerstone-octree_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhipcstone on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import ORACLE

    if ORACLE is None:
        pytest.fail("oracle/liboracle.so is not built (make -C oracle)")
    return ORACLE


@pytest.fixture(scope="session")
def ref():
    from oracle_py import REF

    if REF is None:
        pytest.skip("oracle/_ref/libcstone_ref.so not built (needs the reference sources)")
    return REF


@pytest.fixture(scope="session")
def hip():
    import hipcstone

    return hipcstone


@pytest.fixture(autouse=True)
def _no_device_faults(request):
    """Every GPU test ends with the device's fault word clear: no kernel guard fired (a test that provokes one reads
    and clears the word itself)."""
    yield
    if request.node.get_closest_marker("gpu") is not None:
        import hipcstone

        faults = hipcstone.device_faults()
        assert faults == 0, f"kernel guards fired during the test: fault bits {faults:#x}"


def golden_path(name):
    return os.path.join(ROOT, "tests", "golden", name)


def load_golden(name):
    with np.load(golden_path(name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}
