import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running statistical or full-size checks")


def nist_lj(cfg: int) -> str:
    return os.path.join(GOLDEN, "nist", "lj", f"lj_sample_config_periodic{cfg}.txt")


def nist_spce(cfg: int) -> str:
    return os.path.join(GOLDEN, "nist", "spce", f"spce_sample_config_periodic{cfg}.txt")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle

    if not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build()
    return pyoracle.Oracle(pyoracle.ORACLE_SO)


@pytest.fixture(scope="session")
def ref():
    from oracle import pyoracle

    if not os.path.exists(pyoracle.REF_SO):
        pytest.skip("reference build (oracle/_ref) not present")
    return pyoracle.Oracle(pyoracle.REF_SO)


@pytest.fixture(scope="session")
def gpu_ok():
    import ctypes

    try:
        hip = ctypes.CDLL("libamdhip64.so")
        cnt = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(cnt)) != 0 or cnt.value < 1:
            pytest.skip("no GPU visible")
    except OSError:
        pytest.skip("HIP runtime not loadable")
    return True


@pytest.fixture()
def ctx(gpu_ok):
    from moleculardynamics_amd import native

    c = native.Context()
    yield c
    c.close()
